/*
 * grid_oracle.c -- restatement of lib/grid/grid.cc (single domain).
 * TEST INFRASTRUCTURE ONLY (see saras_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "saras_oracle.h"

/* grid.cc:678-742 setVCDepth, single rank (vcdLoc == vcdGlo). */
int or_vc_depth(const or_grid_desc *g)
{
    int axes[3], na = 0;
    if (g->dim == 2) { axes[na++] = g->n[0]; axes[na++] = g->n[2]; }
    else { axes[na++] = g->n[0]; axes[na++] = g->n[1]; axes[na++] = g->n[2]; }

    int d = 0;
    for (;;) {
        int div = 1 << d, rem = 0;
        for (int a = 0; a < na; a++) if (axes[a] % div) rem = 1;
        if (rem) break;
        d += 1;
    }
    d -= 1;
    int div = 1 << d, mn = axes[0] / div;
    for (int a = 1; a < na; a++) if (axes[a] / div < mn) mn = axes[a] / div;
    if (mn == 1) d -= 1;
    int allwall = g->dim == 2 ? (!g->periodic[0] && !g->periodic[2])
                              : (!g->periodic[0] && !g->periodic[1] && !g->periodic[2]);
    if (allwall) d -= 1;
    return d;
}

/* Double/single-sided tanh mapping and its metrics (grid.cc:437-460). */
static void tanh_point(int two_sided, double L, double beta, double xi,
                       double *x, double *fx, double *fxx, double *fx2)
{
    double thb = tanh(beta);
    double btl = beta * L;
    if (two_sided) {
        *x = L * (1.0 - tanh(beta * (1.0 - 2.0 * xi)) / thb) / 2.0;
        double ndx = *x / L;
        *fx = thb / (btl * (1.0 - pow((1.0 - 2.0 * ndx) * thb, 2)));
        *fxx = -4.0 * pow(thb, 3) * (1.0 - 2.0 * ndx) / (L * btl * pow(1.0 - pow(thb * (1.0 - 2.0 * ndx), 2), 2));
        *fx2 = pow(*fx, 2.0);
    } else {
        *x = L * (1.0 - tanh(beta * (1.0 - xi)) / thb);
        double ndx = *x / L;
        *fx = thb / (btl * (1.0 - pow((1.0 - ndx) * thb, 2)));
        *fxx = -2.0 * pow(thb, 3) * (1.0 - ndx) / (L * btl * pow(1.0 - pow(thb * (1.0 - ndx), 2), 2));
        *fx2 = pow(*fx, 2.0);
    }
}

/* Level-0 arrays over [-p, n+p-1] (createXiEtaZeta grid.cc:318-340,
 * createUniformGrid :358-391, createTanHypGrid :411-508). */
static void level0_axis(const or_grid_desc *g, int axis, int p,
                        double *xi, double *x, double *fx, double *fxx, double *fx2)
{
    int n = g->n[axis];
    double L = g->len[axis];
    double dxi = 1.0 / (double)n;
    int planar_y = (g->dim == 2 && axis == 1);
    if (planar_y) dxi = 1.0;
    for (int i = -p; i < n + p; i++) {
        xi[i + p] = (double)(2 * i + 1) * dxi / 2;
        x[i + p] = L * xi[i + p];
        fx[i + p] = 1.0 / L;
        fxx[i + p] = 0.0;
        fx2[i + p] = pow(1.0 / L, 2.0);
    }
    if (planar_y) { /* createUniformGrid skips y in PLANAR; metrics still 1/Ly */
        return;
    }
    if (g->gtype[axis] > 0) {
        int two = g->gtype[axis] - 1;
        for (int i = 0; i < n; i++)
            tanh_point(two, L, g->beta[axis], xi[i + p], &x[i + p], &fx[i + p], &fxx[i + p], &fx2[i + p]);
        for (int q = 0; q < p; q++) {          /* lftPts = rgtPts - L */
            int l = -p + q, r = n - p + q;
            x[l + p] = x[r + p] - L;
            fx[l + p] = fx[r + p]; fxx[l + p] = fxx[r + p]; fx2[l + p] = fx2[r + p];
        }
        for (int q = 0; q < p; q++) {          /* rgtPts = L + lftPts */
            int r = n + q, l = q;
            x[r + p] = L + x[l + p];
            fx[r + p] = fx[l + p]; fxx[r + p] = fxx[l + p]; fx2[r + p] = fx2[l + p];
        }
    }
}

int or_field_metrics(const or_grid_desc *g, int axis,
                     double *coord, double *xi_x, double *xixx, double *xix2)
{
    int p = g->pad;
    if (g->dim == 2 && axis == 1) p = 1;
    int n = (g->dim == 2 && axis == 1) ? 1 : g->n[axis];
    double *xi = (double *)malloc(sizeof(double) * (size_t)(n + 2 * p));
    or_grid_desc gg = *g;
    if (g->dim == 2 && axis == 1) gg.n[1] = 1;
    level0_axis(&gg, axis, p, xi, coord, xi_x, xixx, xix2);
    free(xi);
    return n + 2 * p;
}

/* mgGridMetrics() grid.cc:521-592 then mgGridMetrics(dim) :603-666. */
int or_level_metrics(const or_grid_desc *g, int level, int axis,
                     double *coord, double *xi, double *xi_x, double *xixx, double *xix2)
{
    int n0 = g->n[axis];
    if (g->dim == 2 && axis == 1) n0 = 1;
    int depth = or_vc_depth(g);
    if (level < 0 || level > depth) return -1;
    double L = g->len[axis];
    int p = g->pad;
    if (g->dim == 2 && axis == 1) p = 1;

    /* level-0 global arrays with the field pad, then sliced to [-1, n]. */
    int n0p = n0 + 2 * p;
    double *gxi = malloc(sizeof(double) * n0p), *gx = malloc(sizeof(double) * n0p);
    double *gf = malloc(sizeof(double) * n0p), *gff = malloc(sizeof(double) * n0p), *gf2 = malloc(sizeof(double) * n0p);
    or_grid_desc gg = *g;
    if (g->dim == 2 && axis == 1) gg.n[1] = 1;
    /* uniform values first: mgGridMetrics() copies xiGlo and the uniform
     * xGlobal for level 0 before createTanHypGrid overwrites them */
    {
        or_grid_desc gu = gg; gu.gtype[axis] = 0;
        level0_axis(&gu, axis, p, gxi, gx, gf, gff, gf2);
    }

    /* walk down the levels computing xi and the uniform coordinates */
    int nl = n0;
    double *cxi = malloc(sizeof(double) * (nl + 2)), *cx = malloc(sizeof(double) * (nl + 2));
    for (int i = -1; i <= nl; i++) { cxi[i + 1] = gxi[i + p]; cx[i + 1] = gx[i + p]; }
    int planar_y = (g->dim == 2 && axis == 1);
    for (int l = 1; l <= level; l++) {
        int cl = planar_y ? nl : nl / 2;
        double *nxi = malloc(sizeof(double) * (cl + 2)), *nx = malloc(sizeof(double) * (cl + 2));
        if (planar_y) {
            /* PLANAR: globalSize(1)/2^l truncates to 0 cells; the y arrays
             * are never read by the 2-D solver.  Keep them well defined. */
            memcpy(nxi, cxi, sizeof(double) * (cl + 2));
            memcpy(nx, cx, sizeof(double) * (cl + 2));
        } else {
            for (int i = 0; i < cl; i++) {
                nxi[i + 1] = (cxi[2 * i + 1] + cxi[2 * i + 2]) / 2.0;
                nx[i + 1] = L * nxi[i + 1];
            }
            nxi[0] = -nxi[1];
            nx[0] = -nx[1];
            nxi[cl + 1] = 1.0 + nxi[1];
            nx[cl + 1] = L + nx[1];
        }
        free(cxi); free(cx);
        cxi = nxi; cx = nx; nl = cl;
    }

    for (int i = 0; i < nl + 2; i++) {
        xi[i] = cxi[i];
        coord[i] = cx[i];
        xi_x[i] = 1.0 / L;
        xixx[i] = 0.0;
        xix2[i] = pow(xi_x[i], 2.0);
    }

    if (!planar_y && g->gtype[axis] > 0) {
        int two = g->gtype[axis] - 1;
        if (level == 0) {
            /* createTanHypGrid writes globalMetrics(5 dim + {1..4}) = stretched [-1, n] */
            double *sx = malloc(sizeof(double) * n0p), *sf = malloc(sizeof(double) * n0p);
            double *sff = malloc(sizeof(double) * n0p), *sf2 = malloc(sizeof(double) * n0p);
            level0_axis(&gg, axis, p, gxi, sx, sf, sff, sf2);
            for (int i = -1; i <= n0; i++) {
                coord[i + 1] = sx[i + p]; xi_x[i + 1] = sf[i + p];
                xixx[i + 1] = sff[i + p]; xix2[i + 1] = sf2[i + p];
            }
            free(sx); free(sf); free(sff); free(sf2);
        } else {
            for (int i = 0; i < nl; i++)
                tanh_point(two, L, g->beta[axis], xi[i + 1], &coord[i + 1], &xi_x[i + 1], &xixx[i + 1], &xix2[i + 1]);
            coord[0] = -coord[1];
            xi_x[0] = xi_x[nl]; xixx[0] = xixx[nl]; xix2[0] = xix2[nl];
            coord[nl + 1] = L + coord[1];
            xi_x[nl + 1] = xi_x[1]; xixx[nl + 1] = xixx[1]; xix2[nl + 1] = xix2[1];
        }
    }
    free(cxi); free(cx);
    free(gxi); free(gx); free(gf); free(gff); free(gf2);
    return nl + 2;
}
