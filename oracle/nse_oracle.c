/*
 * nse_oracle.c -- restatement of the SARAS-CC Navier-Stokes/Boussinesq time
 * step for a single domain: lib/field (derivative.cc, vfield.cc, sfield.cc,
 * plainvf/plainsf algebra), lib/boundary (dirichlet.cc, neumann.cc),
 * lib/parallel/mpidata.cc face sync (periodic self-wrap), lib/timestep
 * (timestep.cc, predictors.cc, lsRK3_d3.cc, eulerCN_d3.cc), lib/force
 * (buoyantForce.cc, constantPGrad in force.h) and the lib/io/tseries.cc
 * diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY (see saras_oracle.h).
 *
 * Fields are full arrays with pad p: (nx+2p)(ny+2p)(nz+2p), z fastest.
 * Blitz++'s normalised stencils central12n/22n/14n/24n (derivative.cc:106-
 * 306) come from an un-vendored library (Blitz++, version unpinned,
 * CMakeLists.txt:83) that is absent here; they are restated from its
 * published stencil operator definitions:
 *   c12n = (F+1 - F-1) * (1/2)          c22n = (F-1 - 2 F) + F+1
 *   c14n = (8 (F+1 - F-1) - (F+2 - F-2)) * (1/12)
 *   c24n = ((-30 F + 16 (F+1 + F-1)) - (F+2 + F-2)) * (1/12)
 * Their rounding order against a real Blitz++ build is unpinned; the
 * formulas are pinned by polynomial exactness and convergence order in the
 * tests, and the second-order edge overrides are the reference's own
 * expressions (derivative.cc:111-112, 215-223).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "saras_oracle.h"

struct or_nse {
    or_nse_desc d;
    int n[3], p;
    int planar, py;      /* PLANAR build (dim 2: ny = 1, y index 0 only, no y pad) */
    long SX, SY;
    size_t full;
    double *x[3], *xi_x[3], *xixx[3], *xix2[3];   /* field metrics over [-p, n+p-1], +p offset */
    double ih[3], ih2[3], i2d[3];                  /* derivative.cc:52-54, vfield.cc i2dx */
    double tsi2h[3], tsih2[3];                     /* timestep::setCoefficients */
    double hwall[6];                               /* neumann.cc:59-75 */
    double *V[3], *P, *T, *Pp;
    double *nse[3], *trhs, *tv[3], *ts, *pg[3], *mgr, *iter, *tmpA, *tmpB;
    or_mg *mg;
    int max_iter;
    double time;
    int jacobi_iters[4];
    or_mg_stats mg_stats;
    double tstp;            /* the input time step ("Time-Step"), for the LES start */
    double *Tsg[6];         /* sub-grid stresses Txx, Tyy, Tzz, Txy, Tyz, Tzx */
    double sgs_diag[3];
};

#define FX(S, i, j, k) ((long)((i) + (S)->p) * (S)->SX + (long)((j) + (S)->py) * (S)->SY + (long)((k) + (S)->p))

static double *zalloc(size_t n) { return calloc(n, sizeof(double)); }

or_nse *or_nse_create(const or_nse_desc *d)
{
    or_nse *s = calloc(1, sizeof(or_nse));
    s->d = *d;
    s->p = d->g.pad;
    for (int a = 0; a < 3; a++) s->n[a] = d->g.n[a];
    int p = s->p;
    /* PLANAR (lsRK3_d2 / eulerCN_d2 with the PLANAR branches of the field
     * classes): the y axis is one cell, never differentiated and never given
     * boundary values, so the arrays carry no y pad */
    s->planar = d->g.dim == 2;
    if (s->planar) s->n[1] = 1;
    s->py = s->planar ? 0 : p;
    s->SY = s->n[2] + 2 * p;
    s->SX = (long)(s->n[1] + 2 * s->py) * s->SY;
    s->full = (size_t)(s->n[0] + 2 * p) * s->SX;
    for (int a = 0; a < 3; a++) {
        int m = s->n[a] + 2 * p;
        s->x[a] = zalloc(m); s->xi_x[a] = zalloc(m); s->xixx[a] = zalloc(m); s->xix2[a] = zalloc(m);
        or_field_metrics(&d->g, a, s->x[a], s->xi_x[a], s->xixx[a], s->xix2[a]);
        double dxi = 1.0 / (double)s->n[a];
        s->ih[a] = 1.0 / dxi;
        s->ih2[a] = pow(s->ih[a], 2.0);
        s->i2d[a] = 1.0 / (2.0 * dxi);
        double h2 = pow(dxi, 2.0);
        s->tsi2h[a] = 0.5 / dxi;
        s->tsih2[a] = 1.0 / h2;
        double L = d->g.len[a];
        s->hwall[2 * a] = 2 * s->x[a][0 + p];
        s->hwall[2 * a + 1] = 2 * (s->x[a][s->n[a] - 1 + p] - L);
    }
    for (int c = 0; c < 3; c++) {
        s->V[c] = zalloc(s->full); s->nse[c] = zalloc(s->full); s->tv[c] = zalloc(s->full); s->pg[c] = zalloc(s->full);
    }
    s->P = zalloc(s->full); s->T = zalloc(s->full); s->Pp = zalloc(s->full);
    s->trhs = zalloc(s->full); s->ts = zalloc(s->full); s->mgr = zalloc(s->full);
    s->iter = zalloc(s->full); s->tmpA = zalloc(s->full); s->tmpB = zalloc(s->full);
    or_mg_params mp = d->mg;
    mp.bc = OR_BC_NEUMANN;   /* allNeumann (poisson.cc:87) */
    /* a PLANAR plume (probType 9: T's top wall plumeOut) solves for Pp with a
     * Dirichlet top wall (poisson_d2.cc:641-647) */
    mp.top_dirichlet = s->planar && d->has_T && d->tbc_type[5] == OR_FBC_PLUME_OUT;
    s->mg = or_mg_create(&d->g, &mp);
    s->max_iter = (int)pow(log((double)s->n[0] * s->n[1] * s->n[2]), 3);  /* timestep.cc:115 */
    s->tstp = d->dt;
    for (int q = 0; q < 6; q++) s->Tsg[q] = zalloc(s->full);
    return s;
}

void or_nse_destroy(or_nse *s)
{
    if (!s) return;
    for (int a = 0; a < 3; a++) { free(s->x[a]); free(s->xi_x[a]); free(s->xixx[a]); free(s->xix2[a]); }
    for (int c = 0; c < 3; c++) { free(s->V[c]); free(s->nse[c]); free(s->tv[c]); free(s->pg[c]); }
    free(s->P); free(s->T); free(s->Pp); free(s->trhs); free(s->ts); free(s->mgr);
    free(s->iter); free(s->tmpA); free(s->tmpB);
    for (int q = 0; q < 6; q++) free(s->Tsg[q]);
    or_mg_destroy(s->mg);
    free(s);
}

int or_nse_max_iterations(const or_nse *s) { return s->max_iter; }

/* ---- boundary conditions ---------------------------------------------- */

/* mpidata::syncFaces on one rank: periodic pad-thick slabs over the full
 * extent of the other axes, in the order x, y, z (mpidata.cc:371-382). */
static void sync_faces(or_nse *s, double *F)
{
    int p = s->p, py = s->py, nx = s->n[0], ny = s->n[1], nz = s->n[2];
    if (s->d.g.periodic[0])
        for (int q = 0; q < p; q++)
            for (int j = -py; j < ny + py; j++)
                for (int k = -p; k < nz + p; k++) {
                    F[FX(s, -p + q, j, k)] = F[FX(s, nx - p + q, j, k)];
                    F[FX(s, nx + q, j, k)] = F[FX(s, q, j, k)];
                }
    if (s->d.g.periodic[1] && !s->planar)
        for (int i = -p; i < nx + p; i++)
            for (int q = 0; q < p; q++)
                for (int k = -p; k < nz + p; k++) {
                    F[FX(s, i, -p + q, k)] = F[FX(s, i, ny - p + q, k)];
                    F[FX(s, i, ny + q, k)] = F[FX(s, i, q, k)];
                }
    if (s->d.g.periodic[2])
        for (int i = -p; i < nx + p; i++)
            for (int j = -py; j < ny + py; j++)
                for (int q = 0; q < p; q++) {
                    F[FX(s, i, j, -p + q)] = F[FX(s, i, j, nz - p + q)];
                    F[FX(s, i, j, nz + q)] = F[FX(s, i, j, q)];
                }
}

/* one wall (first ghost layer over the full extent of the other axes):
 * dirichlet.cc:72-76, neumann.cc:89-93; nullBC::imposeBC (boundary.h:205-212)
 * is empty */
static void wall_bc(or_nse *s, double *F, int wall, int type, double val)
{
    if (type == OR_FBC_NULL) return;
    int p = s->p, n[3] = {s->n[0], s->n[1], s->n[2]};
    int ax = wall / 2, hi = wall & 1;
    int g = hi ? n[ax] : -1, src = hi ? n[ax] - 1 : 0;
    double kh = val * s->hwall[wall];
    int lo[3] = {-p, -s->py, -p}, up[3] = {n[0] + p - 1, n[1] + s->py - 1, n[2] + p - 1};
    lo[ax] = up[ax] = 0;
    for (int a = lo[0]; a <= up[0]; a++)
        for (int b = lo[1]; b <= up[1]; b++)
            for (int c = lo[2]; c <= up[2]; c++) {
                int gi[3] = {a, b, c}, si[3] = {a, b, c};
                gi[ax] = g; si[ax] = src;
                double inner = F[FX(s, si[0], si[1], si[2])];
                long gq = FX(s, gi[0], gi[1], gi[2]);
                if (type == OR_FBC_HOT_PLATE) {
                    /* hotPlate::createPatch (hotPlate.cc:74-111): disc about
                     * (Lx/2, Ly/2, 0), the distance taken to the ghost cell */
                    double cx = s->d.g.len[0] / 2.0, cy = s->d.g.len[1] / 2.0, cz = 0.0;
                    double r = sqrt(pow((s->x[0][gi[0] + p] - cx), 2) + pow((s->x[1][gi[1] + p] - cy), 2) +
                                    pow((s->x[2][gi[2] + p] - cz), 2));
                    double data = r <= s->d.plate_radius ? 1.0 : 0.0;
                    double mask = r <= s->d.plate_radius ? 0.0 : 1.0;
                    /* hotPlate::imposeBC (hotPlate.cc:114-123) */
                    double w = inner;
                    w = w * mask;
                    w = w + data * (2.0 * data - inner);
                    F[gq] = w;
                } else if (type == OR_FBC_PLUME_OUT) {
                    /* plumeOut::imposeBC (plumeOut.cc:71-90): the ghost Vz as it
                     * stands when T's walls are imposed */
                    F[gq] = s->V[2][gq] > 0 ? F[FX(s, gi[0], gi[1], gi[2] - 1)] : 0.0;
                } else {
                    F[gq] = type == OR_FBC_DIRICHLET ? 2.0 * val - inner : inner - kh;
                }
            }
}

/* sfield::imposeBCs / vfield::imposeV?BC: sync, then walls x (Lft, Rgt),
 * y (Frn, Bak), z (Top, Bot) -- vfield.cc:751-847, sfield.cc:520-537 */
static void impose(or_nse *s, double *F, const int *type, const double *val)
{
    sync_faces(s, F);
    for (int ax = 0; ax < 3; ax++) {
        if (s->d.g.periodic[ax] || (ax == 1 && s->planar)) continue;   /* PLANAR: no y walls */
        if (ax < 2) {
            wall_bc(s, F, 2 * ax, type[2 * ax], val[2 * ax]);
            wall_bc(s, F, 2 * ax + 1, type[2 * ax + 1], val[2 * ax + 1]);
        } else {
            wall_bc(s, F, 5, type[5], val[5]);
            wall_bc(s, F, 4, type[4], val[4]);
        }
    }
}

static void impose_v(or_nse *s, int c) { impose(s, s->V[c], s->d.vbc_type[c], s->d.vbc_val[c]); }
static void impose_t(or_nse *s) { impose(s, s->T, s->d.tbc_type, s->d.tbc_val); }
static void impose_p(or_nse *s) { impose(s, s->P, s->d.pbc_type, s->d.pbc_val); }

/* ---- derivatives (derivative.cc:104-306), core cells ------------------- */

static inline double at(const or_nse *s, const double *F, int ax, int i, int j, int k, int off)
{
    int c[3] = {i, j, k};
    c[ax] += off;
    return F[FX(s, c[0], c[1], c[2])];
}

static inline int idx_ax(int ax, int i, int j, int k) { return ax == 0 ? i : (ax == 1 ? j : k); }

static const double recip_12 = 1.0 / 12.0;

static inline double c12n(double fm, double fp) { return (fp - fm) * 0.5; }
static inline double c14n(double fm2, double fm, double fp, double fp2)
{
    return (8.0 * (fp - fm) - (fp2 - fm2)) * recip_12;
}
static inline double c22n(double fm, double f, double fp) { return (fm - 2.0 * f) + fp; }
static inline double c24n(double fm2, double fm, double f, double fp, double fp2)
{
    return ((-30.0 * f + 16.0 * (fp + fm)) - (fp2 + fm2)) * recip_12;
}

/* first derivative along ax at core cell (i,j,k) */
static double d1(const or_nse *s, const double *F, int ax, int i, int j, int k)
{
    int q = idx_ax(ax, i, j, k), n = s->n[ax];
    double fm = at(s, F, ax, i, j, k, -1), fp = at(s, F, ax, i, j, k, 1), v;
    if (s->d.dscheme == 1) v = c12n(fm, fp);
    else if (q == 0 || q == n - 1) v = 0.5 * (fp - fm);
    else v = c14n(at(s, F, ax, i, j, k, -2), fm, fp, at(s, F, ax, i, j, k, 2));
    v *= s->ih[ax];
    return s->xi_x[ax][q + s->p] * v;
}

/* second derivative along ax: xixx * D1 + xix2 * D2 */
static double d2(const or_nse *s, const double *F, int ax, int i, int j, int k)
{
    int q = idx_ax(ax, i, j, k), n = s->n[ax];
    double fm = at(s, F, ax, i, j, k, -1), f = F[FX(s, i, j, k)], fp = at(s, F, ax, i, j, k, 1), t, o;
    if (s->d.dscheme == 1) {
        t = c12n(fm, fp);
        o = c22n(fm, f, fp);
    } else if (q == 0 || q == n - 1) {
        t = 0.5 * (fp - fm);
        o = fp - 2.0 * f + fm;
    } else {
        double fm2 = at(s, F, ax, i, j, k, -2), fp2 = at(s, F, ax, i, j, k, 2);
        t = c14n(fm2, fm, fp, fp2);
        o = c24n(fm2, fm, f, fp, fp2);
    }
    t *= s->ih[ax];
    o *= s->ih2[ax];
    return s->xixx[ax][q + s->p] * t + s->xix2[ax][q + s->p] * o;
}

/* Morinishi difference of one product array: 4th order interior, 2nd order at
 * the first/last cell; z uses the reference's core.ubound(1) edge test
 * (vfield.cc:329,434, sfield.cc:236,295). */
static inline int edge_cell(const or_nse *s, int ax, int q)
{
    int last = ax == 2 ? s->n[1] - 1 : s->n[ax] - 1;
    return q == 0 || q == last;
}

static double mor_adv(const or_nse *s, double u, const double *F, int ax, int i, int j, int k)
{
    int q = idx_ax(ax, i, j, k);
    double xm = s->xi_x[ax][q + s->p];
    if (s->d.dscheme == 1 || edge_cell(s, ax, q))
        return u * xm * (at(s, F, ax, i, j, k, 1) - at(s, F, ax, i, j, k, -1)) * s->i2d[ax];
    return u * xm * (-at(s, F, ax, i, j, k, 2) + 8.0 * at(s, F, ax, i, j, k, 1) - 8.0 * at(s, F, ax, i, j, k, -1) +
                     at(s, F, ax, i, j, k, -2)) * s->i2d[ax] / 6.0;
}

static double mor_div(const or_nse *s, const double *G, int ax, int i, int j, int k)
{
    int q = idx_ax(ax, i, j, k);
    double xm = s->xi_x[ax][q + s->p];
    if (s->d.dscheme == 1 || edge_cell(s, ax, q))
        return xm * (at(s, G, ax, i, j, k, 1) - at(s, G, ax, i, j, k, -1)) * s->i2d[ax];
    return xm * (-at(s, G, ax, i, j, k, 2) + 8.0 * at(s, G, ax, i, j, k, 1) - 8.0 * at(s, G, ax, i, j, k, -1) +
                 at(s, G, ax, i, j, k, -2)) * s->i2d[ax] / 6.0;
}

#define CORE_LOOP(s)                                  \
    _Pragma("omp parallel for schedule(static)")      \
    for (int i = 0; i < (s)->n[0]; i++)               \
        for (int j = 0; j < (s)->n[1]; j++)           \
            for (int k = 0; k < (s)->n[2]; k++)

/* hybrid upwind term of advecting velocity u along axis ax for field F
 * (vfield.cc:475-689, sfield.cc:324-480): central 2nd/4th order where the
 * cell Peclet number |u| dh / diffCoeff is below "Peclet Limit", else biased
 * stencils weighted by "Central Bias"; dScheme 2 edge cells central 2nd. */
static double upwind(const or_nse *s, double u, const double *F, int ax, int i, int j, int k, double diff)
{
    int q = idx_ax(ax, i, j, k), n = s->n[ax];
    double xm = s->xi_x[ax][q + s->p], i2d = s->i2d[ax];
    double om = s->d.central_bias;
    if (s->d.dscheme == 1) {
        double a = 2.0 - om, b = 2.0 * (1.0 - om), c = -om;
        double dh = s->x[ax][q + 1 + s->p] - s->x[ax][q + s->p];
        double pe = fabs(u) * dh / diff;
        double fm = at(s, F, ax, i, j, k, -1), f = F[FX(s, i, j, k)], fp = at(s, F, ax, i, j, k, 1);
        if (pe < s->d.pe_limit) return u * xm * (fp - fm) * i2d;
        if (u > 0) return u * xm * (-a * fm + b * f - c * fp) * i2d;
        return u * xm * (a * fp - b * f + c * fm) * i2d;
    }
    double a = 1.0 - om, b = 4.0 - 3.0 * om, c = 3.0 * (1.0 - om), d = -om;
    double fm = at(s, F, ax, i, j, k, -1), fp = at(s, F, ax, i, j, k, 1);
    if (q == 0 || q == n - 1) return u * xm * (fp - fm) * i2d;
    double dh = s->x[ax][q + 1 + s->p] - s->x[ax][q + s->p];
    double pe = fabs(u) * dh / diff;
    double fm2 = at(s, F, ax, i, j, k, -2), f = F[FX(s, i, j, k)], fp2 = at(s, F, ax, i, j, k, 2);
    if (pe < s->d.pe_limit) return u * xm * (-fp2 + 8.0 * fp - 8.0 * fm + fm2) * i2d / 6.0;
    if (u > 0) return u * xm * (a * fm2 - b * fm + c * f - d * fp) * i2d;
    return u * xm * (-a * fp2 + b * fp - c * f + d * fm) * i2d;
}

/* vfield::computeDiff / sfield::computeDiff: H(core) += d2xx; += d2yy; += d2zz */
static void compute_diff(or_nse *s, const double *F, double *H)
{
    CORE_LOOP(s) {
        long c = FX(s, i, j, k);
        H[c] += d2(s, F, 0, i, j, k);
        if (!s->planar) H[c] += d2(s, F, 1, i, j, k);
        H[c] += d2(s, F, 2, i, j, k);
    }
}

/* vfield::computeNLin (nlScheme 1 central, 3 Morinishi), vfield.cc:205-462 */
static void nlin_v(or_nse *s, double *H[3])
{
    double **V = s->V;
    /* PLANAR (vfield.cc:211-240, 289-349): Vx and Vz only, x and z passes only;
     * the upwind loop has no PLANAR guard, but with Vy = 0 its y pass and its
     * Vy row (never used: solveVy is skipped) change nothing */
    const int pl = s->planar;
    if (s->d.nlscheme == 2) {
        CORE_LOOP(s) {
            long q = FX(s, i, j, k);
            for (int a = 0; a < 3; a++) {
                if (pl && a == 1) continue;
                double u = V[a][q];
                for (int c = 0; c < 3; c++)
                    if (!(pl && c == 1)) H[c][q] -= upwind(s, u, V[c], a, i, j, k, s->d.nu);
            }
        }
        return;
    }
    if (s->d.nlscheme == 1) {
        for (int c = 0; c < 3; c++) {
            if (pl && c == 1) continue;
            CORE_LOOP(s) {
                long q = FX(s, i, j, k);
                H[c][q] -= V[0][q] * d1(s, V[c], 0, i, j, k);
                if (!pl) H[c][q] -= V[1][q] * d1(s, V[c], 1, i, j, k);
                H[c][q] -= V[2][q] * d1(s, V[c], 2, i, j, k);
            }
        }
        return;
    }
    /* products over the full arrays: tempXX = Vx Vx, tempZX = Vz Vx, tempXY = Vx Vy,
     * tempYZ = Vy Vz, tempYY = Vy Vy, tempZZ = Vz Vz (vfield.cc:276-283) */
    double *pr[3][3];
    double *buf = malloc(sizeof(double) * s->full * 6);
    double *XX = buf, *ZZ = buf + s->full, *ZX = buf + 2 * s->full, *XY = buf + 3 * s->full,
           *YZ = buf + 4 * s->full, *YY = buf + 5 * s->full;
    for (size_t q = 0; q < s->full; q++) {
        XX[q] = V[0][q] * V[0][q];
        ZZ[q] = V[2][q] * V[2][q];
        ZX[q] = V[2][q] * V[0][q];
        XY[q] = V[0][q] * V[1][q];
        YZ[q] = V[1][q] * V[2][q];
        YY[q] = V[1][q] * V[1][q];
    }
    /* divergence product of (advecting axis a, component c) */
    pr[0][0] = XX; pr[0][1] = XY; pr[0][2] = ZX;
    pr[1][0] = XY; pr[1][1] = YY; pr[1][2] = YZ;
    pr[2][0] = ZX; pr[2][1] = YZ; pr[2][2] = ZZ;
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        for (int a = 0; a < 3; a++) {
            if (pl && a == 1) continue;
            double u = V[a][q];
            for (int c = 0; c < 3; c++) {
                if (pl && c == 1) continue;
                double adv = mor_adv(s, u, V[c], a, i, j, k);
                double dv = mor_div(s, pr[a][c], a, i, j, k);
                H[c][q] -= (adv + dv) / 2;
            }
        }
    }
    free(buf);
}

/* sfield::computeNLin, sfield.cc:160-311 */
static void nlin_s(or_nse *s, const double *F, double *H)
{
    double **V = s->V;
    const int pl = s->planar;   /* PLANAR: x and z terms only (sfield.cc:164-178, 209-260) */
    if (s->d.nlscheme == 2) {
        CORE_LOOP(s) {
            long q = FX(s, i, j, k);
            for (int a = 0; a < 3; a++)
                if (!(pl && a == 1)) H[q] -= upwind(s, V[a][q], F, a, i, j, k, s->d.kappa);
        }
        return;
    }
    if (s->d.nlscheme == 1) {
        CORE_LOOP(s) {
            long q = FX(s, i, j, k);
            H[q] -= V[0][q] * d1(s, F, 0, i, j, k);
            if (!pl) H[q] -= V[1][q] * d1(s, F, 1, i, j, k);
            H[q] -= V[2][q] * d1(s, F, 2, i, j, k);
        }
        return;
    }
    double *buf = malloc(sizeof(double) * s->full * 3);
    double *TX = buf, *TY = buf + s->full, *TZ = buf + 2 * s->full;
    for (size_t q = 0; q < s->full; q++) {
        TX[q] = V[0][q] * F[q];
        TY[q] = V[1][q] * F[q];
        TZ[q] = V[2][q] * F[q];
    }
    double *pr[3] = {TX, TY, TZ};
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        for (int a = 0; a < 3; a++) {
            if (pl && a == 1) continue;
            double adv = mor_adv(s, V[a][q], F, a, i, j, k);
            double dv = mor_div(s, pr[a], a, i, j, k);
            H[q] -= (adv + dv) / 2;
        }
    }
    free(buf);
}

/* ---- Jacobi Crank-Nicolson (predictors.cc:59-423) ---------------------- */
static int jacobi(or_nse *s, double *F, const double *rhs, double beta, double diff,
                  const int *type, const double *val)
{
    double dt = s->d.dt;
    int it = 0;
    for (;;) {
        if (s->planar) {
            /* predictors.cc:65-77 (PLANAR): no y terms */
            CORE_LOOP(s) {
                double E = F[FX(s, i + 1, j, k)], W = F[FX(s, i - 1, j, k)];
                double Tp = F[FX(s, i, j, k + 1)], B = F[FX(s, i, j, k - 1)];
                double x2 = s->xix2[0][i + s->p], xx = s->xixx[0][i + s->p];
                double z2 = s->xix2[2][k + s->p], zz = s->xixx[2][k + s->p];
                s->iter[FX(s, i, j, k)] =
                    ((s->tsih2[0] * x2 * (E + W) + s->tsi2h[0] * xx * (E - W) + s->tsih2[2] * z2 * (Tp + B) +
                      s->tsi2h[2] * zz * (Tp - B)) *
                         dt * diff * beta +
                     rhs[FX(s, i, j, k)]) /
                    (1.0 + 2.0 * dt * diff * beta * (s->tsih2[0] * x2 + s->tsih2[2] * z2));
            }
        } else {
            CORE_LOOP(s) {
                double E = F[FX(s, i + 1, j, k)], W = F[FX(s, i - 1, j, k)];
                double N = F[FX(s, i, j + 1, k)], Sv = F[FX(s, i, j - 1, k)];
                double Tp = F[FX(s, i, j, k + 1)], B = F[FX(s, i, j, k - 1)];
                double x2 = s->xix2[0][i + s->p], xx = s->xixx[0][i + s->p];
                double y2 = s->xix2[1][j + s->p], yy = s->xixx[1][j + s->p];
                double z2 = s->xix2[2][k + s->p], zz = s->xixx[2][k + s->p];
                s->iter[FX(s, i, j, k)] =
                    ((s->tsih2[0] * x2 * (E + W) + s->tsi2h[0] * xx * (E - W) + s->tsih2[1] * y2 * (N + Sv) +
                      s->tsi2h[1] * yy * (N - Sv) + s->tsih2[2] * z2 * (Tp + B) + s->tsi2h[2] * zz * (Tp - B)) *
                         dt * diff * beta +
                     rhs[FX(s, i, j, k)]) /
                    (1.0 + 2.0 * dt * diff * beta * (s->tsih2[0] * x2 + s->tsih2[1] * y2 + s->tsih2[2] * z2));
            }
        }
        /* the reference copies the whole of iterTemp (predictors.cc:96), whose
         * pads are never written; every pad cell a stencil reads is rebuilt by
         * the BCs that follow, so copying the core is equivalent */
        CORE_LOOP(s) F[FX(s, i, j, k)] = s->iter[FX(s, i, j, k)];
        impose(s, F, type, val);
        double mx = 0.0;
        for (int i = 0; i < s->n[0]; i++)
            for (int j = 0; j < s->n[1]; j++)
                for (int k = 0; k < s->n[2]; k++) {
                    double C = F[FX(s, i, j, k)];
                    double E = F[FX(s, i + 1, j, k)], W = F[FX(s, i - 1, j, k)];
                    double Tp = F[FX(s, i, j, k + 1)], B = F[FX(s, i, j, k - 1)];
                    double x2 = s->xix2[0][i + s->p], xx = s->xixx[0][i + s->p];
                    double z2 = s->xix2[2][k + s->p], zz = s->xixx[2][k + s->p];
                    double r;
                    if (s->planar) {
                        /* predictors.cc:100-109 (PLANAR) */
                        r = C - beta * dt * diff *
                                    (x2 * (E - 2.0 * C + W) * s->tsih2[0] + xx * (E - W) * s->tsi2h[0] +
                                     z2 * (Tp - 2.0 * C + B) * s->tsih2[2] + zz * (Tp - B) * s->tsi2h[2]);
                    } else {
                        double N = F[FX(s, i, j + 1, k)], Sv = F[FX(s, i, j - 1, k)];
                        double y2 = s->xix2[1][j + s->p], yy = s->xixx[1][j + s->p];
                        r = C - beta * dt * diff *
                                    (x2 * (E - 2.0 * C + W) * s->tsih2[0] + xx * (E - W) * s->tsi2h[0] +
                                     y2 * (N - 2.0 * C + Sv) * s->tsih2[1] + yy * (N - Sv) * s->tsi2h[1] +
                                     z2 * (Tp - 2.0 * C + B) * s->tsih2[2] + zz * (Tp - B) * s->tsi2h[2]);
                    }
                    double a = fabs(r - rhs[FX(s, i, j, k)]);
                    if (a > mx) mx = a;
                }
        if (mx < s->d.cn_tol) break;
        it += 1;
        if (it > s->max_iter) return -1;
    }
    return it;
}

/* ---- the RK3 step (lsRK3_d3.cc:190-305, with or without T) ------------- */
static const double alph[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
static const double gamm[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double zeta[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

static void fill(double *a, size_t n, double v) { for (size_t q = 0; q < n; q++) a[q] = v; }

/* plainvf::multAdd (plainvf.cc:89-95): full arrays */
static void mult_add(double *y, const double *x, double k, size_t n) { for (size_t q = 0; q < n; q++) y[q] += k * x[q]; }

/* ---- stretched spiral vortex LES (lib/les/spiral.cc) -------------------- */
#define SGS_EPS (2e-3)   /* spiral.cc:49 */
#define SGS_PI 3.14159265358979323846   /* M_PI */

typedef struct {
    double Sxx, Syy, Szz, Sxy, Syz, Szx;
    double e[3], K;
} sgs_cell;

/* spiral::keIntegral (spiral.cc:590-607) */
static double ke_integral(double k)
{
    double k2 = k * k;
    if (k2 < 2.42806) {
        double pade = (3.0 + 2.5107 * k2 + 0.330357 * k2 * k2 + 0.0295481 * k2 * k2 * k2) /
                      (1.0 + 0.336901 * k2 + 0.0416684 * k2 * k2 + 0.00187191 * k2 * k2 * k2);
        return 0.5 * (pade - 4.06235 * pow(k2, 1.0 / 3.0));
    }
    double pade = (1.26429 + 0.835714 * k2 + 0.0964286 * k2 * k2) /
                  (1.0 + 2.25 * k2 + 0.964286 * k2 * k2 + 0.0964286 * k2 * k2 * k2);
    return 0.5 * pade * exp(-k2);
}

/* spiral::sfIntegral (spiral.cc:620-631) */
static double sf_integral(double d)
{
    double d2 = d * d;
    if (d < 0.873469) return 7.4022 * d2 - 1.82642 * d2 * d2;
    return 12.2946 * pow(d, 2.0 / 3.0) - 6.0 - 0.573159 * pow(d, -1.5) * sin(3.14159 * d - 0.785398);
}

/* spiral::eigenvalueSymm (spiral.cc:643-~690): largest eigenvalue of S;
 * returns 0, or 3 where the reference aborts (q >= 0) */
static int eigval_symm(const sgs_cell *c, double *out)
{
    double ev[3];
    double a = -(c->Sxx + c->Syy + c->Szz);
    double b = c->Sxx * c->Syy - c->Sxy * c->Sxy + c->Syy * c->Szz - c->Syz * c->Syz + c->Szz * c->Sxx -
               c->Szx * c->Szx;
    double cc = -(c->Sxx * (c->Syy * c->Szz - c->Syz * c->Syz) + c->Sxy * (c->Syz * c->Szx - c->Sxy * c->Szz) +
                  c->Szx * (c->Sxy * c->Syz - c->Syy * c->Szx));
    double q = (3.0 * b - a * a) / 9.0;
    double r = (9.0 * a * b - 27.0 * cc - 2.0 * a * a * a) / 54.0;
    if (q >= 0.0) return 3;
    double costheta = r / sqrt(-q * q * q);
    double theta = costheta > 1.0 ? 0.0 : costheta < -1.0 ? SGS_PI : acos(costheta);
    ev[0] = 2.0 * sqrt(-q) * cos((theta) / 3.0) - a / 3.0;
    ev[1] = 2.0 * sqrt(-q) * cos((theta + 2.0 * SGS_PI) / 3.0) - a / 3.0;
    ev[2] = 2.0 * sqrt(-q) * cos((theta + 4.0 * SGS_PI) / 3.0) - a / 3.0;
    double t;
    if (ev[0] > ev[1]) { t = ev[0]; ev[0] = ev[1]; ev[1] = t; }
    if (ev[1] > ev[2]) { t = ev[1]; ev[1] = ev[2]; ev[2] = t; }
    if (ev[0] > ev[1]) { t = ev[0]; ev[0] = ev[1]; ev[1] = t; }
    *out = ev[2];
    return 0;
}

/* spiral::eigenvectorSymm (spiral.cc:702-~742) */
static int eigvec_symm(const sgs_cell *c, double ev, double v[3])
{
    double Sxx = c->Sxx, Syy = c->Syy, Szz = c->Szz, Sxy = c->Sxy, Syz = c->Syz, Szx = c->Szx;
    double fNorm = sqrt(Sxx * Sxx + Syy * Syy + Szz * Szz + Sxy * Sxy + Syz * Syz + Szx * Szx);
    if (fabs((Sxx - ev) * ((Syy - ev) * (Szz - ev) - Syz * Syz) + Sxy * (Syz * Szx - Sxy * (Szz - ev)) +
             Szx * (Sxy * Syz - (Syy - ev) * Szx)) / fabs(fNorm) > SGS_EPS)
        return 3;
    double det[3] = {(Syy - ev) * (Szz - ev) - Syz * Syz, (Szz - ev) * (Sxx - ev) - Szx * Szx,
                     (Sxx - ev) * (Syy - ev) - Sxy * Sxy};
    double f[3] = {fabs(det[0]), fabs(det[1]), fabs(det[2])};
    if (f[0] >= f[1] && f[0] >= f[2]) {
        v[0] = 1.0;
        v[1] = (-Sxy * (Szz - ev) + Szx * Syz) / det[0];
        v[2] = (-Szx * (Syy - ev) + Sxy * Syz) / det[0];
    } else if (f[1] >= f[2] && f[1] >= f[0]) {
        v[0] = (-Sxy * (Szz - ev) + Syz * Szx) / det[1];
        v[1] = 1.0;
        v[2] = (-Syz * (Sxx - ev) + Sxy * Szx) / det[1];
    } else if (f[2] >= f[0] && f[2] >= f[1]) {
        v[0] = (-Szx * (Syy - ev) + Syz * Sxy) / det[2];
        v[1] = (-Syz * (Sxx - ev) + Szx * Sxy) / det[2];
        v[2] = 1.0;
    } else {
        return 3;
    }
    return 0;
}

/* spiral::sgsStress (spiral.cc:472-548) at core cell (i,j,k) */
static int sgs_stress(const or_nse *s, int i, int j, int k, const double dudx[3][3], double del, sgs_cell *c,
                      double T[6])
{
    c->Sxx = 0.5 * (dudx[0][0] + dudx[0][0]);
    c->Syy = 0.5 * (dudx[1][1] + dudx[1][1]);
    c->Szz = 0.5 * (dudx[2][2] + dudx[2][2]);
    c->Sxy = 0.5 * (dudx[0][1] + dudx[1][0]);
    c->Syz = 0.5 * (dudx[1][2] + dudx[2][1]);
    c->Szx = 0.5 * (dudx[2][0] + dudx[0][2]);
    double ev;
    if (eigval_symm(c, &ev)) return 3;
    double *e = c->e;
    if (eigvec_symm(c, ev, e)) return 3;
    double length = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    e[0] /= length;
    e[1] /= length;
    e[2] /= length;
    double a = e[0] * e[0] * c->Sxx + e[0] * e[1] * c->Sxy + e[0] * e[2] * c->Szx + e[1] * e[0] * c->Sxy +
               e[1] * e[1] * c->Syy + e[1] * e[2] * c->Syz + e[2] * e[0] * c->Szx + e[2] * e[1] * c->Syz +
               e[2] * e[2] * c->Szz;
    double lv = sqrt(2.0 * s->d.nu / (3.0 * (fabs(a) + SGS_EPS)));
    /* structure function over the 26 neighbours */
    double F2 = 0.0, Qd = 0.0;
    int cnt = 0;
    int p = s->p;
    const double *U = s->V[0], *Vv = s->V[1], *W = s->V[2];
    long c0 = FX(s, i, j, k);
    for (int di = -1; di <= 1; di++)
        for (int dj = -1; dj <= 1; dj++)
            for (int dk = -1; dk <= 1; dk++) {
                if (!(di || dj || dk)) continue;
                long q = FX(s, i + di, j + dj, k + dk);
                double du = U[q] - U[c0], dv = Vv[q] - Vv[c0], dw = W[q] - W[c0];
                F2 += du * du + dv * dv + dw * dw;
                double dx = s->x[0][i + di + p] - s->x[0][i + p];
                double dy = s->x[1][j + dj + p] - s->x[1][j + p];
                double dz = s->x[2][k + dk + p] - s->x[2][k + p];
                double dx2 = dx * dx + dy * dy + dz * dz;
                double dxe = dx * e[0] + dy * e[1] + dz * e[2];
                double d = sqrt(dx2 - dxe * dxe) / del;
                Qd += sf_integral(d);
                cnt++;
            }
    F2 /= (double)cnt;
    Qd /= (double)cnt;
    double prefac = F2 / Qd;
    double kc = SGS_PI / del;
    c->K = prefac * ke_integral(kc * lv);
    T[0] = (1.0 - e[0] * e[0]) * c->K;
    T[1] = (1.0 - e[1] * e[1]) * c->K;
    T[2] = (1.0 - e[2] * e[2]) * c->K;
    T[3] = (-e[0] * e[1]) * c->K;
    T[4] = (-e[1] * e[2]) * c->K;
    T[5] = (-e[2] * e[0]) * c->K;
    return 0;
}

/* spiral::computeSG(nseRHS, V) (spiral.cc:129-236); the reference leaves
 * two of its diagnostic sums uninitialised, here every sum starts at 0 */
static int compute_sg(or_nse *s, double *H[3])
{
    int p = s->p;
    double dxi[3];
    for (int a = 0; a < 3; a++) dxi[a] = 1.0 / (double)s->n[a];
    double sgke = 0.0, disp = 0.0, nusg = 0.0;
    for (int i = 0; i < s->n[0]; i++) {
        double dx = s->x[0][i + p] - s->x[0][i - 1 + p];
        for (int j = 0; j < s->n[1]; j++) {
            double dy = s->x[1][j + p] - s->x[1][j - 1 + p];
            for (int k = 0; k < s->n[2]; k++) {
                double dz = s->x[2][k + p] - s->x[2][k - 1 + p];
                double del = pow(dx * dy * dz, 1.0 / 3.0);
                double dudx[3][3];
                for (int r = 0; r < 3; r++)
                    for (int a = 0; a < 3; a++) dudx[r][a] = d1(s, s->V[r], a, i, j, k);
                sgs_cell c;
                double T[6];
                if (sgs_stress(s, i, j, k, dudx, del, &c, T)) return 3;
                long q = FX(s, i, j, k);
                for (int m = 0; m < 6; m++) s->Tsg[m][q] = T[m];
                double A = dxi[0] / s->xi_x[0][i + p], B = dxi[1] / s->xi_x[1][j + p], C = dxi[2] / s->xi_x[2][k + p];
                sgke += fabs(c.K) * A * B * C;
                double sgDiss = -(c.Sxx * T[0] + c.Syy * T[1] + c.Szz * T[2] +
                                  2.0 * (c.Sxy * T[3] + c.Syz * T[4] + c.Szx * T[5]));
                disp += sgDiss * A * B * C;
                double sNorm = sqrt(c.Sxx * c.Sxx + c.Syy * c.Syy + c.Szz * c.Szz +
                                    2.0 * (c.Sxy * c.Sxy + c.Syz * c.Syz + c.Szx * c.Szx));
                double tNorm = sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2] +
                                    2.0 * (T[3] * T[3] + T[4] * T[4] + T[5] * T[5]));
                double nuTurb = (tNorm / sNorm) / 2;
                nusg += nuTurb * A * B * C;
            }
        }
    }
    s->sgs_diag[0] = sgke;
    s->sgs_diag[1] = disp;
    s->sgs_diag[2] = nusg;
    for (int m = 0; m < 6; m++) sync_faces(s, s->Tsg[m]);
    double **T = s->Tsg;   /* Txx, Tyy, Tzz, Txy, Tyz, Tzx */
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        double A11 = d1(s, T[0], 0, i, j, k), A12 = d1(s, T[3], 0, i, j, k), A13 = d1(s, T[5], 0, i, j, k);
        double A21 = d1(s, T[3], 1, i, j, k), A22 = d1(s, T[1], 1, i, j, k), A23 = d1(s, T[4], 1, i, j, k);
        double A31 = d1(s, T[5], 2, i, j, k), A32 = d1(s, T[4], 2, i, j, k), A33 = d1(s, T[2], 2, i, j, k);
        H[0][q] = H[0][q] - (A11 + A21 + A31);
        H[1][q] = H[1][q] - (A12 + A22 + A32);
        H[2][q] = H[2][q] - (A13 + A23 + A33);
    }
    return 0;
}

int or_nse_sgs(or_nse *s, double *hx, double *hy, double *hz)
{
    double *H[3] = {hx, hy, hz};
    return compute_sg(s, H);
}

void or_nse_sgs_diag(const or_nse *s, double out[3])
{
    for (int q = 0; q < 3; q++) out[q] = s->sgs_diag[q];
}

double or_nse_time(const or_nse *s) { return s->time; }

void or_nse_get_sgs(const or_nse *s, double *T6)
{
    for (int m = 0; m < 6; m++) memcpy(T6 + (size_t)m * s->full, s->Tsg[m], s->full * sizeof(double));
}

static int les_active(const or_nse *s) { return s->d.les && s->time > 5 * s->tstp; }

/* eulerCN_d3::timeAdvance explicit part (eulerCN_d3.cc:92-121, 181-235):
 * RHS = (nu/2 lap u - H + f - grad P) dt + u, accumulated in the reference's
 * order into one array */
static int euler_rhs(or_nse *s)
{
    size_t n = s->full;
    int hasT = s->d.has_T;
    for (int c = 0; c < 3; c++) {
        fill(s->nse[c], n, 0.0);
        compute_diff(s, s->V[c], s->nse[c]);
        double k = s->d.nu * 0.5;
        for (size_t q = 0; q < n; q++) s->nse[c][q] *= k;
    }
    if (hasT) {
        fill(s->trhs, n, 0.0);
        compute_diff(s, s->T, s->trhs);
        double k = s->d.kappa * 0.5;
        for (size_t q = 0; q < n; q++) s->trhs[q] *= k;
    }
    nlin_v(s, s->nse);
    if (hasT) nlin_s(s, s->T, s->trhs);
    if (hasT)
        for (int c = 0; c < 3; c++)
            for (size_t q = 0; q < n; q++) s->nse[c][q] -= s->d.buoy[c] * s->T[q];
    /* constantPGrad::addForcing: Hv.Vx(core) += meanPGrad (force.h:152) */
    if (s->d.mean_pgrad != 0.0) {
        CORE_LOOP(s) s->nse[0][FX(s, i, j, k)] += s->d.mean_pgrad;
    }
    if (les_active(s) && compute_sg(s, s->nse)) return 3;
    for (int c = 0; c < 3; c++) fill(s->pg[c], n, 0.0);
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        s->pg[0][q] = d1(s, s->P, 0, i, j, k);
        if (!s->planar) s->pg[1][q] = d1(s, s->P, 1, i, j, k);
        s->pg[2][q] = d1(s, s->P, 2, i, j, k);
    }
    for (int c = 0; c < 3; c++) {
        for (size_t q = 0; q < n; q++) s->nse[c][q] -= s->pg[c][q];
        for (size_t q = 0; q < n; q++) s->nse[c][q] *= s->d.dt;
        for (size_t q = 0; q < n; q++) s->nse[c][q] += s->V[c][q];
        sync_faces(s, s->nse[c]);
    }
    if (hasT) {
        for (size_t q = 0; q < n; q++) s->trhs[q] *= s->d.dt;
        for (size_t q = 0; q < n; q++) s->trhs[q] += s->T[q];
        sync_faces(s, s->trhs);
    }
    return 0;
}

int or_nse_rhs(or_nse *s, int stage)
{
    if (s->d.scheme == 1) return euler_rhs(s);
    size_t n = s->full;
    double a = alph[stage], b = alph[stage];
    int hasT = s->d.has_T;
    for (int c = 0; c < 3; c++) {
        fill(s->nse[c], n, 0.0);
        if (stage > 0) mult_add(s->nse[c], s->tv[c], zeta[stage], n);
    }
    if (hasT) {
        fill(s->trhs, n, 0.0);
        if (stage > 0) mult_add(s->trhs, s->ts, zeta[stage], n);
    }
    /* diffusion */
    for (int c = 0; c < 3; c++) { fill(s->tv[c], n, 0.0); compute_diff(s, s->V[c], s->tv[c]); }
    if (hasT) { fill(s->ts, n, 0.0); compute_diff(s, s->T, s->ts); }
    for (int c = 0; c < 3; c++) mult_add(s->nse[c], s->tv[c], a * s->d.nu, n);
    if (hasT) mult_add(s->trhs, s->ts, a * s->d.kappa, n);
    /* forcing and pressure gradient */
    for (int c = 0; c < 3; c++) fill(s->tv[c], n, 0.0);
    if (hasT) {
        fill(s->ts, n, 0.0);
        for (int c = 0; c < 3; c++)
            for (size_t q = 0; q < n; q++) s->tv[c][q] -= s->d.buoy[c] * s->T[q];
    }
    if (s->d.mean_pgrad != 0.0) {
        CORE_LOOP(s) s->tv[0][FX(s, i, j, k)] += s->d.mean_pgrad;
    }
    for (int c = 0; c < 3; c++) fill(s->pg[c], n, 0.0);
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        s->pg[0][q] = d1(s, s->P, 0, i, j, k);
        if (!s->planar) s->pg[1][q] = d1(s, s->P, 1, i, j, k);
        s->pg[2][q] = d1(s, s->P, 2, i, j, k);
    }
    for (int c = 0; c < 3; c++)
        for (size_t q = 0; q < n; q++) s->tv[c][q] -= s->pg[c][q];
    for (int c = 0; c < 3; c++) mult_add(s->nse[c], s->tv[c], a + b, n);
    if (hasT) mult_add(s->trhs, s->ts, a + b, n);
    /* non-linear terms */
    for (int c = 0; c < 3; c++) fill(s->tv[c], n, 0.0);
    nlin_v(s, s->tv);
    if (hasT) { fill(s->ts, n, 0.0); nlin_s(s, s->T, s->ts); }
    if (les_active(s) && compute_sg(s, s->tv)) return 3;
    for (int c = 0; c < 3; c++) mult_add(s->nse[c], s->tv[c], gamm[stage], n);
    if (hasT) mult_add(s->trhs, s->ts, gamm[stage], n);
    for (int c = 0; c < 3; c++) {
        for (size_t q = 0; q < n; q++) s->nse[c][q] *= s->d.dt;
        for (size_t q = 0; q < n; q++) s->nse[c][q] += s->V[c][q];
        sync_faces(s, s->nse[c]);
    }
    if (hasT) {
        for (size_t q = 0; q < n; q++) s->trhs[q] *= s->d.dt;
        for (size_t q = 0; q < n; q++) s->trhs[q] += s->T[q];
        sync_faces(s, s->trhs);
    }
    return 0;
}

int or_nse_stage(or_nse *s, int stage)
{
    size_t n = s->full;
    int euler = s->d.scheme == 1;
    double a = euler ? 0.0 : alph[stage], b = euler ? 0.5 : alph[stage];   /* betaCN2 = 1/2 */
    int rr = or_nse_rhs(s, stage);
    if (rr) return rr;
    for (int c = 0; c < 3; c++) {
        if (s->planar && c == 1) {   /* lsRK3_d2 / eulerCN_d2: no solveVy */
            s->jacobi_iters[1] = 0;
            continue;
        }
        s->jacobi_iters[c] = jacobi(s, s->V[c], s->nse[c], b, s->d.nu, s->d.vbc_type[c], s->d.vbc_val[c]);
        if (s->jacobi_iters[c] < 0) return 2;
    }
    if (s->d.has_T) {
        s->jacobi_iters[3] = jacobi(s, s->T, s->trhs, b, s->d.kappa, s->d.tbc_type, s->d.tbc_val);
        if (s->jacobi_iters[3] < 0) return 2;
    }
    /* projection */
    fill(s->mgr, n, 0.0);
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        s->mgr[q] += d1(s, s->V[0], 0, i, j, k);
        if (!s->planar) s->mgr[q] += d1(s, s->V[1], 1, i, j, k);
        s->mgr[q] += d1(s, s->V[2], 2, i, j, k);
    }
    double sc = euler ? 1.0 / s->d.dt : 1.0 / ((a + b) * s->d.dt);
    for (size_t q = 0; q < n; q++) s->mgr[q] *= sc;
    or_mg_solve(s->mg, s->Pp, s->mgr, s->p, NULL, &s->mg_stats);
    sync_faces(s, s->Pp);
    for (size_t q = 0; q < n; q++) s->P[q] += s->Pp[q];
    for (int c = 0; c < 3; c++) fill(s->pg[c], n, 0.0);   /* holds the zeroed pads of the stage */
    CORE_LOOP(s) {
        long q = FX(s, i, j, k);
        s->pg[0][q] = d1(s, s->Pp, 0, i, j, k);
        if (!s->planar) s->pg[1][q] = d1(s, s->Pp, 1, i, j, k);
        s->pg[2][q] = d1(s, s->Pp, 2, i, j, k);
    }
    double pc = euler ? s->d.dt : (a + b) * s->d.dt;
    for (int c = 0; c < 3; c++) {
        for (size_t q = 0; q < n; q++) s->pg[c][q] *= pc;
        for (size_t q = 0; q < n; q++) s->V[c][q] -= s->pg[c][q];
    }
    for (int c = 0; c < 3; c++) impose_v(s, c);
    impose_p(s);
    if (s->d.has_T) impose_t(s);
    return 0;
}

int or_nse_step(or_nse *s)
{
    for (int stage = 0; stage < (s->d.scheme == 1 ? 1 : 3); stage++) {
        int r = or_nse_stage(s, stage);
        if (r) return r;
    }
    s->time += s->d.dt;
    return 0;
}

void or_nse_jacobi_iters(const or_nse *s, int out[4]) { for (int c = 0; c < 4; c++) out[c] = s->jacobi_iters[c]; }

void or_nse_mg_stats(const or_nse *s, or_mg_stats *st) { *st = s->mg_stats; }

/* fields in/out (full arrays); set applies the initial BCs as the solver
 * constructors do */
void or_nse_set(or_nse *s, const double *vx, const double *vy, const double *vz, const double *p,
                const double *t, const double *pp)
{
    size_t b = s->full * sizeof(double);
    if (vx) memcpy(s->V[0], vx, b);
    if (vy) memcpy(s->V[1], vy, b);
    if (vz) memcpy(s->V[2], vz, b);
    if (p) memcpy(s->P, p, b);
    if (t) memcpy(s->T, t, b);
    if (pp) memcpy(s->Pp, pp, b);
    for (int c = 0; c < 3; c++) impose_v(s, c);
    impose_p(s);
    if (s->d.has_T) impose_t(s);
}

void or_nse_get(const or_nse *s, double *vx, double *vy, double *vz, double *p, double *t, double *pp)
{
    size_t b = s->full * sizeof(double);
    if (vx) memcpy(vx, s->V[0], b);
    if (vy) memcpy(vy, s->V[1], b);
    if (vz) memcpy(vz, s->V[2], b);
    if (p) memcpy(p, s->P, b);
    if (t) memcpy(t, s->T, b);
    if (pp) memcpy(pp, s->Pp, b);
}

void or_nse_get_rhs(const or_nse *s, double *rx, double *ry, double *rz, double *rt, double *hx, double *hy,
                    double *hz, double *ht)
{
    size_t b = s->full * sizeof(double);
    if (rx) memcpy(rx, s->nse[0], b);
    if (ry) memcpy(ry, s->nse[1], b);
    if (rz) memcpy(rz, s->nse[2], b);
    if (rt) memcpy(rt, s->trhs, b);
    if (hx) memcpy(hx, s->tv[0], b);
    if (hy) memcpy(hy, s->tv[1], b);
    if (hz) memcpy(hz, s->tv[2], b);
    if (ht) memcpy(ht, s->ts, b);
}

/* single operators on caller arrays, for stencil tests */
void or_nse_op(or_nse *s, int op, const double *F, double *out)
{
    size_t n = s->full;
    memset(out, 0, n * sizeof(double));
    if (s->planar) return;   /* 3-D operators (as the library's) */
    switch (op) {
        case 0: case 1: case 2:       /* first derivative along op */
            CORE_LOOP(s) out[FX(s, i, j, k)] = d1(s, F, op, i, j, k);
            break;
        case 3: case 4: case 5:       /* second derivative along op-3 */
            CORE_LOOP(s) out[FX(s, i, j, k)] = d2(s, F, op - 3, i, j, k);
            break;
        case 6:                       /* sfield::computeDiff */
            compute_diff(s, F, out);
            break;
        case 7:                       /* sfield::computeNLin with the current V */
            nlin_s(s, F, out);
            break;
    }
}

/* tseries.cc:410-446 (z gravity): volume-mean thermal energy
 * 1/2 (T + z - 1)^2 of the deviation from the conduction profile */
/* The vector operators of vfield/sfield on caller arrays:
 * 0 vfield::divergence (vfield.cc:721-737), 1 sfield::gradient
 * (sfield.cc:496-508), 2 vfield::computeDiff (vfield.cc:150-192),
 * 3 vfield::computeNLin of (fx, fy, fz) advecting itself (vfield.cc:205-689).
 * Ops 0 and 1 assign the core of their outputs; 2 and 3 accumulate. */
void or_nse_vec_op(or_nse *s, int op, const double *fx, const double *fy, const double *fz, double *ox, double *oy,
                   double *oz)
{
    if (s->planar) return;   /* 3-D operators (as the library's) */
    switch (op) {
        case 0:
            CORE_LOOP(s) {
                long q = FX(s, i, j, k);
                double d = 0.0;             /* divV = 0.0; divV += ... */
                d += d1(s, fx, 0, i, j, k);
                d += d1(s, fy, 1, i, j, k);
                d += d1(s, fz, 2, i, j, k);
                ox[q] = d;
            }
            break;
        case 1:
            CORE_LOOP(s) {
                long q = FX(s, i, j, k);
                ox[q] = d1(s, fx, 0, i, j, k);
                oy[q] = d1(s, fx, 1, i, j, k);
                oz[q] = d1(s, fx, 2, i, j, k);
            }
            break;
        case 2:
            compute_diff(s, fx, ox);
            compute_diff(s, fy, oy);
            compute_diff(s, fz, oz);
            break;
        default: {
            /* the field advects itself: run nlin_v on it in place of V */
            double *keep[3] = {s->V[0], s->V[1], s->V[2]};
            double *H[3] = {ox, oy, oz};
            s->V[0] = (double *)fx;
            s->V[1] = (double *)fy;
            s->V[2] = (double *)fz;
            nlin_v(s, H);
            s->V[0] = keep[0];
            s->V[1] = keep[1];
            s->V[2] = keep[2];
            break;
        }
    }
}

double or_nse_thermal_energy(or_nse *s)
{
    int p = s->p;
    double dxi[3];
    for (int a = 0; a < 3; a++) dxi[a] = 1.0 / (double)s->n[a];
    double te = 0.0;
    if (!s->d.has_T) return 0.0;
    for (int i = 0; i < s->n[0]; i++)
        for (int j = 0; j < s->n[1]; j++)
            for (int k = 0; k < s->n[2]; k++) {
                long q = FX(s, i, j, k);
                double dVol = s->planar ? (dxi[0] / s->xi_x[0][i + p]) * (dxi[2] / s->xi_x[2][k + p])
                                        : (dxi[0] / s->xi_x[0][i + p]) * (dxi[1] / s->xi_x[1][j + p]) *
                                              (dxi[2] / s->xi_x[2][k + p]);
                double theta = s->T[q] + s->x[2][k + p] - 1.0;
                te += pow(theta, 2.0) * 0.5 * dVol;
            }
    /* tseries.cc:79-83: totalVol = xLen zLen in PLANAR */
    return te / (s->planar ? s->d.g.len[0] * s->d.g.len[2] : s->d.g.len[0] * s->d.g.len[1] * s->d.g.len[2]);
}

/* tseries.cc:209-290: KE, max |div|, NuVol */
void or_nse_diag(or_nse *s, double *ke, double *divmax, double *nuvol, double *nuwall)
{
    int p = s->p;
    double dxi[3];
    for (int a = 0; a < 3; a++) dxi[a] = 1.0 / (double)s->n[a];
    double tke = 0.0, uzt = 0.0, dm = 0.0;
    for (int i = 0; i < s->n[0]; i++)
        for (int j = 0; j < s->n[1]; j++)
            for (int k = 0; k < s->n[2]; k++) {
                long q = FX(s, i, j, k);
                double dVol, dv = 0.0;
                if (s->planar) {
                    /* tseries.cc:227-233 (PLANAR) */
                    dVol = (dxi[0] / s->xi_x[0][i + p]) * (dxi[2] / s->xi_x[2][k + p]);
                    tke += (pow(s->V[0][q], 2.0) + pow(s->V[2][q], 2.0)) * 0.5 * dVol;
                    dv += d1(s, s->V[0], 0, i, j, k);
                    dv += d1(s, s->V[2], 2, i, j, k);
                } else {
                    dVol = (dxi[0] / s->xi_x[0][i + p]) * (dxi[1] / s->xi_x[1][j + p]) * (dxi[2] / s->xi_x[2][k + p]);
                    tke += (pow(s->V[0][q], 2.0) + pow(s->V[1][q], 2.0) + pow(s->V[2][q], 2.0)) * 0.5 * dVol;
                    dv += d1(s, s->V[0], 0, i, j, k);
                    dv += d1(s, s->V[1], 1, i, j, k);
                    dv += d1(s, s->V[2], 2, i, j, k);
                }
                if (s->d.has_T) uzt += s->V[2][q] * s->T[q] * dVol;
                if (fabs(dv) > dm) dm = fabs(dv);
            }
    double vol = s->planar ? s->d.g.len[0] * s->d.g.len[2] : s->d.g.len[0] * s->d.g.len[1] * s->d.g.len[2];
    if (ke) *ke = tke / vol;
    if (divmax) *divmax = dm;
    if (nuvol) *nuvol = s->d.has_T ? 1.0 + (uzt / vol) / s->d.kappa : 0.0;
    if (nuwall && s->d.has_T) {
        /* z-gravity, one rank holding both walls: wallDn ends as the top-wall
         * distance and serves both sums (tseries.cc:118-128, 374-394); an open
         * top (plumeOut, probType 9) leaves the bottom wall alone, at its own
         * distance, over one wall's area */
        int open_top = s->d.tbc_type[5] == OR_FBC_PLUME_OUT;
        double dt_sum = 0.0;
        double wall_dn = open_top ? s->x[2][0 + p] : s->d.g.len[2] - s->x[2][s->n[2] - 1 + p];
        for (int i = 0; i < s->n[0]; i++)
            for (int j = 0; j < s->n[1]; j++) {
                double dA = s->planar ? dxi[0] / s->xi_x[0][i + p]
                                      : (dxi[0] / s->xi_x[0][i + p]) * (dxi[1] / s->xi_x[1][j + p]);
                double dTdn = (1.0 - fabs(s->T[FX(s, i, j, 0)])) / wall_dn;
                dt_sum += dTdn * dA;
            }
        if (!open_top)
            for (int i = 0; i < s->n[0]; i++)
                for (int j = 0; j < s->n[1]; j++) {
                    double dA = s->planar ? dxi[0] / s->xi_x[0][i + p]
                                          : (dxi[0] / s->xi_x[0][i + p]) * (dxi[1] / s->xi_x[1][j + p]);
                    double dTdn = fabs(s->T[FX(s, i, j, s->n[2] - 1)]) / wall_dn;
                    dt_sum += dTdn * dA;
                }
        /* csArea = xLen in PLANAR (tseries.cc:110-114) */
        double cs = s->planar ? s->d.g.len[0] : s->d.g.len[0] * s->d.g.len[1];
        *nuwall = dt_sum / (cs * (open_top ? 1 : 2));
    }
}

double or_nse_cfl(or_nse *s, double courant, double dt_max)
{
    double avg[3];
    for (int a = 0; a < 3; a++) avg[a] = s->d.g.len[a] / s->n[a];
    double mx = 0.0;
    int first = 1;
    for (int i = 0; i < s->n[0]; i++)
        for (int j = 0; j < s->n[1]; j++)
            for (int k = 0; k < s->n[2]; k++) {
                long q = FX(s, i, j, k);
                /* vfield.cc:694-696 (PLANAR: x and z only) */
                double v = s->planar ? fabs(s->V[0][q]) / avg[0] + fabs(s->V[2][q]) / avg[2]
                                     : (fabs(s->V[0][q]) / avg[0] + fabs(s->V[1][q]) / avg[1]) + fabs(s->V[2][q]) / avg[2];
                if (first || v > mx) mx = v;
                first = 0;
            }
    double dt = courant / mx;
    if (dt > dt_max) dt = dt_max;
    s->d.dt = dt;
    return dt;
}

/* timestep::timestep (timestep.cc:70-98) and buoyantForce::buoyantForce
 * (buoyantForce.cc:48-60), as a table by rbcType: returns 0, or -1 where the
 * reference exits (rbcType outside 1..4) or would divide by zero. */
int or_nse_physics(int prob_type, double re, double ra, double pr, int rbc_type, const double g[3], double *nu,
                   double *kappa, double buoy[3])
{
    buoy[0] = buoy[1] = buoy[2] = 0.0;
    *kappa = 0.0;
    if (prob_type >= 1 && prob_type <= 4) {
        if (re <= 0.0) return -1;
        *nu = 1.0 / re;
        return 0;
    }
    if (prob_type < 5 || prob_type > 9 || rbc_type < 1 || rbc_type > 4) return -1;
    /* rows: rbcType 1..4; columns: nu, kappa, Fb */
    const double tab[4][3] = {{pr, 1.0, ra * pr},
                              {sqrt(pr / ra), 1.0 / sqrt(pr * ra), 1.0},
                              {1.0, 1.0 / pr, ra},
                              {sqrt(pr / ra), 1.0 / sqrt(pr * ra), pr}};
    *nu = tab[rbc_type - 1][0];
    *kappa = tab[rbc_type - 1][1];
    double vn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(vn > 0.0)) return -1;
    for (int a = 0; a < 3; a++) buoy[a] = tab[rbc_type - 1][2] * g[a] / vn;
    return 0;
}
