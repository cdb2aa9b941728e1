/*
 * mg_oracle.c -- restatement of lib/poisson/poisson.cc, poisson_d3.cc and
 * poisson_d2.cc for a single domain (one MPI rank: vcdLoc == vcdGlo,
 * locSolve always true, MPI face exchanges reduce to periodic self-wraps).
 *
 * TEST INFRASTRUCTURE ONLY (see saras_oracle.h).
 *
 * Expressions keep the reference's operand order so that results are
 * reproducible bit-for-bit; the file is compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "saras_oracle.h"

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
static int nthr(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

typedef struct {
    int n[3];            /* cells; n[1] == 1 in 2-D */
    long SX, SY;         /* strides of i and j in the padded array */
    double *lhs, *rhs, *tmp, *smd;
    double *x[3], *xixx[3], *xix2[3];   /* offset +1: index -1 at [0] */
    double ihx2[3], i2hx[3];
} level_t;

struct or_mg {
    or_grid_desc g;
    or_mg_params p;
    int dim, depth, max_count;
    level_t *lv;
    double *xWall, *yWall, *zWall;      /* DIRICHLET_TEST wall values, level 0 */
    int vLevel, zeroBC;
    int last_coarse_iters, coarse_capped;
};

#define IX(L, i, j, k) ((long)((i) + 1) * (L)->SX + (long)((j) + 1) * (L)->SY + (long)((k) + 1))

static size_t level_size(const level_t *L, int dim)
{
    return dim == 3 ? (size_t)(L->n[0] + 2) * (L->n[1] + 2) * (L->n[2] + 2)
                    : (size_t)(L->n[0] + 2) * (L->n[2] + 2);
}

or_mg *or_mg_create(const or_grid_desc *g, const or_mg_params *p)
{
    or_mg *m = calloc(1, sizeof(or_mg));
    m->g = *g;
    m->p = *p;
    m->dim = g->dim;
    m->depth = or_vc_depth(g);
    if (m->depth < 0) { free(m); return NULL; }
    m->lv = calloc((size_t)m->depth + 1, sizeof(level_t));
    double dxi[3];
    for (int a = 0; a < 3; a++) dxi[a] = 1.0 / (double)g->n[a];
    for (int l = 0; l <= m->depth; l++) {
        level_t *L = &m->lv[l];
        for (int a = 0; a < 3; a++) L->n[a] = g->n[a] >> l;
        if (m->dim == 2) L->n[1] = 1;
        if (m->dim == 3) { L->SY = L->n[2] + 2; L->SX = (long)(L->n[1] + 2) * (L->n[2] + 2); }
        else { L->SY = 0; L->SX = L->n[2] + 2; }
        size_t sz = level_size(L, m->dim);
        L->lhs = calloc(sz, sizeof(double)); L->rhs = calloc(sz, sizeof(double));
        L->tmp = calloc(sz, sizeof(double)); L->smd = calloc(sz, sizeof(double));
        for (int a = 0; a < 3; a++) {
            if (m->dim == 2 && a == 1) continue;
            int nl = L->n[a] + 2;
            L->x[a] = malloc(sizeof(double) * nl);
            L->xixx[a] = malloc(sizeof(double) * nl);
            L->xix2[a] = malloc(sizeof(double) * nl);
            double *xi = malloc(sizeof(double) * nl), *xx = malloc(sizeof(double) * nl);
            or_level_metrics(g, l, a, L->x[a], xi, xx, L->xixx[a], L->xix2[a]);
            free(xi); free(xx);
            /* poisson::setCoefficients, poisson.cc:453-483 */
            int hInc = 1 << l;
            double h = hInc * dxi[a];
            double h2 = pow(h, 2);
            L->i2hx[a] = 0.5 / h;
            L->ihx2[a] = 1.0 / h2;
        }
    }
    /* poisson.cc:431-436 */
    {
        const level_t *C = &m->lv[m->depth];
        m->max_count = m->dim == 2 ? (int)(2.5 * C->n[0] * C->n[2])
                                   : (int)(2.5 * C->n[0] * C->n[1] * C->n[2]);
    }
    if (p->bc == OR_BC_DIRICHLET_TEST) {
        /* initDirichlet, poisson_d3.cc:824-873 / poisson_d2.cc:514-543 */
        const level_t *L = &m->lv[0];
        int nx = L->n[0], ny = L->n[1], nz = L->n[2];
        double *fx = malloc(sizeof(double) * (nx + 4 * g->pad + 4));
        double *fy = malloc(sizeof(double) * (ny + 4 * g->pad + 4));
        double *fz = malloc(sizeof(double) * (nz + 4 * g->pad + 4));
        double *t1 = malloc(sizeof(double) * (nx + ny + nz + 8 * g->pad + 8));
        double *t2 = malloc(sizeof(double) * (nx + ny + nz + 8 * g->pad + 8));
        double *t3 = malloc(sizeof(double) * (nx + ny + nz + 8 * g->pad + 8));
        int p0 = g->pad;
        or_field_metrics(g, 0, fx, t1, t2, t3);
        or_field_metrics(g, 2, fz, t1, t2, t3);
        if (m->dim == 3) {
            or_field_metrics(g, 1, fy, t1, t2, t3);
            m->xWall = calloc((size_t)(ny + 2) * (nz + 2), sizeof(double));
            m->yWall = calloc((size_t)(nx + 2) * (nz + 2), sizeof(double));
            m->zWall = calloc((size_t)(nx + 2) * (ny + 2), sizeof(double));
            double xD, yD, zD;
            xD = g->len[0] / 2.0;
            for (int j = 0; j < ny; j++) {
                yD = fy[j + p0] - 0.5;
                for (int k = 0; k < nz; k++) {
                    zD = fz[k + p0] - 0.5;
                    m->xWall[(j + 1) * (nz + 2) + k + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
                }
            }
            yD = g->len[1] / 2.0;
            for (int i = 0; i < nx; i++) {
                xD = fx[i + p0] - 0.5;
                for (int k = 0; k < nz; k++) {
                    zD = fz[k + p0] - 0.5;
                    m->yWall[(i + 1) * (nz + 2) + k + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
                }
            }
            zD = g->len[2] / 2.0;
            for (int i = 0; i < nx; i++) {
                xD = fx[i + p0] - 0.5;
                for (int j = 0; j < ny; j++) {
                    yD = fy[j + p0] - 0.5;
                    m->zWall[(i + 1) * (ny + 2) + j + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
                }
            }
        } else {
            m->xWall = calloc((size_t)(nz + 2), sizeof(double));
            m->zWall = calloc((size_t)(nx + 2), sizeof(double));
            double xD = g->len[0] / 2.0, zD;
            for (int k = 0; k < nz; k++) {
                zD = fz[k + p0] - 0.5;
                m->xWall[k + 1] = (xD * xD + zD * zD) / 4.0;
            }
            zD = g->len[2] / 2.0;
            for (int i = 0; i < nx; i++) {
                xD = fx[i + p0] - 0.5;
                m->zWall[i + 1] = (xD * xD + zD * zD) / 4.0;
            }
        }
        free(fx); free(fy); free(fz); free(t1); free(t2); free(t3);
    }
    return m;
}

void or_mg_destroy(or_mg *m)
{
    if (!m) return;
    for (int l = 0; l <= m->depth; l++) {
        level_t *L = &m->lv[l];
        free(L->lhs); free(L->rhs); free(L->tmp); free(L->smd);
        for (int a = 0; a < 3; a++) { free(L->x[a]); free(L->xixx[a]); free(L->xix2[a]); }
    }
    free(m->lv); free(m->xWall); free(m->yWall); free(m->zWall);
    free(m);
}

int or_mg_levels(const or_mg *m) { return m->depth + 1; }
int or_mg_max_count(const or_mg *m) { return m->max_count; }
void or_mg_level_shape(const or_mg *m, int level, int n[3])
{
    for (int a = 0; a < 3; a++) n[a] = m->lv[level].n[a];
}

/* ------------------------------------------------------------------------
 * Ghost exchange: updateFace (poisson_d3.cc:955-993, poisson_d2.cc:659-690)
 * on a single rank is the periodic self-wrap of full ghost planes, in the
 * order the sends are issued (x, then y, then z).
 * ---------------------------------------------------------------------- */
static void update_face(or_mg *m, level_t *L, double *a)
{
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    if (m->g.periodic[0]) {
        if (m->dim == 3) {
            for (int j = -1; j <= ny; j++) for (int k = -1; k <= nz; k++) a[IX(L, nx, j, k)] = a[IX(L, 0, j, k)];
            for (int j = -1; j <= ny; j++) for (int k = -1; k <= nz; k++) a[IX(L, -1, j, k)] = a[IX(L, nx - 1, j, k)];
        } else {
            for (int k = -1; k <= nz; k++) a[IX(L, nx, 0, k)] = a[IX(L, 0, 0, k)];
            for (int k = -1; k <= nz; k++) a[IX(L, -1, 0, k)] = a[IX(L, nx - 1, 0, k)];
        }
    }
    if (m->dim == 3 && m->g.periodic[1]) {
        for (int i = -1; i <= nx; i++) for (int k = -1; k <= nz; k++) a[IX(L, i, ny, k)] = a[IX(L, i, 0, k)];
        for (int i = -1; i <= nx; i++) for (int k = -1; k <= nz; k++) a[IX(L, i, -1, k)] = a[IX(L, i, ny - 1, k)];
    }
    if (m->g.periodic[2]) {
        int j0 = m->dim == 3 ? -1 : 0, j1 = m->dim == 3 ? ny : 0;
        for (int i = -1; i <= nx; i++) for (int j = j0; j <= j1; j++) a[IX(L, i, j, nz)] = a[IX(L, i, j, 0)];
        for (int i = -1; i <= nx; i++) for (int j = j0; j <= j1; j++) a[IX(L, i, j, -1)] = a[IX(L, i, j, nz - 1)];
    }
}

/* imposeBC, poisson_d3.cc:876-952 / poisson_d2.cc:545-656 (probType 9's
 * Dirichlet top wall of the 2-D Neumann problem by top_dirichlet). */
static void impose_bc(or_mg *m, level_t *L)
{
    double *a = L->lhs;
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    int bc = m->p.bc;
    int inhom = (bc == OR_BC_DIRICHLET_TEST && !m->zeroBC);
    update_face(m, L, a);
    if (!m->g.periodic[0]) {
        int j0 = m->dim == 3 ? -1 : 0, j1 = m->dim == 3 ? ny : 0;
        for (int j = j0; j <= j1; j++)
            for (int k = -1; k <= nz; k++) {
                double wv = 0.0;
                if (inhom) wv = m->dim == 3 ? m->xWall[(j + 1) * (nz + 2) + k + 1] : m->xWall[k + 1];
                if (bc == OR_BC_NEUMANN) {
                    a[IX(L, -1, j, k)] = a[IX(L, 0, j, k)];
                    a[IX(L, nx, j, k)] = a[IX(L, nx - 1, j, k)];
                } else if (inhom) {
                    a[IX(L, -1, j, k)] = 2.0 * wv - a[IX(L, 0, j, k)];
                    a[IX(L, nx, j, k)] = 2.0 * wv - a[IX(L, nx - 1, j, k)];
                } else {
                    a[IX(L, -1, j, k)] = -a[IX(L, 0, j, k)];
                    a[IX(L, nx, j, k)] = -a[IX(L, nx - 1, j, k)];
                }
            }
    }
    if (m->dim == 3 && !m->g.periodic[1]) {
        for (int i = -1; i <= nx; i++)
            for (int k = -1; k <= nz; k++) {
                double wv = inhom ? m->yWall[(i + 1) * (nz + 2) + k + 1] : 0.0;
                if (bc == OR_BC_NEUMANN) {
                    a[IX(L, i, -1, k)] = a[IX(L, i, 0, k)];
                    a[IX(L, i, ny, k)] = a[IX(L, i, ny - 1, k)];
                } else if (inhom) {
                    a[IX(L, i, -1, k)] = 2.0 * wv - a[IX(L, i, 0, k)];
                    a[IX(L, i, ny, k)] = 2.0 * wv - a[IX(L, i, ny - 1, k)];
                } else {
                    a[IX(L, i, -1, k)] = -a[IX(L, i, 0, k)];
                    a[IX(L, i, ny, k)] = -a[IX(L, i, ny - 1, k)];
                }
            }
    }
    if (!m->g.periodic[2]) {
        int j0 = m->dim == 3 ? -1 : 0, j1 = m->dim == 3 ? ny : 0;
        for (int i = -1; i <= nx; i++)
            for (int j = j0; j <= j1; j++) {
                double wv = 0.0;
                if (inhom) wv = m->dim == 3 ? m->zWall[(i + 1) * (ny + 2) + j + 1] : m->zWall[i + 1];
                if (bc == OR_BC_NEUMANN) {
                    a[IX(L, i, j, -1)] = a[IX(L, i, j, 0)];
                    if (m->dim == 2 && m->p.top_dirichlet)
                        a[IX(L, i, j, nz)] = -a[IX(L, i, j, nz - 1)];
                    else
                        a[IX(L, i, j, nz)] = a[IX(L, i, j, nz - 1)];
                } else if (inhom) {
                    a[IX(L, i, j, -1)] = 2.0 * wv - a[IX(L, i, j, 0)];
                    a[IX(L, i, j, nz)] = 2.0 * wv - a[IX(L, i, j, nz - 1)];
                } else {
                    a[IX(L, i, j, -1)] = -a[IX(L, i, j, 0)];
                    a[IX(L, i, j, nz)] = -a[IX(L, i, j, nz - 1)];
                }
            }
    }
}

/* One RB-SOR point update, poisson_d3.cc:109-116 (3-D), poisson_d2.cc:103-108 (2-D). */
static inline double sor3(const level_t *L, const double *x, const double *b, double w, int i, int j, int k)
{
    const double *xix2 = L->xix2[0] + 1, *xixx = L->xixx[0] + 1;
    const double *ety2 = L->xix2[1] + 1, *etyy = L->xixx[1] + 1;
    const double *ztz2 = L->xix2[2] + 1, *ztzz = L->xixx[2] + 1;
    return (1.0 - w) * x[IX(L, i, j, k)] +
           (xix2[i] * L->ihx2[0] * (x[IX(L, i + 1, j, k)] + x[IX(L, i - 1, j, k)]) +
            xixx[i] * L->i2hx[0] * (x[IX(L, i + 1, j, k)] - x[IX(L, i - 1, j, k)]) +
            ety2[j] * L->ihx2[1] * (x[IX(L, i, j + 1, k)] + x[IX(L, i, j - 1, k)]) +
            etyy[j] * L->i2hx[1] * (x[IX(L, i, j + 1, k)] - x[IX(L, i, j - 1, k)]) +
            ztz2[k] * L->ihx2[2] * (x[IX(L, i, j, k + 1)] + x[IX(L, i, j, k - 1)]) +
            ztzz[k] * L->i2hx[2] * (x[IX(L, i, j, k + 1)] - x[IX(L, i, j, k - 1)]) -
            b[IX(L, i, j, k)]) * w / (2.0 * (L->ihx2[0] * xix2[i] + L->ihx2[1] * ety2[j] + L->ihx2[2] * ztz2[k]));
}

static inline double sor2(const level_t *L, const double *x, const double *b, double w, int i, int k)
{
    const double *xix2 = L->xix2[0] + 1, *xixx = L->xixx[0] + 1;
    const double *ztz2 = L->xix2[2] + 1, *ztzz = L->xixx[2] + 1;
    return (1.0 - w) * x[IX(L, i, 0, k)] +
           (xix2[i] * L->ihx2[0] * (x[IX(L, i + 1, 0, k)] + x[IX(L, i - 1, 0, k)]) +
            xixx[i] * L->i2hx[0] * (x[IX(L, i + 1, 0, k)] - x[IX(L, i - 1, 0, k)]) +
            ztz2[k] * L->ihx2[2] * (x[IX(L, i, 0, k + 1)] + x[IX(L, i, 0, k - 1)]) +
            ztzz[k] * L->i2hx[2] * (x[IX(L, i, 0, k + 1)] - x[IX(L, i, 0, k - 1)]) -
            b[IX(L, i, 0, k)]) * w / (2.0 * (L->ihx2[0] * xix2[i] + L->ihx2[2] * ztz2[k]));
}

/* Residual b - A x at one cell, poisson_d3.cc:81-87 / poisson_d2.cc:80-84. */
static inline double res3(const level_t *L, const double *x, const double *b, int i, int j, int k)
{
    const double *xix2 = L->xix2[0] + 1, *xixx = L->xixx[0] + 1;
    const double *ety2 = L->xix2[1] + 1, *etyy = L->xixx[1] + 1;
    const double *ztz2 = L->xix2[2] + 1, *ztzz = L->xixx[2] + 1;
    return b[IX(L, i, j, k)] -
           (xix2[i] * L->ihx2[0] * (x[IX(L, i + 1, j, k)] - 2.0 * x[IX(L, i, j, k)] + x[IX(L, i - 1, j, k)]) +
            xixx[i] * L->i2hx[0] * (x[IX(L, i + 1, j, k)] - x[IX(L, i - 1, j, k)]) +
            ety2[j] * L->ihx2[1] * (x[IX(L, i, j + 1, k)] - 2.0 * x[IX(L, i, j, k)] + x[IX(L, i, j - 1, k)]) +
            etyy[j] * L->i2hx[1] * (x[IX(L, i, j + 1, k)] - x[IX(L, i, j - 1, k)]) +
            ztz2[k] * L->ihx2[2] * (x[IX(L, i, j, k + 1)] - 2.0 * x[IX(L, i, j, k)] + x[IX(L, i, j, k - 1)]) +
            ztzz[k] * L->i2hx[2] * (x[IX(L, i, j, k + 1)] - x[IX(L, i, j, k - 1)]));
}

static inline double res2(const level_t *L, const double *x, const double *b, int i, int k)
{
    const double *xix2 = L->xix2[0] + 1, *xixx = L->xixx[0] + 1;
    const double *ztz2 = L->xix2[2] + 1, *ztzz = L->xixx[2] + 1;
    return b[IX(L, i, 0, k)] -
           (xix2[i] * L->ihx2[0] * (x[IX(L, i + 1, 0, k)] - 2.0 * x[IX(L, i, 0, k)] + x[IX(L, i - 1, 0, k)]) +
            xixx[i] * L->i2hx[0] * (x[IX(L, i + 1, 0, k)] - x[IX(L, i - 1, 0, k)]) +
            ztz2[k] * L->ihx2[2] * (x[IX(L, i, 0, k + 1)] - 2.0 * x[IX(L, i, 0, k)] + x[IX(L, i, 0, k - 1)]) +
            ztzz[k] * L->i2hx[2] * (x[IX(L, i, 0, k + 1)] - x[IX(L, i, 0, k - 1)]));
}

/* The four sub-lattice loops of one colour (red: 000,110,101,011;
 * black: 100,010,001,111), poisson_d3.cc:105-242. */
static void colour3(or_mg *m, level_t *L, int black)
{
    static const int red[4][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    static const int blk[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};
    const int (*s)[3] = black ? blk : red;
    double *x = L->lhs;
    const double *b = L->rhs;
    double w = m->p.sor;
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    for (int q = 0; q < 4; q++) {
        int i0 = s[q][0], j0 = s[q][1], k0 = s[q][2];
#pragma omp parallel for num_threads(nthr()) schedule(static)
        for (int i = i0; i <= nx - 1; i += 2)
            for (int j = j0; j <= ny - 1; j += 2)
                for (int k = k0; k <= nz - 1; k += 2)
                    x[IX(L, i, j, k)] = sor3(L, x, b, w, i, j, k);
    }
}

static void colour2(or_mg *m, level_t *L, int black)
{
    static const int red[2][2] = {{0, 0}, {1, 1}};
    static const int blk[2][2] = {{1, 0}, {0, 1}};
    const int (*s)[2] = black ? blk : red;
    double *x = L->lhs;
    const double *b = L->rhs;
    double w = m->p.sor;
    int nx = L->n[0], nz = L->n[2];
    for (int q = 0; q < 2; q++) {
        int i0 = s[q][0], k0 = s[q][1];
#pragma omp parallel for num_threads(nthr()) schedule(static)
        for (int i = i0; i <= nx - 1; i += 2)
            for (int k = k0; k <= nz - 1; k += 2)
                x[IX(L, i, 0, k)] = sor2(L, x, b, w, i, k);
    }
}

static void sweep(or_mg *m, level_t *L)
{
    impose_bc(m, L);
    if (m->dim == 3) colour3(m, L, 0); else colour2(m, L, 0);
    update_face(m, L, L->lhs);
    if (m->dim == 3) colour3(m, L, 1); else colour2(m, L, 1);
}

/* multigrid_d3::smooth, poisson_d3.cc:96-246 */
static void smooth(or_mg *m, int count)
{
    level_t *L = &m->lv[m->vLevel];
    for (int n = 0; n < count; n++) sweep(m, L);
    impose_bc(m, L);
}

/* multigrid_d3::solve, poisson_d3.cc:249-432 */
static void solve(or_mg *m)
{
    level_t *L = &m->lv[m->vLevel];
    int iterCount = 0;
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    for (;;) {
        sweep(m, L);
        double localMax = -1.0e-10;
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    double t = m->dim == 3 ? fabs(res3(L, L->lhs, L->rhs, i, j, k))
                                           : fabs(res2(L, L->lhs, L->rhs, i, k));
                    if (t > localMax) localMax = t;
                }
        if (localMax < m->p.mg_tol) break;
        iterCount += 1;
        if (iterCount > m->max_count) { m->coarse_capped = 1; break; }
    }
    m->last_coarse_iters = iterCount;
    impose_bc(m, L);
}

/* multigrid_d3::computeResidual, poisson_d3.cc:71-93 (into tmp) */
static void compute_residual(or_mg *m)
{
    level_t *L = &m->lv[m->vLevel];
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    memset(L->tmp, 0, level_size(L, m->dim) * sizeof(double));
#pragma omp parallel for num_threads(nthr()) schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                L->tmp[IX(L, i, j, k)] = m->dim == 3 ? res3(L, L->lhs, L->rhs, i, j, k)
                                                     : res2(L, L->lhs, L->rhs, i, k);
    update_face(m, L, L->tmp);
}

/* multigrid_d3::coarsen :435-513 (regular branch), multigrid_d2::coarsen :258-300 */
static void coarsen(or_mg *m)
{
    int pl = m->vLevel;
    m->vLevel += 1;
    level_t *F = &m->lv[pl], *C = &m->lv[m->vLevel];
    int nx = C->n[0], ny = C->n[1], nz = C->n[2];
    if (m->dim == 2) {
#pragma omp parallel for num_threads(nthr()) schedule(static)
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) {
                int i2 = i * 2, k2 = k * 2;
                C->rhs[IX(C, i, 0, k)] = (F->tmp[IX(F, i2 + 1, 0, k2 + 1)] + F->tmp[IX(F, i2, 0, k2)] +
                                          F->tmp[IX(F, i2 + 1, 0, k2)] + F->tmp[IX(F, i2, 0, k2 + 1)]) / 4;
            }
        return;
    }
    const double *xc = C->x[0] + 1, *yc = C->x[1] + 1, *zc = C->x[2] + 1;
    const double *xf = F->x[0] + 1, *yf = F->x[1] + 1, *zf = F->x[2] + 1;
    const double *t = F->tmp;
#pragma omp parallel for num_threads(nthr()) schedule(static)
    for (int i = 0; i < nx; i++) {
        int i2 = i * 2;
        for (int j = 0; j < ny; j++) {
            int j2 = j * 2;
            for (int k = 0; k < nz; k++) {
                int k2 = k * 2;
                double v000 = (xc[i] - xf[i2]) * (yc[j] - yf[j2]) * (zc[k] - zf[k2]);
                double v100 = (xf[i2 + 1] - xc[i]) * (yc[j] - yf[j2]) * (zc[k] - zf[k2]);
                double v010 = (xc[i] - xf[i2]) * (yf[j2 + 1] - yc[j]) * (zc[k] - zf[k2]);
                double v001 = (xc[i] - xf[i2]) * (yc[j] - yf[j2]) * (zf[k2 + 1] - zc[k]);
                double v011 = (xc[i] - xf[i2]) * (yf[j2 + 1] - yc[j]) * (zf[k2 + 1] - zc[k]);
                double v101 = (xf[i2 + 1] - xc[i]) * (yc[j] - yf[j2]) * (zf[k2 + 1] - zc[k]);
                double v110 = (xf[i2 + 1] - xc[i]) * (yf[j2 + 1] - yc[j]) * (zc[k] - zf[k2]);
                double v111 = (xf[i2 + 1] - xc[i]) * (yf[j2 + 1] - yc[j]) * (zf[k2 + 1] - zc[k]);
                double vTot = (xf[i2 + 1] - xf[i2]) * (yf[j2 + 1] - yf[j2]) * (zf[k2 + 1] - zf[k2]);
                C->rhs[IX(C, i, j, k)] = (v111 * t[IX(F, i2 + 1, j2 + 1, k2 + 1)] + v000 * t[IX(F, i2, j2, k2)] +
                                          v110 * t[IX(F, i2 + 1, j2 + 1, k2)] + v001 * t[IX(F, i2, j2, k2 + 1)] +
                                          v101 * t[IX(F, i2 + 1, j2, k2 + 1)] + v010 * t[IX(F, i2, j2 + 1, k2)] +
                                          v011 * t[IX(F, i2, j2 + 1, k2 + 1)] + v100 * t[IX(F, i2 + 1, j2, k2)]) / vTot;
            }
        }
    }
}

/* multigrid_d3::prolong :516-560 (regular branch) */
static void prolong(or_mg *m)
{
    int pl = m->vLevel;
    m->vLevel -= 1;
    level_t *F = &m->lv[m->vLevel], *C = &m->lv[pl];
    memset(F->lhs, 0, level_size(F, m->dim) * sizeof(double));
    int nx = F->n[0], ny = F->n[1], nz = F->n[2];
#pragma omp parallel for num_threads(nthr()) schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++)
                F->lhs[IX(F, i, j, k)] = C->lhs[IX(C, i / 2, j / 2, k / 2)];
}

/* poisson::vCycle, poisson.cc:203-279 */
static void vcycle(or_mg *m)
{
    m->vLevel = 0;
    m->zeroBC = 0;
    smooth(m, m->p.pre_smooth);
    m->zeroBC = 1;
    for (int i = 0; i < m->depth; i++) {
        compute_residual(m);
        level_t *L = &m->lv[m->vLevel];
        memcpy(L->smd, L->lhs, level_size(L, m->dim) * sizeof(double));
        coarsen(m);
        level_t *C = &m->lv[m->vLevel];
        memset(C->lhs, 0, level_size(C, m->dim) * sizeof(double));
        if (m->vLevel == m->depth) {
            if (m->p.solve_flag) solve(m);
            else smooth(m, m->p.pre_smooth);
        } else smooth(m, m->p.pre_smooth);
    }
    for (int i = 0; i < m->depth; i++) {
        prolong(m);
        level_t *L = &m->lv[m->vLevel];
        size_t sz = level_size(L, m->dim);
        for (size_t q = 0; q < sz; q++) L->lhs[q] += L->smd[q];
        m->zeroBC = (m->vLevel == 0) ? 0 : 1;
        smooth(m, m->p.post_smooth);
    }
}

/* multigrid_d3::computeError, poisson_d3.cc:563-663 (serial sums). */
static double compute_error(or_mg *m, int type)
{
    level_t *L = &m->lv[0];
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    double pointCount = m->dim == 3 ? (double)nx * ny * nz : (double)nx * nz;
    double num = 0.0, den = 0.0, res = 0.0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double r = m->dim == 3 ? res3(L, L->lhs, L->rhs, i, j, k) : res2(L, L->lhs, L->rhs, i, k);
                double b = L->rhs[IX(L, i, j, k)];
                if (type == 0) {
                    double t = fabs(r);
                    if (t > num) num = t;
                    if (fabs(b) > den) den = fabs(b);
                } else if (type == 1) {
                    num += fabs(r);
                    den += fabs(b);
                } else {
                    num += r * r;
                    den += pow(b, 2);
                }
            }
    if (type == 0) res = den != 0.0 ? num / den : num;
    else if (type == 1) res = den != 0.0 ? num / den : num / pointCount;
    else res = den != 0.0 ? sqrt(num / pointCount) / sqrt(den / pointCount) : sqrt(num / pointCount);
    return res;
}

static void set_core_from_field(or_mg *m, double *dst, const double *src, int pad)
{
    level_t *L = &m->lv[0];
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    long fz = nz + 2 * pad, fy = ny + 2 * pad;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long s = m->dim == 3 ? ((long)(i + pad) * fy + (j + pad)) * fz + (k + pad)
                                     : (long)(i + pad) * fz + (k + pad);
                dst[IX(L, i, j, k)] = src[s];
            }
}

/* poisson::mgSolve, poisson.cc:109-189 */
int or_mg_solve(or_mg *m, double *x_full, const double *b_full, int pad,
                double *history, or_mg_stats *st)
{
    m->vLevel = 0;
    m->coarse_capped = 0;
    int cmax = 0;
    for (int l = 0; l <= m->depth; l++) {
        size_t sz = level_size(&m->lv[l], m->dim);
        memset(m->lv[l].lhs, 0, sz * sizeof(double));
        memset(m->lv[l].rhs, 0, sz * sizeof(double));
        memset(m->lv[l].smd, 0, sz * sizeof(double));
    }
    set_core_from_field(m, m->lv[0].rhs, b_full, pad);
    set_core_from_field(m, m->lv[0].lhs, x_full, pad);
    update_face(m, &m->lv[0], m->lv[0].rhs);
    update_face(m, &m->lv[0], m->lv[0].lhs);
    int cycles = 0;
    double res = 0.0;
    for (int c = 0; c < m->p.vc_count; c++) {
        m->last_coarse_iters = 0;
        vcycle(m);
        if (m->last_coarse_iters > cmax) cmax = m->last_coarse_iters;
        res = compute_error(m, m->p.res_type);
        cycles = c + 1;
        if (history) history[c] = res;
        if (!m->p.no_early_exit && res < m->p.vc_tol) break;
    }
    /* outLHS.F(stagFull(0)) = lhs(0)(stagFull(0)) */
    level_t *L = &m->lv[0];
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    long fz = nz + 2 * pad, fy = ny + 2 * pad;
    int j0 = m->dim == 3 ? -1 : 0, j1 = m->dim == 3 ? ny : 0;
    for (int i = -1; i <= nx; i++)
        for (int j = j0; j <= j1; j++)
            for (int k = -1; k <= nz; k++) {
                long s = m->dim == 3 ? ((long)(i + pad) * fy + (j + pad)) * fz + (k + pad)
                                     : (long)(i + pad) * fz + (k + pad);
                x_full[s] = L->lhs[IX(L, i, j, k)];
            }
    if (st) {
        st->cycles = cycles;
        st->residual = res;
        st->coarse_iters_max = cmax;
        st->coarse_capped = m->coarse_capped;
    }
    return 0;
}

/* ---- single-operation entry points ---- */
void or_mg_set_level(or_mg *m, int level, const double *lhs_pad, const double *rhs_pad)
{
    level_t *L = &m->lv[level];
    size_t sz = level_size(L, m->dim);
    if (lhs_pad) memcpy(L->lhs, lhs_pad, sz * sizeof(double));
    if (rhs_pad) memcpy(L->rhs, rhs_pad, sz * sizeof(double));
}
void or_mg_get_level(const or_mg *m, int level, double *lhs_pad, double *rhs_pad)
{
    const level_t *L = &m->lv[level];
    size_t sz = level_size(L, m->dim);
    if (lhs_pad) memcpy(lhs_pad, L->lhs, sz * sizeof(double));
    if (rhs_pad) memcpy(rhs_pad, L->rhs, sz * sizeof(double));
}
void or_mg_smooth(or_mg *m, int level, int count, int zero_bc)
{
    m->vLevel = level; m->zeroBC = zero_bc;
    smooth(m, count);
}
int or_mg_coarse_solve(or_mg *m, int level, int zero_bc)
{
    m->vLevel = level; m->zeroBC = zero_bc;
    solve(m);
    return m->last_coarse_iters;
}
void or_mg_residual_restrict(or_mg *m, int level)
{
    m->vLevel = level;
    compute_residual(m);
    level_t *L = &m->lv[level];
    memcpy(L->smd, L->lhs, level_size(L, m->dim) * sizeof(double));
    coarsen(m);
    level_t *C = &m->lv[level + 1];
    memset(C->lhs, 0, level_size(C, m->dim) * sizeof(double));
}
void or_mg_prolong_correct(or_mg *m, int level)
{
    m->vLevel = level + 1;
    prolong(m);
    level_t *L = &m->lv[level];
    size_t sz = level_size(L, m->dim);
    for (size_t q = 0; q < sz; q++) L->lhs[q] += L->smd[q];
}
double or_mg_error(or_mg *m, int res_type) { return compute_error(m, res_type); }
void or_mg_impose_bc(or_mg *m, int level, int zero_bc)
{
    m->vLevel = level; m->zeroBC = zero_bc;
    impose_bc(m, &m->lv[level]);
}

void or_mg_vcycles(or_mg *m, int count)
{
    for (int c = 0; c < count; c++) vcycle(m);
}

void or_mg_residual(or_mg *m, int level, int zero_bc, double *out_pad)
{
    m->vLevel = level; m->zeroBC = zero_bc;
    level_t *L = &m->lv[level];
    impose_bc(m, L);
    compute_residual(m);
    memcpy(out_pad, L->tmp, level_size(L, m->dim) * sizeof(double));
    int nx = L->n[0], ny = L->n[1], nz = L->n[2];
    /* zero the ghost shell: only the core is a residual */
    for (int i = -1; i <= nx; i++)
        for (int j = (m->dim == 3 ? -1 : 0); j <= (m->dim == 3 ? ny : 0); j++)
            for (int k = -1; k <= nz; k++)
                if (i < 0 || i >= nx || k < 0 || k >= nz || (m->dim == 3 && (j < 0 || j >= ny)))
                    out_pad[IX(L, i, j, k)] = 0.0;
}
