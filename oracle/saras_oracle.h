/*
 * saras_oracle.h -- FP64 CPU restatement of the SARAS-CC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; nothing in saras-cc_amd/
 * links, includes or calls it.
 *
 * Every function restates the single-domain (one MPI rank) branch of the
 * reference algorithm, following the loop order and floating-point
 * expression order of the cited reference lines, so that the GPU FP64 path
 * can be compared bit-for-bit where the reference itself is deterministic.
 *
 * Layout: a level array is (nx+2) x (ny+2) x (nz+2) doubles, row-major with
 * z fastest, lower bound -1 on every axis (the reference's Blitz arrays with
 * pad 1, lib/poisson/poisson.cc:386-391).  Planar (2-D) problems use ny = 1
 * and no y ghost layer: (nx+2) x 1 x (nz+2).
 */
#ifndef SARAS_ORACLE_H
#define SARAS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid types per axis (lib/io/parser.cc:417-448). */
enum { OR_GRID_UNIFORM = 0, OR_GRID_TANH_SINGLE = 1, OR_GRID_TANH_DOUBLE = 2 };

/* Wall treatment of non-periodic axes in the Poisson solver.
 *   NEUMANN     x(-1) =  x(0)          poisson_d3.cc:892-895 (allNeumann, production)
 *   DIRICHLET   x(-1) = -x(0)          poisson_d3.cc:896-900
 *   DIRICHLET_TEST x(-1) = 2 w - x(0), w = r^2/6 (3-D) or (x^2+z^2)/4 (2-D),
 *               homogeneous (-x(0)) below level 0 / when zeroBC
 *               (TEST_POISSON: poisson_d3.cc:881-890, 824-873; poisson_d2.cc:561-637) */
enum { OR_BC_NEUMANN = 0, OR_BC_DIRICHLET = 1, OR_BC_DIRICHLET_TEST = 2 };

typedef struct {
    int dim;          /* 2 (planar, x-z) or 3 */
    int n[3];         /* global cells nx, ny, nz (ny ignored when dim == 2) */
    double len[3];    /* domain lengths Lx, Ly, Lz */
    int periodic[3];  /* "Domain Type" P/N per axis */
    int gtype[3];     /* OR_GRID_* ("Mesh Type" U/S/D) */
    double beta[3];   /* tanh stretching parameters */
    int pad;          /* field pad width: 1 (dScheme 1) or 2 (dScheme 2) */
} or_grid_desc;

typedef struct {
    int vc_count;     /* "V-Cycle Count" */
    double vc_tol;    /* "Residual Tolerance" */
    int solve_flag;   /* "Solve Coarsest" */
    double mg_tol;    /* "Solve Tolerance" (coarsest RB-GS, absolute) */
    double sor;       /* "SOR Parameter" */
    int pre_smooth;   /* "Pre-Smoothing Count" */
    int post_smooth;  /* "Post-Smoothing Count" */
    int res_type;     /* "Residual Type": 0 Linf, 1 L1, 2 L2 */
    int bc;           /* OR_BC_* */
    int no_early_exit;/* TEST_POISSON: run all vc_count cycles (poisson.cc:182-184) */
    int top_dirichlet;/* 2-D all-Neumann problem of probType 9: Dirichlet top wall (poisson_d2.cc:641-647) */
} or_mg_params;

typedef struct {
    int cycles;            /* V-cycles performed */
    double residual;       /* last computeError value */
    int coarse_iters_max;  /* largest coarsest-level iteration count seen */
    int coarse_capped;     /* coarsest solve hit maxCount at least once */
} or_mg_stats;

/* ---- grid (lib/grid/grid.cc) ---- */

/* V-cycle depth vcdGlo (grid.cc:678-742, single domain). */
int or_vc_depth(const or_grid_desc *g);

/* Level-l metric arrays for one axis over indices [-1, n_l]: coordinate,
 * xi, xi_x, xi_xx, xi_x^2 (globalMetrics(15 l + 5 axis + {1,0,2,3,4}),
 * grid.cc:521-666, 411-508).  Each output holds n_l + 2 doubles. */
int or_level_metrics(const or_grid_desc *g, int level, int axis,
                     double *coord, double *xi, double *xi_x, double *xixx, double *xix2);

/* Field-grid coordinate and metrics of one axis over [-pad, n+pad-1]
 * (grid.cc:318-508, the arrays x, xi_x, xixx, xix2 the field stencils use). */
int or_field_metrics(const or_grid_desc *g, int axis,
                     double *coord, double *xi_x, double *xixx, double *xix2);

/* ---- multigrid (lib/poisson) ---- */

typedef struct or_mg or_mg;

or_mg *or_mg_create(const or_grid_desc *g, const or_mg_params *p);
void or_mg_destroy(or_mg *m);
int or_mg_levels(const or_mg *m);                    /* vcdGlo + 1 */
void or_mg_level_shape(const or_mg *m, int level, int n[3]);
int or_mg_max_count(const or_mg *m);                 /* coarse iteration cap */

/* poisson::mgSolve (poisson.cc:109-189).  x_full / b_full are caller
 * field arrays with pad width `pad`: (nx+2p)(ny+2p)(nz+2p) (2-D: ny term = 1).
 * The core of x_full is the initial guess; on return the core and the first
 * ghost layer of x_full hold the solution, exactly as outLHS.F(stagFull(0)).
 * history (optional, >= vc_count entries) receives the residual after each
 * V-cycle. */
int or_mg_solve(or_mg *m, double *x_full, const double *b_full, int pad,
                double *history, or_mg_stats *st);

/* Single-operation entry points on the solver's level arrays, for
 * kernel-level parity tests.  Arrays are padded level arrays as above. */
void or_mg_set_level(or_mg *m, int level, const double *lhs_pad, const double *rhs_pad);
void or_mg_get_level(const or_mg *m, int level, double *lhs_pad, double *rhs_pad);
void or_mg_smooth(or_mg *m, int level, int count, int zero_bc);    /* poisson_d3.cc:96-246 */
int  or_mg_coarse_solve(or_mg *m, int level, int zero_bc);         /* poisson_d3.cc:249-432; returns iterations */
void or_mg_residual_restrict(or_mg *m, int level);                 /* :71-93 then :435-513 into rhs(level+1), lhs(level+1)=0 */
void or_mg_prolong_correct(or_mg *m, int level);                   /* :516-560 + poisson.cc:266, level+1 -> level */
double or_mg_error(or_mg *m, int res_type);                        /* :563-663 at level 0 */
void or_mg_impose_bc(or_mg *m, int level, int zero_bc);            /* :876-952 */
void or_mg_vcycles(or_mg *m, int count);                           /* poisson.cc:203-279, repeated */
/* imposeBC then the residual b - A x of computeResidual (:71-93) at a level,
 * written into a padded array (ghosts zero). */
void or_mg_residual(or_mg *m, int level, int zero_bc, double *out_pad);

/* ---- Navier-Stokes / Boussinesq step (lib/field, lib/timestep) ---- */

/* Field boundary kinds (lib/boundary/dirichlet.cc, neumann.cc).  Walls are
 * numbered 0 x-lo (Lft), 1 x-hi (Rgt), 2 y-lo (Frn), 3 y-hi (Bak),
 * 4 z-lo (Bot), 5 z-hi (Top); periodic axes ignore theirs. */
/* wall kinds: dirichlet.cc, neumann.cc; for T also hotPlate (bottom wall,
 * hotPlate.cc) and plumeOut (top wall, plumeOut.cc); for P also nullBC
 * (boundary.h:205-212), whose imposeBC writes nothing */
enum { OR_FBC_DIRICHLET = 0, OR_FBC_NEUMANN = 1, OR_FBC_HOT_PLATE = 2, OR_FBC_PLUME_OUT = 3, OR_FBC_NULL = 4 };

typedef struct {
    or_grid_desc g;        /* field grid; g.pad = 1 (dScheme 1) or 2 (dScheme 2) */
    or_mg_params mg;       /* pressure-correction solve (bc forced to Neumann, poisson.cc:87) */
    int dscheme;           /* "Differentiation Scheme" 1: 2nd order, 2: 4th order */
    int nlscheme;          /* "Non-Linear Scheme" 1: central, 2: hybrid upwind, 3: Morinishi */
    int has_T;             /* scalar (Boussinesq) run: lsRK3_d3::timeAdvance(V, P, T) */
    double nu, kappa;      /* timestep.cc:70-94 */
    double dt;             /* "Time-Step" */
    double cn_tol;         /* "CN Tolerance" (absolute) */
    double buoy[3];        /* gTerms = Fb g / |g| (buoyantForce.cc:46-61) */
    int vbc_type[3][6];    /* per velocity component, per wall: OR_FBC_* */
    double vbc_val[3][6];
    int tbc_type[6];
    double tbc_val[6];
    int pbc_type[6];
    double pbc_val[6];
    int scheme;            /* "Integration Scheme": 0 lsRK3_d3, 1 eulerCN_d3 */
    double pe_limit;       /* "Peclet Limit" (upwind scheme) */
    double central_bias;   /* "Central Bias" omega (upwind scheme) */
    int les;               /* "LES Model" 1: stretched spiral vortex sub-grid stress (fully periodic runs) */
    double mean_pgrad;     /* "Mean Pressure Gradient" of constantPGrad, force type 5 (force.h:148-153) */
    double plate_radius;   /* "Plate Radius" of the hot plate */
} or_nse_desc;

typedef struct or_nse or_nse;

or_nse *or_nse_create(const or_nse_desc *d);
void or_nse_destroy(or_nse *s);
int or_nse_max_iterations(const or_nse *s);          /* timestep.cc:115 */
/* Load fields (full arrays, pad g.pad; NULL keeps the current one), then
 * impose the BCs of V, P and T. */
void or_nse_set(or_nse *s, const double *vx, const double *vy, const double *vz,
                const double *p, const double *t, const double *pp);
void or_nse_get(const or_nse *s, double *vx, double *vy, double *vz, double *p, double *t, double *pp);
/* The explicit part of one RK3 sub-stage (lsRK3_d3.cc:199-267): fills the
 * stage RHS (nseRHS, tmpRHS) and leaves the non-linear terms in tempVF/tempSF.
 * Euler-CN runs (scheme 1): the explicit part of eulerCN_d3::timeAdvance
 * (eulerCN_d3.cc:92-121, 181-235); stage is ignored. */
int or_nse_rhs(or_nse *s, int stage);
void or_nse_get_rhs(const or_nse *s, double *rx, double *ry, double *rz, double *rt,
                    double *hx, double *hy, double *hz, double *ht);
/* One RK3 sub-stage (lsRK3_d3.cc:199-304) / one full step (three
 * sub-stages, or the single Euler-CN stage of eulerCN_d3.cc:88-165,
 * 178-270); 0, or 2 when a Jacobi solve exceeds maxIterations
 * (predictors.cc:133-141). */
int or_nse_stage(or_nse *s, int stage);
int or_nse_step(or_nse *s);
void or_nse_jacobi_iters(const or_nse *s, int out[4]);
void or_nse_mg_stats(const or_nse *s, or_mg_stats *st);   /* last pressure-correction solve */
/* Single operators over the core (pads of out zero): 0-2 first derivative
 * along x/y/z, 3-5 second derivative, 6 computeDiff (Laplacian), 7 scalar
 * computeNLin with the current V (H -= ...). */
void or_nse_op(or_nse *s, int op, const double *F, double *out);
/* vector operators: 0 divergence, 1 gradient, 2 computeDiff, 3 computeNLin */
void or_nse_vec_op(or_nse *s, int op, const double *fx, const double *fy, const double *fz, double *ox, double *oy,
                   double *oz);
/* tseries.cc:209-290 diagnostics: volume-mean KE, max |div V|, NuVol, and
 * the wall Nusselt number of z-gravity runs. */
void or_nse_diag(or_nse *s, double *ke, double *divmax, double *nuvol, double *nuwall);
double or_nse_thermal_energy(or_nse *s);
/* vfield::computeTStp (vfield.cc:691-706) then the driver's cap
 * (hydro_d3.cc:221-226): dt = min(courant / max(|u|/dx + |v|/dy + |w|/dz), dt_max);
 * the run's dt becomes that value. */
double or_nse_cfl(or_nse *s, double courant, double dt_max);
/* spiral::computeSG(nseRHS, V) (spiral.cc:129-236) on a caller array set
 * H[3] (full arrays): the sub-grid stresses of the current V and
 * H -= div T; returns 0, or 3 where the reference aborts (spiral.cc:643-650,
 * 712-719, 732-737).  Diagnostics of the last evaluation: sums of |K| dV,
 * dissipation dV and nu_t dV (tseries divides them by the volume). */
int or_nse_sgs(or_nse *s, double *hx, double *hy, double *hz);
void or_nse_sgs_diag(const or_nse *s, double out[3]);
/* the six stress fields Txx, Tyy, Tzz, Txy, Tyz, Tzx of the last evaluation
 * (full arrays, one after the other) */
void or_nse_get_sgs(const or_nse *s, double *T6);
double or_nse_time(const or_nse *s);

/* Threads the oracle's OpenMP loops use (0 = runtime default). */
void or_set_threads(int n);

/* nu, kappa and gTerms of a run by the deck (timestep.cc:70-98,
 * buoyantForce.cc:48-60); -1 where the reference aborts. */
int or_nse_physics(int prob_type, double re, double ra, double pr, int rbc_type, const double g[3], double *nu,
                   double *kappa, double buoy[3]);

#ifdef __cplusplus
}
#endif
#endif
