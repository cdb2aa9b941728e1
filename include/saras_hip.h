/*
 * saras_hip.h -- C-ABI of the MI355X-native SARAS-CC hot path.
 *
 * The reference calls its hot path through C++ classes; every entry point
 * below names the reference interface it replaces (file:line in
 * roshansamuel/saras-cc).  Plain pointers and sizes only: no C++ or torch
 * types cross this boundary.
 *
 * Array convention (same as the reference's Blitz arrays): a field of a 3-D
 * grid with nx x ny x nz cells and pad width p is a dense row-major array of
 * (nx+2p)(ny+2p)(nz+2p) values, z fastest, whose element (i, j, k) for
 * i, j, k in [-p, n+p-1] sits at ((i+p)(ny+2p) + (j+p))(nz+2p) + (k+p).
 * Planar (2-D, x-z) problems drop the y axis: (nx+2p)(nz+2p).
 *
 * Threading: a handle is not thread-safe.  Distinct handles may be used from
 * distinct threads on distinct streams.
 */
#ifndef SARAS_HIP_H
#define SARAS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes replace the reference's `MPI_Finalize(); exit(0)` aborts and
 * its printed warnings (poisson_d3.cc:422-428, predictors.cc:137-143,
 * grid.cc:179-184, parser.cc:299-410). */
typedef enum {
    SARAS_OK = 0,
    SARAS_NOT_CONVERGED = 1,   /* V-cycle count exhausted above tolerance */
    SARAS_JACOBI_DIVERGED = 2, /* Crank-Nicolson Jacobi iteration cap hit */
    SARAS_BAD_SHAPE = 3,       /* grid cannot be coarsened / bad extents */
    SARAS_OOM = 4,             /* device allocation failed */
    SARAS_HIP_ERROR = 5,       /* a HIP runtime call failed */
    SARAS_BAD_ARG = 6,         /* invalid parameter or null pointer */
    SARAS_SGS_FAILED = 7       /* LES eigen-decomposition failed where the reference
                                  aborts (spiral.cc:643-650, 712-719, 732-737) */
} saras_status;

/* Grid types per axis ("Mesh Type" U/S/D, lib/io/parser.cc:417-448). */
enum { SARAS_GRID_UNIFORM = 0, SARAS_GRID_TANH_SINGLE = 1, SARAS_GRID_TANH_DOUBLE = 2 };

/* Wall rule of the non-periodic axes of the Poisson problem:
 *   NEUMANN          ghost = inner             (poisson_d3.cc:892-895, the
 *                                               production pressure BC)
 *   DIRICHLET        ghost = -inner            (poisson_d3.cc:896-900)
 *   DIRICHLET_TEST   ghost = 2 w - inner, w = r^2/6 (2-D: (x^2+z^2)/4) on
 *                    the finest level, -inner on coarser levels
 *                    (TEST_POISSON, poisson_d3.cc:881-890, 824-873) */
enum { SARAS_BC_NEUMANN = 0, SARAS_BC_DIRICHLET = 1, SARAS_BC_DIRICHLET_TEST = 2 };

/* Arithmetic of the multigrid solve. */
enum {
    SARAS_PREC_FP64 = 0,   /* plain FP64: the reference algorithm, bit-exact */
    SARAS_PREC_MIXED = 1   /* FP32 V-cycles inside FP64 residual/refinement */
};

/* Where the caller's arrays live. */
enum { SARAS_PTR_HOST = 0, SARAS_PTR_DEVICE = 1 };

/* grid(const parser&, parallel&) (lib/grid/grid.h:165) reduced to the
 * inputs the hot path reads: "Mesh" and "Program: Domain Type" keys. */
typedef struct {
    int dim;            /* 3, or 2 for planar x-z problems */
    int n[3];           /* global cells nx, ny, nz (ny ignored when dim == 2) */
    double len[3];      /* Lx, Ly, Lz */
    int periodic[3];    /* 1 if the axis is periodic */
    int gtype[3];       /* SARAS_GRID_* */
    double beta[3];     /* tanh stretching parameter per axis */
    int pad;            /* pad width of caller field arrays: 1 or 2 */
} saras_grid_desc;

/* The "Multigrid" keys of the reference deck (parser.cc:172-183) plus the
 * precision mode of this implementation. */
typedef struct {
    int vc_count;       /* "V-Cycle Count" */
    double vc_tol;      /* "Residual Tolerance" */
    int solve_flag;     /* "Solve Coarsest" */
    double mg_tol;      /* "Solve Tolerance" (absolute, coarsest level) */
    double sor;         /* "SOR Parameter" */
    int pre_smooth;     /* "Pre-Smoothing Count" */
    int post_smooth;    /* "Post-Smoothing Count" */
    int res_type;       /* "Residual Type": 0 L-inf, 1 L1, 2 L2 (relative) */
    int bc;             /* SARAS_BC_* */
    int precision;      /* SARAS_PREC_* */
    int inner_cycles;   /* MIXED: most FP32 V-cycles per FP64 refinement step; the first step runs at
                           most 3 to measure the convergence factor, later ones the cycles that factor
                           needs to reach the tolerance (or the FP32 accuracy floor) */
    int max_refine;     /* MIXED: cap on refinement steps */
    int top_dirichlet;  /* 1: the top (z-hi) wall of a 2-D Neumann problem is homogeneous Dirichlet, as
                           multigrid_d2 makes it for probType 9 (poisson_d2.cc:641-647); 2-D grids with
                           z walls and SARAS_BC_NEUMANN only */
} saras_mg_params;

typedef struct {
    int cycles;           /* V-cycles performed (FP32 cycles in MIXED mode) */
    int refinements;      /* FP64 refinement steps (MIXED), 0 otherwise */
    double residual;      /* final residual of type res_type */
    int converged;        /* residual < vc_tol */
    int coarse_iters_max; /* largest coarsest-level RB-GS iteration count */
    int coarse_capped;    /* the coarse iteration cap was hit (reference warns) */
    double device_ms;     /* device time of the solve (HIP events) */
} saras_mg_stats;

typedef struct saras_mg saras_mg;

/* poisson::poisson + multigrid_d3::multigrid_d3 / multigrid_d2::multigrid_d2
 * (lib/poisson/poisson.cc:66-92, poisson_d3.cc:60-68, poisson_d2.cc:60-68):
 * builds the level hierarchy (setStagBounds, setCoefficients, copyDerivs,
 * initializeArrays) resident in device memory. */
saras_status saras_mg_create(const saras_grid_desc *grid, const saras_mg_params *params,
                             saras_mg **out);

/* poisson::mgSolve(plainsf &outLHS, const plainsf &inpRHS)
 * (lib/poisson/poisson.cc:109-189).  x_full is in/out: its core is the
 * initial guess and on return its core and first ghost layer hold the
 * solution (outLHS.F(stagFull(0)) = lhs(0)(stagFull(0))).  b_full is the RHS
 * (core read).  ptr_kind: SARAS_PTR_HOST or SARAS_PTR_DEVICE.  stream: a
 * hipStream_t or NULL for the handle's own stream.  The call returns after
 * the stream has finished the solve; with device pointers the arrays must
 * stay unchanged by other streams while it runs (the mixed-precision solve
 * reads b_full and the initial guess in place throughout).  Returns SARAS_OK
 * when the residual fell below vc_tol, SARAS_NOT_CONVERGED when vc_count
 * cycles ran out (the reference is silent there, poisson.cc:127-185). */
saras_status saras_mg_solve(saras_mg *mg, double *x_full, const double *b_full, int ptr_kind,
                            void *stream, saras_mg_stats *stats);

/* poisson::~poisson (poisson.cc:781). */
saras_status saras_mg_destroy(saras_mg *mg);

/* Number of levels (vcdGlo + 1, grid.cc:678-742) and extents of a level. */
int saras_mg_levels(const saras_mg *mg);
saras_status saras_mg_level_shape(const saras_mg *mg, int level, int n[3]);

/* V-cycle depth the reference would pick for this grid (grid::setVCDepth,
 * grid.cc:678-742); negative when the grid cannot be coarsened. */
int saras_vc_depth(const saras_grid_desc *grid);

/* Level metrics of one axis over [-1, n_l] as the reference stores them in
 * grid::globalMetrics (grid.cc:521-666): coordinate, xi, xi_x, xi_xx, xi_x^2.
 * Each output holds n_l + 2 doubles. */
saras_status saras_level_metrics(const saras_grid_desc *grid, int level, int axis, double *coord,
                                 double *xi, double *xi_x, double *xixx, double *xix2);

/* ---- single operations on one level (the private members of poisson the
 * V-cycle is built from), for kernel-level parity testing.  Level arrays are
 * padded (n_l+2)^3 FP64 arrays in host memory, lower bound -1.  These act on
 * the FP64 hierarchy; a mixed-precision saras_mg_solve works from the
 * caller's arrays and leaves its level 0 undefined. ---- */
saras_status saras_mg_set_level(saras_mg *mg, int level, const double *lhs_pad, const double *rhs_pad);
saras_status saras_mg_get_level(saras_mg *mg, int level, double *lhs_pad, double *rhs_pad);
/* multigrid_d3::smooth (poisson_d3.cc:96-246) */
saras_status saras_mg_smooth(saras_mg *mg, int level, int count, int zero_bc);
/* multigrid_d3::solve (poisson_d3.cc:249-432); *iters receives the count */
saras_status saras_mg_coarse_solve(saras_mg *mg, int level, int zero_bc, int *iters);
/* computeResidual (poisson_d3.cc:71-93) + coarsen (:435-513) + lhs(l+1)=0 */
saras_status saras_mg_residual_restrict(saras_mg *mg, int level);
/* prolong (poisson_d3.cc:516-560) + lhs += smd (poisson.cc:266) */
saras_status saras_mg_prolong_correct(saras_mg *mg, int level);
/* computeError (poisson_d3.cc:563-663) at level 0 */
saras_status saras_mg_error(saras_mg *mg, int res_type, double *value);

/* ---- measurement hooks (no counterpart in the reference): run work on the
 * handle's current device state between two HIP events on the handle's
 * stream and return the elapsed device time.  The state is whatever the last
 * solve or set_level left on the device. ---- */
/* `count` V-cycles of the handle's precision (FP32 cycles in MIXED mode), no
 * residual checks or host synchronisation between cycles. */
saras_status saras_mg_bench_cycles(saras_mg *mg, int count, double *ms);
/* `reps` level-0 smoothing sweeps of the handle's precision. */
saras_status saras_mg_bench_sweeps(saras_mg *mg, int reps, double *ms);
/* `reps` launches of a plain streaming copy of n doubles (device buffers
 * allocated and freed inside): the achievable-bandwidth reference and the
 * calibration kernel for HBM byte counters. */
saras_status saras_bench_stream_copy(size_t n, int reps, double *ms);

/* ---- Navier-Stokes / Boussinesq time step (lib/field, lib/timestep) ----
 *
 * One handle owns a run's fields (V, P, T and the pressure correction Pp),
 * the RK3 work arrays (the reference's function-static nseRHS/tempVF/tmpRHS/
 * tempSF, lsRK3_d3.cc:193-197) and the pressure-correction multigrid solver
 * (the time stepper's mgSolver member, timestep.h:142), all resident in
 * device memory.  Caller field arrays use the layout above with pad
 * grid->pad (>= the differentiation order: 1 for dscheme 1, 2 for 2). */

/* Field boundary kinds (lib/boundary/dirichlet.cc, neumann.cc).  Walls are
 * numbered 0 x-lo (Left), 1 x-hi (Right), 2 y-lo (Front), 3 y-hi (Back),
 * 4 z-lo (Bottom), 5 z-hi (Top); periodic axes ignore theirs.  Two kinds
 * serve the temperature of the plume problem (probType 9, scalar.cc:174-191):
 *   SARAS_FBC_HOT_PLATE  bottom wall only (3-D): hotPlate (hotPlate.cc:58-123),
 *                        T = 1 on the disc of radius plate_radius about
 *                        (Lx/2, Ly/2, 0), zero gradient elsewhere;
 *   SARAS_FBC_PLUME_OUT  top wall only: plumeOut (plumeOut.cc:53-90), ghost T
 *                        = T of the cell below where the ghost Vz > 0, else 0.
 *                        In planar runs the pressure correction's top wall then
 *                        turns Dirichlet, as in multigrid_d2
 *                        (poisson_d2.cc:641-647).
 * One kind serves the pressure of the inflow problem (probType 3,
 * hydro.cc:466-468):
 *   SARAS_FBC_NULL       nullBC (boundary.h:205-212): imposeBC leaves the
 *                        ghost layer as it stands, so the wall's first ghost
 *                        layer keeps the caller's initial P ghost plus every
 *                        pressure correction's ghost (P += Pp over the whole
 *                        array: lsRK3_d3.cc:289, sfield.cc:565-569).  Pressure
 *                        walls only. */
enum { SARAS_FBC_DIRICHLET = 0, SARAS_FBC_NEUMANN = 1, SARAS_FBC_HOT_PLATE = 2, SARAS_FBC_PLUME_OUT = 3,
       SARAS_FBC_NULL = 4 };

typedef struct {
    int dscheme;             /* "Differentiation Scheme": 1 (2nd order) or 2 (4th order) */
    int nlscheme;            /* "Non-Linear Term": 1 central, 2 hybrid upwind, 3 Morinishi */
    int has_T;               /* 1: scalar run, lsRK3_d3::timeAdvance(V, P, T) */
    double nu, kappa;        /* diffusivities (timestep.cc:70-94) */
    double dt;               /* "Time-Step" */
    double cn_tol;           /* "Solve Tolerance" of the Jacobi CN solves (absolute) */
    double buoy[3];          /* gTerms = Fb g/|g| of buoyantForce (buoyantForce.cc:46-61) */
    int vbc_type[3][6];      /* SARAS_FBC_* per velocity component and wall */
    double vbc_val[3][6];    /* Dirichlet value or Neumann gradient */
    int tbc_type[6];
    double tbc_val[6];
    int pbc_type[6];
    double pbc_val[6];
    int scheme;              /* "Integration Scheme": 0 lsRK3_d3 (RK3-CN), 1 eulerCN_d3 */
    double pe_limit;         /* "Peclet Limit" of the upwind scheme (2.0 in the reference deck) */
    double central_bias;     /* "Central Bias" of the upwind scheme (0.8 in the reference deck) */
    int les;                 /* "LES Model" 1: stretched spiral vortex sub-grid stress, added once
                                the run time exceeds 5 time steps (lsRK3_d3.cc:238); periodic boxes */
    double mean_pgrad;       /* "Mean Pressure Gradient" of force type 5, constantPGrad (force.h:148-153,
                                hydro.cc:268-271): added to the x-momentum right-hand side of the core.
                                A hydrodynamics forcing: must be 0 when has_T is set */
    double plate_radius;     /* "Plate Radius" of SARAS_FBC_HOT_PLATE (ignored otherwise) */
    int les_sf32;            /* 1: the sub-grid model's structure-function integral (the sfIntegral
                                sum of spiral.cc:157-206, most of its arithmetic) in single
                                precision, the rest of the model in double; 0: all double */
} saras_nse_params;

typedef struct {
    int jacobi_iters[4];     /* iteration count of the last Vx, Vy, Vz, T solves */
    saras_mg_stats mg;       /* last pressure-correction solve */
    double device_ms;        /* device time of the call (HIP events) */
} saras_nse_stats;

typedef struct saras_nse saras_nse;

/* The diffusivities and the buoyancy vector of a run from the deck's
 * "Problem Type", Re, Ra, Pr, "RBC Type" and gravity axis:
 * timestep::timestep (timestep.cc:70-98): prob_type 1..4 (hydrodynamics)
 * nu = 1/Re (kappa = 0, buoy = 0); prob_type 5..9 (with a scalar) by
 * rbc_type 1: (Pr, 1), 2 and 4: (sqrt(Pr/Ra), 1/sqrt(Pr Ra)), 3: (1, 1/Pr);
 * buoyantForce::buoyantForce (buoyantForce.cc:48-60): buoy = Fb g/|g| with
 * Fb = Ra Pr, 1, Ra, Pr for rbc_type 1..4.  SARAS_BAD_ARG where the
 * reference aborts (an rbc_type outside 1..4) or divides by zero. */
saras_status saras_nse_physics(int prob_type, double re, double ra, double pr, int rbc_type,
                               const double g_axis[3], double *nu, double *kappa, double buoy[3]);

/* The wall kinds and values of V, P and T that the drivers wire for a
 * "Problem Type" when "Use Custom BC" is off: hydro::initVBCs
 * (hydro.cc:349-434: lid-driven cavity 1 with Vx = 1 on the top wall,
 * inflow 3 with Vx = 1 on x-lo and zero-gradient outflow on x-hi,
 * free-slip 8, plume 9, no-slip otherwise), hydro::initPBCs
 * (hydro.cc:447-484: nullBC on x-lo for 3, zero gradient elsewhere) and
 * scalar::initTBCs (scalar.cc:163-201: by probType 5..9, the hot plate on
 * the bottom wall when non_hg is set in 3-D, plumeOut on the top for 9).
 * Only the type and value tables of *params are written.  SARAS_BAD_ARG
 * for a prob_type outside 1..9. */
saras_status saras_nse_wiring(int prob_type, int planar, int non_hg, saras_nse_params *params);

/* lsRK3_d3::lsRK3_d3 + timestep::timestep (lsRK3_d3.cc:58-70,
 * timestep.cc:57-122) with the fields of hydro/scalar (vfield, sfield).
 * mg: the "Multigrid" keys of the pressure solve (bc is forced to Neumann,
 * poisson.cc:87).  grid->dim 2 is the PLANAR (x-z) build: n[1] = 1, caller
 * arrays (nx+2p, 1, nz+2p), either dscheme with any non-linear term
 * (lsRK3_d2.cc, eulerCN_d2.cc).  Morinishi runs with dscheme 2 and a wall in z
 * whose z edge test (on core.ubound(1), vfield.cc:329, 434) misses the last z
 * cell -- n_y != n_z, or planar -- read the second ghost layer above the top
 * wall, which keeps the values the caller's arrays hold at set_fields.
 * SARAS_BAD_SHAPE for extents below 4; SARAS_BAD_ARG for pad < dscheme. */
saras_status saras_nse_create(const saras_grid_desc *grid, const saras_mg_params *mg,
                              const saras_nse_params *params, saras_nse **out);
saras_status saras_nse_destroy(saras_nse *s);

/* Device bytes a time-step handle of this grid and parameter set holds
 * (fields, stage arrays, its pressure solver, staging); 0 for an invalid grid.
 * No counterpart in the reference: the request loop admits by it. */
size_t saras_nse_footprint(const saras_grid_desc *grid, const saras_mg_params *mg, const saras_nse_params *params);

/* Set the run time (timestep's solTime, the driver's `time`, which
 * reader::readRestart restores from a restart file, reader.cc:171-226);
 * the LES model enters once it exceeds 5 time steps (lsRK3_d3.cc:238). */
saras_status saras_nse_set_time(saras_nse *s, double time);

/* saras_mg_set_graphs for the handle's pressure solver (no counterpart in
 * the reference). */
saras_status saras_nse_set_graphs(saras_nse *s, int enable);

/* Load fields (NULL keeps the current one; only the core is read) and
 * impose the boundary conditions (the solvers' initial imposeBCs). */
saras_status saras_nse_set_fields(saras_nse *s, const double *vx, const double *vy, const double *vz,
                                  const double *p, const double *t, const double *pp, int ptr_kind);
/* Store fields: the core and every ghost cell with one index outside the
 * core that the reference's imposeBCs/syncFaces define (periodic pads, first
 * wall layer); edge and corner pads and the second wall layer are left
 * untouched. */
saras_status saras_nse_get_fields(saras_nse *s, double *vx, double *vy, double *vz, double *p, double *t,
                                  double *pp, int ptr_kind);

/* lsRK3_d3::timeAdvance(V, P[, T]) (lsRK3_d3.cc:94-177, 190-305): one full
 * time step of three sub-stages; with scheme 1 eulerCN_d3::timeAdvance
 * (eulerCN_d3.cc:88-165, 178-270), one stage.  SARAS_JACOBI_DIVERGED where the reference
 * aborts (predictors.cc:133-141); a pressure solve that runs out of V-cycles
 * is not an error (the reference is silent). */
saras_status saras_nse_step(saras_nse *s, saras_nse_stats *stats);
/* One sub-stage (rkLev = stage in 0..2) of the same loop (Euler-CN: stage 0). */
saras_status saras_nse_stage(saras_nse *s, int stage, saras_nse_stats *stats);
/* The explicit part of a sub-stage alone (lsRK3_d3.cc:199-267): fills the
 * stage RHS and the new non-linear terms; read them with get_rhs (core). */
saras_status saras_nse_rhs(saras_nse *s, int stage);
saras_status saras_nse_get_rhs(saras_nse *s, double *rx, double *ry, double *rz, double *rt, double *hx,
                               double *hy, double *hz, double *ht, int ptr_kind);
/* timestep::solveVx/solveVy/solveVz/solveT (predictors.cc:59-145, 161-219,
 * 235-321, 337-423): Jacobi CN solve of field 0..3 (Vx, Vy, Vz, T) against
 * the current stage RHS; *iters gets the reference's iteration count. */
saras_status saras_nse_solve_field(saras_nse *s, int field, double beta, int *iters);

/* Single field operators over the core of caller arrays whose ghost cells
 * are used as supplied: op 0-2 derivative::calcDerivative1_x/y/z, 3-5
 * calcDerivative2xx/yy/zz (out = value; derivative.cc:104-306), 6
 * sfield::computeDiff (out += lap F; sfield.cc:133-147), 7
 * sfield::computeNLin with the handle's velocity (out += H;
 * sfield.cc:160-311). */
saras_status saras_field_op(saras_nse *s, int op, const double *f, double *out, int ptr_kind);

/* The vector field operators over the core of caller arrays whose ghost
 * cells are used as supplied (layout as above, pad grid->pad):
 *   op 0 vfield::divergence (vfield.cc:721-737): ox = dfx/dx + dfy/dy + dfz/dz
 *   op 1 sfield::gradient (sfield.cc:496-508): (ox, oy, oz) = grad fx
 *        (fy, fz unused)
 *   op 2 vfield::computeDiff (vfield.cc:150-192): o_c += lap f_c
 *   op 3 vfield::computeNLin(V, H) with V = (fx, fy, fz) advecting itself
 *        (vfield.cc:205-689; the handle's non-linear scheme, nu for the
 *        upwind Peclet number): o_c += H_c, i.e. o_c -= (V . grad) f_c
 * Ops 0 and 1 assign the core, 2 and 3 accumulate into it, as the
 * reference's methods do.  3-D handles. */
saras_status saras_field_vec_op(saras_nse *s, int op, const double *fx, const double *fy, const double *fz,
                                double *ox, double *oy, double *oz, int ptr_kind);

/* Sub-grid diagnostics of the last LES evaluation (spiral.cc:186-206):
 * out[0] sum |K| dV, out[1] sum of the SGS dissipation dV, out[2] sum nu_t dV
 * (tseries divides each by the volume); *time = the run time of the handle
 * (advanced by dt after each saras_nse_step). */
saras_status saras_nse_sgs_diag(saras_nse *s, double out[3], double *time);

/* vfield::computeTStp (vfield.cc:691-706) with the driver's cap
 * (hydro_d3.cc:221-226): dt = min(courant / max|u|/dx + |v|/dy + |w|/dz, dt_max)
 * over the core, dx = L/n; the handle's time step becomes *dt. */
saras_status saras_nse_cfl(saras_nse *s, double courant, double dt_max, double *dt);

/* tseries::writeTSData diagnostics (tseries.cc:209-290): out[0] volume-mean
 * kinetic energy, out[1] max |div V|, out[2] NuVol, out[3] NuWall
 * (z gravity). */
saras_status saras_nse_diag(saras_nse *s, double out[4]);

/* The quantities of a TimeSeries.dat line (tseries::writeTSData,
 * tseries.cc:209-290 hydro, 295-490 scalar, z gravity): out[0] volume-mean
 * kinetic energy, out[1] volume-mean thermal energy 1/2 (T + z - 1)^2
 * (0 without a scalar), out[2] max |div V|, out[3] NuVol, out[4] NuWall.
 * Re = sqrt(2 KE) / nu and U_rms = sqrt(2 KE) are formed by the writer. */
saras_status saras_nse_tsdata(saras_nse *s, double out[5]);

/* probes::getData (probes.cc:185-205): the values of Vx, Vy, Vz, P, T at
 * `count` core cells ijk[3q..3q+2] (0-based core indices) into
 * out[5q..5q+4]; T is 0 without a scalar.  BAD_ARG for a point outside the
 * core. */
saras_status saras_nse_probe(saras_nse *s, int count, const int *ijk, double *out);

/* The same values as saras_nse_probe for at most SARAS_PROBE_NOALLOC_MAX
 * points, with no device allocation: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight time loops, harnesses that forbid hipMalloc); saras_nse_probe
 * allocates and frees its staging on every call.  The point indices and
 * values are staged in a fixed device array of the library (about 416 KB per
 * GPU, held until the process ends).  More points: SARAS_BAD_ARG.  Every call
 * on one GPU shares that array; the library serialises its calls in one
 * process with a lock, and a call returns only after its values are copied
 * out, so the array's previous contents never matter.  Calls that overlapped
 * on one GPU without that lock would race. */
#define SARAS_PROBE_NOALLOC_MAX 8192
saras_status saras_nse_probe_noalloc(saras_nse *s, int count, const int *ijk, double *out);

/* Measurement hook: `count` full steps with no host work between them
 * beyond the reference's own convergence checks; *ms = device time. */
saras_status saras_nse_bench_steps(saras_nse *s, int count, double *ms);

/* Measurement hook of the sub-grid stress kernel alone (the per-cell part
 * of spiral::computeSG, spiral.cc:157-206 with sgsStress :472-548): `count`
 * launches on the current velocity, *ms = mean device time per launch.
 * Writes the handle's stress arrays as a step would.  SARAS_BAD_ARG for a
 * handle without the model. */
saras_status saras_nse_bench_sgs(saras_nse *s, int count, double *ms);

/* ---- serving (no counterpart in the reference: these serve many solves
 * per process, the batched request loop the reference does not have). ---- */

/* Replay the handle's V-cycles from captured hipGraphs (enable = 1) or
 * launch them one kernel at a time (0, the default).  Results are identical;
 * graphs remove per-launch host overhead for small grids. */
saras_status saras_mg_set_graphs(saras_mg *mg, int enable);

/* Device bytes a handle of this grid and parameter set holds (hierarchies,
 * staging, workspaces); 0 for an invalid grid. */
size_t saras_mg_footprint(const saras_grid_desc *grid, const saras_mg_params *params);

typedef struct {
    size_t mem_budget;   /* device bytes the server may hold in solver handles */
    int slots;           /* solves in flight at once (one worker thread and stream each) */
    int use_graphs;      /* replay V-cycles from hipGraphs */
    int keep_resident;   /* keep idle hierarchies for reuse (LRU, within the budget) */
    int largest_first;   /* 1: a free slot takes the waiting request of largest footprint (shortens
                            the makespan of a burst); 0: arrival order */
} saras_server_params;

/* One Poisson solve: poisson::mgSolve on caller host arrays (layout as
 * above, pad grid.pad).  b and x must stay valid until the completion with
 * the same tag is returned; x is in/out like outLHS. */
typedef struct {
    saras_grid_desc grid;
    saras_mg_params mg;
    const double *b;
    double *x;
    long long tag;
} saras_request;

typedef struct {
    long long tag;
    int status;            /* saras_status of the solve */
    saras_mg_stats stats;
    double t_submit;       /* seconds on the server clock */
    double t_admit;
    double t_done;
    char message[120];     /* error text when status is not OK / NOT_CONVERGED */
} saras_completion;

typedef struct saras_server saras_server;

/* A short Navier-Stokes run: `steps` calls of saras_nse_step on a handle of
 * (grid, mg, nse) from the caller's fields (host arrays, layout as above,
 * pad grid.pad; t may be NULL when nse.has_T is 0), with the run time and
 * the pressure correction starting at zero, as a fresh lsRK3_d3/eulerCN_d3
 * run does.  On completion vx, vy, vz, p (and t) hold the state after the
 * last step as saras_nse_get_fields writes it.  The arrays must stay valid
 * until the completion with the same tag is returned; its stats are those of
 * the last pressure solve, and its status is that of the first failing step
 * (SARAS_JACOBI_DIVERGED, ...) or SARAS_OK. */
typedef struct {
    saras_grid_desc grid;
    saras_mg_params mg;
    saras_nse_params nse;
    int steps;
    double *vx, *vy, *vz, *p, *t;
    long long tag;
} saras_nse_request;

saras_status saras_server_create(const saras_server_params *params, saras_server **out);
/* Queue a request (returns at once).  Requests are admitted in arrival
 * order as slots and device memory allow; a request that waits for memory
 * lets later, smaller ones pass. */
saras_status saras_server_submit(saras_server *s, const saras_request *req);
/* Queue a short Navier-Stokes run; admitted by its footprint like a solve. */
saras_status saras_server_submit_nse(saras_server *s, const saras_nse_request *req);
/* Move up to max finished requests into out; waits up to wait_ms for the
 * first one.  Returns the number moved (negative on a bad argument). */
int saras_server_poll(saras_server *s, saras_completion *out, int max, int wait_ms);
/* Block until every submitted request has finished. */
saras_status saras_server_drain(saras_server *s);
/* Seconds since the server was created (its clock for t_* fields). */
double saras_server_clock(const saras_server *s);
/* Device bytes currently held by resident hierarchies and running solves. */
size_t saras_server_mem_used(const saras_server *s);
saras_status saras_server_destroy(saras_server *s);

/* ---- device memory helpers, so callers without a GPU framework can hand
 * the solver device pointers. ---- */
saras_status saras_device_alloc(void **ptr, size_t bytes);
saras_status saras_device_free(void *ptr);
saras_status saras_memcpy_h2d(void *dst, const void *src, size_t bytes);
saras_status saras_memcpy_d2h(void *dst, const void *src, size_t bytes);
saras_status saras_device_sync(void);
saras_status saras_device_zero(void *ptr, size_t bytes);   /* synchronous */

/* Text of the last error raised on the calling thread ("" if none). */
const char *saras_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
