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
    SARAS_BAD_ARG = 6          /* invalid parameter or null pointer */
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
    int inner_cycles;   /* MIXED: FP32 V-cycles per FP64 refinement step */
    int max_refine;     /* MIXED: cap on refinement steps */
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
 * hipStream_t or NULL for the handle's own stream.  Returns SARAS_OK when
 * the residual fell below vc_tol, SARAS_NOT_CONVERGED when vc_count cycles
 * ran out (the reference is silent there, poisson.cc:127-185). */
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
 * padded (n_l+2)^3 FP64 arrays in host memory, lower bound -1. ---- */
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

/* ---- device memory helpers, so callers without a GPU framework can hand
 * the solver device pointers. ---- */
saras_status saras_device_alloc(void **ptr, size_t bytes);
saras_status saras_device_free(void *ptr);
saras_status saras_memcpy_h2d(void *dst, const void *src, size_t bytes);
saras_status saras_memcpy_d2h(void *dst, const void *src, size_t bytes);
saras_status saras_device_sync(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char *saras_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
