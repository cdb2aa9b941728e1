// mg_kernels.hpp -- HIP kernels of the geometric-multigrid V-cycle (gfx950).
//
// Storage: every level keeps only its core cells, row-major with z fastest,
// rows padded to a 128-byte pitch (sy elements), sx = sy * ny.  There is no
// ghost layer on the device: boundary neighbours are produced on the fly.
//
// Why that reproduces the reference's ghost timing (poisson_d3.cc:96-246):
//  * a wall ghost x(-1) = f(x(0)) is only ever read by the cell x(0) itself,
//    and the reference refreshes it at the start of each sweep, before that
//    cell's single update of the sweep -> f(current own value) is identical;
//  * a periodic ghost is refreshed before each colour pass (updateFace
//    between red and black).  For an even extent the wrapped neighbour has
//    the other colour, so it does not change during the pass and reading the
//    opposite core cell directly is identical.  Odd periodic extents can only
//    occur on the coarsest level, which the single-workgroup kernel handles
//    with literal ghost arrays.
//
// Row layouts: a level stores each row (i, j) either z-contiguous (element k
// at offset k) or colour-split (hz > 0: even k at k/2, odd k at hz + k/2).
// In the split layout one colour of a row is one contiguous half-row, so a
// red or black pass streams whole vectors (lib/poisson has no such layout;
// the arithmetic per cell is unchanged).
//
// Every expression keeps the reference's operand order; the file is compiled
// with -ffp-contract=off so the FP64 path matches the FP64 CPU build bit for
// bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace saras {

enum BcCode : int { BC_PER = 0, BC_NEU = 1, BC_DIR = 2, BC_DIRV = 3 };

constexpr int kXcds = 8;   // MI355X: 8 XCDs, workgroups dealt round-robin

template <typename T>
struct LevelArgs {
    int n[3];          // cells (n[1] == 1 for planar problems)
    long sx, sy;       // element strides of i and j
    long hz;           // 0: z-contiguous rows; > 0: colour-split rows (odd k at hz + k/2)
    const T *c2[3];    // xi_x^2 / h^2 per core index (poisson.cc:453-483 x copyDerivs)
    const T *c1[3];    // xi_xx / (2 h)
    const T *c2zs, *c1zs;  // split rows: c2[2], c1[2] reordered as [h * nz/2 + m] = c[2m + h]
    int bc_lo[3], bc_hi[3];
    // inhomogeneous wall values (BC_DIRV), padded face arrays with a zero
    // shell: x-walls [(j+1)(nz+2) + k+1], y-walls [(i+1)(nz+2) + k+1],
    // z-walls [(i+1)(ny+2) + j+1]; planar: x-walls [k+1], z-walls [i+1]
    const T *wall[3];
    // split colour passes: 0 for rows in (i, j) order; otherwise rows run in
    // tiles of rtile j-rows (a power of two dividing n[1]), all i of a tile
    // before the next tile, so that the x-neighbour rows a pass reads again
    // two i-planes later are still in L2 on large levels
    int rtile;
};

// offset of element k inside its row
template <typename T>
__device__ __forceinline__ long kof(const LevelArgs<T> &L, int k) {
    return L.hz ? (long)(k & 1) * L.hz + (k >> 1) : (long)k;
}

// t = q * n + rem with 0 <= rem < n: one 32-bit unsigned division whenever t
// fits 32 bits (always, below a few billion cells), the 64-bit sequence
// otherwise -- a 64-bit division is several times the instructions
__device__ __forceinline__ int split_index(long t, int n, long &q) {
    if ((unsigned long)t <= 0xffffffffUL) {
        const unsigned ut = (unsigned)t, un = (unsigned)n, uq = ut / un;
        q = (long)uq;
        return (int)(ut - uq * un);
    }
    q = t / n;
    return (int)(t - q * n);
}

// cell (i, j, k) of the row-major index t over an (., ny, nz) box
__device__ __forceinline__ void cell_of(long t, int ny, int nz, int &i, int &j, int &k) {
    long r, q;
    k = split_index(t, nz, r);
    j = split_index(r, ny, q);
    i = (int)q;
}

// ---------------------------------------------------------------------------
// neighbour values with the boundary rule applied
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T bc_value(int bc, T own, T wrap, T wv) {
    switch (bc) {
        case BC_PER: return wrap;
        case BC_NEU: return own;
        case BC_DIR: return -own;
        default: return T(2.0) * wv - own;
    }
}

template <typename T, int DIM>
struct Stencil {
    T c, xE, xW, yN, yS, zT, zB;
};

template <typename T, int DIM>
__device__ __forceinline__ void gather(const LevelArgs<T> &L, const T *__restrict__ x, int i, int j, int k,
                                       long idx, Stencil<T, DIM> &s) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    s.c = x[idx];
    // x axis
    if (i + 1 < nx) s.xE = x[idx + L.sx];
    else {
        T wv = T(0);
        if (L.bc_hi[0] == BC_DIRV) wv = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + k + 1] : L.wall[0][k + 1];
        T wrap = L.bc_hi[0] == BC_PER ? x[idx - (long)(nx - 1) * L.sx] : T(0);
        s.xE = bc_value(L.bc_hi[0], s.c, wrap, wv);
    }
    if (i > 0) s.xW = x[idx - L.sx];
    else {
        T wv = T(0);
        if (L.bc_lo[0] == BC_DIRV) wv = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + k + 1] : L.wall[0][k + 1];
        T wrap = L.bc_lo[0] == BC_PER ? x[idx + (long)(nx - 1) * L.sx] : T(0);
        s.xW = bc_value(L.bc_lo[0], s.c, wrap, wv);
    }
    if (DIM == 3) {
        if (j + 1 < ny) s.yN = x[idx + L.sy];
        else {
            T wv = L.bc_hi[1] == BC_DIRV ? L.wall[1][(i + 1) * (nz + 2) + k + 1] : T(0);
            T wrap = L.bc_hi[1] == BC_PER ? x[idx - (long)(ny - 1) * L.sy] : T(0);
            s.yN = bc_value(L.bc_hi[1], s.c, wrap, wv);
        }
        if (j > 0) s.yS = x[idx - L.sy];
        else {
            T wv = L.bc_lo[1] == BC_DIRV ? L.wall[1][(i + 1) * (nz + 2) + k + 1] : T(0);
            T wrap = L.bc_lo[1] == BC_PER ? x[idx + (long)(ny - 1) * L.sy] : T(0);
            s.yS = bc_value(L.bc_lo[1], s.c, wrap, wv);
        }
    }
    const long row = idx - kof(L, k);
    if (k + 1 < nz) s.zT = x[row + kof(L, k + 1)];
    else {
        T wv = T(0);
        if (L.bc_hi[2] == BC_DIRV) wv = DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1];
        T wrap = L.bc_hi[2] == BC_PER ? x[row + kof(L, 0)] : T(0);
        s.zT = bc_value(L.bc_hi[2], s.c, wrap, wv);
    }
    if (k > 0) s.zB = x[row + kof(L, k - 1)];
    else {
        T wv = T(0);
        if (L.bc_lo[2] == BC_DIRV) wv = DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1];
        T wrap = L.bc_lo[2] == BC_PER ? x[row + kof(L, nz - 1)] : T(0);
        s.zB = bc_value(L.bc_lo[2], s.c, wrap, wv);
    }
}

// Division of the SOR update: the FP64 path divides exactly as the reference
// does; the FP32 smoothing of the mixed-precision solver, whose result is
// corrected by the FP64 refinement anyway, uses the hardware reciprocal.
__device__ __forceinline__ double sor_div(double a, double d) { return a / d; }
__device__ __forceinline__ float sor_div(float a, float d) { return a * __builtin_amdgcn_rcpf(d); }

// RB-SOR point update (poisson_d3.cc:109-116, poisson_d2.cc:103-108):
// (1 - w) x + sor_quot, the quotient from the neighbours alone
template <typename T, int DIM>
__device__ __forceinline__ T sor_quot(const Stencil<T, DIM> &s, T c2x, T c1x, T c2y, T c1y, T c2z, T c1z, T b,
                                      T w) {
    T sum = c2x * (s.xE + s.xW) + c1x * (s.xE - s.xW);
    T den;
    if (DIM == 3) {
        sum = sum + c2y * (s.yN + s.yS);
        sum = sum + c1y * (s.yN - s.yS);
        sum = sum + c2z * (s.zT + s.zB);
        sum = sum + c1z * (s.zT - s.zB);
        den = T(2.0) * ((c2x + c2y) + c2z);
    } else {
        sum = sum + c2z * (s.zT + s.zB);
        sum = sum + c1z * (s.zT - s.zB);
        den = T(2.0) * (c2x + c2z);
    }
    return sor_div((sum - b) * w, den);
}

template <typename T, int DIM>
__device__ __forceinline__ T sor_update(const Stencil<T, DIM> &s, T c2x, T c1x, T c2y, T c1y, T c2z, T c1z,
                                        T b, T w) {
    return (T(1.0) - w) * s.c + sor_quot<T, DIM>(s, c2x, c1x, c2y, c1y, c2z, c1z, b, w);
}

// residual b - A x (poisson_d3.cc:81-87, poisson_d2.cc:80-84)
template <typename T, int DIM>
__device__ __forceinline__ T residual(const Stencil<T, DIM> &s, T c2x, T c1x, T c2y, T c1y, T c2z, T c1z, T b) {
    T two_c = T(2.0) * s.c;
    T a = c2x * ((s.xE - two_c) + s.xW) + c1x * (s.xE - s.xW);
    if (DIM == 3) {
        a = a + c2y * ((s.yN - two_c) + s.yS);
        a = a + c1y * (s.yN - s.yS);
    }
    a = a + c2z * ((s.zT - two_c) + s.zB);
    a = a + c1z * (s.zT - s.zB);
    return b - a;
}

// ---------------------------------------------------------------------------
// one colour pass of red-black SOR, in place.  colour 0 = red (i+j+k even).
// Thread t handles the pair of z cells (2m, 2m+1) and updates the one of the
// requested colour.
// ---------------------------------------------------------------------------
// FINE only separates the symbol of the finest-level launches in profiles.
template <typename T, int DIM, bool FINE>
__global__ void __launch_bounds__(256) k_rb_pass(LevelArgs<T> L, T *__restrict__ x, const T *__restrict__ b,
                                                 int colour, T w) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int npair = (nz + 1) >> 1;
    const long total = (long)nx * ny * npair;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j, m;
    cell_of(t, ny, npair, i, j, m);
    const int k = 2 * m + ((i + j + colour) & 1);
    if (k >= nz) return;
    const long idx = (long)i * L.sx + (long)j * L.sy + kof(L, k);
    Stencil<T, DIM> s;
    gather<T, DIM>(L, x, i, j, k, idx, s);
    const T c2y = DIM == 3 ? L.c2[1][j] : T(0), c1y = DIM == 3 ? L.c1[1][j] : T(0);
    x[idx] = sor_update<T, DIM>(s, L.c2[0][i], L.c1[0][i], c2y, c1y, L.c2[2][k], L.c1[2][k], b[idx], w);
}

// Per-axis restriction factors of one coarse index I (poisson_d3.cc:494-502):
// lo = X_I - x_2I, hi = x_2I+1 - X_I, wd = x_2I+1 - x_2I.
template <typename T>
struct RestrictArgs {
    const T *lo[3], *hi[3], *wd[3];
};

// ---------------------------------------------------------------------------
// residual at level l, restricted to level l+1 (computeResidual + coarsen,
// poisson_d3.cc:71-93, 484-511; 2-D plain average poisson_d2.cc:288-297), and
// lhs(l+1) = 0 (poisson.cc:245).  One thread per coarse cell.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void __launch_bounds__(256) k_resid_restrict(LevelArgs<T> F, const T *__restrict__ x,
                                                        const T *__restrict__ b, LevelArgs<T> C,
                                                        T *__restrict__ bc, T *__restrict__ xc,
                                                        RestrictArgs<T> R) {
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    const long total = (long)cx * cy * cz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int I, J, K;
    cell_of(t, cy, cz, I, J, K);
    T res[2][2][2];
    const int jmax = DIM == 3 ? 2 : 1;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < jmax; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                const int i = 2 * I + di, j = DIM == 3 ? 2 * J + dj : 0, k = 2 * K + dk;
                const long idx = (long)i * F.sx + (long)j * F.sy + kof(F, k);
                Stencil<T, DIM> s;
                gather<T, DIM>(F, x, i, j, k, idx, s);
                const T c2y = DIM == 3 ? F.c2[1][j] : T(0), c1y = DIM == 3 ? F.c1[1][j] : T(0);
                res[di][dj][dk] = residual<T, DIM>(s, F.c2[0][i], F.c1[0][i], c2y, c1y, F.c2[2][k], F.c1[2][k], b[idx]);
            }
    T out;
    if (DIM == 3) {
        const T lx = R.lo[0][I], hx = R.hi[0][I], ly = R.lo[1][J], hy = R.hi[1][J], lz = R.lo[2][K], hz = R.hi[2][K];
        const T v000 = lx * ly * lz, v100 = hx * ly * lz, v010 = lx * hy * lz, v001 = lx * ly * hz;
        const T v011 = lx * hy * hz, v101 = hx * ly * hz, v110 = hx * hy * lz, v111 = hx * hy * hz;
        const T vtot = R.wd[0][I] * R.wd[1][J] * R.wd[2][K];
        out = (v111 * res[1][1][1] + v000 * res[0][0][0] + v110 * res[1][1][0] + v001 * res[0][0][1] +
               v101 * res[1][0][1] + v010 * res[0][1][0] + v011 * res[0][1][1] + v100 * res[1][0][0]) / vtot;
    } else {
        out = (res[1][0][1] + res[0][0][0] + res[1][0][0] + res[0][0][1]) / T(4);
    }
    const long cidx = (long)I * C.sx + (long)J * C.sy + kof(C, K);
    bc[cidx] = out;
    xc[cidx] = T(0);
}

// ---------------------------------------------------------------------------
// prolongation by injection plus the saved pre-restriction iterate
// (poisson_d3.cc:548-558 then lhs += smd, poisson.cc:266).  The fine iterate
// is left untouched while the coarse levels work, so smd is x itself.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_prolong_correct(int nx, int ny, int nz, long sx, long sy, long hz,
                                                         T *__restrict__ x, const T *__restrict__ e, long csx,
                                                         long csy, long chz) {
    const long total = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    cell_of(t, ny, nz, i, j, k);
    const long idx = (long)i * sx + (long)j * sy + (hz ? (long)(k & 1) * hz + (k >> 1) : (long)k);
    const int kc = k >> 1;
    const T ev = e[(long)(i >> 1) * csx + (long)(j >> 1) * csy + (chz ? (long)(kc & 1) * chz + (kc >> 1) : (long)kc)];
    x[idx] = ev + x[idx];
}

// ---------------------------------------------------------------------------
// level-0 residual norms (computeError, poisson_d3.cc:563-663): per-block
// partials of [max|r|, max|b|] or [sum|r|, sum|b|] or [sum r^2, sum b^2],
// reduced in a fixed order so results are reproducible.
// ---------------------------------------------------------------------------
constexpr int kNormBlocks = 1024;
constexpr int kNormThreads = 256;
static_assert(kNormBlocks % kXcds == 0, "norm grids are dealt to XCDs in equal shares");

template <int TYPE>
__device__ __forceinline__ double norm_acc(double a, double v) {
    if (TYPE == 0) return fmax(a, fabs(v));
    if (TYPE == 1) return a + fabs(v);
    return a + v * v;
}

template <int TYPE>
__device__ __forceinline__ double norm_comb(double a, double b) {
    return TYPE == 0 ? fmax(a, b) : a + b;
}

// block reduction of two values: 64-lane shuffles, then the waves in order
template <int TYPE>
__device__ void block_reduce2(double &a, double &b) {
    __shared__ double sa[kNormThreads / 64], sb[kNormThreads / 64];
    for (int o = 32; o > 0; o >>= 1) {
        a = norm_comb<TYPE>(a, __shfl_xor(a, o, 64));
        b = norm_comb<TYPE>(b, __shfl_xor(b, o, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        sa[wv] = a;
        sb[wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            a = norm_comb<TYPE>(a, sa[q]);
            b = norm_comb<TYPE>(b, sb[q]);
        }
    }
}

// Row-parallel traversal of a level's core: lanes run along z, each pass of
// a block covers `rpp` rows (i, j); rows are dealt to blocks round-robin.
struct RowWalk {
    int kspan, rpp, sub, k0;
    __device__ RowWalk(int nz) {
        kspan = min((int)blockDim.x, (nz + 63) / 64 * 64);
        rpp = blockDim.x / kspan;
        sub = threadIdx.x / kspan;
        k0 = threadIdx.x - sub * kspan;
    }
};

// residual norm partials of computeError (poisson_d3.cc:563-663)
template <int TYPE, int DIM>
__global__ void __launch_bounds__(kNormThreads) k_norm_partial(LevelArgs<double> L, const double *__restrict__ x,
                                                               const double *__restrict__ b,
                                                               double *__restrict__ partial) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const long rows = (long)nx * ny;
    const RowWalk w(nz);
    double ar = 0.0, ab = 0.0;
    if (w.sub < w.rpp) {
        for (long r = (long)blockIdx.x * w.rpp + w.sub; r < rows; r += (long)gridDim.x * w.rpp) {
            const int i = (int)(r / ny);
            const int j = (int)(r - (long)i * ny);
            const double c2x = L.c2[0][i], c1x = L.c1[0][i];
            const double c2y = DIM == 3 ? L.c2[1][j] : 0.0, c1y = DIM == 3 ? L.c1[1][j] : 0.0;
            for (int k = w.k0; k < nz; k += w.kspan) {
                const long idx = (long)i * L.sx + (long)j * L.sy + kof(L, k);
                Stencil<double, DIM> s;
                gather<double, DIM>(L, x, i, j, k, idx, s);
                const double bv = b[idx];
                const double rv = residual<double, DIM>(s, c2x, c1x, c2y, c1y, L.c2[2][k], L.c1[2][k], bv);
                ar = norm_acc<TYPE>(ar, rv);
                ab = norm_acc<TYPE>(ab, bv);
            }
        }
    }
    block_reduce2<TYPE>(ar, ab);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ar;
        partial[2 * blockIdx.x + 1] = ab;
    }
}

// out and seq are pinned host-coherent words: the host polls seq for the
// ticket, which is stored after the two values with release ordering.
template <int TYPE>
__global__ void __launch_bounds__(kNormThreads) k_norm_final(const double *__restrict__ partial, int nblocks,
                                                             double *__restrict__ out, unsigned *seq,
                                                             unsigned ticket) {
    double ar = 0.0, ab = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += blockDim.x) {
        ar = norm_comb<TYPE>(ar, partial[2 * q]);
        ab = norm_comb<TYPE>(ab, partial[2 * q + 1]);
    }
    block_reduce2<TYPE>(ar, ab);
    if (threadIdx.x == 0) {
        out[0] = ar;
        out[1] = ab;
        __hip_atomic_store(seq, ticket, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// coarsest level: one workgroup runs the reference's literal ghost-array
// RB-GS loop (multigrid_d3::solve poisson_d3.cc:249-432, or smooth() when
// "Solve Coarsest" is false), in LDS when it fits, else in a global scratch
// array.  The ghost layer is real here because the coarse residual is taken
// with the ghosts as they stand after the sweep (:398-415), and odd periodic
// extents need the refreshed-between-colours snapshot.
// ---------------------------------------------------------------------------
struct CoarseCtl {
    int mode;          // 0 = solve to tolerance, 1 = fixed number of sweeps
    int sweeps;        // mode 1
    int max_count;     // poisson.cc:431-436
    double tol;        // mgTolerance
    int *iters_out;    // iterations (mode 0)
    // FP32 correction cycles of the mixed solver: the solve also stops once
    // max|r| < rel_floor * max|b| of the level, the FP32 accuracy floor of
    // the current correction equation (an absolute mgTolerance below it would
    // only run the solve to the iteration cap); 0 in FP64, where mgTolerance
    // alone applies, as in the reference
    double rel_floor;
};

// Phases of the ghost-array loops below run either on the whole workgroup
// (separated by barriers) or, for levels of a few hundred cells, on wave 0
// alone: LDS operations of one wavefront execute in issue order, so a
// compiler-level fence is all that separates its phases.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WAVE>
__device__ __forceinline__ void lsync() {
    if (WAVE) wave_lds_sync();
    else __syncthreads();
}
template <bool WAVE>
__device__ __forceinline__ int lthreads() {
    return WAVE ? 64 : (int)blockDim.x;
}

// max over the workgroup (WAVE: over wave 0) of one value per thread, the
// same in every participating thread; `red` holds one slot per wave
template <bool WAVE>
__device__ __forceinline__ double group_max(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if (WAVE) return v;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmax(m, red[q]);
    return m;
}

template <typename T, int DIM>
__device__ __forceinline__ long cidx(int i, int j, int k, int PY, int PZ) {
    return DIM == 3 ? ((long)(i + 1) * PY + (j + 1)) * PZ + (k + 1) : (long)(i + 1) * PZ + (k + 1);
}

template <typename T>
__device__ __forceinline__ T wall_rule(int bc, T inner, T wv) {
    return bc == BC_NEU ? inner : (bc == BC_DIR ? -inner : T(2.0) * wv - inner);
}

// The face ghosts of an LDS level in one pass: periodic axes copy the
// opposite core plane and, with `walls`, the other axes take the wall rule of
// their own core cell (the updateFace + imposeBC pair of a sweep,
// poisson_d3.cc:876-993).  Only ghosts beside core cells are written -- the
// edges and corners of the shell are never read by the 7-point stencil -- so
// every write reads core cells alone and one barrier covers all faces.
template <typename T, int DIM, bool WAVE = false>
__device__ void coarse_ghosts(T *X, const LevelArgs<T> &L, const int *per, bool walls, int PY, int PZ) {
    const int nx = L.n[0], ny = DIM == 3 ? L.n[1] : 1, nz = L.n[2];
    const int tid = threadIdx.x, nt = lthreads<WAVE>();
    const int fx = (per[0] || walls) ? ny * nz : 0;
    const int fy = DIM == 3 && (per[1] || walls) ? nx * nz : 0;
    const int fz = (per[2] || walls) ? nx * ny : 0;
    for (int q = tid; q < fx + fy + fz; q += nt) {
        if (q < fx) {
            const unsigned uq = (unsigned)q, j = uq / (unsigned)nz;
            const int k = (int)(uq - j * (unsigned)nz), jj = (int)j;
            if (per[0]) {
                X[cidx<T, DIM>(nx, jj, k, PY, PZ)] = X[cidx<T, DIM>(0, jj, k, PY, PZ)];
                X[cidx<T, DIM>(-1, jj, k, PY, PZ)] = X[cidx<T, DIM>(nx - 1, jj, k, PY, PZ)];
            } else {
                T wv = T(0);
                if (L.bc_lo[0] == BC_DIRV) wv = DIM == 3 ? L.wall[0][(jj + 1) * (nz + 2) + k + 1] : L.wall[0][k + 1];
                X[cidx<T, DIM>(-1, jj, k, PY, PZ)] = wall_rule(L.bc_lo[0], X[cidx<T, DIM>(0, jj, k, PY, PZ)], wv);
                X[cidx<T, DIM>(nx, jj, k, PY, PZ)] = wall_rule(L.bc_hi[0], X[cidx<T, DIM>(nx - 1, jj, k, PY, PZ)], wv);
            }
        } else if (q < fx + fy) {
            const unsigned uq = (unsigned)(q - fx), i = uq / (unsigned)nz;
            const int k = (int)(uq - i * (unsigned)nz), ii = (int)i;
            if (per[1]) {
                X[cidx<T, DIM>(ii, ny, k, PY, PZ)] = X[cidx<T, DIM>(ii, 0, k, PY, PZ)];
                X[cidx<T, DIM>(ii, -1, k, PY, PZ)] = X[cidx<T, DIM>(ii, ny - 1, k, PY, PZ)];
            } else {
                T wv = L.bc_lo[1] == BC_DIRV ? L.wall[1][(ii + 1) * (nz + 2) + k + 1] : T(0);
                X[cidx<T, DIM>(ii, -1, k, PY, PZ)] = wall_rule(L.bc_lo[1], X[cidx<T, DIM>(ii, 0, k, PY, PZ)], wv);
                X[cidx<T, DIM>(ii, ny, k, PY, PZ)] = wall_rule(L.bc_hi[1], X[cidx<T, DIM>(ii, ny - 1, k, PY, PZ)], wv);
            }
        } else {
            const unsigned uq = (unsigned)(q - fx - fy), i = uq / (unsigned)ny;
            const int j = (int)(uq - i * (unsigned)ny), ii = (int)i;
            if (per[2]) {
                X[cidx<T, DIM>(ii, j, nz, PY, PZ)] = X[cidx<T, DIM>(ii, j, 0, PY, PZ)];
                X[cidx<T, DIM>(ii, j, -1, PY, PZ)] = X[cidx<T, DIM>(ii, j, nz - 1, PY, PZ)];
            } else {
                T wv = T(0);
                if (L.bc_lo[2] == BC_DIRV) wv = DIM == 3 ? L.wall[2][(ii + 1) * (ny + 2) + j + 1] : L.wall[2][ii + 1];
                X[cidx<T, DIM>(ii, j, -1, PY, PZ)] = wall_rule(L.bc_lo[2], X[cidx<T, DIM>(ii, j, 0, PY, PZ)], wv);
                X[cidx<T, DIM>(ii, j, nz, PY, PZ)] = wall_rule(L.bc_hi[2], X[cidx<T, DIM>(ii, j, nz - 1, PY, PZ)], wv);
            }
        }
    }
    lsync<WAVE>();
}

template <typename T, int DIM, bool WAVE = false>
__device__ void coarse_colour(T *X, const T *B, const LevelArgs<T> &L, int colour, T w, int PY, int PZ) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const long total = (long)nx * ny * nz;
    for (long t = threadIdx.x; t < total; t += lthreads<WAVE>()) {
        int i, j, k;
        cell_of(t, ny, nz, i, j, k);
        if (((i + j + k) & 1) != colour) continue;
        const long c = cidx<T, DIM>(i, j, k, PY, PZ);
        Stencil<T, DIM> s;
        s.c = X[c];
        s.xE = X[cidx<T, DIM>(i + 1, j, k, PY, PZ)];
        s.xW = X[cidx<T, DIM>(i - 1, j, k, PY, PZ)];
        if (DIM == 3) {
            s.yN = X[cidx<T, DIM>(i, j + 1, k, PY, PZ)];
            s.yS = X[cidx<T, DIM>(i, j - 1, k, PY, PZ)];
        }
        s.zT = X[c + 1];
        s.zB = X[c - 1];
        const T c2y = DIM == 3 ? L.c2[1][j] : T(0), c1y = DIM == 3 ? L.c1[1][j] : T(0);
        X[c] = sor_update<T, DIM>(s, L.c2[0][i], L.c1[0][i], c2y, c1y, L.c2[2][k], L.c1[2][k], B[t], w);
    }
    lsync<WAVE>();
}

template <typename T, int DIM>
__global__ void __launch_bounds__(1024) k_coarse(LevelArgs<T> L, T *__restrict__ x, const T *__restrict__ b,
                                                 T *__restrict__ scratch, int use_lds, T w, CoarseCtl ctl) {
    extern __shared__ unsigned char lds_raw[];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int PZ = nz + 2, PY = DIM == 3 ? ny + 2 : 1;
    const long padded = DIM == 3 ? (long)(nx + 2) * (ny + 2) * (nz + 2) : (long)(nx + 2) * (nz + 2);
    const long total = (long)nx * ny * nz;
    T *X = use_lds ? reinterpret_cast<T *>(lds_raw) : scratch;
    T *B = use_lds ? X + padded : scratch + padded;
    __shared__ double red[16];
    int per[3] = {L.bc_lo[0] == BC_PER, L.bc_lo[1] == BC_PER, L.bc_lo[2] == BC_PER};
    if (DIM == 2) per[1] = 0;

    for (long q = threadIdx.x; q < padded; q += blockDim.x) X[q] = T(0);
    __syncthreads();
    for (long t = threadIdx.x; t < total; t += blockDim.x) {
        int i, j, k;
        cell_of(t, ny, nz, i, j, k);
        const long g = (long)i * L.sx + (long)j * L.sy + kof(L, k);
        X[cidx<T, DIM>(i, j, k, PY, PZ)] = x[g];
        B[t] = b[g];
    }
    __syncthreads();

    int iter = 0;
    int sweeps = 0;
    double tol = ctl.tol;
    if (ctl.mode == 0 && ctl.rel_floor > 0.0) {
        double bm = 0.0;
        for (long t = threadIdx.x; t < total; t += blockDim.x) bm = fmax(bm, fabs((double)B[t]));
        tol = fmax(tol, ctl.rel_floor * group_max<false>(bm, red));
    }
    for (;;) {
        if (ctl.mode == 1 && sweeps >= ctl.sweeps) break;
        // imposeBC: updateFace then walls
        coarse_ghosts<T, DIM>(X, L, per, true, PY, PZ);
        coarse_colour<T, DIM>(X, B, L, 0, w, PY, PZ);
        coarse_ghosts<T, DIM>(X, L, per, false, PY, PZ);
        coarse_colour<T, DIM>(X, B, L, 1, w, PY, PZ);
        ++sweeps;
        if (ctl.mode == 1) continue;
        // max |r| with the ghosts as they stand (poisson_d3.cc:398-415)
        double lm = -1.0e-10;
        for (long t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            cell_of(t, ny, nz, i, j, k);
            const long c = cidx<T, DIM>(i, j, k, PY, PZ);
            Stencil<T, DIM> s;
            s.c = X[c];
            s.xE = X[cidx<T, DIM>(i + 1, j, k, PY, PZ)];
            s.xW = X[cidx<T, DIM>(i - 1, j, k, PY, PZ)];
            if (DIM == 3) {
                s.yN = X[cidx<T, DIM>(i, j + 1, k, PY, PZ)];
                s.yS = X[cidx<T, DIM>(i, j - 1, k, PY, PZ)];
            }
            s.zT = X[c + 1];
            s.zB = X[c - 1];
            const T c2y = DIM == 3 ? L.c2[1][j] : T(0), c1y = DIM == 3 ? L.c1[1][j] : T(0);
            const T rv = residual<T, DIM>(s, L.c2[0][i], L.c1[0][i], c2y, c1y, L.c2[2][k], L.c1[2][k], B[t]);
            const double a = fabs((double)rv);
            if (a > lm) lm = a;
        }
        const double mx = group_max<false>(lm, red);
        if (mx < tol) break;
        iter += 1;
        if (iter > ctl.max_count) break;
    }
    // final imposeBC (:431 / :245); the ghosts themselves are not kept
    for (long t = threadIdx.x; t < total; t += blockDim.x) {
        int i, j, k;
        cell_of(t, ny, nz, i, j, k);
        x[(long)i * L.sx + (long)j * L.sy + kof(L, k)] = X[cidx<T, DIM>(i, j, k, PY, PZ)];
    }
    if (threadIdx.x == 0 && ctl.mode == 0 && ctl.iters_out) *ctl.iters_out = iter;
}

// ---------------------------------------------------------------------------
// transfers between the caller's padded field arrays and the core layout
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_load_core(const double *__restrict__ full, int pad, int dim, int nx,
                                                   int ny, int nz, long sx, long sy, long hz, T *__restrict__ core) {
    const long total = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    cell_of(t, ny, nz, i, j, k);
    const long fz = nz + 2 * pad, fy = ny + 2 * pad;
    const long f = dim == 3 ? ((long)(i + pad) * fy + (j + pad)) * fz + (k + pad) : (long)(i + pad) * fz + (k + pad);
    core[(long)i * sx + (long)j * sy + (hz ? (long)(k & 1) * hz + (k >> 1) : (long)k)] = (T)full[f];
}

__global__ void __launch_bounds__(256) k_store_core(double *__restrict__ full, int pad, int dim, int nx, int ny,
                                                    int nz, long sx, long sy, long hz,
                                                    const double *__restrict__ core) {
    const long total = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    cell_of(t, ny, nz, i, j, k);
    const long fz = nz + 2 * pad, fy = ny + 2 * pad;
    const long f = dim == 3 ? ((long)(i + pad) * fy + (j + pad)) * fz + (k + pad) : (long)(i + pad) * fz + (k + pad);
    full[f] = core[(long)i * sx + (long)j * sy + (hz ? (long)(k & 1) * hz + (k >> 1) : (long)k)];
}

// One phase of the final imposeBC on the caller's array, over the [-1, n]
// box (poisson.cc:188 copies stagFull(0)).  op 0: periodic wrap of `axis`;
// op 1: wall rule of `axis`.  Phases run in the reference's order: wraps
// x, y, z (updateFace), then walls x, y, z (imposeBC).
__global__ void __launch_bounds__(256) k_ghost_phase(double *__restrict__ full, int pad, int dim, int nx, int ny,
                                                     int nz, int axis, int op, int bc_lo, int bc_hi,
                                                     const double *__restrict__ wall) {
    const int ext[3] = {nx + 2, dim == 3 ? ny + 2 : 1, nz + 2};
    const int n[3] = {nx, ny, nz};
    int o1, o2;  // the two other axes
    if (axis == 0) { o1 = 1; o2 = 2; }
    else if (axis == 1) { o1 = 0; o2 = 2; }
    else { o1 = 0; o2 = 1; }
    const long total = (long)ext[o1] * ext[o2];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int c[3];
    c[o1] = (int)(t / ext[o2]) - 1;
    c[o2] = (int)(t % ext[o2]) - 1;
    if (dim == 2) c[1] = 0;
    const long fz = nz + 2 * pad, fy = ny + 2 * pad;
    auto at = [&](int i, int j, int k) -> long {
        return dim == 3 ? ((long)(i + pad) * fy + (j + pad)) * fz + (k + pad) : (long)(i + pad) * fz + (k + pad);
    };
    int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
    int lsrc[3] = {c[0], c[1], c[2]}, hsrc[3] = {c[0], c[1], c[2]};
    lo[axis] = -1;
    hi[axis] = n[axis];
    if (op == 0) {
        lsrc[axis] = n[axis] - 1;
        hsrc[axis] = 0;
        const double vlo = full[at(lsrc[0], lsrc[1], lsrc[2])];
        const double vhi = full[at(hsrc[0], hsrc[1], hsrc[2])];
        full[at(hi[0], hi[1], hi[2])] = vhi;
        full[at(lo[0], lo[1], lo[2])] = vlo;
    } else {
        lsrc[axis] = 0;
        hsrc[axis] = n[axis] - 1;
        double wv = 0.0;
        if (bc_lo == BC_DIRV || bc_hi == BC_DIRV) {
            long wi;
            if (dim == 3) wi = (axis == 0) ? (long)(c[1] + 1) * (nz + 2) + c[2] + 1
                             : (axis == 1) ? (long)(c[0] + 1) * (nz + 2) + c[2] + 1
                                           : (long)(c[0] + 1) * (ny + 2) + c[1] + 1;
            else wi = (axis == 0) ? c[2] + 1 : c[0] + 1;
            wv = wall[wi];
        }
        full[at(lo[0], lo[1], lo[2])] = wall_rule(bc_lo, full[at(lsrc[0], lsrc[1], lsrc[2])], wv);
        full[at(hi[0], hi[1], hi[2])] = wall_rule(bc_hi, full[at(hsrc[0], hsrc[1], hsrc[2])], wv);
    }
}

// plain streaming copy, one 16-byte vector per lane (the achievable HBM
// bandwidth reference and the counter calibration)
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     long n2) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) dst[t] = src[t];
}

}  // namespace saras
