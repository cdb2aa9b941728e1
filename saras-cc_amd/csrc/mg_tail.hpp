// mg_tail.hpp -- the small bottom of the V-cycle in one workgroup.
//
// Below a few thousand cells per level a V-cycle is all launch latency: the
// levels S..D of the reference V-cycle (poisson.cc:227-277 restricted to
// those levels) run here inside one workgroup, every level held in LDS with
// a real ghost layer and the reference's literal ghost timing (imposeBC at
// the start of each sweep, periodic refresh between colours, final imposeBC;
// poisson_d3.cc:96-246, the coarsest solve :249-432).  Arithmetic follows the
// same expressions as the multi-block kernels, so the FP64 path stays
// bit-identical to the reference ordering.
#pragma once
#include "mg_kernels.hpp"

namespace saras {

constexpr int kTailMaxLevels = 8;

template <typename T>
struct TailArgs {
    int first, last;                     // levels S and D
    LevelArgs<T> L[kTailMaxLevels];      // level args of S..D (index l - S)
    RestrictArgs<T> R[kTailMaxLevels];   // restriction factors of level l (l > S)
    int xoff[kTailMaxLevels], boff[kTailMaxLevels];  // LDS offsets (elements)
    int coff;                            // LDS offset of the cached metric coefficients
    T *x;                                // global core array of level S (in/out)
    const T *b;                          // global core array of level S
    int pre, post, solve_flag, max_count;
    T w;
    double tol;
    int zero_bc_first;                   // 0 only when S is the finest level
    int wave_cells;                      // levels of at most this many padded cells run on wave 0 alone
    int *iters_out;
};

template <typename T, int DIM>
__device__ __forceinline__ T lds_residual(const T *X, const T *B, const LevelArgs<T> &L, int i, int j, int k, int PY,
                                          int PZ) {
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
    const long t = ((long)i * L.n[1] + j) * L.n[2] + k;
    return residual<T, DIM>(s, L.c2[0][i], L.c1[0][i], c2y, c1y, L.c2[2][k], L.c1[2][k], B[t]);
}

// default of TailArgs::wave_cells: levels of at most this many padded cells
// run their loops on wave 0 alone, without barriers
constexpr int kTailWaveCells = 512;

template <typename T, int DIM, bool WAVE>
__device__ void tail_smooth(T *X, const T *B, const LevelArgs<T> &L, int count, T w) {
    const int PZ = L.n[2] + 2, PY = DIM == 3 ? L.n[1] + 2 : 1;
    int per[3] = {L.bc_lo[0] == BC_PER, DIM == 3 && L.bc_lo[1] == BC_PER, L.bc_lo[2] == BC_PER};
    for (int n = 0; n < count; ++n) {
        coarse_ghosts<T, DIM, WAVE>(X, L, per, true, PY, PZ);
        coarse_colour<T, DIM, WAVE>(X, B, L, 0, w, PY, PZ);
        coarse_ghosts<T, DIM, WAVE>(X, L, per, false, PY, PZ);
        coarse_colour<T, DIM, WAVE>(X, B, L, 1, w, PY, PZ);
    }
    coarse_ghosts<T, DIM, WAVE>(X, L, per, true, PY, PZ);
}

// returns the iteration count of the reference's coarsest solve
template <typename T, int DIM, bool WAVE>
__device__ int tail_solve(T *X, const T *B, const LevelArgs<T> &L, T w, double tol, int max_count, double *red) {
    const int PZ = L.n[2] + 2, PY = DIM == 3 ? L.n[1] + 2 : 1;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const long total = (long)nx * ny * nz;
    int per[3] = {L.bc_lo[0] == BC_PER, DIM == 3 && L.bc_lo[1] == BC_PER, L.bc_lo[2] == BC_PER};
    int iter = 0;
    for (;;) {
        coarse_ghosts<T, DIM, WAVE>(X, L, per, true, PY, PZ);
        coarse_colour<T, DIM, WAVE>(X, B, L, 0, w, PY, PZ);
        coarse_ghosts<T, DIM, WAVE>(X, L, per, false, PY, PZ);
        coarse_colour<T, DIM, WAVE>(X, B, L, 1, w, PY, PZ);
        double lm = -1.0e-10;
        for (long t = threadIdx.x; t < total; t += lthreads<WAVE>()) {
            int i, j, k;
            cell_of(t, ny, nz, i, j, k);
            const double a = fabs((double)lds_residual<T, DIM>(X, B, L, i, j, k, PY, PZ));
            if (a > lm) lm = a;
        }
        const double mx = group_max<WAVE>(lm, red);
        if (mx < tol) break;
        iter += 1;
        if (iter > max_count) break;
    }
    coarse_ghosts<T, DIM, WAVE>(X, L, per, true, PY, PZ);
    return iter;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(1024) k_tail(TailArgs<T> A) {
    extern __shared__ __align__(16) unsigned char tail_raw[];
    T *lds = reinterpret_cast<T *>(tail_raw);
    __shared__ double red[16];
    __shared__ LevelArgs<T> SL[kTailMaxLevels];
    const int S = A.first, D = A.last;
    // the metric coefficients of every tail level are read in every phase:
    // keep them in LDS and point the level arguments there
    {
        T *cbuf = lds + A.coff;
        int off = 0;
        for (int q = 0; q <= D - S; ++q) {
            for (int a = 0; a < 3; ++a) {
                const int n = (DIM == 2 && a == 1) ? 1 : A.L[q].n[a];
                for (int i = threadIdx.x; i < n; i += blockDim.x) {
                    cbuf[off + i] = A.L[q].c2[a][i];
                    cbuf[off + n + i] = A.L[q].c1[a][i];
                }
                if (threadIdx.x == 0) {
                    if (a == 0) SL[q] = A.L[q];
                    SL[q].c2[a] = cbuf + off;
                    SL[q].c1[a] = cbuf + off + n;
                }
                off += 2 * n;
            }
        }
        __syncthreads();
    }
    auto X = [&](int l) { return lds + A.xoff[l - S]; };
    auto B = [&](int l) { return lds + A.boff[l - S]; };
    auto padded = [&](int l) {
        const LevelArgs<T> &L = SL[l - S];
        return DIM == 3 ? (long)(L.n[0] + 2) * (L.n[1] + 2) * (L.n[2] + 2) : (long)(L.n[0] + 2) * (L.n[2] + 2);
    };

    // level S: iterate and right-hand side from global memory
    {
        const LevelArgs<T> &L = SL[0];
        const int PZ = L.n[2] + 2, PY = DIM == 3 ? L.n[1] + 2 : 1;
        T *XS = X(S);
        for (long q = threadIdx.x; q < padded(S); q += blockDim.x) XS[q] = T(0);
        __syncthreads();
        const long total = (long)L.n[0] * L.n[1] * L.n[2];
        for (long t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            cell_of(t, L.n[1], L.n[2], i, j, k);
            const long g = (long)i * L.sx + (long)j * L.sy + kof(L, k);
            XS[cidx<T, DIM>(i, j, k, PY, PZ)] = A.x[g];
            B(S)[t] = A.b[g];
        }
        __syncthreads();
    }
    int iters = 0;  // meaningful in thread 0
    // level S: the coarsest treatment when S == D > 0 (the host restricted into
    // it); otherwise pre-smoothing, which is also the whole V-cycle of a
    // single-level hierarchy (poisson.cc:228, no coarse solve)
    // one level's smoothing or coarsest solve, on wave 0 alone when it is small
    auto smooth_lv = [&](int l, T *XL, const T *BL, int count) {
        if (padded(l) <= A.wave_cells) {
            if (threadIdx.x < 64) tail_smooth<T, DIM, true>(XL, BL, SL[l - S], count, A.w);
            __syncthreads();
        } else {
            tail_smooth<T, DIM, false>(XL, BL, SL[l - S], count, A.w);
        }
    };
    auto solve_lv = [&](int l, T *XL, const T *BL) {
        int it = 0;
        if (padded(l) <= A.wave_cells) {
            if (threadIdx.x < 64) it = tail_solve<T, DIM, true>(XL, BL, SL[l - S], A.w, A.tol, A.max_count, red);
            __syncthreads();
        } else {
            it = tail_solve<T, DIM, false>(XL, BL, SL[l - S], A.w, A.tol, A.max_count, red);
        }
        return it;
    };
    if (S == D && S > 0) {
        if (A.solve_flag) iters = solve_lv(S, X(S), B(S));
        else smooth_lv(S, X(S), B(S), A.pre);
    } else {
        smooth_lv(S, X(S), B(S), A.pre);
    }
    // down-leg
    for (int l = S; l < D; ++l) {
        const LevelArgs<T> &F = SL[l - S];
        const LevelArgs<T> &C = SL[l + 1 - S];
        const RestrictArgs<T> &R = A.R[l + 1 - S];
        const int fPZ = F.n[2] + 2, fPY = DIM == 3 ? F.n[1] + 2 : 1;
        T *XC = X(l + 1);
        for (long q = threadIdx.x; q < padded(l + 1); q += blockDim.x) XC[q] = T(0);
        const long ctotal = (long)C.n[0] * C.n[1] * C.n[2];
        for (long t = threadIdx.x; t < ctotal; t += blockDim.x) {
            int I, J, K;
            cell_of(t, C.n[1], C.n[2], I, J, K);
            T res[2][2][2];
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < (DIM == 3 ? 2 : 1); ++dj)
                    for (int dk = 0; dk < 2; ++dk)
                        res[di][dj][dk] = lds_residual<T, DIM>(X(l), B(l), F, 2 * I + di, DIM == 3 ? 2 * J + dj : 0,
                                                               2 * K + dk, fPY, fPZ);
            T out;
            if (DIM == 3) {
                const T lx = R.lo[0][I], hx = R.hi[0][I], ly = R.lo[1][J], hy = R.hi[1][J], lz = R.lo[2][K],
                        hz = R.hi[2][K];
                const T v000 = lx * ly * lz, v100 = hx * ly * lz, v010 = lx * hy * lz, v001 = lx * ly * hz;
                const T v011 = lx * hy * hz, v101 = hx * ly * hz, v110 = hx * hy * lz, v111 = hx * hy * hz;
                const T vtot = R.wd[0][I] * R.wd[1][J] * R.wd[2][K];
                out = (v111 * res[1][1][1] + v000 * res[0][0][0] + v110 * res[1][1][0] + v001 * res[0][0][1] +
                       v101 * res[1][0][1] + v010 * res[0][1][0] + v011 * res[0][1][1] + v100 * res[1][0][0]) / vtot;
            } else {
                out = (res[1][0][1] + res[0][0][0] + res[1][0][0] + res[0][0][1]) / T(4);
            }
            B(l + 1)[t] = out;
        }
        __syncthreads();
        if (l + 1 == D) {
            if (A.solve_flag) iters = solve_lv(D, XC, B(D));
            else smooth_lv(D, XC, B(D), A.pre);
        } else {
            smooth_lv(l + 1, XC, B(l + 1), A.pre);
        }
    }
    // up-leg: injection plus the saved iterate, then post-smoothing
    for (int l = D - 1; l >= S; --l) {
        const LevelArgs<T> &F = SL[l - S];
        const LevelArgs<T> &C = SL[l + 1 - S];
        const int fPZ = F.n[2] + 2, fPY = DIM == 3 ? F.n[1] + 2 : 1;
        const int cPZ = C.n[2] + 2, cPY = DIM == 3 ? C.n[1] + 2 : 1;
        const long total = (long)F.n[0] * F.n[1] * F.n[2];
        T *XF = X(l);
        const T *XC = X(l + 1);
        for (long t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            cell_of(t, F.n[1], F.n[2], i, j, k);
            const long c = cidx<T, DIM>(i, j, k, fPY, fPZ);
            XF[c] = XC[cidx<T, DIM>(i >> 1, j >> 1, k >> 1, cPY, cPZ)] + XF[c];
        }
        __syncthreads();
        smooth_lv(l, XF, B(l), A.post);
    }
    // level S back to global memory
    {
        const LevelArgs<T> &L = SL[0];
        const int PZ = L.n[2] + 2, PY = DIM == 3 ? L.n[1] + 2 : 1;
        const long total = (long)L.n[0] * L.n[1] * L.n[2];
        for (long t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            cell_of(t, L.n[1], L.n[2], i, j, k);
            A.x[(long)i * L.sx + (long)j * L.sy + kof(L, k)] = X(S)[cidx<T, DIM>(i, j, k, PY, PZ)];
        }
    }
    if (threadIdx.x == 0 && A.iters_out && A.solve_flag) *A.iters_out = iters;
}

}  // namespace saras
