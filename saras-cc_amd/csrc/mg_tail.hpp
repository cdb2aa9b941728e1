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
//
// Each phase of a level is a chain of LDS round trips, so the kernel keeps
// that chain short: a level is described by LDS offsets (every access is a
// plain LDS read, never a pointer fetched from memory), and when a level has
// no more cells than threads each thread loads its cell's indices,
// coefficients and right-hand side once per smoothing call, leaving a colour
// pass with the seven neighbour reads and one write.
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
    int coff;                            // LDS offset of the staged coefficients and restriction factors
    const T *coef;                       // their packed image in global memory (the LDS region's order)
    T *x;                                // global core array of level S (in/out)
    const T *b;                          // global core array of level S
    int pre, post, solve_flag, max_count;
    T w;
    double tol;
    double rel_floor;                    // as CoarseCtl::rel_floor
    int zero_bc_first;                   // 0 only when S is the finest level
    int *iters_out;
};


// One tail level, every array an LDS element offset.
template <typename T>
struct TailLevel {
    int n[3];
    int PY, PZ;                // padded extents of y (1 in 2-D) and z
    int x, b;                  // padded iterate, core right-hand side
    int c2[3], c1[3];          // metric coefficients per axis
    int lo[3], hi[3], wd[3];   // restriction factors of this (coarse) level, 3-D
    int bc_lo[3], bc_hi[3];
    const T *wall[3];
};

template <int DIM>
__device__ __forceinline__ int tidx(int PY, int PZ, int i, int j, int k) {
    return DIM == 3 ? ((i + 1) * PY + (j + 1)) * PZ + (k + 1) : (i + 1) * PZ + (k + 1);
}

// (i, j, k) of core index t of a small level (32-bit unsigned arithmetic)
__device__ __forceinline__ void tcoords(int t, int ny, int nz, int &i, int &j, int &k) {
    const unsigned r = (unsigned)t / (unsigned)nz;
    k = t - (int)(r * (unsigned)nz);
    const unsigned q = r / (unsigned)ny;
    j = (int)(r - q * (unsigned)ny);
    i = (int)q;
}

// one core cell with everything but its neighbours
template <typename T>
struct TCell {
    int c;     // padded offset (relative to the iterate)
    int par;   // (i + j + k) & 1
    T c2x, c1x, c2y, c1y, c2z, c1z, b;
};

template <typename T, int DIM>
__device__ __forceinline__ void tcell_load(const T *lds, const TailLevel<T> &V, int t, TCell<T> &q) {
    int i, j, k;
    tcoords(t, V.n[1], V.n[2], i, j, k);
    q.c = tidx<DIM>(V.PY, V.PZ, i, j, k);
    q.par = (i + j + k) & 1;
    q.c2x = lds[V.c2[0] + i];
    q.c1x = lds[V.c1[0] + i];
    q.c2y = DIM == 3 ? lds[V.c2[1] + j] : T(0);
    q.c1y = DIM == 3 ? lds[V.c1[1] + j] : T(0);
    q.c2z = lds[V.c2[2] + k];
    q.c1z = lds[V.c1[2] + k];
    q.b = lds[V.b + t];
}

template <typename T, int DIM>
__device__ __forceinline__ void tgather(const T *X, int c, int sx, int sy, Stencil<T, DIM> &s) {
    s.c = X[c];
    s.xE = X[c + sx];
    s.xW = X[c - sx];
    if (DIM == 3) {
        s.yN = X[c + sy];
        s.yS = X[c - sy];
    }
    s.zT = X[c + 1];
    s.zB = X[c - 1];
}

template <typename T, int DIM>
__device__ __forceinline__ T tresidual(const T *X, const TCell<T> &q, int sx, int sy) {
    Stencil<T, DIM> s;
    tgather<T, DIM>(X, q.c, sx, sy, s);
    return residual<T, DIM>(s, q.c2x, q.c1x, q.c2y, q.c1y, q.c2z, q.c1z, q.b);
}

// One face-ghost pair: the ghosts at both ends of a line of cells along an
// axis (lo beside cell 0, hi beside cell n-1).
template <typename T>
struct TGhost {
    int glo, ghi, s0, s1;
    int bclo, bchi;
    bool per;
    T wv;
};

// item q of the face lines: x faces (j, k), then y faces (i, k), then z faces (i, j)
template <typename T, int DIM>
__device__ __forceinline__ void tghost_load(const TailLevel<T> &V, int q, TGhost<T> &g) {
    const int nx = V.n[0], ny = DIM == 3 ? V.n[1] : 1, nz = V.n[2];
    const int fx = ny * nz, fy = DIM == 3 ? nx * nz : 0;
    if (q < fx) {
        const unsigned j = (unsigned)q / (unsigned)nz;
        const int k = q - (int)(j * (unsigned)nz), jj = (int)j;
        g.glo = tidx<DIM>(V.PY, V.PZ, -1, jj, k);
        g.ghi = tidx<DIM>(V.PY, V.PZ, nx, jj, k);
        g.s0 = tidx<DIM>(V.PY, V.PZ, 0, jj, k);
        g.s1 = tidx<DIM>(V.PY, V.PZ, nx - 1, jj, k);
        g.bclo = V.bc_lo[0];
        g.bchi = V.bc_hi[0];
        g.wv = V.bc_lo[0] == BC_DIRV ? (DIM == 3 ? V.wall[0][(jj + 1) * (nz + 2) + k + 1] : V.wall[0][k + 1]) : T(0);
    } else if (q < fx + fy) {
        const unsigned uq = (unsigned)(q - fx), i = uq / (unsigned)nz;
        const int k = (int)(uq - i * (unsigned)nz), ii = (int)i;
        g.glo = tidx<DIM>(V.PY, V.PZ, ii, -1, k);
        g.ghi = tidx<DIM>(V.PY, V.PZ, ii, ny, k);
        g.s0 = tidx<DIM>(V.PY, V.PZ, ii, 0, k);
        g.s1 = tidx<DIM>(V.PY, V.PZ, ii, ny - 1, k);
        g.bclo = V.bc_lo[1];
        g.bchi = V.bc_hi[1];
        g.wv = V.bc_lo[1] == BC_DIRV ? V.wall[1][(ii + 1) * (nz + 2) + k + 1] : T(0);
    } else {
        const unsigned uq = (unsigned)(q - fx - fy), i = uq / (unsigned)ny;
        const int j = (int)(uq - i * (unsigned)ny), ii = (int)i;
        g.glo = tidx<DIM>(V.PY, V.PZ, ii, j, -1);
        g.ghi = tidx<DIM>(V.PY, V.PZ, ii, j, nz);
        g.s0 = tidx<DIM>(V.PY, V.PZ, ii, j, 0);
        g.s1 = tidx<DIM>(V.PY, V.PZ, ii, j, nz - 1);
        g.bclo = V.bc_lo[2];
        g.bchi = V.bc_hi[2];
        g.wv = V.bc_lo[2] == BC_DIRV ? (DIM == 3 ? V.wall[2][(ii + 1) * (ny + 2) + j + 1] : V.wall[2][ii + 1]) : T(0);
    }
    g.per = g.bclo == BC_PER;
}

// periodic lines copy the opposite core cell; with `walls` the others take
// the wall rule of their own core cell (updateFace / imposeBC,
// poisson_d3.cc:876-993).  Only ghosts beside core cells are written: edges
// and corners of the shell are never read by the 7-point stencil.
template <typename T>
__device__ __forceinline__ void tghost_apply(T *X, const TGhost<T> &g, bool walls) {
    if (g.per) {
        X[g.ghi] = X[g.s0];
        X[g.glo] = X[g.s1];
    } else if (walls) {
        X[g.glo] = wall_rule(g.bclo, X[g.s0], g.wv);
        X[g.ghi] = wall_rule(g.bchi, X[g.s1], g.wv);
    }
}

// cells of one colour per thread held in a colour list (see tail_level)
constexpr int kTailColourCells = 4;
// face-ghost items per thread held in registers
constexpr int kTailGhostItems = 2;

// The smoothing sweeps (or the coarsest solve) of one level: `count` RB-SOR
// sweeps, then the final imposeBC; with SOLVE, sweeps until max|r| < tol or
// the iteration cap (poisson_d3.cc:249-432), returning the iteration count.
// A level with no more cells than threads keeps each thread's cell in
// registers; a larger level with an even z extent and at most
// kTailColourCells cells of a colour per thread keeps colour lists: the
// packed (i, j, k) of the thread's cells of each colour, so a colour pass
// visits only its own cells with every lane busy and loads all of a
// thread's stencils before the first write (the cells of a colour never
// read each other).  Other levels walk all cells.
template <typename T, int DIM, bool WAVE, bool SOLVE>
__device__ int tail_level(T *lds, const TailLevel<T> &V, int count, T w, double tol, double rel_floor, int max_count,
                          double *red) {
    constexpr int CPT = kTailColourCells, GPT = kTailGhostItems;
    const int tid = threadIdx.x, nthr = lthreads<WAVE>();
    const int total = V.n[0] * V.n[1] * V.n[2];
    const int nface = DIM == 3 ? V.n[1] * V.n[2] + V.n[0] * V.n[2] + V.n[0] * V.n[1] : V.n[2] + V.n[0];
    const int sx = DIM == 3 ? V.PY * V.PZ : V.PZ, sy = V.PZ;
    T *X = lds + V.x;
    const bool cached = total <= nthr, gcached = nface <= GPT * nthr;
    const int half = total >> 1, hz2 = V.n[2] >> 1;
    const bool lists = !cached && (V.n[2] & 1) == 0 && half <= CPT * nthr;
    TCell<T> q{};
    TGhost<T> g[GPT]{};
    const bool act = cached && tid < total;
    if (act) tcell_load<T, DIM>(lds, V, tid, q);
    if (gcached)
#pragma unroll
        for (int e = 0; e < GPT; ++e)
            if (tid + e * nthr < nface) tghost_load<T, DIM>(V, tid + e * nthr, g[e]);
    int cell[2][CPT];   // packed i << 20 | j << 10 | k, or -1
#pragma unroll
    for (int col = 0; col < 2; ++col)
#pragma unroll
        for (int e = 0; e < CPT; ++e) {
            cell[col][e] = -1;
            const int u = tid + e * nthr;
            if (lists && u < half) {
                const unsigned row = (unsigned)u / (unsigned)hz2;
                const int m = u - (int)(row * (unsigned)hz2);
                const unsigned i = row / (unsigned)V.n[1];
                const int j = (int)(row - i * (unsigned)V.n[1]);
                const int k = 2 * m + (((int)i + j + col) & 1);
                cell[col][e] = ((int)i << 20) | (j << 10) | k;
            }
        }

    auto ghosts = [&](bool walls) {
        if (gcached) {
#pragma unroll
            for (int e = 0; e < GPT; ++e)
                if (tid + e * nthr < nface) tghost_apply(X, g[e], walls);
        } else {
            for (int t = tid; t < nface; t += nthr) {
                TGhost<T> h;
                tghost_load<T, DIM>(V, t, h);
                tghost_apply(X, h, walls);
            }
        }
        lsync<WAVE>();
    };
    auto update = [&](const TCell<T> &r) {
        Stencil<T, DIM> s;
        tgather<T, DIM>(X, r.c, sx, sy, s);
        X[r.c] = sor_update<T, DIM>(s, r.c2x, r.c1x, r.c2y, r.c1y, r.c2z, r.c1z, r.b, w);
    };
    auto colour = [&](int col) {
        if (cached) {
            if (act && q.par == col) update(q);
        } else if (lists) {
            T r[CPT];
            int c[CPT];
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                const int p = col ? cell[1][e] : cell[0][e];
                c[e] = -1;
                if (p < 0) continue;
                const int i = p >> 20, j = (p >> 10) & 1023, k = p & 1023;
                TCell<T> t;
                t.c = tidx<DIM>(V.PY, V.PZ, i, j, k);
                t.c2x = lds[V.c2[0] + i];
                t.c1x = lds[V.c1[0] + i];
                t.c2y = DIM == 3 ? lds[V.c2[1] + j] : T(0);
                t.c1y = DIM == 3 ? lds[V.c1[1] + j] : T(0);
                t.c2z = lds[V.c2[2] + k];
                t.c1z = lds[V.c1[2] + k];
                t.b = lds[V.b + (i * V.n[1] + j) * V.n[2] + k];
                Stencil<T, DIM> s;
                tgather<T, DIM>(X, t.c, sx, sy, s);
                r[e] = sor_update<T, DIM>(s, t.c2x, t.c1x, t.c2y, t.c1y, t.c2z, t.c1z, t.b, w);
                c[e] = t.c;
            }
#pragma unroll
            for (int e = 0; e < CPT; ++e)
                if (c[e] >= 0) X[c[e]] = r[e];
        } else {
            for (int t = tid; t < total; t += nthr) {
                int i, j, k;
                tcoords(t, V.n[1], V.n[2], i, j, k);
                if (((i + j + k) & 1) != col) continue;
                TCell<T> r;
                tcell_load<T, DIM>(lds, V, t, r);
                update(r);
            }
        }
        lsync<WAVE>();
    };
    // imposeBC, red, the periodic refresh, black: one loop so that the
    // colour pass's code appears once
    auto sweep = [&]() {
        for (int col = 0; col < 2; ++col) {
            ghosts(col == 0);
            colour(col);
        }
    };

    int iter = 0;
    if (SOLVE) {
        if (rel_floor > 0.0) {
            // the FP32 floor of the correction equation (CoarseCtl::rel_floor)
            double bm = 0.0;
            for (int t = tid; t < total; t += nthr) bm = fmax(bm, fabs((double)lds[V.b + t]));
            tol = fmax(tol, rel_floor * group_max<WAVE>(bm, red));
        }
        for (;;) {
            sweep();
            // max |r| with the ghosts as they stand (poisson_d3.cc:398-415)
            double lm = -1.0e-10;
            if (cached) {
                if (act) {
                    const double a = fabs((double)tresidual<T, DIM>(X, q, sx, sy));
                    if (a > lm) lm = a;
                }
            } else {
                for (int t = tid; t < total; t += nthr) {
                    TCell<T> r;
                    tcell_load<T, DIM>(lds, V, t, r);
                    const double a = fabs((double)tresidual<T, DIM>(X, r, sx, sy));
                    if (a > lm) lm = a;
                }
            }
            const double mx = group_max<WAVE>(lm, red);
            if (mx < tol) break;
            iter += 1;
            if (iter > max_count) break;
        }
    } else {
        for (int n = 0; n < count; ++n) sweep();
    }
    ghosts(true);
    return iter;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(512) k_tail(TailArgs<T> A) {
    extern __shared__ __align__(16) unsigned char tail_raw[];
    T *lds = reinterpret_cast<T *>(tail_raw);
    __shared__ double red[16];
    __shared__ TailLevel<T> TL[kTailMaxLevels];
    const int S = A.first, D = A.last, nl = D - S + 1;
    auto ext = [&](int q, int a) { return (DIM == 2 && a == 1) ? 1 : A.L[q].n[a]; };

    // Staging, with every global load issued before the first barrier: the
    // metric coefficients of each tail level and (3-D) the restriction
    // factors of levels S+1..D, which every phase reads, are copied to LDS
    // from their packed image (one flat coalesced copy) and described by LDS
    // offsets; level S's iterate and right-hand side come in beside them.
    // Level S's ghost shell is not cleared: every face ghost is written
    // before it is read.
    {
        if (threadIdx.x == 0) {
            int off = A.coff;
            for (int q = 0; q < nl; ++q) {
                TailLevel<T> &V = TL[q];
                const LevelArgs<T> &L = A.L[q];
                for (int a = 0; a < 3; ++a) {
                    V.n[a] = L.n[a];
                    V.bc_lo[a] = L.bc_lo[a];
                    V.bc_hi[a] = L.bc_hi[a];
                    V.wall[a] = L.wall[a];
                    V.c2[a] = off;
                    V.c1[a] = off + ext(q, a);
                    off += 2 * ext(q, a);
                }
                V.PY = DIM == 3 ? L.n[1] + 2 : 1;
                V.PZ = L.n[2] + 2;
                V.x = A.xoff[q];
                V.b = A.boff[q];
            }
            if (DIM == 3)
                for (int q = 1; q < nl; ++q)
                    for (int a = 0; a < 3; ++a) {
                        const int n = A.L[q].n[a];
                        TL[q].lo[a] = off;
                        TL[q].hi[a] = off + n;
                        TL[q].wd[a] = off + 2 * n;
                        off += 3 * n;
                    }
        }
        int ncoef = 0, nres = 0;
        for (int q = 0; q < nl; ++q)
            for (int a = 0; a < 3; ++a) {
                ncoef += 2 * ext(q, a);
                if (DIM == 3 && q > 0) nres += 3 * A.L[q].n[a];
            }
        for (int e = threadIdx.x; e < ncoef + nres; e += blockDim.x) lds[A.coff + e] = A.coef[e];
        const LevelArgs<T> &L = A.L[0];
        const int PZ = L.n[2] + 2, PY = DIM == 3 ? L.n[1] + 2 : 1;
        T *XS = lds + A.xoff[0];
        T *BS = lds + A.boff[0];
        const int total = L.n[0] * L.n[1] * L.n[2];
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            tcoords(t, L.n[1], L.n[2], i, j, k);
            const long g = (long)i * L.sx + (long)j * L.sy + kof(L, k);
            XS[tidx<DIM>(PY, PZ, i, j, k)] = A.x[g];
            BS[t] = A.b[g];
        }
        __syncthreads();
    }
    auto padded = [&](const TailLevel<T> &V) { return DIM == 3 ? (V.n[0] + 2) * V.PY * V.PZ : (V.n[0] + 2) * V.PZ; };
    // one level's smoothing or coarsest solve by the whole workgroup (a
    // wave-only variant for the smallest levels measured no faster and
    // doubled the kernel's code)
    auto level = [&](int l, bool solve, int count) {
        const TailLevel<T> V = TL[l - S];
        if (solve) return tail_level<T, DIM, false, true>(lds, V, 0, A.w, A.tol, A.rel_floor, A.max_count, red);
        tail_level<T, DIM, false, false>(lds, V, count, A.w, A.tol, A.rel_floor, A.max_count, red);
        return 0;
    };

    // The V-cycle over levels S..D as one loop of steps, so that each phase's
    // code appears once in the kernel (a single CU runs this kernel, and its
    // instruction cache holds far less than one copy per call site):
    //   step 0: level S -- the coarsest treatment when S == D > 0 (the host
    //     restricted into it), otherwise pre-smoothing, which is also the
    //     whole V-cycle of a single-level hierarchy (poisson.cc:228);
    //   down-leg, per l = S..D-1: the residual of l restricted into the
    //     right-hand side of l + 1 (whose iterate starts at zero), then
    //     pre-smoothing of l + 1 or the coarsest solve;
    //   up-leg, per l = D-1..S: injection plus the saved iterate, then
    //     post-smoothing of l.
    int iters = 0;  // meaningful in thread 0
    const int nsteps = 1 + 4 * (D - S);
    for (int st = 0; st < nsteps; ++st) {
        int kind = 0, l = S, count = A.pre;   // kind 0: smooth/solve l, 1: restrict l, 2: prolong into l
        bool solve = false;
        if (st == 0) {
            solve = S == D && S > 0 && A.solve_flag;
        } else if (st <= 2 * (D - S)) {
            const int d = st - 1;
            l = S + d / 2;
            if ((d & 1) == 0) {
                kind = 1;
            } else {
                l += 1;
                solve = l == D && A.solve_flag;
            }
        } else {
            const int u = st - 1 - 2 * (D - S);
            l = D - 1 - u / 2;
            kind = (u & 1) == 0 ? 2 : 0;
            count = A.post;
        }
        if (kind == 1) {
            const TailLevel<T> F = TL[l - S];
            const TailLevel<T> C = TL[l + 1 - S];
            const int fsx = DIM == 3 ? F.PY * F.PZ : F.PZ, fsy = F.PZ;
            const T *XF = lds + F.x;
            T *XC = lds + C.x;
            for (int q = threadIdx.x; q < padded(C); q += blockDim.x) XC[q] = T(0);
            const int ctotal = C.n[0] * C.n[1] * C.n[2];
            for (int t = threadIdx.x; t < ctotal; t += blockDim.x) {
                int I, J, K;
                tcoords(t, C.n[1], C.n[2], I, J, K);
                T res[2][2][2];
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < (DIM == 3 ? 2 : 1); ++dj)
                        for (int dk = 0; dk < 2; ++dk) {
                            const int i = 2 * I + di, j = DIM == 3 ? 2 * J + dj : 0, k = 2 * K + dk;
                            TCell<T> r;
                            r.c = tidx<DIM>(F.PY, F.PZ, i, j, k);
                            r.c2x = lds[F.c2[0] + i];
                            r.c1x = lds[F.c1[0] + i];
                            r.c2y = DIM == 3 ? lds[F.c2[1] + j] : T(0);
                            r.c1y = DIM == 3 ? lds[F.c1[1] + j] : T(0);
                            r.c2z = lds[F.c2[2] + k];
                            r.c1z = lds[F.c1[2] + k];
                            r.b = lds[F.b + (i * F.n[1] + j) * F.n[2] + k];
                            res[di][dj][dk] = tresidual<T, DIM>(XF, r, fsx, fsy);
                        }
                T out;
                if (DIM == 3) {
                    const T lx = lds[C.lo[0] + I], hx = lds[C.hi[0] + I], ly = lds[C.lo[1] + J],
                            hy = lds[C.hi[1] + J], lz = lds[C.lo[2] + K], hz = lds[C.hi[2] + K];
                    const T v000 = lx * ly * lz, v100 = hx * ly * lz, v010 = lx * hy * lz, v001 = lx * ly * hz;
                    const T v011 = lx * hy * hz, v101 = hx * ly * hz, v110 = hx * hy * lz, v111 = hx * hy * hz;
                    const T vtot = lds[C.wd[0] + I] * lds[C.wd[1] + J] * lds[C.wd[2] + K];
                    out = (v111 * res[1][1][1] + v000 * res[0][0][0] + v110 * res[1][1][0] + v001 * res[0][0][1] +
                           v101 * res[1][0][1] + v010 * res[0][1][0] + v011 * res[0][1][1] + v100 * res[1][0][0]) /
                          vtot;
                } else {
                    out = (res[1][0][1] + res[0][0][0] + res[1][0][0] + res[0][0][1]) / T(4);
                }
                lds[C.b + t] = out;
            }
            __syncthreads();
        } else if (kind == 2) {
            const TailLevel<T> F = TL[l - S];
            const TailLevel<T> C = TL[l + 1 - S];
            const int total = F.n[0] * F.n[1] * F.n[2];
            T *XF = lds + F.x;
            const T *XC = lds + C.x;
            for (int t = threadIdx.x; t < total; t += blockDim.x) {
                int i, j, k;
                tcoords(t, F.n[1], F.n[2], i, j, k);
                const int c = tidx<DIM>(F.PY, F.PZ, i, j, k);
                XF[c] = XC[tidx<DIM>(C.PY, C.PZ, i >> 1, j >> 1, k >> 1)] + XF[c];
            }
            __syncthreads();
        } else {
            const int it = level(l, solve, count);
            if (solve) iters = it;
        }
    }
    // level S back to global memory
    {
        const LevelArgs<T> &L = A.L[0];
        const TailLevel<T> V = TL[0];
        const int total = L.n[0] * L.n[1] * L.n[2];
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j, k;
            tcoords(t, L.n[1], L.n[2], i, j, k);
            A.x[(long)i * L.sx + (long)j * L.sy + kof(L, k)] = lds[V.x + tidx<DIM>(V.PY, V.PZ, i, j, k)];
        }
    }
    if (threadIdx.x == 0 && A.iters_out && A.solve_flag) *A.iters_out = iters;
}

}  // namespace saras
