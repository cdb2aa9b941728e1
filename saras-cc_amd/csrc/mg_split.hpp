// mg_split.hpp -- red-black SOR colour pass on colour-split rows.
//
// In a split row the cells of one colour are one contiguous half-row (even k
// in [0, nz/2), odd k at hz + [0, nz/2)), so one colour pass (poisson_d3.cc:
// 105-170 for red, 177-242 for black) reads and writes whole 16-byte vectors:
// each lane updates W = 16 / sizeof(T) consecutive cells of its row's
// colour.  Per cell it loads the same-half vectors of the four x/y
// neighbour rows, the other half of its own row (the two z neighbours of
// each cell; the one that lies in the adjacent vector comes from the
// neighbouring lane), the right-hand side and the reordered z metric
// coefficients.  The update is the same sor_update expression as the
// z-contiguous pass, so both layouts give identical values.
//
// Blocks are remapped so that each XCD works through one contiguous range of
// rows: the neighbour rows of a row were fetched by the same XCD a moment
// earlier and are read from its L2.
#pragma once
#include "mg_kernels.hpp"

namespace saras {

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using V = float4;
    static constexpr int W = 4;
    __device__ static float get(const float4 &v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
    __device__ static float4 make(const float *a) { return make_float4(a[0], a[1], a[2], a[3]); }
};
template <>
struct Vec16<double> {
    using V = double2;
    static constexpr int W = 2;
    __device__ static double get(const double2 &v, int e) { return e == 0 ? v.x : v.y; }
    __device__ static double2 make(const double *a) { return make_double2(a[0], a[1]); }
};

constexpr int kSplitThreads = 256;

// W consecutive coarse cells K0..K0+W-1 of coarse row (I, J) in the coarse
// level's own layout (split or z-contiguous), as one vector
template <typename T>
__device__ __forceinline__ void coarse_vec(const T *__restrict__ e, long base, long chz, int K0, T (&out)[Vec16<T>::W]) {
    using VT = Vec16<T>;
    constexpr int W = VT::W;
    if (chz == 0) {
        const typename VT::V v = *reinterpret_cast<const typename VT::V *>(e + base + K0);
#pragma unroll
        for (int q = 0; q < W; ++q) out[q] = VT::get(v, q);
    } else {
        if constexpr (W == 4) {
            // even K at K/2, odd K at chz + K/2 (K0 is a multiple of 4): two pairs
            const float2 ev = *reinterpret_cast<const float2 *>(e + base + (K0 >> 1));
            const float2 od = *reinterpret_cast<const float2 *>(e + base + chz + (K0 >> 1));
            out[0] = ev.x;
            out[1] = od.x;
            out[2] = ev.y;
            out[3] = od.y;
        } else {
#pragma unroll
            for (int q = 0; q < W; ++q) out[q] = e[base + (q & 1) * chz + ((K0 + q) >> 1)];
        }
    }
}

template <typename T>
__device__ __forceinline__ void coarse_store(T *__restrict__ dst, long base, long chz, int K0,
                                             const T (&in)[Vec16<T>::W]) {
    using VT = Vec16<T>;
    constexpr int W = VT::W;
    if (chz == 0) {
        *reinterpret_cast<typename VT::V *>(dst + base + K0) = VT::make(in);
    } else {
#pragma unroll
        for (int q = 0; q < W; ++q) dst[base + (q & 1) * chz + ((K0 + q) >> 1)] = in[q];
    }
}


// N consecutive values of a row at p: one vector load where N fills 16 bytes
template <typename T, int N>
__device__ __forceinline__ void load_n(const T *__restrict__ p, T (&out)[N]) {
    if constexpr (N == Vec16<T>::W) {
        const typename Vec16<T>::V v = *reinterpret_cast<const typename Vec16<T>::V *>(p);
#pragma unroll
        for (int q = 0; q < N; ++q) out[q] = Vec16<T>::get(v, q);
    } else if constexpr (N == 2 && sizeof(T) == 4) {
        const float2 v = *reinterpret_cast<const float2 *>(p);
        out[0] = v.x;
        out[1] = v.y;
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q) out[q] = p[q];
    }
}

// N consecutive coarse cells K0.. of a coarse row in the coarse level's layout
template <typename T, int N>
__device__ __forceinline__ void coarse_store_n(T *__restrict__ dst, long base, long chz, int K0, const T (&in)[N]) {
    if (chz == 0 && N == Vec16<T>::W) {
        *reinterpret_cast<typename Vec16<T>::V *>(dst + base + K0) = Vec16<T>::make(in);
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            const int K = K0 + q;
            dst[base + (chz ? (long)(K & 1) * chz + (K >> 1) : (long)K)] = in[q];
        }
    }
}

// one coarse cell (I, J, K) of a level in its own layout
template <typename T>
__device__ __forceinline__ T coarse_at(const T *__restrict__ e, long base, long chz, int K) {
    return e[base + (chz ? (long)(K & 1) * chz + (K >> 1) : (long)K)];
}

// One colour pass over a split level; the grid is a multiple of kXcds
// blocks, `nblocks` of them carry work.  PRO: the pass belongs to the first
// sweep after the prolongation, whose input is x + P e with P the injection
// from the coarse iterate e (poisson_d3.cc:548-558, poisson.cc:266): the red
// pass reads every cell as e(coarse) + x, the black pass its own cells (its
// neighbours are the red values just written), so the separate
// prolongation pass is not needed.
// ZPRE: the z neighbour that no adjacent lane holds (periodic wrap, or a row
// split between waves) is loaded together with the row vectors; without it
// that load is issued only where needed, which is cheaper when it never is.
// W1: w == 1 (the "SOR Parameter" of every reference deck).  The update
// (1 - w) x + q then equals q whenever q != 0 and x is finite (0 x is a
// signed zero), so the cells' own values are read only where q is a zero
// (to give the sign the reference's sum has) or where a wall rule needs them:
// a pass reads 1.5 words per updated cell pair instead of 2.
// Z: the level's iterate is known to be zero (the FP32 correction at the
// start of a refinement); the red pass reads no x at all, uses +0 wherever
// it would have read a cell, and writes +0 to the black half of its vectors,
// which the iterate would otherwise have been cleared to beforehand: the
// same values as clearing x and running the ordinary pass.
template <typename T, int DIM, bool PRO, bool ZPRE, bool W1, bool Z = false>
__device__ __forceinline__ void rb_split_body(const LevelArgs<T> &L, T *__restrict__ x, const T *__restrict__ b,
                                              int colour, T w, int nblocks, const T *__restrict__ e, long esx,
                                              long esy, long ehz) {
    static_assert(!(Z && PRO), "a zero iterate has no prolongation to apply");
    using VT = Vec16<T>;
    using V = typename VT::V;
    constexpr int W = VT::W;
    const int per = gridDim.x / kXcds;
    const int logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (logical >= nblocks) return;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int half = nz >> 1, q = half / W;
    const long total = (long)nx * ny * q;
    const long t = (long)logical * kSplitThreads + threadIdx.x;
    if (t >= total) return;
    const int row = (int)(t / q);
    const int v = (int)(t - (long)row * q);
    int i, j;
    if (L.rtile) {
        const int jt = L.rtile, tile = row / (nx * jt), rem = row - tile * (nx * jt);
        i = rem >> __builtin_ctz(jt);
        j = tile * jt + (rem & (jt - 1));
    } else {
        i = row / ny;
        j = row - i * ny;
    }
    const int h = (i + j + colour) & 1;   // half of this row holding the colour
    const int m0 = v * W;
    const long rb = (long)i * L.sx + (long)j * L.sy;
    const long own = rb + h * L.hz + m0;
    const long oth = rb + (1 - h) * L.hz + m0;
    // the cell's own value: always without W1; with it only on rows whose x/y
    // neighbour lies beyond a wall and at the z-wall ends of a row
    bool need_c = !W1;
    if (W1) {
        need_c = (i + 1 >= nx && L.bc_hi[0] != BC_PER) || (i == 0 && L.bc_lo[0] != BC_PER);
        if (DIM == 3) need_c = need_c || (j + 1 >= ny && L.bc_hi[1] != BC_PER) || (j == 0 && L.bc_lo[1] != BC_PER);
        need_c = need_c || (h == 0 && m0 == 0 && L.bc_lo[2] != BC_PER) ||
                 (h == 1 && m0 + W == half && L.bc_hi[2] != BC_PER);
    }
    V cv{}, ov{};
    if (!Z && need_c) cv = *reinterpret_cast<const V *>(x + own);
    if (!Z) ov = *reinterpret_cast<const V *>(x + oth);
    const V bv = *reinterpret_cast<const V *>(b + own);
    const V c2v = *reinterpret_cast<const V *>(L.c2zs + h * half + m0);
    const V c1v = *reinterpret_cast<const V *>(L.c1zs + h * half + m0);
    // The z neighbour beyond this lane's vector that no adjacent lane of the
    // wave holds (the periodic wrap at a row end, or the row's previous / next
    // vector when it sits in another wave), loaded with the row vectors so
    // that its latency overlaps theirs.  h == 0 (k even): element m0 - 1 of
    // the other half; h == 1 (k odd): element m0 + W of the other half.
    const int lane = threadIdx.x & 63;
    long zoff = -1;
    int zK = 0;
    if (ZPRE) {
    if (h == 0) {
        if (m0 == 0) {
            if (L.bc_lo[2] == BC_PER) {
                zoff = rb + L.hz + (half - 1);
                zK = half - 1;
            }
        } else if (lane == 0) {
            zoff = oth - 1;
            zK = m0 - 1;
        }
    } else {
        if (m0 + W == half) {
            if (L.bc_hi[2] == BC_PER) {
                zoff = rb;
                zK = 0;
            }
        } else if (lane == 63) {
            zoff = oth + W;
            zK = m0 + W;
        }
    }
    }
    const T zx = (!Z && ZPRE && zoff >= 0) ? x[zoff] : T(0);
    T c[W], o[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        c[q] = VT::get(cv, q);
        o[q] = VT::get(ov, q);
    }
    // own coarse row of the prolongation: cells m0.. of both halves map to K = m0..
    // The black pass of W1 reads no cell of its own colour (its neighbours are
    // the red values just written), so it needs the correction only where it
    // reads its own cells: on wall rows, or for a zero quotient (loaded there).
    const long ebase = (long)(i >> 1) * esx + (long)(j >> 1) * esy;
    T evo[W];
    const bool evo_now = PRO && (colour == 0 || !W1 || need_c);
    if (evo_now) {
        coarse_vec(e, ebase, ehz, m0, evo);
#pragma unroll
        for (int q = 0; q < W; ++q) {
            c[q] = evo[q] + c[q];
            if (colour == 0) o[q] = evo[q] + o[q];
        }
    }
    // x and y neighbour rows: whole vectors, or the wall rule per cell
    auto kk = [&](int e) { return 2 * (m0 + e) + h; };
    T xe[W], xw[W], yn[W], ys[W];
    // nbr_i, nbr_j: the neighbour row's indices (wrapped), for its coarse row
    auto side = [&](T(&out)[W], bool inside, long off, int bc, int axis, int nbr_i, int nbr_j) {
        if (inside || bc == BC_PER) {
            V nv{};
            if (!Z) nv = *reinterpret_cast<const V *>(x + own + off);
#pragma unroll
            for (int q = 0; q < W; ++q) out[q] = VT::get(nv, q);
            if (PRO && colour == 0) {
                // a neighbour row in the same coarse row shares this lane's correction
                if ((nbr_i >> 1) == (i >> 1) && (nbr_j >> 1) == (j >> 1)) {
#pragma unroll
                    for (int q = 0; q < W; ++q) out[q] = evo[q] + out[q];
                } else {
                    T ev[W];
                    coarse_vec(e, (long)(nbr_i >> 1) * esx + (long)(nbr_j >> 1) * esy, ehz, m0, ev);
#pragma unroll
                    for (int q = 0; q < W; ++q) out[q] = ev[q] + out[q];
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < W; ++q) {
                T wv = T(0);
                if (bc == BC_DIRV) {
                    const int k = kk(q);
                    if (axis == 0) wv = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + k + 1] : L.wall[0][k + 1];
                    else wv = L.wall[1][(i + 1) * (nz + 2) + k + 1];
                }
                out[q] = bc_value(bc, c[q], T(0), wv);
            }
        }
    };
    side(xe, i + 1 < nx, i + 1 < nx ? L.sx : -(long)(nx - 1) * L.sx, L.bc_hi[0], 0, i + 1 < nx ? i + 1 : 0, j);
    side(xw, i > 0, i > 0 ? -L.sx : (long)(nx - 1) * L.sx, L.bc_lo[0], 0, i > 0 ? i - 1 : nx - 1, j);
    if (DIM == 3) {
        side(yn, j + 1 < ny, j + 1 < ny ? L.sy : -(long)(ny - 1) * L.sy, L.bc_hi[1], 1, i, j + 1 < ny ? j + 1 : 0);
        side(ys, j > 0, j > 0 ? -L.sy : (long)(ny - 1) * L.sy, L.bc_lo[1], 1, i, j > 0 ? j - 1 : ny - 1);
    }
    // the z value beyond the vector, with the correction when the red pass of
    // a prolongation sweep reads it (coarse cell K): prefetched (ZPRE) or read here
    auto zread = [&](long off, int K) -> T {
        const T v = Z ? T(0) : x[off];
        return (PRO && colour == 0) ? coarse_at(e, ebase, ehz, K) + v : v;
    };
    auto zfix = [&](long off, int K) -> T {
        if (ZPRE) return (PRO && colour == 0) ? coarse_at(e, ebase, ehz, zK) + zx : zx;
        return zread(off, K);
    };
    // z neighbours: k - 1 and k + 1 lie in the other half.  h == 0 (k even):
    // other[m - 1], other[m]; h == 1 (k odd): other[m], other[m + 1].  The
    // element beyond this lane's vector comes from the adjacent lane when it
    // holds the same row.
    const T from_prev = __shfl_up(o[W - 1], 1, 64);
    const T from_next = __shfl_down(o[0], 1, 64);
    T zlo0 = T(0), zhi_last = T(0);
    if (h == 0) {
        if (m0 == 0) {
            if (L.bc_lo[2] == BC_PER) zlo0 = zfix(rb + L.hz + (half - 1), half - 1);
            else {
                const T wv = L.bc_lo[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                   : T(0);
                zlo0 = bc_value(L.bc_lo[2], c[0], T(0), wv);
            }
        } else {
            zlo0 = lane > 0 ? from_prev : zfix(oth - 1, m0 - 1);
        }
    } else {
        if (m0 + W == half) {
            if (L.bc_hi[2] == BC_PER) zhi_last = zfix(rb, 0);
            else {
                const T wv = L.bc_hi[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                   : T(0);
                zhi_last = bc_value(L.bc_hi[2], c[W - 1], T(0), wv);
            }
        } else {
            zhi_last = lane < 63 ? from_next : zfix(oth + W, m0 + W);
        }
    }
    const T c2x = L.c2[0][i], c1x = L.c1[0][i];
    const T c2y = DIM == 3 ? L.c2[1][j] : T(0), c1y = DIM == 3 ? L.c1[1][j] : T(0);
    T r[W];
    bool zero = false;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        Stencil<T, DIM> s;
        s.c = c[q];
        s.xE = xe[q];
        s.xW = xw[q];
        if (DIM == 3) {
            s.yN = yn[q];
            s.yS = ys[q];
        }
        if (h == 0) {
            s.zB = q > 0 ? o[q - 1] : zlo0;
            s.zT = o[q];
        } else {
            s.zB = o[q];
            s.zT = q < W - 1 ? o[q + 1] : zhi_last;
        }
        r[q] = sor_quot<T, DIM>(s, c2x, c1x, c2y, c1y, VT::get(c2v, q), VT::get(c1v, q), VT::get(bv, q), w);
        zero = zero || r[q] == T(0);
    }
    if (W1 && zero && !need_c) {
        // a zero quotient takes the sign of (1 - w) x + q: read the cells now
        V lv{};
        if (!Z) lv = *reinterpret_cast<const V *>(x + own);
        if (PRO && !evo_now) coarse_vec(e, ebase, ehz, m0, evo);
#pragma unroll
        for (int q = 0; q < W; ++q) c[q] = PRO ? evo[q] + VT::get(lv, q) : VT::get(lv, q);
    }
#pragma unroll
    for (int q = 0; q < W; ++q)
        if (!W1 || r[q] == T(0)) r[q] = (T(1.0) - w) * c[q] + r[q];
    *reinterpret_cast<V *>(x + own) = VT::make(r);
    if (Z) *reinterpret_cast<V *>(x + oth) = V{};
}

template <typename T, int DIM, bool FINE, bool PRO, bool W1, bool Z = false>
__global__ void __launch_bounds__(kSplitThreads) k_rb_split(LevelArgs<T> L, T *__restrict__ x,
                                                            const T *__restrict__ b, int colour, T w, int nblocks,
                                                            const T *__restrict__ e, long esx, long esy, long ehz) {
    rb_split_body<T, DIM, PRO, false, W1, Z>(L, x, b, colour, w, nblocks, e, esx, esy, ehz);
}

template <typename T, int DIM, bool FINE, bool PRO, bool W1, bool Z = false>
__global__ void __launch_bounds__(kSplitThreads) k_rb_split_zp(LevelArgs<T> L, T *__restrict__ x,
                                                               const T *__restrict__ b, int colour, T w, int nblocks,
                                                               const T *__restrict__ e, long esx, long esy, long ehz) {
    rb_split_body<T, DIM, PRO, true, W1, Z>(L, x, b, colour, w, nblocks, e, esx, esy, ehz);
}

// ---------------------------------------------------------------------------
// residual of a split level restricted to the next level (the same
// computeResidual + coarsen expressions as k_resid_restrict,
// poisson_d3.cc:71-93, 484-511; poisson_d2.cc:288-297), and x(l+1) = 0.  A
// lane owns W consecutive coarse cells K0.. of coarse row (I, J): their fine
// children are the even and odd half-rows at K0.. of the 2 x 2 child rows.
// RED: the coarse level's first pre-smoothing red pass is applied as well.
// Its input iterate is the zero just written, so each red cell's update reads
// only zeros (wall rules of zero give zero) and its own new right-hand side:
// the same sor_update expression, evaluated here instead of in a launch of
// its own.
// ---------------------------------------------------------------------------
// Coarse cells per lane: 2 in FP32 (4 measured slower); 1 in FP64, where two per lane need
// more than 256 registers.
template <typename T>
constexpr int rr_width() {
    return sizeof(T) == 4 ? 2 : 1;
}

template <typename T, int DIM, bool RED = false>
__global__ void __launch_bounds__(kSplitThreads) k_rr_split(LevelArgs<T> F, const T *__restrict__ x,
                                                            const T *__restrict__ b, LevelArgs<T> C,
                                                            T *__restrict__ bc, T *__restrict__ xc, RestrictArgs<T> R,
                                                            int nblocks, T w) {
    constexpr int W = rr_width<T>();
    const int per = gridDim.x / kXcds;
    const int logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (logical >= nblocks) return;
    const int nx = F.n[0], ny = F.n[1], nz = F.n[2];
    const int half = nz >> 1, q = half / W;
    const int cy = C.n[1];
    const long total = (long)C.n[0] * cy * q;
    const long t = (long)logical * kSplitThreads + threadIdx.x;
    if (t >= total) return;
    const int crow = (int)(t / q);
    const int v = (int)(t - (long)crow * q);
    const int I = crow / cy, J = crow - (crow / cy) * cy;
    const int K0 = v * W;
    const long hz = F.hz;
    const int lane = threadIdx.x & 63;
    auto ldv = [&](const T *p, T (&out)[W]) { load_n<T, W>(p, out); };
    // residuals of the children: [di][dj][dk][e], dk = 0 the even cell 2K, 1 the odd cell 2K+1
    T res[2][DIM == 3 ? 2 : 1][2][W];
    const T c2e0 = T(0);
    (void)c2e0;
    T cz2e[W], cz1e[W], cz2o[W], cz1o[W];
    ldv(F.c2zs + K0, cz2e);
    ldv(F.c1zs + K0, cz1e);
    ldv(F.c2zs + half + K0, cz2o);
    ldv(F.c1zs + half + K0, cz1o);
#pragma unroll
    for (int di = 0; di < 2; ++di) {
#pragma unroll
        for (int dj = 0; dj < (DIM == 3 ? 2 : 1); ++dj) {
            const int i = 2 * I + di, j = DIM == 3 ? 2 * J + dj : 0;
            const long rb = (long)i * F.sx + (long)j * F.sy;
            T xe[W], xo[W], be[W], bo[W];
            ldv(x + rb + K0, xe);
            ldv(x + rb + hz + K0, xo);
            ldv(b + rb + K0, be);
            ldv(b + rb + hz + K0, bo);
            // neighbour rows: [0] E, [1] W, [2] N, [3] S; even and odd halves
            T ne[4][W], no[4][W];
            auto nbr = [&](int slot, bool inside, long off, int bc, int axis) {
                if (inside || bc == BC_PER) {
                    ldv(x + rb + off + K0, ne[slot]);
                    ldv(x + rb + off + hz + K0, no[slot]);
                } else {
#pragma unroll
                    for (int e = 0; e < W; ++e) {
                        T wve = T(0), wvo = T(0);
                        if (bc == BC_DIRV) {
                            const int ke = 2 * (K0 + e), ko = ke + 1;
                            if (axis == 0) {
                                wve = DIM == 3 ? F.wall[0][(j + 1) * (nz + 2) + ke + 1] : F.wall[0][ke + 1];
                                wvo = DIM == 3 ? F.wall[0][(j + 1) * (nz + 2) + ko + 1] : F.wall[0][ko + 1];
                            } else {
                                wve = F.wall[1][(i + 1) * (nz + 2) + ke + 1];
                                wvo = F.wall[1][(i + 1) * (nz + 2) + ko + 1];
                            }
                        }
                        ne[slot][e] = bc_value(bc, xe[e], T(0), wve);
                        no[slot][e] = bc_value(bc, xo[e], T(0), wvo);
                    }
                }
            };
            nbr(0, i + 1 < nx, i + 1 < nx ? F.sx : -(long)(nx - 1) * F.sx, F.bc_hi[0], 0);
            nbr(1, i > 0, i > 0 ? -F.sx : (long)(nx - 1) * F.sx, F.bc_lo[0], 0);
            if (DIM == 3) {
                nbr(2, j + 1 < ny, j + 1 < ny ? F.sy : -(long)(ny - 1) * F.sy, F.bc_hi[1], 1);
                nbr(3, j > 0, j > 0 ? -F.sy : (long)(ny - 1) * F.sy, F.bc_lo[1], 1);
            }
            // z: even cell 2K has zB = odd[K-1], zT = odd[K]; odd cell has zB = even[K], zT = even[K+1]
            const T from_prev = __shfl_up(xo[W - 1], 1, 64);
            const T from_next = __shfl_down(xe[0], 1, 64);
            T zlo0, zhi_last;
            if (K0 == 0) {
                if (F.bc_lo[2] == BC_PER) zlo0 = x[rb + hz + (half - 1)];
                else {
                    const T wv = F.bc_lo[2] == BC_DIRV ? (DIM == 3 ? F.wall[2][(i + 1) * (ny + 2) + j + 1] : F.wall[2][i + 1])
                                                       : T(0);
                    zlo0 = bc_value(F.bc_lo[2], xe[0], T(0), wv);
                }
            } else {
                zlo0 = lane > 0 ? from_prev : x[rb + hz + K0 - 1];
            }
            if (K0 + W == half) {
                if (F.bc_hi[2] == BC_PER) zhi_last = x[rb];
                else {
                    const T wv = F.bc_hi[2] == BC_DIRV ? (DIM == 3 ? F.wall[2][(i + 1) * (ny + 2) + j + 1] : F.wall[2][i + 1])
                                                       : T(0);
                    zhi_last = bc_value(F.bc_hi[2], xo[W - 1], T(0), wv);
                }
            } else {
                zhi_last = lane < 63 ? from_next : x[rb + K0 + W];
            }
            const T c2x = F.c2[0][i], c1x = F.c1[0][i];
            const T c2y = DIM == 3 ? F.c2[1][j] : T(0), c1y = DIM == 3 ? F.c1[1][j] : T(0);
#pragma unroll
            for (int e = 0; e < W; ++e) {
                Stencil<T, DIM> s;
                s.c = xe[e];
                s.xE = ne[0][e];
                s.xW = ne[1][e];
                if (DIM == 3) {
                    s.yN = ne[2][e];
                    s.yS = ne[3][e];
                }
                s.zB = e > 0 ? xo[e - 1] : zlo0;
                s.zT = xo[e];
                res[di][dj][0][e] = residual<T, DIM>(s, c2x, c1x, c2y, c1y, cz2e[e], cz1e[e], be[e]);
                s.c = xo[e];
                s.xE = no[0][e];
                s.xW = no[1][e];
                if (DIM == 3) {
                    s.yN = no[2][e];
                    s.yS = no[3][e];
                }
                s.zB = xe[e];
                s.zT = e < W - 1 ? xe[e + 1] : zhi_last;
                res[di][dj][1][e] = residual<T, DIM>(s, c2x, c1x, c2y, c1y, cz2o[e], cz1o[e], bo[e]);
            }
        }
    }
    T out[W], zero[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        zero[e] = T(0);
        if (DIM == 3) {
            const T lx = R.lo[0][I], hx = R.hi[0][I], ly = R.lo[1][J], hy = R.hi[1][J];
            const T lz = R.lo[2][K0 + e], hz2 = R.hi[2][K0 + e];
            const T v000 = lx * ly * lz, v100 = hx * ly * lz, v010 = lx * hy * lz, v001 = lx * ly * hz2;
            const T v011 = lx * hy * hz2, v101 = hx * ly * hz2, v110 = hx * hy * lz, v111 = hx * hy * hz2;
            const T vtot = R.wd[0][I] * R.wd[1][J] * R.wd[2][K0 + e];
            out[e] = (v111 * res[1][DIM == 3 ? 1 : 0][1][e] + v000 * res[0][0][0][e] +
                      v110 * res[1][DIM == 3 ? 1 : 0][0][e] + v001 * res[0][0][1][e] + v101 * res[1][0][1][e] +
                      v010 * res[0][DIM == 3 ? 1 : 0][0][e] + v011 * res[0][DIM == 3 ? 1 : 0][1][e] +
                      v100 * res[1][0][0][e]) /
                     vtot;
        } else {
            out[e] = (res[1][0][1][e] + res[0][0][0][e] + res[1][0][0][e] + res[0][0][1][e]) / T(4);
        }
    }
    const long cb = (long)I * C.sx + (long)J * C.sy;
    coarse_store_n<T, W>(bc, cb, C.hz, K0, out);
    if (RED) {
        T xr[W];
        const T c2x = C.c2[0][I], c1x = C.c1[0][I];
        const T c2y = DIM == 3 ? C.c2[1][J] : T(0), c1y = DIM == 3 ? C.c1[1][J] : T(0);
#pragma unroll
        for (int e = 0; e < W; ++e) {
            const int K = K0 + e;
            if (((I + J + K) & 1) == 0) {
                Stencil<T, DIM> s;
                s.c = T(0);
                s.xE = bc_value(C.bc_hi[0], T(0), T(0), T(0));
                s.xW = bc_value(C.bc_lo[0], T(0), T(0), T(0));
                if (DIM == 3) {
                    s.yN = bc_value(C.bc_hi[1], T(0), T(0), T(0));
                    s.yS = bc_value(C.bc_lo[1], T(0), T(0), T(0));
                }
                s.zT = bc_value(C.bc_hi[2], T(0), T(0), T(0));
                s.zB = bc_value(C.bc_lo[2], T(0), T(0), T(0));
                xr[e] = sor_update<T, DIM>(s, c2x, c1x, c2y, c1y, C.c2[2][K], C.c1[2][K], out[e], w);
            } else {
                xr[e] = T(0);
            }
        }
        coarse_store_n<T, W>(xc, cb, C.hz, K0, xr);
    } else {
        coarse_store_n<T, W>(xc, cb, C.hz, K0, zero);
    }
}

// ---------------------------------------------------------------------------
// Residual + restriction of a split 3-D level with the eight children of a
// coarse cell spread over the four waves of a block.  A block owns 64
// consecutive (coarse row, vector) tasks; wave w = (di, dj) = (w / 2, w % 2)
// computes, for every task, the residuals of fine row (2I + di, 2J + dj) at
// the 2W fine cells under the task's W coarse cells -- whole 16-byte vectors,
// as in a colour pass -- and leaves the products v_{di dj dk} r in LDS.  After
// one barrier each thread adds the eight products of one coarse cell in the
// reference's order and divides by the volume (poisson_d3.cc:494-511).  Every
// product and partial sum is the same rounding as in the one-lane form, so
// the values are identical.  RED as in k_rr_split.
// ---------------------------------------------------------------------------
constexpr int kRrTasks = 64;

// position of v_{di dj dk} r_{di dj dk} in the reference's sum
// v111 r111 + v000 r000 + v110 r110 + v001 r001 + v101 r101 + v010 r010 + v011 r011 + v100 r100
__device__ __forceinline__ int rr_slot(int di, int dj, int dk) {
    switch (di * 4 + dj * 2 + dk) {   // codes 7, 0, 6, 1, 5, 2, 3, 4 in sum order
        case 7: return 0;
        case 0: return 1;
        case 6: return 2;
        case 1: return 3;
        case 5: return 4;
        case 2: return 5;
        case 3: return 6;
        default: return 7;
    }
}

template <typename T, bool RED>
__global__ void __launch_bounds__(kSplitThreads) k_rr_split4(LevelArgs<T> F, const T *__restrict__ x,
                                                             const T *__restrict__ b, LevelArgs<T> C,
                                                             T *__restrict__ bc, T *__restrict__ xc, RestrictArgs<T> R,
                                                             int nblocks, T w) {
    using VT = Vec16<T>;
    constexpr int W = VT::W;
    __shared__ __align__(16) T prod[8][kRrTasks * W];
    const int per = gridDim.x / kXcds;
    const int logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (logical >= nblocks) return;
    const int nx = F.n[0], ny = F.n[1], nz = F.n[2];
    const int half = nz >> 1, q = half / W;
    const int cy = C.n[1];
    const long ntask = (long)C.n[0] * cy * q;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int di = wave >> 1, dj = wave & 1;
    const long task = (long)logical * kRrTasks + lane;
    const bool live = task < ntask;
    const long hz = F.hz;

    int I = 0, J = 0, K0 = 0, i = 0, j = 0;
    long rb = 0;
    T xe[W], xo[W], zl = T(0), zh = T(0);
    for (int e = 0; e < W; ++e) xe[e] = xo[e] = T(0);
    if (live) {
        const long crow = task / q;
        const int v = (int)(task - crow * q);
        I = (int)(crow / cy);
        J = (int)(crow - (long)I * cy);
        K0 = v * W;
        i = 2 * I + di;
        j = 2 * J + dj;
        rb = (long)i * F.sx + (long)j * F.sy;
        load_n<T, W>(x + rb + K0, xe);
        load_n<T, W>(x + rb + hz + K0, xo);
        // z neighbours no adjacent lane holds, loaded with the row vectors
        if (K0 == 0 ? F.bc_lo[2] == BC_PER : lane == 0) zl = K0 == 0 ? x[rb + hz + (half - 1)] : x[rb + hz + K0 - 1];
        if (K0 + W == half ? F.bc_hi[2] == BC_PER : lane == 63) zh = K0 + W == half ? x[rb] : x[rb + K0 + W];
    }
    // cross-vector z neighbours from the adjacent lanes (all lanes take part;
    // a live lane's predecessor is live, and a live lane whose successor is
    // not ends its row, where the boundary rule applies instead)
    const T from_prev = __shfl_up(xo[W - 1], 1, 64);
    const T from_next = __shfl_down(xe[0], 1, 64);
    if (live) {
        T be[W], bo[W], cz2e[W], cz1e[W], cz2o[W], cz1o[W];
        load_n<T, W>(b + rb + K0, be);
        load_n<T, W>(b + rb + hz + K0, bo);
        load_n<T, W>(F.c2zs + K0, cz2e);
        load_n<T, W>(F.c1zs + K0, cz1e);
        load_n<T, W>(F.c2zs + half + K0, cz2o);
        load_n<T, W>(F.c1zs + half + K0, cz1o);
        // neighbour rows: [0] E, [1] W, [2] N, [3] S; even and odd halves
        T ne[4][W], no[4][W];
        auto nbr = [&](int slot, bool inside, long off, int bcode, int axis) {
            if (inside || bcode == BC_PER) {
                load_n<T, W>(x + rb + off + K0, ne[slot]);
                load_n<T, W>(x + rb + off + hz + K0, no[slot]);
            } else {
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    T wve = T(0), wvo = T(0);
                    if (bcode == BC_DIRV) {
                        const int ke = 2 * (K0 + e), ko = ke + 1;
                        if (axis == 0) {
                            wve = F.wall[0][(j + 1) * (nz + 2) + ke + 1];
                            wvo = F.wall[0][(j + 1) * (nz + 2) + ko + 1];
                        } else {
                            wve = F.wall[1][(i + 1) * (nz + 2) + ke + 1];
                            wvo = F.wall[1][(i + 1) * (nz + 2) + ko + 1];
                        }
                    }
                    ne[slot][e] = bc_value(bcode, xe[e], T(0), wve);
                    no[slot][e] = bc_value(bcode, xo[e], T(0), wvo);
                }
            }
        };
        nbr(0, i + 1 < nx, i + 1 < nx ? F.sx : -(long)(nx - 1) * F.sx, F.bc_hi[0], 0);
        nbr(1, i > 0, i > 0 ? -F.sx : (long)(nx - 1) * F.sx, F.bc_lo[0], 0);
        nbr(2, j + 1 < ny, j + 1 < ny ? F.sy : -(long)(ny - 1) * F.sy, F.bc_hi[1], 1);
        nbr(3, j > 0, j > 0 ? -F.sy : (long)(ny - 1) * F.sy, F.bc_lo[1], 1);
        // z: even cell 2K has zB = odd[K-1], zT = odd[K]; odd cell has zB = even[K], zT = even[K+1]
        T zlo0, zhi_last;
        if (K0 == 0) {
            if (F.bc_lo[2] == BC_PER) zlo0 = zl;
            else {
                const T wv = F.bc_lo[2] == BC_DIRV ? F.wall[2][(i + 1) * (ny + 2) + j + 1] : T(0);
                zlo0 = bc_value(F.bc_lo[2], xe[0], T(0), wv);
            }
        } else {
            zlo0 = lane > 0 ? from_prev : zl;
        }
        if (K0 + W == half) {
            if (F.bc_hi[2] == BC_PER) zhi_last = zh;
            else {
                const T wv = F.bc_hi[2] == BC_DIRV ? F.wall[2][(i + 1) * (ny + 2) + j + 1] : T(0);
                zhi_last = bc_value(F.bc_hi[2], xo[W - 1], T(0), wv);
            }
        } else {
            zhi_last = lane < 63 ? from_next : zh;
        }
        const T c2x = F.c2[0][i], c1x = F.c1[0][i], c2y = F.c2[1][j], c1y = F.c1[1][j];
        const T ax = di ? R.hi[0][I] : R.lo[0][I], ay = dj ? R.hi[1][J] : R.lo[1][J];
        const T axy = ax * ay;
        T lz[W], hzf[W];
        load_n<T, W>(R.lo[2] + K0, lz);
        load_n<T, W>(R.hi[2] + K0, hzf);
        T p0[W], p1[W];
#pragma unroll
        for (int e = 0; e < W; ++e) {
            Stencil<T, 3> s;
            s.c = xe[e];
            s.xE = ne[0][e];
            s.xW = ne[1][e];
            s.yN = ne[2][e];
            s.yS = ne[3][e];
            s.zB = e > 0 ? xo[e - 1] : zlo0;
            s.zT = xo[e];
            const T r0 = residual<T, 3>(s, c2x, c1x, c2y, c1y, cz2e[e], cz1e[e], be[e]);
            s.c = xo[e];
            s.xE = no[0][e];
            s.xW = no[1][e];
            s.yN = no[2][e];
            s.yS = no[3][e];
            s.zB = xe[e];
            s.zT = e < W - 1 ? xe[e + 1] : zhi_last;
            const T r1 = residual<T, 3>(s, c2x, c1x, c2y, c1y, cz2o[e], cz1o[e], bo[e]);
            p0[e] = (axy * lz[e]) * r0;
            p1[e] = (axy * hzf[e]) * r1;
        }
        *reinterpret_cast<typename VT::V *>(&prod[rr_slot(di, dj, 0)][lane * W]) = VT::make(p0);
        *reinterpret_cast<typename VT::V *>(&prod[rr_slot(di, dj, 1)][lane * W]) = VT::make(p1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < kRrTasks * W; c += blockDim.x) {
        const long tk = (long)logical * kRrTasks + c / W;
        if (tk >= ntask) continue;
        const long crow = tk / q;
        const int v = (int)(tk - crow * q);
        const int CI = (int)(crow / cy), CJ = (int)(crow - (long)CI * cy);
        const int K = v * W + c % W;
        T sum = prod[0][c] + prod[1][c];
        sum = sum + prod[2][c];
        sum = sum + prod[3][c];
        sum = sum + prod[4][c];
        sum = sum + prod[5][c];
        sum = sum + prod[6][c];
        sum = sum + prod[7][c];
        const T vtot = R.wd[0][CI] * R.wd[1][CJ] * R.wd[2][K];
        const T out = sum / vtot;
        const long cidx = (long)CI * C.sx + (long)CJ * C.sy + kof(C, K);
        bc[cidx] = out;
        T xr = T(0);
        if (RED && ((CI + CJ + K) & 1) == 0) {
            Stencil<T, 3> s;
            s.c = T(0);
            s.xE = bc_value(C.bc_hi[0], T(0), T(0), T(0));
            s.xW = bc_value(C.bc_lo[0], T(0), T(0), T(0));
            s.yN = bc_value(C.bc_hi[1], T(0), T(0), T(0));
            s.yS = bc_value(C.bc_lo[1], T(0), T(0), T(0));
            s.zT = bc_value(C.bc_hi[2], T(0), T(0), T(0));
            s.zB = bc_value(C.bc_lo[2], T(0), T(0), T(0));
            xr = sor_update<T, 3>(s, C.c2[0][CI], C.c1[0][CI], C.c2[1][CJ], C.c1[1][CJ], C.c2[2][K], C.c1[2][K], out, w);
        }
        xc[cidx] = xr;
    }
}

// ---------------------------------------------------------------------------
// prolongation by injection plus the saved iterate on a split level
// (poisson_d3.cc:548-558, poisson.cc:266): both halves of W fine positions
// m0.. of row (i, j) take the correction of coarse cells K = m0.. of coarse
// row (i/2, j/2).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kSplitThreads) k_prolong_split(int nx, int ny, int nz, long sx, long sy, long hz,
                                                                 T *__restrict__ x, const T *__restrict__ e,
                                                                 long csx, long csy, long chz, int nblocks) {
    using VT = Vec16<T>;
    using V = typename VT::V;
    constexpr int W = VT::W;
    const int per = gridDim.x / kXcds;
    const int logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (logical >= nblocks) return;
    const int half = nz >> 1, q = half / W;
    const long total = (long)nx * ny * q;
    const long t = (long)logical * kSplitThreads + threadIdx.x;
    if (t >= total) return;
    const int row = (int)(t / q);
    const int v = (int)(t - (long)row * q);
    const int i = row / ny, j = row - (row / ny) * ny;
    const int m0 = v * W;
    T ev[W];
    coarse_vec(e, (long)(i >> 1) * csx + (long)(j >> 1) * csy, chz, m0, ev);
    const long rb = (long)i * sx + (long)j * sy;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        V *p = reinterpret_cast<V *>(x + rb + hh * hz + m0);
        const V xv = *p;
        T r[W];
#pragma unroll
        for (int q2 = 0; q2 < W; ++q2) r[q2] = ev[q2] + VT::get(xv, q2);
        *p = VT::make(r);
    }
}

// ---------------------------------------------------------------------------
// FP64 refinement step of the mixed-precision solve on split rows (the same
// expressions as k_refine): x_new = x_old + e, r = b - A x_new with the wall
// rules applied to x_new, r32 = (float) r, per-block norm partials of r and
// b.  A lane owns W = 2 positions m0.. of both halves of one row, i.e. the
// four cells k = 2 m0 .. 2 m0 + 3; FP64 arrays use the FP64 level's split
// rows (hz), the FP32 correction and residual the FP32 level's (ehz).
// ---------------------------------------------------------------------------
// NORM_ONLY: the residual norm partials of computeError alone
// (poisson_d3.cc:563-663): no correction, nothing written but the partials.
// CX / CB: x_old / b are read from the caller's padded array (FullRows: the
// core cell (i, j, k) at off + i si + j sj + k) instead of the level's split
// rows, so the solve needs no separate layout conversion of its inputs; the
// values, and so the results, are the same.  CXO: x_new is written into the
// caller's array (xo) instead of the split rows -- the refinement expected
// to be the last, which then needs no layout conversion of its output either
// (never together with CX: x_old and x_new are then the same array).
struct FullRows {
    const double *x = nullptr, *b = nullptr;
    double *xo = nullptr;
    long si = 0, sj = 0, off = 0;
};

// Four consecutive doubles of a caller's row, whose start need not be
// 16-byte aligned (an odd pad puts the core at an odd element): two 16-byte
// accesses when it is, else 8 + 16 + 8 bytes.  The parity is the same along
// a row, so the lanes of a wave take one branch.
__device__ __forceinline__ void load4(const double *__restrict__ p, double (&v)[4]) {
    if ((reinterpret_cast<unsigned long>(p) & 15) == 0) {
        const double2 a = *reinterpret_cast<const double2 *>(p), c = *reinterpret_cast<const double2 *>(p + 2);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = c.x;
        v[3] = c.y;
    } else {
        const double2 m = *reinterpret_cast<const double2 *>(p + 1);
        v[0] = p[0];
        v[1] = m.x;
        v[2] = m.y;
        v[3] = p[3];
    }
}
__device__ __forceinline__ void store4(double *__restrict__ p, double v0, double v1, double v2, double v3) {
    if ((reinterpret_cast<unsigned long>(p) & 15) == 0) {
        *reinterpret_cast<double2 *>(p) = make_double2(v0, v1);
        *reinterpret_cast<double2 *>(p + 2) = make_double2(v2, v3);
    } else {
        p[0] = v0;
        *reinterpret_cast<double2 *>(p + 1) = make_double2(v1, v2);
        p[3] = v3;
    }
}

// NOXW: r32 and the norm only, x_new not written (the initial residual of a
// solve whose iterate stays in the caller's array until the first refinement).
template <int TYPE, int DIM, bool NORM_ONLY = false, bool CX = false, bool CB = false, bool CXO = false,
          bool NOXW = false>
__global__ void __launch_bounds__(kNormThreads) k_refine_split(LevelArgs<double> L, const double *__restrict__ xold,
                                                               const float *__restrict__ e, long esx, long esy,
                                                               long ehz, const double *__restrict__ b,
                                                               double *__restrict__ xnew, float *__restrict__ r32,
                                                               double *__restrict__ partial, FullRows F) {
    static_assert(!(CX && CXO), "x_old and x_new cannot both be the caller's array");
    constexpr int W = 2;
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int half = nz >> 1, q = half / W;
    const long total = (long)nx * ny * q;
    const long hz = L.hz;
    const int lane = threadIdx.x & 63;
    double ar = 0.0, ab = 0.0;
    // rows of the x_old + e iterate: even and odd halves at m0
    auto row_vals = [&](int ii, int jj, int m0, double (&ve)[W], double (&vo)[W]) {
        if (CX) {
            double xr[4];
            load4(F.x + F.off + (long)ii * F.si + (long)jj * F.sj + 2 * m0, xr);
            ve[0] = xr[0];
            vo[0] = xr[1];
            ve[1] = xr[2];
            vo[1] = xr[3];
            if (e) {
                const long eb = (long)ii * esx + (long)jj * esy;
                const float2 fe = *reinterpret_cast<const float2 *>(e + eb + m0);
                const float2 fo = *reinterpret_cast<const float2 *>(e + eb + ehz + m0);
                ve[0] = ve[0] + (double)fe.x;
                ve[1] = ve[1] + (double)fe.y;
                vo[0] = vo[0] + (double)fo.x;
                vo[1] = vo[1] + (double)fo.y;
            }
            return;
        }
        const long rb = (long)ii * L.sx + (long)jj * L.sy;
        const double2 xe = *reinterpret_cast<const double2 *>(xold + rb + m0);
        const double2 xo = *reinterpret_cast<const double2 *>(xold + rb + hz + m0);
        ve[0] = xe.x;
        ve[1] = xe.y;
        vo[0] = xo.x;
        vo[1] = xo.y;
        if (e) {
            const long eb = (long)ii * esx + (long)jj * esy;
            const float2 fe = *reinterpret_cast<const float2 *>(e + eb + m0);
            const float2 fo = *reinterpret_cast<const float2 *>(e + eb + ehz + m0);
            ve[0] = ve[0] + (double)fe.x;
            ve[1] = ve[1] + (double)fe.y;
            vo[0] = vo[0] + (double)fo.x;
            vo[1] = vo[1] + (double)fo.y;
        }
    };
    // Rows are walked in one of two orders.  Where a level's neighbour
    // planes exceed half an XCD's L2 (L.rtile, set for the colour passes too),
    // each XCD walks one contiguous range of j-tiled rows, its blocks striding
    // through it together, so a row's x and y neighbours were fetched into
    // the same L2 a few rows before; otherwise the grid strides over all rows
    // in order (the level's x and b are then served from the Infinity Cache
    // and this order measured faster).  The grid is a multiple of kXcds.
    const bool xwalk = L.rtile != 0;
    const int per = xwalk ? gridDim.x / kXcds : gridDim.x;
    const int xcd = xwalk ? blockIdx.x % kXcds : 0, kb = xwalk ? blockIdx.x / kXcds : blockIdx.x;
    const long rows = (long)nx * ny;
    const long xrows = xwalk ? (rows + kXcds - 1) / kXcds : rows;
    const long lo = (long)xcd * xrows * q, hi = min(total, lo + xrows * q);
    for (long t = lo + (long)kb * blockDim.x + threadIdx.x; t < hi; t += (long)per * blockDim.x) {
        const int row = (int)(t / q);
        const int v = (int)(t - (long)row * q);
        int i, j;
        if (xwalk) {
            const int jt = L.rtile, tile = row / (nx * jt), rem = row - tile * (nx * jt);
            i = rem >> __builtin_ctz(jt);
            j = tile * jt + (rem & (jt - 1));
        } else {
            i = row / ny;
            j = row - i * ny;
        }
        const int m0 = v * W;
        const long rb = (long)i * L.sx + (long)j * L.sy;
        const long eb = (long)i * esx + (long)j * esy;
        // the iterate at cell k of this row
        auto cell = [&](int k) {
            const long h = k & 1, m = k >> 1;
            double x = CX ? F.x[F.off + (long)i * F.si + (long)j * F.sj + k] : xold[rb + h * hz + m];
            if (e) x = x + (double)e[eb + h * ehz + m];
            return x;
        };
        // the z neighbours beyond this lane's vectors that no adjacent lane
        // holds (periodic wrap, or the row's vector in another wave), loaded
        // first so that their latency overlaps the row loads
        const bool need_lo = m0 == 0 ? L.bc_lo[2] == BC_PER : lane == 0;
        const bool need_hi = m0 + W == half ? L.bc_hi[2] == BC_PER : lane == 63;
        double zl = 0.0, zh = 0.0;
        if (need_lo) zl = cell(m0 == 0 ? nz - 1 : 2 * m0 - 1);
        if (need_hi) zh = cell(m0 + W == half ? 0 : 2 * (m0 + W));
        double ce[W], co[W];
        row_vals(i, j, m0, ce, co);
        // x/y neighbours: [0] E, [1] W, [2] N, [3] S
        double ne[4][W], no[4][W];
        auto nbr = [&](int slot, bool inside, int ii, int jj, int bc, int axis) {
            if (inside || bc == BC_PER) {
                row_vals(ii, jj, m0, ne[slot], no[slot]);
            } else {
#pragma unroll
                for (int p = 0; p < W; ++p) {
                    double wve = 0.0, wvo = 0.0;
                    if (bc == BC_DIRV) {
                        const int ke = 2 * (m0 + p), ko = ke + 1;
                        if (axis == 0) {
                            wve = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + ke + 1] : L.wall[0][ke + 1];
                            wvo = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + ko + 1] : L.wall[0][ko + 1];
                        } else {
                            wve = L.wall[1][(i + 1) * (nz + 2) + ke + 1];
                            wvo = L.wall[1][(i + 1) * (nz + 2) + ko + 1];
                        }
                    }
                    ne[slot][p] = bc_value<double>(bc, ce[p], 0.0, wve);
                    no[slot][p] = bc_value<double>(bc, co[p], 0.0, wvo);
                }
            }
        };
        nbr(0, i + 1 < nx, i + 1 < nx ? i + 1 : 0, j, L.bc_hi[0], 0);
        nbr(1, i > 0, i > 0 ? i - 1 : nx - 1, j, L.bc_lo[0], 0);
        if (DIM == 3) {
            nbr(2, j + 1 < ny, i, j + 1 < ny ? j + 1 : 0, L.bc_hi[1], 1);
            nbr(3, j > 0, i, j > 0 ? j - 1 : ny - 1, L.bc_lo[1], 1);
        }
        // z: even cell 2m: zB = odd[m-1], zT = odd[m]; odd cell 2m+1: zB = even[m], zT = even[m+1]
        const double from_prev = __shfl_up(co[W - 1], 1, 64);
        const double from_next = __shfl_down(ce[0], 1, 64);
        double zlo0, zhi_last;
        if (m0 == 0) {
            if (L.bc_lo[2] == BC_PER) zlo0 = zl;
            else {
                const double wv = L.bc_lo[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                        : 0.0;
                zlo0 = bc_value<double>(L.bc_lo[2], ce[0], 0.0, wv);
            }
        } else {
            zlo0 = lane > 0 ? from_prev : zl;
        }
        if (m0 + W == half) {
            if (L.bc_hi[2] == BC_PER) zhi_last = zh;
            else {
                const double wv = L.bc_hi[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                        : 0.0;
                zhi_last = bc_value<double>(L.bc_hi[2], co[W - 1], 0.0, wv);
            }
        } else {
            zhi_last = lane < 63 ? from_next : zh;
        }
        const double c2x = L.c2[0][i], c1x = L.c1[0][i];
        const double c2y = DIM == 3 ? L.c2[1][j] : 0.0, c1y = DIM == 3 ? L.c1[1][j] : 0.0;
        double bev[W], bov[W];
        if (CB) {
            double br[4];
            load4(F.b + F.off + (long)i * F.si + (long)j * F.sj + 2 * m0, br);
            bev[0] = br[0];
            bov[0] = br[1];
            bev[1] = br[2];
            bov[1] = br[3];
        } else {
            const double2 be = *reinterpret_cast<const double2 *>(b + rb + m0);
            const double2 bo = *reinterpret_cast<const double2 *>(b + rb + hz + m0);
            bev[0] = be.x;
            bev[1] = be.y;
            bov[0] = bo.x;
            bov[1] = bo.y;
        }
        double re[W], ro[W];
#pragma unroll
        for (int p = 0; p < W; ++p) {
            Stencil<double, DIM> s;
            s.c = ce[p];
            s.xE = ne[0][p];
            s.xW = ne[1][p];
            if (DIM == 3) {
                s.yN = ne[2][p];
                s.yS = ne[3][p];
            }
            s.zB = p > 0 ? co[p - 1] : zlo0;
            s.zT = co[p];
            re[p] = residual<double, DIM>(s, c2x, c1x, c2y, c1y, L.c2zs[m0 + p], L.c1zs[m0 + p], bev[p]);
            s.c = co[p];
            s.xE = no[0][p];
            s.xW = no[1][p];
            if (DIM == 3) {
                s.yN = no[2][p];
                s.yS = no[3][p];
            }
            s.zB = ce[p];
            s.zT = p < W - 1 ? ce[p + 1] : zhi_last;
            ro[p] = residual<double, DIM>(s, c2x, c1x, c2y, c1y, L.c2zs[half + m0 + p], L.c1zs[half + m0 + p], bov[p]);
        }
        if (!NORM_ONLY) {
            if (NOXW) {
            } else if (CXO) {
                store4(F.xo + F.off + (long)i * F.si + (long)j * F.sj + 2 * m0, ce[0], co[0], ce[1], co[1]);
            } else {
                *reinterpret_cast<double2 *>(xnew + rb + m0) = make_double2(ce[0], ce[1]);
                *reinterpret_cast<double2 *>(xnew + rb + hz + m0) = make_double2(co[0], co[1]);
            }
            *reinterpret_cast<float2 *>(r32 + eb + m0) = make_float2((float)re[0], (float)re[1]);
            *reinterpret_cast<float2 *>(r32 + eb + ehz + m0) = make_float2((float)ro[0], (float)ro[1]);
        }
#pragma unroll
        for (int p = 0; p < W; ++p) {
            ar = norm_acc<TYPE>(ar, re[p]);
            ab = norm_acc<TYPE>(ab, bev[p]);
            ar = norm_acc<TYPE>(ar, ro[p]);
            ab = norm_acc<TYPE>(ab, bov[p]);
        }
    }
    block_reduce2<TYPE>(ar, ab);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ar;
        partial[2 * blockIdx.x + 1] = ab;
    }
}

}  // namespace saras
