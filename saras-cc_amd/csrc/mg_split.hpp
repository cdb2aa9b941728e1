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

// one colour pass over a split level; the grid is a multiple of kXcds
// blocks, `nblocks` of them carry work
template <typename T, int DIM, bool FINE>
__global__ void __launch_bounds__(kSplitThreads) k_rb_split(LevelArgs<T> L, T *__restrict__ x,
                                                            const T *__restrict__ b, int colour, T w, int nblocks) {
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
    const int i = row / ny, j = row - (row / ny) * ny;
    const int h = (i + j + colour) & 1;   // half of this row holding the colour
    const int m0 = v * W;
    const long rb = (long)i * L.sx + (long)j * L.sy;
    const long own = rb + h * L.hz + m0;
    const long oth = rb + (1 - h) * L.hz + m0;
    const V cv = *reinterpret_cast<const V *>(x + own);
    const V ov = *reinterpret_cast<const V *>(x + oth);
    const V bv = *reinterpret_cast<const V *>(b + own);
    const V c2v = *reinterpret_cast<const V *>(L.c2zs + h * half + m0);
    const V c1v = *reinterpret_cast<const V *>(L.c1zs + h * half + m0);
    T c[W], o[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        c[e] = VT::get(cv, e);
        o[e] = VT::get(ov, e);
    }
    // x and y neighbour rows: whole vectors, or the wall rule per cell
    auto kk = [&](int e) { return 2 * (m0 + e) + h; };
    T xe[W], xw[W], yn[W], ys[W];
    auto side = [&](T(&out)[W], bool inside, long off, int bc, int axis) {
        if (inside || bc == BC_PER) {
            const V nv = *reinterpret_cast<const V *>(x + own + off);
#pragma unroll
            for (int e = 0; e < W; ++e) out[e] = VT::get(nv, e);
        } else {
#pragma unroll
            for (int e = 0; e < W; ++e) {
                T wv = T(0);
                if (bc == BC_DIRV) {
                    const int k = kk(e);
                    if (axis == 0) wv = DIM == 3 ? L.wall[0][(j + 1) * (nz + 2) + k + 1] : L.wall[0][k + 1];
                    else wv = L.wall[1][(i + 1) * (nz + 2) + k + 1];
                }
                out[e] = bc_value(bc, c[e], T(0), wv);
            }
        }
    };
    side(xe, i + 1 < nx, i + 1 < nx ? L.sx : -(long)(nx - 1) * L.sx, L.bc_hi[0], 0);
    side(xw, i > 0, i > 0 ? -L.sx : (long)(nx - 1) * L.sx, L.bc_lo[0], 0);
    if (DIM == 3) {
        side(yn, j + 1 < ny, j + 1 < ny ? L.sy : -(long)(ny - 1) * L.sy, L.bc_hi[1], 1);
        side(ys, j > 0, j > 0 ? -L.sy : (long)(ny - 1) * L.sy, L.bc_lo[1], 1);
    }
    // z neighbours: k - 1 and k + 1 lie in the other half.  h == 0 (k even):
    // other[m - 1], other[m]; h == 1 (k odd): other[m], other[m + 1].  The
    // element beyond this lane's vector comes from the adjacent lane when it
    // holds the same row.
    const int lane = threadIdx.x & 63;
    const T from_prev = __shfl_up(o[W - 1], 1, 64);
    const T from_next = __shfl_down(o[0], 1, 64);
    T zlo0 = T(0), zhi_last = T(0);
    if (h == 0) {
        if (m0 == 0) {
            if (L.bc_lo[2] == BC_PER) zlo0 = x[rb + L.hz + (half - 1)];
            else {
                const T wv = L.bc_lo[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                   : T(0);
                zlo0 = bc_value(L.bc_lo[2], c[0], T(0), wv);
            }
        } else {
            zlo0 = lane > 0 ? from_prev : x[oth - 1];
        }
    } else {
        if (m0 + W == half) {
            if (L.bc_hi[2] == BC_PER) zhi_last = x[rb];
            else {
                const T wv = L.bc_hi[2] == BC_DIRV ? (DIM == 3 ? L.wall[2][(i + 1) * (ny + 2) + j + 1] : L.wall[2][i + 1])
                                                   : T(0);
                zhi_last = bc_value(L.bc_hi[2], c[W - 1], T(0), wv);
            }
        } else {
            zhi_last = lane < 63 ? from_next : x[oth + W];
        }
    }
    const T c2x = L.c2[0][i], c1x = L.c1[0][i];
    const T c2y = DIM == 3 ? L.c2[1][j] : T(0), c1y = DIM == 3 ? L.c1[1][j] : T(0);
    T r[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        Stencil<T, DIM> s;
        s.c = c[e];
        s.xE = xe[e];
        s.xW = xw[e];
        if (DIM == 3) {
            s.yN = yn[e];
            s.yS = ys[e];
        }
        if (h == 0) {
            s.zB = e > 0 ? o[e - 1] : zlo0;
            s.zT = o[e];
        } else {
            s.zB = o[e];
            s.zT = e < W - 1 ? o[e + 1] : zhi_last;
        }
        r[e] = sor_update<T, DIM>(s, c2x, c1x, c2y, c1y, VT::get(c2v, e), VT::get(c1v, e), VT::get(bv, e), w);
    }
    *reinterpret_cast<V *>(x + own) = VT::make(r);
}

}  // namespace saras
