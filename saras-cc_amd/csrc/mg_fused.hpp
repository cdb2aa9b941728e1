// mg_fused.hpp -- the fused red-black sweep and the FP64 refinement kernel.
//
// k_sweep_march: one launch performs a whole RB-SOR sweep (red pass then
// black pass, poisson_d3.cc:99-243) out of place: x_out = sweep(x_in).
//   red value   R(c) = sor(c; x_in of its (black) neighbours)
//   black value      = sor(c; R of its (red) neighbours, own x_in)
// which is exactly what the two in-place passes produce, so the result is
// bit-identical to the reference ordering.  Each workgroup owns a TJ x TK
// tile of (j, k) columns and marches along i through a chunk of planes,
// keeping a 3-plane ring of x_in (tile + 2-cell halo) and a 3-plane ring of
// half-updated planes H (red cells = R, black cells = x_in; tile + 1-cell
// halo) in LDS.  The red values of the halo ring are recomputed redundantly
// instead of being exchanged, so workgroups never wait on each other.  HBM
// sees each x_in, b and x_out word once per sweep; halo re-reads hit L2.
// Workgroups are ordered so that consecutive tiles land on one XCD (the
// dispatcher deals workgroup ids round-robin over the 8 XCDs) and share
// their halos in that XCD's L2.
//
// Requirements checked by the host: 3-D, every periodic extent even.
#pragma once
#include "mg_kernels.hpp"

namespace saras {

template <typename T>
__device__ __forceinline__ T wall_val(const LevelArgs<T> &L, int axis, int i, int j, int k) {
    const int ny = L.n[1], nz = L.n[2];
    if (axis == 0) return L.wall[0][(j + 1) * (nz + 2) + k + 1];
    if (axis == 1) return L.wall[1][(i + 1) * (nz + 2) + k + 1];
    return L.wall[2][(i + 1) * (ny + 2) + j + 1];
}

// neighbour across a wall (non-periodic boundary) of a cell whose own value is `own`
template <typename T>
__device__ __forceinline__ T wall_nb(const LevelArgs<T> &L, int bc, int axis, T own, int i, int j, int k) {
    if (bc == BC_NEU) return own;
    if (bc == BC_DIR) return -own;
    return T(2.0) * wall_val(L, axis, i, j, k) - own;
}

template <typename T>
struct Pair;
template <>
struct Pair<float> {
    using V = float2;
};
template <>
struct Pair<double> {
    using V = double2;
};

// One z-pair (k, k+1), k even, of a (j, k) region around the tile: wrapped
// j and first k, and its in-plane offset (-1 when it lies beyond a wall).
struct PairElem {
    int goff;
    int row, pc;  // region row and pair column
    int j, kps;
};

template <int PX>
__device__ __forceinline__ PairElem pair_elem(int e, int j_first, int k_first, int ny, int nz, bool pery, bool perz,
                                              long sy) {
    PairElem r;
    r.row = e / PX;
    r.pc = e - r.row * PX;
    int j = j_first + r.row, k = k_first + 2 * r.pc;
    bool ok = true;
    if (j < 0 || j >= ny) { if (pery) j = j < 0 ? j + ny : j - ny; else ok = false; }
    if (k < 0 || k >= nz) { if (perz) k = k < 0 ? k + nz : k - nz; else ok = false; }
    r.j = j;
    r.kps = k;
    r.goff = ok ? (int)((long)j * sy + k) : -1;
    return r;
}

// Fused out-of-place RB-SOR sweep.  Every lane owns z-pairs (k, k+1): one
// red and one black cell, so both colour updates keep all lanes busy.
// Regions in LDS (pair-aligned in k, k in [k0-2, k0+TK+2)):
//   x ring: rows j0-2 .. j0+TJ+1        H ring and b ring: rows j0-1 .. j0+TJ
// When `e` is given the sweep's input is x_in + e(i/2, j/2, k/2): the
// prolongation by injection plus the saved iterate (poisson_d3.cc:548-558,
// poisson.cc:266) folded into the first post-smoothing sweep.
template <typename T, int NP, int TJ, bool FINE>
__global__ void __launch_bounds__(NP *TJ) k_sweep_march(LevelArgs<T> L, const T *__restrict__ xin,
                                                        const T *__restrict__ b, T *__restrict__ xout, T w,
                                                        int chunk, int ntk, int ntj, int ntiles,
                                                        const T *__restrict__ e, long esx, long esy) {
    using V = typename Pair<T>::V;
    constexpr int TK = 2 * NP;
    constexpr int PX = NP + 2;  // pair columns per region row
    constexpr int CW = 2 * PX;  // scalars per region row
    constexpr int RX = TJ + 4, RH = TJ + 2;
    constexpr int NT = NP * TJ;
    constexpr int NXP = RX * PX, NHP = RH * PX;
    static_assert(NXP <= 2 * NT && NHP <= 2 * NT, "two pair elements per thread");
    // FP32: 4-plane rings and one barrier per plane (the black update lags the
    // red one by two planes, so no ring slot is written while it is read);
    // FP64: 3-plane rings and two barriers (LDS per workgroup stays ~50 KB).
    constexpr bool ONEBAR = sizeof(T) == 4;
    constexpr int NR = ONEBAR ? 4 : 3;
    __shared__ __align__(16) T xs[NR][RX * CW];
    __shared__ __align__(16) T hs[NR][RH * CW];
    __shared__ __align__(16) T bs[NR][RH * CW];

    // XCD-aware logical tile id (grid padded to a multiple of kXcds)
    const int per = gridDim.x / kXcds;
    const int logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (logical >= ntiles) return;
    const int tk = logical % ntk;
    const int rest = logical / ntk;
    const int tj = rest % ntj;
    const int ch = rest / ntj;

    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const bool perx = L.bc_lo[0] == BC_PER, pery = L.bc_lo[1] == BC_PER, perz = L.bc_lo[2] == BC_PER;
    const int j0 = tj * TJ, k0 = tk * TK;
    const int p0 = ch * chunk;
    const int p1 = min(nx, p0 + chunk);
    const int tid = threadIdx.x;
    const long sx = L.sx;

    // fixed per-thread geometry
    const PairElem xe0 = pair_elem<PX>(tid, j0 - 2, k0 - 2, ny, nz, pery, perz, L.sy);
    const bool has_x1 = tid + NT < NXP;
    const PairElem xe1 = pair_elem<PX>(has_x1 ? tid + NT : tid, j0 - 2, k0 - 2, ny, nz, pery, perz, L.sy);
    const PairElem he0 = pair_elem<PX>(tid, j0 - 1, k0 - 2, ny, nz, pery, perz, L.sy);
    const bool has_h1 = tid + NT < NHP;
    const PairElem he1 = pair_elem<PX>(has_h1 ? tid + NT : tid, j0 - 1, k0 - 2, ny, nz, pery, perz, L.sy);
    const int ty = tid / NP, tp = tid - (tid / NP) * NP;
    const int jt = j0 + ty, kt = k0 + 2 * tp;
    const int t_off = (int)((long)jt * L.sy + kt);

    struct HCoef {
        T c2y, c1y, c2za, c1za, c2zb, c1zb;
    };
    auto hcoef = [&](const PairElem &e) -> HCoef {
        HCoef c{T(0), T(0), T(0), T(0), T(0), T(0)};
        if (e.goff >= 0) {
            c.c2y = L.c2[1][e.j];
            c.c1y = L.c1[1][e.j];
            c.c2za = L.c2[2][e.kps];
            c.c1za = L.c1[2][e.kps];
            c.c2zb = L.c2[2][e.kps + 1];
            c.c1zb = L.c1[2][e.kps + 1];
        }
        return c;
    };
    const HCoef hc0 = hcoef(he0), hc1 = hcoef(he1);
    const HCoef tc = {L.c2[1][jt], L.c1[1][jt], L.c2[2][kt], L.c1[2][kt], L.c2[2][kt + 1], L.c1[2][kt + 1]};

    auto plane_of = [&](int q) -> int {
        if (q >= 0 && q < nx) return q;
        if (!perx) return -1;
        return q < 0 ? q + nx : q - nx;
    };
    auto ld2 = [&](const T *base, int qq, const PairElem &e) -> V {
        V v;
        if (qq >= 0 && e.goff >= 0) {
            v = *reinterpret_cast<const V *>(base + (long)qq * sx + e.goff);
        } else {
            v.x = T(0);
            v.y = T(0);
        }
        return v;
    };
    auto st2 = [&](T *lds, int off, V v) { *reinterpret_cast<V *>(lds + off) = v; };
    // x_in pair, with the coarse correction added when prolongation is fused
    auto ldx = [&](int qq, const PairElem &pe) -> V {
        V v = ld2(xin, qq, pe);
        if (e && qq >= 0 && pe.goff >= 0) {
            const T ev = e[(long)(qq >> 1) * esx + (long)(pe.j >> 1) * esy + (pe.kps >> 1)];
            v.x = ev + v.x;
            v.y = ev + v.y;
        }
        return v;
    };

    // H of one pair: the red cell gets R (if it lies in the 1-cell ring), the
    // black cell keeps x_in
    auto h_pair = [&](const PairElem &e, const HCoef &c, int hslot, int sm, int sc, int sp, int qq, bool lo_wall,
                      bool hi_wall, T c2x, T c1x) {
        const int hoff = e.row * CW + 2 * e.pc;
        if (e.goff < 0) {
            V z;
            z.x = T(0);
            z.y = T(0);
            st2(hs[hslot], hoff, z);
            return;
        }
        const int xoff = hoff + CW;  // same pair one row down in the x region
        V own = *reinterpret_cast<const V *>(xs[sc] + xoff);
        const int r = (qq + e.j) & 1;  // red offset inside the pair (k even)
        const int col = 2 * e.pc + r;
        const bool needed = !(col == 0 || col == CW - 1);
        if (needed) {
            const int xi = xoff + r;
            const int kr = e.kps + r;
            const T o = r ? own.y : own.x;
            Stencil<T, 3> st;
            st.c = o;
            st.xW = lo_wall ? wall_nb(L, L.bc_lo[0], 0, o, qq, e.j, kr) : xs[sm][xi];
            st.xE = hi_wall ? wall_nb(L, L.bc_hi[0], 0, o, qq, e.j, kr) : xs[sp][xi];
            st.yS = (e.j == 0 && !pery) ? wall_nb(L, L.bc_lo[1], 1, o, qq, e.j, kr) : xs[sc][xi - CW];
            st.yN = (e.j == ny - 1 && !pery) ? wall_nb(L, L.bc_hi[1], 1, o, qq, e.j, kr) : xs[sc][xi + CW];
            st.zB = (kr == 0 && !perz) ? wall_nb(L, L.bc_lo[2], 2, o, qq, e.j, kr) : xs[sc][xi - 1];
            st.zT = (kr == nz - 1 && !perz) ? wall_nb(L, L.bc_hi[2], 2, o, qq, e.j, kr) : xs[sc][xi + 1];
            const T rv = sor_update<T, 3>(st, c2x, c1x, c.c2y, c.c1y, r ? c.c2zb : c.c2za, r ? c.c1zb : c.c1za,
                                          bs[hslot][hoff + r], w);
            if (r) own.y = rv;
            else own.x = rv;
        }
        st2(hs[hslot], hoff, own);
    };

    auto make_h = [&](int hslot, int sm, int sc, int sp, int q) {
        const int qq = plane_of(q);
        const bool lo_wall = (q == 0 && !perx), hi_wall = (q == nx - 1 && !perx);
        const T c2x = L.c2[0][qq], c1x = L.c1[0][qq];
        h_pair(he0, hc0, hslot, sm, sc, sp, qq, lo_wall, hi_wall, c2x, c1x);
        if (has_h1) h_pair(he1, hc1, hslot, sm, sc, sp, qq, lo_wall, hi_wall, c2x, c1x);
    };

    // black update of the own tile pair of plane p, then the pair is written
    auto finish = [&](int hm, int hc, int hp, int p, bool lo_wall, bool hi_wall, int sb) {
        const int hoff = (ty + 1) * CW + 2 * (tp + 1);
        V pr = *reinterpret_cast<const V *>(hs[hc] + hoff);
        const int bo = 1 - ((p + jt) & 1);  // black offset inside the pair
        const int hi = hoff + bo;
        const int kb = kt + bo;
        const T o = bo ? pr.y : pr.x;
        Stencil<T, 3> st;
        st.c = o;
        st.xW = lo_wall ? wall_nb(L, L.bc_lo[0], 0, o, p, jt, kb) : hs[hm][hi];
        st.xE = hi_wall ? wall_nb(L, L.bc_hi[0], 0, o, p, jt, kb) : hs[hp][hi];
        st.yS = (jt == 0 && !pery) ? wall_nb(L, L.bc_lo[1], 1, o, p, jt, kb) : hs[hc][hi - CW];
        st.yN = (jt == ny - 1 && !pery) ? wall_nb(L, L.bc_hi[1], 1, o, p, jt, kb) : hs[hc][hi + CW];
        st.zB = (kb == 0 && !perz) ? wall_nb(L, L.bc_lo[2], 2, o, p, jt, kb) : hs[hc][hi - 1];
        st.zT = (kb == nz - 1 && !perz) ? wall_nb(L, L.bc_hi[2], 2, o, p, jt, kb) : hs[hc][hi + 1];
        const T nv = sor_update<T, 3>(st, L.c2[0][p], L.c1[0][p], tc.c2y, tc.c1y, bo ? tc.c2zb : tc.c2za,
                                      bo ? tc.c1zb : tc.c1za, bs[sb][hi], w);
        if (bo) pr.y = nv;
        else pr.x = nv;
        *reinterpret_cast<V *>(xout + (long)p * sx + t_off) = pr;
    };

    // H planes qs..qe; black planes p0..p1-1.  Ring slots: x(p) -> (p - xbase) % 3,
    // H(q) and b(q) -> (q - qs) % 3.  Each iteration q stores a payload
    // {x(q+1), b(q)} loaded two iterations earlier (two register sets, loop
    // unrolled by two) so that two planes of loads are always in flight.
    const int qs = (p0 == 0 && !perx) ? 0 : p0 - 1;
    const int qe = (p1 == nx && !perx) ? nx - 1 : p1;
    const int xbase = qs - 1;
    auto sx_of = [&](int pl) { return (pl - xbase) % NR; };
    auto sq_of = [&](int q) { return ((q - qs) % NR + NR) % NR; };
    struct Payload {
        V x0, x1, b0, b1;
    };
    auto fetch = [&](Payload &r, int q) {
        if (q > qe) return;
        const int qx = plane_of(q + 1), qb = plane_of(q);
        r.x0 = ldx(qx, xe0);
        if (has_x1) r.x1 = ldx(qx, xe1);
        r.b0 = ld2(b, qb, he0);
        if (has_h1) r.b1 = ld2(b, qb, he1);
    };
    const int xo0 = xe0.row * CW + 2 * xe0.pc, xo1 = xe1.row * CW + 2 * xe1.pc;
    const int ho0 = he0.row * CW + 2 * he0.pc, ho1 = he1.row * CW + 2 * he1.pc;
    {
        const int a = plane_of(qs - 1), c = plane_of(qs);
        st2(xs[sx_of(qs - 1)], xo0, ldx(a, xe0));
        st2(xs[sx_of(qs)], xo0, ldx(c, xe0));
        if (has_x1) {
            st2(xs[sx_of(qs - 1)], xo1, ldx(a, xe1));
            st2(xs[sx_of(qs)], xo1, ldx(c, xe1));
        }
    }
    Payload A, B;
    fetch(A, qs);
    fetch(B, qs + 1);
    auto body = [&](int q, Payload &r) {
        const int sxp = sx_of(q + 1), sb = sq_of(q);
        st2(xs[sxp], xo0, r.x0);
        if (has_x1) st2(xs[sxp], xo1, r.x1);
        st2(bs[sb], ho0, r.b0);
        if (has_h1) st2(bs[sb], ho1, r.b1);
        fetch(r, q + 2);
        __syncthreads();
        make_h(sb, sx_of(q - 1), sx_of(q), sxp, q);
        if (ONEBAR) {
            const int p = q - 2;
            if (p >= p0 && p < p1)
                finish(sq_of(q - 3), sq_of(q - 2), sq_of(q - 1), p, p == 0 && !perx, false, sq_of(q - 2));
        } else {
            __syncthreads();
            const int p = q - 1;
            if (p >= p0 && p < p1) finish(sq_of(q - 2), sq_of(q - 1), sb, p, p == 0 && !perx, false, sq_of(q - 1));
        }
    };
    for (int q = qs; q <= qe; q += 2) {
        body(q, A);
        if (q + 1 <= qe) body(q + 1, B);
    }
    if (ONEBAR) {
        // the black update of the last H plane's predecessor
        __syncthreads();
        const int p = qe - 1;
        if (p >= p0 && p < p1) finish(sq_of(qe - 2), sq_of(qe - 1), sq_of(qe), p, p == 0 && !perx, false, sq_of(qe - 1));
    }
    if (qe == nx - 1 && p1 == nx && !perx) {
        // top wall: the last plane's x+ neighbour is the wall rule
        finish(sq_of(nx - 2), sq_of(nx - 1), sq_of(nx - 1), nx - 1, nx - 1 == 0, true, sq_of(nx - 1));
    }
}

// ---------------------------------------------------------------------------
// FP64 refinement step of the mixed-precision solve:
//   x_new = x_old + e            (e = FP32 correction, may be null)
//   r     = b - A x_new          (FP64, wall rules applied to x_new)
//   r32   = (float) r            (the next FP32 right-hand side)
// plus per-block partial norms of r and b (types as computeError).
// Out of place in x so that neighbours are read before any update.
// ---------------------------------------------------------------------------
template <int TYPE, int DIM>
__global__ void __launch_bounds__(kNormThreads) k_refine(LevelArgs<double> L, const double *__restrict__ xold,
                                                         const float *__restrict__ e, long esx, long esy,
                                                         const double *__restrict__ b, double *__restrict__ xnew,
                                                         float *__restrict__ r32, long rsx, long rsy, long ehz,
                                                         double *__restrict__ partial) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const long rows = (long)nx * ny;
    const RowWalk wk(nz);
    double ar = 0.0, ab = 0.0;
    for (long row = (long)blockIdx.x * wk.rpp + wk.sub; wk.sub < wk.rpp && row < rows; row += (long)gridDim.x * wk.rpp)
    for (int k = wk.k0; k < nz; k += wk.kspan) {
        const int i = (int)(row / ny);
        const int j = (int)(row - (long)i * ny);
        const long idx = (long)i * L.sx + (long)j * L.sy + kof(L, k);
        auto xv = [&](int ii, int jj, int kk) -> double {
            const double v = xold[(long)ii * L.sx + (long)jj * L.sy + kof(L, kk)];
            return e ? v + (double)e[(long)ii * esx + (long)jj * esy + (ehz ? (long)(kk & 1) * ehz + (kk >> 1) : (long)kk)]
                     : v;
        };
        Stencil<double, DIM> s;
        s.c = xv(i, j, k);
        auto nb = [&](int axis, int lohi, int ii, int jj, int kk, bool inside) -> double {
            if (inside) return xv(ii, jj, kk);
            const int bc = lohi ? L.bc_hi[axis] : L.bc_lo[axis];
            if (bc == BC_PER) {
                if (axis == 0) ii = lohi ? 0 : nx - 1;
                else if (axis == 1) jj = lohi ? 0 : ny - 1;
                else kk = lohi ? 0 : nz - 1;
                return xv(ii, jj, kk);
            }
            double wv = 0.0;
            if (bc == BC_DIRV) {
                if (DIM == 3) wv = axis == 0 ? L.wall[0][(j + 1) * (nz + 2) + k + 1]
                                 : axis == 1 ? L.wall[1][(i + 1) * (nz + 2) + k + 1]
                                             : L.wall[2][(i + 1) * (ny + 2) + j + 1];
                else wv = axis == 0 ? L.wall[0][k + 1] : L.wall[2][i + 1];
            }
            return bc_value<double>(bc, s.c, 0.0, wv);
        };
        s.xE = nb(0, 1, i + 1, j, k, i + 1 < nx);
        s.xW = nb(0, 0, i - 1, j, k, i > 0);
        if (DIM == 3) {
            s.yN = nb(1, 1, i, j + 1, k, j + 1 < ny);
            s.yS = nb(1, 0, i, j - 1, k, j > 0);
        }
        s.zT = nb(2, 1, i, j, k + 1, k + 1 < nz);
        s.zB = nb(2, 0, i, j, k - 1, k > 0);
        const double c2y = DIM == 3 ? L.c2[1][j] : 0.0, c1y = DIM == 3 ? L.c1[1][j] : 0.0;
        const double bv = b[idx];
        const double rv = residual<double, DIM>(s, L.c2[0][i], L.c1[0][i], c2y, c1y, L.c2[2][k], L.c1[2][k], bv);
        xnew[idx] = s.c;
        r32[(long)i * rsx + (long)j * rsy + (ehz ? (long)(k & 1) * ehz + (k >> 1) : (long)k)] = (float)rv;
        ar = norm_acc<TYPE>(ar, rv);
        ab = norm_acc<TYPE>(ab, bv);
    }
    block_reduce2<TYPE>(ar, ab);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ar;
        partial[2 * blockIdx.x + 1] = ab;
    }
}

}  // namespace saras
