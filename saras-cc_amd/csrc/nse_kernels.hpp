// nse_kernels.hpp -- FP64 kernels of the Navier-Stokes/Boussinesq time step
// (lib/field, lib/boundary, lib/timestep of the reference).
//
// Device layout: every field is stored core-only, nx x ny x nz doubles with
// z fastest and a row pitch of 16 doubles (128 bytes), the same layout as the
// multigrid level arrays, so the divergence can be written straight into the
// pressure solver's level-0 RHS and the pressure correction read straight out
// of its level-0 iterate.  Ghost cells are never stored: a stencil that steps
// out of the core gets the value the reference's imposeBCs() would have put
// there (field.cc:128-148, dirichlet.cc:72-76, neumann.cc:89-93, and the
// pad-thick periodic syncFaces of mpidata.cc):
//   periodic axis  F(-q) = F(n-q), F(n-1+q) = F(q-1)
//   Dirichlet wall ghost = 2 v - F(edge)
//   Neumann wall   ghost = F(edge) - v h_wall
// Every stencil of the step stays within one ghost layer on a wall axis
// (first and last cells use second-order stencils), so the second layer the
// reference leaves stale is not needed -- except above a z-hi wall for
// Morinishi's term when n_y != n_z (its edge test reads the y extent), where
// that layer is stored.  The other stored ghost layer is that of a nullBC
// pressure wall (boundary.h:205-212), whose ghosts no rule gives: they keep
// what the whole-array updates of P leave there.
//
// Floating point: the expressions below keep the operand order of the cited
// reference lines and are built with -ffp-contract=off, so the FP64 results
// match an FP64 CPU build of the same expressions bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace saras {
namespace nse {

constexpr double kRecip12 = 1.0 / 12.0;

// wall rules of one field (6 walls: x-lo, x-hi, y-lo, y-hi, z-lo, z-hi)
struct FieldBC {
    int type[6];     // 0 Dirichlet, 1 Neumann, 2 hot plate (z-lo), 3 plume outflow (z-hi), 4 null (stored ghosts)
    double val[6];   // Dirichlet value
    double kh[6];    // Neumann value * h_wall (neumann.cc:59-75)
    // kind 4 (nullBC, boundary.h:205-212): the wall's first ghost layer as
    // stored values, indexed by the two other core indices (x walls: j nz + k,
    // y walls: i nz + k, z walls: i ny + j)
    double *ghost[6];
    // the second ghost layer above a z-hi wall (i ny + j), stored where a
    // stencil reads it: Morinishi's dScheme 2 edge test on z uses the y extent
    // (vfield.cc:329, 434; sfield.cc:236, 295), so when n_y != n_z the last
    // z cell takes the 4th-order difference.  No imposeBCs and no step
    // writes that layer, so it keeps the caller's values.
    const double *top2;
};

// index of the ghost of core cell (c0, c1, c2) across a wall of axis ax in
// that wall's stored plane
__device__ __forceinline__ long ghost_plane_index(const int *n, int ax, int c0, int c1, int c2) {
    return ax == 0 ? (long)c1 * n[2] + c2 : (ax == 1 ? (long)c0 * n[2] + c2 : (long)c0 * n[1] + c1);
}

struct Geom {
    int n[3];
    long sx, sy;         // core pitch (elements)
    int per[3];
    int edge_last[3];    // last index of the 2nd-order edge test per axis (Morinishi z uses ny-1)
    int dscheme;
    double ih[3], ih2[3];          // derivative.cc:52-54
    double i2d[3];                 // vfield.cc:133-135
    double tsih2[3], tsi2h[3];     // timestep.cc:156-173
    const double *xi_x[3], *xixx[3], *xix2[3];   // core cells
    const double *xc[3];           // cell-centre coordinates over [-1, n] (xc[a][-1] valid)
    double pe_limit;               // "Peclet Limit"
    double up1[3], up2[4];         // biased-stencil weights from "Central Bias" (vfield.cc:115-130)
    // the wall kinds 2 and 3 (temperature only): one flag per (i, j) column,
    // nx*ny for the bottom wall then nx*ny for the top wall
    const unsigned char *wmask;
};

// The cell of this thread in the (ceil(nz/threads), ny, nx) row grids: a
// block is a segment of row (i, j).
__device__ __forceinline__ void row_cell(const Geom &G, int &i, int &j, int &k) {
    (void)G;
    j = blockIdx.y;
    i = blockIdx.z;
    k = blockIdx.x * blockDim.x + threadIdx.x;
}
// The z-wall kinds whose ghost depends on the column:
//   hot plate (hotPlate.cc:114-123 with the patch of :74-111): F(wall) =
//     F(edge) * mask + data * (2 data - F(edge)), data = 1 inside the disc,
//     i.e. 2 - F(edge) on the disc and F(edge) off it;
//   plume outflow (plumeOut.cc:71-90): F(edge) where the flag (ghost Vz > 0
//     at the last imposeBCs) is set, else 0.
__device__ __forceinline__ double z_wall_mask(const Geom &G, int type, bool hi, int i, int j, double inner) {
    const bool on = G.wmask[(hi ? (long)G.n[0] * G.n[1] : 0L) + (long)i * G.n[1] + j] != 0;
    if (type == 2) return on ? 2.0 - inner : inner;
    return on ? inner : 0.0;
}

__device__ __forceinline__ int pick(int ax, int i, int j, int k) { return ax == 0 ? i : (ax == 1 ? j : k); }

// The core index a neighbour load of index q (at most 2 outside [0, n)) reads:
// the periodic image on a periodic axis, else the edge cell a wall rule
// starts from.  Always inside [0, n): on the one-row y axis of a planar grid
// (n = 1, periodic) the image of q = +-2 is row 0 as well -- the second-order
// edge stencil there never uses it, but the load is issued unconditionally.
__device__ __forceinline__ int core_index(int per, int q, int n) {
    const int w = per ? (q < 0 ? q + n : (q >= n ? q - n : q)) : q;
    return w < 0 ? 0 : (w >= n ? n - 1 : w);
}

// Core-only field with on-the-fly ghost cells.
struct CoreField {
    const double *f;
    FieldBC bc;
    __device__ __forceinline__ double at(const Geom &G, int i, int j, int k) const {
        return f[(long)i * G.sx + (long)j * G.sy + k];
    }
    // The cell's own load is issued unconditionally from an index that is
    // always inside the core (the wrapped one on a periodic axis, else the
    // edge cell a wall rule reads), and the wall rule is applied to the loaded
    // value afterwards: no load waits behind a lane-dependent branch, so a
    // stencil's loads are issued together.
    __device__ __forceinline__ double nb(const Geom &G, int ax, int i, int j, int k, int off) const {
        int c[3] = {i, j, k};
        const int q = c[ax] + off;
        const int n = G.n[ax];
        const bool lo = q < 0, hi = q >= n;
        c[ax] = core_index(G.per[ax], q, n);
        const double v = f[(long)c[0] * G.sx + (long)c[1] * G.sy + c[2]];
        if (!(lo || hi) || G.per[ax]) return v;
        return wall(G, ax, hi, q, c, v);
    }
    // the ghost across a wall of axis ax (c: the edge cell, inner: its value);
    // select between the two walls without a run-time index into the rule
    // table (keeps the table in kernel-argument memory)
    __device__ __forceinline__ double wall(const Geom &G, int ax, bool hi, int q, const int (&c)[3],
                                           double inner) const {
        const int n = G.n[ax];
        if (ax == 2 && q == n + 1 && bc.top2) return bc.top2[(long)c[0] * G.n[1] + c[1]];
        const int type = hi ? bc.type[2 * ax + 1] : bc.type[2 * ax];
        const double val = hi ? bc.val[2 * ax + 1] : bc.val[2 * ax];
        const double kh = hi ? bc.kh[2 * ax + 1] : bc.kh[2 * ax];
        if (type == 4) {
            const double *gp = hi ? bc.ghost[2 * ax + 1] : bc.ghost[2 * ax];
            return gp[ghost_plane_index(G.n, ax, c[0], c[1], c[2])];
        }
        if (type > 1) return z_wall_mask(G, type, hi, c[0], c[1], inner);
        return type == 0 ? 2.0 * val - inner : inner - kh;
    }
};

// Core-only field of a fully periodic box (the LES stresses): every
// neighbour wraps, no wall rule.
struct PerField {
    const double *f;
    __device__ __forceinline__ double at(const Geom &G, int i, int j, int k) const {
        return f[(long)i * G.sx + (long)j * G.sy + k];
    }
    __device__ __forceinline__ double nb(const Geom &G, int ax, int i, int j, int k, int off) const {
        const int n = ax == 0 ? G.n[0] : (ax == 1 ? G.n[1] : G.n[2]);
        int q = (ax == 0 ? i : (ax == 1 ? j : k)) + off;
        q = q < 0 ? q + n : (q >= n ? q - n : q);
        const long o = ax == 0 ? (long)q * G.sx + (long)j * G.sy + k
                               : (ax == 1 ? (long)i * G.sx + (long)q * G.sy + k : (long)i * G.sx + (long)j * G.sy + q);
        return f[o];
    }
};

// Core-only field on colour-split rows (the pressure solver's level-0
// iterate): element k of a row at (k & 1) hz + k / 2, otherwise as CoreField.
struct SplitField {
    const double *f;
    FieldBC bc;
    long hz;
    __device__ __forceinline__ long zo(int k) const { return (long)(k & 1) * hz + (k >> 1); }
    __device__ __forceinline__ double at(const Geom &G, int i, int j, int k) const {
        return f[(long)i * G.sx + (long)j * G.sy + zo(k)];
    }
    // loads first, wall rule after, as CoreField::nb
    __device__ __forceinline__ double nb(const Geom &G, int ax, int i, int j, int k, int off) const {
        int c[3] = {i, j, k};
        const int q = c[ax] + off;
        const int n = G.n[ax];
        const bool lo = q < 0, hi = q >= n;
        c[ax] = core_index(G.per[ax], q, n);
        const double v = f[(long)c[0] * G.sx + (long)c[1] * G.sy + zo(c[2])];
        if (!(lo || hi) || G.per[ax]) return v;
        const int type = hi ? bc.type[2 * ax + 1] : bc.type[2 * ax];
        const double val = hi ? bc.val[2 * ax + 1] : bc.val[2 * ax];
        const double kh = hi ? bc.kh[2 * ax + 1] : bc.kh[2 * ax];
        return type == 0 ? 2.0 * val - v : v - kh;
    }
};

// Caller-layout field with stored pads (pad p, (n+2p)^3): the single-operator
// entry points read whatever ghost values the caller supplies, as the
// reference's derivative class reads its array.
struct PadField {
    const double *f;
    long SX, SY;
    int p;
    __device__ __forceinline__ double at(const Geom &, int i, int j, int k) const {
        return f[(long)(i + p) * SX + (long)(j + p) * SY + (k + p)];
    }
    __device__ __forceinline__ double nb(const Geom &, int ax, int i, int j, int k, int off) const {
        int c[3] = {i, j, k};
        c[ax] += off;
        return f[(long)(c[0] + p) * SX + (long)(c[1] + p) * SY + (c[2] + p)];
    }
};

// View of a field for a row at least two rows from the x and y faces: every
// x or y neighbour a stencil reads lies in the core and is read straight from
// memory (the value the field's own rule would return, without its wrap and
// wall tests); z neighbours still go through the field's rule.  The test is
// per row, so a wavefront (one row segment) takes one path.
template <class FB>
struct FastXY {
    FB base;
    __device__ __forceinline__ double at(const Geom &G, int i, int j, int k) const { return base.at(G, i, j, k); }
    __device__ __forceinline__ double nb(const Geom &G, int ax, int i, int j, int k, int off) const {
        if (ax == 0) return base.at(G, i + off, j, k);
        if (ax == 1) return base.at(G, i, j + off, k);
        return base.nb(G, 2, i, j, k, off);
    }
};
template <class FB>
__device__ __forceinline__ FastXY<FB> fast_of(const FB &c) { return FastXY<FB>{c}; }
// rows whose x/y stencils (reach 2) stay in the core
__device__ __forceinline__ bool interior2(const Geom &G, int i, int j, int k) {
    (void)k;
    return i >= 2 && i < G.n[0] - 2 && j >= 2 && j < G.n[1] - 2;
}

// derivative::calcDerivative1_x/y/z (derivative.cc:104-192)
template <class F>
__device__ __forceinline__ double d1(const Geom &G, const F &A, int ax, int i, int j, int k) {
    const int q = pick(ax, i, j, k);
    const double fm = A.nb(G, ax, i, j, k, -1), fp = A.nb(G, ax, i, j, k, 1);
    double v;
    if (G.dscheme == 1) {
        v = (fp - fm) * 0.5;
    } else {
        // loaded before the edge test (which differs between lanes); an edge
        // cell does not use them
        const double fm2 = A.nb(G, ax, i, j, k, -2), fp2 = A.nb(G, ax, i, j, k, 2);
        if (q == 0 || q == G.n[ax] - 1) v = 0.5 * (fp - fm);
        else v = (8.0 * (fp - fm) - (fp2 - fm2)) * kRecip12;
    }
    v = v * G.ih[ax];
    return G.xi_x[ax][q] * v;
}

// derivative::calcDerivative2xx/yy/zz (derivative.cc:206-306)
template <class F>
__device__ __forceinline__ double d2(const Geom &G, const F &A, int ax, int i, int j, int k) {
    const int q = pick(ax, i, j, k);
    const double fm = A.nb(G, ax, i, j, k, -1), f = A.at(G, i, j, k), fp = A.nb(G, ax, i, j, k, 1);
    double t, o;
    if (G.dscheme == 1) {
        t = (fp - fm) * 0.5;
        o = (fm - 2.0 * f) + fp;
    } else {
        const double fm2 = A.nb(G, ax, i, j, k, -2), fp2 = A.nb(G, ax, i, j, k, 2);
        if (q == 0 || q == G.n[ax] - 1) {
            t = 0.5 * (fp - fm);
            o = fp - 2.0 * f + fm;
        } else {
            t = (8.0 * (fp - fm) - (fp2 - fm2)) * kRecip12;
            o = ((-30.0 * f + 16.0 * (fp + fm)) - (fp2 + fm2)) * kRecip12;
        }
    }
    t = t * G.ih[ax];
    o = o * G.ih2[ax];
    return G.xixx[ax][q] * t + G.xix2[ax][q] * o;
}

// computeDiff: H += d2xx; += d2yy; += d2zz (vfield.cc:150-192, sfield.cc:133-147)
template <class F>
__device__ __forceinline__ double laplacian(const Geom &G, const F &A, int i, int j, int k) {
    double h = 0.0;
    h += d2(G, A, 0, i, j, k);
    h += d2(G, A, 1, i, j, k);
    h += d2(G, A, 2, i, j, k);
    return h;
}

__device__ __forceinline__ bool mor_edge(const Geom &G, int ax, int q) {
    return G.dscheme == 1 || q == 0 || q == G.edge_last[ax];
}

// Morinishi skew-symmetric term of one (advecting axis a, advected field C)
// pair: (adv + div) / 2 with adv = u_a xi_x dC, div = xi_x d(U_a C)
// (vfield.cc:351-459, sfield.cc:236-307).  U is the advecting component.
template <class FU, class FC>
__device__ __forceinline__ double mor_pair(const Geom &G, const FU &U, const FC &Cf, int a, int i, int j, int k) {
    const int q = pick(a, i, j, k);
    const double xm = G.xi_x[a][q];
    const double u = U.at(G, i, j, k);
    const double cp = Cf.nb(G, a, i, j, k, 1), cm = Cf.nb(G, a, i, j, k, -1);
    const double up = U.nb(G, a, i, j, k, 1), um = U.nb(G, a, i, j, k, -1);
    const double pp = up * cp, pm = um * cm;
    // the reach-2 values are loaded before the lane-dependent edge test
    double cp2 = 0.0, cm2 = 0.0, up2 = 0.0, um2 = 0.0;
    if (G.dscheme != 1) {
        cp2 = Cf.nb(G, a, i, j, k, 2);
        cm2 = Cf.nb(G, a, i, j, k, -2);
        up2 = U.nb(G, a, i, j, k, 2);
        um2 = U.nb(G, a, i, j, k, -2);
    }
    double adv, dv;
    if (mor_edge(G, a, q)) {
        adv = u * xm * (cp - cm) * G.i2d[a];
        dv = xm * (pp - pm) * G.i2d[a];
    } else {
        const double pp2 = up2 * cp2, pm2 = um2 * cm2;
        adv = u * xm * (-cp2 + 8.0 * cp - 8.0 * cm + cm2) * G.i2d[a] / 6.0;
        dv = xm * (-pp2 + 8.0 * pp - 8.0 * pm + pm2) * G.i2d[a] / 6.0;
    }
    return (adv + dv) / 2;
}

// hybrid upwind term along axis `a` with advecting velocity u (vfield.cc:
// 475-689, sfield.cc:324-480): central differences where the cell Peclet
// number |u| dh / diff is below the limit, else biased stencils
template <class FC>
__device__ __forceinline__ double upwind_term(const Geom &G, double u, const FC &Cf, int a, int i, int j, int k,
                                              double diff) {
    const int q = pick(a, i, j, k);
    const double xm = G.xi_x[a][q], i2d = G.i2d[a];
    const double fm = Cf.nb(G, a, i, j, k, -1), fp = Cf.nb(G, a, i, j, k, 1);
    // every load is issued before the lane-dependent edge and Peclet tests
    const double dh = G.xc[a][q + 1] - G.xc[a][q];
    const double f = Cf.at(G, i, j, k);
    if (G.dscheme == 1) {
        const double pe = fabs(u) * dh / diff;
        if (pe < G.pe_limit) return u * xm * (fp - fm) * i2d;
        const double A = G.up1[0], B = G.up1[1], Cc = G.up1[2];
        if (u > 0) return u * xm * (-A * fm + B * f - Cc * fp) * i2d;
        return u * xm * (A * fp - B * f + Cc * fm) * i2d;
    }
    const double fm2 = Cf.nb(G, a, i, j, k, -2), fp2 = Cf.nb(G, a, i, j, k, 2);
    if (q == 0 || q == G.n[a] - 1) return u * xm * (fp - fm) * i2d;
    const double pe = fabs(u) * dh / diff;
    if (pe < G.pe_limit) return u * xm * (-fp2 + 8.0 * fp - 8.0 * fm + fm2) * i2d / 6.0;
    const double A = G.up2[0], B = G.up2[1], Cc = G.up2[2], D = G.up2[3];
    if (u > 0) return u * xm * (A * fm2 - B * fm + Cc * f - D * fp) * i2d;
    return u * xm * (-A * fp2 + B * fp - Cc * f + D * fm) * i2d;
}

// computeNLin of advected field C with velocity V, subtracted from `acc` in
// the reference's order (x, y, z terms); `diff` is the field's diffusion
// coefficient (the upwind scheme's Peclet number)
template <class FV, class FC>
__device__ __forceinline__ double nlin_acc(const Geom &G, int nlscheme, const FV *V, const FC &Cf, int i, int j, int k,
                                           double acc, double diff = 1.0) {
    if (nlscheme == 2) {
        acc -= upwind_term(G, V[0].at(G, i, j, k), Cf, 0, i, j, k, diff);
        acc -= upwind_term(G, V[1].at(G, i, j, k), Cf, 1, i, j, k, diff);
        acc -= upwind_term(G, V[2].at(G, i, j, k), Cf, 2, i, j, k, diff);
    } else if (nlscheme == 1) {
        // central: H -= V_a dC/dx_a (vfield.cc:207-249, sfield.cc:162-178)
        acc -= V[0].at(G, i, j, k) * d1(G, Cf, 0, i, j, k);
        acc -= V[1].at(G, i, j, k) * d1(G, Cf, 1, i, j, k);
        acc -= V[2].at(G, i, j, k) * d1(G, Cf, 2, i, j, k);
    } else {
        acc -= mor_pair(G, V[0], Cf, 0, i, j, k);
        acc -= mor_pair(G, V[1], Cf, 1, i, j, k);
        acc -= mor_pair(G, V[2], Cf, 2, i, j, k);
    }
    return acc;
}

// H_c of computeNLin into a zeroed array
template <class FV, class FC>
__device__ __forceinline__ double nlin(const Geom &G, int nlscheme, const FV *V, const FC &Cf, int i, int j, int k,
                                       double diff) {
    return nlin_acc(G, nlscheme, V, Cf, i, j, k, 0.0, diff);
}

struct StageCoef {
    double zeta;       // zetaRK3(rkLev)
    double a_nu;       // alphRK3 * nu
    double a_kappa;    // alphRK3 * kappa
    double ab;         // alphRK3 + betaRK3
    double gamma;      // gammRK3
    double dt;
    double buoy[3];    // gTerms
    double mpg;        // constantPGrad's mean pressure gradient (x component)
    int stage;
    int has_T;
    int nlscheme;
    double nu, kappa;  // the Peclet numbers of the upwind scheme
};

// The explicit part of one RK3 sub-stage (lsRK3_d3.cc:199-267) for velocity
// component C: R = (zeta H_old + alpha nu lap u + (alpha+beta)(f - grad P)
// + gamma H) dt + u, with H the new non-linear term, written back for the
// next sub-stage.  One launch per component keeps the register footprint of
// the Morinishi stencils spill-free.
template <int C, class F>
__device__ __forceinline__ void rhs_v_cell(const Geom &G, const StageCoef &c, const F &V0, const F &V1, const F &V2,
                                           const F &T, const F &P, double *__restrict__ H, double *__restrict__ R,
                                           int les, const PerField &Sa, const PerField &Sb, const PerField &Sc, int i,
                                           int j, int k) {
    const long idx = (long)i * G.sx + (long)j * G.sy + k;
    const F V[3] = {V0, V1, V2};
    const F &U = V[C];
    double nse = 0.0;
    if (c.stage > 0) nse += c.zeta * H[idx];
    const double diff = laplacian(G, U, i, j, k);
    nse += c.a_nu * diff;
    double f = 0.0;
    if (c.has_T) f -= c.buoy[C] * T.at(G, i, j, k);
    if (C == 0 && c.mpg != 0.0) f += c.mpg;   // constantPGrad::addForcing (force.h:152)
    f -= d1(G, P, C, i, j, k);
    nse += c.ab * f;
    double h = nlin(G, c.nlscheme, V, U, i, j, k, c.nu);
    // spiral::computeSG: H = H - (dSa/dx + dSb/dy + dSc/dz) (spiral.cc:225-235)
    if (les) h = h - (d1(G, Sa, 0, i, j, k) + d1(G, Sb, 1, i, j, k) + d1(G, Sc, 2, i, j, k));
    nse += c.gamma * h;
    nse *= c.dt;
    nse += U.at(G, i, j, k);
    H[idx] = h;
    R[idx] = nse;
}

template <int C>
__global__ void __launch_bounds__(256) k_rhs_v(Geom G, StageCoef c, CoreField V0, CoreField V1, CoreField V2,
                                               CoreField T, CoreField P, double *__restrict__ H,
                                               double *__restrict__ R, int les, PerField Sa, PerField Sb,
                                               PerField Sc) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    if (interior2(G, i, j, k))
        rhs_v_cell<C>(G, c, fast_of(V0), fast_of(V1), fast_of(V2), fast_of(T), fast_of(P), H, R, les, Sa, Sb, Sc, i,
                      j, k);
    else
        rhs_v_cell<C>(G, c, V0, V1, V2, T, P, H, R, les, Sa, Sb, Sc, i, j, k);
}

// the same for the scalar (tmpRHS / tempSF); its forcing term is zero
template <class F>
__device__ __forceinline__ void rhs_t_cell(const Geom &G, const StageCoef &c, const F &V0, const F &V1, const F &V2,
                                           const F &T, double *__restrict__ H, double *__restrict__ R, int i, int j,
                                           int k) {
    const long idx = (long)i * G.sx + (long)j * G.sy + k;
    const F V[3] = {V0, V1, V2};
    double tr = 0.0;
    if (c.stage > 0) tr += c.zeta * H[idx];
    const double diff = laplacian(G, T, i, j, k);
    tr += c.a_kappa * diff;
    tr += c.ab * 0.0;
    const double h = nlin(G, c.nlscheme, V, T, i, j, k, c.kappa);
    tr += c.gamma * h;
    tr *= c.dt;
    tr += T.at(G, i, j, k);
    H[idx] = h;
    R[idx] = tr;
}

__global__ void __launch_bounds__(256) k_rhs_t(Geom G, StageCoef c, CoreField V0, CoreField V1, CoreField V2,
                                               CoreField T, double *__restrict__ H, double *__restrict__ R) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    if (interior2(G, i, j, k)) rhs_t_cell(G, c, fast_of(V0), fast_of(V1), fast_of(V2), fast_of(T), H, R, i, j, k);
    else rhs_t_cell(G, c, V0, V1, V2, T, H, R, i, j, k);
}

// The explicit part of the Euler-CN step for velocity component C
// (eulerCN_d3.cc:181-226): R = ((nu/2) lap u - H - g T - dP/dx_C) dt + u,
// accumulated in one value in the reference's order.
template <int C, class F>
__device__ __forceinline__ void euler_v_cell(const Geom &G, double k_nu, double nu, double dt, double buoy, int has_T,
                                             double mpg, int nlscheme, const F &V0, const F &V1, const F &V2, const F &T,
                                             const F &P, double *__restrict__ R, int les, const PerField &Sa,
                                             const PerField &Sb, const PerField &Sc, int i, int j, int k) {
    const F V[3] = {V0, V1, V2};
    const F &U = V[C];
    double r = laplacian(G, U, i, j, k);
    r *= k_nu;
    r = nlin_acc(G, nlscheme, V, U, i, j, k, r, nu);
    if (has_T) r -= buoy * T.at(G, i, j, k);
    if (mpg != 0.0) r += mpg;   // constantPGrad::addForcing (force.h:152), x component only
    if (les) r = r - (d1(G, Sa, 0, i, j, k) + d1(G, Sb, 1, i, j, k) + d1(G, Sc, 2, i, j, k));
    r -= d1(G, P, C, i, j, k);
    r *= dt;
    r += U.at(G, i, j, k);
    R[(long)i * G.sx + (long)j * G.sy + k] = r;
}

template <int C>
__global__ void __launch_bounds__(256) k_euler_v(Geom G, double k_nu, double nu, double dt, double buoy, int has_T,
                                                 double mpg, int nlscheme, CoreField V0, CoreField V1, CoreField V2,
                                                 CoreField T, CoreField P, double *__restrict__ R, int les,
                                                 PerField Sa, PerField Sb, PerField Sc) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    if (interior2(G, i, j, k))
        euler_v_cell<C>(G, k_nu, nu, dt, buoy, has_T, mpg, nlscheme, fast_of(V0), fast_of(V1), fast_of(V2), fast_of(T),
                        fast_of(P), R, les, Sa, Sb, Sc, i, j, k);
    else
        euler_v_cell<C>(G, k_nu, nu, dt, buoy, has_T, mpg, nlscheme, V0, V1, V2, T, P, R, les, Sa, Sb, Sc, i, j, k);
}

// ... and for the scalar (eulerCN_d3.cc:188-232); its forcing is zero
template <class F>
__device__ __forceinline__ void euler_t_cell(const Geom &G, double k_kappa, double kappa, double dt, int nlscheme,
                                             const F &V0, const F &V1, const F &V2, const F &T, double *__restrict__ R,
                                             int i, int j, int k) {
    const F V[3] = {V0, V1, V2};
    double r = laplacian(G, T, i, j, k);
    r *= k_kappa;
    r = nlin_acc(G, nlscheme, V, T, i, j, k, r, kappa);
    r *= dt;
    r += T.at(G, i, j, k);
    R[(long)i * G.sx + (long)j * G.sy + k] = r;
}

__global__ void __launch_bounds__(256) k_euler_t(Geom G, double k_kappa, double kappa, double dt, int nlscheme, CoreField V0,
                                                 CoreField V1, CoreField V2, CoreField T, double *__restrict__ R) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    if (interior2(G, i, j, k))
        euler_t_cell(G, k_kappa, kappa, dt, nlscheme, fast_of(V0), fast_of(V1), fast_of(V2), fast_of(T), R, i, j, k);
    else
        euler_t_cell(G, k_kappa, kappa, dt, nlscheme, V0, V1, V2, T, R, i, j, k);
}

// ---- stretched spiral vortex sub-grid stress (lib/les/spiral.cc) ----------
constexpr double kSgsEps = 2e-3;                    // spiral.cc:49
constexpr double kPi = 3.14159265358979323846;      // M_PI

__device__ __forceinline__ double ke_integral(double k) {   // spiral.cc:590-607
    const double k2 = k * k;
    if (k2 < 2.42806) {
        const double pade = (3.0 + 2.5107 * k2 + 0.330357 * k2 * k2 + 0.0295481 * k2 * k2 * k2) /
                            (1.0 + 0.336901 * k2 + 0.0416684 * k2 * k2 + 0.00187191 * k2 * k2 * k2);
        return 0.5 * (pade - 4.06235 * cbrt(k2));   // k2^(1/3)
    }
    const double pade = (1.26429 + 0.835714 * k2 + 0.0964286 * k2 * k2) /
                        (1.0 + 2.25 * k2 + 0.964286 * k2 * k2 + 0.0964286 * k2 * k2 * k2);
    return 0.5 * pade * exp(-k2);
}

__device__ __forceinline__ double sf_integral(double d) {   // spiral.cc:620-631
    const double d2 = d * d;
    if (d < 0.873469) return 7.4022 * d2 - 1.82642 * d2 * d2;
    // d^(2/3) and d^(-3/2) as cbrt(d)^2 and 1 / (d sqrt(d)): the same values to
    // a few units in the last place, at a fraction of pow's cost
    const double c = cbrt(d);
    return 12.2946 * (c * c) - 6.0 - 0.573159 / (d * sqrt(d)) * sin(3.14159 * d - 0.785398);
}

// sf_integral in single precision (the same expression)
__device__ __forceinline__ float sf_integral_f(float d) {
    const float d2 = d * d;
    if (d < 0.873469f) return 7.4022f * d2 - 1.82642f * d2 * d2;
    const float c = cbrtf(d);
    return 12.2946f * (c * c) - 6.0f - 0.573159f / (d * sqrtf(d)) * sinf(3.14159f * d - 0.785398f);
}

// One core cell of spiral::computeSG (spiral.cc:157-206 with sgsStress
// :472-548, eigenvalueSymm :643-~690, eigenvectorSymm :702-~742) on a fully
// periodic box: writes the six stresses, returns false where the reference
// aborts.  SF32: the structure-function integral Qd (the 13 or 26
// sfIntegral terms, most of the cell's arithmetic) in single precision; the
// velocity gradients, eigen-decomposition, F2 and the energy integral stay
// in double, so the stress carries the single-precision rounding of Qd only.
template <bool SF32 = false>
__device__ bool sgs_cell(const Geom &G, const CoreField *V, double nu, int i, int j, int k, double T[6],
                         double &K, double S[6]) {
    double dudx[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        dudx[r][0] = d1(G, V[r], 0, i, j, k);
        dudx[r][1] = d1(G, V[r], 1, i, j, k);
        dudx[r][2] = d1(G, V[r], 2, i, j, k);
    }
    const double Sxx = 0.5 * (dudx[0][0] + dudx[0][0]), Syy = 0.5 * (dudx[1][1] + dudx[1][1]);
    const double Szz = 0.5 * (dudx[2][2] + dudx[2][2]), Sxy = 0.5 * (dudx[0][1] + dudx[1][0]);
    const double Syz = 0.5 * (dudx[1][2] + dudx[2][1]), Szx = 0.5 * (dudx[2][0] + dudx[0][2]);
    S[0] = Sxx; S[1] = Syy; S[2] = Szz; S[3] = Sxy; S[4] = Syz; S[5] = Szx;
    // eigenvalueSymm
    const double a = -(Sxx + Syy + Szz);
    const double b = Sxx * Syy - Sxy * Sxy + Syy * Szz - Syz * Syz + Szz * Sxx - Szx * Szx;
    const double cc = -(Sxx * (Syy * Szz - Syz * Syz) + Sxy * (Syz * Szx - Sxy * Szz) + Szx * (Sxy * Syz - Syy * Szx));
    const double q = (3.0 * b - a * a) / 9.0;
    const double r = (9.0 * a * b - 27.0 * cc - 2.0 * a * a * a) / 54.0;
    if (q >= 0.0) return false;
    const double costheta = r / sqrt(-q * q * q);
    const double theta = costheta > 1.0 ? 0.0 : costheta < -1.0 ? kPi : acos(costheta);
    double ev[3];
    ev[0] = 2.0 * sqrt(-q) * cos((theta) / 3.0) - a / 3.0;
    ev[1] = 2.0 * sqrt(-q) * cos((theta + 2.0 * kPi) / 3.0) - a / 3.0;
    ev[2] = 2.0 * sqrt(-q) * cos((theta + 4.0 * kPi) / 3.0) - a / 3.0;
    double tt;
    if (ev[0] > ev[1]) { tt = ev[0]; ev[0] = ev[1]; ev[1] = tt; }
    if (ev[1] > ev[2]) { tt = ev[1]; ev[1] = ev[2]; ev[2] = tt; }
    if (ev[0] > ev[1]) { tt = ev[0]; ev[0] = ev[1]; ev[1] = tt; }
    const double lam = ev[2];
    // eigenvectorSymm
    const double fNorm = sqrt(Sxx * Sxx + Syy * Syy + Szz * Szz + Sxy * Sxy + Syz * Syz + Szx * Szx);
    if (fabs((Sxx - lam) * ((Syy - lam) * (Szz - lam) - Syz * Syz) + Sxy * (Syz * Szx - Sxy * (Szz - lam)) +
             Szx * (Sxy * Syz - (Syy - lam) * Szx)) / fabs(fNorm) > kSgsEps)
        return false;
    const double det0 = (Syy - lam) * (Szz - lam) - Syz * Syz, det1 = (Szz - lam) * (Sxx - lam) - Szx * Szx;
    const double det2 = (Sxx - lam) * (Syy - lam) - Sxy * Sxy;
    const double f0 = fabs(det0), f1 = fabs(det1), f2 = fabs(det2);
    double e[3];
    if (f0 >= f1 && f0 >= f2) {
        e[0] = 1.0;
        e[1] = (-Sxy * (Szz - lam) + Szx * Syz) / det0;
        e[2] = (-Szx * (Syy - lam) + Sxy * Syz) / det0;
    } else if (f1 >= f2 && f1 >= f0) {
        e[0] = (-Sxy * (Szz - lam) + Syz * Szx) / det1;
        e[1] = 1.0;
        e[2] = (-Syz * (Sxx - lam) + Sxy * Szx) / det1;
    } else if (f2 >= f0 && f2 >= f1) {
        e[0] = (-Szx * (Syy - lam) + Syz * Sxy) / det2;
        e[1] = (-Syz * (Sxx - lam) + Szx * Sxy) / det2;
        e[2] = 1.0;
    } else {
        return false;
    }
    const double length = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    e[0] /= length;
    e[1] /= length;
    e[2] /= length;
    const double ax = e[0] * e[0] * Sxx + e[0] * e[1] * Sxy + e[0] * e[2] * Szx + e[1] * e[0] * Sxy +
                      e[1] * e[1] * Syy + e[1] * e[2] * Syz + e[2] * e[0] * Szx + e[2] * e[1] * Syz +
                      e[2] * e[2] * Szz;
    const double lv = sqrt(2.0 * nu / (3.0 * (fabs(ax) + kSgsEps)));
    // structure function over the 26 neighbours (periodic wrap)
    const double dxv = G.xc[0][i] - G.xc[0][i - 1], dyv = G.xc[1][j] - G.xc[1][j - 1];
    const double dzv = G.xc[2][k] - G.xc[2][k - 1];
    const double del = cbrt(dxv * dyv * dzv);
    const double u0 = V[0].at(G, i, j, k), v0 = V[1].at(G, i, j, k), w0 = V[2].at(G, i, j, k);
    double F2 = 0.0, Qd = 0.0;
    int cnt = 0;
#pragma unroll 1
    for (int di = -1; di <= 1; ++di) {
        const int ii = i + di < 0 ? i + di + G.n[0] : (i + di >= G.n[0] ? i + di - G.n[0] : i + di);
        for (int dj = -1; dj <= 1; ++dj) {
            const int jj = j + dj < 0 ? j + dj + G.n[1] : (j + dj >= G.n[1] ? j + dj - G.n[1] : j + dj);
            for (int dk = -1; dk <= 1; ++dk) {
                if (!(di || dj || dk)) continue;
                const int kk = k + dk < 0 ? k + dk + G.n[2] : (k + dk >= G.n[2] ? k + dk - G.n[2] : k + dk);
                const double du = V[0].at(G, ii, jj, kk) - u0, dv = V[1].at(G, ii, jj, kk) - v0;
                const double dw = V[2].at(G, ii, jj, kk) - w0;
                F2 += du * du + dv * dv + dw * dw;
                ++cnt;
            }
        }
    }
    // sfIntegral of each neighbour offset's distance from the vortex axis.  The
    // distance is even in the offset, so where the spacing on the two sides of
    // the cell agrees to rounding (uniform axes) each of the 13 offsets of one
    // half stands for its mirror image too; elsewhere all 26 are evaluated.
    const double dxp = G.xc[0][i + 1] - G.xc[0][i], dxm = G.xc[0][i - 1] - G.xc[0][i];
    const double dyp = G.xc[1][j + 1] - G.xc[1][j], dym = G.xc[1][j - 1] - G.xc[1][j];
    const double dzp = G.xc[2][k + 1] - G.xc[2][k], dzm = G.xc[2][k - 1] - G.xc[2][k];
    auto close = [](double a, double b) { return fabs(a + b) <= 1e-12 * fabs(a); };
    const bool mirror = close(dxp, dxm) && close(dyp, dym) && close(dzp, dzm);
    const float fdxm = (float)dxm, fdxp = (float)dxp, fdym = (float)dym, fdyp = (float)dyp;
    const float fdzm = (float)dzm, fdzp = (float)dzp, fe0 = (float)e[0], fe1 = (float)e[1], fe2 = (float)e[2];
    const float fdel = (float)del;
    auto sf_at = [&](int di, int dj, int dk) -> double {
        if (SF32) {
            const float dx = di < 0 ? fdxm : (di > 0 ? fdxp : 0.0f);
            const float dy = dj < 0 ? fdym : (dj > 0 ? fdyp : 0.0f);
            const float dz = dk < 0 ? fdzm : (dk > 0 ? fdzp : 0.0f);
            const float dx2 = dx * dx + dy * dy + dz * dz;
            const float dxe = dx * fe0 + dy * fe1 + dz * fe2;
            return (double)sf_integral_f(sqrtf(fmaxf(dx2 - dxe * dxe, 0.0f)) / fdel);
        }
        const double dx = di < 0 ? dxm : (di > 0 ? dxp : 0.0);
        const double dy = dj < 0 ? dym : (dj > 0 ? dyp : 0.0);
        const double dz = dk < 0 ? dzm : (dk > 0 ? dzp : 0.0);
        const double dx2 = dx * dx + dy * dy + dz * dz;
        const double dxe = dx * e[0] + dy * e[1] + dz * e[2];
        return sf_integral(sqrt(dx2 - dxe * dxe) / del);
    };
    if (mirror) {
        // the 13 offsets before (0, 0, 0) in (di, dj, dk) order
#pragma unroll 1
        for (int t = 0; t < 13; ++t) {
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            Qd += 2.0 * sf_at(di, dj, dk);
        }
    } else {
#pragma unroll 1
        for (int t = 0; t < 27; ++t) {
            if (t == 13) continue;
            Qd += sf_at(t / 9 - 1, (t / 3) % 3 - 1, t % 3 - 1);
        }
    }
    F2 /= (double)cnt;
    Qd /= (double)cnt;
    const double prefac = F2 / Qd;
    const double kc = kPi / del;
    K = prefac * ke_integral(kc * lv);
    T[0] = (1.0 - e[0] * e[0]) * K;
    T[1] = (1.0 - e[1] * e[1]) * K;
    T[2] = (1.0 - e[2] * e[2]) * K;
    T[3] = (-e[0] * e[1]) * K;
    T[4] = (-e[1] * e[2]) * K;
    T[5] = (-e[2] * e[0]) * K;
    return true;
}

// sub-grid stresses of every core cell, per-block partial diagnostics
// [sum |K| dV, sum dissipation dV, sum nu_t dV], failure flag
template <bool SF32>
__global__ void __launch_bounds__(256) k_sgs_stress(Geom G, CoreField V0, CoreField V1, CoreField V2, double nu,
                                                    double *__restrict__ Txx, double *__restrict__ Tyy,
                                                    double *__restrict__ Tzz, double *__restrict__ Txy,
                                                    double *__restrict__ Tyz, double *__restrict__ Tzx,
                                                    double *__restrict__ part, int *__restrict__ fail) {
    int i, j, k;
    row_cell(G, i, j, k);
    double v[3] = {0.0, 0.0, 0.0};
    if (k < G.n[2]) {
        const CoreField V[3] = {V0, V1, V2};
        double T[6], K, S[6];
        if (!sgs_cell<SF32>(G, V, nu, i, j, k, T, K, S)) {
            atomicOr(fail, 1);
            T[0] = T[1] = T[2] = T[3] = T[4] = T[5] = 0.0;
            K = 0.0;
        }
        const long idx = (long)i * G.sx + (long)j * G.sy + k;
        Txx[idx] = T[0];
        Tyy[idx] = T[1];
        Tzz[idx] = T[2];
        Txy[idx] = T[3];
        Tyz[idx] = T[4];
        Tzx[idx] = T[5];
        const double A = (1.0 / G.n[0]) / G.xi_x[0][i], B = (1.0 / G.n[1]) / G.xi_x[1][j];
        const double C = (1.0 / G.n[2]) / G.xi_x[2][k];
        v[0] = fabs(K) * A * B * C;
        const double sgDiss = -(S[0] * T[0] + S[1] * T[1] + S[2] * T[2] + 2.0 * (S[3] * T[3] + S[4] * T[4] + S[5] * T[5]));
        v[1] = sgDiss * A * B * C;
        const double sNorm = sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.0 * (S[3] * S[3] + S[4] * S[4] + S[5] * S[5]));
        const double tNorm = sqrt(T[0] * T[0] + T[1] * T[1] + T[2] * T[2] + 2.0 * (T[3] * T[3] + T[4] * T[4] + T[5] * T[5]));
        v[2] = ((tNorm / sNorm) / 2) * A * B * C;
    }
    __shared__ double red[3][4];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double x = v[q];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double x = red[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x += red[threadIdx.x][w];
        part[(((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = x;
    }
}

// sums of n triples -> out[3] (one block of 1024 threads, fixed order)
__global__ void __launch_bounds__(1024) k_sum3_final(const double *__restrict__ part, long n, double *__restrict__ out) {
    __shared__ double red[3][16];
    double a[3] = {0.0, 0.0, 0.0};
    // unrolled so that several loads are in flight; the additions keep their order
#pragma unroll 8
    for (long q = threadIdx.x; q < n; q += blockDim.x)
        for (int c = 0; c < 3; ++c) a[c] += part[3 * q + c];
    for (int c = 0; c < 3; ++c) {
        double x = a[c];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double x = red[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x += red[threadIdx.x][w];
        out[threadIdx.x] = x;
    }
}

// Block max of non-negative doubles (NaN wins, as in a failed comparison
// chain) written to part[block].
__device__ __forceinline__ void block_max_store(double v, double *part, long block) {
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (w > v || w != w) ? w : v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = (red[q] > m || red[q] != red[q]) ? red[q] : m;
        part[block] = m;
    }
}

// max over n partials -> out[0] (one block of 1024 threads)
// out and seq are pinned host-coherent words: the host polls seq for the
// ticket, which is stored after the maximum with release ordering.
__global__ void __launch_bounds__(1024) k_max_final(const double *__restrict__ part, long n, double *__restrict__ out,
                                                    unsigned *seq, unsigned ticket) {
    __shared__ double red[16];
    double v = 0.0;
    // unrolled so that several loads are in flight (the maximum does not
    // depend on the order)
#pragma unroll 8
    for (long q = threadIdx.x; q < n; q += blockDim.x) {
        const double w = part[q];
        v = (w > v || w != w) ? w : v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (w > v || w != w) ? w : v;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = (red[q] > m || red[q] != red[q]) ? red[q] : m;
        out[0] = m;
        __hip_atomic_store(seq, ticket, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// One Jacobi iteration of the Crank-Nicolson solve (predictors.cc:59-145):
// out = update(F).  With `slot`, also the residual max|F - beta dt nu lap2 F - rhs|
// of the input F (the reference's check after its previous update).
template <bool RES, class FF>
__device__ __forceinline__ double jacobi_cell(const Geom &G, const FF &F, const double *__restrict__ rhs,
                                              double *__restrict__ out, double dt, double diff, double beta, int i,
                                              int j, int k) {
    double res = 0.0;
    {
        const long idx = (long)i * G.sx + (long)j * G.sy + k;
        const double C = F.at(G, i, j, k);
        const double E = F.nb(G, 0, i, j, k, 1), W = F.nb(G, 0, i, j, k, -1);
        const double N = F.nb(G, 1, i, j, k, 1), S = F.nb(G, 1, i, j, k, -1);
        const double Tp = F.nb(G, 2, i, j, k, 1), B = F.nb(G, 2, i, j, k, -1);
        const double x2 = G.xix2[0][i], xx = G.xixx[0][i];
        const double y2 = G.xix2[1][j], yy = G.xixx[1][j];
        const double z2 = G.xix2[2][k], zz = G.xixx[2][k];
        const double r = rhs[idx];
        out[idx] = ((G.tsih2[0] * x2 * (E + W) + G.tsi2h[0] * xx * (E - W) + G.tsih2[1] * y2 * (N + S) +
                     G.tsi2h[1] * yy * (N - S) + G.tsih2[2] * z2 * (Tp + B) + G.tsi2h[2] * zz * (Tp - B)) *
                        dt * diff * beta +
                    r) /
                   (1.0 + 2.0 * dt * diff * beta * (G.tsih2[0] * x2 + G.tsih2[1] * y2 + G.tsih2[2] * z2));
        if (RES) {
            const double t = C - beta * dt * diff *
                                     (x2 * (E - 2.0 * C + W) * G.tsih2[0] + xx * (E - W) * G.tsi2h[0] +
                                      y2 * (N - 2.0 * C + S) * G.tsih2[1] + yy * (N - S) * G.tsi2h[1] +
                                      z2 * (Tp - 2.0 * C + B) * G.tsih2[2] + zz * (Tp - B) * G.tsi2h[2]);
            res = fabs(t - r);
        }
    }
    return res;
}

template <bool RES>
__global__ void __launch_bounds__(256) k_jacobi(Geom G, CoreField F, const double *__restrict__ rhs,
                                                double *__restrict__ out, double dt, double diff, double beta,
                                                double *__restrict__ part) {
    int i, j, k;
    row_cell(G, i, j, k);
    double res = 0.0;
    if (k < G.n[2]) {
        if (interior2(G, i, j, k)) res = jacobi_cell<RES>(G, fast_of(F), rhs, out, dt, diff, beta, i, j, k);
        else res = jacobi_cell<RES>(G, F, rhs, out, dt, diff, beta, i, j, k);
    }
    if (RES) block_max_store(res, part, ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
}

// vfield::divergence (vfield.cc:721-737) scaled into the Poisson RHS:
// b = (((0 + dVx/dx) + dVy/dy) + dVz/dz) * sc   (lsRK3_d3.cc:278-279)
__global__ void __launch_bounds__(256) k_div(Geom G, CoreField V0, CoreField V1, CoreField V2, double sc,
                                             double *__restrict__ b, long bsx, long bsy, long bhz) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    double d = 0.0;
    if (interior2(G, i, j, k)) {
        d += d1(G, fast_of(V0), 0, i, j, k);
        d += d1(G, fast_of(V1), 1, i, j, k);
        d += d1(G, fast_of(V2), 2, i, j, k);
    } else {
        d += d1(G, V0, 0, i, j, k);
        d += d1(G, V1, 1, i, j, k);
        d += d1(G, V2, 2, i, j, k);
    }
    b[(long)i * bsx + (long)j * bsy + (bhz ? (long)(k & 1) * bhz + (k >> 1) : (long)k)] = d * sc;
}

// P += Pp; V -= ((alpha+beta) dt) grad Pp   (lsRK3_d3.cc:285-295)
template <class PF>
__global__ void __launch_bounds__(256) k_correct(Geom G, PF Pp, double *__restrict__ P, FieldBC pb,
                                                 double *__restrict__ V0, double *__restrict__ V1,
                                                 double *__restrict__ V2, double pc) {
    int i, j, k;
    row_cell(G, i, j, k);
    if (k >= G.n[2]) return;
    const long idx = (long)i * G.sx + (long)j * G.sy + k;
    P[idx] += Pp.at(G, i, j, k);
    // P += Pp covers the ghost layers too (sfield.cc:565-569): a nullBC wall of
    // P keeps that sum, so the ghost beside a wall cell takes Pp's ghost there
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int q = pick(ax, i, j, k);
        if (q == 0 && pb.type[2 * ax] == 4)
            pb.ghost[2 * ax][ghost_plane_index(G.n, ax, i, j, k)] += Pp.nb(G, ax, i, j, k, -1);
        if (q == G.n[ax] - 1 && pb.type[2 * ax + 1] == 4)
            pb.ghost[2 * ax + 1][ghost_plane_index(G.n, ax, i, j, k)] += Pp.nb(G, ax, i, j, k, 1);
    }
    double g0, g1, g2;
    if (interior2(G, i, j, k)) {
        const auto Pf = fast_of(Pp);
        g0 = d1(G, Pf, 0, i, j, k);
        g1 = d1(G, Pf, 1, i, j, k);
        g2 = d1(G, Pf, 2, i, j, k);
    } else {
        g0 = d1(G, Pp, 0, i, j, k);
        g1 = d1(G, Pp, 1, i, j, k);
        g2 = d1(G, Pp, 2, i, j, k);
    }
    g0 *= pc;
    g1 *= pc;
    g2 *= pc;
    V0[idx] -= g0;
    V1[idx] -= g1;
    V2[idx] -= g2;
}

// tseries diagnostics, per-block partial sums (tseries.cc:209-290):
// [0] sum KE dVol, [1] sum Vz T dVol, [2] max |div V|, [3] sum bottom-wall dTdn dA,
// [4] sum top-wall dTdn dA.  dxi = 1/n per axis; wall_dn = the top-wall distance.
constexpr int kDiagBlocks = 1024;
constexpr int kDiagSums = 6;   // KE, uz T, max|div|, bottom and top wall flux, thermal energy
constexpr int kDiagThreads = 256;

__global__ void __launch_bounds__(kDiagThreads) k_diag(Geom G, CoreField V0, CoreField V1, CoreField V2,
                                                       CoreField T, int has_T, double dxi0, double dxi1,
                                                       double dxi2, double wall_dn, double *__restrict__ part) {
    __shared__ double red[kDiagSums][kDiagThreads / 64];
    const long total = (long)G.n[0] * G.n[1] * G.n[2];
    double ke = 0.0, uzt = 0.0, dm = 0.0, nb = 0.0, nt = 0.0, te = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
        const int k = (int)(c % G.n[2]);
        const int j = (int)((c / G.n[2]) % G.n[1]);
        const int i = (int)(c / ((long)G.n[2] * G.n[1]));
        const double dVol = (dxi0 / G.xi_x[0][i]) * (dxi1 / G.xi_x[1][j]) * (dxi2 / G.xi_x[2][k]);
        const double u = V0.at(G, i, j, k), v = V1.at(G, i, j, k), w = V2.at(G, i, j, k);
        ke += (u * u + v * v + w * w) * 0.5 * dVol;
        double d = 0.0;
        d += d1(G, V0, 0, i, j, k);
        d += d1(G, V1, 1, i, j, k);
        d += d1(G, V2, 2, i, j, k);
        d = fabs(d);
        dm = (d > dm || d != d) ? d : dm;
        if (has_T) {
            const double t = T.at(G, i, j, k);
            uzt += w * t * dVol;
            // theta = T + z - 1: the deviation from the conduction profile (tseries.cc:423-429)
            const double th = t + G.xc[2][k] - 1.0;
            te += th * th * 0.5 * dVol;
            const double dA = (dxi0 / G.xi_x[0][i]) * (dxi1 / G.xi_x[1][j]);
            if (k == 0) nb += (1.0 - fabs(t)) / wall_dn * dA;
            if (k == G.n[2] - 1) nt += fabs(t) / wall_dn * dA;
        }
    }
    double vals[kDiagSums] = {ke, uzt, dm, nb, nt, te};
    for (int q = 0; q < kDiagSums; ++q) {
        double v = vals[q];
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_xor(v, o, 64);
            if (q == 2) v = (w > v || w != w) ? w : v;
            else v += w;
        }
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < kDiagSums) {
        const int q = threadIdx.x;
        double v = red[q][0];
        for (int w = 1; w < kDiagThreads / 64; ++w) {
            if (q == 2) v = (red[q][w] > v || red[q][w] != red[q][w]) ? red[q][w] : v;
            else v += red[q][w];
        }
        part[(long)blockIdx.x * kDiagSums + q] = v;
    }
}

// probes::getData (probes.cc:185-205): Vx, Vy, Vz, P, T at core cells
// (i, j, k) of the list; T reads 0 without a scalar field
__global__ void __launch_bounds__(256) k_probe(Geom G, const double *__restrict__ u, const double *__restrict__ v,
                                               const double *__restrict__ w, const double *__restrict__ p,
                                               const double *__restrict__ t, int n, const int *__restrict__ ijk,
                                               double *__restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = (long)ijk[3 * q] * G.sx + (long)ijk[3 * q + 1] * G.sy + ijk[3 * q + 2];
    out[5 * q + 0] = u[o];
    out[5 * q + 1] = v[o];
    out[5 * q + 2] = w[o];
    out[5 * q + 3] = p[o];
    out[5 * q + 4] = t ? t[o] : 0.0;
}

// vfield::computeTStp (vfield.cc:691-706): per-block maxima of
// |u|/dx + |v|/dy + |w|/dz over the core
__global__ void __launch_bounds__(256) k_cfl(Geom G, const double *__restrict__ u, const double *__restrict__ v,
                                             const double *__restrict__ w, double ax, double ay, double az,
                                             double *__restrict__ part) {
    int i, j, k;
    row_cell(G, i, j, k);
    double m = 0.0;
    if (k < G.n[2]) {
        const long idx = (long)i * G.sx + (long)j * G.sy + k;
        m = (fabs(u[idx]) / ax + fabs(v[idx]) / ay) + fabs(w[idx]) / az;
    }
    block_max_store(m, part, ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x);
}

// caller array (pad p) -> core
// py: the caller array's y pad (0 for a planar run, whose arrays have no y extent)
__global__ void k_field_load(const double *__restrict__ src, long SX, long SY, int p, int py, double *__restrict__ dst,
                             long sx, long sy, long hz, int nx, int ny, int nz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= nz) return;
    dst[(long)i * sx + (long)j * sy + (hz ? (long)(k & 1) * hz + (k >> 1) : (long)k)] =
        src[(long)(i + p) * SX + (long)(j + py) * SY + (k + p)];
}

// core + face ghost cells -> caller array (pad p): every cell with at most one
// index outside the core gets the value the reference's field holds after
// imposeBCs (periodic pads p deep, first wall layer); edge/corner pads and the
// second wall layer are left as the caller had them.
template <class FF>
__global__ void k_field_store(Geom G, FF F, double *__restrict__ dst, long SX, long SY, int p, int py) {
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;   // over nz + 2p
    const int jj = blockIdx.y, ii = blockIdx.z;               // over ny + 2py, nx + 2p
    if (kk >= G.n[2] + 2 * p) return;
    const int c[3] = {ii - p, jj - py, kk - p};
    int outside = 0, ax = -1;
    for (int a = 0; a < 3; ++a)
        if (c[a] < 0 || c[a] >= G.n[a]) {
            ++outside;
            ax = a;
        }
    const long o = (long)ii * SX + (long)jj * SY + kk;
    if (outside == 0) {
        dst[o] = F.at(G, c[0], c[1], c[2]);
        return;
    }
    if (outside > 1) return;
    const int q = c[ax];
    if (!G.per[ax] && (q < -1 || q > G.n[ax])) return;
    int base[3] = {c[0], c[1], c[2]};
    base[ax] = q < 0 ? 0 : G.n[ax] - 1;
    const int off = q < 0 ? q : q - (G.n[ax] - 1);
    dst[o] = F.nb(G, ax, base[0], base[1], base[2], off);
}

// Single field operators on caller arrays (ghosts as supplied):
// 0-2 first derivative along x/y/z, 3-5 second derivative (out = value),
// 6 computeDiff (out += lap F), 7 sfield::computeNLin with V (out += H).
template <int OP>
__global__ void k_field_op(Geom G, PadField F, const CoreField V0, const CoreField V1, const CoreField V2,
                           int nlscheme, double kappa, double *__restrict__ out, long SX, long SY, int p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= G.n[2]) return;
    const long o = (long)(i + p) * SX + (long)(j + p) * SY + (k + p);
    if (OP <= 2) {
        out[o] = d1(G, F, OP, i, j, k);
    } else if (OP <= 5) {
        out[o] = d2(G, F, OP - 3, i, j, k);
    } else if (OP == 6) {
        double h = out[o];
        h += d2(G, F, 0, i, j, k);
        h += d2(G, F, 1, i, j, k);
        h += d2(G, F, 2, i, j, k);
        out[o] = h;
    } else {
        const CoreField V[3] = {V0, V1, V2};
        out[o] = nlin_acc(G, nlscheme, V, F, i, j, k, out[o], kappa);
    }
}

// The vector operators over caller arrays (ghost values as supplied), one
// core cell per thread: 0 vfield::divergence (vfield.cc:721-737: ox = 0 +
// dVx/dx + dVy/dy + dVz/dz), 1 sfield::gradient of A (sfield.cc:496-508:
// (ox, oy, oz) = grad A), 2 vfield::computeDiff (vfield.cc:150-192: o_c +=
// d2xx; += d2yy; += d2zz of component c), 3 vfield::computeNLin of the field
// by itself (vfield.cc:205-689: o_c -= the x, y, z terms of component c for
// the non-linear scheme; diffusivity nu for the upwind Peclet number).
template <int OP>
__global__ void k_field_vec_op(Geom G, PadField A, PadField B, PadField Cc, int nlscheme, double nu,
                               double *__restrict__ ox, double *__restrict__ oy, double *__restrict__ oz, long SX,
                               long SY, int p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= G.n[2]) return;
    const long o = (long)(i + p) * SX + (long)(j + p) * SY + (k + p);
    if (OP == 0) {
        double d = 0.0;
        d += d1(G, A, 0, i, j, k);
        d += d1(G, B, 1, i, j, k);
        d += d1(G, Cc, 2, i, j, k);
        ox[o] = d;
    } else if (OP == 1) {
        ox[o] = d1(G, A, 0, i, j, k);
        oy[o] = d1(G, A, 1, i, j, k);
        oz[o] = d1(G, A, 2, i, j, k);
    } else if (OP == 2) {
        const PadField F[3] = {A, B, Cc};
        double *out[3] = {ox, oy, oz};
        for (int c = 0; c < 3; ++c) {
            double h = out[c][o];
            h += d2(G, F[c], 0, i, j, k);
            h += d2(G, F[c], 1, i, j, k);
            h += d2(G, F[c], 2, i, j, k);
            out[c][o] = h;
        }
    } else {
        const PadField V[3] = {A, B, Cc};
        ox[o] = nlin_acc(G, nlscheme, V, A, i, j, k, ox[o], nu);
        oy[o] = nlin_acc(G, nlscheme, V, B, i, j, k, oy[o], nu);
        oz[o] = nlin_acc(G, nlscheme, V, Cc, i, j, k, oz[o], nu);
    }
}

// plumeOut's flags at T.imposeBCs: top-wall column (i, j) is set where the
// ghost Vz (Vz's own top-wall rule) is positive (plumeOut.cc:71-90).
__global__ void __launch_bounds__(256) k_plume_flags(Geom G, CoreField vz, unsigned char *__restrict__ flags) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)G.n[0] * G.n[1]) return;
    const int i = (int)(c / G.n[1]), j = (int)(c % G.n[1]);
    flags[c] = vz.nb(G, 2, i, j, G.n[2] - 1, 1) > 0 ? 1 : 0;
}

}  // namespace nse
}  // namespace saras
