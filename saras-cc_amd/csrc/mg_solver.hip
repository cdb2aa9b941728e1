// mg_solver.hip -- device-resident multigrid hierarchy and the C-ABI of
// include/saras_hip.h.
//
// The host drives the reference's V-cycle (lib/poisson/poisson.cc:203-279)
// level by level on one HIP stream; all level arrays stay in HBM between
// calls, as the reference's poisson object keeps its hierarchy between time
// steps (timestep.h:142).  Two arithmetic modes:
//   SARAS_PREC_FP64  - the reference algorithm in FP64, bit-exact with the
//                      FP64 CPU build for the L-inf residual type;
//   SARAS_PREC_MIXED - FP64 residual and iterative refinement around FP32
//                      V-cycles on the correction equation A e = r.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <array>
#include <vector>

#include "grid_host.hpp"
#include "mg_fused.hpp"
#include "mg_internal.hpp"
#include "mg_kernels.hpp"
#include "mg_split.hpp"
#include "mg_tail.hpp"
#include "saras_hip.h"

namespace {

thread_local std::string g_last_error;

void set_error(const std::string &s) { g_last_error = s; }

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));               \
            return e_ == hipErrorOutOfMemory ? SARAS_OOM : SARAS_HIP_ERROR;             \
        }                                                                               \
    } while (0)

constexpr int kThreads = 256;
// default dynamic LDS budget of the single-workgroup tail (SARAS_TAIL_LDS_KB
// overrides it, up to 150 KB of the 160 KB a workgroup may hold: the tail's
// static LDS is under 2 KB); the tail also takes only levels its threads
// cover with a cell each or with colour lists (see tail_first_level)
constexpr size_t kTailLdsDefault = 128 * 1024;
constexpr int kTailThreads = 512;

inline unsigned blocks_for(long n) { return (unsigned)std::max<long>(1, (n + kThreads - 1) / kThreads); }

template <typename T>
long row_pitch(int nz) {
    const long a = 128 / (long)sizeof(T);
    return (nz + a - 1) / a * a;
}

template <typename T>
struct DevLevel {
    int n[3] = {0, 0, 0};
    long sx = 0, sy = 0;
    size_t elems = 0;  // allocation size in elements
    T *x = nullptr, *b = nullptr;
    T *xalt = nullptr;  // second iterate buffer: fused sweeps run out of place
    long hz = 0;        // colour-split rows when > 0 (offset of the odd half)
    T *c2zs = nullptr, *c1zs = nullptr;  // z coefficients in split order
    T *c2[3] = {nullptr, nullptr, nullptr}, *c1[3] = {nullptr, nullptr, nullptr};
    T *rlo[3] = {nullptr, nullptr, nullptr}, *rhi[3] = {nullptr, nullptr, nullptr}, *rwd[3] = {nullptr, nullptr, nullptr};
};

template <typename T>
saras_status upload(T **dst, const std::vector<double> &src) {
    std::vector<T> tmp(src.begin(), src.end());
    HIP_TRY(hipMalloc((void **)dst, std::max<size_t>(1, tmp.size()) * sizeof(T)));
    HIP_TRY(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return SARAS_OK;
}

template <typename T>
struct Hierarchy {
    int dim = 3;
    std::vector<DevLevel<T>> lv;
    T *wall[3] = {nullptr, nullptr, nullptr};   // DIRICHLET_TEST level-0 wall values
    T *coarse_scratch = nullptr;
    bool coarse_lds = true;
    size_t coarse_lds_bytes = 0;
    bool split_rows = false;   // store levels whose rows split into whole vectors colour-split
    // host copies of the metric coefficients (c2, c1) and restriction
    // factors (lo, hi, wd) per level and axis, for the tail's packed image
    std::vector<std::array<std::vector<double>, 3>> hc2, hc1, hlo, hhi, hwd;
    // the tail kernel's coefficient region of LDS for tail levels
    // tail_first..D, packed in its order (see k_tail)
    T *tail_coef = nullptr;
    int tail_first = -1;
    size_t tail_coef_n = 0;

    // a level takes split rows when each half-row is a whole number of 16-byte vectors
    static bool splittable(int nz) {
        constexpr int W = 16 / (int)sizeof(T);
        return nz >= 2 * W && nz % (2 * W) == 0;
    }

    ~Hierarchy() {
        for (auto &L : lv) {
            (void)hipFree(L.x); (void)hipFree(L.b); (void)hipFree(L.xalt);
            (void)hipFree(L.c2zs); (void)hipFree(L.c1zs);
            for (int a = 0; a < 3; ++a) {
                (void)hipFree(L.c2[a]); (void)hipFree(L.c1[a]);
                (void)hipFree(L.rlo[a]); (void)hipFree(L.rhi[a]); (void)hipFree(L.rwd[a]);
            }
        }
        for (int a = 0; a < 3; ++a) (void)hipFree(wall[a]);
        (void)hipFree(coarse_scratch);
        (void)hipFree(tail_coef);
    }

    // The tail's coefficient image for levels S..D: per level and axis c2
    // then c1 (one zero each for the planar y axis), then for levels S+1..D
    // of a 3-D hierarchy per axis the restriction factors lo, hi, wd.
    saras_status pack_tail(int S) {
        const int D = (int)lv.size() - 1;
        if (S > D) return SARAS_OK;
        std::vector<double> flat;
        for (int l = S; l <= D; ++l)
            for (int a = 0; a < 3; ++a) {
                flat.insert(flat.end(), hc2[l][a].begin(), hc2[l][a].end());
                flat.insert(flat.end(), hc1[l][a].begin(), hc1[l][a].end());
            }
        if (dim == 3)
            for (int l = S + 1; l <= D; ++l)
                for (int a = 0; a < 3; ++a) {
                    flat.insert(flat.end(), hlo[l][a].begin(), hlo[l][a].end());
                    flat.insert(flat.end(), hhi[l][a].begin(), hhi[l][a].end());
                    flat.insert(flat.end(), hwd[l][a].begin(), hwd[l][a].end());
                }
        (void)hipFree(tail_coef);
        tail_coef = nullptr;
        saras_status st = upload(&tail_coef, flat);
        if (st) return st;
        tail_first = S;
        tail_coef_n = flat.size();
        return SARAS_OK;
    }

    saras_status build(const saras::GridLevels &G, const std::vector<std::vector<double>> &walls) {
        dim = G.dim;
        const int D = G.depth;
        lv.resize(D + 1);
        hc2.assign(D + 1, {});
        hc1.assign(D + 1, {});
        hlo.assign(D + 1, {});
        hhi.assign(D + 1, {});
        hwd.assign(D + 1, {});
        for (int l = 0; l <= D; ++l) {
            DevLevel<T> &L = lv[l];
            for (int a = 0; a < 3; ++a) L.n[a] = G.axes[l][a].n;
            if (dim == 2) L.n[1] = 1;
            L.sy = row_pitch<T>(L.n[2]);
            L.sx = L.sy * (long)L.n[1];
            L.elems = (size_t)L.sx * L.n[0];
            L.hz = split_rows && splittable(L.n[2]) ? L.sy / 2 : 0;
            HIP_TRY(hipMalloc((void **)&L.x, L.elems * sizeof(T)));
            HIP_TRY(hipMalloc((void **)&L.b, L.elems * sizeof(T)));
            HIP_TRY(hipMemset(L.x, 0, L.elems * sizeof(T)));
            if (l < D && dim == 3 && L.hz == 0) {
                HIP_TRY(hipMalloc((void **)&L.xalt, L.elems * sizeof(T)));
                HIP_TRY(hipMemset(L.xalt, 0, L.elems * sizeof(T)));
            }
            HIP_TRY(hipMemset(L.b, 0, L.elems * sizeof(T)));
            for (int a = 0; a < 3; ++a) {
                if (dim == 2 && a == 1) {
                    std::vector<double> z(1, 0.0);
                    if (upload(&L.c2[a], z) || upload(&L.c1[a], z)) return SARAS_HIP_ERROR;
                    hc2[l][a] = z;
                    hc1[l][a] = z;
                    continue;
                }
                const saras::AxisLevel &A = G.axes[l][a];
                // poisson::setCoefficients (poisson.cc:463-482)
                const int hInc = 1 << l;
                const double dxi = 1.0 / double(G.axes[0][a].n);
                const double h = hInc * dxi;
                const double h2 = std::pow(h, 2);
                const double i2h = 0.5 / h, ih2 = 1.0 / h2;
                std::vector<double> c2(A.n), c1(A.n);
                for (int i = 0; i < A.n; ++i) {
                    c2[i] = A.xix2[i + 1] * ih2;
                    c1[i] = A.xixx[i + 1] * i2h;
                }
                if (upload(&L.c2[a], c2) || upload(&L.c1[a], c1)) return SARAS_HIP_ERROR;
                hc2[l][a] = c2;
                hc1[l][a] = c1;
                if (a == 2 && L.hz) {
                    const int half = A.n / 2;
                    std::vector<double> c2s(A.n), c1s(A.n);
                    for (int hh = 0; hh < 2; ++hh)
                        for (int mm = 0; mm < half; ++mm) {
                            c2s[hh * half + mm] = c2[2 * mm + hh];
                            c1s[hh * half + mm] = c1[2 * mm + hh];
                        }
                    if (upload(&L.c2zs, c2s) || upload(&L.c1zs, c1s)) return SARAS_HIP_ERROR;
                }
                if (l > 0) {
                    // restriction factors of coarse index I (poisson_d3.cc:494-502)
                    const saras::AxisLevel &Fa = G.axes[l - 1][a];
                    std::vector<double> lo(A.n), hi(A.n), wd(A.n);
                    for (int I = 0; I < A.n; ++I) {
                        const double X = A.coord[I + 1];
                        const double f0 = Fa.coord[2 * I + 1], f1 = Fa.coord[2 * I + 2];
                        lo[I] = X - f0;
                        hi[I] = f1 - X;
                        wd[I] = f1 - f0;
                    }
                    if (upload(&L.rlo[a], lo) || upload(&L.rhi[a], hi) || upload(&L.rwd[a], wd))
                        return SARAS_HIP_ERROR;
                    hlo[l][a] = lo;
                    hhi[l][a] = hi;
                    hwd[l][a] = wd;
                }
            }
        }
        for (int a = 0; a < 3; ++a)
            if (!walls[a].empty() && upload(&wall[a], walls[a])) return SARAS_HIP_ERROR;
        // coarsest-level workspace
        const DevLevel<T> &C = lv[D];
        const size_t padded = dim == 3 ? (size_t)(C.n[0] + 2) * (C.n[1] + 2) * (C.n[2] + 2)
                                       : (size_t)(C.n[0] + 2) * (C.n[2] + 2);
        const size_t need = (padded + (size_t)C.n[0] * C.n[1] * C.n[2]) * sizeof(T);
        // a workgroup may hold up to 64 KB of dynamic LDS without opting in
        // (k_coarse's static LDS is 128 bytes): 14^3 in FP64 fits
        coarse_lds = need <= 64 * 1024 - 256;
        coarse_lds_bytes = coarse_lds ? need : 0;
        if (!coarse_lds) HIP_TRY(hipMalloc((void **)&coarse_scratch, need));
        return SARAS_OK;
    }
};

int bc_code(const saras_mg_params &p, const saras_grid_desc &g, int axis, int level, bool inhom_allowed, bool hi) {
    if (g.periodic[axis]) return saras::BC_PER;
    switch (p.bc) {
        // multigrid_d2::imposeBC, all-Neumann branch: probType 9 turns the
        // top wall Dirichlet (poisson_d2.cc:641-647)
        case SARAS_BC_NEUMANN: return (hi && axis == 2 && p.top_dirichlet) ? saras::BC_DIR : saras::BC_NEU;
        case SARAS_BC_DIRICHLET: return saras::BC_DIR;
        default: return (level == 0 && inhom_allowed) ? saras::BC_DIRV : saras::BC_DIR;
    }
}

}  // namespace

struct saras_mg {
    saras_grid_desc g{};
    saras_mg_params p{};
    saras::GridLevels G;
    int depth = 0;
    int max_count = 0;
    Hierarchy<double> h64;
    std::unique_ptr<Hierarchy<float>> h32;
    double *x64b = nullptr;            // second FP64 iterate buffer (MIXED)
    double *partial = nullptr;         // norm partials
    // The norm's last reduction and the coarse solves write these pinned,
    // host-coherent words directly, with no copy kernel in between: the host
    // reads the norm once its ticket shows it written, and a coarse count
    // once a later norm has been read (kernels of a stream run in order).
    double *norm_host = nullptr;       // [numerator, denominator] of computeError
    unsigned *norm_seq = nullptr;      // ticket stored after norm_host (see wait_ticket)
    unsigned norm_ticket = 0;          // ticket of the last norm launched
    int *iters_host = nullptr;         // coarse iteration count of the last coarse solve
    double *stage_x = nullptr, *stage_b = nullptr;  // device staging for host pointers
    size_t stage_elems = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int coarse_iters_max = 0;
    int coarse_capped = 0;
    bool fused_off = false;            // SARAS_NO_FUSED_SWEEP=1 forces the two-pass smoother
    bool tail_off = false;             // SARAS_NO_TAIL=1 runs every level with multi-block kernels
    int march_wgs = 0;                 // target workgroups of a march sweep (SARAS_MARCH_WGS; 0 = per type)
    bool march_tj8 = false;            // FP32 march with 64 x 8 tiles (SARAS_MARCH_TJ8=1)
    bool split_off = false;            // SARAS_NO_SPLIT=1 keeps every level z-contiguous
    size_t tail_lds = kTailLdsDefault;  // dynamic LDS budget of the tail kernel
    bool red_fold = true;              // SARAS_NO_RED_FOLD=1: the restriction leaves the first red pass to the sweep
    bool rr_lane = false;              // SARAS_RR_LANE=1: 3-D split restriction with all eight children in one lane
    // V-cycles replayed from captured hipGraphs (saras_mg_set_graphs).  A
    // graph bakes in the level buffers, and out-of-place sweeps swap them, so
    // graphs are cached per buffer state (at most two states per precision).
    bool graphs = false;
    struct GraphEntry {
        std::vector<void *> key;    // (x, xalt) of every level before the cycle
        std::vector<void *> post;   // ... and after it
        hipGraphExec_t exec;
        hipStream_t stream;
    };
    std::vector<GraphEntry> graph_cache[2];   // [0] FP64 cycles, [1] FP32 cycles

    ~saras_mg() {
        for (auto &c : graph_cache)
            for (auto &e : c) (void)hipGraphExecDestroy(e.exec);
        (void)hipFree(x64b); (void)hipFree(partial);
        (void)hipHostFree(norm_host); (void)hipHostFree(iters_host);
        (void)hipHostFree(norm_seq);
        (void)hipFree(stage_x); (void)hipFree(stage_b);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
saras::LevelArgs<T> level_args(const saras_mg *m, const Hierarchy<T> &H, int l, bool inhom) {
    const DevLevel<T> &L = H.lv[l];
    saras::LevelArgs<T> a{};
    for (int q = 0; q < 3; ++q) {
        a.n[q] = L.n[q];
        a.c2[q] = L.c2[q];
        a.c1[q] = L.c1[q];
        a.bc_lo[q] = bc_code(m->p, m->g, q, l, inhom, false);
        a.bc_hi[q] = bc_code(m->p, m->g, q, l, inhom, true);
        a.wall[q] = H.wall[q];
    }
    a.sx = L.sx;
    a.sy = L.sy;
    a.hz = L.hz;
    a.c2zs = L.c2zs;
    a.c1zs = L.c1zs;
    a.rtile = 0;
    // A colour pass reads row (i+1, j) again two i-planes after its first
    // read.  Where those two planes of x and b exceed half of an XCD's 4 MB
    // L2, the rows run in j-tiles of about 256 KB of such reuse instead.
    if (L.hz && H.dim == 3) {
        const size_t half = (size_t)L.n[2] / 2 * sizeof(T);
        if (4 * (size_t)L.n[1] * half > (2u << 20)) {
            int jt = 1;
            while (jt < 256 && (size_t)(2 * jt) * 4 * half <= (256u << 10)) jt *= 2;
            while (jt > 1 && L.n[1] % jt) jt /= 2;
            if (jt >= 8) a.rtile = jt;
        }
    }
    return a;
}

// first level of the single-workgroup tail: the smallest S whose levels S..D
// fit in the tail's LDS budget (D + 1, i.e. no tail, when even the coarsest
// level does not fit)
template <typename T>
int tail_first_level(const saras_mg *m, const Hierarchy<T> &H) {
    const int D = (int)H.lv.size() - 1;
    size_t bytes = 0;
    int S = D + 1;
    for (int l = D; l >= 0; --l) {
        const DevLevel<T> &L = H.lv[l];
        const size_t core = (size_t)L.n[0] * L.n[1] * L.n[2];
        const size_t pad = H.dim == 3 ? (size_t)(L.n[0] + 2) * (L.n[1] + 2) * (L.n[2] + 2)
                                      : (size_t)(L.n[0] + 2) * (L.n[2] + 2);
        // metric coefficients (2 per index) and restriction factors (3)
        const size_t coef = 5 * ((size_t)L.n[0] + L.n[1] + L.n[2]);
        bytes += ((pad + 3) / 4 * 4 + (core + 3) / 4 * 4 + coef) * sizeof(T);
        if (bytes > m->tail_lds || D - l + 1 > saras::kTailMaxLevels) break;
        // a cell per thread, or colour lists (even z extent, at most
        // kTailColourCells cells of a colour per thread); larger levels
        // would walk every cell in each colour pass and run faster as
        // multi-workgroup passes
        const bool one = core <= (size_t)kTailThreads;
        const bool lists = L.n[2] % 2 == 0 && core / 2 <= (size_t)saras::kTailColourCells * kTailThreads;
        if (l < D && !one && !lists) break;
        S = l;
    }
    return S;
}

// Level operations for one precision.  `inhom0` selects the inhomogeneous
// DIRICHLET_TEST wall values on level 0 (false for the correction equation
// of the mixed-precision refinement, whose walls are homogeneous).
template <typename T>
struct Ops {
    saras_mg *m;
    Hierarchy<T> &H;
    hipStream_t s;
    bool inhom0;
    T w;

    int pend_level = -1;  // level whose next sweep also applies the prolongation
    // level 0's iterate is zero and has not been written: set by the mixed
    // solve instead of clearing it, for a V-cycle whose first operation is a
    // split-row sweep of level 0 (the conditions are its `zero_start`)
    bool zero0 = false;
    int skip_red = -1;    // level whose next sweep's red pass the restriction already applied
    // a host-side guard that refused a launch: the caller returns this status
    // (with the message in saras_last_error) instead of the process aborting
    saras_status err = SARAS_OK;

    bool refuse(const char *what) {
        if (err == SARAS_OK) set_error(std::string("internal error: ") + what);
        err = SARAS_HIP_ERROR;
        return false;
    }

    bool inhom(int l, bool zero_bc) const { return inhom0 && l == 0 && !zero_bc; }

    // FP32 cycles solve the coarsest level to max(mgTolerance, kCoarseFp32Rel
    // max|b| of that level): the FP32 residual of a level of up to a few
    // thousand cells stalls near 1e-5 of its right-hand side
    static constexpr double kCoarseFp32Rel = 1e-4;
    double coarse_rel_floor() const { return sizeof(T) == 4 ? kCoarseFp32Rel : 0.0; }

    // one RB-SOR sweep at level l: the fused out-of-place march kernel where
    // the level tiles exactly, else the two in-place colour passes
    template <int NP, int TJ>
    void march(int l, const saras::LevelArgs<T> &a) {
        DevLevel<T> &L = H.lv[l];
        const int ntk = L.n[2] / (2 * NP), ntj = L.n[1] / TJ;
        const int tiles2d = ntk * ntj;
        const int target = m->march_wgs > 0 ? m->march_wgs : (sizeof(T) == 4 ? 512 : 2048);
        int nchunks = (target + tiles2d - 1) / tiles2d;
        nchunks = std::max(1, std::min(nchunks, L.n[0] / 8));
        const int chunk = (L.n[0] + nchunks - 1) / nchunks;
        nchunks = (L.n[0] + chunk - 1) / chunk;
        const int ntiles = nchunks * tiles2d;
        const int grid = (ntiles + saras::kXcds - 1) / saras::kXcds * saras::kXcds;
        const T *e = nullptr;
        long esx = 0, esy = 0;
        if (pend_level == l) {  // prolongation folded into this first post-smoothing sweep
            const DevLevel<T> &C = H.lv[l + 1];
            e = C.x;
            esx = C.sx;
            esy = C.sy;
            pend_level = -1;
        }
        if (l == 0)
            hipLaunchKernelGGL((saras::k_sweep_march<T, NP, TJ, true>), dim3(grid), dim3(NP * TJ), 0, s, a, L.x, L.b,
                               L.xalt, w, chunk, ntk, ntj, ntiles, e, esx, esy);
        else
            hipLaunchKernelGGL((saras::k_sweep_march<T, NP, TJ, false>), dim3(grid), dim3(NP * TJ), 0, s, a, L.x,
                               L.b, L.xalt, w, chunk, ntk, ntj, ntiles, e, esx, esy);
        std::swap(L.x, L.xalt);
    }

    bool march_ok(int l) const {
        const DevLevel<T> &L = H.lv[l];
        if (H.dim != 3 || !L.xalt || m->fused_off || L.hz) return false;
        for (int q = 0; q < 3; ++q)
            if (m->g.periodic[q] && (L.n[q] & 1)) return false;
        // the march needs many tiles and planes to pay off; smaller levels use
        // the two colour passes
        return L.n[1] % 16 == 0 && L.n[2] % 64 == 0 && (long)L.n[0] * L.n[1] * L.n[2] >= (8L << 20);
    }

    template <int DIM, bool ZP, bool FINE, bool PRO>
    void split_launch(const dim3 &g, const saras::LevelArgs<T> &a, DevLevel<T> &L, int colour, int nb, const T *e,
                      long esx, long esy, long ehz, bool z = false) {
        const dim3 bl(saras::kSplitThreads);
        // z: the finest level's first red pass on an iterate known to be zero
        if constexpr (FINE && !PRO) {
            if (z) {
                if (w == T(1)) {
                    if (ZP) hipLaunchKernelGGL((saras::k_rb_split_zp<T, DIM, FINE, PRO, true, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
                    else hipLaunchKernelGGL((saras::k_rb_split<T, DIM, FINE, PRO, true, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
                } else {
                    if (ZP) hipLaunchKernelGGL((saras::k_rb_split_zp<T, DIM, FINE, PRO, false, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
                    else hipLaunchKernelGGL((saras::k_rb_split<T, DIM, FINE, PRO, false, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
                }
                return;
            }
        }
        // w == 1 (every reference deck): the pass that skips the cells' own values
        if (w == T(1)) {
            if (ZP) hipLaunchKernelGGL((saras::k_rb_split_zp<T, DIM, FINE, PRO, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
            else hipLaunchKernelGGL((saras::k_rb_split<T, DIM, FINE, PRO, true>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
        } else {
            if (ZP) hipLaunchKernelGGL((saras::k_rb_split_zp<T, DIM, FINE, PRO, false>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
            else hipLaunchKernelGGL((saras::k_rb_split<T, DIM, FINE, PRO, false>), g, bl, 0, s, a, L.x, L.b, colour, w, nb, e, esx, esy, ehz);
        }
    }

    template <int DIM>
    void split_pass(bool fine, bool pro, int grid, const saras::LevelArgs<T> &a, DevLevel<T> &L, int colour, int nb,
                    const T *e, long esx, long esy, long ehz, bool z = false) {
        const dim3 g(grid);
        // rows that wrap in z, or rows split between waves, need the
        // z-prefetching variant (see rb_split_body)
        const int q = L.n[2] / 2 / saras::Vec16<T>::W;
        const bool zp = a.bc_lo[2] == saras::BC_PER || !(q % 64 == 0 || 64 % q == 0);
        if (zp) {
            if (fine) {
                if (pro) split_launch<DIM, true, true, true>(g, a, L, colour, nb, e, esx, esy, ehz);
                else split_launch<DIM, true, true, false>(g, a, L, colour, nb, e, esx, esy, ehz, z);
            } else {
                if (pro) split_launch<DIM, true, false, true>(g, a, L, colour, nb, e, esx, esy, ehz);
                else split_launch<DIM, true, false, false>(g, a, L, colour, nb, e, esx, esy, ehz);
            }
            return;
        }
        if (fine) {
            if (pro) split_launch<DIM, false, true, true>(g, a, L, colour, nb, e, esx, esy, ehz);
            else split_launch<DIM, false, true, false>(g, a, L, colour, nb, e, esx, esy, ehz, z);
        } else {
            if (pro) split_launch<DIM, false, false, true>(g, a, L, colour, nb, e, esx, esy, ehz);
            else split_launch<DIM, false, false, false>(g, a, L, colour, nb, e, esx, esy, ehz);
        }
    }

    void rb_sweep(int l, bool zero_bc) {
        DevLevel<T> &L = H.lv[l];
        auto a = level_args(m, H, l, inhom(l, zero_bc));
        if (L.hz) {
            constexpr int W = saras::Vec16<T>::W;
            const long total = (long)L.n[0] * L.n[1] * (L.n[2] / 2 / W);
            const int nb = (int)((total + saras::kSplitThreads - 1) / saras::kSplitThreads);
            const int grid = (nb + saras::kXcds - 1) / saras::kXcds * saras::kXcds;
            // the first sweep after the prolongation applies it (see k_rb_split)
            const bool pro = pend_level == l;
            const T *e = nullptr;
            long esx = 0, esy = 0, ehz = 0;
            if (pro) {
                const DevLevel<T> &C = H.lv[l + 1];
                e = C.x;
                esx = C.sx;
                esy = C.sy;
                ehz = C.hz;
                pend_level = -1;
            }
            const int first = skip_red == l ? 1 : 0;
            if (first) skip_red = -1;
            // the finest level's iterate is known to be zero: its first red
            // pass reads no x and clears the black half itself
            const bool z = l == 0 && zero0 && first == 0 && !pro;
            if (l == 0) zero0 = false;
            for (int colour = first; colour < 2; ++colour) {
                const bool zc = z && colour == 0;
                if (H.dim == 3) split_pass<3>(l == 0, pro, grid, a, L, colour, nb, e, esx, esy, ehz, zc);
                else split_pass<2>(l == 0, pro, grid, a, L, colour, nb, e, esx, esy, ehz, zc);
            }
            return;
        }
        if (march_ok(l)) {
            // FP32: 64 x 16 tiles (512 lanes); FP64: 64 x 8 tiles to keep LDS per workgroup ~50 KB
            constexpr int TJ = sizeof(T) == 4 ? 16 : 8;
            if (sizeof(T) == 4 && m->march_tj8) march<32, 8>(l, a);
            else if (L.n[2] % 64 == 0 && L.n[1] % TJ == 0) march<32, TJ>(l, a);
            else if (L.n[2] % 32 == 0) march<16, 8>(l, a);
            else march<8, 8>(l, a);
            return;
        }
        const long n = (long)L.n[0] * L.n[1] * ((L.n[2] + 1) / 2);
        for (int colour = 0; colour < 2; ++colour) {
            if (H.dim == 3) {
                if (l == 0)
                    hipLaunchKernelGGL((saras::k_rb_pass<T, 3, true>), dim3(blocks_for(n)), dim3(kThreads), 0, s, a, L.x, L.b, colour, w);
                else
                    hipLaunchKernelGGL((saras::k_rb_pass<T, 3, false>), dim3(blocks_for(n)), dim3(kThreads), 0, s, a, L.x, L.b, colour, w);
            } else {
                if (l == 0)
                    hipLaunchKernelGGL((saras::k_rb_pass<T, 2, true>), dim3(blocks_for(n)), dim3(kThreads), 0, s, a, L.x, L.b, colour, w);
                else
                    hipLaunchKernelGGL((saras::k_rb_pass<T, 2, false>), dim3(blocks_for(n)), dim3(kThreads), 0, s, a, L.x, L.b, colour, w);
            }
        }
    }

    void coarse(int l, bool zero_bc, int mode, int sweeps) {
        DevLevel<T> &L = H.lv[l];
        auto a = level_args(m, H, l, inhom(l, zero_bc));
        saras::CoarseCtl ctl{mode, sweeps, m->max_count, m->p.mg_tol, m->iters_host, coarse_rel_floor()};
        const long cells = (long)L.n[0] * L.n[1] * L.n[2];
        int threads = 64;  // power of two: the kernel's tree reduction relies on it
        while (threads < cells && threads < 1024) threads <<= 1;
        if (H.dim == 3)
            hipLaunchKernelGGL((saras::k_coarse<T, 3>), dim3(1), dim3(threads), H.coarse_lds_bytes, s, a, L.x, L.b,
                               H.coarse_scratch, H.coarse_lds ? 1 : 0, w, ctl);
        else
            hipLaunchKernelGGL((saras::k_coarse<T, 2>), dim3(1), dim3(threads), H.coarse_lds_bytes, s, a, L.x, L.b,
                               H.coarse_scratch, H.coarse_lds ? 1 : 0, w, ctl);
    }

    // poisson smooth(count) at level l (coarsest level: literal kernel)
    void smooth(int l, int count, bool zero_bc) {
        if (count <= 0) return;
        if (l == (int)H.lv.size() - 1) {
            coarse(l, zero_bc, 1, count);
            return;
        }
        for (int c = 0; c < count; ++c) rb_sweep(l, zero_bc);
    }

    // red_next: the restriction also applies the coarse level's first
    // pre-smoothing red pass (split levels; see k_rr_split)
    void resid_restrict(int l, bool red_next = false) {
        // host-side guard: never launch past the hierarchy
        if (l < 0 || l + 1 >= (int)H.lv.size()) {
            refuse("restriction from a level outside the hierarchy");
            return;
        }
        DevLevel<T> &F = H.lv[l], &C = H.lv[l + 1];
        auto fa = level_args(m, H, l, inhom(l, false));
        auto ca = level_args(m, H, l + 1, false);
        saras::RestrictArgs<T> r{};
        for (int q = 0; q < 3; ++q) {
            r.lo[q] = C.rlo[q];
            r.hi[q] = C.rhi[q];
            r.wd[q] = C.rwd[q];
        }
        const long n = (long)C.n[0] * C.n[1] * C.n[2];
        if (F.hz && H.dim == 3 && !m->rr_lane) {
            const long ntask = (long)C.n[0] * C.n[1] * (F.n[2] / 2 / saras::Vec16<T>::W);
            const int nb = (int)((ntask + saras::kRrTasks - 1) / saras::kRrTasks);
            const int grid = (nb + saras::kXcds - 1) / saras::kXcds * saras::kXcds;
            const bool red = red_next && C.hz;
            if (red) skip_red = l + 1;
            if (red)
                hipLaunchKernelGGL((saras::k_rr_split4<T, true>), dim3(grid), dim3(saras::kSplitThreads), 0, s, fa,
                                   F.x, F.b, ca, C.b, C.x, r, nb, w);
            else
                hipLaunchKernelGGL((saras::k_rr_split4<T, false>), dim3(grid), dim3(saras::kSplitThreads), 0, s, fa,
                                   F.x, F.b, ca, C.b, C.x, r, nb, w);
            return;
        }
        if (F.hz) {
            constexpr int W = saras::rr_width<T>();
            const long total = (long)C.n[0] * C.n[1] * (F.n[2] / 2 / W);
            const int nb = (int)((total + saras::kSplitThreads - 1) / saras::kSplitThreads);
            const int grid = (nb + saras::kXcds - 1) / saras::kXcds * saras::kXcds;
            const bool red = red_next && C.hz;
            if (red) skip_red = l + 1;
            if (H.dim == 3) {
                if (red)
                    hipLaunchKernelGGL((saras::k_rr_split<T, 3, true>), dim3(grid), dim3(saras::kSplitThreads), 0, s,
                                       fa, F.x, F.b, ca, C.b, C.x, r, nb, w);
                else
                    hipLaunchKernelGGL((saras::k_rr_split<T, 3, false>), dim3(grid), dim3(saras::kSplitThreads), 0, s,
                                       fa, F.x, F.b, ca, C.b, C.x, r, nb, w);
            } else {
                if (red)
                    hipLaunchKernelGGL((saras::k_rr_split<T, 2, true>), dim3(grid), dim3(saras::kSplitThreads), 0, s,
                                       fa, F.x, F.b, ca, C.b, C.x, r, nb, w);
                else
                    hipLaunchKernelGGL((saras::k_rr_split<T, 2, false>), dim3(grid), dim3(saras::kSplitThreads), 0, s,
                                       fa, F.x, F.b, ca, C.b, C.x, r, nb, w);
            }
            return;
        }
        if (H.dim == 3)
            hipLaunchKernelGGL((saras::k_resid_restrict<T, 3>), dim3(blocks_for(n)), dim3(kThreads), 0, s, fa, F.x, F.b,
                               ca, C.b, C.x, r);
        else
            hipLaunchKernelGGL((saras::k_resid_restrict<T, 2>), dim3(blocks_for(n)), dim3(kThreads), 0, s, fa, F.x, F.b,
                               ca, C.b, C.x, r);
    }

    void prolong_correct(int l) {
        // host-side guard: never launch past the hierarchy
        if (l < 0 || l + 1 >= (int)H.lv.size()) {
            refuse("prolongation onto a level outside the hierarchy");
            return;
        }
        DevLevel<T> &F = H.lv[l], &C = H.lv[l + 1];
        const long n = (long)F.n[0] * F.n[1] * F.n[2];
        if (F.hz) {
            constexpr int W = saras::Vec16<T>::W;
            const long total = (long)F.n[0] * F.n[1] * (F.n[2] / 2 / W);
            const int nb = (int)((total + saras::kSplitThreads - 1) / saras::kSplitThreads);
            const int grid = (nb + saras::kXcds - 1) / saras::kXcds * saras::kXcds;
            hipLaunchKernelGGL((saras::k_prolong_split<T>), dim3(grid), dim3(saras::kSplitThreads), 0, s, F.n[0],
                               F.n[1], F.n[2], F.sx, F.sy, F.hz, F.x, C.x, C.sx, C.sy, C.hz, nb);
            return;
        }
        hipLaunchKernelGGL((saras::k_prolong_correct<T>), dim3(blocks_for(n)), dim3(kThreads), 0, s, F.n[0], F.n[1],
                           F.n[2], F.sx, F.sy, F.hz, F.x, C.x, C.sx, C.sy, C.hz);
    }

    int tail_start() const { return tail_first_level(m, H); }

    void tail(int S) {
        const int D = (int)H.lv.size() - 1;
        saras::TailArgs<T> A{};
        A.first = S;
        A.last = D;
        size_t off = 0;
        for (int l = S; l <= D; ++l) {
            const DevLevel<T> &L = H.lv[l];
            A.L[l - S] = level_args(m, H, l, inhom(l, false));
            for (int q = 0; q < 3; ++q) {
                A.R[l - S].lo[q] = L.rlo[q];
                A.R[l - S].hi[q] = L.rhi[q];
                A.R[l - S].wd[q] = L.rwd[q];
            }
            const size_t core = (size_t)L.n[0] * L.n[1] * L.n[2];
            const size_t pad = H.dim == 3 ? (size_t)(L.n[0] + 2) * (L.n[1] + 2) * (L.n[2] + 2)
                                          : (size_t)(L.n[0] + 2) * (L.n[2] + 2);
            A.xoff[l - S] = (int)off;
            off += (pad + 3) / 4 * 4;
            A.boff[l - S] = (int)off;
            off += (core + 3) / 4 * 4;
        }
        A.coff = (int)off;
        for (int l = S; l <= D; ++l) {
            const DevLevel<T> &L = H.lv[l];
            off += 2 * ((size_t)L.n[0] + (H.dim == 3 ? L.n[1] : 1) + L.n[2]);
            if (H.dim == 3 && l > S) off += 3 * ((size_t)L.n[0] + L.n[1] + L.n[2]);   // restriction factors
        }
        A.x = H.lv[S].x;
        A.b = H.lv[S].b;
        A.pre = m->p.pre_smooth;
        A.post = m->p.post_smooth;
        A.solve_flag = m->p.solve_flag;
        A.max_count = m->max_count;
        A.w = w;
        A.tol = m->p.mg_tol;
        A.rel_floor = coarse_rel_floor();
        A.zero_bc_first = S != 0;
        A.iters_out = m->iters_host;
        size_t ncoef = 0;
        for (int l = S; l <= D; ++l)
            for (int q = 0; q < 3; ++q) {
                ncoef += 2 * (size_t)(H.dim == 2 && q == 1 ? 1 : H.lv[l].n[q]);
                if (H.dim == 3 && l > S) ncoef += 3 * (size_t)H.lv[l].n[q];
            }
        if (H.tail_first != S || !H.tail_coef || H.tail_coef_n != ncoef) {
            refuse("tail coefficients not packed for this tail");
            return;
        }
        A.coef = H.tail_coef;
        // host-side guard: tail_start() keeps the levels within tail_lds
        if (off * sizeof(T) > m->tail_lds) {
            refuse("tail levels exceed the LDS budget");
            return;
        }
        if (H.dim == 3)
            hipLaunchKernelGGL((saras::k_tail<T, 3>), dim3(1), dim3(kTailThreads), off * sizeof(T), s, A);
        else
            hipLaunchKernelGGL((saras::k_tail<T, 2>), dim3(1), dim3(kTailThreads), off * sizeof(T), s, A);
    }

    // poisson::vCycle (poisson.cc:203-279); levels S..D run in the tail kernel
    void vcycle() {
        const int D = (int)H.lv.size() - 1;
        const int S = m->tail_off ? D + 1 : tail_start();
        if (S == 0) {
            tail(0);
            return;
        }
        smooth(0, m->p.pre_smooth, false);
        for (int l = 0; l < D; ++l) {
            resid_restrict(l, m->red_fold && l + 1 < S && l + 1 < D && m->p.pre_smooth > 0);
            if (l + 1 == S) break;
            if (l + 1 == D) {
                if (m->p.solve_flag) coarse(D, true, 0, 0);
                else smooth(D, m->p.pre_smooth, true);
            } else {
                smooth(l + 1, m->p.pre_smooth, true);
            }
        }
        if (S <= D) tail(S);
        for (int l = std::min(S, D) - 1; l >= 0; --l) {
            if (m->p.post_smooth > 0 && (march_ok(l) || H.lv[l].hz)) pend_level = l;
            else prolong_correct(l);
            smooth(l, m->p.post_smooth, l != 0);
        }
    }
};

// One V-cycle of `ops`, eager or replayed from the handle's graph cache.
template <typename T>
saras_status run_vcycle(saras_mg *m, Ops<T> &ops) {
    if (!m->graphs) {
        ops.vcycle();
        return ops.err;
    }
    auto &cache = m->graph_cache[sizeof(T) == 4 ? 1 : 0];
    std::vector<void *> key;
    for (auto &L : ops.H.lv) {
        key.push_back((void *)L.x);
        key.push_back((void *)L.xalt);
    }
    for (auto &e : cache)
        if (e.key == key && e.stream == ops.s) {
            HIP_TRY(hipGraphLaunch(e.exec, ops.s));
            // the replayed cycle's buffer swaps, applied on the host side
            for (size_t l = 0; l < ops.H.lv.size(); ++l) {
                ops.H.lv[l].x = (T *)e.post[2 * l];
                ops.H.lv[l].xalt = (T *)e.post[2 * l + 1];
            }
            return SARAS_OK;
        }
    hipGraph_t g = nullptr;
    std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
    HIP_TRY(hipStreamBeginCapture(ops.s, hipStreamCaptureModeThreadLocal));
    ops.vcycle();
    hipError_t ce = hipStreamEndCapture(ops.s, &g);
    if (ops.err) {
        if (g) (void)hipGraphDestroy(g);
        return ops.err;
    }
    if (ce != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        return SARAS_HIP_ERROR;
    }
    hipGraphExec_t exec = nullptr;
    hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        return SARAS_HIP_ERROR;
    }
    std::vector<void *> post;
    for (auto &L : ops.H.lv) {
        post.push_back((void *)L.x);
        post.push_back((void *)L.xalt);
    }
    cache.push_back({key, post, exec, ops.s});
    HIP_TRY(hipGraphLaunch(exec, ops.s));
    return SARAS_OK;
}

template <int DIM>
void launch_norm(int type, saras_mg *m, const saras::LevelArgs<double> &a, const double *x, const double *b,
                 hipStream_t s) {
    if (a.hz) {
        // split rows: the vectorised residual pass, norms only
        const unsigned nb = saras::kNormBlocks, nt = saras::kNormThreads;
        if (type == 0)
            hipLaunchKernelGGL((saras::k_refine_split<0, DIM, true>), dim3(nb), dim3(nt), 0, s, a, x, nullptr, 0L, 0L,
                               0L, b, nullptr, nullptr, m->partial, saras::FullRows{});
        else if (type == 1)
            hipLaunchKernelGGL((saras::k_refine_split<1, DIM, true>), dim3(nb), dim3(nt), 0, s, a, x, nullptr, 0L, 0L,
                               0L, b, nullptr, nullptr, m->partial, saras::FullRows{});
        else
            hipLaunchKernelGGL((saras::k_refine_split<2, DIM, true>), dim3(nb), dim3(nt), 0, s, a, x, nullptr, 0L, 0L,
                               0L, b, nullptr, nullptr, m->partial, saras::FullRows{});
        if (type == 0)
            hipLaunchKernelGGL((saras::k_norm_final<0>), dim3(1), dim3(nt), 0, s, m->partial, saras::kNormBlocks,
                               m->norm_host, m->norm_seq, ++m->norm_ticket);
        else if (type == 1)
            hipLaunchKernelGGL((saras::k_norm_final<1>), dim3(1), dim3(nt), 0, s, m->partial, saras::kNormBlocks,
                               m->norm_host, m->norm_seq, ++m->norm_ticket);
        else
            hipLaunchKernelGGL((saras::k_norm_final<2>), dim3(1), dim3(nt), 0, s, m->partial, saras::kNormBlocks,
                               m->norm_host, m->norm_seq, ++m->norm_ticket);
        return;
    }
    switch (type) {
        case 0:
            hipLaunchKernelGGL((saras::k_norm_partial<0, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0,
                               s, a, x, b, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<0>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
        case 1:
            hipLaunchKernelGGL((saras::k_norm_partial<1, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0,
                               s, a, x, b, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<1>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
        default:
            hipLaunchKernelGGL((saras::k_norm_partial<2, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0,
                               s, a, x, b, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<2>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
    }
}

// the norm of the last k_norm_final launched on s, once it is on the host
static hipError_t wait_norm(saras_mg *m, hipStream_t s) {
    return saras::wait_ticket(m->norm_seq, m->norm_ticket, s);
}

// computeError (poisson_d3.cc:563-663): ratio formed on the host exactly as
// the reference does.
saras_status compute_error(saras_mg *m, int type, const double *x, hipStream_t s, double *out) {
    auto a = level_args(m, m->h64, 0, true);
    if (m->G.dim == 3) launch_norm<3>(type, m, a, x, m->h64.lv[0].b, s);
    else launch_norm<2>(type, m, a, x, m->h64.lv[0].b, s);
    HIP_TRY(wait_norm(m, s));
    const double num = m->norm_host[0], den = m->norm_host[1];
    const auto &L = m->h64.lv[0];
    const int pointCount = L.n[0] * L.n[1] * L.n[2];
    double r;
    if (type == 0) r = den != 0.0 ? num / den : num;
    else if (type == 1) r = den != 0.0 ? num / den : num / pointCount;
    else r = den != 0.0 ? std::sqrt(num / pointCount) / std::sqrt(den / pointCount) : std::sqrt(num / pointCount);
    *out = r;
    return SARAS_OK;
}

// the last coarse solve's iteration count, once a norm launched after it has
// been read
void note_coarse_iters(saras_mg *m) {
    const int it = *m->iters_host;
    m->coarse_iters_max = std::max(m->coarse_iters_max, it);
    if (it > m->max_count) m->coarse_capped = 1;
}


long full_elems(const saras_grid_desc &g) {
    const long p = g.pad;
    if (g.dim == 3) return (g.n[0] + 2 * p) * (long)(g.n[1] + 2 * p) * (g.n[2] + 2 * p);
    return (g.n[0] + 2 * p) * (long)(g.n[2] + 2 * p);
}

template <typename T>
void load_core(saras_mg *m, const double *full, const DevLevel<T> &L, T *dst, hipStream_t s) {
    const long n = (long)L.n[0] * L.n[1] * L.n[2];
    hipLaunchKernelGGL((saras::k_load_core<T>), dim3(blocks_for(n)), dim3(kThreads), 0, s, full, m->g.pad, m->g.dim,
                       L.n[0], L.n[1], L.n[2], L.sx, L.sy, L.hz, dst);
}

// core + the [-1, n] ghost shell, phased as the reference's final imposeBC;
// core == nullptr: the core is in `full` already, only the shell is written
void store_full(saras_mg *m, double *full, const double *core, hipStream_t s) {
    const auto &L = m->h64.lv[0];
    const long n = (long)L.n[0] * L.n[1] * L.n[2];
    if (core)
        hipLaunchKernelGGL(saras::k_store_core, dim3(blocks_for(n)), dim3(kThreads), 0, s, full, m->g.pad, m->g.dim,
                           L.n[0], L.n[1], L.n[2], L.sx, L.sy, L.hz, core);
    const int ext[3] = {L.n[0] + 2, m->g.dim == 3 ? L.n[1] + 2 : 1, L.n[2] + 2};
    for (int op = 0; op < 2; ++op)
        for (int a = 0; a < 3; ++a) {
            if (m->g.dim == 2 && a == 1) continue;
            const bool per = m->g.periodic[a] != 0;
            if ((op == 0) != per) continue;
            const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
            const long cnt = (long)ext[o1] * ext[o2];
            const int bc_lo = bc_code(m->p, m->g, a, 0, true, false), bc_hi = bc_code(m->p, m->g, a, 0, true, true);
            hipLaunchKernelGGL(saras::k_ghost_phase, dim3(blocks_for(cnt)), dim3(kThreads), 0, s, full, m->g.pad,
                               m->g.dim, L.n[0], L.n[1], L.n[2], a, op, bc_lo, bc_hi, (const double *)m->h64.wall[a]);
        }
}

std::vector<std::vector<double>> make_walls(const saras_grid_desc &g, const saras_mg_params &p) {
    std::vector<std::vector<double>> w(3);
    if (p.bc != SARAS_BC_DIRICHLET_TEST) return w;
    // initDirichlet (poisson_d3.cc:824-873, poisson_d2.cc:514-543): r^2/6 on
    // the walls from the field-grid cell centres; zero on the ghost shell
    std::vector<double> fx, fy, fz, t1, t2, t3;
    saras::field_axis(g, 0, fx, t1, t2, t3);
    saras::field_axis(g, 2, fz, t1, t2, t3);
    const int p0 = g.pad;
    const int nx = g.n[0], nz = g.n[2];
    if (g.dim == 3) {
        saras::field_axis(g, 1, fy, t1, t2, t3);
        const int ny = g.n[1];
        w[0].assign((size_t)(ny + 2) * (nz + 2), 0.0);
        w[1].assign((size_t)(nx + 2) * (nz + 2), 0.0);
        w[2].assign((size_t)(nx + 2) * (ny + 2), 0.0);
        double xD, yD, zD;
        xD = g.len[0] / 2.0;
        for (int j = 0; j < ny; ++j) {
            yD = fy[j + p0] - 0.5;
            for (int k = 0; k < nz; ++k) {
                zD = fz[k + p0] - 0.5;
                w[0][(size_t)(j + 1) * (nz + 2) + k + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
            }
        }
        yD = g.len[1] / 2.0;
        for (int i = 0; i < nx; ++i) {
            xD = fx[i + p0] - 0.5;
            for (int k = 0; k < nz; ++k) {
                zD = fz[k + p0] - 0.5;
                w[1][(size_t)(i + 1) * (nz + 2) + k + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
            }
        }
        zD = g.len[2] / 2.0;
        for (int i = 0; i < nx; ++i) {
            xD = fx[i + p0] - 0.5;
            for (int j = 0; j < ny; ++j) {
                yD = fy[j + p0] - 0.5;
                w[2][(size_t)(i + 1) * (ny + 2) + j + 1] = (xD * xD + yD * yD + zD * zD) / 6.0;
            }
        }
    } else {
        w[0].assign((size_t)(nz + 2), 0.0);
        w[2].assign((size_t)(nx + 2), 0.0);
        double xD = g.len[0] / 2.0, zD;
        for (int k = 0; k < nz; ++k) {
            zD = fz[k + p0] - 0.5;
            w[0][k + 1] = (xD * xD + zD * zD) / 4.0;
        }
        zD = g.len[2] / 2.0;
        for (int i = 0; i < nx; ++i) {
            xD = fx[i + p0] - 0.5;
            w[2][i + 1] = (xD * xD + zD * zD) / 4.0;
        }
    }
    return w;
}

hipStream_t pick_stream(saras_mg *m, void *stream) { return stream ? (hipStream_t)stream : m->stream; }


}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *saras_last_error(void) { return g_last_error.c_str(); }

int saras_vc_depth(const saras_grid_desc *grid) { return grid ? saras::vc_depth(*grid) : -1; }

saras_status saras_level_metrics(const saras_grid_desc *grid, int level, int axis, double *coord, double *xi,
                                 double *xi_x, double *xixx, double *xix2) {
    if (!grid || axis < 0 || axis > 2) return SARAS_BAD_ARG;
    saras::GridLevels G;
    if (!saras::build_levels(*grid, G)) {
        set_error("grid cannot be coarsened");
        return SARAS_BAD_SHAPE;
    }
    if (level < 0 || level > G.depth) return SARAS_BAD_ARG;
    const saras::AxisLevel &A = G.axes[level][axis];
    const size_t m = A.coord.size();
    if (coord) std::memcpy(coord, A.coord.data(), m * sizeof(double));
    if (xi) std::memcpy(xi, A.xi.data(), m * sizeof(double));
    if (xi_x) std::memcpy(xi_x, A.xi_x.data(), m * sizeof(double));
    if (xixx) std::memcpy(xixx, A.xixx.data(), m * sizeof(double));
    if (xix2) std::memcpy(xix2, A.xix2.data(), m * sizeof(double));
    return SARAS_OK;
}

saras_status saras_mg_create(const saras_grid_desc *grid, const saras_mg_params *params, saras_mg **out) {
    if (!grid || !params || !out) {
        set_error("null argument");
        return SARAS_BAD_ARG;
    }
    *out = nullptr;
    const saras_grid_desc &g = *grid;
    const saras_mg_params &p = *params;
    if (g.dim != 2 && g.dim != 3) { set_error("dim must be 2 or 3"); return SARAS_BAD_ARG; }
    for (int a = 0; a < 3; ++a) {
        if (g.dim == 2 && a == 1) continue;
        if (g.n[a] < 2) { set_error("every axis needs at least 2 cells"); return SARAS_BAD_SHAPE; }
        if (!(g.len[a] > 0)) { set_error("domain lengths must be positive"); return SARAS_BAD_ARG; }
    }
    if (p.res_type < 0 || p.res_type > 2) { set_error("Residual Type must be 0, 1 or 2"); return SARAS_BAD_ARG; }
    if (p.bc < 0 || p.bc > 2 || p.precision < 0 || p.precision > 1) { set_error("bad bc/precision"); return SARAS_BAD_ARG; }
    if (p.top_dirichlet && !(g.dim == 2 && p.bc == SARAS_BC_NEUMANN && !g.periodic[2])) {
        set_error("top_dirichlet: 2-D Neumann problems with z walls (poisson_d2.cc:641-647)");
        return SARAS_BAD_ARG;
    }
    if (p.vc_count < 0 || p.pre_smooth < 0 || p.post_smooth < 0) { set_error("negative counts"); return SARAS_BAD_ARG; }
    std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
    std::unique_ptr<saras_mg> m(new saras_mg());
    m->g = g;
    m->p = p;
    {
        const char *env = std::getenv("SARAS_NO_FUSED_SWEEP");
        m->fused_off = env && env[0] == '1';
        const char *env2 = std::getenv("SARAS_NO_TAIL");
        m->tail_off = env2 && env2[0] == '1';
        const char *env3 = std::getenv("SARAS_MARCH_WGS");
        if (env3 && std::atoi(env3) > 0) m->march_wgs = std::atoi(env3);
        const char *env4 = std::getenv("SARAS_MARCH_TJ8");
        m->march_tj8 = env4 && env4[0] == '1';
        const char *env5 = std::getenv("SARAS_NO_SPLIT");
        m->split_off = env5 && env5[0] == '1';
        const char *env9 = std::getenv("SARAS_RR_LANE");
        m->rr_lane = env9 && env9[0] == '1';
        const char *env8 = std::getenv("SARAS_NO_RED_FOLD");
        m->red_fold = !(env8 && env8[0] == '1');
        const char *env6 = std::getenv("SARAS_TAIL_LDS_KB");
        if (env6 && std::atoi(env6) > 0) m->tail_lds = (size_t)std::min(150, std::atoi(env6)) * 1024;
    }
    if (m->p.inner_cycles <= 0) m->p.inner_cycles = 3;
    if (m->p.max_refine <= 0) m->p.max_refine = 20;
    if (!saras::build_levels(g, m->G)) {
        set_error("grid cannot be coarsened (setVCDepth gives a negative depth)");
        return SARAS_BAD_SHAPE;
    }
    m->depth = m->G.depth;
    {
        const auto &C = m->G.axes[m->depth];
        const int cx = C[0].n, cy = g.dim == 3 ? C[1].n : 1, cz = C[2].n;
        m->max_count = g.dim == 2 ? (int)(2.5 * cx * cz) : (int)(2.5 * cx * cy * cz);
    }
    const auto walls = make_walls(g, p);
    m->h64.split_rows = !m->split_off;
    saras_status st = m->h64.build(m->G, walls);
    if (st) return st;
    if (p.precision == SARAS_PREC_MIXED) {
        m->h32.reset(new Hierarchy<float>());
        m->h32->split_rows = !m->split_off;
        std::vector<std::vector<double>> nowall(3);
        st = m->h32->build(m->G, nowall);
        if (st) return st;
        HIP_TRY(hipMalloc((void **)&m->x64b, m->h64.lv[0].elems * sizeof(double)));
        st = m->h32->pack_tail(tail_first_level(m.get(), *m->h32));
        if (st) return st;
    }
    st = m->h64.pack_tail(tail_first_level(m.get(), m->h64));
    if (st) return st;
    HIP_TRY(hipMalloc((void **)&m->partial, 2 * saras::kNormBlocks * sizeof(double)));
    HIP_TRY(hipHostMalloc((void **)&m->norm_host, 2 * sizeof(double), hipHostMallocCoherent));
    HIP_TRY(hipHostMalloc((void **)&m->norm_seq, sizeof(unsigned), hipHostMallocCoherent));
    *m->norm_seq = 0;
    HIP_TRY(hipHostMalloc((void **)&m->iters_host, sizeof(int), hipHostMallocCoherent));
    *m->iters_host = 0;
    HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&m->ev0));
    HIP_TRY(hipEventCreate(&m->ev1));
    // the level arrays were cleared on the null stream; the handle's stream
    // does not synchronise with it
    HIP_TRY(hipDeviceSynchronize());
    *out = m.release();
    return SARAS_OK;
}

saras_status saras_mg_destroy(saras_mg *mg) {
    if (!mg) return SARAS_BAD_ARG;
    (void)hipStreamSynchronize(mg->stream);
    std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
    delete mg;
    return SARAS_OK;
}

int saras_mg_levels(const saras_mg *mg) { return mg ? mg->depth + 1 : -1; }

saras_status saras_mg_level_shape(const saras_mg *mg, int level, int n[3]) {
    if (!mg || !n || level < 0 || level > mg->depth) return SARAS_BAD_ARG;
    for (int a = 0; a < 3; ++a) n[a] = mg->h64.lv[level].n[a];
    return SARAS_OK;
}

static saras_status solve_fp64(saras_mg *m, hipStream_t s, saras_mg_stats *st) {
    Ops<double> ops{m, m->h64, s, true, m->p.sor};
    auto &L0 = m->h64.lv[0];
    int cycles = 0;
    double res = 0.0;
    bool conv = false;
    for (int c = 0; c < m->p.vc_count; ++c) {
        saras_status ve = run_vcycle(m, ops);
        if (ve) return ve;
        saras_status e = compute_error(m, m->p.res_type, L0.x, s, &res);
        if (e) return e;
        // compute_error read a norm launched after this cycle's coarse solve
        if (m->p.solve_flag) note_coarse_iters(m);
        cycles = c + 1;
        if (res < m->p.vc_tol) {
            conv = true;
            break;
        }
    }
    st->cycles = cycles;
    st->refinements = 0;
    st->residual = res;
    st->converged = conv;
    return SARAS_OK;
}

}  // extern "C"

template <int TYPE, int DIM, bool CX, bool CB, bool CXO = false, bool NOXW = false>
void launch_refine_split(saras_mg *m, const saras::LevelArgs<double> &a, const float *e, long esx, long esy, long ehz,
                         float *r32, const saras::FullRows &fr, hipStream_t s) {
    auto &L0 = m->h64.lv[0];
    const unsigned nb = saras::kNormBlocks, nt = saras::kNormThreads;
    hipLaunchKernelGGL((saras::k_refine_split<TYPE, DIM, false, CX, CB, CXO, NOXW>), dim3(nb), dim3(nt), 0, s, a,
                       L0.x, e, esx, esy, ehz, L0.b, m->x64b, r32, m->partial, fr);
    hipLaunchKernelGGL((saras::k_norm_final<TYPE>), dim3(1), dim3(nt), 0, s, m->partial, saras::kNormBlocks,
                       m->norm_host, m->norm_seq, ++m->norm_ticket);
}

template <int TYPE, int DIM>
void launch_refine_split(saras_mg *m, const saras::LevelArgs<double> &a, const float *e, long esx, long esy, long ehz,
                         float *r32, const saras::FullRows &fr, hipStream_t s) {
    // x_old in the caller's array with no correction: the initial residual,
    // whose iterate stays there (nothing to write but r32)
    if (fr.x && !e) launch_refine_split<TYPE, DIM, true, true, false, true>(m, a, e, esx, esy, ehz, r32, fr, s);
    else if (fr.x) launch_refine_split<TYPE, DIM, true, true>(m, a, e, esx, esy, ehz, r32, fr, s);
    else if (fr.xo) launch_refine_split<TYPE, DIM, false, true, true>(m, a, e, esx, esy, ehz, r32, fr, s);
    else if (fr.b) launch_refine_split<TYPE, DIM, false, true>(m, a, e, esx, esy, ehz, r32, fr, s);
    else launch_refine_split<TYPE, DIM, false, false>(m, a, e, esx, esy, ehz, r32, fr, s);
}

// fr: the caller's arrays when x_old and/or b are read from them instead of
// level 0, or x_new written there (fr.xo, with b from the caller and x_old
// from level 0); empty otherwise
template <int DIM>
void launch_refine(int type, saras_mg *m, const saras::LevelArgs<double> &a, const float *e, long esx, long esy,
                   long ehz, float *r32, const saras::FullRows &fr, hipStream_t s) {
    auto &L0 = m->h64.lv[0];
    if (L0.hz && ehz) {
        if (type == 0) launch_refine_split<0, DIM>(m, a, e, esx, esy, ehz, r32, fr, s);
        else if (type == 1) launch_refine_split<1, DIM>(m, a, e, esx, esy, ehz, r32, fr, s);
        else launch_refine_split<2, DIM>(m, a, e, esx, esy, ehz, r32, fr, s);
        return;
    }
    switch (type) {
        case 0:
            hipLaunchKernelGGL((saras::k_refine<0, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0, s, a,
                               L0.x, e, esx, esy, L0.b, m->x64b, r32, esx, esy, ehz, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<0>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
        case 1:
            hipLaunchKernelGGL((saras::k_refine<1, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0, s, a,
                               L0.x, e, esx, esy, L0.b, m->x64b, r32, esx, esy, ehz, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<1>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
        default:
            hipLaunchKernelGGL((saras::k_refine<2, DIM>), dim3(saras::kNormBlocks), dim3(saras::kNormThreads), 0, s, a,
                               L0.x, e, esx, esy, L0.b, m->x64b, r32, esx, esy, ehz, m->partial);
            hipLaunchKernelGGL((saras::k_norm_final<2>), dim3(1), dim3(saras::kNormThreads), 0, s, m->partial,
                               saras::kNormBlocks, m->norm_host, m->norm_seq, ++m->norm_ticket);
            break;
    }
}

// x64 <- x64 + e32 (if with_e), r32 <- b - A x64, residual norm of type res_type
static saras_status refine_step(saras_mg *m, hipStream_t s, bool with_e, const saras::FullRows &fr, double *res) {
    auto &L0 = m->h64.lv[0];
    auto &F0 = m->h32->lv[0];
    auto a = level_args(m, m->h64, 0, true);
    const float *e = with_e ? F0.x : nullptr;
    // host-side guards: the caller-layout reads and writes exist on split rows
    // only, always with b from the caller, and never read and write the
    // caller's iterate in one pass
    if ((fr.x || fr.b || fr.xo) && !(L0.hz && F0.hz && fr.b)) {
        set_error("internal error: caller-layout refinement on a level without split rows");
        return SARAS_HIP_ERROR;
    }
    if (fr.x && fr.xo) {
        set_error("internal error: refinement reading and writing the caller's iterate");
        return SARAS_HIP_ERROR;
    }
    if (m->G.dim == 3) launch_refine<3>(m->p.res_type, m, a, e, F0.sx, F0.sy, F0.hz, F0.b, fr, s);
    else launch_refine<2>(m->p.res_type, m, a, e, F0.sx, F0.sy, F0.hz, F0.b, fr, s);
    // x_new is in level 0's other buffer, or (fr.xo) in the caller's array, or
    // (x_old from the caller's array, no correction) still only there
    if (!fr.xo && !(fr.x && !with_e)) std::swap(L0.x, m->x64b);
    HIP_TRY(wait_norm(m, s));
    const double num = m->norm_host[0], den = m->norm_host[1];
    const int pointCount = L0.n[0] * L0.n[1] * L0.n[2];
    const int type = m->p.res_type;
    if (type == 0) *res = den != 0.0 ? num / den : num;
    else if (type == 1) *res = den != 0.0 ? num / den : num / pointCount;
    else *res = den != 0.0 ? std::sqrt(num / pointCount) / std::sqrt(den / pointCount) : std::sqrt(num / pointCount);
    return SARAS_OK;
}

// Mixed precision: FP64 iterate, residual and stopping test; FP32 V-cycles
// (same smoother, transfers and coarse solve as the reference, homogeneous
// walls) on the correction equation A e = r.
// full: the caller's x (initial guess) and b, read in place by the FP64
// passes when both levels 0 have split rows (empty: level 0 holds them).
// With full.xo, the refinement expected to end the solve writes its iterate
// straight into the caller's array; *x_in_caller says where the solution is.
static saras_status solve_mixed(saras_mg *m, hipStream_t s, saras_mg_stats *st,
                                const saras::FullRows &full = saras::FullRows{}, bool *x_in_caller = nullptr) {
    Ops<float> ops{m, *m->h32, s, false, (float)m->p.sor};
    auto &F0 = m->h32->lv[0];
    double res = 0.0;
    saras::FullRows first = full, later = full;
    first.xo = nullptr;
    later.x = nullptr;
    later.xo = nullptr;
    // the FP64 iterate lives in the caller's array: the initial guess until
    // the first refinement reads it from there
    bool in_caller = first.x != nullptr;
    saras_status e = refine_step(m, s, false, first, &res);
    if (e) return e;
    int cycles = 0, refs = 0;
    const int cap = m->p.vc_count > 0 ? m->p.vc_count : (1 << 30);
    // inner_cycles is the most FP32 cycles per refinement.  The first
    // refinement runs at most kProbeCycles of them, few enough that FP32
    // rounding (a correction good to about kFp32Floor of the residual) does
    // not mask the cycle's convergence factor rho, which it measures.  Each
    // later refinement runs the cycles rho needs to take the residual to half
    // the tolerance, or down to the FP32 floor when that comes first (cycles
    // past it gain nothing).  A short estimate costs one more refinement,
    // never accuracy: the FP64 residual alone decides convergence.
    constexpr int kProbeCycles = 3;
    constexpr double kFp32Floor = 1e-6;
    // level 0 is not the coarsest (whose smoothing is the coarse solve) and
    // the V-cycle opens with its pre-smoothing sweep, not the tail kernel
    const bool zero_start = !m->graphs && F0.hz && m->p.pre_smooth > 0 && m->depth > 0 &&
                            (m->tail_off || ops.tail_start() > 0);
    double rho = 1.0;
    bool measured = false;
    while (!(res < m->p.vc_tol) && refs < m->p.max_refine && cycles < cap) {
        int k = m->p.inner_cycles;
        if (!measured) {
            k = std::min(k, kProbeCycles);
        } else if (rho > 0.0 && rho < 0.9 && res > 0.0) {
            const double target = std::max(0.5 * m->p.vc_tol / res, kFp32Floor);
            const double need = std::ceil(std::log(target) / std::log(rho));
            if (need < 1.0) k = 1;
            else if (need < (double)k) k = (int)need;
        }
        // the correction starts from zero: cleared by the first red pass of
        // the round's first cycle where that pass runs on split rows (eager
        // launches; a replayed graph keeps the separate clear)
        if (zero_start) ops.zero0 = true;
        else HIP_TRY(hipMemsetAsync(F0.x, 0, F0.elems * sizeof(float), s));
        int ran = 0;
        for (int c = 0; c < k && cycles < cap; ++c) {
            e = run_vcycle(m, ops);
            if (e) return e;
            ++cycles;
            ++ran;
        }
        const double before = res;
        // where this refinement reads and writes the FP64 iterate: after a
        // refinement that wrote the caller's array, read it from there; else
        // write there when this one is expected to end the solve (the
        // measured rate predicts the tolerance, or a cap ends the loop).  A
        // wrong guess costs nothing: the next refinement reads it in place.
        saras::FullRows fr = later;
        if (in_caller) {
            fr.x = full.x;
        } else if (full.xo) {
            const bool capped = refs + 1 >= m->p.max_refine || cycles >= cap;
            const bool expect = measured && rho > 0.0 && rho < 1.0 && res * std::pow(rho, ran) < m->p.vc_tol;
            if (capped || expect) fr.xo = full.xo;
        }
        e = refine_step(m, s, true, fr, &res);
        if (e) return e;
        in_caller = fr.xo != nullptr;
        ++refs;
        if (!std::isfinite(res)) break;
        rho = (ran > 0 && before > 0.0 && res < before) ? std::pow(res / before, 1.0 / ran) : 1.0;
        measured = true;
    }
    // the last refinement's norm came after the last cycle's coarse solve
    if (m->p.solve_flag && cycles > 0) note_coarse_iters(m);
    st->cycles = cycles;
    st->refinements = refs;
    st->residual = res;
    st->converged = res < m->p.vc_tol;
    if (x_in_caller) *x_in_caller = in_caller;
    return SARAS_OK;
}

namespace saras {

void mg_level0(saras_mg *m, double **x, double **b, long *sx, long *sy, long *hz) {
    auto &L0 = m->h64.lv[0];
    *x = L0.x;
    *b = L0.b;
    *sx = L0.sx;
    *sy = L0.sy;
    if (hz) *hz = L0.hz;
}

hipStream_t mg_stream(saras_mg *m) { return m->stream; }

void set_last_error(const std::string &msg) { set_error(msg); }

std::recursive_mutex &alloc_mutex() {
    static std::recursive_mutex mu;
    return mu;
}

saras_status mg_solve_resident(saras_mg *m, hipStream_t s, saras_mg_stats *st) {
    m->coarse_iters_max = 0;
    m->coarse_capped = 0;
    saras_mg_stats r{};
    saras_status e = m->p.precision == SARAS_PREC_FP64 ? solve_fp64(m, s, &r) : solve_mixed(m, s, &r);
    if (e) return e;
    r.coarse_iters_max = m->coarse_iters_max;
    r.coarse_capped = m->coarse_capped;
    if (st) *st = r;
    return r.converged ? SARAS_OK : SARAS_NOT_CONVERGED;
}

}  // namespace saras

extern "C" {

saras_status saras_mg_solve(saras_mg *mg, double *x_full, const double *b_full, int ptr_kind, void *stream,
                            saras_mg_stats *stats) {
    if (!mg || !x_full || !b_full) {
        set_error("null argument");
        return SARAS_BAD_ARG;
    }
    if (ptr_kind != SARAS_PTR_HOST && ptr_kind != SARAS_PTR_DEVICE) return SARAS_BAD_ARG;
    saras_mg *m = mg;
    hipStream_t s = pick_stream(m, stream);
    const long fe = full_elems(m->g);
    double *dx = x_full;
    const double *db = b_full;
    if (ptr_kind == SARAS_PTR_HOST) {
        if (m->stage_elems < (size_t)fe) {
            std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
            (void)hipFree(m->stage_x);
            (void)hipFree(m->stage_b);
            m->stage_x = m->stage_b = nullptr;
            HIP_TRY(hipMalloc((void **)&m->stage_x, fe * sizeof(double)));
            HIP_TRY(hipMalloc((void **)&m->stage_b, fe * sizeof(double)));
            m->stage_elems = fe;
        }
        HIP_TRY(hipMemcpyAsync(m->stage_x, x_full, fe * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->stage_b, b_full, fe * sizeof(double), hipMemcpyHostToDevice, s));
        dx = m->stage_x;
        db = m->stage_b;
    }
    m->coarse_iters_max = 0;
    m->coarse_capped = 0;
    HIP_TRY(hipEventRecord(m->ev0, s));
    auto &L0 = m->h64.lv[0];
    saras_mg_stats st{};
    saras_status e = SARAS_OK;
    bool in_caller = false;   // the solution's core is in dx already
    if (m->p.precision == SARAS_PREC_FP64) {
        load_core<double>(m, db, L0, L0.b, s);
        load_core<double>(m, dx, L0, L0.x, s);
        e = solve_fp64(m, s, &st);
    } else if (L0.hz && m->h32->lv[0].hz) {
        // the FP64 passes of the mixed solve read b and the initial guess
        // from the caller's arrays; level 0 receives x from the first pass
        const long p = m->g.pad, fz = L0.n[2] + 2 * p, fy = m->g.dim == 3 ? L0.n[1] + 2 * p : 1;
        saras::FullRows fr;
        fr.x = dx;
        fr.b = db;
        fr.xo = dx;
        fr.si = m->g.dim == 3 ? fy * fz : fz;
        fr.sj = m->g.dim == 3 ? fz : 0;
        fr.off = m->g.dim == 3 ? p * fy * fz + p * fz + p : p * fz + p;
        e = solve_mixed(m, s, &st, fr, &in_caller);
    } else {
        load_core<double>(m, db, L0, L0.b, s);
        load_core<double>(m, dx, L0, L0.x, s);
        e = solve_mixed(m, s, &st);
    }
    if (e) return e;
    store_full(m, dx, in_caller ? nullptr : L0.x, s);
    HIP_TRY(hipEventRecord(m->ev1, s));
    if (ptr_kind == SARAS_PTR_HOST)
        HIP_TRY(hipMemcpyAsync(x_full, m->stage_x, fe * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    st.device_ms = ms;
    st.coarse_iters_max = m->coarse_iters_max;
    st.coarse_capped = m->coarse_capped;
    if (stats) *stats = st;
    return st.converged ? SARAS_OK : SARAS_NOT_CONVERGED;
}

saras_status saras_mg_set_graphs(saras_mg *mg, int enable) {
    if (!mg) return SARAS_BAD_ARG;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    if (!enable)
        for (auto &c : mg->graph_cache) {
            for (auto &e : c) (void)hipGraphExecDestroy(e.exec);
            c.clear();
        }
    mg->graphs = enable != 0;
    return SARAS_OK;
}

size_t saras_mg_footprint(const saras_grid_desc *grid, const saras_mg_params *params) {
    if (!grid || !params) return 0;
    saras::GridLevels G;
    if (!saras::build_levels(*grid, G)) return 0;
    const int D = G.depth;
    auto level_bytes = [&](size_t word) {
        size_t tot = 0;
        for (int l = 0; l <= D; ++l) {
            int n[3];
            for (int a = 0; a < 3; ++a) n[a] = G.axes[l][a].n;
            if (grid->dim == 2) n[1] = 1;
            const long a128 = 128 / (long)word;
            const size_t sy = (size_t)((n[2] + a128 - 1) / a128 * a128);
            const size_t elems = sy * n[1] * n[0];
            tot += elems * word * ((l < D && grid->dim == 3) ? 3 : 2);
            tot += 5 * 3 * (size_t)(n[0] + n[1] + n[2]) * word;
        }
        return tot;
    };
    size_t bytes = level_bytes(8);
    const size_t full = (size_t)full_elems(*grid);
    if (params->precision == SARAS_PREC_MIXED) {
        bytes += level_bytes(4);
        const auto &A = G.axes[0];
        const size_t sy = (size_t)((A[2].n + 15) / 16 * 16);
        bytes += sy * (grid->dim == 2 ? 1 : A[1].n) * A[0].n * 8;   // second FP64 iterate
    }
    bytes += 2 * full * 8;                      // host-pointer staging
    bytes += 2 * 1024 * 1024;                   // norms, counters, coarse workspace
    return bytes;
}

// ---- single-operation entry points ----
saras_status saras_mg_set_level(saras_mg *mg, int level, const double *lhs_pad, const double *rhs_pad) {
    if (!mg || level < 0 || level > mg->depth) return SARAS_BAD_ARG;
    auto &L = mg->h64.lv[level];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    std::vector<double> core(L.elems, 0.0);
    const bool d3 = mg->g.dim == 3;
    for (int which = 0; which < 2; ++which) {
        const double *src = which == 0 ? lhs_pad : rhs_pad;
        if (!src) continue;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const long f = d3 ? ((long)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + (k + 1)
                                      : (long)(i + 1) * (nz + 2) + (k + 1);
                    core[(size_t)i * L.sx + (size_t)j * L.sy + (L.hz ? (k & 1) * L.hz + (k >> 1) : k)] = src[f];
                }
        HIP_TRY(hipMemcpy(which == 0 ? L.x : L.b, core.data(), L.elems * sizeof(double), hipMemcpyHostToDevice));
    }
    return SARAS_OK;
}

saras_status saras_mg_get_level(saras_mg *mg, int level, double *lhs_pad, double *rhs_pad) {
    if (!mg || level < 0 || level > mg->depth) return SARAS_BAD_ARG;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    auto &L = mg->h64.lv[level];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    std::vector<double> core(L.elems);
    const bool d3 = mg->g.dim == 3;
    const size_t padded = d3 ? (size_t)(nx + 2) * (ny + 2) * (nz + 2) : (size_t)(nx + 2) * (nz + 2);
    for (int which = 0; which < 2; ++which) {
        double *dst = which == 0 ? lhs_pad : rhs_pad;
        if (!dst) continue;
        HIP_TRY(hipMemcpy(core.data(), which == 0 ? L.x : L.b, L.elems * sizeof(double), hipMemcpyDeviceToHost));
        std::memset(dst, 0, padded * sizeof(double));
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const long f = d3 ? ((long)(i + 1) * (ny + 2) + (j + 1)) * (nz + 2) + (k + 1)
                                      : (long)(i + 1) * (nz + 2) + (k + 1);
                    dst[f] = core[(size_t)i * L.sx + (size_t)j * L.sy + (L.hz ? (k & 1) * L.hz + (k >> 1) : k)];
                }
    }
    return SARAS_OK;
}

saras_status saras_mg_smooth(saras_mg *mg, int level, int count, int zero_bc) {
    if (!mg || level < 0 || level > mg->depth || count < 0) return SARAS_BAD_ARG;
    Ops<double> ops{mg, mg->h64, mg->stream, true, mg->p.sor};
    ops.smooth(level, count, zero_bc != 0);
    if (ops.err) return ops.err;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    HIP_TRY(hipGetLastError());
    return SARAS_OK;
}

saras_status saras_mg_coarse_solve(saras_mg *mg, int level, int zero_bc, int *iters) {
    if (!mg || level < 0 || level > mg->depth) return SARAS_BAD_ARG;
    if (level != mg->depth) {
        set_error("the single-workgroup solve runs on the coarsest level only");
        return SARAS_BAD_ARG;
    }
    Ops<double> ops{mg, mg->h64, mg->stream, true, mg->p.sor};
    ops.coarse(level, zero_bc != 0, 0, 0);
    if (ops.err) return ops.err;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    if (iters) *iters = *mg->iters_host;
    return SARAS_OK;
}

saras_status saras_mg_residual_restrict(saras_mg *mg, int level) {
    if (!mg || level < 0 || level >= mg->depth) return SARAS_BAD_ARG;
    Ops<double> ops{mg, mg->h64, mg->stream, true, mg->p.sor};
    ops.resid_restrict(level);
    if (ops.err) return ops.err;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    return SARAS_OK;
}

saras_status saras_mg_prolong_correct(saras_mg *mg, int level) {
    if (!mg || level < 0 || level >= mg->depth) return SARAS_BAD_ARG;
    Ops<double> ops{mg, mg->h64, mg->stream, true, mg->p.sor};
    ops.prolong_correct(level);
    if (ops.err) return ops.err;
    HIP_TRY(hipStreamSynchronize(mg->stream));
    return SARAS_OK;
}

saras_status saras_mg_error(saras_mg *mg, int res_type, double *value) {
    if (!mg || !value || res_type < 0 || res_type > 2) return SARAS_BAD_ARG;
    return compute_error(mg, res_type, mg->h64.lv[0].x, mg->stream, value);
}

// ---- measurement hooks ----
saras_status saras_mg_bench_cycles(saras_mg *mg, int count, double *ms) {
    if (!mg || count < 0 || !ms) return SARAS_BAD_ARG;
    hipStream_t s = mg->stream;
    HIP_TRY(hipEventRecord(mg->ev0, s));
    if (mg->p.precision == SARAS_PREC_FP64) {
        Ops<double> ops{mg, mg->h64, s, true, mg->p.sor};
        for (int c = 0; c < count; ++c) {
            saras_status e = run_vcycle(mg, ops);
            if (e) return e;
        }
    } else {
        Ops<float> ops{mg, *mg->h32, s, false, (float)mg->p.sor};
        for (int c = 0; c < count; ++c) {
            saras_status e = run_vcycle(mg, ops);
            if (e) return e;
        }
    }
    HIP_TRY(hipEventRecord(mg->ev1, s));
    HIP_TRY(hipEventSynchronize(mg->ev1));
    HIP_TRY(hipGetLastError());
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, mg->ev0, mg->ev1));
    *ms = f;
    return SARAS_OK;
}

saras_status saras_mg_bench_sweeps(saras_mg *mg, int reps, double *ms) {
    if (!mg || reps < 0 || !ms) return SARAS_BAD_ARG;
    if (mg->depth == 0) {
        set_error("level 0 is the coarsest level");
        return SARAS_BAD_ARG;
    }
    hipStream_t s = mg->stream;
    HIP_TRY(hipEventRecord(mg->ev0, s));
    if (mg->p.precision == SARAS_PREC_FP64) {
        Ops<double> ops{mg, mg->h64, s, true, mg->p.sor};
        for (int c = 0; c < reps; ++c) ops.rb_sweep(0, false);
    } else {
        Ops<float> ops{mg, *mg->h32, s, false, (float)mg->p.sor};
        for (int c = 0; c < reps; ++c) ops.rb_sweep(0, true);
    }
    HIP_TRY(hipEventRecord(mg->ev1, s));
    HIP_TRY(hipEventSynchronize(mg->ev1));
    HIP_TRY(hipGetLastError());
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, mg->ev0, mg->ev1));
    *ms = f;
    return SARAS_OK;
}

saras_status saras_bench_stream_copy(size_t n, int reps, double *ms) {
    if (!ms || n < 2 || reps < 1) return SARAS_BAD_ARG;
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const long n2 = (long)(n / 2);
    HIP_TRY(hipMalloc((void **)&a, n2 * 2 * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&b, n2 * 2 * sizeof(double)));
    HIP_TRY(hipMemset(a, 0, n2 * 2 * sizeof(double)));
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const unsigned blocks = (unsigned)((n2 + 255) / 256);
    hipLaunchKernelGGL(saras::k_stream_copy, dim3(blocks), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n2);
    HIP_TRY(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(saras::k_stream_copy, dim3(blocks), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n2);
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    return SARAS_OK;
}

// ---- device memory helpers ----
saras_status saras_device_alloc(void **ptr, size_t bytes) {
    if (!ptr) return SARAS_BAD_ARG;
    HIP_TRY(hipMalloc(ptr, bytes));
    return SARAS_OK;
}
saras_status saras_device_free(void *ptr) {
    HIP_TRY(hipFree(ptr));
    return SARAS_OK;
}
saras_status saras_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SARAS_OK;
}
saras_status saras_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SARAS_OK;
}
saras_status saras_device_sync(void) {
    HIP_TRY(hipDeviceSynchronize());
    return SARAS_OK;
}
saras_status saras_device_zero(void *ptr, size_t bytes) {
    if (!ptr) return SARAS_BAD_ARG;
    HIP_TRY(hipMemset(ptr, 0, bytes));
    // hipMemset on device memory may return before the fill has run; the
    // solver handles' streams do not wait for the null stream
    HIP_TRY(hipDeviceSynchronize());
    return SARAS_OK;
}

}  // extern "C"
