// nse_solver.hip -- device-resident RK3 / Crank-Nicolson time stepper of the
// incompressible Navier-Stokes (+ Boussinesq scalar) equations and the
// field-operator entry points of include/saras_hip.h.
//
// One sub-stage (lsRK3_d3.cc:199-304) is
//   k_rhs        explicit terms of every field in one pass (diffusion, forcing,
//                pressure gradient, Morinishi non-linear term, RK3 weights)
//   k_jacobi     Jacobi CN iterations per field; each launch updates the field
//                and returns the residual of its input iterate, so the
//                reference's update + residual pair costs one pass
//   k_div        divergence into the pressure solver's level-0 RHS
//   multigrid    the FP64 (or mixed) V-cycle solver on resident arrays
//   k_correct    P += Pp, V -= (alpha+beta) dt grad Pp
// with every boundary condition applied on the fly by the stencils.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "grid_host.hpp"
#include "mg_internal.hpp"
#include "nse_kernels.hpp"
#include "saras_hip.h"

using saras::nse::CoreField;
using saras::nse::FieldBC;
using saras::nse::Geom;
using saras::nse::PadField;

namespace {

#define NSE_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            saras::set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            return e_ == hipErrorOutOfMemory ? SARAS_OOM : SARAS_HIP_ERROR;             \
        }                                                                               \
    } while (0)

// lsRK3_d3.cc:63-66 (beta = alpha)
const double kAlpha[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
const double kGamma[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
const double kZeta[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};

long pitch16(int nz) { return (nz + 15) / 16 * 16; }

}  // namespace

struct saras_nse {
    saras_grid_desc g{};
    saras_nse_params p{};
    saras_mg *mg = nullptr;
    Geom G{};
    double *metrics = nullptr;          // 9 core metric arrays
    FieldBC bc[4]{};                     // Vx, Vy, Vz, T
    FieldBC pbc{};                       // P
    double *pghost[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // nullBC walls of P
    // second ghost layer above the z-hi wall of Vx, Vy, Vz, T (FieldBC::top2)
    double *top2[4] = {nullptr, nullptr, nullptr, nullptr};
    FieldBC ppbc{};                      // Pp: the Poisson solver's walls (Neumann; a planar plume's top: Dirichlet)
    double *F[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // Vx, Vy, Vz, T, P
    double *H[4] = {nullptr, nullptr, nullptr, nullptr};
    double *R[4] = {nullptr, nullptr, nullptr, nullptr};
    double *scratch = nullptr;
    // per-column flags of the hot-plate (bottom) and plume-outflow (top) walls of T
    unsigned char *wmask = nullptr;
    bool plume = false;
    size_t elems = 0;
    double *part = nullptr;             // per-block residual maxima of a Jacobi launch
    long nblocks = 0;
    double *res_host = nullptr;         // pinned, host-coherent [1]: written by k_max_final
    unsigned *res_seq = nullptr;        // its ticket, stored after it (saras::wait_ticket)
    unsigned res_ticket = 0;
    int max_iter = 0;
    double *diag_part = nullptr, *diag_host = nullptr;
    double *stage_a = nullptr, *stage_b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int jac[4] = {0, 0, 0, 0};
    saras_mg_stats mg_stats{};
    double wall_dn = 0.0;
    long full = 0;
    // planar (2-D, x-z) run: the reference's PLANAR build, y index fixed at 0.
    // Stored as one y row with y wrapping onto itself and every y metric and
    // y difference factor zero, so each y term of a stencil is an exact zero.
    bool planar = false;
    int ypad() const { return planar ? 0 : g.pad; }
    // tseries totalVol and the wall cross-section (tseries.cc:76-80, 96-104)
    double volume() const { return planar ? g.len[0] * g.len[2] : g.len[0] * g.len[1] * g.len[2]; }
    double wall_area() const { return planar ? g.len[0] : g.len[0] * g.len[1]; }
    int threads = 256;
    // LES (spiral.cc): sub-grid stresses Txx, Tyy, Tzz, Txy, Tyz, Tzx
    double *Tsg[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *sgs_part = nullptr;          // per-block [sum |K| dV, dissipation dV, nu_t dV]
    double *sgs_dev = nullptr;           // device: 3 sums, then the failure flag (as int)
    double *sgs_host = nullptr;          // pinned copy of sgs_dev
    double sgs_diag[3] = {0.0, 0.0, 0.0};
    double time = 0.0, tstp = 0.0;       // run time and the input time step (lsRK3_d3.cc:238)

    ~saras_nse() {
        for (auto *q : F) (void)hipFree(q);
        for (auto *q : H) (void)hipFree(q);
        for (auto *q : R) (void)hipFree(q);
        (void)hipFree(scratch);
        (void)hipFree(wmask);
        (void)hipFree(metrics);
        (void)hipFree(part);
        (void)hipHostFree(res_host);
        (void)hipHostFree(res_seq);
        (void)hipFree(diag_part);
        (void)hipHostFree(diag_host);
        (void)hipFree(stage_a);
        (void)hipFree(stage_b);
        for (auto *q : Tsg) (void)hipFree(q);
        for (auto *q : pghost) (void)hipFree(q);
        for (auto *q : top2) (void)hipFree(q);
        (void)hipFree(sgs_part);
        (void)hipFree(sgs_dev);
        (void)hipHostFree(sgs_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (mg) saras_mg_destroy(mg);
    }

    int stages() const { return p.scheme == 1 ? 1 : 3; }
    // cells of wall w's ghost plane: the core extents of the two other axes
    size_t plane_cells(int w) const {
        const int ax = w / 2;
        return ax == 0 ? (size_t)G.n[1] * G.n[2] : (ax == 1 ? (size_t)G.n[0] * G.n[2] : (size_t)G.n[0] * G.n[1]);
    }
    dim3 grid3() const { return dim3((G.n[2] + threads - 1) / threads, G.n[1], G.n[0]); }
    dim3 grid3_full(int p) const {
        return dim3((G.n[2] + 2 * p + threads - 1) / threads, G.n[1] + 2 * ypad(), G.n[0] + 2 * p);
    }
    CoreField field(int q) const { return CoreField{F[q], q == 4 ? pbc : bc[q]}; }
    // stress m as a field of the periodic box
    saras::nse::PerField tsg(int m) const { return saras::nse::PerField{Tsg[m]}; }
    bool les_active() const { return p.les && time > 5 * tstp; }
    // the three stresses whose x, y, z derivatives enter component c
    // (spiral.cc:225-235): x: Txx Txy Tzx, y: Txy Tyy Tyz, z: Tzx Tyz Tzz
    saras::nse::PerField sg(int c, int ax) const {
        static const int m[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
        return tsg(m[c][ax]);
    }
    // the pressure solver's level-0 iterate (Pp), in its row layout
    saras::nse::SplitField pp() const {
        double *x, *b;
        long sx, sy, hz;
        saras::mg_level0(mg, &x, &b, &sx, &sy, &hz);
        return saras::nse::SplitField{x, ppbc, hz};
    }
};

namespace {

saras_status plume_flags(saras_nse *s);

// the stresses and per-block diagnostics of every core cell (les_sf32: the
// structure-function integral in single precision)
void launch_sgs(saras_nse *s) {
    if (s->p.les_sf32)
        hipLaunchKernelGGL(saras::nse::k_sgs_stress<true>, s->grid3(), dim3(s->threads), 0, s->s, s->G, s->field(0),
                           s->field(1), s->field(2), s->p.nu, s->Tsg[0], s->Tsg[1], s->Tsg[2], s->Tsg[3], s->Tsg[4],
                           s->Tsg[5], s->sgs_part, reinterpret_cast<int *>(s->sgs_dev + 3));
    else
        hipLaunchKernelGGL(saras::nse::k_sgs_stress<false>, s->grid3(), dim3(s->threads), 0, s->s, s->G, s->field(0),
                           s->field(1), s->field(2), s->p.nu, s->Tsg[0], s->Tsg[1], s->Tsg[2], s->Tsg[3], s->Tsg[4],
                           s->Tsg[5], s->sgs_part, reinterpret_cast<int *>(s->sgs_dev + 3));
}

// spiral::computeSG stresses and diagnostics (spiral.cc:129-206); the
// divergence of the stresses is taken inside the RHS kernels
saras_status compute_sg(saras_nse *s) {
    NSE_TRY(hipMemsetAsync(s->sgs_dev, 0, 4 * sizeof(double), s->s));
    launch_sgs(s);
    hipLaunchKernelGGL(saras::nse::k_sum3_final, dim3(1), dim3(1024), 0, s->s, s->sgs_part, s->nblocks, s->sgs_dev);
    NSE_TRY(hipGetLastError());
    NSE_TRY(hipMemcpyAsync(s->sgs_host, s->sgs_dev, 4 * sizeof(double), hipMemcpyDeviceToHost, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    if (*reinterpret_cast<const int *>(s->sgs_host + 3)) {
        saras::set_last_error("spiral SGS: eigenvalue or eigenvector of the strain tensor failed (spiral.cc:643-737)");
        return SARAS_SGS_FAILED;
    }
    for (int q = 0; q < 3; ++q) s->sgs_diag[q] = s->sgs_host[q];
    return SARAS_OK;
}

// eulerCN_d3 explicit part (eulerCN_d3.cc:92-121, 181-235)
saras_status rhs_euler(saras_nse *s) {
    const dim3 gr = s->grid3(), bl(s->threads);
    const double k_nu = s->p.nu * 0.5, dt = s->p.dt;   // alphCN2 = 1/2
    const int les = s->les_active();
    if (les) {
        saras_status e = compute_sg(s);
        if (e) return e;
    }
    hipLaunchKernelGGL(saras::nse::k_euler_v<0>, gr, bl, 0, s->s, s->G, k_nu, s->p.nu, dt, s->p.buoy[0], s->p.has_T,
                       s->p.mean_pgrad, s->p.nlscheme, s->field(0), s->field(1), s->field(2), s->field(3), s->field(4),
                       s->R[0], les, s->sg(0, 0), s->sg(0, 1), s->sg(0, 2));
    hipLaunchKernelGGL(saras::nse::k_euler_v<1>, gr, bl, 0, s->s, s->G, k_nu, s->p.nu, dt, s->p.buoy[1], s->p.has_T,
                       0.0, s->p.nlscheme, s->field(0), s->field(1), s->field(2), s->field(3), s->field(4),
                       s->R[1], les, s->sg(1, 0), s->sg(1, 1), s->sg(1, 2));
    hipLaunchKernelGGL(saras::nse::k_euler_v<2>, gr, bl, 0, s->s, s->G, k_nu, s->p.nu, dt, s->p.buoy[2], s->p.has_T,
                       0.0, s->p.nlscheme, s->field(0), s->field(1), s->field(2), s->field(3), s->field(4),
                       s->R[2], les, s->sg(2, 0), s->sg(2, 1), s->sg(2, 2));
    if (s->p.has_T)
        hipLaunchKernelGGL(saras::nse::k_euler_t, gr, bl, 0, s->s, s->G, s->p.kappa * 0.5, s->p.kappa, dt, s->p.nlscheme,
                           s->field(0), s->field(1), s->field(2), s->field(3), s->R[3]);
    NSE_TRY(hipGetLastError());
    return SARAS_OK;
}

saras_status rhs(saras_nse *s, int stage) {
    if (s->p.scheme == 1) return rhs_euler(s);
    saras::nse::StageCoef c{};
    const double a = kAlpha[stage], b = kAlpha[stage];
    c.zeta = kZeta[stage];
    c.a_nu = a * s->p.nu;
    c.a_kappa = a * s->p.kappa;
    c.ab = a + b;
    c.gamma = kGamma[stage];
    c.dt = s->p.dt;
    for (int q = 0; q < 3; ++q) c.buoy[q] = s->p.buoy[q];
    c.mpg = s->p.mean_pgrad;
    c.stage = stage;
    c.has_T = s->p.has_T;
    c.nlscheme = s->p.nlscheme;
    c.nu = s->p.nu;
    c.kappa = s->p.kappa;
    const int les = s->les_active();
    if (les) {
        saras_status e = compute_sg(s);
        if (e) return e;
    }
    const dim3 gr = s->grid3(), bl(s->threads);
    hipLaunchKernelGGL(saras::nse::k_rhs_v<0>, gr, bl, 0, s->s, s->G, c, s->field(0), s->field(1), s->field(2),
                       s->field(3), s->field(4), s->H[0], s->R[0], les, s->sg(0, 0), s->sg(0, 1),
                       s->sg(0, 2));
    hipLaunchKernelGGL(saras::nse::k_rhs_v<1>, gr, bl, 0, s->s, s->G, c, s->field(0), s->field(1), s->field(2),
                       s->field(3), s->field(4), s->H[1], s->R[1], les, s->sg(1, 0), s->sg(1, 1),
                       s->sg(1, 2));
    hipLaunchKernelGGL(saras::nse::k_rhs_v<2>, gr, bl, 0, s->s, s->G, c, s->field(0), s->field(1), s->field(2),
                       s->field(3), s->field(4), s->H[2], s->R[2], les, s->sg(2, 0), s->sg(2, 1),
                       s->sg(2, 2));
    if (s->p.has_T)
        hipLaunchKernelGGL(saras::nse::k_rhs_t, gr, bl, 0, s->s, s->G, c, s->field(0), s->field(1), s->field(2),
                           s->field(3), s->H[3], s->R[3]);
    NSE_TRY(hipGetLastError());
    return SARAS_OK;
}

// timestep::solveVx/Vy/Vz/solveT: Jacobi iterations until the residual of the
// updated field is below cn_tol, at most maxIterations extra iterations.
saras_status solve_field(saras_nse *s, int f, double beta, int *iters) {
    const double diff = f == 3 ? s->p.kappa : s->p.nu;
    const double dt = s->p.dt;
    double *cur = s->scratch, *other = s->F[f];
    hipLaunchKernelGGL(saras::nse::k_jacobi<false>, s->grid3(), dim3(s->threads), 0, s->s, s->G,
                       CoreField{s->F[f], s->bc[f]}, s->R[f], cur, dt, diff, beta, nullptr);
    NSE_TRY(hipGetLastError());
    // solveT imposes T's walls after each update (predictors.cc:376): from the
    // first one on, the plume flags see the predicted Vz
    if (f == 3) {
        saras_status e = plume_flags(s);
        if (e) return e;
    }
    int it = 0;
    for (;;) {
        hipLaunchKernelGGL(saras::nse::k_jacobi<true>, s->grid3(), dim3(s->threads), 0, s->s, s->G,
                           CoreField{cur, s->bc[f]}, s->R[f], other, dt, diff, beta, s->part);
        hipLaunchKernelGGL(saras::nse::k_max_final, dim3(1), dim3(1024), 0, s->s, s->part, s->nblocks, s->res_host,
                           s->res_seq, ++s->res_ticket);
        NSE_TRY(hipGetLastError());
        NSE_TRY(saras::wait_ticket(s->res_seq, s->res_ticket, s->s));
        const double r = s->res_host[0];
        if (r < s->p.cn_tol) break;
        it += 1;
        if (it > s->max_iter) {
            *iters = -1;
            saras::set_last_error("Jacobi iterations not converging (predictors.cc:133-141)");
            // leave the field at its latest iterate, as the reference's V holds it
            if (cur != s->F[f]) {
                s->scratch = s->F[f];
                s->F[f] = cur;
            }
            return SARAS_JACOBI_DIVERGED;
        }
        std::swap(cur, other);
    }
    if (cur != s->F[f]) {
        s->scratch = s->F[f];
        s->F[f] = cur;
    }
    *iters = it;
    return SARAS_OK;
}

saras_status stage_run(saras_nse *s, int stage) {
    const bool euler = s->p.scheme == 1;
    const double a = euler ? 0.0 : kAlpha[stage], b = euler ? 0.5 : kAlpha[stage];   // betaCN2 = 1/2
    saras_status e = rhs(s, stage);
    if (e) return e;
    const int nf = s->p.has_T ? 4 : 3;
    for (int f = 0; f < nf; ++f) {
        e = solve_field(s, f, b, &s->jac[f]);
        if (e) return e;
    }
    double *x0, *b0;
    long sx0, sy0, hz0;
    saras::mg_level0(s->mg, &x0, &b0, &sx0, &sy0, &hz0);
    const double sc = euler ? 1.0 / s->p.dt : 1.0 / ((a + b) * s->p.dt);
    hipLaunchKernelGGL(saras::nse::k_div, s->grid3(), dim3(s->threads), 0, s->s, s->G, s->field(0), s->field(1),
                       s->field(2), sc, b0, sx0, sy0, hz0);
    NSE_TRY(hipGetLastError());
    e = saras::mg_solve_resident(s->mg, s->s, &s->mg_stats);
    if (e && e != SARAS_NOT_CONVERGED) return e;
    const double pc = euler ? s->p.dt : (a + b) * s->p.dt;
    const saras::nse::SplitField pp = s->pp();
    if (pp.hz)
        hipLaunchKernelGGL(saras::nse::k_correct<saras::nse::SplitField>, s->grid3(), dim3(s->threads), 0, s->s, s->G,
                           pp, s->F[4], s->pbc, s->F[0], s->F[1], s->F[2], pc);
    else
        hipLaunchKernelGGL(saras::nse::k_correct<CoreField>, s->grid3(), dim3(s->threads), 0, s->s, s->G,
                           CoreField{pp.f, pp.bc}, s->F[4], s->pbc, s->F[0], s->F[1], s->F[2], pc);
    NSE_TRY(hipGetLastError());
    // V.imposeBCs(); P.imposeBCs(); T.imposeBCs() closing the stage
    return plume_flags(s);
}

void fill_stats(saras_nse *s, saras_nse_stats *st, float ms) {
    if (!st) return;
    for (int q = 0; q < 4; ++q) st->jacobi_iters[q] = s->jac[q];
    st->mg = s->mg_stats;
    st->device_ms = ms;
}

saras_status ensure_stage(saras_nse *s) {
    if (s->stage_a) return SARAS_OK;
    std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
    NSE_TRY(hipMalloc((void **)&s->stage_a, s->full * sizeof(double)));
    NSE_TRY(hipMalloc((void **)&s->stage_b, s->full * sizeof(double)));
    return SARAS_OK;
}

FieldBC make_bc(const int *type, const double *val, const double *hwall) {
    FieldBC b{};
    for (int w = 0; w < 6; ++w) {
        b.type[w] = type[w];   // SARAS_FBC_* values are the kernels' kinds
        b.val[w] = val[w];
        b.kh[w] = val[w] * hwall[w];
    }
    return b;
}

// plumeOut's flags follow every T.imposeBCs of the reference: they are
// taken from the ghost Vz at that moment (plumeOut.cc:71-90)
saras_status plume_flags(saras_nse *s) {
    if (!s->plume) return SARAS_OK;
    const long cols = (long)s->G.n[0] * s->G.n[1];
    hipLaunchKernelGGL(saras::nse::k_plume_flags, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s->s, s->G,
                       s->field(2), s->wmask + cols);
    NSE_TRY(hipGetLastError());
    return SARAS_OK;
}

// caller array (host or device, pad p) -> core buffer dst
saras_status load_field(saras_nse *s, const double *src, int ptr_kind, double *dst, long hz = 0) {
    const double *d = src;
    if (ptr_kind == SARAS_PTR_HOST) {
        saras_status e = ensure_stage(s);
        if (e) return e;
        NSE_TRY(hipMemcpyAsync(s->stage_a, src, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        d = s->stage_a;
    }
    const int p = s->g.pad, py = s->ypad();
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->G.n[1] + 2 * py);
    hipLaunchKernelGGL(saras::nse::k_field_load, s->grid3(), dim3(s->threads), 0, s->s, d, SX, SY, p, py, dst,
                       s->G.sx, s->G.sy, hz, s->G.n[0], s->G.n[1], s->G.n[2]);
    NSE_TRY(hipGetLastError());
    if (ptr_kind == SARAS_PTR_HOST) NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

// the first ghost layer of wall w of a caller array (pad p) -> its plane
__global__ void k_ghost_plane_load(const double *__restrict__ src, long SX, long SY, int p, int py, int w, int nx,
                                   int ny, int nz, double *__restrict__ dst) {
    const int ax = w / 2, hi = w & 1;
    const int na = ax == 0 ? ny : nx, nb = ax == 2 ? ny : nz;   // plane extents, fastest last
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)na * nb) return;
    const int a = (int)(t / nb), b = (int)(t - (long)a * nb);
    int c[3];
    if (ax == 0) {
        c[0] = hi ? nx : -1;
        c[1] = a;
        c[2] = b;
    } else if (ax == 1) {
        c[0] = a;
        c[1] = hi ? ny : -1;
        c[2] = b;
    } else {
        c[0] = a;
        c[1] = b;
        c[2] = hi ? nz : -1;
    }
    dst[t] = src[(long)(c[0] + p) * SX + (long)(c[1] + py) * SY + (c[2] + p)];
}

// P's nullBC walls take the caller's first ghost layer
saras_status load_ghost_planes(saras_nse *s, const double *src, int ptr_kind) {
    bool any = false;
    for (int w = 0; w < 6; ++w) any = any || s->pghost[w];
    if (!any) return SARAS_OK;
    const double *d = src;
    if (ptr_kind == SARAS_PTR_HOST) {
        saras_status e = ensure_stage(s);
        if (e) return e;
        NSE_TRY(hipMemcpyAsync(s->stage_a, src, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        d = s->stage_a;
    }
    const int p = s->g.pad, py = s->ypad();
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->G.n[1] + 2 * py);
    for (int w = 0; w < 6; ++w) {
        if (!s->pghost[w]) continue;
        const long cells = (long)s->plane_cells(w);
        hipLaunchKernelGGL(k_ghost_plane_load, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->s, d, SX, SY,
                           p, py, w, s->G.n[0], s->G.n[1], s->G.n[2], s->pghost[w]);
    }
    NSE_TRY(hipGetLastError());
    if (ptr_kind == SARAS_PTR_HOST) NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

// the second ghost layer above the z-hi wall of a caller array (pad p >= 2)
// -> plane (i ny + j)
__global__ void k_top2_load(const double *__restrict__ src, long SX, long SY, int p, int py, int nx, int ny, int nz,
                            double *__restrict__ dst) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nx * ny) return;
    const int i = (int)(t / ny), j = (int)(t - (long)i * ny);
    dst[t] = src[(long)(i + p) * SX + (long)(j + py) * SY + (nz + 1 + p)];
}

saras_status load_top2(saras_nse *s, const double *src, int ptr_kind, double *dst) {
    const double *d = src;
    if (ptr_kind == SARAS_PTR_HOST) {
        saras_status e = ensure_stage(s);
        if (e) return e;
        NSE_TRY(hipMemcpyAsync(s->stage_a, src, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        d = s->stage_a;
    }
    const int p = s->g.pad, py = s->ypad();
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->G.n[1] + 2 * py);
    const long cells = (long)s->G.n[0] * s->G.n[1];
    hipLaunchKernelGGL(k_top2_load, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s->s, d, SX, SY, p, py,
                       s->G.n[0], s->G.n[1], s->G.n[2], dst);
    NSE_TRY(hipGetLastError());
    if (ptr_kind == SARAS_PTR_HOST) NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

__global__ void k_store_core(const double *__restrict__ src, long sx, long sy, double *__restrict__ dst, long SX,
                             long SY, int p, int py, int nz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if (k >= nz) return;
    dst[(long)(i + p) * SX + (long)(j + py) * SY + (k + p)] = src[(long)i * sx + (long)j * sy + k];
}

// core buffer (+ face ghosts unless core_only) -> caller array
template <class FF>
saras_status store_field(saras_nse *s, FF f, double *dst, int ptr_kind, bool core_only) {
    double *d = dst;
    if (ptr_kind == SARAS_PTR_HOST) {
        saras_status e = ensure_stage(s);
        if (e) return e;
        NSE_TRY(hipMemcpyAsync(s->stage_a, dst, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        d = s->stage_a;
    }
    const int p = s->g.pad, py = s->ypad();
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->G.n[1] + 2 * py);
    if (core_only)
        hipLaunchKernelGGL(k_store_core, s->grid3(), dim3(s->threads), 0, s->s, f.f, s->G.sx, s->G.sy, d, SX, SY, p,
                           py, s->G.n[2]);
    else
        hipLaunchKernelGGL(saras::nse::k_field_store<FF>, s->grid3_full(p), dim3(s->threads), 0, s->s, s->G, f, d, SX,
                           SY, p, py);
    NSE_TRY(hipGetLastError());
    if (ptr_kind == SARAS_PTR_HOST) {
        NSE_TRY(hipMemcpyAsync(dst, s->stage_a, s->full * sizeof(double), hipMemcpyDeviceToHost, s->s));
        NSE_TRY(hipStreamSynchronize(s->s));
    }
    return SARAS_OK;
}

}  // namespace

extern "C" {

saras_status saras_nse_create(const saras_grid_desc *grid, const saras_mg_params *mgp,
                              const saras_nse_params *params, saras_nse **out) {
    if (!grid || !mgp || !params || !out) return SARAS_BAD_ARG;
    *out = nullptr;
    const saras_grid_desc &g = *grid;
    const saras_nse_params &P = *params;
    const bool planar = g.dim == 2;
    if (g.dim != 3 && !planar) {
        saras::set_last_error("grid dim must be 2 (planar x-z) or 3");
        return SARAS_BAD_SHAPE;
    }
    // Every PLANAR loop (vfield.cc:211-349, sfield.cc:164-260, predictors.cc)
    // is the 3-D one without its y terms; the planar handle keeps the 3-D
    // operand order with every y term an exact zero.  The PLANAR grid's y
    // pad of 1 (grid.cc:92-95) is never read, so planar caller arrays carry
    // no y pad.  With dScheme 2 the Morinishi z edge test reads
    // core.ubound(1) = 0 (vfield.cc:329), so the last z cell takes the 4th-order
    // stencil: on a z wall that reads the second ghost layer above it, stored
    // as for 3-D grids with ny != nz.
    if (planar && (P.les || g.n[1] != 1)) {
        saras::set_last_error("the planar path covers n_y = 1 without the LES model (3-D only in the reference)");
        return SARAS_BAD_ARG;
    }
    if ((P.dscheme != 1 && P.dscheme != 2) || P.nlscheme < 1 || P.nlscheme > 3 || g.pad < P.dscheme ||
        (P.scheme != 0 && P.scheme != 1)) {
        saras::set_last_error("dscheme must be 1 or 2 with pad >= dscheme; nlscheme 1, 2 or 3; scheme 0 or 1");
        return SARAS_BAD_ARG;
    }
    for (int a = 0; a < 3; ++a)
        if (g.n[a] < 4 && !(planar && a == 1)) return SARAS_BAD_SHAPE;
    for (int w = 0; w < 6; ++w) {
        // nullBC: a pressure wall of a wall axis (the planar y walls are none)
        const bool null_ok = P.pbc_type[w] == SARAS_FBC_NULL && !g.periodic[w / 2] && !(planar && w / 2 == 1);
        bool ok = P.pbc_type[w] == SARAS_FBC_DIRICHLET || P.pbc_type[w] == SARAS_FBC_NEUMANN || null_ok;
        for (int c = 0; c < 3; ++c)
            ok = ok && (P.vbc_type[c][w] == SARAS_FBC_DIRICHLET || P.vbc_type[c][w] == SARAS_FBC_NEUMANN);
        const int t = P.tbc_type[w];
        // hotPlate: the bottom wall of 3-D runs (scalar.cc:178-185; PLANAR builds
        // ignore the heating plate); plumeOut: the top wall
        const bool special = (t == SARAS_FBC_HOT_PLATE && w == 4 && !planar) || (t == SARAS_FBC_PLUME_OUT && w == 5);
        ok = ok && (t == SARAS_FBC_DIRICHLET || t == SARAS_FBC_NEUMANN || (special && P.has_T && !g.periodic[2]));
        if (!ok) {
            saras::set_last_error("wall kinds: Dirichlet or Neumann; for T of a run with z walls also the hot "
                                  "plate (bottom, 3-D) and the plume outflow (top); for P on a wall axis also null");
            return SARAS_BAD_ARG;
        }
    }
    if (P.has_T && P.mean_pgrad != 0.0) {
        saras::set_last_error("mean_pgrad (constantPGrad, force type 5) is a hydrodynamics forcing: has_T must be 0");
        return SARAS_BAD_ARG;
    }
    if (P.les && !(g.periodic[0] && g.periodic[1] && g.periodic[2])) {
        saras::set_last_error("the spiral LES path covers fully periodic boxes");
        return SARAS_BAD_SHAPE;
    }
    saras_nse *s = new saras_nse();
    s->g = g;
    s->p = P;
    s->planar = planar;
    saras_mg_params mp = *mgp;
    mp.bc = SARAS_BC_NEUMANN;
    // a planar plume (probType 9) solves for Pp with a Dirichlet top wall
    // (poisson_d2.cc:641-647); the 3-D solver keeps all walls Neumann
    const bool open_top_2d = planar && P.has_T && P.tbc_type[5] == SARAS_FBC_PLUME_OUT;
    mp.top_dirichlet = open_top_2d ? 1 : 0;
    saras_status e = saras_mg_create(&g, &mp, &s->mg);
    if (e) {
        delete s;
        return e;
    }
    std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
    s->s = saras::mg_stream(s->mg);
    Geom &G = s->G;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = g.n[a];
        G.per[a] = g.periodic[a];
    }
    if (planar) {
        G.n[1] = 1;
        G.per[1] = 1;
    }
    G.sy = pitch16(g.n[2]);
    G.sx = G.sy * G.n[1];
    {
        double *x0, *b0;
        long sx0, sy0;
        saras::mg_level0(s->mg, &x0, &b0, &sx0, &sy0);
        if (sx0 != G.sx || sy0 != G.sy) {
            saras::set_last_error("pitch mismatch with the Poisson level-0 arrays");
            delete s;
            return SARAS_HIP_ERROR;
        }
    }
    G.edge_last[0] = g.n[0] - 1;
    G.edge_last[1] = G.n[1] - 1;
    G.edge_last[2] = G.n[1] - 1;   // vfield.cc:434, sfield.cc:295
    G.dscheme = P.dscheme;
    double hwall[6];
    std::vector<double> coords[3];   // field coordinates over [-pad, n + pad)
    const size_t stride = (size_t)std::max(g.n[0], std::max(g.n[1], g.n[2])) + 2;
    std::vector<double> host(12 * stride, 0.0);
    for (int a = 0; a < 3; ++a) {
        if (planar && a == 1) {
            // no y direction: xi_y = 1 (a unit cell width in the volume
            // element), every y metric and y difference factor zero
            host[(3 * a + 0) * stride] = 1.0;
            G.ih[a] = G.ih2[a] = G.i2d[a] = G.tsi2h[a] = G.tsih2[a] = 0.0;
            hwall[2 * a] = hwall[2 * a + 1] = 0.0;
            continue;
        }
        std::vector<double> &coord = coords[a];
        std::vector<double> xi_x, xixx, xix2;
        saras::field_axis(g, a, coord, xi_x, xixx, xix2);
        const int p = g.pad, n = g.n[a];
        for (int i = 0; i < n; ++i) {
            host[(3 * a + 0) * stride + i] = xi_x[i + p];
            host[(3 * a + 1) * stride + i] = xixx[i + p];
            host[(3 * a + 2) * stride + i] = xix2[i + p];
        }
        for (int i = -1; i <= n; ++i) host[(9 + a) * stride + 1 + i] = coord[i + p];
        hwall[2 * a] = 2 * coord[p];
        hwall[2 * a + 1] = 2 * (coord[n - 1 + p] - g.len[a]);
        // derivative.cc:52-54, vfield.cc:133-135, timestep.cc:156-173
        const double dxi = 1.0 / double(n);
        G.ih[a] = 1.0 / dxi;
        G.ih2[a] = std::pow(G.ih[a], 2.0);
        G.i2d[a] = 1.0 / (2.0 * dxi);
        const double h2 = std::pow(dxi, 2.0);
        G.tsi2h[a] = 0.5 / dxi;
        G.tsih2[a] = 1.0 / h2;
        // tseries.cc:118-128: the top wall's distance when both walls count;
        // with an open top (plumeOut, probType 9) the bottom wall's
        if (a == 2)
            s->wall_dn = P.has_T && P.tbc_type[5] == SARAS_FBC_PLUME_OUT ? coord[p] : g.len[2] - coord[n - 1 + p];
    }
    auto fail = [&](saras_status st) {
        delete s;
        return st;
    };
    if (hipMalloc((void **)&s->metrics, host.size() * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
    if (hipMemcpy(s->metrics, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return fail(SARAS_HIP_ERROR);
    for (int a = 0; a < 3; ++a) {
        G.xi_x[a] = s->metrics + (3 * a + 0) * stride;
        G.xixx[a] = s->metrics + (3 * a + 1) * stride;
        G.xix2[a] = s->metrics + (3 * a + 2) * stride;
        G.xc[a] = s->metrics + (9 + a) * stride + 1;
    }
    // upwind weights (vfield.cc:115-130)
    {
        const double om = P.central_bias;
        G.pe_limit = P.pe_limit;
        G.up1[0] = 2.0 - om;
        G.up1[1] = 2.0 * (1.0 - om);
        G.up1[2] = -om;
        G.up2[0] = 1.0 - om;
        G.up2[1] = 4.0 - 3.0 * om;
        G.up2[2] = 3.0 * (1.0 - om);
        G.up2[3] = -om;
    }
    for (int c = 0; c < 3; ++c) s->bc[c] = make_bc(P.vbc_type[c], P.vbc_val[c], hwall);
    s->bc[3] = make_bc(P.tbc_type, P.tbc_val, hwall);
    if (P.has_T && (P.tbc_type[4] == SARAS_FBC_HOT_PLATE || P.tbc_type[5] == SARAS_FBC_PLUME_OUT)) {
        const long cols = (long)g.n[0] * G.n[1];
        std::vector<unsigned char> flags(2 * cols, 0);
        if (P.tbc_type[4] == SARAS_FBC_HOT_PLATE) {
            // hotPlate::createPatch (hotPlate.cc:74-111): the disc about
            // (Lx/2, Ly/2, 0), distances taken to the ghost cell (z(-1))
            const int p = g.pad;
            const double cx = g.len[0] / 2.0, cy = g.len[1] / 2.0, cz = 0.0;
            const double zg = coords[2][p - 1];
            for (int i = 0; i < g.n[0]; ++i)
                for (int j = 0; j < g.n[1]; ++j) {
                    const double r = std::sqrt(std::pow((coords[0][i + p] - cx), 2) +
                                               std::pow((coords[1][j + p] - cy), 2) + std::pow((zg - cz), 2));
                    flags[(long)i * g.n[1] + j] = r <= P.plate_radius ? 1 : 0;
                }
        }
        if (hipMalloc((void **)&s->wmask, flags.size()) != hipSuccess) return fail(SARAS_OOM);
        if (hipMemcpy(s->wmask, flags.data(), flags.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(SARAS_HIP_ERROR);
        G.wmask = s->wmask;
        s->plume = P.tbc_type[5] == SARAS_FBC_PLUME_OUT;
    }
    s->pbc = make_bc(P.pbc_type, P.pbc_val, hwall);
    for (int w = 0; w < 6; ++w) {
        if (P.pbc_type[w] != SARAS_FBC_NULL) continue;
        const size_t cells = s->plane_cells(w);
        if (hipMalloc((void **)&s->pghost[w], cells * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
        (void)hipMemset(s->pghost[w], 0, cells * sizeof(double));
        s->pbc.ghost[w] = s->pghost[w];
    }
    // Morinishi (dScheme 2) on a z-walled grid whose z edge test misses the
    // last z cell: that cell reads the second layer above the top wall
    if (P.nlscheme == 3 && P.dscheme == 2 && !g.periodic[2] && G.edge_last[2] != g.n[2] - 1) {
        const size_t cells = (size_t)g.n[0] * G.n[1];
        for (int q = 0; q < 4; ++q) {
            if (q == 3 && !P.has_T) continue;
            if (hipMalloc((void **)&s->top2[q], cells * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
            (void)hipMemset(s->top2[q], 0, cells * sizeof(double));
            s->bc[q].top2 = s->top2[q];
        }
    }
    {
        // Pp's ghosts as the solver's last imposeBC leaves them: Neumann, and
        // -Pp(n-1) at a Dirichlet top
        const int neu[6] = {1, 1, 1, 1, 1, open_top_2d ? 0 : 1};
        const double zero[6] = {0, 0, 0, 0, 0, 0};
        s->ppbc = make_bc(neu, zero, hwall);
    }
    s->elems = (size_t)G.sx * g.n[0];
    const int nf = 5;
    for (int q = 0; q < nf; ++q) {
        if (hipMalloc((void **)&s->F[q], s->elems * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
        (void)hipMemset(s->F[q], 0, s->elems * sizeof(double));
    }
    for (int q = 0; q < 4; ++q) {
        if (hipMalloc((void **)&s->H[q], s->elems * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
        if (hipMalloc((void **)&s->R[q], s->elems * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
        (void)hipMemset(s->H[q], 0, s->elems * sizeof(double));
        (void)hipMemset(s->R[q], 0, s->elems * sizeof(double));
    }
    if (hipMalloc((void **)&s->scratch, s->elems * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
    // timestep.cc:115
    s->max_iter = int(std::pow(std::log(double(g.n[0]) * g.n[1] * g.n[2]), 3));
    // one thread per cell of a row: the block of 64..256 threads (whole
    // waves) that leaves the fewest idle lanes in a row's last block, the
    // larger on a tie (384 cells: 2 blocks of 192, not 256 + half of 256)
    {
        int best = 64;
        long best_idle = -1;
        for (int t = 256; t >= 64; t -= 64) {
            const long idle = ((long)g.n[2] + t - 1) / t * t - g.n[2];
            if (best_idle < 0 || idle < best_idle) {
                best = t;
                best_idle = idle;
            }
        }
        s->threads = best;
    }
    {
        const dim3 gr = s->grid3();
        s->nblocks = (long)gr.x * gr.y * gr.z;
    }
    if (hipMalloc((void **)&s->part, sizeof(double) * s->nblocks) != hipSuccess) return fail(SARAS_OOM);
    // written by the maximum's last reduction directly (pinned, host-coherent)
    if (hipHostMalloc((void **)&s->res_host, sizeof(double), hipHostMallocCoherent) != hipSuccess) return fail(SARAS_OOM);
    if (hipHostMalloc((void **)&s->res_seq, sizeof(unsigned), hipHostMallocCoherent) != hipSuccess) return fail(SARAS_OOM);
    *s->res_seq = 0;
    if (hipMalloc((void **)&s->diag_part, sizeof(double) * saras::nse::kDiagSums * saras::nse::kDiagBlocks) != hipSuccess)
        return fail(SARAS_OOM);
    if (hipHostMalloc((void **)&s->diag_host, sizeof(double) * saras::nse::kDiagSums * saras::nse::kDiagBlocks) != hipSuccess)
        return fail(SARAS_OOM);
    if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) return fail(SARAS_HIP_ERROR);
    s->full = (long)(g.n[0] + 2 * g.pad) * (G.n[1] + 2 * s->ypad()) * (g.n[2] + 2 * g.pad);
    s->tstp = P.dt;
    if (P.les) {
        for (int m = 0; m < 6; ++m) {
            if (hipMalloc((void **)&s->Tsg[m], s->elems * sizeof(double)) != hipSuccess) return fail(SARAS_OOM);
            (void)hipMemset(s->Tsg[m], 0, s->elems * sizeof(double));
        }
        if (hipMalloc((void **)&s->sgs_part, sizeof(double) * 3 * s->nblocks) != hipSuccess) return fail(SARAS_OOM);
        if (hipMalloc((void **)&s->sgs_dev, sizeof(double) * 4) != hipSuccess) return fail(SARAS_OOM);
        if (hipHostMalloc((void **)&s->sgs_host, sizeof(double) * 4) != hipSuccess) return fail(SARAS_OOM);
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(SARAS_HIP_ERROR);
    *out = s;
    return SARAS_OK;
}

size_t saras_nse_footprint(const saras_grid_desc *grid, const saras_mg_params *mg, const saras_nse_params *params) {
    if (!grid || !mg || !params) return 0;
    saras_mg_params mp = *mg;
    mp.bc = SARAS_BC_NEUMANN;
    const size_t solver = saras_mg_footprint(grid, &mp);
    if (solver == 0) return 0;
    const bool planar = grid->dim == 2;
    const size_t ny = planar ? 1 : (size_t)grid->n[1];
    const size_t elems = (size_t)pitch16(grid->n[2]) * ny * grid->n[0];
    // V, T, P, H(4), R(4), the Jacobi buffer, and the six LES stresses
    const size_t arrays = 5 + 4 + 4 + 1 + (params->les ? 6 : 0);
    const size_t p = grid->pad, yp = planar ? 0 : p;
    const size_t full = (grid->n[0] + 2 * p) * (ny + 2 * yp) * (grid->n[2] + 2 * p);
    return solver + arrays * elems * sizeof(double) + 2 * full * sizeof(double) + (1 << 20);
}

saras_status saras_nse_set_graphs(saras_nse *s, int enable) {
    if (!s) return SARAS_BAD_ARG;
    return saras_mg_set_graphs(s->mg, enable);
}

saras_status saras_nse_set_time(saras_nse *s, double time) {
    if (!s || !std::isfinite(time)) return SARAS_BAD_ARG;
    s->time = time;
    return SARAS_OK;
}

saras_status saras_nse_destroy(saras_nse *s) {
    if (!s) return SARAS_BAD_ARG;
    (void)hipStreamSynchronize(s->s);
    saras_mg *mg = s->mg;
    s->mg = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(saras::alloc_mutex());
        delete s;
    }
    return saras_mg_destroy(mg);
}

saras_status saras_nse_set_fields(saras_nse *s, const double *vx, const double *vy, const double *vz,
                                  const double *p, const double *t, const double *pp, int ptr_kind) {
    if (!s || (ptr_kind != SARAS_PTR_HOST && ptr_kind != SARAS_PTR_DEVICE)) return SARAS_BAD_ARG;
    const double *src[5] = {vx, vy, vz, t, p};
    for (int q = 0; q < 5; ++q) {
        if (!src[q]) continue;
        saras_status e = load_field(s, src[q], ptr_kind, s->F[q]);
        if (e) return e;
    }
    if (p) {
        saras_status e = load_ghost_planes(s, p, ptr_kind);
        if (e) return e;
    }
    for (int q = 0; q < 4; ++q) {
        if (!src[q] || !s->top2[q]) continue;
        saras_status e = load_top2(s, src[q], ptr_kind, s->top2[q]);
        if (e) return e;
    }
    if (pp) {
        double *x0, *b0;
        long sx0, sy0, hz0;
        saras::mg_level0(s->mg, &x0, &b0, &sx0, &sy0, &hz0);
        saras_status e = load_field(s, pp, ptr_kind, x0, hz0);
        if (e) return e;
    }
    // the drivers' initial imposeBCs of V, P and then T
    saras_status e = plume_flags(s);
    if (e) return e;
    NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

saras_status saras_nse_get_fields(saras_nse *s, double *vx, double *vy, double *vz, double *p, double *t, double *pp,
                                  int ptr_kind) {
    if (!s || (ptr_kind != SARAS_PTR_HOST && ptr_kind != SARAS_PTR_DEVICE)) return SARAS_BAD_ARG;
    double *dst[5] = {vx, vy, vz, t, p};
    for (int q = 0; q < 5; ++q) {
        if (!dst[q]) continue;
        saras_status e = store_field(s, s->field(q), dst[q], ptr_kind, false);
        if (e) return e;
    }
    if (pp) {
        const saras::nse::SplitField f = s->pp();
        saras_status e = f.hz ? store_field(s, f, pp, ptr_kind, false)
                              : store_field(s, CoreField{f.f, f.bc}, pp, ptr_kind, false);
        if (e) return e;
    }
    NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

saras_status saras_nse_rhs(saras_nse *s, int stage) {
    if (!s || stage < 0 || stage > 2) return SARAS_BAD_ARG;
    saras_status e = rhs(s, stage);
    if (e) return e;
    NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

saras_status saras_nse_get_rhs(saras_nse *s, double *rx, double *ry, double *rz, double *rt, double *hx, double *hy,
                               double *hz, double *ht, int ptr_kind) {
    if (!s) return SARAS_BAD_ARG;
    double *dst[8] = {rx, ry, rz, rt, hx, hy, hz, ht};
    for (int q = 0; q < 8; ++q) {
        if (!dst[q]) continue;
        double *src = q < 4 ? s->R[q] : s->H[q - 4];
        saras_status e = store_field(s, CoreField{src, s->bc[0]}, dst[q], ptr_kind, true);
        if (e) return e;
    }
    return SARAS_OK;
}

saras_status saras_nse_solve_field(saras_nse *s, int field, double beta, int *iters) {
    if (!s || field < 0 || field > 3 || (field == 3 && !s->p.has_T)) return SARAS_BAD_ARG;
    int it = 0;
    saras_status e = solve_field(s, field, beta, &it);
    s->jac[field] = it;
    if (iters) *iters = it;
    return e;
}

saras_status saras_nse_stage(saras_nse *s, int stage, saras_nse_stats *st) {
    if (!s || stage < 0 || stage >= s->stages()) return SARAS_BAD_ARG;
    NSE_TRY(hipEventRecord(s->ev0, s->s));
    saras_status e = stage_run(s, stage);
    NSE_TRY(hipEventRecord(s->ev1, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    float ms = 0.f;
    NSE_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    fill_stats(s, st, ms);
    return e;
}

saras_status saras_nse_step(saras_nse *s, saras_nse_stats *st) {
    if (!s) return SARAS_BAD_ARG;
    NSE_TRY(hipEventRecord(s->ev0, s->s));
    saras_status e = SARAS_OK;
    for (int k = 0; k < s->stages() && !e; ++k) e = stage_run(s, k);
    NSE_TRY(hipEventRecord(s->ev1, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    float ms = 0.f;
    NSE_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    fill_stats(s, st, ms);
    if (!e) s->time += s->p.dt;
    return e;
}

saras_status saras_nse_bench_steps(saras_nse *s, int count, double *ms) {
    if (!s || count < 0 || !ms) return SARAS_BAD_ARG;
    NSE_TRY(hipStreamSynchronize(s->s));
    NSE_TRY(hipEventRecord(s->ev0, s->s));
    saras_status e = SARAS_OK;
    for (int c = 0; c < count && !e; ++c) {
        for (int k = 0; k < s->stages() && !e; ++k) e = stage_run(s, k);
        if (!e) s->time += s->p.dt;
    }
    NSE_TRY(hipEventRecord(s->ev1, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    float t = 0.f;
    NSE_TRY(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms = t;
    return e;
}

saras_status saras_nse_bench_sgs(saras_nse *s, int count, double *ms) {
    if (!s || count < 1 || !ms) return SARAS_BAD_ARG;
    if (!s->p.les) {
        saras::set_last_error("the handle has no sub-grid model (les = 0)");
        return SARAS_BAD_ARG;
    }
    NSE_TRY(hipStreamSynchronize(s->s));
    NSE_TRY(hipEventRecord(s->ev0, s->s));
    for (int c = 0; c < count; ++c) launch_sgs(s);
    NSE_TRY(hipGetLastError());
    NSE_TRY(hipEventRecord(s->ev1, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    float t = 0.f;
    NSE_TRY(hipEventElapsedTime(&t, s->ev0, s->ev1));
    *ms = t / count;
    return SARAS_OK;
}

saras_status saras_field_op(saras_nse *s, int op, const double *f, double *out, int ptr_kind) {
    if (!s || !f || !out || op < 0 || op > 7) return SARAS_BAD_ARG;
    if (ptr_kind != SARAS_PTR_HOST && ptr_kind != SARAS_PTR_DEVICE) return SARAS_BAD_ARG;
    if (s->planar) {
        saras::set_last_error("single field operators cover 3-D handles");
        return SARAS_BAD_SHAPE;
    }
    const int p = s->g.pad;
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->g.n[1] + 2 * p);
    const double *df = f;
    double *dout = out;
    if (ptr_kind == SARAS_PTR_HOST) {
        saras_status e = ensure_stage(s);
        if (e) return e;
        NSE_TRY(hipMemcpyAsync(s->stage_a, f, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        NSE_TRY(hipMemcpyAsync(s->stage_b, out, s->full * sizeof(double), hipMemcpyHostToDevice, s->s));
        df = s->stage_a;
        dout = s->stage_b;
    }
    PadField pf{df, SX, SY, p};
    const CoreField v0 = s->field(0), v1 = s->field(1), v2 = s->field(2);
    const dim3 gr = s->grid3(), bl(s->threads);
    switch (op) {
        case 0: hipLaunchKernelGGL(saras::nse::k_field_op<0>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 1: hipLaunchKernelGGL(saras::nse::k_field_op<1>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 2: hipLaunchKernelGGL(saras::nse::k_field_op<2>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 3: hipLaunchKernelGGL(saras::nse::k_field_op<3>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 4: hipLaunchKernelGGL(saras::nse::k_field_op<4>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 5: hipLaunchKernelGGL(saras::nse::k_field_op<5>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        case 6: hipLaunchKernelGGL(saras::nse::k_field_op<6>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
        default: hipLaunchKernelGGL(saras::nse::k_field_op<7>, gr, bl, 0, s->s, s->G, pf, v0, v1, v2, s->p.nlscheme, s->p.kappa, dout, SX, SY, p); break;
    }
    NSE_TRY(hipGetLastError());
    if (ptr_kind == SARAS_PTR_HOST)
        NSE_TRY(hipMemcpyAsync(out, s->stage_b, s->full * sizeof(double), hipMemcpyDeviceToHost, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

saras_status saras_field_vec_op(saras_nse *s, int op, const double *fx, const double *fy, const double *fz,
                                double *ox, double *oy, double *oz, int ptr_kind) {
    if (!s || op < 0 || op > 3) return SARAS_BAD_ARG;
    if (ptr_kind != SARAS_PTR_HOST && ptr_kind != SARAS_PTR_DEVICE) return SARAS_BAD_ARG;
    // inputs: V (ops 0, 2, 3) or the scalar fx (op 1); outputs: ox (op 0) or all three
    const bool vin = op != 1, vout = op != 0;
    if (!fx || (vin && (!fy || !fz)) || !ox || (vout && (!oy || !oz))) return SARAS_BAD_ARG;
    if (s->planar) {
        saras::set_last_error("single field operators cover 3-D handles");
        return SARAS_BAD_SHAPE;
    }
    const int p = s->g.pad;
    const long SY = s->g.n[2] + 2 * p, SX = SY * (s->g.n[1] + 2 * p);
    const size_t bytes = s->full * sizeof(double);
    const double *in[3] = {fx, vin ? fy : fx, vin ? fz : fx};
    double *out[3] = {ox, vout ? oy : nullptr, vout ? oz : nullptr};
    const double *din[3] = {in[0], in[1], in[2]};
    double *dout[3] = {out[0], out[1], out[2]};
    double *stage = nullptr;
    if (ptr_kind == SARAS_PTR_HOST) {
        // six full arrays for this call (a parity and tooling entry point)
        NSE_TRY(hipMalloc((void **)&stage, 6 * bytes));
        for (int c = 0; c < 3; ++c) {
            double *d = stage + c * s->full;
            if (hipMemcpyAsync(d, in[c], bytes, hipMemcpyHostToDevice, s->s) != hipSuccess) {
                (void)hipFree(stage);
                return SARAS_HIP_ERROR;
            }
            din[c] = d;
            if (out[c]) {
                double *e = stage + (3 + c) * s->full;
                if (hipMemcpyAsync(e, out[c], bytes, hipMemcpyHostToDevice, s->s) != hipSuccess) {
                    (void)hipFree(stage);
                    return SARAS_HIP_ERROR;
                }
                dout[c] = e;
            }
        }
    }
    const PadField A{din[0], SX, SY, p}, B{din[1], SX, SY, p}, Cc{din[2], SX, SY, p};
    const dim3 gr = s->grid3(), bl(s->threads);
    const double nu = s->p.nu;
    switch (op) {
        case 0: hipLaunchKernelGGL(saras::nse::k_field_vec_op<0>, gr, bl, 0, s->s, s->G, A, B, Cc, s->p.nlscheme, nu, dout[0], dout[1], dout[2], SX, SY, p); break;
        case 1: hipLaunchKernelGGL(saras::nse::k_field_vec_op<1>, gr, bl, 0, s->s, s->G, A, B, Cc, s->p.nlscheme, nu, dout[0], dout[1], dout[2], SX, SY, p); break;
        case 2: hipLaunchKernelGGL(saras::nse::k_field_vec_op<2>, gr, bl, 0, s->s, s->G, A, B, Cc, s->p.nlscheme, nu, dout[0], dout[1], dout[2], SX, SY, p); break;
        default: hipLaunchKernelGGL(saras::nse::k_field_vec_op<3>, gr, bl, 0, s->s, s->G, A, B, Cc, s->p.nlscheme, nu, dout[0], dout[1], dout[2], SX, SY, p); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ptr_kind == SARAS_PTR_HOST)
        for (int c = 0; c < 3 && e == hipSuccess; ++c)
            if (out[c]) e = hipMemcpyAsync(out[c], dout[c], bytes, hipMemcpyDeviceToHost, s->s);
    if (e == hipSuccess) e = hipStreamSynchronize(s->s);
    if (stage) (void)hipFree(stage);
    if (e != hipSuccess) {
        saras::set_last_error(std::string("saras_field_vec_op: ") + hipGetErrorString(e));
        return SARAS_HIP_ERROR;
    }
    return SARAS_OK;
}

saras_status saras_nse_sgs_diag(saras_nse *s, double out[3], double *time) {
    if (!s) return SARAS_BAD_ARG;
    if (out)
        for (int q = 0; q < 3; ++q) out[q] = s->sgs_diag[q];
    if (time) *time = s->time;
    return SARAS_OK;
}

saras_status saras_nse_cfl(saras_nse *s, double courant, double dt_max, double *dt) {
    if (!s || !dt || !(courant > 0)) return SARAS_BAD_ARG;
    // avgDx = xLen / globalSize(0) (vfield.cc:91-95)
    const double ax = s->g.len[0] / s->g.n[0], ay = s->g.len[1] / s->g.n[1], az = s->g.len[2] / s->g.n[2];
    hipLaunchKernelGGL(saras::nse::k_cfl, s->grid3(), dim3(s->threads), 0, s->s, s->G, s->F[0], s->F[1], s->F[2],
                       ax, ay, az, s->part);
    hipLaunchKernelGGL(saras::nse::k_max_final, dim3(1), dim3(1024), 0, s->s, s->part, s->nblocks, s->res_host,
                       s->res_seq, ++s->res_ticket);
    NSE_TRY(hipGetLastError());
    NSE_TRY(hipStreamSynchronize(s->s));
    double d = courant / s->res_host[0];
    if (d > dt_max) d = dt_max;
    s->p.dt = d;
    *dt = d;
    return SARAS_OK;
}

saras_status saras_nse_diag(saras_nse *s, double out[4]) {
    if (!s || !out) return SARAS_BAD_ARG;
    double ts[5];
    const saras_status e = saras_nse_tsdata(s, ts);
    if (e) return e;
    out[0] = ts[0];
    out[1] = ts[2];
    out[2] = ts[3];
    out[3] = ts[4];
    return SARAS_OK;
}

saras_status saras_nse_tsdata(saras_nse *s, double out[5]) {
    if (!s || !out) return SARAS_BAD_ARG;
    constexpr int NS = saras::nse::kDiagSums;
    const long total = (long)s->G.n[0] * s->G.n[1] * s->G.n[2];
    const int blocks = (int)std::min<long>(saras::nse::kDiagBlocks, (total + 255) / 256);
    hipLaunchKernelGGL(saras::nse::k_diag, dim3(blocks), dim3(saras::nse::kDiagThreads), 0, s->s, s->G, s->field(0),
                       s->field(1), s->field(2), s->field(3), s->p.has_T, 1.0 / double(s->G.n[0]),
                       1.0 / double(s->G.n[1]), 1.0 / double(s->G.n[2]), s->wall_dn, s->diag_part);
    NSE_TRY(hipGetLastError());
    NSE_TRY(hipMemcpyAsync(s->diag_host, s->diag_part, sizeof(double) * NS * blocks, hipMemcpyDeviceToHost, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    double acc[NS] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < blocks; ++b) {
        for (int q = 0; q < NS; ++q) {
            const double v = s->diag_host[NS * b + q];
            if (q == 2) acc[q] = (v > acc[q] || v != v) ? v : acc[q];
            else acc[q] += v;
        }
    }
    const double vol = s->volume();
    out[0] = acc[0] / vol;
    out[1] = s->p.has_T ? acc[5] / vol : 0.0;
    out[2] = acc[2];
    out[3] = s->p.has_T ? 1.0 + (acc[1] / vol) / s->p.kappa : 0.0;
    // wall Nusselt number over both walls, or the bottom one alone under an
    // open top (tseries.cc:114-128: wallCount 1 and no top sum for probType 9)
    if (!s->p.has_T)
        out[4] = 0.0;
    else if (s->p.tbc_type[5] == SARAS_FBC_PLUME_OUT)
        out[4] = acc[3] / (s->wall_area() * 1);
    else
        out[4] = (acc[3] + acc[4]) / (s->wall_area() * 2);
    return SARAS_OK;
}

saras_status saras_nse_probe(saras_nse *s, int count, const int *ijk, double *out) {
    if (!s || count < 0 || (count > 0 && (!ijk || !out))) return SARAS_BAD_ARG;
    if (count == 0) return SARAS_OK;
    for (int q = 0; q < count; ++q)
        for (int a = 0; a < 3; ++a)
            if (ijk[3 * q + a] < 0 || ijk[3 * q + a] >= s->G.n[a]) {
                saras::set_last_error("probe point outside the core");
                return SARAS_BAD_ARG;
            }
    int *dij = nullptr;
    double *dout = nullptr;
    if (hipMalloc((void **)&dij, sizeof(int) * 3 * count) != hipSuccess) return SARAS_OOM;
    if (hipMalloc((void **)&dout, sizeof(double) * 5 * count) != hipSuccess) {
        (void)hipFree(dij);
        return SARAS_OOM;
    }
    saras_status st = SARAS_OK;
    if (hipMemcpyAsync(dij, ijk, sizeof(int) * 3 * count, hipMemcpyHostToDevice, s->s) != hipSuccess) st = SARAS_HIP_ERROR;
    if (!st) {
        // Vx, Vy, Vz, P, T (T only with a scalar field)
        const double *f[5] = {s->F[0], s->F[1], s->F[2], s->F[4], s->p.has_T ? s->F[3] : nullptr};
        hipLaunchKernelGGL(saras::nse::k_probe, dim3((count + 255) / 256), dim3(256), 0, s->s, s->G, f[0], f[1], f[2],
                           f[3], f[4], count, dij, dout);
        if (hipGetLastError() != hipSuccess) st = SARAS_HIP_ERROR;
    }
    if (!st && hipMemcpyAsync(out, dout, sizeof(double) * 5 * count, hipMemcpyDeviceToHost, s->s) != hipSuccess)
        st = SARAS_HIP_ERROR;
    if (!st && hipStreamSynchronize(s->s) != hipSuccess) st = SARAS_HIP_ERROR;
    (void)hipFree(dij);
    (void)hipFree(dout);
    return st;
}

}  // extern "C"

namespace {

// Probe staging of saras_nse_probe_noalloc: the point indices (3 ints per
// point) followed by the values (5 doubles per point) of at most
// SARAS_PROBE_NOALLOC_MAX points.  Each GPU holds one copy for the life of
// the process.  Host code never passes `scratch` itself to a kernel or a copy
// (on the host the name is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns for the current GPU.
constexpr size_t kProbeIdxBytes = (size_t)SARAS_PROBE_NOALLOC_MAX * 3 * sizeof(int);
constexpr size_t kProbeScratchBytes = kProbeIdxBytes + (size_t)SARAS_PROBE_NOALLOC_MAX * 5 * sizeof(double);
static_assert(kProbeIdxBytes % 256 == 0, "the value block starts 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kProbeScratchBytes];

constexpr int kMaxDevices = 64;

struct ProbeScratch {
    std::mutex mu;                     // one call at a time per GPU: calls share the GPU's array
    unsigned char *ptr = nullptr;      // device address of `scratch` on this GPU
    bool ready = false;
};

ProbeScratch &probe_scratch(int dev) {
    static ProbeScratch per_dev[kMaxDevices];
    return per_dev[dev];
}

}  // namespace

extern "C" {

saras_status saras_nse_probe_noalloc(saras_nse *s, int count, const int *ijk, double *out) {
    if (!s || count < 0 || (count > 0 && (!ijk || !out))) return SARAS_BAD_ARG;
    if (count > SARAS_PROBE_NOALLOC_MAX) {
        saras::set_last_error("saras_nse_probe_noalloc: " + std::to_string(count) + " points exceed the limit of " +
                              std::to_string(SARAS_PROBE_NOALLOC_MAX));
        return SARAS_BAD_ARG;
    }
    if (count == 0) return SARAS_OK;
    for (int q = 0; q < count; ++q)
        for (int a = 0; a < 3; ++a)
            if (ijk[3 * q + a] < 0 || ijk[3 * q + a] >= s->G.n[a]) {
                saras::set_last_error("probe point outside the core");
                return SARAS_BAD_ARG;
            }
    int dev = 0;
    NSE_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        saras::set_last_error("saras_nse_probe_noalloc: device index out of range");
        return SARAS_BAD_ARG;
    }
    ProbeScratch &ps = probe_scratch(dev);
    std::lock_guard<std::mutex> lk(ps.mu);
    if (!ps.ready) {
        // once per GPU, with that GPU current
        NSE_TRY(hipGetSymbolAddress((void **)&ps.ptr, HIP_SYMBOL(scratch)));
        if (!ps.ptr) {
            saras::set_last_error("saras_nse_probe_noalloc: no device address for the probe scratch");
            return SARAS_HIP_ERROR;
        }
        ps.ready = true;
    }
    int *dij = reinterpret_cast<int *>(ps.ptr);
    double *dout = reinterpret_cast<double *>(ps.ptr + kProbeIdxBytes);
    NSE_TRY(hipMemcpyAsync(dij, ijk, sizeof(int) * 3 * count, hipMemcpyHostToDevice, s->s));
    const double *f[5] = {s->F[0], s->F[1], s->F[2], s->F[4], s->p.has_T ? s->F[3] : nullptr};
    hipLaunchKernelGGL(saras::nse::k_probe, dim3((count + 255) / 256), dim3(256), 0, s->s, s->G, f[0], f[1], f[2], f[3],
                       f[4], count, dij, dout);
    NSE_TRY(hipGetLastError());
    NSE_TRY(hipMemcpyAsync(out, dout, sizeof(double) * 5 * count, hipMemcpyDeviceToHost, s->s));
    NSE_TRY(hipStreamSynchronize(s->s));
    return SARAS_OK;
}

}  // extern "C"
