// saras_dropin.hpp -- header-only C++ owners of the C-ABI handles of
// saras_hip.h, with the reference's class roles and method names, so the
// reference's callers change one member type and keep their call sites.
//
//   saras_amd::multigrid  stands for multigrid_d3 / multigrid_d2
//                         (lib/poisson/poisson.h:125-205); mgSolve() is
//                         poisson::mgSolve (poisson.cc:109-189)
//   saras_amd::lsRK3      stands for lsRK3_d3 (lib/timestep/lsRK3_d3.cc,
//                         timestep.h:57-58) together with the fields it
//                         advances; timeAdvance() is lsRK3_d3::timeAdvance
//
// Errors: the reference prints and calls MPI_Finalize(); exit(0) where
// these throw saras_amd::error (status + saras_last_error() text).  A
// V-cycle count that runs out is silent in the reference and is reported by
// stats() here, not thrown.
#pragma once
#include <stdexcept>
#include <string>

#include "saras_hip.h"

namespace saras_amd {

struct error : std::runtime_error {
    saras_status status;
    error(saras_status s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void check(saras_status s, const char *where) {
    if (s != SARAS_OK) throw error(s, std::string(where) + ": " + saras_last_error());
}

class multigrid {
  public:
    multigrid(const saras_grid_desc &grid, const saras_mg_params &params) {
        check(saras_mg_create(&grid, &params, &h_), "saras_mg_create");
    }
    ~multigrid() {
        if (h_) saras_mg_destroy(h_);
    }
    multigrid(const multigrid &) = delete;
    multigrid &operator=(const multigrid &) = delete;

    // outLHS: in/out full array (pad from the grid descriptor), core read as
    // the initial guess, core + first ghost layer written; inpRHS: core read.
    void mgSolve(double *outLHS, const double *inpRHS, int ptr_kind = SARAS_PTR_HOST, void *stream = nullptr) {
        const saras_status s = saras_mg_solve(h_, outLHS, inpRHS, ptr_kind, stream, &stats_);
        if (s != SARAS_NOT_CONVERGED) check(s, "saras_mg_solve");
    }
    const saras_mg_stats &stats() const { return stats_; }
    saras_mg *handle() const { return h_; }

  private:
    saras_mg *h_ = nullptr;
    saras_mg_stats stats_{};
};

class lsRK3 {
  public:
    lsRK3(const saras_grid_desc &grid, const saras_mg_params &poisson, const saras_nse_params &params) {
        check(saras_nse_create(&grid, &poisson, &params, &h_), "saras_nse_create");
    }
    ~lsRK3() {
        if (h_) saras_nse_destroy(h_);
    }
    lsRK3(const lsRK3 &) = delete;
    lsRK3 &operator=(const lsRK3 &) = delete;

    // Full arrays in the reference layout; nullptr keeps a field.
    void setFields(const double *vx, const double *vy, const double *vz, const double *p, const double *t,
                   int ptr_kind = SARAS_PTR_HOST) {
        check(saras_nse_set_fields(h_, vx, vy, vz, p, t, nullptr, ptr_kind), "saras_nse_set_fields");
    }
    void getFields(double *vx, double *vy, double *vz, double *p, double *t, int ptr_kind = SARAS_PTR_HOST) {
        check(saras_nse_get_fields(h_, vx, vy, vz, p, t, nullptr, ptr_kind), "saras_nse_get_fields");
    }
    // lsRK3_d3::timeAdvance(V, P, T) / (V, P): one time step
    void timeAdvance() { check(saras_nse_step(h_, &stats_), "saras_nse_step"); }
    // tseries diagnostics: kinetic energy, max |div V|, NuVol, NuWall
    void diagnostics(double out[4]) { check(saras_nse_diag(h_, out), "saras_nse_diag"); }
    // vfield::computeTStp with the driver's cap (vfield.cc:691-706); the
    // returned time step is the one the following steps use
    double computeTStp(double courant, double dt_max) {
        double dt = 0.0;
        check(saras_nse_cfl(h_, courant, dt_max, &dt), "saras_nse_cfl");
        return dt;
    }
    // spiral LES diagnostics of the last step (sum |K| dV, SGS dissipation dV,
    // nu_t dV; divide by the volume as tseries does) and the run time
    void sgsDiagnostics(double out[3], double *time) {
        check(saras_nse_sgs_diag(h_, out, time), "saras_nse_sgs_diag");
    }
    const saras_nse_stats &stats() const { return stats_; }
    saras_nse *handle() const { return h_; }

  private:
    saras_nse *h_ = nullptr;
    saras_nse_stats stats_{};
};

}  // namespace saras_amd
