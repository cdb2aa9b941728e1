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
//   saras_amd::tseries    stands for tseries (lib/io/tseries.h): the
//                         TimeSeries.dat lines of writeTSHeader/writeTSData
//   saras_amd::probes     stands for probes (lib/io/probes.h): the
//                         ProbesData.dat lines of probeData
//
// Errors: the reference prints and calls MPI_Finalize(); exit(0) where
// these throw saras_amd::error (status + saras_last_error() text).  A
// V-cycle count that runs out is silent in the reference and is reported by
// stats() here, not thrown.
#pragma once
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

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

// tseries (tseries.cc:60-195, 209-490): the header of writeTSHeader and one
// line per writeTSData, in the file and on the console, in the reference's
// fixed-width format.  `scalar` selects the scalar solver's columns (probType
// above 4: Re, Nu, KE, TE), `les` the sub-grid columns; nu is timestep's
// mDiff (timestep.cc:120), volume the box volume.
// writeTSData's divergence guard (max |div V| above 1e3 and rising ends the
// run, tseries.cc:217-223) throws saras_amd::error instead of exit(0).
class tseries {
  public:
    tseries(lsRK3 &solver, const std::string &path, double nu, double volume, bool scalar, bool les,
            bool append = false, std::ostream *console = &std::cout)
        : ts_(solver), nu_(nu), vol_(volume), scalar_(scalar), les_(les), con_(console) {
        file_.open(path, append ? std::fstream::out | std::fstream::app : std::fstream::out);
        if (!file_) throw error(SARAS_BAD_ARG, "tseries: cannot open " + path);
    }

    void writeTSHeader() {
        if (!scalar_) {
            if (con_) *con_ << std::setw(9) << "Time" << std::setw(20) << "Total KE" << std::setw(20) << "Divergence\n";
            file_ << (les_ ? "#VARIABLES = Time, Total KE, U_rms, Divergence, Subgrid KE, SG Dissipation, Turb. Viscosity, dt\n"
                           : "#VARIABLES = Time, Total KE, U_rms, Divergence, dt\n");
        } else {
            if (con_)
                *con_ << std::setw(9) << "Time" << std::setw(20) << "Re (Urms)" << std::setw(20) << "Nu (Wall)"
                      << std::setw(20) << "Divergence" << "\n";
            file_ << (les_ ? "#VARIABLES = Time, Re, Nu (wall), Nu (vol), Total KE, Total TE, Divergence, Subgrid KE, "
                             "SG Dissipation, Turb. Viscosity, dt\n"
                           : "#VARIABLES = Time, Re, Nu (wall), Nu (vol), Total KE, Total TE, Divergence, dt\n");
        }
        file_.flush();
    }

    // one line for the fields as they stand at `time`, with time step dt
    void writeTSData(double time, double dt) {
        double d[5];
        check(saras_nse_tsdata(ts_.handle(), d), "saras_nse_tsdata");
        const double ke = d[0], te = d[1], div = d[2], nuvol = d[3], nuwall = d[4];
        if (div > 1e3 && div - old_div_ > 0)
            throw error(SARAS_BAD_ARG, "tseries: divergence exceeds permissible limits");
        old_div_ = div;
        double sg[3] = {0.0, 0.0, 0.0};
        if (les_) {
            double t = 0.0;
            ts_.sgsDiagnostics(sg, &t);
            for (double &v : sg) v /= vol_;
        }
        auto cols = [&](std::ostream &o, std::initializer_list<double> v) {
            o << std::fixed << std::setprecision(4) << std::setw(9) << time << std::setprecision(8);
            for (double x : v) o << std::setw(20) << x;
        };
        if (!scalar_) {
            if (con_) {
                cols(*con_, {ke, div});
                *con_ << "\n";
            }
            if (les_) cols(file_, {ke, std::sqrt(2.0 * ke), div, sg[0], sg[1], sg[2], dt});
            else cols(file_, {ke, std::sqrt(2.0 * ke), div, dt});
        } else {
            const double re = std::sqrt(2.0 * ke) / nu_;
            if (con_) {
                cols(*con_, {re, nuwall, div});
                *con_ << "\n";
            }
            if (les_) cols(file_, {re, nuwall, nuvol, ke, te, div, sg[0], sg[1], sg[2], dt});
            else cols(file_, {re, nuwall, nuvol, ke, te, div, dt});
        }
        file_ << "\n";
        file_.flush();
    }

  private:
    lsRK3 &ts_;
    double nu_, vol_;
    bool scalar_, les_;
    std::ostream *con_;
    std::ofstream file_;
    double old_div_ = 1.0e10;
};

// probes (probes.cc:61-176): ProbesData.dat with the header of the
// constructor and one line per probe per probeData call.  Points are the
// probes' global core indices (0-based, as the reference's probe list);
// fields Vx, Vy, Vz, P and, with a scalar, T.
class probes {
  public:
    probes(lsRK3 &solver, const std::string &path, const std::vector<int> &ijk, bool scalar, bool restart = false)
        : ts_(solver), ijk_(ijk), scalar_(scalar) {
        if (ijk.size() % 3) throw error(SARAS_BAD_ARG, "probes: ijk holds (i, j, k) triples");
        file_.open(path, std::fstream::out | std::fstream::app);
        if (!file_) throw error(SARAS_BAD_ARG, "probes: cannot open " + path);
        if (!restart) file_ << (scalar ? "#VARIABLES = Time, X, Y, Z, Vx, Vy, Vz, P, T" : "#VARIABLES = Time, X, Y, Z, Vx, Vy, Vz, P")
                            << std::endl;
    }

    void probeData(double time) {
        const int n = (int)(ijk_.size() / 3);
        std::vector<double> v(5 * (size_t)n);
        check(saras_nse_probe(ts_.handle(), n, ijk_.data(), v.data()), "saras_nse_probe");
        const int nf = scalar_ ? 5 : 4;
        for (int q = 0; q < n; ++q) {
            file_ << std::fixed << std::setw(6) << std::setprecision(4) << time << "\t";
            file_ << std::setw(4) << std::setprecision(0) << ijk_[3 * q] << "\t" << ijk_[3 * q + 1] << "\t"
                  << ijk_[3 * q + 2] << "\t";
            file_ << std::setw(16) << std::setprecision(8) << v[5 * (size_t)q];
            for (int f = 1; f < nf; ++f) file_ << "\t" << std::setw(16) << std::setprecision(8) << v[5 * (size_t)q + f];
            file_ << std::endl;
        }
    }

  private:
    lsRK3 &ts_;
    std::vector<int> ijk_;
    bool scalar_;
    std::ofstream file_;
};

}  // namespace saras_amd
