// Exercises the C++ drop-in classes of include/saras_dropin.hpp the way the
// reference's callers use multigrid_d3 and lsRK3_d3: one mgSolve on a
// caller-owned array, then time steps of a Rayleigh-Benard run.
//
// usage: dropin_run <n> <steps> <in.bin> <out.bin> [<ts.dat> <probes.dat> [<h5dir>]]
//   in.bin : RHS b, then T, each (n+4)^3 doubles (pad 2)
//   out.bin: mgSolve result x, then Vx, Vy, Vz, P, T after the steps
//   ts.dat, probes.dat: TimeSeries.dat and ProbesData.dat of the steps
//   (tseries / probes), probes at (1, 2, 3), (n-1, 0, n/2), (5, 5, 0)
//   h5dir: (built with HDF5) the solution and restart files after the steps;
//   the restart file is read into a second solver, whose fields must equal
//   the first one's bit for bit
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "saras_dropin.hpp"
#ifdef SARAS_WITH_HDF5
#include "saras_h5.hpp"
#endif

int main(int argc, char **argv) {
    if (argc != 5 && argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s n steps in.bin out.bin [ts.dat probes.dat [h5dir]]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]), pad = 2;
    const size_t full = (size_t)(n + 2 * pad) * (n + 2 * pad) * (n + 2 * pad);
    std::vector<double> b(full), T(full);
    FILE *f = std::fopen(argv[3], "rb");
    if (!f || std::fread(b.data(), sizeof(double), full, f) != full || std::fread(T.data(), sizeof(double), full, f) != full)
        return 3;
    std::fclose(f);

    saras_grid_desc g{};
    g.dim = 3;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = n;
        g.len[a] = 1.0;
        g.periodic[a] = a < 2;
        g.gtype[a] = a == 2 ? SARAS_GRID_TANH_DOUBLE : SARAS_GRID_UNIFORM;
        g.beta[a] = a == 2 ? 1.5 : 1.0;
    }
    g.pad = pad;
    saras_mg_params mp{};
    mp.vc_count = 20;
    mp.vc_tol = 1e-6;
    mp.solve_flag = 0;
    mp.mg_tol = 1e-2;
    mp.sor = 1.0;
    mp.pre_smooth = 3;
    mp.post_smooth = 6;
    mp.res_type = 0;
    mp.bc = SARAS_BC_NEUMANN;
    mp.precision = SARAS_PREC_FP64;
    mp.inner_cycles = 2;
    mp.max_refine = 10;

    std::vector<double> x(full, 0.0);
    try {
        saras_amd::multigrid mg(g, mp);
        mg.mgSolve(x.data(), b.data());

        saras_nse_params p{};
        p.dscheme = 2;
        p.nlscheme = 3;
        p.has_T = 1;
        p.nu = p.kappa = std::sqrt(1.0 / 1e7);
        p.dt = 1e-3;
        p.cn_tol = 1e-6;
        p.buoy[2] = -1.0;
        for (int w = 0; w < 6; ++w) {
            for (int c = 0; c < 3; ++c) p.vbc_type[c][w] = SARAS_FBC_DIRICHLET;
            p.tbc_type[w] = w >= 4 ? SARAS_FBC_DIRICHLET : SARAS_FBC_NEUMANN;
            p.pbc_type[w] = SARAS_FBC_NEUMANN;
        }
        p.tbc_val[4] = 1.0;
        std::vector<double> zero(full, 0.0), one(full, 1.0);
        saras_amd::lsRK3 ts(g, mp, p);
        ts.setFields(zero.data(), zero.data(), zero.data(), one.data(), T.data());
        if (argc >= 7) {
            saras_amd::tseries tw(ts, argv[5], p.nu, 1.0, true, false, false, nullptr);
            saras_amd::probes pw(ts, argv[6], {1, 2, 3, n - 1, 0, n / 2, 5, 5, 0}, true);
            tw.writeTSHeader();
            for (int s = 0; s < steps; ++s) {
                ts.timeAdvance();
                tw.writeTSData((s + 1) * p.dt, p.dt);
                pw.probeData((s + 1) * p.dt);
            }
        } else {
            for (int s = 0; s < steps; ++s) ts.timeAdvance();
        }
        std::vector<std::vector<double>> out(5, std::vector<double>(full, 0.0));
        ts.getFields(out[0].data(), out[1].data(), out[2].data(), out[3].data(), out[4].data());
        FILE *o = std::fopen(argv[4], "wb");
        if (!o) return 4;
        std::fwrite(x.data(), sizeof(double), full, o);
        for (auto &v : out) std::fwrite(v.data(), sizeof(double), full, o);
        std::fclose(o);
        if (argc == 8) {
#ifdef SARAS_WITH_HDF5
            const double time = steps * p.dt;
            saras_amd::writer w(ts, g, true, argv[7]);
            w.writeSolution(time);
            w.writeRestart(time);
            saras_amd::lsRK3 ts2(g, mp, p);
            saras_amd::reader r(ts2, g, true, argv[7]);
            if (r.readRestart() != time) return 6;
            std::vector<std::vector<double>> back(5, std::vector<double>(full, 0.0));
            ts2.getFields(back[0].data(), back[1].data(), back[2].data(), back[3].data(), back[4].data());
            for (int q = 0; q < 5; ++q)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j)
                        for (int k = 0; k < n; ++k) {
                            const size_t at = ((size_t)(i + pad) * (n + 2 * pad) + (j + pad)) * (n + 2 * pad) + (k + pad);
                            if (back[q][at] != out[q][at]) return 7;
                        }
#else
            return 8;
#endif
        }
    } catch (const saras_amd::error &e) {
        std::fprintf(stderr, "saras error %d: %s\n", (int)e.status, e.what());
        return 5;
    }
    return 0;
}
