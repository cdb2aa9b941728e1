// Restart of a planar (x-z) run through the HDF5 files of
// include/saras_h5.hpp, as the reference's PLANAR build writes and reads
// them (hydro_d2.cc:64-76, 130-140): a lid-driven cavity at Re 1000 on an
// n x 1 x n grid runs `steps` Euler-CN steps, writes its solution and restart
// files, and a second solver reads the restart file back.  The fields must
// agree bit for bit, and so must both solvers after `steps` more steps each:
// the restart file holds everything a step reads except the warm start of
// the pressure-correction solve (Pp, not written by writer.cc:417-478), which
// a restarted run begins at zero, so the first solver's Pp is cleared too.
//
// usage: planar_h5 <n> <steps> <h5dir>
// exit status: 0 ok, 2 usage, 5 library error, 6 restart time differs,
// 7 fields differ after the read, 9 fields differ after the further steps
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "saras_dropin.hpp"
#include "saras_h5.hpp"

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s n steps h5dir\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), steps = std::atoi(argv[2]), pad = 1;
    saras_grid_desc g{};
    g.dim = 2;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = a == 1 ? 1 : n;
        g.len[a] = 1.0;
        g.gtype[a] = a == 1 ? SARAS_GRID_UNIFORM : SARAS_GRID_TANH_DOUBLE;
        g.beta[a] = 1.0;
    }
    g.pad = pad;
    saras_mg_params mp{};
    mp.vc_count = 7;
    mp.vc_tol = 1e-6;
    mp.mg_tol = 1e-6;
    mp.sor = 1.0;
    mp.pre_smooth = 3;
    mp.post_smooth = 3;
    mp.res_type = 1;
    mp.bc = SARAS_BC_NEUMANN;
    mp.precision = SARAS_PREC_FP64;
    saras_nse_params p{};
    p.dscheme = 1;
    p.nlscheme = 1;
    p.nu = p.kappa = 1e-3;
    p.dt = 1e-3;
    p.cn_tol = 1e-6;
    p.scheme = 1;
    for (int w = 0; w < 6; ++w) {
        for (int c = 0; c < 3; ++c) p.vbc_type[c][w] = SARAS_FBC_DIRICHLET;
        p.pbc_type[w] = SARAS_FBC_NEUMANN;
    }
    p.vbc_val[0][5] = 1.0;   // the lid: Vx = 1 on the top wall (hydro.cc:365-368)
    const size_t full = (size_t)(n + 2 * pad) * (n + 2 * pad);
    auto fields = [&](saras_amd::lsRK3 &ts) {
        std::vector<std::vector<double>> f(4, std::vector<double>(full, 0.0));
        ts.getFields(f[0].data(), f[1].data(), f[2].data(), f[3].data(), nullptr);
        return f;
    };
    auto same = [&](const std::vector<std::vector<double>> &a, const std::vector<std::vector<double>> &b) {
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k) {
                    const size_t at = (size_t)(i + pad) * (n + 2 * pad) + (k + pad);
                    if (a[q][at] != b[q][at]) return false;
                }
        return true;
    };
    try {
        std::vector<double> zero(full, 0.0);
        saras_amd::lsRK3 ts(g, mp, p);
        ts.setFields(zero.data(), zero.data(), zero.data(), zero.data(), nullptr);
        for (int s = 0; s < steps; ++s) ts.timeAdvance();
        const double time = steps * p.dt;
        saras_amd::writer w(ts, g, false, argv[3]);
        w.writeSolution(time);
        w.writeRestart(time);
        saras_amd::lsRK3 ts2(g, mp, p);
        saras_amd::reader r(ts2, g, false, argv[3]);
        if (r.readRestart() != time) return 6;
        if (!same(fields(ts), fields(ts2))) return 7;
        if (saras_nse_set_fields(ts.handle(), nullptr, nullptr, nullptr, nullptr, nullptr, zero.data(),
                                 SARAS_PTR_HOST) != SARAS_OK)
            return 5;
        for (int s = 0; s < steps; ++s) {
            ts.timeAdvance();
            ts2.timeAdvance();
        }
        if (!same(fields(ts), fields(ts2))) return 9;
    } catch (const saras_amd::error &e) {
        std::fprintf(stderr, "saras error %d: %s\n", (int)e.status, e.what());
        return 5;
    }
    return 0;
}
