// grid_host.cpp -- see grid_host.hpp.  Compiled with -ffp-contract=off so that
// every expression rounds exactly as the reference's x86-64 build does.
#include "grid_host.hpp"

#include <cmath>

namespace saras {

namespace {

bool is_planar_y(const saras_grid_desc &g, int axis) { return g.dim == 2 && axis == 1; }

// One point of the tanh mapping and its metrics (grid.cc:437-460, 621-641).
void tanh_map(bool two_sided, double L, double beta, double xi, double &x, double &fx,
              double &fxx, double &fx2) {
    const double thb = std::tanh(beta);
    const double btl = beta * L;
    if (two_sided) {
        x = L * (1.0 - std::tanh(beta * (1.0 - 2.0 * xi)) / thb) / 2.0;
        const double ndx = x / L;
        fx = thb / (btl * (1.0 - std::pow((1.0 - 2.0 * ndx) * thb, 2)));
        fxx = -4.0 * std::pow(thb, 3) * (1.0 - 2.0 * ndx) /
              (L * btl * std::pow(1.0 - std::pow(thb * (1.0 - 2.0 * ndx), 2), 2));
    } else {
        x = L * (1.0 - std::tanh(beta * (1.0 - xi)) / thb);
        const double ndx = x / L;
        fx = thb / (btl * (1.0 - std::pow((1.0 - ndx) * thb, 2)));
        fxx = -2.0 * std::pow(thb, 3) * (1.0 - ndx) /
              (L * btl * std::pow(1.0 - std::pow(thb * (1.0 - ndx), 2), 2));
    }
    fx2 = std::pow(fx, 2.0);
}

// Level-0 arrays of one axis over [-p, n+p-1]; `uniform_only` skips the
// stretching (mgGridMetrics() runs before createTanHypGrid).
void level0(const saras_grid_desc &g, int axis, int p, bool uniform_only, std::vector<double> &xi,
            std::vector<double> &x, std::vector<double> &fx, std::vector<double> &fxx,
            std::vector<double> &fx2) {
    const bool py = is_planar_y(g, axis);
    const int n = py ? 1 : g.n[axis];
    const double L = g.len[axis];
    const double dxi = py ? 1.0 : 1.0 / double(n);
    const int tot = n + 2 * p;
    xi.assign(tot, 0.0); x.assign(tot, 0.0); fx.assign(tot, 0.0);
    fxx.assign(tot, 0.0); fx2.assign(tot, 0.0);
    const double inv_len = 1.0 / L;
    const double inv_len2 = std::pow(inv_len, 2.0);
    for (int i = -p; i < n + p; ++i) {
        const int q = i + p;
        xi[q] = double(2 * i + 1) * dxi / 2;
        x[q] = L * xi[q];
        fx[q] = inv_len;
        fxx[q] = 0.0;
        fx2[q] = inv_len2;
    }
    if (py || uniform_only || g.gtype[axis] == SARAS_GRID_UNIFORM) return;
    const bool two = g.gtype[axis] == SARAS_GRID_TANH_DOUBLE;
    for (int i = 0; i < n; ++i)
        tanh_map(two, L, g.beta[axis], xi[i + p], x[i + p], fx[i + p], fxx[i + p], fx2[i + p]);
    // ghost points continue the grid periodically (grid.cc:466-480)
    for (int q = 0; q < p; ++q) {
        const int dst = q, src = n + q;           // index -p+q <- n-p+q
        x[dst] = x[src] - L;
        fx[dst] = fx[src]; fxx[dst] = fxx[src]; fx2[dst] = fx2[src];
    }
    for (int q = 0; q < p; ++q) {
        const int dst = n + p + q, src = p + q;   // index n+q <- q
        x[dst] = L + x[src];
        fx[dst] = fx[src]; fxx[dst] = fxx[src]; fx2[dst] = fx2[src];
    }
}

}  // namespace

int vc_depth(const saras_grid_desc &g) {
    std::vector<int> ext;
    if (g.dim == 2) ext = {g.n[0], g.n[2]};
    else ext = {g.n[0], g.n[1], g.n[2]};
    for (int e : ext)
        if (e < 1) return -1;
    int d = 0;
    for (;;) {
        bool rem = false;
        for (int e : ext) rem = rem || (e % (1 << d)) != 0;
        if (rem) break;
        ++d;
    }
    --d;
    int smallest = ext[0] >> d;
    for (int e : ext) smallest = std::min(smallest, e >> d);
    if (smallest == 1) --d;
    bool all_walls = g.dim == 2 ? (!g.periodic[0] && !g.periodic[2])
                                : (!g.periodic[0] && !g.periodic[1] && !g.periodic[2]);
    if (all_walls) --d;
    return d;
}

void field_axis(const saras_grid_desc &g, int axis, std::vector<double> &coord,
                std::vector<double> &xi_x, std::vector<double> &xixx, std::vector<double> &xix2) {
    const int p = is_planar_y(g, axis) ? 1 : g.pad;
    std::vector<double> xi;
    level0(g, axis, p, false, xi, coord, xi_x, xixx, xix2);
}

bool build_levels(const saras_grid_desc &g, GridLevels &out) {
    if (g.dim != 2 && g.dim != 3) return false;
    if (g.pad < 1 || g.pad > 2) return false;
    const int depth = vc_depth(g);
    if (depth < 0) return false;
    out.dim = g.dim;
    out.depth = depth;
    out.axes.assign(depth + 1, std::vector<AxisLevel>(3));
    for (int a = 0; a < 3; ++a) {
        const bool py = is_planar_y(g, a);
        const int p = py ? 1 : g.pad;
        const int n0 = py ? 1 : g.n[a];
        const double L = g.len[a];
        std::vector<double> uxi, ux, uf, uff, uf2;
        level0(g, a, p, true, uxi, ux, uf, uff, uf2);
        // level 0, [-1, n]
        AxisLevel cur;
        cur.n = n0;
        cur.xi.assign(uxi.begin() + (p - 1), uxi.begin() + (p + n0 + 1));
        cur.coord.assign(ux.begin() + (p - 1), ux.begin() + (p + n0 + 1));
        for (int l = 0; l <= depth; ++l) {
            if (l > 0 && !py) {
                // mgGridMetrics(): xi and the uniform coordinate on level l
                const int cn = cur.n / 2;
                AxisLevel nx;
                nx.n = cn;
                nx.xi.assign(cn + 2, 0.0);
                nx.coord.assign(cn + 2, 0.0);
                for (int i = 0; i < cn; ++i) {
                    nx.xi[i + 1] = (cur.xi[2 * i + 1] + cur.xi[2 * i + 2]) / 2.0;
                    nx.coord[i + 1] = L * nx.xi[i + 1];
                }
                nx.xi[0] = -nx.xi[1];
                nx.coord[0] = -nx.coord[1];
                nx.xi[cn + 1] = 1.0 + nx.xi[1];
                nx.coord[cn + 1] = L + nx.coord[1];
                cur = nx;
            }
            AxisLevel lev = cur;
            const int m = lev.n + 2;
            lev.xi_x.assign(m, 1.0 / L);
            lev.xixx.assign(m, 0.0);
            lev.xix2.assign(m, std::pow(1.0 / L, 2.0));
            if (!py && g.gtype[a] != SARAS_GRID_UNIFORM) {
                if (l == 0) {
                    std::vector<double> sxi, sx, sf, sff, sf2;
                    level0(g, a, p, false, sxi, sx, sf, sff, sf2);
                    for (int i = -1; i <= n0; ++i) {
                        lev.coord[i + 1] = sx[i + p];
                        lev.xi_x[i + 1] = sf[i + p];
                        lev.xixx[i + 1] = sff[i + p];
                        lev.xix2[i + 1] = sf2[i + p];
                    }
                } else {
                    const bool two = g.gtype[a] == SARAS_GRID_TANH_DOUBLE;
                    const int nl = lev.n;
                    for (int i = 0; i < nl; ++i)
                        tanh_map(two, L, g.beta[a], lev.xi[i + 1], lev.coord[i + 1], lev.xi_x[i + 1],
                                 lev.xixx[i + 1], lev.xix2[i + 1]);
                    lev.coord[0] = -lev.coord[1];
                    lev.xi_x[0] = lev.xi_x[nl]; lev.xixx[0] = lev.xixx[nl]; lev.xix2[0] = lev.xix2[nl];
                    lev.coord[nl + 1] = L + lev.coord[1];
                    lev.xi_x[nl + 1] = lev.xi_x[1]; lev.xixx[nl + 1] = lev.xixx[1];
                    lev.xix2[nl + 1] = lev.xix2[1];
                }
            }
            out.axes[l][a] = lev;
        }
    }
    return true;
}

}  // namespace saras

extern "C" {

// timestep::timestep (timestep.cc:70-98): nu = 1/Re for hydrodynamics
// (probType <= 4); for the scalar problems (probType 5..9) the pair picked by
// rbcType.  buoyantForce::buoyantForce (buoyantForce.cc:48-60): Fb by
// rbcType, gTerms = Fb g / |g|.
saras_status saras_nse_physics(int prob_type, double re, double ra, double pr, int rbc_type,
                               const double g_axis[3], double *nu, double *kappa, double buoy[3]) {
    if (!nu || !kappa || !buoy) return SARAS_BAD_ARG;
    if (prob_type < 1 || prob_type > 9) return SARAS_BAD_ARG;
    if (prob_type <= 4) {
        if (!(re > 0.0)) return SARAS_BAD_ARG;
        *nu = 1.0 / re;
        *kappa = 0.0;     // no scalar field: the reference leaves kappa unset
        buoy[0] = buoy[1] = buoy[2] = 0.0;
        return SARAS_OK;
    }
    double Fb = 0.0;
    switch (rbc_type) {
        case 1:
            *nu = pr;
            *kappa = 1.0;
            Fb = ra * pr;
            break;
        case 2:
            *nu = std::sqrt(pr / ra);
            *kappa = 1.0 / std::sqrt(pr * ra);
            Fb = 1.0;
            break;
        case 3:
            *nu = 1.0;
            *kappa = 1.0 / pr;
            Fb = ra;
            break;
        case 4:
            *nu = std::sqrt(pr / ra);
            *kappa = 1.0 / std::sqrt(pr * ra);
            Fb = pr;
            break;
        default:
            return SARAS_BAD_ARG;   // the reference prints an error and exits
    }
    if (!g_axis) return SARAS_BAD_ARG;
    const double vnorm = std::sqrt(g_axis[0] * g_axis[0] + g_axis[1] * g_axis[1] + g_axis[2] * g_axis[2]);
    if (!(vnorm > 0.0)) return SARAS_BAD_ARG;
    for (int a = 0; a < 3; ++a) buoy[a] = Fb * g_axis[a] / vnorm;
    return SARAS_OK;
}

// The drivers' wall wiring without "Use Custom BC".  Walls: 0 Left (x-lo),
// 1 Right, 2 Front (y-lo), 3 Back, 4 Bottom (z-lo), 5 Top.
saras_status saras_nse_wiring(int prob_type, int planar, int non_hg, saras_nse_params *params) {
    if (!params || prob_type < 1 || prob_type > 9) return SARAS_BAD_ARG;
    const bool scalar = prob_type >= 5;
    // PLANAR builds print a warning for the heating plate and wire no bottom
    // wall for T at all (scalar.cc:178-185): no run to mirror
    if (scalar && prob_type != 6 && prob_type != 7 && non_hg && planar) return SARAS_BAD_ARG;
    saras_nse_params &p = *params;
    auto set = [](int *type, double *val, int w, int t, double v) {
        type[w] = t;
        val[w] = v;
    };
    const int D = SARAS_FBC_DIRICHLET, N = SARAS_FBC_NEUMANN;
    // hydro::initVBCs (hydro.cc:349-434): no-slip / no-penetration by default
    for (int c = 0; c < 3; ++c)
        for (int w = 0; w < 6; ++w) set(p.vbc_type[c], p.vbc_val[c], w, D, 0.0);
    int *ut = p.vbc_type[0], *vt = p.vbc_type[1], *wt = p.vbc_type[2];
    double *uv = p.vbc_val[0], *vv = p.vbc_val[1], *wv = p.vbc_val[2];
    if (prob_type == 3) {   // inflow on the left, outflow on the right
        set(ut, uv, 0, D, 1.0);
        set(ut, uv, 1, N, 0.0);
        set(vt, vv, 1, N, 0.0);
        set(wt, wv, 1, N, 0.0);
    }
    if (prob_type == 1) set(ut, uv, 5, D, 1.0);   // lid-driven cavity
    if (prob_type == 8)                            // free-slip bottom and top
        for (int w = 4; w < 6; ++w) {
            set(ut, uv, w, N, 0.0);
            set(vt, vv, w, N, 0.0);
        }
    if (prob_type == 9) {   // plume: open side walls and top for Vz
        set(wt, wv, 0, N, 0.0);
        set(wt, wv, 1, N, 0.0);
        set(wt, wv, 5, N, 0.0);
    }
    // hydro::initPBCs (hydro.cc:447-484)
    for (int w = 0; w < 6; ++w) set(p.pbc_type, p.pbc_val, w, N, 0.0);
    if (prob_type == 3) set(p.pbc_type, p.pbc_val, 0, SARAS_FBC_NULL, 0.0);
    // scalar::initTBCs (scalar.cc:163-201); hydrodynamics runs carry no T and
    // get zero-gradient walls.  probType 7 wires no bottom or top wall (its
    // runs need a periodic z axis); those walls are zero-gradient here.
    for (int w = 0; w < 6; ++w) set(p.tbc_type, p.tbc_val, w, N, 0.0);
    if (prob_type == 7) {
        set(p.tbc_type, p.tbc_val, 0, D, 1.0);
        set(p.tbc_type, p.tbc_val, 1, D, 0.0);
    }
    if (prob_type == 5 || prob_type == 8 || prob_type == 9) {
        if (non_hg) set(p.tbc_type, p.tbc_val, 4, SARAS_FBC_HOT_PLATE, 0.0);
        else set(p.tbc_type, p.tbc_val, 4, D, 1.0);
        if (prob_type == 9) set(p.tbc_type, p.tbc_val, 5, SARAS_FBC_PLUME_OUT, 0.0);
        else set(p.tbc_type, p.tbc_val, 5, D, 0.0);
    } else if (prob_type == 6) {
        set(p.tbc_type, p.tbc_val, 4, D, 0.0);
        set(p.tbc_type, p.tbc_val, 5, D, 1.0);
    }
    return SARAS_OK;
}

}  // extern "C"
