// grid_host.hpp -- host-side grid and multigrid-level metrics.
//
// Restates lib/grid/grid.cc for one domain: the transformed uniform xi grid
// (createXiEtaZeta :318-340), uniform or tanh-stretched physical grid
// (createUniformGrid :358-391, createTanHypGrid :411-508), the per-level
// metric arrays of the multigrid hierarchy (mgGridMetrics :521-666) and the
// V-cycle depth (setVCDepth :678-742).  These are O(n) 1-D arrays computed
// once on the host and uploaded; no kernel is involved.
#pragma once
#include <vector>

#include "saras_hip.h"

namespace saras {

struct AxisLevel {
    int n = 0;                       // cells on this level
    std::vector<double> coord;       // [-1, n] -> index + 1
    std::vector<double> xi, xi_x, xixx, xix2;
};

struct GridLevels {
    int dim = 3;
    int depth = -1;                  // vcdGlo
    // axes[level][axis]; planar problems leave axis 1 with n = 1
    std::vector<std::vector<AxisLevel>> axes;
};

// Returns vcdGlo, or a negative value if the grid cannot be coarsened.
int vc_depth(const saras_grid_desc &g);

// Builds every level's metric arrays.  Returns false on bad input.
bool build_levels(const saras_grid_desc &g, GridLevels &out);

// Field-grid arrays of one axis over [-pad, n+pad-1] (grid::x, xi_x, xixx,
// xix2 used by the field stencils).
void field_axis(const saras_grid_desc &g, int axis, std::vector<double> &coord,
                std::vector<double> &xi_x, std::vector<double> &xixx, std::vector<double> &xix2);

}  // namespace saras
