// mg_internal.hpp -- hooks other translation units of the library use to run
// the multigrid solver on arrays that are already resident on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "saras_hip.h"

namespace saras {

// Level-0 iterate and RHS of the FP64 hierarchy (core-only, row pitch sy,
// plane pitch sx; rows colour-split when hz > 0: even k at k/2, odd k at
// hz + k/2).  The iterate pointer changes across solves (out-of-place sweeps
// swap buffers): fetch it again after every solve.
void mg_level0(saras_mg *m, double **x, double **b, long *sx, long *sy, long *hz = nullptr);

// poisson::mgSolve on the resident level-0 arrays: the iterate is the
// initial guess, the RHS must be in place.  Runs on stream s.
saras_status mg_solve_resident(saras_mg *m, hipStream_t s, saras_mg_stats *st);

hipStream_t mg_stream(saras_mg *m);

// Serialises device allocation, deallocation and hipGraph capture across
// host threads: a free (which synchronises the device) or a null-stream
// memset issued while another thread captures a stream would invalidate the
// capture.
std::recursive_mutex &alloc_mutex();

// Text returned by saras_last_error() on the calling thread.
void set_last_error(const std::string &msg);

}  // namespace saras
