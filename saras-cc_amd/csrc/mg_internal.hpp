// mg_internal.hpp -- hooks other translation units of the library use to run
// the multigrid solver on arrays that are already resident on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
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

// Waits until the pinned host-coherent word *seq holds `want`, the ticket the
// last reduction kernel launched on s stores after its result (a system-scope
// release, so the result is visible once the ticket is).  Polling the word
// returns as soon as the result has landed, without waiting for the
// stream's completion signal.  The poll stops once the stream reports
// completion or an error, so a failed launch cannot leave it spinning: the
// stream is then synchronised, which returns the error.  SARAS_NORM_SYNC=1
// synchronises the stream instead of polling.
inline hipError_t wait_ticket(const unsigned *seq, unsigned want, hipStream_t s) {
    static const bool sync_only = [] {
        const char *env = std::getenv("SARAS_NORM_SYNC");
        return env && env[0] == '1';
    }();
    for (unsigned spin = 1; !sync_only; ++spin) {
        if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == want) return hipSuccess;
        if ((spin & 63) == 0 && hipStreamQuery(s) != hipErrorNotReady) break;
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    return __atomic_load_n(seq, __ATOMIC_ACQUIRE) == want ? hipSuccess : hipErrorUnknown;
}

}  // namespace saras
