// Device-side CG divergence cleaning (ieq_cg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/inductioneq.h"

namespace ieq {

struct Params;
struct CgWork;

// Runs the CG loop (r = div B, iterations, final residual) with device-side scalars.
// Returns 0 on success, 1 when the configuration is not covered (caller falls back to
// the reference kernel sequence), negative on error.  Does not apply the projection.
template <typename R>
int cg_run_t(const Params& p, const ieq_program_info& info, const double* host_tab, int nbl, int wl, hipStream_t s,
             CgWork** work, R* b, R* phi, double abs_thr, double rel_tol, int max_it, int chunk, ieq_cg_result* res,
             std::string& err);

void cg_release(CgWork* w);

}  // namespace ieq
