// Tiled z-marching stage kernel for programs without the Hall term: the reference-schedule
// stage (out = cP*P + cQ*Q + cD*DB(Y), kernel_time_integrator.cl:9-31 with the SBP boundary
// rows, the SAT inflow terms, a velocity field and high-order or first-order dissipation) for
// bounded and periodic grids.  Drop-in for the one-thread-per-node stage kernel of
// ieq_generic.hip whenever the NDRange covers the grid.
#pragma once

#include <hip/hip_runtime.h>

namespace ieq {

struct Params;

// Whether the tiled stage kernel covers a program with these settings (REAL of `real_bytes`
// bytes): no Hall term, dissipation none / high order / first order, every dimension at least
// the stencil width, 32-bit byte offsets for three component planes.
bool tiled_stage_ok(const Params& p, int ops, int real_bytes);

// Enqueue one stage on the whole grid: out = cP*P + cQ*Q + cD*DB(Y) (P or Q null: absent), the
// velocity field u, the stage time read from time[1] on the device, as the reference stage does.
template <typename R>
hipError_t tiled_stage(const Params& p, int ops, hipStream_t s, R* out, const R* P, R cP, const R* Q, R cQ, const R* Y, R cD,
                       const R* u, const R* time);

}  // namespace ieq
