// FP64 diagnostics (energy, divergence norm, error norm) of a vector field.
#pragma once

#include <hip/hip_runtime.h>

namespace ieq {

struct Params;

constexpr int kDiagBlocks = 1024, kDiagThreads = 256;

// Blocks of the partial pass for the program's grid (at most kDiagBlocks): the partial buffer
// holds 3 doubles per block.
int diag_blocks(const Params& p);

// Device address, on the current device, of the library's static partial buffer (3 * kDiagBlocks
// doubles) used by ieq_diagnostics_noalloc; looked up once per device.
hipError_t diag_scratch(double** ptr);

// Enqueue the partial pass into `part` (3 * diag_blocks(p) doubles) and the fixed-order sum into
// out_dev[0..2] = (energy, divergence norm, error norm); `analytic` selects the error term
// against b_analytical at time t.
template <typename R>
hipError_t diag_launch(const Params& p, int ops, hipStream_t s, const R* b, double t, int analytic, double* part,
                       double* out_dev);

}  // namespace ieq
