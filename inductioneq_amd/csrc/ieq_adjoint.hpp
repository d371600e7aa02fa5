// Discrete adjoint of the fused periodic SSPRK33 solve with binomial checkpointing.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/inductioneq.h"

namespace ieq {

struct Params;
struct Adjoint;

int adjoint_create(const Params& p, const ieq_program_info& info, long long nsteps, int snapshots, Adjoint** out,
                   std::string& err);
int adjoint_forward(Adjoint* A, hipStream_t s, const void* b0, void* bT, double c_hall, const double* u, double* seconds,
                    std::string& err);
int adjoint_backward(Adjoint* A, hipStream_t s, const void* lamT, void* grad_b0, double* grads, double* seconds,
                     std::string& err);
void adjoint_counts(const Adjoint* A, long long* fwd, long long* adj);
void adjoint_release(Adjoint* A);

}  // namespace ieq
