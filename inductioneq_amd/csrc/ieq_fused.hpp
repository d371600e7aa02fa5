// Fused time stepping for periodic grids with uniform velocity and density.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/inductioneq.h"

namespace ieq {

struct Params;
struct FusedWork;

// Kernel names of one time step of `integrator` (compute_numerical_solution_setup.m:140-371).
// Returns 0, or -1 for an unknown integrator.
int schedule(const char* integrator, int hall, int variable_u, std::vector<std::string>& names, int& nf);

struct FusedOpts {
  // (u_x, u_y, u_z, rho) of the uniform coefficient fields as stated by the caller; null: the
  // fields are checked on the device (one host synchronisation)
  const double* uniform = nullptr;
  // leave curlB_rho = curl B / rho of the last stage's input, as the reference schedule does
  // (otherwise it holds a work field)
  bool final_curl = false;
};

// Advance nsteps with the fused kernels when the program qualifies (periodic grid,
// uniform u and rho, no time-dependent velocity, no dissipation other than high
// order).  Sets *used = 0 and enqueues nothing when it does not qualify.
int fused_step(const Params& p, const ieq_program_info& info, hipStream_t s, FusedWork** work, const char* integrator,
               long long nsteps, int variable_u, void* const* args, const FusedOpts& opt, int* used, std::string& err);

void fused_release(FusedWork* w);

struct Ensemble;
int ensemble_create(const Params& p, const ieq_program_info& info, const ieq_member* members, int n, const char* integrator,
                    Ensemble** out, std::string& err);
int ensemble_step(Ensemble* e, long long nsteps, double* seconds, std::string& err);
void ensemble_config(const Ensemble* e, int* ty, int* zc, int* order);
void ensemble_release(Ensemble* e);

// Host copy of a -Laplacian table: lap 0 WS_D0, 1 NS_D0, 2 WS_LN; nullptr when absent.
const double* host_lap_table(int opset, int lap, int* nbl, int* wl);

}  // namespace ieq
