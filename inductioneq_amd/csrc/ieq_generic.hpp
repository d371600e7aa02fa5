// Host launchers of the reference-schedule kernels (ieq_generic.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ieq {

struct Params;

template <typename R>
void gen_stage(const Params& p, int ops, hipStream_t s, const uint32_t g[3], R* out, const R* P, R cP, const R* Q, R cQ,
               const R* Y, R cD, const R* J, const R* u, const R* time);
template <typename R>
void gen_update_2r(const Params& p, hipStream_t s, const uint32_t g[3], R* X, R* O, const R* D, R a, R b, R* time, double tc);
template <typename R>
void gen_update_2n(const Params& p, hipStream_t s, const uint32_t g[3], R* X, const R* D, R b, R* time, double tc);
template <typename R> void gen_ssprk104_06(const Params& p, hipStream_t s, const uint32_t g[3], const R* b1, R* b2, R* b3, R* time);
template <typename R> void gen_time(hipStream_t s, R* time, double c, double dt, int mode);
template <typename R> void gen_curl_rho(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* J, const R* rho);
template <typename R> void gen_calc_u(const Params& p, hipStream_t s, const uint32_t g[3], R* u, const R* time);
template <typename R> void gen_init(const Params& p, hipStream_t s, const uint32_t g[3], R* f, int which, const R* time);
template <typename R> void gen_div(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* out, int off);
template <typename R> void gen_curl(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* out);
template <typename R> void gen_grad(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* f, R* out);
template <typename R> void gen_lap(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* f, R* out);
template <typename R> void gen_proj(const Params& p, int ops, hipStream_t s, const uint32_t g[3], R* b, const R* phi);
template <typename R> void gen_axpy(const Params& p, hipStream_t s, const uint32_t g[3], R* x, const R* y, const R* a, int xpay);
template <typename R>
void gen_group_reduce(const Params& p, int ops, hipStream_t s, const R* a, const R* b, R* out, int which, long long wsize,
                      long long ngroups, long long nnodes);
template <typename R> void gen_aos_to_soa(hipStream_t s, const R* aos, R* soa, long long n, long long stride);
template <typename R> void gen_soa_to_aos(hipStream_t s, const R* soa, R* aos, long long n, long long stride);

}  // namespace ieq
