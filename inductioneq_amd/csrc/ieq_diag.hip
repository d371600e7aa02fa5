// FP64 diagnostics of a vector field for either REAL type: the SBP-norm energy (norm2,
// kernel_operator.cl:11-44), the norm of the SBP divergence (calc_div + norm2_S,
// kernel_operator.cl:98-155) and the error against the testcase's analytic field
// (analytical_b + norm2_diff, kernel_init.cl:36-48, kernel_operator.cl:50-92), evaluated in
// one pass with every product and sum in double.  A fixed number of blocks strides over the
// nodes; each block writes one partial triple and a second one-block kernel adds the
// partials in a fixed order, so the result is deterministic and independent of the caller.
#include <mutex>

#include "ieq_device.hpp"
#include "ieq_diag.hpp"

namespace ieq {
namespace {

// Partial sums of ieq_diagnostics_noalloc: kDiagBlocks triples (24 KiB), one copy per GPU, held
// until the process ends.  Host code never passes this name to a kernel or a copy: it uses the
// device address that diag_scratch() looked up for the current device.
__device__ __align__(256) double scratch[3 * kDiagBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

template <int OPS>
__device__ inline double ddiv_dir(const Params& p, const double* hd, int d, uint32_t ix, uint32_t iy, uint32_t iz,
                                  const float* f32, const double* f64) {
  using O = Op<OPS>;
  constexpr int C = (O::W - 1) / 2;
  const uint32_t n3[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz}, i3[3] = {ix, iy, iz};
  const int row = bound_row(i3[d], n3[d], O::NB, p.periodic);
  double acc = 0.0;
  for (int k = 0; k < O::W; ++k) {
    const long long j = at_dir(p, ix, iy, iz, d, k - C) + d * p.stride;
    acc += O::d(row, k) * (f64 ? f64[j] : (double)f32[j]);
  }
  return acc / hd[d];
}

template <typename R, int OPS>
__global__ __launch_bounds__(kDiagThreads) void k_diag(Params p, const R* b, double t, int analytic, double* part) {
  using O = Op<OPS>;
  const long long n = (long long)p.nx * p.ny * p.nz, S = p.stride, nxy = (long long)p.nx * p.ny;
  const double hd[3] = {p.dx, p.dy, p.dz};
  const float* f32 = sizeof(R) == 4 ? (const float*)b : nullptr;
  const double* f64 = sizeof(R) == 8 ? (const double*)b : nullptr;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const uint32_t iz = (uint32_t)(i / nxy), iy = (uint32_t)((i - iz * nxy) / p.nx), ix = (uint32_t)(i - iy * p.nx - iz * nxy);
    const double fac = (p.dx / O::minv(bound_row(ix, p.nx, O::NB, p.periodic))) *
                       (p.dy / O::minv(bound_row(iy, p.ny, O::NB, p.periodic))) *
                       (p.dz / O::minv(bound_row(iz, p.nz, O::NB, p.periodic)));
    const double bx = (double)b[i], by = (double)b[S + i], bz = (double)b[2 * S + i];
    acc[0] += fac * (bx * bx + by * by + bz * bz);
    double dv = 0.0;
    for (int d = 0; d < 3; ++d) dv += ddiv_dir<OPS>(p, hd, d, ix, iy, iz, f32, f64);
    acc[1] += fac * (dv * dv);
    if (analytic) {
      const V3<double> a = b_analytical<double>(p, ix, iy, iz, t);
      const double ex = bx - a.x, ey = by - a.y, ez = bz - a.z;
      acc[2] += fac * (ex * ex + ey * ey + ez * ez);
    }
  }
  __shared__ double sm[3][kDiagThreads];
  for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

// out = (energy, divergence norm, error norm) from nblk partial triples, summed in block order;
// the error norm is NaN without an analytic field.
__global__ __launch_bounds__(64) void k_diag_sum(const double* part, int nblk, int analytic, double* out) {
  if (threadIdx.x >= 3) return;
  double s = 0.0;
  for (int k = 0; k < nblk; ++k) s += part[k * 3 + threadIdx.x];
  out[threadIdx.x] = threadIdx.x == 0 ? s : (threadIdx.x == 2 && !analytic) ? __longlong_as_double(0x7ff8000000000000LL)
                                                                            : sqrt(fmax(s, 0.0));
}

}  // namespace

int diag_blocks(const Params& p) {
  const long long n = (long long)p.nx * p.ny * p.nz;
  const long long need = (n + kDiagThreads - 1) / kDiagThreads;
  return (int)(need < kDiagBlocks ? need : kDiagBlocks);
}

template <typename R>
hipError_t diag_launch(const Params& p, int ops, hipStream_t s, const R* b, double t, int analytic, double* part,
                       double* out_dev) {
  const int nblk = diag_blocks(p);
  switch (ops) {
    case 0: k_diag<R, 0><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
    case 1: k_diag<R, 1><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
    case 2: k_diag<R, 2><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
    case 3: k_diag<R, 3><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
    case 4: k_diag<R, 4><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
    default: k_diag<R, 5><<<nblk, kDiagThreads, 0, s>>>(p, b, t, analytic, part); break;
  }
  k_diag_sum<<<1, 64, 0, s>>>(part, nblk, analytic, out_dev);
  return hipGetLastError();
}

hipError_t diag_scratch(double** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (double*)p;
  }
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

template hipError_t diag_launch<double>(const Params&, int, hipStream_t, const double*, double, int, double*, double*);
template hipError_t diag_launch<float>(const Params&, int, hipStream_t, const float*, double, int, double*, double*);

}  // namespace ieq
