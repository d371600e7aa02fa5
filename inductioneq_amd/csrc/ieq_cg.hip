// Conjugate-gradient divergence cleaning with device-side alpha/beta.
//
// Restates the host loop of matlab/+divergence_cleaning/clean_div_CG.m:31-86
// (calc_div_divcleaning -> r; repeat { u = beta u + r; q = -Lap u; alpha = rr/(u.q);
// phi += alpha u; r -= alpha q; rr = r.r } while rr > thr^2 and it < max; projection)
// as two kernels per iteration with all scalars kept on the device:
//   k_cg_lap     u' = beta*u + r (beta = rr/old, old = 1 at start), q = -Lap u' on the
//                BNODES interior, partial u'.q  (u' goes to the other direction register)
//   k_cg_update  phi += alpha u', r -= alpha q, partial r.r
// Dot products accumulate in FP64 whatever the field precision: every block writes
// its partial and a one-workgroup kernel sums the partials in index order, so the
// result is deterministic.  Once the stop test holds, every kernel
// returns immediately, so any number of iterations can be enqueued without a host
// round trip and the result equals stopping exactly at convergence.
#include <cmath>
#include <cstring>

#include "ieq_cg.hpp"
#include "ieq_device.hpp"
#include "ieq_generic.hpp"
#include "ieq_stencil.hpp"

namespace ieq {

struct CgState {
  double rr, old, alpha, rr0;
  double tol2;       // stop when rr <= tol2 ...
  double rel2;       // ... or rr <= rel2 * rr0 (rel2 = 0: absolute test only)
  int it, max_it, done, pad;
};

namespace {

constexpr int kBlk = 256;
constexpr int kLapRows = 32;  // table rows held in LDS by k_cg_lap (2 * NUM_BOUNDS + 1 of the Laplacian)

// Deterministic block reduction of one double (wave64 shuffles, then LDS).
__device__ inline double block_sum(double v, double* smem) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (wid == 0) {
    t = lane < nw ? smem[lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  }
  __syncthreads();
  return t;  // valid in wave 0
}

__device__ inline double sum_partials(const double* part, unsigned int n, double* smem) {
  double v = 0.0;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
  for (unsigned int i = tid; i < n; i += nt) v += part[i];
  return block_sum(v, smem);
}

struct LapArgs {
  int nx, ny, nz, off;  // grid and BNODES
  int zc;               // planes per block
  int nbl, wl, hw;      // table rows per side, width, half-width
  const double* tab;    // (2nbl+1) x wl, device copy
  double idx2, idy2, idz2;
  unsigned int nblocks;
};

// Fused direction update and Laplacian on [off, n-off)^3:
//   u' = beta u + r (beta = rr/old),  q = -Lap u',  partial u'.q
// u' is written to a second buffer (neighbouring blocks still read u), so the direction
// registers alternate between iterations.  Outside the interior u and r are 0 (never
// written), so u' is 0 there as in the reference, where calc_axpy runs on the interior
// only; neighbours outside the grid read 0 (their coefficients are 0).  Planes are
// double-buffered in LDS and the loads of plane z+1 are in flight while plane z is computed.
template <typename R, int HW, int TY>
__global__ __launch_bounds__(64 * TY) void k_cg_lap(LapArgs a, const R* __restrict__ u, const R* __restrict__ r,
                                                    R* __restrict__ un, R* __restrict__ q, CgState* st, double* part) {
  constexpr int LX = 64 + 2 * HW, LY = TY + 2 * HW, Q = 2 * HW + 1;
  __shared__ R s[2][LY][LX];
  __shared__ double red[16];
  if (st->done) return;
  const R beta = (R)(st->rr / st->old);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = a.nx - 2 * a.off, gy = a.ny - 2 * a.off, gz = a.nz - 2 * a.off;
  const int x = a.off + blockIdx.x * 64 + tx, y = a.off + blockIdx.y * TY + ty;
  const int x0 = a.off + blockIdx.x * 64, y0 = a.off + blockIdx.y * TY;
  const bool active = (x - a.off) < gx && (y - a.off) < gy;
  // u' at a node (0 outside the grid)
  auto dirv = [&](int xx, int yy, int zz) -> R {
    if ((unsigned)xx < (unsigned)a.nx && (unsigned)yy < (unsigned)a.ny && (unsigned)zz < (unsigned)a.nz) {
      const unsigned o = (unsigned)(xx + a.nx * (yy + a.ny * zz)) * (unsigned)sizeof(R);
      return beta * ldo(u, o) + ldo(r, o);
    }
    return R(0);
  };
  // per-thread x/y rows
  R cx[Q], cy[Q];
  {
    const int rx = bound_row((uint32_t)x, (uint32_t)a.nx, a.nbl, 0), ry = bound_row((uint32_t)y, (uint32_t)a.ny, a.nbl, 0);
    const int c0 = (a.wl - 1) / 2;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int j = c0 - HW + k;
      cx[k] = (j >= 0 && j < a.wl) ? (R)(a.tab[rx * a.wl + j] * a.idx2) : R(0);
      cy[k] = (j >= 0 && j < a.wl) ? (R)(a.tab[ry * a.wl + j] * a.idy2) : R(0);
    }
  }
  // z rows of the table, scaled by 1/dz^2 and rounded to REAL once (uniform per plane)
  __shared__ R sz[kLapRows][Q];
  {
    const int rows = 2 * a.nbl + 1, c0 = (a.wl - 1) / 2;
    for (int i = ty * 64 + tx; i < rows * Q; i += 64 * TY) {
      const int row = i / Q, k = i % Q, j = c0 - HW + k;
      sz[row][k] = (j >= 0 && j < a.wl) ? (R)(a.tab[row * a.wl + j] * a.idz2) : R(0);
    }
  }
  double dot = 0.0;
  R qz[Q + 1];
  R h[4];  // halo values of the plane being staged: x left, x right, y below, y above
  const int z0 = a.off + blockIdx.z * a.zc;
  const int zn = min(a.zc, gz - (int)blockIdx.z * a.zc);
  auto load_halo = [&](int z) {
    if (tx < HW) h[0] = dirv(x0 - HW + tx, y, z);
    if (tx >= 64 - HW) h[1] = dirv(x + HW, y, z);
    if (ty < HW) h[2] = dirv(x, y0 - HW + ty, z);
    if (ty >= TY - HW) h[3] = dirv(x, y + HW, z);
  };
#pragma unroll
  for (int j = 0; j < Q; ++j) qz[j] = dirv(x, y, z0 - HW + j);
  load_halo(z0);
  for (int zi = 0; zi < zn; ++zi) {
    const int z = z0 + zi;
    auto& sp = s[zi & 1];
    sp[ty + HW][tx + HW] = qz[HW];
    if (tx < HW) sp[ty + HW][tx] = h[0];
    if (tx >= 64 - HW) sp[ty + HW][tx + 2 * HW] = h[1];
    if (ty < HW) sp[ty][tx + HW] = h[2];
    if (ty >= TY - HW) sp[ty + 2 * HW][tx + HW] = h[3];
    // loads of the next plane (halo first), in flight across the barrier and the compute
    if (zi + 1 < zn) load_halo(z + 1);
    qz[Q] = dirv(x, y, z + HW + 1);
    __syncthreads();
    if (active) {
      const int rz = bound_row((uint32_t)z, (uint32_t)a.nz, a.nbl, 0);
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < Q; ++k) acc = acc + cx[k] * sp[ty + HW][tx + k] + cy[k] * sp[ty + k][tx + HW] + sz[rz][k] * qz[k];
      const unsigned o = (unsigned)(x + a.nx * (y + a.ny * z)) * (unsigned)sizeof(R);
      sto(un, o, qz[HW]);
      sto(q, o, acc);
      dot += (double)qz[HW] * (double)acc;
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) qz[j] = qz[j + 1];
  }
  const double bs = block_sum(dot, red);
  const unsigned int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (tx == 0 && ty == 0) part[bid] = bs;
}

// One workgroup: alpha = rr / u'.q from the block partials (summed in block order).  A
// separate launch instead of a last-block finish: a device-scope fence in every block
// writes back and invalidates the L2 of its XCD.
__global__ __launch_bounds__(256) void k_cg_alpha(const double* part, unsigned int n, CgState* st) {
  __shared__ double red[16];
  if (st->done) return;
  const double uq = sum_partials(part, n, red);
  if (threadIdx.x == 0) st->alpha = st->rr / uq;
}

// One workgroup: rr = r.r, the stop test and the iteration count.
__global__ __launch_bounds__(256) void k_cg_rr(const double* part, unsigned int n, CgState* st) {
  __shared__ double red[16];
  if (st->done) return;
  const double rr = sum_partials(part, n, red);
  if (threadIdx.x == 0) {
    st->old = st->rr;
    st->rr = rr;
    st->it += 1;
    const bool go = rr > st->tol2 && (st->rel2 <= 0.0 || rr > st->rel2 * st->rr0) && st->it < st->max_it;
    st->done = go ? 0 : 1;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlk) void k_cg_update(R* __restrict__ phi, const R* __restrict__ u, R* __restrict__ r,
                                                    const R* __restrict__ q, long long n, CgState* st, double* part) {
  __shared__ double red[16];
  if (st->done) return;
  const R alpha = (R)st->alpha;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    phi[i] = phi[i] + alpha * u[i];
    const R rn = r[i] + (-alpha) * q[i];
    r[i] = rn;
    acc += (double)rn * (double)rn;
  }
  const double bs = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// r.r partials of the initial residual.
template <typename R>
__global__ __launch_bounds__(kBlk) void k_cg_init(const R* __restrict__ r, long long n, double* part) {
  __shared__ double red[16];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += (double)r[i] * (double)r[i];
  const double bs = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// State initialisation from the partials of k_cg_init.
__global__ __launch_bounds__(256) void k_cg_init_state(const double* part, unsigned int n, CgState* st, double tol2,
                                                       double rel2, int max_it) {
  __shared__ double red[16];
  const double rr = sum_partials(part, n, red);
  if (threadIdx.x == 0) {
    st->rr = rr;
    st->rr0 = rr;
    st->old = 1.0;  // clean_div_CG.m:43
    st->alpha = 0.0;
    st->it = 0;
    st->max_it = max_it;
    st->tol2 = tol2;
    st->rel2 = rel2;
    st->done = (rr > tol2 && (rel2 <= 0.0 || rr > rel2 * rr) && max_it > 0) ? 0 : 1;
  }
}

template <typename R, int HW, int TY>
void launch_lap(const LapArgs& a, const R* u, const R* r, R* un, R* q, CgState* st, double* part, hipStream_t s) {
  const int gx = a.nx - 2 * a.off, gy = a.ny - 2 * a.off, gz = a.nz - 2 * a.off;
  dim3 grid((gx + 63) / 64, (gy + TY - 1) / TY, (gz + a.zc - 1) / a.zc);
  k_cg_lap<R, HW, TY><<<grid, dim3(64, TY), 0, s>>>(a, u, r, un, q, st, part);
}

}  // namespace

struct CgWork {
  void* buf = nullptr;  // r, u, q (num_nodes each) + partials + state + table
  size_t bytes = 0;
};

void cg_release(CgWork* w) {
  if (!w) return;
  if (w->buf) (void)hipFree(w->buf);
  delete w;
}

template <typename R>
int cg_run_t(const Params& p, const ieq_program_info& info, const double* host_tab, int nbl, int wl, hipStream_t s, CgWork** work,
             R* b, R* phi, double abs_thr, double rel_tol, int max_it, int chunk, ieq_cg_result* res, std::string& err) {
  const long long n = info.num_nodes;
  const int hw = (wl - 1) / 2;
  if (hw < 1 || hw > 7 || info.periodic || 2 * nbl + 1 > kLapRows || !offsets_fit(p.nx, p.ny, p.nz, sizeof(R)))
    return 1;  // not covered by the fused loop
  const int TY = 8;
  const int gx = p.nx - 2 * p.bnodes, gy = p.ny - 2 * p.bnodes, gz = p.nz - 2 * p.bnodes;
  const int zc = gz >= 128 ? 16 : 8;
  const unsigned int lap_blocks = (unsigned)(((gx + 63) / 64) * ((gy + TY - 1) / TY) * ((gz + zc - 1) / zc));
  const unsigned int vec_blocks = 1024;
  const size_t fbytes = ((size_t)info.stride * sizeof(R) + 255) / 256 * 256;
  const size_t part_bytes = (size_t)std::max(lap_blocks, vec_blocks) * sizeof(double);
  const size_t tab_bytes = (size_t)(2 * nbl + 1) * wl * sizeof(double);
  const size_t need = 4 * fbytes + ((part_bytes + 255) / 256 * 256) + 256 + tab_bytes;
  if (!*work) *work = new CgWork();
  CgWork* w = *work;
  if (w->bytes < need) {
    if (w->buf) (void)hipFree(w->buf);
    w->buf = nullptr;
    w->bytes = 0;
    if (hipMalloc(&w->buf, need) != hipSuccess) return err = "hipMalloc failed", -4;
    w->bytes = need;
  }
  char* base = (char*)w->buf;
  R* r = (R*)base;
  R* ub[2] = {(R*)(base + fbytes), (R*)(base + 3 * fbytes)};  // direction registers, alternating
  R* q = (R*)(base + 2 * fbytes);
  double* part = (double*)(base + 4 * fbytes);
  CgState* st = (CgState*)(base + 4 * fbytes + (part_bytes + 255) / 256 * 256);
  double* tab = (double*)((char*)st + 256);
  if (hipMemsetAsync(base, 0, 4 * fbytes, s) != hipSuccess || hipMemsetAsync(phi, 0, fbytes, s) != hipSuccess ||
      hipMemsetAsync(st, 0, sizeof(CgState), s) != hipSuccess ||
      hipMemcpyAsync(tab, host_tab, tab_bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return err = "CG set-up failed", -4;

  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, s);
  // r = div B on the interior (calc_div_divcleaning, kernel_div_cleaning.cl:60-70)
  const uint32_t g_dc[3] = {(uint32_t)(p.nx - 2 * p.bnodes), (uint32_t)(p.ny - 2 * p.bnodes), (uint32_t)(p.nz - 2 * p.bnodes)};
  gen_div<R>(p, info.opset, s, g_dc, b, r, p.bnodes);
  k_cg_init<R><<<vec_blocks, kBlk, 0, s>>>(r, n, part);
  k_cg_init_state<<<1, 256, 0, s>>>(part, vec_blocks, st, abs_thr * abs_thr, rel_tol * rel_tol, max_it);

  LapArgs la;
  la.nx = p.nx;
  la.ny = p.ny;
  la.nz = p.nz;
  la.off = p.bnodes;
  la.nbl = nbl;
  la.wl = wl;
  la.hw = hw;
  la.tab = tab;
  la.idx2 = 1.0 / (p.dx * p.dx);
  la.idy2 = 1.0 / (p.dy * p.dy);
  la.idz2 = 1.0 / (p.dz * p.dz);
  la.nblocks = lap_blocks;
  la.zc = zc;
  CgState hst;
  int launched = 0;
  while (launched < max_it) {
    const int k = std::min(chunk, max_it - launched);
    for (int i = 0; i < k; ++i) {
      const int it = launched + i;
      const R* u = ub[it & 1];
      R* un = ub[(it + 1) & 1];
      switch (hw) {
        case 1: launch_lap<R, 1, 8>(la, u, r, un, q, st, part, s); break;
        case 2: launch_lap<R, 2, 8>(la, u, r, un, q, st, part, s); break;
        case 3: launch_lap<R, 3, 8>(la, u, r, un, q, st, part, s); break;
        case 4: launch_lap<R, 4, 8>(la, u, r, un, q, st, part, s); break;
        case 5: launch_lap<R, 5, 8>(la, u, r, un, q, st, part, s); break;
        case 6: launch_lap<R, 6, 8>(la, u, r, un, q, st, part, s); break;
        default: launch_lap<R, 7, 8>(la, u, r, un, q, st, part, s); break;
      }
      k_cg_alpha<<<1, 256, 0, s>>>(part, lap_blocks, st);
      k_cg_update<R><<<vec_blocks, kBlk, 0, s>>>(phi, un, r, q, n, st, part);
      k_cg_rr<<<1, 256, 0, s>>>(part, vec_blocks, st);
    }
    launched += k;
    if (hipMemcpyAsync(&hst, st, sizeof(CgState), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return err = "CG state read failed", -4;
    if (hst.done) break;
  }
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (hipGetLastError() != hipSuccess) return err = "CG kernel launch failed", -4;
  if (hipMemcpy(&hst, st, sizeof(CgState), hipMemcpyDeviceToHost) != hipSuccess) return err = "CG state read failed", -4;
  res->iterations = hst.it;
  res->residual_norm = std::sqrt(hst.rr);
  res->divergence_norm_initial = std::sqrt(hst.rr0);
  res->seconds = ms * 1e-3;
  return 0;
}

template int cg_run_t<double>(const Params&, const ieq_program_info&, const double*, int, int, hipStream_t, CgWork**, double*,
                              double*, double, double, int, int, ieq_cg_result*, std::string&);
template int cg_run_t<float>(const Params&, const ieq_program_info&, const double*, int, int, hipStream_t, CgWork**, float*,
                             float*, double, double, int, int, ieq_cg_result*, std::string&);

}  // namespace ieq
