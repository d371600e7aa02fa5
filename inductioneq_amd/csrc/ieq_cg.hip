// Conjugate-gradient divergence cleaning with device-side alpha/beta.
//
// Restates the host loop of matlab/+divergence_cleaning/clean_div_CG.m:31-86
// (calc_div_divcleaning -> r; repeat { u = beta u + r; q = -Lap u; alpha = rr/(u.q);
// phi += alpha u; r -= alpha q; rr = r.r } while rr > thr^2 and it < max; projection)
// as two kernels per iteration with all scalars kept on the device:
//   k_cg_lap     rr = r.r (sum of the previous update's partials), the stop test,
//                u' = beta*u + r (beta = rr/old, old = 1 at start), q = -Lap u' on the
//                BNODES interior, partial u'.q  (u' goes to the other direction register)
//   k_cg_update  alpha = rr / u'.q (sum of the Laplacian's partials), phi += alpha u',
//                r -= alpha q, partial r.r
// Dot products accumulate in FP64 whatever the field precision: every block writes its
// partial, and every block of the next kernel sums all partials in the same fixed order, so
// all blocks agree on alpha (beta) bit for bit and the result is deterministic.  rr of
// iteration i is kept in slot i % 2 of the device state, so the block that records it never
// overwrites the value the other blocks of the same launch read.  Once the stop test holds,
// every kernel returns immediately, so any number of iterations can be enqueued without a
// host round trip and the result equals stopping exactly at convergence.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "ieq_cg.hpp"
#include "ieq_device.hpp"
#include "ieq_generic.hpp"
#include "ieq_stencil.hpp"

namespace ieq {

struct CgState {
  double rr, old, alpha, rr0;
  double tol2;       // stop when rr <= tol2 ...
  double rel2;       // ... or rr <= rel2 * rr0 (rel2 = 0: absolute test only)
  double rrs[2];     // r.r of iteration i in slot i % 2 (slot 1 starts at 1: old of iteration 0)
  int it, max_it, done, pad;
};

namespace {

constexpr int kBlk = 256;
constexpr int kLapRows = 32;  // table rows held in LDS by k_cg_lap (2 * NUM_BOUNDS + 1 of the Laplacian)

// Deterministic block reduction of one double (wave64 shuffles, then LDS).
__device__ inline double block_sum(double v, double* smem) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int tid = threadIdx.y * blockDim.x + threadIdx.x, lane = tid & 63, wid = tid >> 6;
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

// The same sum, returned to every thread of the block (smem: at least 17 doubles).
__device__ inline double sum_partials_all(const double* part, unsigned int n, double* smem) {
  const double t = sum_partials(part, n, smem);
  if (threadIdx.x == 0 && threadIdx.y == 0) smem[16] = t;
  __syncthreads();
  const double v = smem[16];
  __syncthreads();
  return v;
}

struct LapArgs {
  int nx, ny, nz, off;  // grid and BNODES
  int zc;               // planes per block
  int nbl, wl, hw;      // table rows per side, width, half-width
  const double* tab;    // (2nbl+1) x wl, device copy
  double idx2, idy2, idz2;
  unsigned int nblocks;
};

// Minimum waves per SIMD the register allocation of a 1024-thread Laplacian tile must allow.  A
// workgroup of 64 x 16 threads puts 4 waves on each SIMD, so it needs at least 4 (the
// unconstrained FP64 kernel at half-width 4 takes 133 VGPRs, 3 waves per SIMD, and would not
// launch).  With one plane of loads in flight the kernels up to half-width 5 fit without
// spilling (half-width 4: FP64 120 VGPRs, 4 waves; FP32 74 VGPRs, 6 waves).
#ifndef IEQ_LAP_WAVES64
#define IEQ_LAP_WAVES64 4
#endif
#ifndef IEQ_LAP_WAVES32
#define IEQ_LAP_WAVES32 6
#endif

// Fused direction update and Laplacian on [off, n-off)^3:
//   u' = beta u + r (beta = rr/old),  q = -Lap u',  partial u'.q
// u' is written to a second buffer (neighbouring blocks still read u), so the direction
// registers alternate between iterations.  Outside the interior u and r are 0 (never
// written), so u' is 0 there as in the reference, where calc_axpy runs on the interior
// only; neighbours outside the grid count as 0 (their coefficients are 0).
// z-march: u' of the own column's z-stencil lives in a register window, one xy-plane plus
// halo at a time in LDS (double-buffered).  Loads are issued PD planes ahead of their use:
// they always target a valid node (coordinates clamped into the grid) and land in raw
// (u, r) registers; the masking and beta u + r happen only when a plane is used, so no
// load is waited for before it is needed.  The per-plane barrier orders LDS traffic only.
// Tiles are handed out per XCD in contiguous ranges (x fastest, then y, then z-chunk), so
// the halo rows of y-neighbours are read through one L2.
template <typename R, int HW, int TY, int PD, int TX>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(TX * TY >= 1024 ? (sizeof(R) == 4 ? IEQ_LAP_WAVES32 : IEQ_LAP_WAVES64) : 1, 8))) void k_cg_lap(LapArgs a, const R* __restrict__ u, const R* __restrict__ r,
                                                    R* __restrict__ un, R* __restrict__ q, CgState* st,
                                                    const double* rr_part, unsigned int n_rr, double* part, int iter) {
  constexpr int LX = TX + 2 * HW, LY = TY + 2 * HW, Q = 2 * HW + 1;
  __shared__ R s[2][LY][LX];
  __shared__ double red[17];
  if (st->done) return;
  // rr of this iteration from the previous update's partials (the initial residual's for
  // iteration 0), the stop test, and beta = rr / old (clean_div_CG.m:45-48)
  const double rr = sum_partials_all(rr_part, n_rr, red);
  const double old = st->rrs[(iter + 1) & 1];
  const bool go = rr > st->tol2 && (st->rel2 <= 0.0 || rr > st->rel2 * st->rr0) && iter < st->max_it;
  if (threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x == 0) {
    st->rr = rr;
    st->old = old;
    st->it = iter;
    if (go) st->rrs[iter & 1] = rr;
    else st->done = 1;
  }
  if (!go) return;
  const R beta = (R)(rr / old);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int gx = a.nx - 2 * a.off, gy = a.ny - 2 * a.off, gz = a.nz - 2 * a.off;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ntx = (gx + TX - 1) / TX, nty = (gy + TY - 1) / TY;
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int x = a.off + bx * TX + tx, y = a.off + by * TY + ty;
  const int x0 = a.off + bx * TX, y0 = a.off + by * TY;
  const bool active = (x - a.off) < gx && (y - a.off) < gy;
  const unsigned RB = sizeof(R);
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;
  auto cl = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  auto inb = [](int v, int n) { return (unsigned)v < (unsigned)n; };
  // In-plane byte offsets (clamped) and in-grid masks of the column and its halo cells.
  // Every thread issues the same loads each plane: one x-halo cell (left or right; threads
  // off the x edges load their own column again), one y-halo cell per NYH (below / above,
  // or the own column), so the memory counter of a plane's loads is the same on every path.
  constexpr int NYH = 2 * HW <= TY ? 1 : 2;
  const unsigned oc = (unsigned)(cl(x, a.nx) + a.nx * cl(y, a.ny)) * RB;
  const bool colin = inb(x, a.nx) && inb(y, a.ny);
  const bool hl = tx < HW, hr = tx >= TX - HW, hb = ty < HW, ht = ty >= TY - HW;
  const int hxx = hl ? x0 - HW + tx : x + HW;  // x-halo cell (used when hl || hr)
  const int hyb = y0 - HW + ty, hya = y + HW;    // y-halo cells below / above
  const unsigned ohx = (hl || hr) ? (unsigned)(cl(hxx, a.nx) + a.nx * cl(y, a.ny)) * RB : oc;
  unsigned ohy[NYH];
  bool mhy[NYH];
  if constexpr (NYH == 1) {
    const int yy = hb ? hyb : hya;
    ohy[0] = (hb || ht) ? (unsigned)(cl(x, a.nx) + a.nx * cl(yy, a.ny)) * RB : oc;
    mhy[0] = inb(x, a.nx) && inb(yy, a.ny);
  } else {
    ohy[0] = hb ? (unsigned)(cl(x, a.nx) + a.nx * cl(hyb, a.ny)) * RB : oc;
    ohy[1] = ht ? (unsigned)(cl(x, a.nx) + a.nx * cl(hya, a.ny)) * RB : oc;
    mhy[0] = inb(x, a.nx) && inb(hyb, a.ny);
    mhy[1] = inb(x, a.nx) && inb(hya, a.ny);
  }
  const bool mhx = inb(hxx, a.nx) && inb(y, a.ny);
  // LDS destinations of the halo cells
  const int lxc = hl ? tx : tx + 2 * HW;                   // x-halo column
  const int lyb = ty, lya = ty + 2 * HW;                   // y-halo rows below / above
  // per-thread x/y rows
  R cx[Q], cy[Q];
  {
    const int rx = bound_row((uint32_t)x, (uint32_t)a.nx, a.nbl, 0), ry = bound_row((uint32_t)y, (uint32_t)a.ny, a.nbl, 0);
    const int c0 = (a.wl - 1) / 2;
    double tcx[Q], tcy[Q];  // unconditional loads (clamped column), masked below
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int j = min(max(c0 - HW + k, 0), a.wl - 1);
      tcx[k] = a.tab[rx * a.wl + j];
      tcy[k] = a.tab[ry * a.wl + j];
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const int j = c0 - HW + k;
      cx[k] = (j >= 0 && j < a.wl) ? (R)(tcx[k] * a.idx2) : R(0);
      cy[k] = (j >= 0 && j < a.wl) ? (R)(tcy[k] * a.idy2) : R(0);
    }
  }
  // z rows of the table, scaled by 1/dz^2 and rounded to REAL once (uniform per plane)
  __shared__ R sz[kLapRows][Q];
  {
    const int rows = 2 * a.nbl + 1, c0 = (a.wl - 1) / 2;
    for (int i = ty * TX + tx; i < rows * Q; i += TX * TY) {
      const int row = i / Q, k = i % Q, j = c0 - HW + k;
      sz[row][k] = (j >= 0 && j < a.wl) ? (R)(a.tab[row * a.wl + j] * a.idz2) : R(0);
    }
  }
  double dot = 0.0;
  const int z0 = a.off + bz * a.zc;
  const int zn = min(a.zc, gz - bz * a.zc);
  // u' from raw values (0 outside the grid)
  auto dir = [&](R uu, R rr, bool in) { return in ? beta * uu + rr : R(0); };
  // Raw loads ride in NS = PD+1 register slots that rotate with the plane index: the
  // column value of plane z0+HW+1+m and the halo of plane z0+m sit in slot m % NS.  The
  // z-loop is unrolled NS times so every slot index is a compile-time constant: a register
  // with a load in flight is never copied (a copy would wait for the load).
  constexpr int NS = PD + 1;
  static_assert(NS >= 2 && NS <= 4, "1 <= PD <= 3");
  constexpr int NH = 1 + NYH;      // halo cells per thread and plane: x, then y
  R w[Q];                          // u' of the column, planes z-HW .. z+HW
  R lu[NS], lr[NS];                // raw column values
  R hu[NS][NH], hrw[NS][NH];       // raw halo values
  auto load_col = [&](R& uu, R& rr, int zz) {
    const unsigned o = oc + (unsigned)cl(zz, a.nz) * plb;
    uu = ldo(u, o);
    rr = ldo(r, o);
  };
  auto load_halo = [&](R* uu, R* rr, int zz) {  // unconditional: every path issues 2*NH loads
    const unsigned pz = (unsigned)cl(zz, a.nz) * plb;
    uu[0] = ldo(u, ohx + pz);
    rr[0] = ldo(r, ohx + pz);
#pragma unroll
    for (int k = 0; k < NYH; ++k) {
      uu[1 + k] = ldo(u, ohy[k] + pz);
      rr[1 + k] = ldo(r, ohy[k] + pz);
    }
  };
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    R uu, rr;
    load_col(uu, rr, z0 - HW + j);
    w[j] = dir(uu, rr, colin && inb(z0 - HW + j, a.nz));
  }
  // first PD planes in flight, in the order the z-loop issues them (halo, then column); the
  // scheduling barriers keep that order, which the memory-counter waits in the loop rely on
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < PD; ++m) {
    load_halo(hu[m], hrw[m], z0 + m);
    load_col(lu[m], lr[m], z0 + HW + 1 + m);
    __builtin_amdgcn_sched_barrier(0);
  }
  auto body = [&](auto slot, int zi) {
    constexpr int S = decltype(slot)::value, SN = (S + PD) % NS;
    const int z = z0 + zi;
    auto& sp = s[zi & 1];
    sp[ty + HW][tx + HW] = w[HW];
    if (hl || hr) sp[ty + HW][lxc] = dir(hu[S][0], hrw[S][0], mhx);
    if constexpr (NYH == 1) {
      if (hb || ht) sp[hb ? lyb : lya][tx + HW] = dir(hu[S][1], hrw[S][1], mhy[0]);
    } else {
      if (hb) sp[lyb][tx + HW] = dir(hu[S][1], hrw[S][1], mhy[0]);
      if (ht) sp[lya][tx + HW] = dir(hu[S][2], hrw[S][2], mhy[1]);
    }
    // loads PD planes ahead (halo first), in flight across PD barriers and computes; past the
    // chunk they read a clamped plane and are never used
    load_halo(hu[SN], hrw[SN], z + PD);
    load_col(lu[SN], lr[SN], z + HW + 1 + PD);
    lds_barrier();
    if (active && zi < zn) {
      const int rz = bound_row((uint32_t)z, (uint32_t)a.nz, a.nbl, 0);
      R acc = R(0);
#pragma unroll
      for (int k = 0; k < Q; ++k) acc = acc + cx[k] * sp[ty + HW][tx + k] + cy[k] * sp[ty + k][tx + HW] + sz[rz][k] * w[k];
      const unsigned o = oc + (unsigned)z * plb;
      sto(un, o, w[HW]);
      sto(q, o, acc);
      dot += (double)w[HW] * (double)acc;
    }
    // advance: the window takes plane z+HW+1 (slot S, issued PD iterations ago)
#pragma unroll
    for (int j = 0; j < Q - 1; ++j) w[j] = w[j + 1];
    w[Q - 1] = dir(lu[S], lr[S], colin && inb(z + HW + 1, a.nz));
  };
  // Whole groups of NS planes (the host makes the chunk a multiple of NS; a shorter last
  // chunk runs its trailing planes without output), so every path issues the same loads.
  for (int zb = 0; zb < zn; zb += NS) {
    body(std::integral_constant<int, 0>{}, zb);
    body(std::integral_constant<int, 1 % NS>{}, zb + 1);
    if constexpr (NS > 2) body(std::integral_constant<int, 2 % NS>{}, zb + 2);
    if constexpr (NS > 3) body(std::integral_constant<int, 3 % NS>{}, zb + 3);
  }
  const double bs = block_sum(dot, red);
  if (tx == 0 && ty == 0) part[tile] = bs;
}

// One workgroup, after the last enqueued iteration when the stop test never held: rr of the
// final update and the iteration count (what the next Laplacian launch would record).
__global__ __launch_bounds__(256) void k_cg_finish(const double* rr_part, unsigned int n, CgState* st, int iter) {
  __shared__ double red[16];
  if (st->done) return;
  const double rr = sum_partials(rr_part, n, red);
  if (threadIdx.x == 0) {
    st->old = st->rrs[(iter + 1) & 1];
    st->rr = rr;
    st->it = iter;
    st->done = 1;
  }
}

// Every block sums the Laplacian's u'.q partials in the same order (alpha = rr / u'.q).  The
// reductions run inside the bulk kernels: a "last block finishes" scheme would need a
// device-scope fence in every block, which on gfx950 writes back and invalidates the L2 of
// its XCD.
template <typename R>
__global__ __launch_bounds__(kBlk) void k_cg_update(R* __restrict__ phi, const R* __restrict__ u, R* __restrict__ r,
                                                    const R* __restrict__ q, long long n, CgState* st,
                                                    const double* uq_part, unsigned int n_uq, double* part, int iter) {
  __shared__ double red[17];
  if (st->done) return;
  const double uq = sum_partials_all(uq_part, n_uq, red);
  const double alpha_d = st->rrs[iter & 1] / uq;
  if (threadIdx.x == 0 && blockIdx.x == 0) st->alpha = alpha_d;
  const R alpha = (R)alpha_d;
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
    st->rrs[0] = 0.0;
    st->rrs[1] = 1.0;  // old of iteration 0
    st->alpha = 0.0;
    st->it = 0;
    st->max_it = max_it;
    st->tol2 = tol2;
    st->rel2 = rel2;
    st->done = (rr > tol2 && (rel2 <= 0.0 || rr > rel2 * rr) && max_it > 0) ? 0 : 1;
  }
}

#ifndef IEQ_LAP_PD
#define IEQ_LAP_PD 1
#endif
constexpr int kLapPD = IEQ_LAP_PD;  // planes of the Laplacian's z-march in flight per thread
#ifndef IEQ_LAP_TY
#define IEQ_LAP_TY 16
#endif
#ifndef IEQ_LAP_TX
#define IEQ_LAP_TX 64
#endif
// Shape of a Laplacian tile (TX x TY threads) for stencil half-width hw: 64 x 16 up to hw = 5;
// the wider stencils hold more registers (they would spill at 4 waves per SIMD) and run 64 x 8.
constexpr int lap_ty(int hw) { return hw <= 5 ? IEQ_LAP_TY : 8; }
constexpr int lap_tx(int hw) { return hw <= 5 ? IEQ_LAP_TX : 64; }

template <typename R, int HW, int TY>
void launch_lap(const LapArgs& a, const R* u, const R* r, R* un, R* q, CgState* st, const double* rr_part, unsigned int n_rr,
                double* part, int iter, hipStream_t s) {
  const int gx = a.nx - 2 * a.off, gy = a.ny - 2 * a.off, gz = a.nz - 2 * a.off;
  constexpr int TX = lap_tx(HW);
  const int grid = ((gx + TX - 1) / TX) * ((gy + TY - 1) / TY) * ((gz + a.zc - 1) / a.zc);
  k_cg_lap<R, HW, TY, kLapPD, TX><<<grid, dim3(TX, TY), 0, s>>>(a, u, r, un, q, st, rr_part, n_rr, part, iter);
}

// Planes per z-chunk: the smallest chunk whose grid fits in one round of resident
// workgroups (so no partial second round), at least 8 planes, rounded up to the kernel's
// group of PD+1 planes.
template <typename R, int HW, int TY> int lap_zc(int gx, int gy, int gz) {
  constexpr int TX = lap_tx(HW);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_lap<R, HW, TY, kLapPD, TX>, TX * TY, 0) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 16;
  const long long xy = (long long)((gx + TX - 1) / TX) * ((gy + TY - 1) / TY);
  const long long resident = (long long)cus * per_cu;
  const long long chunks = std::max(1LL, resident / std::max(1LL, xy));
  const int ns = kLapPD + 1;  // the kernel's z-loop advances in groups of ns planes
  const int zc = std::max(8, (int)((gz + chunks - 1) / chunks));
  return (zc + ns - 1) / ns * ns;
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
  const int TY = lap_ty(hw);
  const int gx = p.nx - 2 * p.bnodes, gy = p.ny - 2 * p.bnodes, gz = p.nz - 2 * p.bnodes;
  int zc = 8;
  if (const char* e = getenv("IEQ_LAP_ZC")) zc = std::max(1, atoi(e));
  else switch (hw) {
    case 1: zc = lap_zc<R, 1, lap_ty(1)>(gx, gy, gz); break;
    case 2: zc = lap_zc<R, 2, lap_ty(2)>(gx, gy, gz); break;
    case 3: zc = lap_zc<R, 3, lap_ty(3)>(gx, gy, gz); break;
    case 4: zc = lap_zc<R, 4, lap_ty(4)>(gx, gy, gz); break;
    case 5: zc = lap_zc<R, 5, lap_ty(5)>(gx, gy, gz); break;
    case 6: zc = lap_zc<R, 6, lap_ty(6)>(gx, gy, gz); break;
    default: zc = lap_zc<R, 7, lap_ty(7)>(gx, gy, gz); break;
  }
  const int TX = lap_tx(hw);
  const unsigned int lap_blocks = (unsigned)(((gx + TX - 1) / TX) * ((gy + TY - 1) / TY) * ((gz + zc - 1) / zc));
  const unsigned int vec_blocks = 1024;
  const size_t fbytes = ((size_t)info.stride * sizeof(R) + 255) / 256 * 256;
  const size_t part_bytes = (size_t)(lap_blocks + vec_blocks) * sizeof(double);  // u'.q partials, then r.r partials
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
  double* part_uq = (double*)(base + 4 * fbytes);
  double* part_rr = part_uq + lap_blocks;
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
  k_cg_init<R><<<vec_blocks, kBlk, 0, s>>>(r, n, part_rr);
  k_cg_init_state<<<1, 256, 0, s>>>(part_rr, vec_blocks, st, abs_thr * abs_thr, rel_tol * rel_tol, max_it);

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
        case 1: launch_lap<R, 1, lap_ty(1)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        case 2: launch_lap<R, 2, lap_ty(2)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        case 3: launch_lap<R, 3, lap_ty(3)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        case 4: launch_lap<R, 4, lap_ty(4)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        case 5: launch_lap<R, 5, lap_ty(5)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        case 6: launch_lap<R, 6, lap_ty(6)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
        default: launch_lap<R, 7, lap_ty(7)>(la, u, r, un, q, st, part_rr, vec_blocks, part_uq, it, s); break;
      }
      k_cg_update<R><<<vec_blocks, kBlk, 0, s>>>(phi, un, r, q, n, st, part_uq, lap_blocks, part_rr, it);
    }
    launched += k;
    if (hipMemcpyAsync(&hst, st, sizeof(CgState), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return err = "CG state read failed", -4;
    if (hst.done) break;
  }
  k_cg_finish<<<1, 256, 0, s>>>(part_rr, vec_blocks, st, launched);  // no-op once the stop test held
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
