// Discrete adjoint of the fused periodic SSPRK33 solve (uniform u, Hall coefficient c_H = 1/rho).
//
// Forward stage (periodic central SBP operators, D_d^T = -D_d):
//   out = cP P + cQ Q + cD dt RHS(Y),
//   RHS(Y) = C(Y) + c_H H0(Y) + Diss(Y),
//   C(Y)   = -(u.grad_D) Y + s u (div_D Y)          (s = 1 for the ZERO source form)
//   H0(Y)  = -curl_D(J x Y),  J = curl_D Y
//   Diss   = symmetric HO dissipation.
// Transposed RHS applied to an adjoint field L (curl_D is symmetric on periodic grids):
//   RHS'(Y)^T L = (u.grad_D) L - s grad_D(u.L) - c_H [curl_D(Y x M) + M x J] + Diss(L),  M = curl_D L
// and the parameter derivatives of <L, RHS(Y)>:
//   d/dc_H = -<M, J x Y>,   d/du_k = sum [ -L.(D_k Y)/dx_k + s L_k div_D Y ].
// k_adj_pre evaluates F = c_H (Y x M), G = c_H (M x J) and the parameter partials in one
// z-marching pass over Y and L; the transposed stage is the forward stage kernel run on
// (Y := L, E := F) with the pointwise -G term, the transposed source form and u -> -u.
// Each block of k_adj_pre adds its scaled partial sums to its own slot of a running FP64
// accumulator (the grid, and so the block of every tile, is the same for every launch of a
// sweep); the slots are summed in block order once, when the sweep ends.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ieq_adjoint.hpp"
#include "ieq_device.hpp"
#include "ieq_stencil.hpp"

namespace ieq {

namespace {


template <typename R> struct PreArgs {
  int nx, ny, nz, zc;
  long long S;
  StencilCoef<R> c;
  R inv_h[3];  // (R)(1/dx_d)
  R c_hall;
  double scale;  // weight of this launch's partial sums
  int src_zero;
  int hall;  // 0: no Hall term (F, G not written, no c_H partial)
  const R* Y;
  const R* L;
  R* F;
  R* G;
  double* partial;  // [gridDim.x][4] running sums of d/dc_H, d/du_x, d/du_y, d/du_z
};

template <typename R, int RR, int TY, int TX>
__global__ __launch_bounds__(TX * TY) void k_adj_pre(PreArgs<R> a) {
  constexpr int Q = 2 * RR + 1;
  constexpr int LX = TX + 2 * RR, LY = TY + 2 * RR, NW = TX * TY / 64;
  __shared__ R sYb[2][3][LY][LX];  // double-buffered planes: plane z in buffer (z - z0) & 1
  __shared__ R sLb[2][3][LY][LX];
  __shared__ double red[4][NW];
  const int tx = threadIdx.x, ty = threadIdx.y;
  int x0, y0, z0;
  tile_origin<TY, TX>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  const bool active = x0 + tx < a.nx && y0 + ty < a.ny;
  const long long S = a.S;
  const unsigned RB = sizeof(R);
  const unsigned colb = (unsigned)(wrap(x0 + tx, a.nx) + a.nx * wrap(y0 + ty, a.ny)) * RB;
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;
  const int zend = min(z0 + a.zc, a.nz);
  PlaneStage<R, RR, TY, 3, 0b111, 0b111, TX> hy;
  PlaneStage<R, RR, TY, 3, 0b110, 0b101, TX> hl;  // L: x halo L_y, L_z;  y halo L_x, L_z
  hy.init(tx, ty, x0, y0, a.nx, a.ny);
  hl.init(tx, ty, x0, y0, a.nx, a.ny);
  const Comps<R, 3> Yc = comps<R, 3>(a.Y, S), Lc = comps<R, 3>(a.L, S);

  R qy[3][Q + 1], ql[2][Q + 1];  // z-queues (planes z-RR .. z+RR, slot Q: plane z+RR+1 in flight)
  R lz;                          // L_z of the own column at the plane staged next
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const unsigned o = colb + (unsigned)wrap(z0 - RR + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) qy[c][j] = ldo(Yc.c[c], o);
    ql[0][j] = ldo(Lc.c[0], o);
    ql[1][j] = ldo(Lc.c[1], o);
  }
  hy.load(Yc, (unsigned)z0 * plb);
  hl.load(Lc, (unsigned)z0 * plb);
  lz = ldo(Lc.c[2], colb + (unsigned)z0 * plb);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int z = z0; z < zend; ++z) {
    const int buf = (z - z0) & 1;
    auto& sY = sYb[buf];
    auto& sL = sLb[buf];
#pragma unroll
    for (int c = 0; c < 3; ++c) sY[c][ty + RR][tx + RR] = qy[c][RR];
    sL[0][ty + RR][tx + RR] = ql[0][RR];
    sL[1][ty + RR][tx + RR] = ql[1][RR];
    sL[2][ty + RR][tx + RR] = lz;
    hy.store(sY, tx, ty);
    hl.store(sL, tx, ty);
    // next plane's loads, in the order they are consumed: halos and L_z, then the column; on
    // the last plane of the chunk they re-read that plane (unconditional loads keep the waits
    // for them exact)
    {
      const int zn = min(z + 1, zend - 1);
      hy.load(Yc, (unsigned)zn * plb);
      hl.load(Lc, (unsigned)zn * plb);
      lz = ldo(Lc.c[2], colb + (unsigned)zn * plb);
      const unsigned o = colb + (unsigned)wrap(z + RR + 1, a.nz) * plb;
#pragma unroll
      for (int c = 0; c < 3; ++c) qy[c][Q] = ldo(Yc.c[c], o);
      ql[0][Q] = ldo(Lc.c[0], o);
      ql[1][Q] = ldo(Lc.c[1], o);
    }
    const unsigned ic = colb + (unsigned)z * plb;
    lds_barrier();
    if (active) {
      R v[Q];
      R dY[3][3];  // dY[d][c] = (D_d Y_c) / h_d
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = sY[c][ty + RR][tx + k];
        dY[0][c] = deriv<R, RR>(a.c, v) * a.inv_h[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = sY[c][ty + k][tx + RR];
        dY[1][c] = deriv<R, RR>(a.c, v) * a.inv_h[1];
        dY[2][c] = deriv<R, RR>(a.c, qy[c]) * a.inv_h[2];
      }
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sL[1][ty + RR][tx + k];
      const R dxLy = deriv<R, RR>(a.c, v) * a.inv_h[0];
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sL[2][ty + RR][tx + k];
      const R dxLz = deriv<R, RR>(a.c, v) * a.inv_h[0];
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sL[0][ty + k][tx + RR];
      const R dyLx = deriv<R, RR>(a.c, v) * a.inv_h[1];
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sL[2][ty + k][tx + RR];
      const R dyLz = deriv<R, RR>(a.c, v) * a.inv_h[1];
      const R dzLx = deriv<R, RR>(a.c, ql[0]) * a.inv_h[2];
      const R dzLy = deriv<R, RR>(a.c, ql[1]) * a.inv_h[2];
      const R Mx = dyLz - dzLy, My = dzLx - dxLz, Mz = dxLy - dyLx;
      const R Jx = dY[1][2] - dY[2][1], Jy = dY[2][0] - dY[0][2], Jz = dY[0][1] - dY[1][0];
      const R Yx = qy[0][RR], Yy = qy[1][RR], Yz = qy[2][RR];
      const R Lx = ql[0][RR], Ly = ql[1][RR], Lz = sL[2][ty + RR][tx + RR];
      if (a.hall) {
        const R ch = a.c_hall;
        sto(a.F, ic, ch * (Yy * Mz - Yz * My));
        sto(a.F + S, ic, ch * (Yz * Mx - Yx * Mz));
        sto(a.F + 2 * S, ic, ch * (Yx * My - Yy * Mx));
        sto(a.G, ic, ch * (My * Jz - Mz * Jy));
        sto(a.G + S, ic, ch * (Mz * Jx - Mx * Jz));
        sto(a.G + 2 * S, ic, ch * (Mx * Jy - My * Jx));
        // -<M, J x Y>
        const double jxy0 = (double)Jy * Yz - (double)Jz * Yy, jxy1 = (double)Jz * Yx - (double)Jx * Yz,
                     jxy2 = (double)Jx * Yy - (double)Jy * Yx;
        acc[0] -= (double)Mx * jxy0 + (double)My * jxy1 + (double)Mz * jxy2;
      }
      const double divY = a.src_zero ? (double)dY[0][0] + dY[1][1] + dY[2][2] : 0.0;
      const double Ld[3] = {(double)Lx, (double)Ly, (double)Lz};
#pragma unroll
      for (int k = 0; k < 3; ++k)
        acc[1 + k] += -(Ld[0] * dY[k][0] + Ld[1] * dY[k][1] + Ld[2] * dY[k][2]) + Ld[k] * divY;
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) {
#pragma unroll
      for (int c = 0; c < 3; ++c) qy[c][j] = qy[c][j + 1];
      ql[0][j] = ql[0][j + 1];
      ql[1][j] = ql[1][j + 1];
    }
  }
  // block reduction in a fixed order (waves by flat thread index)
  const int tid = tx + TX * ty, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double v = acc[j];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (ty == 0 && tx < 4) {
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[tx][w];
    a.partial[(long long)blockIdx.x * 4 + tx] += a.scale * s;
  }
}

// acc[j] += scale * sum_b partial[b][j], summed in block order by one workgroup.
__global__ __launch_bounds__(256) void k_adj_accumulate(const double* __restrict__ partial, int nb, double scale,
                                                        double* __restrict__ acc) {
  __shared__ double sm[4][256];
  const int t = threadIdx.x;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = t; b < nb; b += 256)
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += partial[(long long)b * 4 + j];
#pragma unroll
  for (int j = 0; j < 4; ++j) sm[j][t] = s[j];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int j = 0; j < 4; ++j) sm[j][t] += sm[j][t + w];
    __syncthreads();
  }
  if (t < 4) acc[t] += scale * sm[t][0];
}

__global__ void k_adj_zero(double* acc, int n) {
  for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) acc[i] = 0.0;
}

template <typename R> __global__ void k_copy(R* __restrict__ dst, const R* __restrict__ src, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Optimal binomial checkpointing: cost[n][s] = min_m m + cost[n-m][s-1] + cost[m][s] forward steps
// to reverse n steps holding s free snapshots besides the start state (cost[1][s] = 0,
// cost[n][0] = n(n-1)/2); split[n][s] is the minimising m.
struct Binomial {
  int n = 0, s = 0;
  std::vector<long long> cost;
  std::vector<int> split;
  long long& C(int i, int j) { return cost[(size_t)i * (s + 1) + j]; }
  int& M(int i, int j) { return split[(size_t)i * (s + 1) + j]; }
  void build(int n_, int s_) {
    n = n_;
    s = s_;
    cost.assign((size_t)(n + 1) * (s + 1), 0);
    split.assign((size_t)(n + 1) * (s + 1), 0);
    for (int i = 1; i <= n; ++i) C(i, 0) = (long long)i * (i - 1) / 2;
    for (int j = 1; j <= s; ++j)
      for (int i = 2; i <= n; ++i) {
        long long best = -1;
        int bm = 1;
        for (int m = 1; m < i; ++m) {
          const long long c = m + C(i - m, j - 1) + C(m, j);
          if (best < 0 || c < best) {
            best = c;
            bm = m;
          }
        }
        C(i, j) = best;
        M(i, j) = bm;
      }
  }
};

}  // namespace

struct Adjoint {
  Params p{};
  ieq_program_info info{};
  int rb = 8, rr = 3;
  long long nsteps = 0;
  int snapshots = 0;
  hipStream_t s = nullptr;
  size_t vbytes = 0;
  long long nvals = 0;  // 3 * stride
  // device work registers
  void *cur = nullptr, *b2 = nullptr, *b3 = nullptr, *E = nullptr;
  void *lam = nullptr, *lam3 = nullptr, *lam2 = nullptr, *F = nullptr, *G = nullptr;
  std::vector<void*> slots;     // snapshot pool (slot 0: the initial state)
  std::vector<int> slot_step;   // step held by each slot, -1 = free
  double* partial = nullptr;
  double* acc = nullptr;
  int nblk = 0;
  int ty = 8, zc = 8;  // tile rows of k_adj_pre and z-chunk of every launch
  int pre_tx = 64;     // tile width of k_adj_pre
  int shape = 8;       // tile shape code of the stage and EMF launches (as in ieq_step)
  int zc_emf = 8;      // z-chunk of the EMF launches
  Binomial bin;
  double c_hall = 1.0, u[3] = {0, 0, 0};
  bool have_forward = false;
  long long fwd_steps = 0, adj_steps = 0;
  // snapshots >= nsteps: every step's stage inputs b1, b2, b3 are kept (3 nsteps + 1 fields)
  // and the backward recomputes nothing
  bool stages = false;
  void* store = nullptr;
  void* st(int which, long long k) const {  // which: 0 b1, 1 b2, 2 b3
    return (char*)store + (size_t)(which == 0 ? k : nsteps + 1 + (which - 1) * nsteps + k) * vbytes;
  }
};

namespace {

template <typename R> FArgs<R> base_args(const Adjoint& A, const StencilCoef<R>& coef) {
  FArgs<R> a{};
  a.nx = A.p.nx;
  a.ny = A.p.ny;
  a.nz = A.p.nz;
  a.zc = A.zc;
  a.S = A.info.stride;
  a.c = coef;
  const double h[3] = {A.p.dx, A.p.dy, A.p.dz};
  for (int d = 0; d < 3; ++d) {
    a.inv_h[d] = (R)(1.0 / h[d]);
    a.two_h[d] = (R)(2.0 / h[d]);
  }
  a.rho = (R)(1.0 / A.c_hall);
  for (int d = 0; d < 3; ++d) a.jsc[d] = (R)(A.c_hall / h[d]);
  for (int d = 0; d < 3; ++d) a.u[d] = (R)A.u[d];
  a.conv = (A.u[0] != 0.0 || A.u[1] != 0.0 || A.u[2] != 0.0);
  a.src_zero = A.p.form_source == 0;
  return a;
}

template <typename R> StencilCoef<R> coefs(const Adjoint& A) {
  StencilCoef<R> c;
  central_coefs<R>(A.info.opset, A.p.ho_factor, c);
  return c;
}

template <typename R> void emf_launch(const Adjoint& A, FArgs<R> a) {
  a.zc = A.zc_emf;
  launch_emf<R>(A.rr, A.shape, a, A.s);
}

// One forward SSPRK33 step from b1 with stage registers b2, b3 (and E); the last stage writes
// `out` (which may be b1 itself).  The reference stage arithmetic (kernel_time_integrator.cl:224-303).
template <typename R> void fwd_step_to(Adjoint& A, R* b1, R* b2, R* b3, R* out) {
  const StencilCoef<R> coef = coefs<R>(A);
  FArgs<R> a = base_args<R>(A, coef);
  R* F[4] = {b1, b2, b3, out};
  const bool hall = A.p.hall, hod = A.p.diss == 1;
  // (out, P, cP, Q, cQ, Y, cD) of the three stages
  const int outs[3] = {1, 2, 3}, Ps[3] = {0, 0, 0}, Qs[3] = {-1, 1, 2}, Ys[3] = {0, 1, 2};
  const double cPs[3] = {1.0, 0.75, 1.0 / 3.0}, cQs[3] = {0.0, 0.25, 2.0 / 3.0}, cDs[3] = {1.0, 0.25, 2.0 / 3.0};
  for (int k = 0; k < 3; ++k) {
    if (hall) {
      a.Y = F[Ys[k]];
      a.Eout = (R*)A.E;
      emf_launch<R>(A, a);
    }
    a.Y = F[Ys[k]];
    a.E = (R*)A.E;
    a.mode = 0;
    a.out = F[outs[k]];
    a.P = F[Ps[k]];
    a.Q = Qs[k] >= 0 ? F[Qs[k]] : nullptr;
    a.cP = (R)cPs[k];
    a.cQ = (R)cQs[k];
    a.cD = (R)(cDs[k] * A.p.dt);
    launch_stage<R>(A.rr, A.shape, a, hall, hod, A.s);
  }
  ++A.fwd_steps;
}

template <typename R> void fwd_step(Adjoint& A, R* b) { fwd_step_to<R>(A, b, (R*)A.b2, (R*)A.b3, b); }

template <typename R> void pre(Adjoint& A, const R* Y, const R* L, double scale) {
  PreArgs<R> q{};
  q.nx = A.p.nx;
  q.ny = A.p.ny;
  q.nz = A.p.nz;
  q.zc = A.zc;
  q.S = A.info.stride;
  q.c = coefs<R>(A);
  q.inv_h[0] = (R)(1.0 / A.p.dx);
  q.inv_h[1] = (R)(1.0 / A.p.dy);
  q.inv_h[2] = (R)(1.0 / A.p.dz);
  q.scale = scale;
  q.c_hall = (R)A.c_hall;
  q.src_zero = A.p.form_source == 0;
  q.hall = A.p.hall != 0;
  q.Y = Y;
  q.L = L;
  q.F = (R*)A.F;
  q.G = (R*)A.G;
  q.partial = A.partial;
  const int grid = A.nblk;
  dim3 blk(A.pre_tx, A.ty);
  if (A.ty == 4) {
    if (A.rr == 1) k_adj_pre<R, 1, 4, 64><<<grid, blk, 0, A.s>>>(q);
    else if (A.rr == 2) k_adj_pre<R, 2, 4, 64><<<grid, blk, 0, A.s>>>(q);
    else k_adj_pre<R, 3, 4, 64><<<grid, blk, 0, A.s>>>(q);
  } else if (A.pre_tx == 32) {
    if (A.rr == 1) k_adj_pre<R, 1, 8, 32><<<grid, blk, 0, A.s>>>(q);
    else if (A.rr == 2) k_adj_pre<R, 2, 8, 32><<<grid, blk, 0, A.s>>>(q);
    else k_adj_pre<R, 3, 8, 32><<<grid, blk, 0, A.s>>>(q);
  } else {
    if (A.rr == 1) k_adj_pre<R, 1, 8, 64><<<grid, blk, 0, A.s>>>(q);
    else if (A.rr == 2) k_adj_pre<R, 2, 8, 64><<<grid, blk, 0, A.s>>>(q);
    else k_adj_pre<R, 3, 8, 64><<<grid, blk, 0, A.s>>>(q);
  }
}

// Transposed stage: out = cP P + cQ Q + cY L + cD RHS'(.)^T L, with F, G from pre().
template <typename R> void adj_stage(Adjoint& A, const R* L, R* out, const R* P, double cP, const R* Q, double cQ,
                                     double cY, double cD) {
  const StencilCoef<R> coef = coefs<R>(A);
  FArgs<R> a = base_args<R>(A, coef);
  for (int d = 0; d < 3; ++d) a.u[d] = (R)(-A.u[d]);
  a.conv = 1;
  a.src_adj = a.src_zero;
  a.Y = L;
  a.E = (const R*)A.F;
  a.G = (const R*)A.G;
  a.mode = 0;
  a.out = out;
  a.P = P;
  a.Q = Q;
  a.cP = (R)cP;
  a.cQ = (R)cQ;
  a.cY = (R)cY;
  a.cD = (R)(cD * A.p.dt);
  if (!A.p.hall) {
    a.E = nullptr;
    a.G = nullptr;
  }
  launch_stage<R>(A.rr, A.shape, a, A.p.hall != 0, A.p.diss == 1, A.s);
}

// Reverse one SSPRK33 step given its stage inputs b1, b2, b3: lam <- step^T lam, parameter partials.
template <typename R> void adj_from(Adjoint& A, const R* b1, const R* b2, const R* b3) {
  R* lam = (R*)A.lam;
  R* lam3 = (R*)A.lam3;
  R* lam2 = (R*)A.lam2;
  // stage 3: b1' = 1/3 b1 + 2/3 b3 + 2/3 dt R(b3)
  pre<R>(A, b3, lam, 2.0 / 3.0 * A.p.dt);
  adj_stage<R>(A, lam, lam3, lam, 2.0 / 3.0, nullptr, 0.0, 0.0, 2.0 / 3.0);  // lam3 = 2/3 lam + 2/3 dt R^T lam
  // stage 2: b3 = 3/4 b1 + 1/4 b2 + 1/4 dt R(b2)
  pre<R>(A, b2, lam3, 0.25 * A.p.dt);
  adj_stage<R>(A, lam3, lam2, lam3, 0.25, nullptr, 0.0, 0.0, 0.25);  // lam2 = 1/4 lam3 + 1/4 dt R^T lam3
  // stage 1: b2 = b1 + dt R(b1); lam_b1 = 1/3 lam + 3/4 lam3 + lam2 + dt R^T lam2
  pre<R>(A, b1, lam2, A.p.dt);
  adj_stage<R>(A, lam2, lam, lam, 1.0 / 3.0, lam3, 0.75, 1.0, 1.0);
  ++A.adj_steps;
}

// Reverse one step whose start state is b1 (read only): recompute b2, b3 (the first two
// stages), then the transposed stages.
template <typename R> void adj_step(Adjoint& A, const R* b1) {
  const StencilCoef<R> coef = coefs<R>(A);
  FArgs<R> a = base_args<R>(A, coef);
  const bool hall = A.p.hall, hod = A.p.diss == 1;
  R* b2 = (R*)A.b2;
  R* b3 = (R*)A.b3;
  if (hall) {
    a.Y = b1;
    a.Eout = (R*)A.E;
    emf_launch<R>(A, a);
  }
  a.Y = b1; a.E = (R*)A.E; a.mode = 0; a.out = b2; a.P = b1; a.Q = nullptr; a.cP = R(1); a.cQ = R(0);
  a.cD = (R)(1.0 * A.p.dt);
  launch_stage<R>(A.rr, A.shape, a, hall, hod, A.s);
  if (hall) {
    a.Y = b2;
    a.Eout = (R*)A.E;
    emf_launch<R>(A, a);
  }
  a.Y = b2; a.E = (R*)A.E; a.out = b3; a.P = b1; a.Q = b2; a.cP = (R)0.75; a.cQ = (R)0.25; a.cD = (R)(0.25 * A.p.dt);
  launch_stage<R>(A.rr, A.shape, a, hall, hod, A.s);
  adj_from<R>(A, b1, b2, b3);
}

template <typename R> void copy(Adjoint& A, void* dst, const void* src) {
  const long long n = A.nvals;
  k_copy<R><<<(int)std::min<long long>((n + 255) / 256, 4096), 256, 0, A.s>>>((R*)dst, (const R*)src, n);
}

int free_slot(Adjoint& A) {
  for (size_t k = 1; k < A.slots.size(); ++k)
    if (A.slot_step[k] < 0) return (int)k;
  return -1;
}

// Reverse steps [i, i+n) whose start state is in slot `si`; `free` snapshot slots remain.
template <typename R> void reverse(Adjoint& A, long long i, long long n, int si, int nfree) {
  if (n <= 0) return;
  if (n == 1) {
    adj_step<R>(A, (const R*)A.slots[si]);
    return;
  }
  if (nfree == 0) {
    for (long long k = n - 1; k >= 0; --k) {
      copy<R>(A, A.cur, A.slots[si]);
      for (long long j = 0; j < k; ++j) fwd_step<R>(A, (R*)A.cur);
      adj_step<R>(A, (const R*)A.cur);
    }
    return;
  }
  const long long m = A.bin.M((int)n, std::min(nfree, A.bin.s));
  int sj = -1;
  for (size_t k = 1; k < A.slots.size(); ++k)
    if (A.slot_step[k] == i + m) sj = (int)k;  // stored by the forward sweep
  if (sj < 0) {
    sj = free_slot(A);
    if (sj < 0) {  // no slot left: recompute from the start state
      reverse<R>(A, i, n, si, 0);
      return;
    }
    copy<R>(A, A.cur, A.slots[si]);
    for (long long j = 0; j < m; ++j) fwd_step<R>(A, (R*)A.cur);
    copy<R>(A, A.slots[sj], A.cur);
    A.slot_step[sj] = (int)(i + m);
  }
  reverse<R>(A, i + m, n - m, sj, nfree - 1);
  A.slot_step[sj] = -1;
  reverse<R>(A, i, m, si, nfree);
}

template <typename R> int forward_t(Adjoint& A, const void* b0, void* bT) {
  if (A.stages) {
    copy<R>(A, A.st(0, 0), b0);
    for (long long k = 0; k < A.nsteps; ++k)
      fwd_step_to<R>(A, (R*)A.st(0, k), (R*)A.st(1, k), (R*)A.st(2, k), (R*)A.st(0, k + 1));
    copy<R>(A, bT, A.st(0, A.nsteps));
    return 0;
  }
  for (size_t k = 1; k < A.slots.size(); ++k) A.slot_step[k] = -1;
  copy<R>(A, A.slots[0], b0);
  A.slot_step[0] = 0;
  // snapshots along the first descent of the reversal: positions i + split(n - i, free)
  std::vector<long long> store_at;
  {
    long long i = 0, n = A.nsteps;
    int nfree = A.snapshots;
    while (n > 1 && nfree > 0) {
      const long long m = A.bin.M((int)n, std::min(nfree, A.bin.s));
      i += m;
      n -= m;
      --nfree;
      store_at.push_back(i);
    }
  }
  copy<R>(A, A.cur, A.slots[0]);
  size_t next = 0;
  for (long long k = 0; k < A.nsteps; ++k) {
    fwd_step<R>(A, (R*)A.cur);
    if (next < store_at.size() && store_at[next] == k + 1) {
      const int sj = free_slot(A);
      copy<R>(A, A.slots[sj], A.cur);
      A.slot_step[sj] = (int)(k + 1);
      ++next;
    }
  }
  copy<R>(A, bT, A.cur);
  return 0;
}

template <typename R> int backward_t(Adjoint& A, const void* lamT, void* grad_b0) {
  copy<R>(A, A.lam, lamT);
  k_adj_zero<<<1, 64, 0, A.s>>>(A.acc, 4);
  k_adj_zero<<<1, 256, 0, A.s>>>(A.partial, A.nblk * 4);
  if (A.stages) {
    for (long long k = A.nsteps - 1; k >= 0; --k)
      adj_from<R>(A, (const R*)A.st(0, k), (const R*)A.st(1, k), (const R*)A.st(2, k));
  } else {
    reverse<R>(A, 0, A.nsteps, 0, A.snapshots);
  }
  copy<R>(A, grad_b0, A.lam);
  k_adj_accumulate<<<1, 256, 0, A.s>>>(A.partial, A.nblk, 1.0, A.acc);  // the blocks' running sums, in block order
  return 0;
}

}  // namespace

int adjoint_create(const Params& p, const ieq_program_info& info, long long nsteps, int snapshots, Adjoint** out,
                   std::string& err) {
  *out = nullptr;
  if (!p.periodic) return err = "the adjoint needs a periodic program", IEQ_ERR_UNSUPPORTED;
  if (p.diss != 0 && p.diss != 1) return err = "the adjoint supports high-order dissipation only", IEQ_ERR_UNSUPPORTED;
  if (nsteps < 1 || nsteps > 1000000 || snapshots < 0) return err = "invalid step or snapshot count", IEQ_ERR_ARGS;
  Adjoint* A = new Adjoint();
  A->p = p;
  A->info = info;
  A->rb = info.real_bytes;
  StencilCoef<double> c;
  A->rr = central_coefs<double>(info.opset, p.ho_factor, c);
  if (A->rr < 1 || A->rr > 3) {
    delete A;
    return err = "the adjoint needs an operator whose central rows are at most 7 wide", IEQ_ERR_UNSUPPORTED;
  }
  if (!offsets_fit(p.nx, p.ny, p.nz, info.real_bytes)) {
    delete A;
    return err = "the grid is too large for the adjoint (one component must stay below 2 GiB)", IEQ_ERR_UNSUPPORTED;
  }
  A->nsteps = nsteps;
  A->snapshots = (int)std::min<long long>(snapshots, nsteps);
  A->nvals = 3 * info.stride;
  A->vbytes = (size_t)A->nvals * A->rb;
  // binomial table: the DP is O(n^2 s); beyond 4096 steps the split is taken in halves
  if (nsteps > 4096) {
    delete A;
    return err = "the adjoint supports at most 4096 steps per solve", IEQ_ERR_UNSUPPORTED;
  }
  A->bin.build((int)nsteps, std::max(A->snapshots, 1));
  void** regs[9] = {&A->cur, &A->b2, &A->b3, &A->E, &A->lam, &A->lam3, &A->lam2, &A->F, &A->G};
  for (void** r : regs)
    if (hipMalloc(r, A->vbytes) != hipSuccess) {
      adjoint_release(A);
      return err = "hipMalloc failed (adjoint registers)", IEQ_ERR_HIP;
    }
  if (snapshots >= nsteps) {  // keep every stage input when the device has room for it
    size_t freeb = 0, totalb = 0;
    const size_t need = (size_t)(3 * nsteps + 1) * A->vbytes;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && need < freeb / 2 && hipMalloc(&A->store, need) == hipSuccess) {
      A->stages = true;
      A->snapshots = 0;  // the snapshot pool below only holds the start state
    }
  }
  A->slots.assign(A->snapshots + 1, nullptr);
  A->slot_step.assign(A->snapshots + 1, -1);
  for (auto& sl : A->slots)
    if (hipMalloc(&sl, A->vbytes) != hipSuccess) {
      adjoint_release(A);
      return err = "hipMalloc failed (snapshots)", IEQ_ERR_HIP;
    }
  // tiles as in ieq_step: stage and EMF kernels FP64 32 x 8, FP32 64 x 8 where 64 divides nx and
  // 32 x 8 otherwise (no idle lanes); k_adj_pre FP64 64 x 4, FP32 64 x 8 or 32 x 8 alike; the
  // z-chunk that fills one round of resident workgroups of the forward stage kernel (SSPRK33
  // stage 2 form, convection on)
  const bool narrow = info.real_bytes == 4 && p.nx % 64 != 0;
  A->ty = info.real_bytes == 8 ? 4 : 8;
  A->pre_tx = narrow ? 32 : 64;
  A->shape = info.real_bytes == 8 || narrow ? 32 : 8;
  {
    int fake[3];  // distinct addresses standing for b1, b2, E: only compared, never read
    if (info.real_bytes == 8) {
      StencilCoef<double> cc;
      central_coefs<double>(info.opset, p.ho_factor, cc);
      FArgs<double> t = base_args<double>(*A, cc);
      t.conv = 1;
      t.Y = t.Q = (const double*)&fake[1];
      t.P = (const double*)&fake[0];
      t.E = (const double*)&fake[2];
      A->zc = pick_zc<double>(A->rr, A->shape, t, p.hall != 0, p.diss == 1);
      A->zc_emf = pick_zc_emf<double>(A->rr, A->shape, t);
    } else {
      StencilCoef<float> cc;
      central_coefs<float>(info.opset, p.ho_factor, cc);
      FArgs<float> t = base_args<float>(*A, cc);
      t.conv = 1;
      t.Y = t.Q = (const float*)&fake[1];
      t.P = (const float*)&fake[0];
      t.E = (const float*)&fake[2];
      A->zc = pick_zc<float>(A->rr, A->shape, t, p.hall != 0, p.diss == 1);
      A->zc_emf = pick_zc_emf<float>(A->rr, A->shape, t);
    }
  }
  A->nblk = ((p.nx + A->pre_tx - 1) / A->pre_tx) * ((p.ny + A->ty - 1) / A->ty) * ((p.nz + A->zc - 1) / A->zc);
  if (hipMalloc(&A->partial, (size_t)A->nblk * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&A->acc, 4 * sizeof(double)) != hipSuccess) {
    adjoint_release(A);
    return err = "hipMalloc failed (partials)", IEQ_ERR_HIP;
  }
  *out = A;
  return IEQ_OK;
}

int adjoint_forward(Adjoint* A, hipStream_t s, const void* b0, void* bT, double c_hall, const double* u, double* seconds,
                    std::string& err) {
  A->s = s;
  A->c_hall = c_hall;
  for (int d = 0; d < 3; ++d) A->u[d] = u ? u[d] : 0.0;
  if (!(c_hall == c_hall) || c_hall == 0.0) return err = "the Hall coefficient must be non-zero", IEQ_ERR_ARGS;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  (void)hipEventRecord(t0, s);
  A->fwd_steps = 0;
  if (A->rb == 8) forward_t<double>(*A, b0, bT);
  else forward_t<float>(*A, b0, bT);
  (void)hipEventRecord(t1, s);
  hipError_t e = hipEventSynchronize(t1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (seconds) *seconds = ms * 1e-3;
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return err = "adjoint forward failed", IEQ_ERR_HIP;
  A->have_forward = true;
  return IEQ_OK;
}

int adjoint_backward(Adjoint* A, hipStream_t s, const void* lamT, void* grad_b0, double* grads, double* seconds,
                     std::string& err) {
  if (!A->have_forward) return err = "ieq_adjoint_forward must run before ieq_adjoint_backward", IEQ_ERR_ARGS;
  A->s = s;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  (void)hipEventRecord(t0, s);
  A->adj_steps = 0;
  if (A->rb == 8) backward_t<double>(*A, lamT, grad_b0);
  else backward_t<float>(*A, lamT, grad_b0);
  (void)hipEventRecord(t1, s);
  hipError_t e = hipEventSynchronize(t1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (seconds) *seconds = ms * 1e-3;
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return err = "adjoint backward failed", IEQ_ERR_HIP;
  double g[4];
  if (hipMemcpy(g, A->acc, sizeof(g), hipMemcpyDeviceToHost) != hipSuccess) return err = "copy failed", IEQ_ERR_HIP;
  for (int j = 0; j < 4; ++j) grads[j] = g[j];
  return IEQ_OK;
}

void adjoint_counts(const Adjoint* A, long long* fwd, long long* adj) {
  *fwd = A->fwd_steps;
  *adj = A->adj_steps;
}

void adjoint_release(Adjoint* A) {
  if (!A) return;
  void* regs[9] = {A->cur, A->b2, A->b3, A->E, A->lam, A->lam3, A->lam2, A->F, A->G};
  for (void* r : regs)
    if (r) (void)hipFree(r);
  for (void* sl : A->slots)
    if (sl) (void)hipFree(sl);
  if (A->partial) (void)hipFree(A->partial);
  if (A->acc) (void)hipFree(A->acc);
  if (A->store) (void)hipFree(A->store);
  delete A;
}

}  // namespace ieq
