// Reference-schedule kernels: one work-item per node, every option of the program
// (order/closure, periodic or bounded grid, flux forms, Hall, SATs, the three
// dissipation variants, the -Laplacians).  These back every reference kernel name
// accepted by ieq_run(); the fused kernels in ieq_fused.hip replace the hot ones.
#include "ieq_device.hpp"
#include "ieq_generic.hpp"

namespace ieq {
namespace {

constexpr int BX = 64, BY = 4;

// kappa4 of include/artificial_dissipation.h:23-45 (k enters only through k*k).
template <typename R>
__device__ inline V3<R> kappa3(const Params& p, uint32_t k, R h, V3<R> b1, V3<R> b2, V3<R> u1, V3<R> u2) {
  k = k + (k == 0u);
  R kk = (R)(uint32_t)(k * k);
  R lo = kk * (R)p.mp2min * (R)p.cmin * h, hi = kk * (R)p.mp2max * (R)p.cmax * h;
  R out[3];
  for (int c = 0; c < 3; ++c) {
    R db = cmp(b1, c) - cmp(b2, c);
    R s = fmax(fabs(cmp(u1, c)), fabs(cmp(u2, c))) * db * db;
    out[c] = (R)((lo < s) && (s < hi)) * sqrt(fabs(s - lo) / fabs(hi - lo)) + (R)(s >= hi);
  }
  return {out[0], out[1], out[2]};
}

// One HOD partial sum, diff_HOD_* of include/artificial_dissipation.h:58-101.
template <typename R, int OPS>
__device__ inline V3<R> hod_part(const Params& p, const R* b, uint32_t ix, uint32_t iy, uint32_t iz, int d, int row, int s) {
  using O = Op<OPS>;
  const long long S = p.stride;
  const R sgn = ((O::ORDER / 2 + 1) & 1) ? R(-1) : R(1);
  V3<R> h = zero3<R>();
  for (int i = 0; i < O::WH; ++i) {
    R coef = (R)O::minv(row) * (R)O::dho(row, s + O::WH - 1) * (R)O::dho(O::NB, i) * sgn;
    h = h + coef * ld3(b, S, at_dir(p, ix, iy, iz, d, i + s));
  }
  return h;
}

// The semi-discrete right-hand side DB (kernel_time_integrator.cl:9-31).
template <typename R, int OPS>
__device__ V3<R> rhs_node(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, const R* b, const R* J, const R* u, R t) {
  using O = Op<OPS>;
  constexpr int C = (O::W - 1) / 2;
  const long long S = p.stride;
  const long long i0 = lin(p, ix, iy, iz);
  const uint32_t n3[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz}, i3[3] = {ix, iy, iz};
  int rows[3];
  for (int d = 0; d < 3; ++d) rows[d] = bound_row(i3[d], n3[d], O::NB, p.periodic);

  // convec_volume: induction_Eq_volume.cl:168-212
  V3<R> Bm = ld3(b, S, i0), um = ld3(u, S, i0);
  V3<R> db = zero3<R>();
  for (int d = 0; d < 3; ++d) {
    V3<R> acc = zero3<R>();
    for (int i = 0; i < O::W; ++i) {
      R c = (R)O::d(rows[d], i);
      long long j = at_dir(p, ix, iy, iz, d, i - C);
      acc = acc + c * volume_flux(p, d, Bm, um, ld3(b, S, j), ld3(u, S, j));
    }
    db = db + (R)(2.0 / hdir(p, d)) * acc;
  }

  const R m0 = (R)O::minv(0);
  if (p.hall) {
    if (!p.periodic) {  // convec_Hall_surface: induction_Eq_surface.cl:35-81
      V3<R> Jc = ld3(J, S, i0);
      V3<R> tv = R(0.5) * um - Jc;
      V3<R> sat = zero3<R>();
      for (int d = 0; d < 3; ++d) {
        int l = i3[d] == 0, r = i3[d] == n3[d] - 1;
        R tvd = cmp(tv, d);
        R s1 = (R)(l * (tvd > R(0))) * R(-1) + (R)(r * (tvd < R(0)));
        R s2 = (R)l * R(-1) + (R)r;
        sat = sat + (R)(m0 / (R)hdir(p, d)) * ((s1 * tvd) * Bm + (s2 * cmp(Bm, d)) * Jc);
      }
      db = db + sat;
    }
    // Hall_volume: induction_Eq_volume.cl:215-256
    V3<R> hv = zero3<R>();
    for (int d = 0; d < 3; ++d) {
      V3<R> acc = zero3<R>();
      for (int i = 0; i < O::W; ++i) {
        R c = (R)O::d(rows[d], i);
        long long j = at_dir(p, ix, iy, iz, d, i - C);
        V3<R> B = ld3(b, S, j), Jk = ld3(J, S, j);
        acc = acc + c * (cmp(Jk, d) * B - cmp(B, d) * Jk);
      }
      hv = hv + (R)(1.0 / hdir(p, d)) * acc;
    }
    db = db + hv;
  } else if (!p.periodic) {  // convec_surface: induction_Eq_surface.cl:7-33
    V3<R> g = b_boundary<R>(p, ix, iy, iz, t);
    V3<R> diff = Bm - g;
    V3<R> sat = zero3<R>();
    for (int d = 0; d < 3; ++d) {
      R ud = cmp(um, d);
      R s = (R)((i3[d] == 0) * (ud > R(0))) * R(-1) + (R)((i3[d] == n3[d] - 1) * (ud < R(0)));
      sat = sat + ((R)(m0 / (R)hdir(p, d)) * s * ud) * diff;
    }
    db = db + sat;
  }

  if (p.diss == 1) {  // high_order_dissipation: artificial_dissipation.cl:13-32
    const R F = (R)p.ho_factor;
    V3<R> h = zero3<R>();
    for (int s = -O::WH + 1; s < 1; ++s)
      for (int d = 0; d < 3; ++d) h = h + F * hod_part<R, OPS>(p, b, ix, iy, iz, d, rows[d], s);
    db = db + h;
  } else if (p.diss == 2) {  // first_order_dissipation: artificial_dissipation.cl:36-79
    V3<R> fo = zero3<R>();
    for (int d = 0; d < 3; ++d) {
      V3<R> acc = zero3<R>();
      for (int i = 0; i < O::W; ++i) {
        long long j = at_dir(p, ix, iy, iz, d, i - C);
        V3<R> sg = sigma3(ld3(u, S, j), um), dB = ld3(b, S, j) - Bm;
        R c = (R)fabs(O::d(rows[d], i));
        acc = acc + mk(c * sg.x * dB.x, c * sg.y * dB.y, c * sg.z * dB.z);
      }
      R h = (R)hdir(p, d);
      fo = fo + mk(acc.x / h, acc.y / h, acc.z / h);
    }
    db = db + fo;
  } else if (p.diss == 3) {  // adaptive_dissipation: artificial_dissipation.cl:83-160
    V3<R> ad = zero3<R>();
    for (int d = 0; d < 3; ++d) {
      R h = (R)hdir(p, d);
      V3<R> acc = zero3<R>();
      for (int i = 0; i < O::W; ++i) {
        long long j = at_dir(p, ix, iy, iz, d, i - C);
        V3<R> Bk = ld3(b, S, j), uk = ld3(u, S, j);
        V3<R> kap = kappa3<R>(p, (uint32_t)i - (uint32_t)C, h, Bk, Bm, uk, um);
        V3<R> sg = sigma3(uk, um), dB = Bk - Bm;
        R c = (R)fabs(O::d(rows[d], i));
        acc = acc + mk(c * kap.x * sg.x * dB.x, c * kap.y * sg.y * dB.y, c * kap.z * sg.z * dB.z);
      }
      ad = ad + mk(acc.x / h, acc.y / h, acc.z / h);
    }
    const R F = (R)p.ho_factor;
    for (int s = -O::WH + 1; s <= 0; ++s)
      for (int d = 0; d < 3; ++d) {
        long long j1 = at_dir(p, ix, iy, iz, d, s + O::WH - 1), j2 = at_dir(p, ix, iy, iz, d, s + O::WH - 2);
        V3<R> kap = kappa3<R>(p, 1u, (R)hdir(p, d), ld3(b, S, j1), ld3(b, S, j2), ld3(u, S, j1), ld3(u, S, j2));
        V3<R> hp = hod_part<R, OPS>(p, b, ix, iy, iz, d, rows[d], s);
        ad = ad + mk((R(1) - kap.x) * F * hp.x, (R(1) - kap.y) * F * hp.y, (R(1) - kap.z) * F * hp.z);
      }
    db = db + ad;
  }
  return db;
}

// d/dx_d of a vector field (diff_x/y/z, SBP_operator.cl:12-51) or of its M-adjoint.
template <typename R, int OPS, bool ADJ>
__device__ inline V3<R> dvec(const Params& p, const R* f, uint32_t ix, uint32_t iy, uint32_t iz, int d) {
  using O = Op<OPS>;
  constexpr int C = (O::W - 1) / 2;
  const uint32_t n3[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz}, i3[3] = {ix, iy, iz};
  int row = bound_row(i3[d], n3[d], O::NB, p.periodic);
  V3<R> acc = zero3<R>();
  for (int i = 0; i < O::W; ++i)
    acc = acc + (R)(ADJ ? O::dadj(row, i) : O::d(row, i)) * ld3(f, p.stride, at_dir(p, ix, iy, iz, d, i - C));
  R h = (R)hdir(p, d);
  return mk(acc.x / h, acc.y / h, acc.z / h);
}

template <typename R, int OPS, bool ADJ>
__device__ inline R dscal(const Params& p, const R* f, uint32_t ix, uint32_t iy, uint32_t iz, int d) {
  using O = Op<OPS>;
  constexpr int C = (O::W - 1) / 2;
  const uint32_t n3[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz}, i3[3] = {ix, iy, iz};
  int row = bound_row(i3[d], n3[d], O::NB, p.periodic);
  R acc = R(0);
  for (int i = 0; i < O::W; ++i) acc = acc + (R)(ADJ ? O::dadj(row, i) : O::d(row, i)) * f[at_dir(p, ix, iy, iz, d, i - C)];
  return acc / (R)hdir(p, d);
}

template <typename R, int OPS> __device__ inline V3<R> curl_node(const Params& p, const R* f, uint32_t ix, uint32_t iy, uint32_t iz) {
  V3<R> dx = dvec<R, OPS, false>(p, f, ix, iy, iz, 0), dy = dvec<R, OPS, false>(p, f, ix, iy, iz, 1),
        dz = dvec<R, OPS, false>(p, f, ix, iy, iz, 2);
  return {dy.z - dz.y, dz.x - dx.z, dx.y - dy.x};  // SBP_operator.cl:126-135
}

template <typename R, int OPS> __device__ inline R div_node(const Params& p, const R* f, uint32_t ix, uint32_t iy, uint32_t iz) {
  return dvec<R, OPS, false>(p, f, ix, iy, iz, 0).x + dvec<R, OPS, false>(p, f, ix, iy, iz, 1).y +
         dvec<R, OPS, false>(p, f, ix, iy, iz, 2).z;  // SBP_operator.cl:101-110
}

#define IEQ_NODE_PREAMBLE(OFF)                                            \
  uint32_t ix = blockIdx.x * BX + threadIdx.x, iy = blockIdx.y * BY + threadIdx.y, iz = blockIdx.z; \
  if (ix >= g.x || iy >= g.y) return;                                     \
  ix += (OFF); iy += (OFF); iz += (OFF);                                  \
  const long long idx = lin(p, ix, iy, iz);                               \
  const long long S = p.stride;                                           \
  (void)S; (void)idx;

struct Range { uint32_t x, y, z; };

// out = cP*P + cQ*Q + cD*DB(Y); null P/Q contribute nothing.
template <typename R, int OPS>
__global__ __launch_bounds__(BX* BY) void k_stage(Params p, Range g, R* out, const R* P, R cP, const R* Q, R cQ, const R* Y,
                                                  R cD, const R* J, const R* u, const R* time) {
  IEQ_NODE_PREAMBLE(0)
  const R t = time[1];
  V3<R> d = rhs_node<R, OPS>(p, ix, iy, iz, Y, J, u, t);
  V3<R> r = cD * d;
  if (Q) r = cQ * ld3(Q, S, idx) + r;
  if (P) r = cP * ld3(P, S, idx) + r;
  st3(out, S, idx, r);
}

template <typename R> __device__ inline void set_time(const Range& g, R* time, double c, double dt, uint32_t ix, uint32_t iy, uint32_t iz) {
  if (ix == 0 && iy == 0 && iz == 0 && c >= 0) time[1] = time[0] + (R)c * (R)dt;
}

// 2R "b" stage: X = X + a*D; O = X + b*D  (kernel_time_integrator.cl:642-660)
template <typename R>
__global__ __launch_bounds__(BX* BY) void k_update_2r(Params p, Range g, R* X, R* O, const R* D, R a, R b, R* time, double tc) {
  IEQ_NODE_PREAMBLE(0)
  for (int c = 0; c < 3; ++c) {
    R x = X[c * S + idx] + a * D[c * S + idx];
    X[c * S + idx] = x;
    O[c * S + idx] = x + b * D[c * S + idx];
  }
  set_time(g, time, tc, p.dt, ix, iy, iz);
}

// 2N "b" stage: X = X + b*D (kernel_time_integrator.cl:1036-1057)
template <typename R>
__global__ __launch_bounds__(BX* BY) void k_update_2n(Params p, Range g, R* X, const R* D, R b, R* time, double tc) {
  IEQ_NODE_PREAMBLE(0)
  for (int c = 0; c < 3; ++c) X[c * S + idx] = X[c * S + idx] + b * D[c * S + idx];
  set_time(g, time, tc, p.dt, ix, iy, iz);
}

// SSPRK104_06 (kernel_time_integrator.cl:458-477)
template <typename R>
__global__ __launch_bounds__(BX* BY) void k_ssprk104_06(Params p, Range g, const R* b1, R* b2, R* b3, R* time) {
  IEQ_NODE_PREAMBLE(0)
  for (int c = 0; c < 3; ++c) {
    R x1 = b1[c * S + idx], x2 = b2[c * S + idx];
    b3[c * S + idx] = R(0.04) * x1 + R(0.36) * x2;
    b2[c * S + idx] = R(0.60) * x1 + R(0.40) * x2;
  }
  if (ix == 0 && iy == 0 && iz == 0) time[1] = time[0] + (R)(p.dt / 3.0);
}

// time-only kernels: mode 0 sets time[1] = time[0] + c*DT; mode 1 is calc_time_* (time[0] += DT).
template <typename R> __global__ void k_time(R* time, double c, double dt, int mode) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (mode == 1) {
      time[0] = time[0] + (R)dt;
      time[1] = time[0];
    } else {
      time[1] = time[0] + (R)(c * dt);
    }
  }
}

template <typename R, int OPS>
__global__ __launch_bounds__(BX* BY) void k_curl_rho(Params p, Range g, const R* b, R* J, const R* rho) {
  IEQ_NODE_PREAMBLE(0)
  V3<R> c = curl_node<R, OPS>(p, b, ix, iy, iz);
  R r = rho[idx];
  st3(J, S, idx, mk(c.x / r, c.y / r, c.z / r));
}

template <typename R> __global__ __launch_bounds__(BX* BY) void k_calc_u(Params p, Range g, R* u, const R* time) {
  IEQ_NODE_PREAMBLE(0)
  st3(u, S, idx, u_analytical<R>(p, ix, iy, iz, time[1]));
}

// init_b / init_u / init_rho / analytical_b (kernel_init.cl:3-48)
template <typename R> __global__ __launch_bounds__(BX* BY) void k_init(Params p, Range g, R* f, int which, const R* time) {
  IEQ_NODE_PREAMBLE(0)
  if (which == 0) st3(f, S, idx, b_analytical<R>(p, ix, iy, iz, R(0)));
  else if (which == 1) st3(f, S, idx, u_analytical<R>(p, ix, iy, iz, R(0)));
  else if (which == 2) f[idx] = R(1);
  else st3(f, S, idx, b_analytical<R>(p, ix, iy, iz, time[0]));
}

template <typename R, int OPS> __global__ __launch_bounds__(BX* BY) void k_div(Params p, Range g, const R* b, R* out, int off) {
  IEQ_NODE_PREAMBLE(off)
  out[idx] = div_node<R, OPS>(p, b, ix, iy, iz);
}

template <typename R, int OPS> __global__ __launch_bounds__(BX* BY) void k_curl(Params p, Range g, const R* b, R* out) {
  IEQ_NODE_PREAMBLE(0)
  st3(out, S, idx, curl_node<R, OPS>(p, b, ix, iy, iz));
}

template <typename R, int OPS> __global__ __launch_bounds__(BX* BY) void k_grad(Params p, Range g, const R* f, R* out) {
  IEQ_NODE_PREAMBLE(0)
  st3(out, S, idx, mk(dscal<R, OPS, false>(p, f, ix, iy, iz, 0), dscal<R, OPS, false>(p, f, ix, iy, iz, 1),
                      dscal<R, OPS, false>(p, f, ix, iy, iz, 2)));
}

// -Laplacian of kernel_div_cleaning.cl:11-56 on the range offset by BNODES.
template <typename R, int OPS> __device__ inline R lap_node(const Params& p, const R* f, uint32_t ix, uint32_t iy, uint32_t iz) {
  using O = Op<OPS>;
  int nbl = p.laplace == 0 ? O::LNB0 : p.laplace == 1 ? O::LNB1 : O::LNB2;
  int wl = p.laplace == 0 ? O::LW0 : p.laplace == 1 ? O::LW1 : O::LW2;
  int c = (wl - 1) / 2;
  int bx = bound_row(ix, p.nx, nbl, p.periodic), by = bound_row(iy, p.ny, nbl, p.periodic), bz = bound_row(iz, p.nz, nbl, p.periodic);
  R dx2 = (R)p.dx * (R)p.dx, dy2 = (R)p.dy * (R)p.dy, dz2 = (R)p.dz * (R)p.dz;
  R val = R(0);
  for (int i = 0; i < wl; ++i) {
    R cx, cy, cz;
    if (p.laplace == 0) { cx = (R)O::lap0(bx, i); cy = (R)O::lap0(by, i); cz = (R)O::lap0(bz, i); }
    else if (p.laplace == 1) { cx = (R)O::lap1(bx, i); cy = (R)O::lap1(by, i); cz = (R)O::lap1(bz, i); }
    else { cx = (R)O::lap2(bx, i); cy = (R)O::lap2(by, i); cz = (R)O::lap2(bz, i); }
    val = val + cx * f[at_dir(p, ix, iy, iz, 0, i - c)] / dx2 + cy * f[at_dir(p, ix, iy, iz, 1, i - c)] / dy2 +
          cz * f[at_dir(p, ix, iy, iz, 2, i - c)] / dz2;
  }
  return val;
}

template <typename R, int OPS> __global__ __launch_bounds__(BX* BY) void k_lap(Params p, Range g, const R* f, R* out) {
  IEQ_NODE_PREAMBLE(p.bnodes)
  out[idx] = lap_node<R, OPS>(p, f, ix, iy, iz);
}

// projector_divcleaning (kernel_div_cleaning.cl:96-109): B +/- grad(phi), .xyz only
template <typename R, int OPS> __global__ __launch_bounds__(BX* BY) void k_proj(Params p, Range g, R* b, const R* phi) {
  IEQ_NODE_PREAMBLE(0)
  if (p.laplace == 2) {
    for (int d = 0; d < 3; ++d) b[d * S + idx] = b[d * S + idx] - dscal<R, OPS, true>(p, phi, ix, iy, iz, d);
  } else {
    for (int d = 0; d < 3; ++d) b[d * S + idx] = b[d * S + idx] + dscal<R, OPS, false>(p, phi, ix, iy, iz, d);
  }
}

// calc_axpy (x = a x + y) / calc_xpay (x = x + a y), kernel_CG.cl:8-30
template <typename R> __global__ __launch_bounds__(BX* BY) void k_axpy(Params p, Range g, R* x, const R* y, const R* a, int xpay) {
  IEQ_NODE_PREAMBLE(p.bnodes)
  R av = a[0];
  x[idx] = xpay ? x[idx] + av * y[idx] : av * x[idx] + y[idx];
}

// Group reductions: output[g] = sum over the W_SIZE nodes of group g of
//   which 0: fac*|B|^2 (norm2), 1: fac*|B1-B2|^2 (norm2_diff), 2: fac*s^2 (norm2_S), 3: a*b (dot_product)
// with fac the SBP norm weight prod_d dx_d / M_INV[row_d] (kernel_operator.cl:11-137), accumulated in
// double.
template <typename R, int OPS>
__global__ __launch_bounds__(256) void k_group_reduce(Params p, const R* a, const R* b, R* out, int which, long long wsize,
                                                      long long nnodes) {
  using O = Op<OPS>;
  __shared__ double part[256];
  const long long g0 = (long long)blockIdx.x * wsize;
  double acc = 0.0;
  const long long S = p.stride;
  const long long nxy = (long long)p.nx * p.ny;
  for (long long l = threadIdx.x; l < wsize; l += blockDim.x) {
    long long gid = g0 + l;
    if (gid >= nnodes) continue;  // padding nodes hold zeros
    // every product formed in double (FP32 programs widen the nodal values first); only the group
    // partial is rounded to REAL, as the reference's output array is REAL
    double v;
    if (which == 3) {
      v = (double)a[gid] * (double)b[gid];
    } else {
      uint32_t iz = (uint32_t)(gid / nxy), iy = (uint32_t)((gid - iz * nxy) / p.nx), ix = (uint32_t)(gid - iy * p.nx - iz * nxy);
      const double fac = (p.dx / O::minv(bound_row(ix, p.nx, O::NB, p.periodic))) *
                         (p.dy / O::minv(bound_row(iy, p.ny, O::NB, p.periodic))) *
                         (p.dz / O::minv(bound_row(iz, p.nz, O::NB, p.periodic)));
      double q;
      if (which == 0) {
        const double fx = a[gid], fy = a[S + gid], fz = a[2 * S + gid];
        q = fx * fx + fy * fy + fz * fz;
      } else if (which == 1) {
        const double fx = (double)a[gid] - (double)b[gid], fy = (double)a[S + gid] - (double)b[S + gid],
                     fz = (double)a[2 * S + gid] - (double)b[2 * S + gid];
        q = fx * fx + fy * fy + fz * fz;
      } else {
        const double f = a[gid];
        q = f * f;
      }
      v = fac * q;
    }
    acc += v;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = (R)part[0];
}

// AoS REAL4 <-> SoA conversion (the reference host layout <-> this library's).
template <typename R> __global__ void k_aos_to_soa(const R* aos, R* soa, long long n, long long stride) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  soa[i] = aos[4 * i];
  soa[stride + i] = aos[4 * i + 1];
  soa[2 * stride + i] = aos[4 * i + 2];
}
template <typename R> __global__ void k_soa_to_aos(const R* soa, R* aos, long long n, long long stride) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[4 * i] = soa[i];
  aos[4 * i + 1] = soa[stride + i];
  aos[4 * i + 2] = soa[2 * stride + i];
  aos[4 * i + 3] = R(0);
}

dim3 grid_of(Range g) { return dim3((g.x + BX - 1) / BX, (g.y + BY - 1) / BY, g.z); }

}  // namespace

// ----------------------------------------------------------------------------------
// Host launchers (dispatch the operator set at run time).

#define IEQ_OPS_SWITCH(ops, CALL) \
  switch (ops) {                  \
    case 0: { constexpr int OPS = 0; CALL; } break; \
    case 1: { constexpr int OPS = 1; CALL; } break; \
    case 2: { constexpr int OPS = 2; CALL; } break; \
    case 3: { constexpr int OPS = 3; CALL; } break; \
    case 4: { constexpr int OPS = 4; CALL; } break; \
    default: { constexpr int OPS = 5; CALL; } break; \
  }

template <typename R>
void gen_stage(const Params& p, int ops, hipStream_t s, const uint32_t g[3], R* out, const R* P, R cP, const R* Q, R cQ,
               const R* Y, R cD, const R* J, const R* u, const R* time) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_stage<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, out, P, cP, Q, cQ, Y, cD, J, u, time)))
}

template <typename R>
void gen_update_2r(const Params& p, hipStream_t s, const uint32_t g[3], R* X, R* O, const R* D, R a, R b, R* time, double tc) {
  Range r{g[0], g[1], g[2]};
  k_update_2r<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, X, O, D, a, b, time, tc);
}

template <typename R>
void gen_update_2n(const Params& p, hipStream_t s, const uint32_t g[3], R* X, const R* D, R b, R* time, double tc) {
  Range r{g[0], g[1], g[2]};
  k_update_2n<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, X, D, b, time, tc);
}

template <typename R> void gen_ssprk104_06(const Params& p, hipStream_t s, const uint32_t g[3], const R* b1, R* b2, R* b3, R* time) {
  Range r{g[0], g[1], g[2]};
  k_ssprk104_06<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, b1, b2, b3, time);
}

template <typename R> void gen_time(hipStream_t s, R* time, double c, double dt, int mode) {
  k_time<R><<<1, 64, 0, s>>>(time, c, dt, mode);
}

template <typename R> void gen_curl_rho(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* J, const R* rho) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_curl_rho<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, b, J, rho)))
}

template <typename R> void gen_calc_u(const Params& p, hipStream_t s, const uint32_t g[3], R* u, const R* time) {
  Range r{g[0], g[1], g[2]};
  k_calc_u<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, u, time);
}

template <typename R> void gen_init(const Params& p, hipStream_t s, const uint32_t g[3], R* f, int which, const R* time) {
  Range r{g[0], g[1], g[2]};
  k_init<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, f, which, time);
}

template <typename R> void gen_div(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* out, int off) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_div<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, b, out, off)))
}

template <typename R> void gen_curl(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* b, R* out) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_curl<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, b, out)))
}

template <typename R> void gen_grad(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* f, R* out) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_grad<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, f, out)))
}

template <typename R> void gen_lap(const Params& p, int ops, hipStream_t s, const uint32_t g[3], const R* f, R* out) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_lap<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, f, out)))
}

template <typename R> void gen_proj(const Params& p, int ops, hipStream_t s, const uint32_t g[3], R* b, const R* phi) {
  Range r{g[0], g[1], g[2]};
  IEQ_OPS_SWITCH(ops, (k_proj<R, OPS><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, b, phi)))
}

template <typename R> void gen_axpy(const Params& p, hipStream_t s, const uint32_t g[3], R* x, const R* y, const R* a, int xpay) {
  Range r{g[0], g[1], g[2]};
  k_axpy<R><<<grid_of(r), dim3(BX, BY), 0, s>>>(p, r, x, y, a, xpay);
}

template <typename R>
void gen_group_reduce(const Params& p, int ops, hipStream_t s, const R* a, const R* b, R* out, int which, long long wsize,
                      long long ngroups, long long nnodes) {
  IEQ_OPS_SWITCH(ops, (k_group_reduce<R, OPS><<<(unsigned)ngroups, 256, 0, s>>>(p, a, b, out, which, wsize, nnodes)))
}

template <typename R> void gen_aos_to_soa(hipStream_t s, const R* aos, R* soa, long long n, long long stride) {
  if (n > 0) k_aos_to_soa<R><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(aos, soa, n, stride);
}
template <typename R> void gen_soa_to_aos(hipStream_t s, const R* soa, R* aos, long long n, long long stride) {
  if (n > 0) k_soa_to_aos<R><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(soa, aos, n, stride);
}

#define IEQ_INSTANTIATE(R)                                                                                                   \
  template void gen_stage<R>(const Params&, int, hipStream_t, const uint32_t*, R*, const R*, R, const R*, R, const R*, R,  \
                             const R*, const R*, const R*);                                                                  \
  template void gen_update_2r<R>(const Params&, hipStream_t, const uint32_t*, R*, R*, const R*, R, R, R*, double);           \
  template void gen_update_2n<R>(const Params&, hipStream_t, const uint32_t*, R*, const R*, R, R*, double);                  \
  template void gen_ssprk104_06<R>(const Params&, hipStream_t, const uint32_t*, const R*, R*, R*, R*);                        \
  template void gen_time<R>(hipStream_t, R*, double, double, int);                                                           \
  template void gen_curl_rho<R>(const Params&, int, hipStream_t, const uint32_t*, const R*, R*, const R*);                   \
  template void gen_calc_u<R>(const Params&, hipStream_t, const uint32_t*, R*, const R*);                                    \
  template void gen_init<R>(const Params&, hipStream_t, const uint32_t*, R*, int, const R*);                                 \
  template void gen_div<R>(const Params&, int, hipStream_t, const uint32_t*, const R*, R*, int);                             \
  template void gen_curl<R>(const Params&, int, hipStream_t, const uint32_t*, const R*, R*);                                 \
  template void gen_grad<R>(const Params&, int, hipStream_t, const uint32_t*, const R*, R*);                                 \
  template void gen_lap<R>(const Params&, int, hipStream_t, const uint32_t*, const R*, R*);                                  \
  template void gen_proj<R>(const Params&, int, hipStream_t, const uint32_t*, R*, const R*);                                 \
  template void gen_axpy<R>(const Params&, hipStream_t, const uint32_t*, R*, const R*, const R*, int);                       \
  template void gen_group_reduce<R>(const Params&, int, hipStream_t, const R*, const R*, R*, int, long long, long long,      \
                                    long long);                                                                              \
  template void gen_aos_to_soa<R>(hipStream_t, const R*, R*, long long, long long);                                         \
  template void gen_soa_to_aos<R>(hipStream_t, const R*, R*, long long, long long);

IEQ_INSTANTIATE(double)
IEQ_INSTANTIATE(float)

}  // namespace ieq
