// Device-side building blocks shared by all kernels of libinductioneq_hip.
//
// Data layout in HBM: a vector field is SoA, three component planes of `stride`
// REALs each (component c of node idx at f[c*stride + idx]); the node index is the
// reference's ix + NX*(iy + NY*iz) (include/utils.h:66-68).  Scalar fields are one
// plane.  `stride` is the node count rounded up to 64 elements, so every plane
// starts 256-byte aligned.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IEQ_TABLE_STORAGE static __constant__ const
#include "ieq_tables.h"

namespace ieq {

// ---------------------------------------------------------------------------------
// Run-time parameters of a program (the reference's compile-time -D settings).
struct Params {
  int nx, ny, nz;
  int periodic;
  long long stride;
  double dx, dy, dz;
  double xmin, ymin, zmin;
  double dt;
  int hall;         // USE_HALL
  int diss;         // 0 none, 1 high order, 2 first order, 3 adaptive
  int form_uibj;    // 0 CENTRAL, 1 SPLIT, 2 PRODUCT
  int form_source;  // 0 ZERO, 1 CENTRAL, 2 SPLIT
  int form_ujbi;    // 0 CENTRAL, 1 SPLIT, 2 PRODUCT
  int testcase;     // see TC_*
  int laplace;      // 0 WS_D0, 1 NS_D0, 2 WS_LN
  int bnodes;       // BNODES
  double ho_factor, mp2min, mp2max, cmin, cmax;
  double ux, uy, uz;  // uniform velocity of the uniform_flow testcase
};

enum { TC_ROTATION_2D = 0, TC_ROTATION_3D, TC_HALL_PERIODIC, TC_HALL_TRAVELLING_WAVE, TC_CONFINED_DOMAIN,
       TC_ALFVEN_PERIODIC_2D, TC_GAUSSIAN_ROTATION_3D, TC_UNIFORM_FLOW };

// ---------------------------------------------------------------------------------
// Operator-set traits over the __constant__ tables.
template <int OPS> struct Op;

#define IEQ_DEFINE_OP(ID, lo, UP, NSD0)                                                           \
  template <> struct Op<ID> {                                                                     \
    static constexpr int ORDER = IEQ_##UP##_ORDER, NB = IEQ_##UP##_NB, W = IEQ_##UP##_W;          \
    static constexpr int WH = IEQ_##UP##_W_HOD;                                                   \
    static constexpr int LNB0 = IEQ_##UP##_LAP_WS_D0_NB, LW0 = IEQ_##UP##_LAP_WS_D0_W;            \
    static constexpr int LNB1 = IEQ_##UP##_LAP_NS_D0_NB, LW1 = IEQ_##UP##_LAP_NS_D0_W;            \
    static constexpr int LNB2 = IEQ_##UP##_LAP_WS_LN_NB, LW2 = IEQ_##UP##_LAP_WS_LN_W;            \
    __device__ static double d(int r, int i) { return ieq_##lo##_diff[r * W + i]; }               \
    __device__ static double dadj(int r, int i) { return ieq_##lo##_diff_adj[r * W + i]; }        \
    __device__ static double minv(int r) { return ieq_##lo##_minv[r]; }                           \
    __device__ static double dho(int r, int i) { return ieq_##lo##_dho[r * WH + i]; }             \
    __device__ static double lap0(int r, int i) { return ieq_##lo##_lap_ws_d0[r * LW0 + i]; }     \
    __device__ static double lap1(int r, int i) { return NSD0; }                                  \
    __device__ static double lap2(int r, int i) { return ieq_##lo##_lap_ws_ln[r * LW2 + i]; }     \
  };

IEQ_DEFINE_OP(0, 2nd, 2ND, ieq_2nd_lap_ns_d0[r * LW1 + i])
IEQ_DEFINE_OP(1, 4th, 4TH, ieq_4th_lap_ns_d0[r * LW1 + i])
IEQ_DEFINE_OP(2, 6th, 6TH, ieq_6th_lap_ns_d0[r * LW1 + i])
IEQ_DEFINE_OP(3, 2nd_ext, 2ND_EXT, 0.0)
IEQ_DEFINE_OP(4, 4th_ext, 4TH_EXT, 0.0)
IEQ_DEFINE_OP(5, 6th_ext, 6TH_EXT, 0.0)
#undef IEQ_DEFINE_OP

// ---------------------------------------------------------------------------------
// Small 3-vector.
template <typename R> struct V3 {
  R x, y, z;
  __device__ __host__ R operator[](int c) const { return c == 0 ? x : c == 1 ? y : z; }
};
template <typename R> __device__ inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> __device__ inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> __device__ inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> __device__ inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename R> __device__ inline V3<R> zero3() { return {R(0), R(0), R(0)}; }

template <typename R> __device__ inline V3<R> ld3(const R* __restrict__ f, long long s, long long i) {
  return {f[i], f[s + i], f[2 * s + i]};
}
template <typename R> __device__ inline void st3(R* __restrict__ f, long long s, long long i, V3<R> v) {
  f[i] = v.x;
  f[s + i] = v.y;
  f[2 * s + i] = v.z;
}

// ---------------------------------------------------------------------------------
// Two-point fluxes shared by the reference-schedule and the tiled stage kernels.
template <typename R> __device__ inline R cmp(V3<R> v, int d) { return d == 0 ? v.x : d == 1 ? v.y : v.z; }

// Two-point volume flux of the linear term along direction d
// (induction_Eq_volume.cl:13-165): f = f(u_i B_j) - f(u_i d_j B_j) - f(u_j B_i).
template <typename R>
__device__ inline V3<R> volume_flux(const Params& p, int d, V3<R> Bm, V3<R> um, V3<R> Bk, V3<R> uk) {
  R bmd = cmp(Bm, d), bkd = cmp(Bk, d), umd = cmp(um, d), ukd = cmp(uk, d);
  V3<R> f1, f2, f3;
  switch (p.form_uibj) {
    case 0: f1 = R(0.5) * (bmd * um + bkd * uk); break;
    case 1: f1 = (R(0.25) * (bmd + bkd)) * (um + uk); break;
    default: f1 = R(0.5) * (bkd * um + bmd * uk); break;
  }
  switch (p.form_source) {
    case 0: f2 = zero3<R>(); break;
    case 1: f2 = (R(0.5) * (bkd - bmd)) * um; break;
    default: f2 = (R(0.25) * (bkd - bmd)) * (um + uk); break;
  }
  switch (p.form_ujbi) {
    case 0: f3 = R(0.5) * (umd * Bm + ukd * Bk); break;
    case 1: f3 = (R(0.25) * (umd + ukd)) * (Bm + Bk); break;
    default: f3 = R(0.5) * (umd * Bk + ukd * Bm); break;
  }
  return (f1 - f2) - f3;
}

template <typename R> __device__ inline V3<R> sigma3(V3<R> a, V3<R> b) {
  return {fmax(fabs(a.x), fabs(b.x)), fmax(fabs(a.y), fabs(b.y)), fmax(fabs(a.z), fabs(b.z))};
}

// ---------------------------------------------------------------------------------
// Boundary row selection: row of a (2nb+1)-row table for node i of n nodes.  Equals
// the reference's accumulation over check_bound_l / check_bound_*r (utils.h:172-201),
// including its behaviour on grids shorter than 2*nb.
__device__ inline int bound_row(uint32_t i, uint32_t n, int nb, int periodic) {
  if (periodic) return nb;
  int b = 0;
  for (int k = 0; k < nb; ++k) b += (nb - k) * ((int)(i == n - (uint32_t)(k + 1)) - (int)(i == (uint32_t)k));
  return nb + b;
}

// Neighbour along one axis: periodic wrap (utils.h:91-98) or, on a bounded grid,
// the node itself when the offset leaves the grid (utils.h:144-148).
__device__ inline uint32_t nbr(uint32_t i, int off, uint32_t n, int periodic) {
  uint32_t a = (uint32_t)(off < 0 ? -off : off);
  if (periodic) {
    uint32_t j = i + (uint32_t)off;
    if (off < 0 && i < a) j += n;
    if (off > 0 && i >= n - a) j -= n;
    return j;
  }
  if (off < 0) return i >= a ? i + (uint32_t)off : i;
  if (off > 0) return i < n - a ? i + (uint32_t)off : i;
  return i;
}

__device__ inline long long lin(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz) {
  return (long long)ix + (long long)p.nx * ((long long)iy + (long long)p.ny * iz);
}

// Index of the node at offset o along axis d from (ix, iy, iz).
__device__ inline long long at_dir(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, int d, int o) {
  if (d == 0) ix = nbr(ix, o, (uint32_t)p.nx, p.periodic);
  else if (d == 1) iy = nbr(iy, o, (uint32_t)p.ny, p.periodic);
  else iz = nbr(iz, o, (uint32_t)p.nz, p.periodic);
  return lin(p, ix, iy, iz);
}

__device__ inline double hdir(const Params& p, int d) { return d == 0 ? p.dx : d == 1 ? p.dy : p.dz; }

// ---------------------------------------------------------------------------------
// Testcases (include/<testcase>.h): analytic field, boundary data, velocity.
template <typename R>
__device__ inline void node_xyz(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, R& x, R& y, R& z) {
  x = (R)p.xmin + (R)ix * (R)p.dx;
  y = (R)p.ymin + (R)iy * (R)p.dy;
  z = (R)p.zmin + (R)iz * (R)p.dz;
}

template <typename R> __device__ V3<R> b_analytical(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, R t) {
  R x, y, z;
  node_xyz(p, ix, iy, iz, x, y, z);
  switch (p.testcase) {
    case TC_ROTATION_2D: {  // rotation_2D.h:15-32
      R s = sin(t), c = cos(t);
      R xx = c * x + s * y, yy = -s * x + c * y;
      R fac = exp(R(-20) * ((xx - R(0.5)) * (xx - R(0.5)) + yy * yy));
      R bx0 = R(-4) * yy * fac, by0 = (R(4) * xx - R(2)) * fac;
      return {c * bx0 - s * by0, s * bx0 + c * by0, R(0)};
    }
    case TC_GAUSSIAN_ROTATION_3D: {  // curl(0,0,exp(-20((x-1/2)^2+y^2+z^2))) rotated about z
      R s = sin(t), c = cos(t);
      R xx = c * x + s * y, yy = -s * x + c * y;
      R fac = exp(R(-20) * ((xx - R(0.5)) * (xx - R(0.5)) + yy * yy + z * z));
      R bx0 = R(-40) * yy * fac, by0 = (R(40) * xx - R(20)) * fac;
      return {c * bx0 - s * by0, s * bx0 + c * by0, R(0)};
    }
    case TC_ROTATION_3D: {  // rotation_3D.h:7-33
      R s = sin(t), c = cos(t), q = sqrt(R(3));
      R xx = ((1 + 2 * c) * x + (1 - c + q * s) * y + (1 - c - q * s) * z) / R(3);
      R yy = ((1 - c - q * s) * x + (1 + 2 * c) * y + (1 - c + q * s) * z) / R(3);
      R zz = ((1 - c + q * s) * x + (1 - c - q * s) * y + (1 + 2 * c) * z) / R(3);
      R al = exp(-R(5.0 / 3.0) * (3 - 2 * (3 + q) * xx + 12 * xx * xx - 2 * (-3 + q) * yy + 12 * yy * yy + 4 * q * zz +
                                 12 * zz * zz));
      R b0x = al * (3 - q - 4 * q * yy + 4 * q * zz) / R(48);
      R b0y = al * (-3 - q + 4 * q * xx - 4 * q * zz) / R(48);
      R b0z = al * (1 - 2 * xx + 2 * yy) / (8 * q);
      return {((1 + 2 * c) * b0x + (1 - c - q * s) * b0y + (1 - c + q * s) * b0z) / R(3),
              ((1 - c + q * s) * b0x + (1 + 2 * c) * b0y + (1 - c - q * s) * b0z) / R(3),
              ((1 - c - q * s) * b0x + (1 - c + q * s) * b0y + (1 + 2 * c) * b0z) / R(3)};
    }
    case TC_HALL_PERIODIC: {  // hall_periodic.h:15-34
      const R n1 = R(0.5773502691896258), al = R(0.5);
      R kk = (R(1) - al * al) / al;
      R u1 = cos(kk * y + al * kk * t * n1) + sin(kk * z + al * kk * t * n1);
      R u2 = cos(kk * z + al * kk * t * n1) + sin(kk * x + al * kk * t * n1);
      R u3 = cos(kk * x + al * kk * t * n1) + sin(kk * y + al * kk * t * n1);
      return {al * u1 + n1, al * u2 + n1, al * u3 + n1};
    }
    case TC_HALL_TRAVELLING_WAVE: {  // hall_travelling_wave.h:15-30
      const R uz = R(0.1);
      R xi = z - t;
      R om = ((uz - 1) * (uz - 1) - 1) / (uz - 1);
      return {cos(om * xi), sin(om * xi), R(1)};
    }
    case TC_CONFINED_DOMAIN: {  // confined_domain.h:7-24
      const R pi = R(3.141592653589793);
      R sx = sin(pi * x), cx = cos(pi * x), sy = sin(pi * y), cy = cos(pi * y), sz = sin(pi * z), cz = cos(pi * z);
      return {sx * cy * cz, cx * sy * cz, R(-2) * cx * cy * sz};
    }
    case TC_ALFVEN_PERIODIC_2D: {  // alfven_periodic_2D.h:8-26
      const R A = R(0.1), al = R(0.5235987755982988);
      R ca = cos(al), sa = sin(al);
      R ph = R(6.283185307179586) * (x * ca + y * sa + t);
      R cp = cos(ph), sp = sin(ph);
      return {ca - A * sa * sp, sa - A * ca * sp, A * cp};
    }
    default:
      return zero3<R>();
  }
}

template <typename R> __device__ V3<R> b_boundary(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, R t) {
  if (p.testcase == TC_HALL_PERIODIC || p.testcase == TC_HALL_TRAVELLING_WAVE || p.testcase == TC_UNIFORM_FLOW)
    return zero3<R>();
  return b_analytical<R>(p, ix, iy, iz, t);
}

template <typename R> __device__ V3<R> u_analytical(const Params& p, uint32_t ix, uint32_t iy, uint32_t iz, R t) {
  R x, y, z;
  node_xyz(p, ix, iy, iz, x, y, z);
  switch (p.testcase) {
    case TC_ROTATION_2D:
    case TC_GAUSSIAN_ROTATION_3D:
      return {-y, x, R(0)};
    case TC_ROTATION_3D: {
      R q = sqrt(R(3));
      return {(z - y) / q, (x - z) / q, (y - x) / q};
    }
    case TC_HALL_PERIODIC: {  // hall_periodic.h:75-91
      const R n1 = R(0.5773502691896258), al = R(0.5);
      R kk = (R(1) - al * al) / al;
      return {cos(kk * y + al * kk * t * n1) + sin(kk * z + al * kk * t * n1),
              cos(kk * z + al * kk * t * n1) + sin(kk * x + al * kk * t * n1),
              cos(kk * x + al * kk * t * n1) + sin(kk * y + al * kk * t * n1)};
    }
    case TC_HALL_TRAVELLING_WAVE: {  // hall_travelling_wave.h:141-157
      const R uz = R(0.1);
      R xi = z - t;
      R om = ((uz - 1) * (uz - 1) - 1) / (uz - 1);
      R fac = R(1) / (uz - 1);
      return {fac * cos(om * xi) + R(1), fac * sin(om * xi) + R(1), uz};
    }
    case TC_CONFINED_DOMAIN:
      return b_analytical<R>(p, ix, iy, iz, R(0));
    case TC_ALFVEN_PERIODIC_2D: {  // alfven_periodic_2D.h:222-240
      const R A = R(0.1), al = R(0.5235987755982988);
      R ca = cos(al), sa = sin(al);
      R ph = R(6.283185307179586) * (x * ca + y * sa + t);
      return {-A * sa * sin(ph), A * ca * sin(ph), A * cos(ph)};
    }
    case TC_UNIFORM_FLOW:
      return {(R)p.ux, (R)p.uy, (R)p.uz};
    default:
      return zero3<R>();
  }
}

}  // namespace ieq
