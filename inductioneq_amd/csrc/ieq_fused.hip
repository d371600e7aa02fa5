// Fused stage kernels for periodic grids with uniform velocity u and density rho.
//
// On a periodic grid only the central operator rows are used, so the right-hand
// side reduces to fixed stencils:
//   convection  (induction_Eq_volume.cl:168-212, two-point fluxes with u_m = u_k = u):
//               sum_d (2/dx_d) * 0.5*([source==ZERO] u (D_d B)_d - u_d (D_d B))
//   Hall term   (induction_Eq_volume.cl:215-256):  sum_d (1/dx_d) D_d (J_d B - J B_d)
//               = -curl_D(E),  E = J x B,  J = curl_D(B)/rho  -- the same pairwise
//               products J_a B_b - J_b B_a, regrouped per component of E
//   dissipation (artificial_dissipation.cl:13-32): F * sum_d sum_m w_m B(x + m e_d),
//               w = sign * (central D_HO row convolved with itself)
// With the Hall term a stage is one kernel, k_hall_stage (ieq_hall_tiles.hpp), which forms E
// on its tile and halo itself; IEQ_FUSED_SPLIT=1 selects the two-kernel form instead (k_emf
// writes E, k_stage reads B and E).  Without the Hall term a stage is one k_stage.  All
// kernels march along z with register queues of the z-stencil and stage one xy-plane (plus
// halo) at a time in LDS for the x/y stencils.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "ieq_device.hpp"
#include "ieq_fused.hpp"
#include "ieq_fused_tiles.hpp"
#include "ieq_generic.hpp"
#include "ieq_hall.hpp"
#include "ieq_stencil.hpp"

namespace ieq {

struct FusedWork {
  void* scratch = nullptr;  // one vector field: ping-pong register of the 2N schemes
  size_t scratch_bytes = 0;
  double* red = nullptr;    // uniformity check partials
  double* red_host = nullptr;
};

void fused_release(FusedWork* w) {
  if (!w) return;
  if (w->scratch) (void)hipFree(w->scratch);
  if (w->red) (void)hipFree(w->red);
  if (w->red_host) (void)hipHostFree(w->red_host);
  delete w;
}

int schedule(const char* integrator, int hall, int variable_u, std::vector<std::string>& s, int& nf) {
  s.clear();
  std::string I(integrator);
  auto pre = [&](const char* curl, const char* u) {
    if (hall) s.push_back(curl);
    if (variable_u) s.push_back(u);
  };
  if (I == "SSPRK33") {
    pre("calc_curlB_rho_1_3args", "calc_u_3_args");
    s.push_back("SSPRK33_1");
    s.push_back("SSPRK33_2a");
    pre("calc_curlB_rho_2_3args", "calc_u_3_args");
    s.push_back("SSPRK33_2b");
    s.push_back("SSPRK33_3a");
    pre("calc_curlB_rho_3_3args", "calc_u_3_args");
    s.push_back("SSPRK33_3b");
    s.push_back("calc_time_3_args");
    nf = 3;
    return 0;
  }
  if (I == "SSPRK104") {
    const char* tk[10] = {nullptr, "SSPRK104_02a", "SSPRK104_03a", "SSPRK104_04a", "SSPRK104_05a",
                          "SSPRK104_06", "SSPRK104_08a", "SSPRK104_09a", "SSPRK104_10a", "SSPRK104_11a"};
    const int cf[10] = {1, 2, 3, 2, 3, 2, 1, 2, 1, 2};
    const char* sk[10] = {"SSPRK104_01", "SSPRK104_02b", "SSPRK104_03b", "SSPRK104_04b", "SSPRK104_05b",
                          "SSPRK104_07", "SSPRK104_08b", "SSPRK104_09b", "SSPRK104_10b", "SSPRK104_11b"};
    for (int i = 0; i < 10; ++i) {
      if (tk[i]) s.push_back(tk[i]);
      std::string c = "calc_curlB_rho_" + std::to_string(cf[i]) + "_3args";
      pre(c.c_str(), "calc_u_3_args");
      s.push_back(sk[i]);
    }
    s.push_back("calc_time_3_args");
    nf = 3;
    return 0;
  }
  if (I == "KennedyCarpenterLewis2R54C" || I == "CalvoFrancoRandez2R64") {
    const bool kcl = I == "KennedyCarpenterLewis2R54C";
    const int n = kcl ? 5 : 6;
    const int kcl_curl[5] = {1, 1, 2, 1, 2}, cfr_curl[6] = {1, 2, 2, 2, 2, 2};
    for (int i = 1; i <= n; ++i) {
      std::string c = "calc_curlB_rho_" + std::to_string(kcl ? kcl_curl[i - 1] : cfr_curl[i - 1]) + "_3args";
      pre(c.c_str(), "calc_u_3_args");
      if (i < n) {
        s.push_back(I + "_" + std::to_string(i) + "a");
        s.push_back(I + "_" + std::to_string(i) + "b");
      } else {
        s.push_back(I + "_" + std::to_string(i));
        s.push_back("calc_time_3_args");
      }
    }
    nf = 3;
    return 0;
  }
  if (I == "CarpenterKennedy2N54" || I == "ToulorgeDesmet2N84F") {
    const int n = I == "CarpenterKennedy2N54" ? 5 : 8;
    for (int i = 1; i <= n; ++i) {
      pre("calc_curlB_rho_1_2args", "calc_u_2_args");
      s.push_back(I + "_" + std::to_string(i) + "a");
      s.push_back(I + "_" + std::to_string(i) + "b");
    }
    s.push_back("calc_time_2_args");
    nf = 2;
    return 0;
  }
  return -1;
}

namespace {
// min/max of u (3 components) and rho over the nodes: partials per block.
template <typename R> __global__ void k_uniform_check(const R* u, const R* rho, long long n, long long S, double* out) {
  double mn[4] = {1e300, 1e300, 1e300, 1e300}, mx[4] = {-1e300, -1e300, -1e300, -1e300};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v[4] = {(double)u[i], (double)u[S + i], (double)u[2 * S + i], (double)rho[i]};
    for (int c = 0; c < 4; ++c) {
      if (v[c] != v[c]) {  // NaN: never uniform (fmin/fmax would skip it)
        mn[c] = -INFINITY;
        mx[c] = INFINITY;
      } else {
        mn[c] = fmin(mn[c], v[c]);
        mx[c] = fmax(mx[c], v[c]);
      }
    }
  }
  __shared__ double smn[4][256], smx[4][256];
  for (int c = 0; c < 4; ++c) {
    smn[c][threadIdx.x] = mn[c];
    smx[c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 4; ++c) {
        smn[c][threadIdx.x] = fmin(smn[c][threadIdx.x], smn[c][threadIdx.x + s]);
        smx[c][threadIdx.x] = fmax(smx[c][threadIdx.x], smx[c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) {
      out[blockIdx.x * 8 + c] = smn[c][0];
      out[blockIdx.x * 8 + 4 + c] = smx[c][0];
    }
}

// nsteps applications of calc_time_* (kernel_time_integrator.cl:42-82).
template <typename R> __global__ void k_advance_time(R* time, R dt, long long nsteps) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    R t = time[0];
    for (long long k = 0; k < nsteps; ++k) t = t + dt;
    time[0] = t;
    time[1] = t;
  }
}

template <typename R, int RR, int TY, int NBUF, int TX = 64> const void* emf_kernel_ty() {
  return reinterpret_cast<const void*>(&k_emf<R, RR, TY, NBUF, TX>);
}

template <typename R, int RR, int TY, bool ADJ, int NBUF, int MODE, int TX>
const void* stage_kernel_hdc(bool hall, bool hod, bool conv) {
  if (conv) {
    if (hall && hod) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, true, true, true, ADJ, NBUF, MODE, TX>);
    if (hall) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, true, false, true, ADJ, NBUF, MODE, TX>);
    if (hod) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, false, true, true, ADJ, NBUF, MODE, TX>);
    return reinterpret_cast<const void*>(&k_stage<R, RR, TY, false, false, true, ADJ, NBUF, MODE, TX>);
  }
  if (hall && hod) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, true, true, false, ADJ, NBUF, MODE, TX>);
  if (hall) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, true, false, false, ADJ, NBUF, MODE, TX>);
  if (hod) return reinterpret_cast<const void*>(&k_stage<R, RR, TY, false, true, false, ADJ, NBUF, MODE, TX>);
  return reinterpret_cast<const void*>(&k_stage<R, RR, TY, false, false, false, ADJ, NBUF, MODE, TX>);
}

// LDS plane buffers of the single-grid stage kernel: 2 (one barrier per plane) or 1 (half the
// LDS, a second barrier per plane)
#ifndef IEQ_STAGE_NBUF
#define IEQ_STAGE_NBUF 2
#endif
// Stage kernel for the arguments: the transposed stages of the adjoint sweeps (G, src_adj or
// cY set) use the three-register form with convection always evaluated; otherwise the form
// follows stage_mode.
template <typename R, int RR, int TY, int TX = 64> const void* stage_kernel_ty(const FArgs<R>& a, bool hall, bool hod) {
  if (a.G || a.src_adj || a.cY != R(0)) return stage_kernel_hdc<R, RR, TY, true, IEQ_STAGE_NBUF, 0, TX>(hall, hod, true);
  switch (stage_mode(a.mode == 1, a.P, a.Q, a.Y)) {
    case 1: return stage_kernel_hdc<R, RR, TY, false, IEQ_STAGE_NBUF, 1, TX>(hall, hod, a.conv);
    case 2: return stage_kernel_hdc<R, RR, TY, false, IEQ_STAGE_NBUF, 2, TX>(hall, hod, a.conv);
    case 3: return stage_kernel_hdc<R, RR, TY, false, IEQ_STAGE_NBUF, 3, TX>(hall, hod, a.conv);
    default: return stage_kernel_hdc<R, RR, TY, false, IEQ_STAGE_NBUF, 0, TX>(hall, hod, a.conv);
  }
}

// Tile shapes of the single-grid launches, by code: 4 -> 64 x 4, 8 -> 64 x 8, 32 -> 32 x 8 (a
// narrower tile whose y halo is a smaller share of the loads).
// The narrow shape is a build knob (measured at 128^3 FP64: 32 x 8 and 16 x 16 equal, 32 x 16
// 5 % slower).
#ifndef IEQ_NARROW_TX
#define IEQ_NARROW_TX 32
#define IEQ_NARROW_TY 8
#endif
constexpr int shape_tx(int code) { return code == 32 ? IEQ_NARROW_TX : 64; }
constexpr int shape_ty(int code) { return code == 32 ? IEQ_NARROW_TY : code; }

template <typename R> const void* emf_kernel(int rr, int code) {
  if (code == 32) return rr == 1 ? emf_kernel_ty<R, 1, IEQ_NARROW_TY, 2, IEQ_NARROW_TX>()
                   : rr == 2 ? emf_kernel_ty<R, 2, IEQ_NARROW_TY, 2, IEQ_NARROW_TX>() : emf_kernel_ty<R, 3, IEQ_NARROW_TY, 2, IEQ_NARROW_TX>();
  if (code == 8) return rr == 1 ? emf_kernel_ty<R, 1, 8, 2>() : rr == 2 ? emf_kernel_ty<R, 2, 8, 2>() : emf_kernel_ty<R, 3, 8, 2>();
  return rr == 1 ? emf_kernel_ty<R, 1, 4, 2>() : rr == 2 ? emf_kernel_ty<R, 2, 4, 2>() : emf_kernel_ty<R, 3, 4, 2>();
}
template <typename R> const void* stage_kernel(int rr, int code, const FArgs<R>& a, bool hall, bool hod) {
  if (code == 32)
    return rr == 1   ? stage_kernel_ty<R, 1, IEQ_NARROW_TY, IEQ_NARROW_TX>(a, hall, hod)
           : rr == 2 ? stage_kernel_ty<R, 2, IEQ_NARROW_TY, IEQ_NARROW_TX>(a, hall, hod)
                     : stage_kernel_ty<R, 3, IEQ_NARROW_TY, IEQ_NARROW_TX>(a, hall, hod);
  if (code == 8)
    return rr == 1 ? stage_kernel_ty<R, 1, 8>(a, hall, hod) : rr == 2 ? stage_kernel_ty<R, 2, 8>(a, hall, hod) : stage_kernel_ty<R, 3, 8>(a, hall, hod);
  return rr == 1 ? stage_kernel_ty<R, 1, 4>(a, hall, hod) : rr == 2 ? stage_kernel_ty<R, 2, 4>(a, hall, hod) : stage_kernel_ty<R, 3, 4>(a, hall, hod);
}

template <typename R> int tile_grid(const FArgs<R>& a, int code) {
  const int tx = shape_tx(code), ty = shape_ty(code);
  return ((a.nx + tx - 1) / tx) * ((a.ny + ty - 1) / ty) * ((a.nz + a.zc - 1) / a.zc);
}
template <typename R> void launch_k(const void* k, int code, const FArgs<R>& a, hipStream_t s) {
  FArgs<R> arg = a;
  void* kargs[1] = {(void*)&arg};
  (void)hipLaunchKernel(k, dim3(tile_grid(a, code)), dim3(shape_tx(code), shape_ty(code)), kargs, 0, s);
}
template <typename R> void emf(int rr, int code, const FArgs<R>& a, hipStream_t s) { launch_k(emf_kernel<R>(rr, code), code, a, s); }
template <typename R> void stage(int rr, int code, const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  launch_k(stage_kernel<R>(rr, code, a, hall, hod), code, a, s);
}

int env_int(const char* k, int dflt) {
  const char* v = std::getenv(k);
  return v && *v ? std::atoi(v) : dflt;
}

}  // namespace

int hall_shape_for(long long nx, long long ny, long long nz, int real_bytes) {
  const int forced = env_int("IEQ_HALL_TILE", -1);
  if (forced == 0 || forced == 1) return forced;
  if (real_bytes == 4) return nx % 64 != 0 ? 1 : 0;  // 32 x 8 where 64-wide tiles would leave lanes idle
  return 3LL * 3 * nx * ny * nz * real_bytes <= 1000LL * 1000 * 1000 ? 1 : 0;
}

namespace {

// Central-row coefficients of the program's operator set (host copies of the tables).
#undef IEQ_TABLE_STORAGE
struct HostOp {
  int nb, w, wh, order;
  const double *d, *minv, *dho;
};

}  // namespace

template <typename R> void launch_emf(int rr, int ty, const FArgs<R>& a, hipStream_t s) { emf<R>(rr, ty, a, s); }
template <typename R> void launch_stage(int rr, int ty, const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  stage<R>(rr, ty, a, hall, hod, s);
}

// z-chunk for a grid: the smallest chunk (at least 4 planes) whose tile grid fits one round of
// resident workgroups of the stage kernel the arguments select, so no partial second round
// and the shortest possible z lead-in for the available parallelism.
static int zc_for(const void* kernel, int code, int nx, int ny, int nz) {
  const int tx = shape_tx(code), ty = shape_ty(code);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, tx * ty, 0) != hipSuccess || cus <= 0 || per_cu <= 0)
    return 8;
  const long long xy = (long long)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
  const long long chunks = std::max(1LL, (long long)cus * per_cu / std::max(1LL, xy));
  return std::max(4, (int)((nz + chunks - 1) / chunks));
}
template <typename R> int pick_zc(int rr, int ty, const FArgs<R>& a, bool hall, bool hod) {
  return zc_for(stage_kernel<R>(rr, ty, a, hall, hod), ty, a.nx, a.ny, a.nz);
}
// The same for the EMF kernel (fewer registers, so more resident workgroups and shorter chunks).
template <typename R> int pick_zc_emf(int rr, int ty, const FArgs<R>& a) { return zc_for(emf_kernel<R>(rr, ty), ty, a.nx, a.ny, a.nz); }
template void launch_emf<double>(int, int, const FArgs<double>&, hipStream_t);
template void launch_emf<float>(int, int, const FArgs<float>&, hipStream_t);
template void launch_stage<double>(int, int, const FArgs<double>&, bool, bool, hipStream_t);
template void launch_stage<float>(int, int, const FArgs<float>&, bool, bool, hipStream_t);
template int pick_zc<double>(int, int, const FArgs<double>&, bool, bool);
template int pick_zc<float>(int, int, const FArgs<float>&, bool, bool);
template int pick_zc_emf<double>(int, int, const FArgs<double>&);
template int pick_zc_emf<float>(int, int, const FArgs<float>&);

}  // namespace ieq

// host-side copies of the tables for coefficient extraction
namespace ieq_host_tables {
#undef IEQ_TABLES_H
#define IEQ_TABLE_STORAGE static const
#include "ieq_tables.h"
}  // namespace ieq_host_tables

namespace ieq {
namespace {

HostOp host_op(int ops) {
  namespace T = ieq_host_tables;
  switch (ops) {
    case 0: return {IEQ_2ND_NB, IEQ_2ND_W, IEQ_2ND_W_HOD, IEQ_2ND_ORDER, T::ieq_2nd_diff, T::ieq_2nd_minv, T::ieq_2nd_dho};
    case 1: return {IEQ_4TH_NB, IEQ_4TH_W, IEQ_4TH_W_HOD, IEQ_4TH_ORDER, T::ieq_4th_diff, T::ieq_4th_minv, T::ieq_4th_dho};
    case 2: return {IEQ_6TH_NB, IEQ_6TH_W, IEQ_6TH_W_HOD, IEQ_6TH_ORDER, T::ieq_6th_diff, T::ieq_6th_minv, T::ieq_6th_dho};
    case 3: return {IEQ_2ND_EXT_NB, IEQ_2ND_EXT_W, IEQ_2ND_EXT_W_HOD, IEQ_2ND_EXT_ORDER, T::ieq_2nd_ext_diff, T::ieq_2nd_ext_minv, T::ieq_2nd_ext_dho};
    case 4: return {IEQ_4TH_EXT_NB, IEQ_4TH_EXT_W, IEQ_4TH_EXT_W_HOD, IEQ_4TH_EXT_ORDER, T::ieq_4th_ext_diff, T::ieq_4th_ext_minv, T::ieq_4th_ext_dho};
    default: return {IEQ_6TH_EXT_NB, IEQ_6TH_EXT_W, IEQ_6TH_EXT_W_HOD, IEQ_6TH_EXT_ORDER, T::ieq_6th_ext_diff, T::ieq_6th_ext_minv, T::ieq_6th_ext_dho};
  }
}

}  // namespace

const double* host_lap_table(int opset, int lap, int* nbl, int* wl) {
  namespace T = ieq_host_tables;
  struct L { const double* t; int nb, w; };
  static const L tabs[6][3] = {
      {{T::ieq_2nd_lap_ws_d0, IEQ_2ND_LAP_WS_D0_NB, IEQ_2ND_LAP_WS_D0_W}, {T::ieq_2nd_lap_ns_d0, IEQ_2ND_LAP_NS_D0_NB, IEQ_2ND_LAP_NS_D0_W}, {T::ieq_2nd_lap_ws_ln, IEQ_2ND_LAP_WS_LN_NB, IEQ_2ND_LAP_WS_LN_W}},
      {{T::ieq_4th_lap_ws_d0, IEQ_4TH_LAP_WS_D0_NB, IEQ_4TH_LAP_WS_D0_W}, {T::ieq_4th_lap_ns_d0, IEQ_4TH_LAP_NS_D0_NB, IEQ_4TH_LAP_NS_D0_W}, {T::ieq_4th_lap_ws_ln, IEQ_4TH_LAP_WS_LN_NB, IEQ_4TH_LAP_WS_LN_W}},
      {{T::ieq_6th_lap_ws_d0, IEQ_6TH_LAP_WS_D0_NB, IEQ_6TH_LAP_WS_D0_W}, {T::ieq_6th_lap_ns_d0, IEQ_6TH_LAP_NS_D0_NB, IEQ_6TH_LAP_NS_D0_W}, {T::ieq_6th_lap_ws_ln, IEQ_6TH_LAP_WS_LN_NB, IEQ_6TH_LAP_WS_LN_W}},
      {{T::ieq_2nd_ext_lap_ws_d0, IEQ_2ND_EXT_LAP_WS_D0_NB, IEQ_2ND_EXT_LAP_WS_D0_W}, {nullptr, 0, 0}, {T::ieq_2nd_ext_lap_ws_ln, IEQ_2ND_EXT_LAP_WS_LN_NB, IEQ_2ND_EXT_LAP_WS_LN_W}},
      {{T::ieq_4th_ext_lap_ws_d0, IEQ_4TH_EXT_LAP_WS_D0_NB, IEQ_4TH_EXT_LAP_WS_D0_W}, {nullptr, 0, 0}, {T::ieq_4th_ext_lap_ws_ln, IEQ_4TH_EXT_LAP_WS_LN_NB, IEQ_4TH_EXT_LAP_WS_LN_W}},
      {{T::ieq_6th_ext_lap_ws_d0, IEQ_6TH_EXT_LAP_WS_D0_NB, IEQ_6TH_EXT_LAP_WS_D0_W}, {nullptr, 0, 0}, {T::ieq_6th_ext_lap_ws_ln, IEQ_6TH_EXT_LAP_WS_LN_NB, IEQ_6TH_EXT_LAP_WS_LN_W}},
  };
  if (opset < 0 || opset > 5 || lap < 0 || lap > 2) return nullptr;
  *nbl = tabs[opset][lap].nb;
  *wl = tabs[opset][lap].w;
  return tabs[opset][lap].t;
}

// Stencil half-width and coefficients of the central rows; returns -1 when the
// central row has non-zero entries beyond the dissipation half-width.
template <typename R> int central_coefs(int ops, double ho_factor, StencilCoef<R>& c) {
  HostOp o = host_op(ops);
  const int rr = o.wh - 1;
  const int cw = (o.w - 1) / 2;
  const double* row = o.d + o.nb * o.w;
  for (int i = 0; i < o.w; ++i)
    if (std::abs(i - cw) > rr && row[i] != 0.0) return -1;
  std::memset(&c, 0, sizeof(c));
  for (int k = 1; k <= rr; ++k) {
    c.dp[k] = (R)row[cw + k];
    c.dm[k] = (R)row[cw - k];
  }
  // w_m = minv * sum_{s+i=m} dho[s+wh-1] * dho[i] * (-1)^(order/2+1)   (artificial_dissipation.h:58-101)
  const double sgn = ((o.order / 2 + 1) % 2) ? -1.0 : 1.0;
  double w[9] = {0};
  const double* dr = o.dho + o.nb * o.wh;
  for (int s = -o.wh + 1; s <= 0; ++s)
    for (int i = 0; i < o.wh; ++i) w[i + s + rr] += o.minv[o.nb] * dr[s + o.wh - 1] * dr[i] * sgn;
  for (int k = 0; k <= rr; ++k) c.w[k] = (R)(ho_factor * w[rr + k]);
  return rr;
}

// Fused plan of one step (same arithmetic as the schedule, time kernels dropped).
bool fused_plan(const std::string& I, int hall, std::vector<StageSpec>& v, int& nf) {
  v.clear();
  auto E = [&](int y) {
    if (hall) v.push_back({3, -1, -1, -1, y, 0, 0, 0});
  };
  auto ST = [&](int out, int P, double cP, int Q, double cQ, int Y, double cD) { v.push_back({0, out, P, Q, Y, cP, cQ, cD}); };
  auto U2R = [&](int X, int O, double a, double b) { v.push_back({1, O, X, 2, -1, a, b, 0}); };
  if (I == "SSPRK33") {
    nf = 3;
    E(0); ST(1, 0, 1.0, -1, 0, 0, 1.0);
    E(1); ST(2, 0, 0.75, 1, 0.25, 1, 0.25);
    E(2); ST(0, 0, 1.0 / 3.0, 2, 2.0 / 3.0, 2, 2.0 / 3.0);
    return true;
  }
  if (I == "SSPRK104") {
    nf = 3;
    const int outs[10] = {1, 2, 1, 2, 1, 0, 1, 0, 1, 0}, ins[10] = {0, 1, 2, 1, 2, 1, 0, 1, 0, 1};
    for (int i = 0; i < 10; ++i) {
      if (i == 5) v.push_back({2, -1, -1, -1, -1, 0, 0, 0});
      E(ins[i]);
      if (i == 9) ST(0, 2, 1.0, 1, 0.6, 1, 0.6 / 6.0);
      else ST(outs[i], ins[i], 1.0, -1, 0, ins[i], 1.0 / 6.0);
    }
    return true;
  }
  if (I == "KennedyCarpenterLewis2R54C") {
    nf = 3;
    E(0); ST(2, -1, 0, -1, 0, 0, 1.0); U2R(0, 1, 0.22502245872571303, -0.17379315208537388);
    E(0); ST(2, -1, 0, -1, 0, 0, 1.0); U2R(1, 0, 0.5440433129514047, -0.1630884872250029);
    E(1); ST(2, -1, 0, -1, 0, 1, 1.0); U2R(0, 1, 0.14456824349399464, -0.5179208398863779);
    E(0); ST(2, -1, 0, -1, 0, 0, 1.0); U2R(1, 0, 0.7866643421983568, -0.19416305717199442);
    E(1); ST(0, 0, 1.0, -1, 0, 1, 0.34866717899927996);
    return true;
  }
  if (I == "CalvoFrancoRandez2R64") {
    nf = 3;
    const double B[5] = {0.10893125722541, 0.13201701492152, 0.38911623225517, -0.59203884581148, 0.47385028714844};
    const double A[5] = {0.17985400977138, 0.14081893152111, 0.08255631629428, 0.65804425034331, 0.31862993413251};
    for (int i = 0; i < 5; ++i) {
      E(i == 0 ? 0 : 1);
      ST(2, -1, 0, -1, 0, i == 0 ? 0 : 1, 1.0);
      U2R(0, 1, B[i], A[i]);
    }
    E(1); ST(0, 0, 1.0, -1, 0, 1, 0.48812405426094);
    return true;
  }
  if (I == "CarpenterKennedy2N54" || I == "ToulorgeDesmet2N84F") {
    nf = 2;
    static const double ckA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                  -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
    static const double ckB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                  1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                  2277821191437.0 / 14882151754819.0};
    static const double tdA[8] = {0.0, -0.5534431294501569, 0.01065987570203490, -0.5515812888932000,
                                  -1.885790377558741, -5.701295742793264, 2.113903965664793, -0.5339578826675280};
    static const double tdB[8] = {0.08037936882736950, 0.5388497458569843, 0.01974974409031960, 0.09911841297339970,
                                  0.7466920411064123, 1.679584245618894, 0.2433728067008188, 0.1422730459001373};
    const bool ck = I == "CarpenterKennedy2N54";
    const int n = ck ? 5 : 8;
    for (int i = 0; i < n; ++i) {
      E(0);
      v.push_back({4, 0, 1, -1, 0, ck ? ckA[i] : tdA[i], ck ? ckB[i] : tdB[i], 1.0});
    }
    return true;
  }
  return false;
}

template int central_coefs<double>(int, double, StencilCoef<double>&);
template int central_coefs<float>(int, double, StencilCoef<float>&);

namespace {

template <typename R>
int fused_step_t(const Params& p, const ieq_program_info& info, hipStream_t s, FusedWork** work, const char* integrator,
                 long long nsteps, void* const* args, const FusedOpts& opt, int* used, std::string& err) {
  const double* uniform = opt.uniform;
  std::vector<StageSpec> plan;
  int nf = 0;
  if (!fused_plan(integrator, p.hall, plan, nf)) return IEQ_OK;
  StencilCoef<R> coef;
  const int rr = central_coefs<R>(info.opset, p.ho_factor, coef);
  if (rr < 1 || rr > 3) return IEQ_OK;
  if (!offsets_fit(p.nx, p.ny, p.nz, sizeof(R))) return IEQ_OK;  // 32-bit byte offsets per component

  R* F[3] = {(R*)args[0], (R*)args[1], nf == 3 ? (R*)args[2] : nullptr};
  R* J = (R*)args[nf];
  const R* u = (const R*)args[nf + 1];
  const R* rho = (const R*)args[nf + 2];
  R* time = (R*)args[nf + 3];

  if (!*work) *work = new FusedWork();
  FusedWork* w = *work;
  double mn[4], mx[4];
  if (uniform) {  // the caller states the uniform values: no device check, no host synchronisation
    for (int c = 0; c < 4; ++c) mn[c] = mx[c] = uniform[c];
  } else {
    const int nblk = 256;
    if (!w->red || !w->red_host) {
      if (w->red) (void)hipFree(w->red);
      w->red = nullptr;
      if (hipMalloc(&w->red, nblk * 8 * sizeof(double)) != hipSuccess) return w->red = nullptr, err = "hipMalloc failed", IEQ_ERR_HIP;
      if (hipHostMalloc(&w->red_host, nblk * 8 * sizeof(double)) != hipSuccess) {
        (void)hipFree(w->red);
        w->red = nullptr;
        w->red_host = nullptr;
        return err = "hipHostMalloc failed", IEQ_ERR_HIP;
      }
    }
    k_uniform_check<R><<<nblk, 256, 0, s>>>(u, rho, info.num_nodes, info.stride, w->red);
    if (hipMemcpyAsync(w->red_host, w->red, nblk * 8 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return err = "uniformity check failed", IEQ_ERR_HIP;
    for (int c = 0; c < 4; ++c) {
      mn[c] = 1e300;
      mx[c] = -1e300;
      for (int b = 0; b < nblk; ++b) {
        mn[c] = std::fmin(mn[c], w->red_host[b * 8 + c]);
        mx[c] = std::fmax(mx[c], w->red_host[b * 8 + 4 + c]);
      }
    }
  }
  for (int c = 0; c < 4; ++c)
    if (!(mn[c] == mx[c])) return IEQ_OK;  // not uniform (or NaN): reference kernel sequence
  if (!(mn[3] != 0.0)) return IEQ_OK;

  const size_t vbytes = (size_t)3 * info.stride * sizeof(R);
  if (nf == 2 && w->scratch_bytes < vbytes) {
    if (w->scratch) (void)hipFree(w->scratch);
    w->scratch = nullptr;
    w->scratch_bytes = 0;
    if (hipMalloc(&w->scratch, vbytes) != hipSuccess) return err = "hipMalloc failed", IEQ_ERR_HIP;
    w->scratch_bytes = vbytes;
  }

  FArgs<R> a{};
  a.nx = p.nx;
  a.ny = p.ny;
  a.nz = p.nz;
  // Tile shape measured at 128^3: FP64 32 x 8 (2 waves per SIMD, so small workgroups; the
  // narrow tile keeps the y halo a smaller share of the loads than 64 x 4), FP32 64 x 8 where
  // 64 divides nx and 32 x 8 otherwise (no idle lanes: 96^3 SSPRK33 with convection, 33.6 ->
  // 28.5 us per stage).  The z-chunk is set below from the stage kernel's occupancy.
  const bool f64 = sizeof(R) == 8;
  const int ty_env = env_int("IEQ_FUSED_TY", f64 || p.nx % 64 != 0 ? 32 : 8);  // tile shape code: 4, 8 or 32 (32 x 8)
  const int ty = ty_env == 4 || ty_env == 32 ? ty_env : 8;
  a.S = info.stride;
  a.c = coef;
  const double h[3] = {p.dx, p.dy, p.dz};
  for (int d = 0; d < 3; ++d) {
    a.inv_h[d] = (R)(1.0 / h[d]);
    a.two_h[d] = (R)(2.0 / h[d]);
  }
  a.rho = (R)mn[3];
  for (int d = 0; d < 3; ++d) a.jsc[d] = (R)(1.0 / (h[d] * mn[3]));
  a.u[0] = (R)mn[0];
  a.u[1] = (R)mn[1];
  a.u[2] = (R)mn[2];
  a.conv = (mn[0] != 0.0 || mn[1] != 0.0 || mn[2] != 0.0);
  a.src_zero = p.form_source == 0;
  const bool hod = p.diss == 1;
  const R DT = (R)p.dt;

  R* b1 = F[0];
  R* b1alt = nf == 2 ? (R*)w->scratch : nullptr;
  a.zc = env_int("IEQ_FUSED_ZC", 0);
  if (a.zc < 1) {  // one round of resident workgroups for the plan's first stage kernel
    FArgs<R> t = a;
    t.zc = 1;
    for (const StageSpec& st : plan) {
      if (st.kind == 4) {
        t.mode = 1;
        t.Y = b1;
        t.b2 = F[1];
        break;
      }
      if (st.kind == 0) {
        R* FF[3] = {b1, F[1], F[2]};
        t.mode = 0;
        t.Y = FF[st.Y];
        t.P = st.P >= 0 ? FF[st.P] : nullptr;
        t.Q = st.Q >= 0 ? FF[st.Q] : nullptr;
        break;
      }
    }
    a.zc = pick_zc<R>(rr, ty, t, p.hall, hod);
  }
  // One kernel per Hall stage unless split on request or the 32-bit byte offsets of the staged
  // halo (which include the component) would not fit.
  // FP32 stages with convection run faster as the EMF kernel + stage kernel pair (96^3 SSPRK33:
  // 28.5 against 35.2 us per stage; 128^3: equal), every other Hall stage as one kernel.
  const int split = env_int("IEQ_FUSED_SPLIT", (!f64 && a.conv) ? 1 : 0);
  const bool one_kernel = p.hall && split == 0 && (long long)3 * info.stride * sizeof(R) < (1LL << 32);
  int zc_emf = 0, zc_hall = 0;
  if (p.hall && !one_kernel) {
    zc_emf = env_int("IEQ_EMF_ZC", 0);
    if (zc_emf < 1) zc_emf = pick_zc_emf<R>(rr, ty, a);
  }
  const int hshape = hall_shape_for(p.nx, p.ny, p.nz, (int)sizeof(R));
  if (one_kernel) {
    zc_hall = env_int("IEQ_FUSED_ZC", 0);
    if (zc_hall < 1) {
      FArgs<R> t = a;
      t.zc = 1;
      t.mode = plan.back().kind == 4 ? 1 : 0;
      t.Y = b1;
      t.b2 = F[1];
      zc_hall = pick_zc_hall<R>(rr, t, hod, hshape);
    }
  }
  const R* y_last = nullptr;  // field differentiated by the last stage (curl B / rho of it on request)
  for (long long step = 0; step < nsteps; ++step) {
    for (const StageSpec& st : plan) {
      a.P = nullptr;
      a.Q = nullptr;
      if (st.kind == 3 && one_kernel) continue;  // the stage kernel forms E itself
      if (st.kind == 3) {  // EMF of the field the next stage differentiates
        a.Y = st.Y == 0 ? b1 : F[st.Y];
        a.Eout = J;
        FArgs<R> ae = a;
        ae.zc = zc_emf;
        emf<R>(rr, ty, ae, s);
      } else if (st.kind == 0) {
        R* FF[3] = {b1, F[1], F[2]};
        a.Y = FF[st.Y];
        a.E = J;
        a.mode = 0;
        a.out = FF[st.out];
        a.P = st.P >= 0 ? FF[st.P] : nullptr;
        a.Q = st.Q >= 0 ? FF[st.Q] : nullptr;
        a.cP = (R)st.cP;
        a.cQ = (R)st.cQ;
        a.cD = (R)(st.cD * p.dt);
        y_last = a.Y;
        if (one_kernel) {
          FArgs<R> ah = a;
          ah.zc = zc_hall;
          launch_hall_stage<R>(rr, ah, hod, s, hshape);
        } else {
          stage<R>(rr, ty, a, p.hall, hod, s);
        }
      } else if (st.kind == 1) {
        uint32_t g[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz};
        R* FF[3] = {b1, F[1], F[2]};
        gen_update_2r<R>(p, s, g, FF[st.P], FF[st.out], FF[2], (R)st.cP, (R)st.cQ, time, -1.0);
      } else if (st.kind == 2) {
        uint32_t g[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz};
        gen_ssprk104_06<R>(p, s, g, b1, F[1], F[2], time);
      } else {  // 2N: b2 = A*b2 + DT*RHS(b1); b1' = b1 + B*b2 into the other buffer
        a.Y = b1;
        a.E = J;
        a.mode = 1;
        a.b2 = F[1];
        a.out = b1alt;
        a.cP = (R)st.cP;
        a.cQ = (R)st.cQ;
        a.cD = DT;
        y_last = a.Y;
        if (one_kernel) {
          FArgs<R> ah = a;
          ah.zc = zc_hall;
          launch_hall_stage<R>(rr, ah, hod, s, hshape);
        } else {
          stage<R>(rr, ty, a, p.hall, hod, s);
        }
        R* t = b1;
        b1 = b1alt;
        b1alt = t;
      }
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return err = std::string("fused launch: ") + hipGetErrorString(e), IEQ_ERR_HIP;
    }
  }
  if (opt.final_curl && p.hall && y_last) {
    // curlB_rho as the reference schedule leaves it: calc_curlB_rho of the last stage's input
    // (still intact: a register of the 3-register schemes, the ping-pong copy of the 2N ones)
    const uint32_t g[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz};
    gen_curl_rho<R>(p, info.opset, s, g, y_last, J, rho);
  }
  if (b1 != F[0]) {
    if (hipMemcpyAsync(F[0], b1, vbytes, hipMemcpyDeviceToDevice, s) != hipSuccess) return err = "copy-back failed", IEQ_ERR_HIP;
  }
  k_advance_time<R><<<1, 64, 0, s>>>(time, DT, nsteps);
  *used = 1;
  return IEQ_OK;
}

}  // namespace


int fused_step(const Params& p, const ieq_program_info& info, hipStream_t s, FusedWork** work, const char* integrator,
               long long nsteps, int variable_u, void* const* args, const FusedOpts& opt, int* used, std::string& err) {
  *used = 0;
  if (!p.periodic || variable_u || (p.diss != 0 && p.diss != 1)) return IEQ_OK;
  if (info.real_bytes == 8) return fused_step_t<double>(p, info, s, work, integrator, nsteps, args, opt, used, err);
  return fused_step_t<float>(p, info, s, work, integrator, nsteps, args, opt, used, err);
}

}  // namespace ieq
