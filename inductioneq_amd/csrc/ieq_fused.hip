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
// Two kernels per stage: k_emf writes E; k_stage reads B and E and applies the RK
// update.  Both march along z with register queues of the z-stencil and stage one
// xy-plane (plus halo) at a time in LDS for the x/y stencils.
#include <cmath>
#include <cstring>

#include "ieq_device.hpp"
#include "ieq_fused.hpp"
#include "ieq_generic.hpp"

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

// ------------------------------------------------------------------------------------
template <typename R> struct StencilCoef {
  R dp[4];  // central first-derivative row: coefficient at offset +k (k = 1..r)
  R dm[4];  // ... at offset -k
  R w[4];   // dissipation weights at offsets +-k (k = 0..r), HO_DISSIPATION_FACTOR folded in
};

template <typename R> struct FArgs {
  int nx, ny, nz, zc;
  long long S;
  StencilCoef<R> c;
  R inv_h[3];  // (R)(1/dx_d): Hall_volume scaling
  R two_h[3];  // (R)(2/dx_d): convec_volume scaling
  R h[3];      // (R)dx_d: derivative divisor of curl
  R rho;       // uniform density (J = curl B / rho)
  R u[3];      // uniform velocity
  int conv;    // velocity non-zero: evaluate the convection term
  int src_zero;
  const R* Y;  // field differentiated (B of the stage)
  const R* E;  // EMF (k_stage, Hall)
  R* Eout;     // EMF output (k_emf)
  int mode;    // 0: out = cP*P + cQ*Q + cD*RHS;  1 (2N): b2 = cP*b2 + cD*RHS, out = Y + cQ*b2
  R* out;
  const R* P;
  const R* Q;
  R* b2;
  R cP, cQ, cD;
};

__device__ inline int wrap(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

template <typename R, int RR> __device__ inline R deriv(const StencilCoef<R>& c, const R* v) {
  // v[RR + k] holds the value at offset k
  R acc = R(0);
#pragma unroll
  for (int k = RR; k >= 1; --k) acc = acc + c.dm[k] * v[RR - k];
#pragma unroll
  for (int k = 1; k <= RR; ++k) acc = acc + c.dp[k] * v[RR + k];
  return acc;
}

template <typename R, int RR> __device__ inline R hodsum(const StencilCoef<R>& c, const R* v) {
  R acc = c.w[0] * v[RR];
#pragma unroll
  for (int k = 1; k <= RR; ++k) acc = acc + c.w[k] * (v[RR - k] + v[RR + k]);
  return acc;
}

// EMF kernel: E = (curl_D B / rho) x B.
template <typename R, int RR, int TY>
__global__ __launch_bounds__(64 * TY) void k_emf(FArgs<R> a) {
  constexpr int LX = 64 + 2 * RR, LY = TY + 2 * RR, Q = 2 * RR + 1;
  __shared__ R sB[3][LY][LX];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * TY, z0 = blockIdx.z * a.zc;
  const int xg = x0 + tx, yg = y0 + ty;
  const bool active = xg < a.nx && yg < a.ny;
  const int xw = wrap(xg, a.nx), yw = wrap(yg, a.ny);
  const long long S = a.S, NX = a.nx, NXY = (long long)a.nx * a.ny;
  const int zend = min(z0 + a.zc, a.nz);
  const R* B = a.Y;

  R qx[Q], qy[Q];  // B_x, B_y along z at planes z-RR .. z+RR
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    long long id = xw + NX * yw + NXY * wrap(z0 - RR + j, a.nz);
    qx[j] = B[id];
    qy[j] = B[S + id];
  }
  for (int z = z0; z < zend; ++z) {
    {
      long long id = xw + NX * yw + NXY * wrap(z + RR, a.nz);
      qx[Q - 1] = B[id];
      qy[Q - 1] = B[S + id];
    }
    const long long pz = NXY * z;
    const long long ic = xw + NX * yw + pz;
    sB[0][ty + RR][tx + RR] = qx[RR];
    sB[1][ty + RR][tx + RR] = qy[RR];
    sB[2][ty + RR][tx + RR] = B[2 * S + ic];
    if (tx < RR) {
      long long id = wrap(x0 - RR + tx, a.nx) + NX * yw + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + RR][tx] = B[c * S + id];
    } else if (tx >= 64 - RR) {
      long long id = wrap(xg + RR, a.nx) + NX * yw + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + RR][tx + 2 * RR] = B[c * S + id];
    }
    if (ty < RR) {
      long long id = xw + NX * wrap(y0 - RR + ty, a.ny) + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty][tx + RR] = B[c * S + id];
    } else if (ty >= TY - RR) {
      long long id = xw + NX * wrap(yg + RR, a.ny) + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + 2 * RR][tx + RR] = B[c * S + id];
    }
    __syncthreads();
    if (active) {
      R vx[Q], vy[Q];
      // d/dx of B_y, B_z; d/dy of B_x, B_z; d/dz of B_x, B_y
#pragma unroll
      for (int k = 0; k < Q; ++k) vx[k] = sB[1][ty + RR][tx + k];
      R dxBy = deriv<R, RR>(a.c, vx) / a.h[0];
#pragma unroll
      for (int k = 0; k < Q; ++k) vx[k] = sB[2][ty + RR][tx + k];
      R dxBz = deriv<R, RR>(a.c, vx) / a.h[0];
#pragma unroll
      for (int k = 0; k < Q; ++k) vy[k] = sB[0][ty + k][tx + RR];
      R dyBx = deriv<R, RR>(a.c, vy) / a.h[1];
#pragma unroll
      for (int k = 0; k < Q; ++k) vy[k] = sB[2][ty + k][tx + RR];
      R dyBz = deriv<R, RR>(a.c, vy) / a.h[1];
      R dzBx = deriv<R, RR>(a.c, qx) / a.h[2];
      R dzBy = deriv<R, RR>(a.c, qy) / a.h[2];
      // J = curl B / rho  (SBP_operator.cl:126-135, kernel_time_integrator.cl:147)
      R Jx = (dyBz - dzBy) / a.rho, Jy = (dzBx - dxBz) / a.rho, Jz = (dxBy - dyBx) / a.rho;
      R Bx = qx[RR], By = qy[RR], Bz = sB[2][ty + RR][tx + RR];
      a.Eout[ic] = Jy * Bz - Jz * By;
      a.Eout[S + ic] = Jz * Bx - Jx * Bz;
      a.Eout[2 * S + ic] = Jx * By - Jy * Bx;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Q - 1; ++j) {
      qx[j] = qx[j + 1];
      qy[j] = qy[j + 1];
    }
  }
}

// Stage kernel: RHS = convection + (-curl E) + dissipation, then the RK update.
template <typename R, int RR, int TY, bool HALL, bool HOD>
__global__ __launch_bounds__(64 * TY) void k_stage(FArgs<R> a) {
  constexpr int LX = 64 + 2 * RR, LY = TY + 2 * RR, Q = 2 * RR + 1;
  constexpr int NE = HALL ? 3 : 1, LYE = HALL ? LY : 1, LXE = HALL ? LX : 1;
  __shared__ R sB[3][LY][LX];
  __shared__ R sE[NE][LYE][LXE];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * TY, z0 = blockIdx.z * a.zc;
  const int xg = x0 + tx, yg = y0 + ty;
  const bool active = xg < a.nx && yg < a.ny;
  const int xw = wrap(xg, a.nx), yw = wrap(yg, a.ny);
  const long long S = a.S, NX = a.nx, NXY = (long long)a.nx * a.ny;
  const int zend = min(z0 + a.zc, a.nz);
  const R* B = a.Y;
  const R* E = a.E;

  R qb[3][Q];
  R qe[2][HALL ? Q : 1];
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    long long id = xw + NX * yw + NXY * wrap(z0 - RR + j, a.nz);
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][j] = B[c * S + id];
    if constexpr (HALL) {
      qe[0][j] = E[id];
      qe[1][j] = E[S + id];
    }
  }
  for (int z = z0; z < zend; ++z) {
    {
      long long id = xw + NX * yw + NXY * wrap(z + RR, a.nz);
#pragma unroll
      for (int c = 0; c < 3; ++c) qb[c][Q - 1] = B[c * S + id];
      if constexpr (HALL) {
        qe[0][Q - 1] = E[id];
        qe[1][Q - 1] = E[S + id];
      }
    }
    const long long pz = NXY * z;
    const long long ic = xw + NX * yw + pz;
#pragma unroll
    for (int c = 0; c < 3; ++c) sB[c][ty + RR][tx + RR] = qb[c][RR];
    if constexpr (HALL) {
      sE[0][ty + RR][tx + RR] = qe[0][RR];
      sE[1][ty + RR][tx + RR] = qe[1][RR];
      sE[2][ty + RR][tx + RR] = E[2 * S + ic];
    }
    if (tx < RR) {
      long long id = wrap(x0 - RR + tx, a.nx) + NX * yw + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + RR][tx] = B[c * S + id];
      if constexpr (HALL)
        for (int c = 1; c < 3; ++c) sE[c][ty + RR][tx] = E[c * S + id];
    } else if (tx >= 64 - RR) {
      long long id = wrap(xg + RR, a.nx) + NX * yw + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + RR][tx + 2 * RR] = B[c * S + id];
      if constexpr (HALL)
        for (int c = 1; c < 3; ++c) sE[c][ty + RR][tx + 2 * RR] = E[c * S + id];
    }
    if (ty < RR) {
      long long id = xw + NX * wrap(y0 - RR + ty, a.ny) + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty][tx + RR] = B[c * S + id];
      if constexpr (HALL) {
        sE[0][ty][tx + RR] = E[id];
        sE[2][ty][tx + RR] = E[2 * S + id];
      }
    } else if (ty >= TY - RR) {
      long long id = xw + NX * wrap(yg + RR, a.ny) + pz;
      for (int c = 0; c < 3; ++c) sB[c][ty + 2 * RR][tx + RR] = B[c * S + id];
      if constexpr (HALL) {
        sE[0][ty + 2 * RR][tx + RR] = E[id];
        sE[2][ty + 2 * RR][tx + RR] = E[2 * S + id];
      }
    }
    __syncthreads();
    if (active) {
      R rhs[3] = {R(0), R(0), R(0)};
      R hod[3] = {R(0), R(0), R(0)};
      R der[3][3];  // der[d][c] = sum_i D_i B_c along d (no 1/dx)
      R v[Q];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = sB[c][ty + RR][tx + k];
        if constexpr (HOD) hod[c] = hodsum<R, RR>(a.c, v);
        der[0][c] = a.conv ? deriv<R, RR>(a.c, v) : R(0);
#pragma unroll
        for (int k = 0; k < Q; ++k) v[k] = sB[c][ty + k][tx + RR];
        if constexpr (HOD) hod[c] = hod[c] + hodsum<R, RR>(a.c, v);
        der[1][c] = a.conv ? deriv<R, RR>(a.c, v) : R(0);
        if constexpr (HOD) hod[c] = hod[c] + hodsum<R, RR>(a.c, qb[c]);
        der[2][c] = a.conv ? deriv<R, RR>(a.c, qb[c]) : R(0);
      }
      if (a.conv) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            rhs[c] = rhs[c] + a.two_h[d] * (R(0.5) * ((a.src_zero ? a.u[c] * der[d][d] : R(0)) - a.u[d] * der[d][c]));
      }
      if constexpr (HALL) {
        R ev[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) ev[k] = sE[1][ty + RR][tx + k];
        R dxEy = deriv<R, RR>(a.c, ev);
#pragma unroll
        for (int k = 0; k < Q; ++k) ev[k] = sE[2][ty + RR][tx + k];
        R dxEz = deriv<R, RR>(a.c, ev);
#pragma unroll
        for (int k = 0; k < Q; ++k) ev[k] = sE[0][ty + k][tx + RR];
        R dyEx = deriv<R, RR>(a.c, ev);
#pragma unroll
        for (int k = 0; k < Q; ++k) ev[k] = sE[2][ty + k][tx + RR];
        R dyEz = deriv<R, RR>(a.c, ev);
        R dzEx = deriv<R, RR>(a.c, qe[0]);
        R dzEy = deriv<R, RR>(a.c, qe[1]);
        rhs[0] = rhs[0] + (a.inv_h[1] * (-dyEz) + a.inv_h[2] * dzEy);
        rhs[1] = rhs[1] + (a.inv_h[0] * dxEz + a.inv_h[2] * (-dzEx));
        rhs[2] = rhs[2] + (a.inv_h[0] * (-dxEy) + a.inv_h[1] * dyEx);
      }
      if constexpr (HOD) {
#pragma unroll
        for (int c = 0; c < 3; ++c) rhs[c] = rhs[c] + hod[c];
      }
      if (a.mode == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          R r = a.cD * rhs[c];
          if (a.Q) r = a.cQ * a.Q[c * S + ic] + r;
          if (a.P) r = a.cP * a.P[c * S + ic] + r;
          a.out[c * S + ic] = r;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          R b2 = a.cP * a.b2[c * S + ic] + a.cD * rhs[c];
          a.b2[c * S + ic] = b2;
          a.out[c * S + ic] = qb[c][RR] + a.cQ * b2;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < Q - 1; ++j) {
#pragma unroll
      for (int c = 0; c < 3; ++c) qb[c][j] = qb[c][j + 1];
      if constexpr (HALL) {
        qe[0][j] = qe[0][j + 1];
        qe[1][j] = qe[1][j + 1];
      }
    }
  }
}

// min/max of u (3 components) and rho over the nodes: partials per block.
template <typename R> __global__ void k_uniform_check(const R* u, const R* rho, long long n, long long S, double* out) {
  double mn[4] = {1e300, 1e300, 1e300, 1e300}, mx[4] = {-1e300, -1e300, -1e300, -1e300};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v[4] = {(double)u[i], (double)u[S + i], (double)u[2 * S + i], (double)rho[i]};
    for (int c = 0; c < 4; ++c) {
      mn[c] = fmin(mn[c], v[c]);
      mx[c] = fmax(mx[c], v[c]);
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

constexpr int kTY = 4;

template <typename R, int RR>
void launch_emf(const FArgs<R>& a, hipStream_t s) {
  dim3 grid((a.nx + 63) / 64, (a.ny + kTY - 1) / kTY, (a.nz + a.zc - 1) / a.zc);
  k_emf<R, RR, kTY><<<grid, dim3(64, kTY), 0, s>>>(a);
}

template <typename R, int RR>
void launch_stage(const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  dim3 grid((a.nx + 63) / 64, (a.ny + kTY - 1) / kTY, (a.nz + a.zc - 1) / a.zc);
  dim3 blk(64, kTY);
  if (hall && hod) k_stage<R, RR, kTY, true, true><<<grid, blk, 0, s>>>(a);
  else if (hall) k_stage<R, RR, kTY, true, false><<<grid, blk, 0, s>>>(a);
  else if (hod) k_stage<R, RR, kTY, false, true><<<grid, blk, 0, s>>>(a);
  else k_stage<R, RR, kTY, false, false><<<grid, blk, 0, s>>>(a);
}

template <typename R> void emf(int rr, const FArgs<R>& a, hipStream_t s) {
  if (rr == 1) launch_emf<R, 1>(a, s);
  else if (rr == 2) launch_emf<R, 2>(a, s);
  else launch_emf<R, 3>(a, s);
}
template <typename R> void stage(int rr, const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  if (rr == 1) launch_stage<R, 1>(a, hall, hod, s);
  else if (rr == 2) launch_stage<R, 2>(a, hall, hod, s);
  else launch_stage<R, 3>(a, hall, hod, s);
}

// Central-row coefficients of the program's operator set (host copies of the tables).
#undef IEQ_TABLE_STORAGE
struct HostOp {
  int nb, w, wh, order;
  const double *d, *minv, *dho;
};

}  // namespace
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

struct StageSpec {
  int kind;  // 0: K_STAGE, 1: 2R update, 2: SSPRK104_06, 3: EMF, 4: 2N fused (a+b)
  int out, P, Q, Y;
  double cP, cQ, cD;  // K_STAGE: cD multiplies DT; 2R: (a, b); 2N: A (cP), B (cQ)
};

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

template <typename R>
int fused_step_t(const Params& p, const ieq_program_info& info, hipStream_t s, FusedWork** work, const char* integrator,
                 long long nsteps, void* const* args, int* used, std::string& err) {
  std::vector<StageSpec> plan;
  int nf = 0;
  if (!fused_plan(integrator, p.hall, plan, nf)) return IEQ_OK;
  StencilCoef<R> coef;
  const int rr = central_coefs<R>(info.opset, p.ho_factor, coef);
  if (rr < 1 || rr > 3) return IEQ_OK;

  R* F[3] = {(R*)args[0], (R*)args[1], nf == 3 ? (R*)args[2] : nullptr};
  R* J = (R*)args[nf];
  const R* u = (const R*)args[nf + 1];
  const R* rho = (const R*)args[nf + 2];
  R* time = (R*)args[nf + 3];

  if (!*work) *work = new FusedWork();
  FusedWork* w = *work;
  const int nblk = 256;
  if (!w->red) {
    if (hipMalloc(&w->red, nblk * 8 * sizeof(double)) != hipSuccess) return err = "hipMalloc failed", IEQ_ERR_HIP;
    if (hipHostMalloc(&w->red_host, nblk * 8 * sizeof(double)) != hipSuccess) return err = "hipHostMalloc failed", IEQ_ERR_HIP;
  }
  k_uniform_check<R><<<nblk, 256, 0, s>>>(u, rho, info.num_nodes, info.stride, w->red);
  if (hipMemcpyAsync(w->red_host, w->red, nblk * 8 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return err = "uniformity check failed", IEQ_ERR_HIP;
  double mn[4], mx[4];
  for (int c = 0; c < 4; ++c) {
    mn[c] = 1e300;
    mx[c] = -1e300;
    for (int b = 0; b < nblk; ++b) {
      mn[c] = std::fmin(mn[c], w->red_host[b * 8 + c]);
      mx[c] = std::fmax(mx[c], w->red_host[b * 8 + 4 + c]);
    }
  }
  for (int c = 0; c < 4; ++c)
    if (mn[c] != mx[c]) return IEQ_OK;  // not uniform: reference kernel sequence
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
  a.zc = p.nz >= 64 ? 16 : 8;
  a.S = info.stride;
  a.c = coef;
  const double h[3] = {p.dx, p.dy, p.dz};
  for (int d = 0; d < 3; ++d) {
    a.inv_h[d] = (R)(1.0 / h[d]);
    a.two_h[d] = (R)(2.0 / h[d]);
    a.h[d] = (R)h[d];
  }
  a.rho = (R)mn[3];
  a.u[0] = (R)mn[0];
  a.u[1] = (R)mn[1];
  a.u[2] = (R)mn[2];
  a.conv = (mn[0] != 0.0 || mn[1] != 0.0 || mn[2] != 0.0);
  a.src_zero = p.form_source == 0;
  const bool hod = p.diss == 1;
  const R DT = (R)p.dt;

  R* b1 = F[0];
  R* b1alt = nf == 2 ? (R*)w->scratch : nullptr;
  for (long long step = 0; step < nsteps; ++step) {
    for (const StageSpec& st : plan) {
      a.P = nullptr;
      a.Q = nullptr;
      if (st.kind == 3) {  // EMF of the field the next stage differentiates
        a.Y = st.Y == 0 ? b1 : F[st.Y];
        a.Eout = J;
        emf<R>(rr, a, s);
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
        stage<R>(rr, a, p.hall, hod, s);
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
        stage<R>(rr, a, p.hall, hod, s);
        R* t = b1;
        b1 = b1alt;
        b1alt = t;
      }
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return err = std::string("fused launch: ") + hipGetErrorString(e), IEQ_ERR_HIP;
    }
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
               long long nsteps, int variable_u, void* const* args, int* used, std::string& err) {
  *used = 0;
  if (!p.periodic || variable_u || (p.diss != 0 && p.diss != 1)) return IEQ_OK;
  if (info.real_bytes == 8) return fused_step_t<double>(p, info, s, work, integrator, nsteps, args, used, err);
  return fused_step_t<float>(p, info, s, work, integrator, nsteps, args, used, err);
}

}  // namespace ieq
