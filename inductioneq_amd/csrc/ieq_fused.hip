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
#include <cstdlib>
#include <cstring>

#include "ieq_device.hpp"
#include "ieq_fused.hpp"
#include "ieq_generic.hpp"
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

// EMF kernel: E = (curl_D B / rho) x B.
//
// Lead-aligned z-march: iteration `it` stages plane p = it + RR (the newest plane of the
// z-queue) into LDS together with its x/y halo and evaluates the x/y-derivative terms of
// J for that plane; the z-derivative terms of plane `it` come from the register queue and
// complete J(it).  The column value and the halo of a plane are therefore read in the same
// iteration, which is also when the neighbouring tiles (dispatched together to the same
// XCD) read those rows as their own column values, so the halo rows are L2 hits.
template <typename R, int RR, int TY, int NBUF = 2>
__device__ __forceinline__ void emf_tile(const FArgs<R>& a, int x0, int y0, int z0, R (*sB)[3][TY + 2 * RR][64 + 2 * RR]) {
  constexpr int Q = 2 * RR + 1;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const bool active = x0 + tx < a.nx && y0 + ty < a.ny;
  const unsigned RB = sizeof(R);
  const unsigned colb = (unsigned)(wrap(x0 + tx, a.nx) + a.nx * wrap(y0 + ty, a.ny)) * RB;  // own column
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;                                       // plane stride
  const int zend = min(z0 + a.zc, a.nz);
  const Comps<R, 3> B = comps<R, 3>(a.Y, a.S);
  PlaneStage<R, RR, TY, 3> hs;
  hs.init(tx, ty, x0, y0, a.nx, a.ny);

  R q[3][Q];       // own column: planes it-RR .. it+RR
  R ql[3];         // landing register: plane it+RR+1
  R jr[3][RR + 1];  // x/y parts of curl_D B for planes it .. it+RR: (d_y B_z, -d_x B_z, d_x B_y - d_y B_x)
  const int it0 = z0 - RR;
#pragma unroll
  for (int j = RR; j < Q; ++j) {  // planes z0-RR .. z0 (older slots are never read)
    const unsigned o = colb + (unsigned)wrap(it0 - RR + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c][j] = ldo(B.c[c], o);
  }
  hs.load(B, (unsigned)z0 * plb);
  for (int it = it0; it < zend; ++it) {
    const int p = it + RR;  // plane staged in this iteration
    const int buf = NBUF == 2 ? (it - it0) & 1 : 0;
    if constexpr (NBUF == 1) {
      if (it > it0) lds_barrier();
    }
    if (p < zend) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sB[buf][c][ty + RR][tx + RR] = q[c][Q - 1];
      hs.store(sB[buf], tx, ty);
    }
    // next iteration's loads, halo first (the memory counter retires in issue order)
    if (p + 1 < zend) hs.load(B, (unsigned)(p + 1) * plb);
    if (it + 1 < zend) {
      const unsigned o = colb + (unsigned)wrap(p + 1, a.nz) * plb;
#pragma unroll
      for (int c = 0; c < 3; ++c) ql[c] = ldo(B.c[c], o);
    }
    lds_barrier();
    if (active) {
      if (p < zend) {
        R vx[Q], vy[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) vx[k] = sB[buf][1][ty + RR][tx + k];
        const R dxBy = deriv<R, RR>(a.c, vx) / a.h[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) vx[k] = sB[buf][2][ty + RR][tx + k];
        const R dxBz = deriv<R, RR>(a.c, vx) / a.h[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) vy[k] = sB[buf][0][ty + k][tx + RR];
        const R dyBx = deriv<R, RR>(a.c, vy) / a.h[1];
#pragma unroll
        for (int k = 0; k < Q; ++k) vy[k] = sB[buf][2][ty + k][tx + RR];
        const R dyBz = deriv<R, RR>(a.c, vy) / a.h[1];
        jr[0][RR] = dyBz;
        jr[1][RR] = -dxBz;
        jr[2][RR] = dxBy - dyBx;
      }
      if (it >= z0) {
        const R dzBx = deriv<R, RR>(a.c, q[0]) / a.h[2];
        const R dzBy = deriv<R, RR>(a.c, q[1]) / a.h[2];
        // J = curl B / rho  (SBP_operator.cl:126-135, kernel_time_integrator.cl:147)
        const R Jx = (jr[0][0] - dzBy) / a.rho, Jy = (dzBx + jr[1][0]) / a.rho, Jz = jr[2][0] / a.rho;
        const R Bx = q[0][RR], By = q[1][RR], Bz = q[2][RR];
        const unsigned o = colb + (unsigned)it * plb;
        sto(a.Eout, o, Jy * Bz - Jz * By);
        sto(a.Eout + a.S, o, Jz * Bx - Jx * Bz);
        sto(a.Eout + 2 * a.S, o, Jx * By - Jy * Bx);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) q[c][j] = q[c][j + 1];
      q[c][Q - 1] = ql[c];
#pragma unroll
      for (int k = 0; k < RR; ++k) jr[c][k] = jr[c][k + 1];
    }
  }
}

template <typename R, int RR, int TY, int NBUF>
__global__ __launch_bounds__(64 * TY) void k_emf(FArgs<R> a) {
  __shared__ R sB[NBUF][3][TY + 2 * RR][64 + 2 * RR];
  int x0, y0, z0;
  tile_origin<TY>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  emf_tile<R, RR, TY, NBUF>(a, x0, y0, z0, sB);
}

// Persistent batched form: blocks stride over a tile list covering every member.
template <typename R, int RR, int TY>
__global__ __launch_bounds__(64 * TY) void k_emf_batched(const FArgs<R>* __restrict__ args, const int4* __restrict__ tiles,
                                                         int ntiles) {
  __shared__ R sB[2][3][TY + 2 * RR][64 + 2 * RR];
  const int chunk = (ntiles + kXcds - 1) / kXcds, slots = gridDim.x / kXcds;  // gridDim.x is a multiple of 8
  const int first = (blockIdx.x % kXcds) * chunk, last = min(first + chunk, ntiles);
  for (int t = first + (int)blockIdx.x / kXcds; t < last; t += slots) {
    const int4 d = tiles[t];
    emf_tile<R, RR, TY>(args[d.w], d.x, d.y, d.z, sB);
    lds_barrier();
  }
}

// Stage kernel: RHS = convection + (-curl E) + dissipation, then the RK update.
// Same pipelining as k_emf, with B (3 components) and E (3 in LDS, x/y in the queue).
#ifndef IEQ_STAGE_WAVES
#define IEQ_STAGE_WAVES 2
#endif
template <typename R, int RR, int TY, bool HALL, int NBUF = 2> struct StageLds {
  static constexpr int LX = 64 + 2 * RR, LY = TY + 2 * RR;
  static constexpr int NE = HALL ? 3 : 1, LYE = HALL ? LY : 1, LXE = HALL ? LX : 1;
  R sB[NBUF][3][LY][LX];
  R sE[NBUF][NE][LYE][LXE];
};

// Pointwise operands of the RK update at node ic: P (or b2 in 2N mode), Q, and the adjoint's G.
template <bool ADJ, typename R>
__device__ __forceinline__ void load_point(const FArgs<R>& a, unsigned o, R* pv, R* qv, R* gv) {
  const long long S = a.S;
  const R* P = a.mode == 0 ? a.P : a.b2;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pv[c] = P ? ldo(P + c * S, o) : R(0);
    qv[c] = (a.mode == 0 && a.Q) ? ldo(a.Q + c * S, o) : R(0);
  }
  if constexpr (ADJ) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gv[c] = a.G ? ldo(a.G + c * S, o) : R(0);
  }
}

// Stage kernel body, lead-aligned like emf_tile: iteration `it` stages plane p = it + RR of B
// (and E) into LDS and evaluates the x/y-derivative terms of the RHS of plane p; the output
// plane `it` adds the z-derivative terms from the register queues and applies the RK update.
template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV, bool ADJ = false, int NBUF = 2>
__device__ __forceinline__ void stage_tile(const FArgs<R>& a, int x0, int y0, int z0, StageLds<R, RR, TY, HALL, NBUF>& L) {
  constexpr int Q = 2 * RR + 1;
  auto& sB = L.sB;
  auto& sE = L.sE;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const bool active = x0 + tx < a.nx && y0 + ty < a.ny;
  const long long S = a.S;
  const unsigned RB = sizeof(R);
  const unsigned colb = (unsigned)(wrap(x0 + tx, a.nx) + a.nx * wrap(y0 + ty, a.ny)) * RB;  // own column
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;                                       // plane stride
  const int zend = min(z0 + a.zc, a.nz);
  const Comps<R, 3> B = comps<R, 3>(a.Y, S);
  const Comps<R, 3> E = comps<R, 3>(HALL ? a.E : a.Y, S);
  PlaneStage<R, RR, TY, 3> hsb;
  PlaneStage<R, RR, TY, 3> hse;
  hsb.init(tx, ty, x0, y0, a.nx, a.ny);
  if constexpr (HALL) hse.init(tx, ty, x0, y0, a.nx, a.ny);

  R qb[3][Q], qbl[3];                  // B of the own column: planes it-RR .. it+RR; landing it+RR+1
  R qe[2][HALL ? Q : 1], qel[2];       // E_x, E_y of the own column (z-derivatives of curl E)
  R ez = R(0);                         // E_z of the own column at the plane staged next
  R rx[3][RR + 1];                     // x/y terms of the RHS for planes it .. it+RR
  const int it0 = z0 - RR;
#pragma unroll
  for (int j = RR; j < Q; ++j) {  // planes z0-RR .. z0 (older slots are never read)
    const unsigned o = colb + (unsigned)wrap(it0 - RR + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][j] = ldo(B.c[c], o);
    if constexpr (HALL) {
      qe[0][j] = ldo(E.c[0], o);
      qe[1][j] = ldo(E.c[1], o);
    }
  }
  hsb.load(B, (unsigned)z0 * plb);
  if constexpr (HALL) {
    hse.load(E, (unsigned)z0 * plb);
    ez = ldo(E.c[2], colb + (unsigned)z0 * plb);
  }
  // pointwise RK operands of the output plane (pv, qv, gv) and of the next one (pn, qn, gn)
  R pv[3], qv[3], gv[ADJ ? 3 : 1], pn[3], qn[3], gn[ADJ ? 3 : 1];
  if (active) load_point<ADJ>(a, colb + (unsigned)z0 * plb, pn, qn, gn);
  for (int it = it0; it < zend; ++it) {
    const int p = it + RR;
    const int buf = NBUF == 2 ? (it - it0) & 1 : 0;
    if constexpr (NBUF == 1) {
      if (it > it0) lds_barrier();
    }
    if (p < zend) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sB[buf][c][ty + RR][tx + RR] = qb[c][Q - 1];
      hsb.store(sB[buf], tx, ty);
      if constexpr (HALL) {
        sE[buf][0][ty + RR][tx + RR] = qe[0][Q - 1];
        sE[buf][1][ty + RR][tx + RR] = qe[1][Q - 1];
        sE[buf][2][ty + RR][tx + RR] = ez;
        hse.store(sE[buf], tx, ty);
      }
    }
    if (it >= z0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pv[c] = pn[c];
        qv[c] = qn[c];
      }
      if constexpr (ADJ) {
#pragma unroll
        for (int c = 0; c < 3; ++c) gv[c] = gn[c];
      }
    }
    // next iteration's loads in the order they are consumed (the memory counter retires in
    // issue order and also counts stores): halo of plane p+1, pointwise operands of output
    // plane it+1, column plane p+1
    if (p + 1 < zend) {
      hsb.load(B, (unsigned)(p + 1) * plb);
      if constexpr (HALL) {
        hse.load(E, (unsigned)(p + 1) * plb);
        ez = ldo(E.c[2], colb + (unsigned)(p + 1) * plb);
      }
    }
    if (it + 1 > z0 && it + 1 < zend && active) load_point<ADJ>(a, colb + (unsigned)(it + 1) * plb, pn, qn, gn);
    if (it + 1 < zend) {
      const unsigned o = colb + (unsigned)wrap(p + 1, a.nz) * plb;
#pragma unroll
      for (int c = 0; c < 3; ++c) qbl[c] = ldo(B.c[c], o);
      if constexpr (HALL) {
        qel[0] = ldo(E.c[0], o);
        qel[1] = ldo(E.c[1], o);
      }
    }
    lds_barrier();
    if (active) {
      if (p < zend) {  // x/y terms of plane p
        R r[3] = {R(0), R(0), R(0)};
        R v[Q];
        if constexpr (CONV) {
          // uniform-velocity convection: sum_d (2/dx_d) 0.5 ([ZERO] u (D_d B)_d - u_d (D_d B))
#pragma unroll
          for (int d = 0; d < 2; ++d) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
#pragma unroll
              for (int k = 0; k < Q; ++k) v[k] = d == 0 ? sB[buf][c][ty + RR][tx + k] : sB[buf][c][ty + k][tx + RR];
              const R der = deriv<R, RR>(a.c, v);
              const R f = a.two_h[d] * R(0.5);
              r[c] = r[c] - f * (a.u[d] * der);
              if constexpr (ADJ) {
                if (a.src_adj) r[d] = r[d] + f * (a.u[c] * der);
              }
              if (c == d && a.src_zero && !(ADJ && a.src_adj)) {
#pragma unroll
                for (int e = 0; e < 3; ++e) r[e] = r[e] + f * (a.u[e] * der);
              }
            }
          }
        }
        if constexpr (HALL) {  // -curl_D E, x/y terms
#pragma unroll
          for (int k = 0; k < Q; ++k) v[k] = sE[buf][1][ty + RR][tx + k];
          const R dxEy = deriv<R, RR>(a.c, v);
#pragma unroll
          for (int k = 0; k < Q; ++k) v[k] = sE[buf][2][ty + RR][tx + k];
          const R dxEz = deriv<R, RR>(a.c, v);
#pragma unroll
          for (int k = 0; k < Q; ++k) v[k] = sE[buf][0][ty + k][tx + RR];
          const R dyEx = deriv<R, RR>(a.c, v);
#pragma unroll
          for (int k = 0; k < Q; ++k) v[k] = sE[buf][2][ty + k][tx + RR];
          const R dyEz = deriv<R, RR>(a.c, v);
          r[0] = r[0] + a.inv_h[1] * (-dyEz);
          r[1] = r[1] + a.inv_h[0] * dxEz;
          r[2] = r[2] + (a.inv_h[0] * (-dxEy) + a.inv_h[1] * dyEx);
        }
        if constexpr (HOD) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int k = 0; k < Q; ++k) v[k] = sB[buf][c][ty + RR][tx + k];
            R h = hodsum<R, RR>(a.c, v);
#pragma unroll
            for (int k = 0; k < Q; ++k) v[k] = sB[buf][c][ty + k][tx + RR];
            h = h + hodsum<R, RR>(a.c, v);
            r[c] = r[c] + h;
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) rx[c][RR] = r[c];
      }
      if (it >= z0) {  // output plane it: z terms and the RK update
        R rhs[3] = {rx[0][0], rx[1][0], rx[2][0]};
        if constexpr (CONV) {
          const R f = a.two_h[2] * R(0.5);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const R der = deriv<R, RR>(a.c, qb[c]);
            rhs[c] = rhs[c] - f * (a.u[2] * der);
            if constexpr (ADJ) {
              if (a.src_adj) rhs[2] = rhs[2] + f * (a.u[c] * der);
            }
            if (c == 2 && a.src_zero && !(ADJ && a.src_adj)) {
#pragma unroll
              for (int e = 0; e < 3; ++e) rhs[e] = rhs[e] + f * (a.u[e] * der);
            }
          }
        }
        if constexpr (HALL) {
          const R dzEx = deriv<R, RR>(a.c, qe[0]);
          const R dzEy = deriv<R, RR>(a.c, qe[1]);
          rhs[0] = rhs[0] + a.inv_h[2] * dzEy;
          rhs[1] = rhs[1] + a.inv_h[2] * (-dzEx);
        }
        if constexpr (HOD) {
#pragma unroll
          for (int c = 0; c < 3; ++c) rhs[c] = rhs[c] + hodsum<R, RR>(a.c, qb[c]);
        }
        if constexpr (ADJ) {
#pragma unroll
          for (int c = 0; c < 3; ++c) rhs[c] = rhs[c] - gv[c];
        }
        const unsigned oc = colb + (unsigned)it * plb;
        if (a.mode == 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            R o = a.cD * rhs[c];
            if constexpr (ADJ) o = a.cY * qb[c][RR] + o;
            if (a.Q) o = a.cQ * qv[c] + o;
            if (a.P) o = a.cP * pv[c] + o;
            sto(a.out + c * S, oc, o);
          }
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const R b2 = a.cP * pv[c] + a.cD * rhs[c];
            sto(a.b2 + c * S, oc, b2);
            sto(a.out + c * S, oc, qb[c][RR] + a.cQ * b2);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) qb[c][j] = qb[c][j + 1];
      qb[c][Q - 1] = qbl[c];
#pragma unroll
      for (int k = 0; k < RR; ++k) rx[c][k] = rx[c][k + 1];
    }
    if constexpr (HALL) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int j = 0; j < Q - 1; ++j) qe[c][j] = qe[c][j + 1];
        qe[c][Q - 1] = qel[c];
      }
    }
  }
}

// FP64: 2 waves per SIMD (the stage state needs ~200 VGPRs); FP32 fits 4 waves without spills
template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV, bool ADJ, int NBUF>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : IEQ_STAGE_WAVES, 8))) void k_stage(
    FArgs<R> a) {
  __shared__ StageLds<R, RR, TY, HALL, NBUF> L;
  int x0, y0, z0;
  tile_origin<TY>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  stage_tile<R, RR, TY, HALL, HOD, CONV, ADJ, NBUF>(a, x0, y0, z0, L);
}

template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV>
__global__ __launch_bounds__(64 * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : IEQ_STAGE_WAVES, 8))) void k_stage_batched(
    const FArgs<R>* __restrict__ args, const int4* __restrict__ tiles, int ntiles) {
  __shared__ StageLds<R, RR, TY, HALL> L;
  const int chunk = (ntiles + kXcds - 1) / kXcds, slots = gridDim.x / kXcds;  // gridDim.x is a multiple of 8
  const int first = (blockIdx.x % kXcds) * chunk, last = min(first + chunk, ntiles);
  for (int t = first + (int)blockIdx.x / kXcds; t < last; t += slots) {
    const int4 d = tiles[t];
    stage_tile<R, RR, TY, HALL, HOD, CONV>(args[d.w], d.x, d.y, d.z, L);
    lds_barrier();
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

template <typename R, int RR, int TY, int NBUF>
void launch_emf_ty(const FArgs<R>& a, hipStream_t s) {
  const int grid = ((a.nx + 63) / 64) * ((a.ny + TY - 1) / TY) * ((a.nz + a.zc - 1) / a.zc);
  k_emf<R, RR, TY, NBUF><<<grid, dim3(64, TY), 0, s>>>(a);
}

template <typename R, int RR, int TY, bool ADJ, int NBUF>
void launch_stage_hdc(const FArgs<R>& a, bool hall, bool hod, bool conv, hipStream_t s) {
  const int grid = ((a.nx + 63) / 64) * ((a.ny + TY - 1) / TY) * ((a.nz + a.zc - 1) / a.zc);
  dim3 blk(64, TY);
  if (conv) {
    if (hall && hod) k_stage<R, RR, TY, true, true, true, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else if (hall) k_stage<R, RR, TY, true, false, true, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else if (hod) k_stage<R, RR, TY, false, true, true, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else k_stage<R, RR, TY, false, false, true, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
  } else {
    if (hall && hod) k_stage<R, RR, TY, true, true, false, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else if (hall) k_stage<R, RR, TY, true, false, false, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else if (hod) k_stage<R, RR, TY, false, true, false, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
    else k_stage<R, RR, TY, false, false, false, ADJ, NBUF><<<grid, blk, 0, s>>>(a);
  }
}

template <typename R, int RR, int TY>
void launch_stage_ty(const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  if (a.G || a.src_adj || a.cY != R(0))  // transposed stage (adjoint sweeps): convection always evaluated
    launch_stage_hdc<R, RR, TY, true, 2>(a, hall, hod, true, s);
  else
    launch_stage_hdc<R, RR, TY, false, 2>(a, hall, hod, a.conv, s);
}

template <typename R, int TY> void emf_t(int rr, const FArgs<R>& a, hipStream_t s) {
  if (rr == 1) launch_emf_ty<R, 1, TY, 2>(a, s);
  else if (rr == 2) launch_emf_ty<R, 2, TY, 2>(a, s);
  else launch_emf_ty<R, 3, TY, 2>(a, s);
}
template <typename R> void emf(int rr, int ty, const FArgs<R>& a, hipStream_t s) {
  if (ty == 8) emf_t<R, 8>(rr, a, s);
  else emf_t<R, 4>(rr, a, s);
}
template <typename R> void stage(int rr, int ty, const FArgs<R>& a, bool hall, bool hod, hipStream_t s) {
  if (ty == 8) {
    if (rr == 1) launch_stage_ty<R, 1, 8>(a, hall, hod, s);
    else if (rr == 2) launch_stage_ty<R, 2, 8>(a, hall, hod, s);
    else launch_stage_ty<R, 3, 8>(a, hall, hod, s);
  } else {
    if (rr == 1) launch_stage_ty<R, 1, 4>(a, hall, hod, s);
    else if (rr == 2) launch_stage_ty<R, 2, 4>(a, hall, hod, s);
    else launch_stage_ty<R, 3, 4>(a, hall, hod, s);
  }
}

int env_int(const char* k, int dflt) {
  const char* v = std::getenv(k);
  return v && *v ? std::atoi(v) : dflt;
}

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
template void launch_emf<double>(int, int, const FArgs<double>&, hipStream_t);
template void launch_emf<float>(int, int, const FArgs<float>&, hipStream_t);
template void launch_stage<double>(int, int, const FArgs<double>&, bool, bool, hipStream_t);
template void launch_stage<float>(int, int, const FArgs<float>&, bool, bool, hipStream_t);

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
                 long long nsteps, void* const* args, int* used, std::string& err) {
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
  a.zc = env_int("IEQ_FUSED_ZC", 8);
  if (a.zc < 1) a.zc = 1;
  const int ty = env_int("IEQ_FUSED_TY", 8) == 4 ? 4 : 8;
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
        emf<R>(rr, ty, a, s);
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
        stage<R>(rr, ty, a, p.hall, hod, s);
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
        stage<R>(rr, ty, a, p.hall, hod, s);
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


// ---------------------------------------------------------------------------------
// Ensembles: many independent periodic solves stepped together.  Every fused kernel of
// an RK stage is one persistent launch whose blocks stride over a tile list of all
// members; each tile carries its member index and the member's stage arguments are an
// FArgs entry in device memory.  One step (two for the 2N schemes, whose ping-pong
// register alternates) is captured once in a hipGraph and replayed.

struct Ensemble {
  int real_bytes = 8;
  int rr = 3, hall = 0, hod = 0, conv = 0, src_zero = 0;
  StencilCoef<double> coef;  // central-row coefficients (converted to REAL per member)
  std::vector<ieq_member> m;
  std::vector<StageSpec> plan;
  int nf = 0, steps_per_graph = 1;
  std::vector<std::vector<void*>> b1_at;  // per unrolled stage: b1 pointers after swaps (2N)
  void* dargs = nullptr;                  // FArgs<R>[nstage_unrolled][nmembers]
  size_t args_bytes = 0;
  int4* dtiles = nullptr;
  int ntiles = 0, ty = 8, zc = 8, order = 0, grid_emf = 0, grid_stage = 0;
  bool tuned = false;
  bool use_graph = true;  // replay the captured step, or launch directly (chosen by timing)
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int ncu = 256;
  bool final_in_alt = false;  // 2N with an odd number of stages per graph
};

namespace {

template <typename R> FArgs<R> member_base(const Ensemble& e, const ieq_member& mb) {
  FArgs<R> a{};
  a.nx = mb.nx;
  a.ny = mb.ny;
  a.nz = mb.nz;
  a.zc = e.zc;
  a.S = (((long long)mb.nx * mb.ny * mb.nz) + 63) / 64 * 64;
  for (int k = 0; k < 4; ++k) {
    a.c.dp[k] = (R)e.coef.dp[k];
    a.c.dm[k] = (R)e.coef.dm[k];
    a.c.w[k] = (R)e.coef.w[k];
  }
  const double h[3] = {mb.dx, mb.dy, mb.dz};
  for (int d = 0; d < 3; ++d) {
    a.inv_h[d] = (R)(1.0 / h[d]);
    a.two_h[d] = (R)(2.0 / h[d]);
    a.h[d] = (R)h[d];
  }
  a.rho = (R)mb.rho;
  a.u[0] = (R)mb.ux;
  a.u[1] = (R)mb.uy;
  a.u[2] = (R)mb.uz;
  a.conv = e.conv;
  a.src_zero = e.src_zero;
  return a;
}

// FArgs of every unrolled stage entry for every member (zc taken from e.zc).
template <typename R> std::vector<FArgs<R>> ensemble_args(Ensemble& e) {
  const int nm = (int)e.m.size();
  std::vector<FArgs<R>> out;
  std::vector<R*> b1(nm), b1alt(nm);
  for (int i = 0; i < nm; ++i) {
    b1[i] = (R*)e.m[i].b1;
    b1alt[i] = (R*)e.m[i].b3;
  }
  for (int step = 0; step < e.steps_per_graph; ++step)
    for (const StageSpec& st : e.plan) {
      for (int i = 0; i < nm; ++i) {
        const ieq_member& mb = e.m[i];
        R* F[3] = {b1[i], (R*)mb.b2, (R*)mb.b3};
        FArgs<R> a = member_base<R>(e, mb);
        if (st.kind == 3) {
          a.Y = st.Y == 0 ? b1[i] : F[st.Y];
          a.Eout = (R*)mb.emf;
        } else if (st.kind == 0) {
          a.Y = F[st.Y];
          a.E = (R*)mb.emf;
          a.mode = 0;
          a.out = F[st.out];
          a.P = st.P >= 0 ? F[st.P] : nullptr;
          a.Q = st.Q >= 0 ? F[st.Q] : nullptr;
          a.cP = (R)st.cP;
          a.cQ = (R)st.cQ;
          a.cD = (R)(st.cD * mb.dt);
        } else {  // 2N
          a.Y = b1[i];
          a.E = (R*)mb.emf;
          a.mode = 1;
          a.b2 = (R*)mb.b2;
          a.out = b1alt[i];
          a.cP = (R)st.cP;
          a.cQ = (R)st.cQ;
          a.cD = (R)mb.dt;
        }
        out.push_back(a);
      }
      if (st.kind == 4)
        for (int i = 0; i < nm; ++i) std::swap(b1[i], b1alt[i]);
    }
  e.final_in_alt = b1[0] != (R*)e.m[0].b1;
  return out;
}

std::vector<int4> ensemble_tiles(const Ensemble& e, int ty, int zc, int order) {
  std::vector<int> idx(e.m.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  if (order == 1)
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      return (long long)e.m[a].nx * e.m[a].ny * e.m[a].nz > (long long)e.m[b].nx * e.m[b].ny * e.m[b].nz;
    });
  std::vector<int4> t;
  for (int i : idx) {
    const ieq_member& mb = e.m[i];
    for (int z = 0; z < mb.nz; z += zc)
      for (int y = 0; y < mb.ny; y += ty)
        for (int x = 0; x < mb.nx; x += 64) t.push_back(make_int4(x, y, z, i));
  }
  return t;
}

template <typename R, int RR, int TY> const void* ens_kernel_t(const Ensemble& e, int kind) {
  if (kind == 3) return reinterpret_cast<const void*>(&k_emf_batched<R, RR, TY>);
  const bool h = e.hall, d = e.hod;
  if (e.conv) {
    if (h && d) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, true, true, true>);
    if (h) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, true, false, true>);
    if (d) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, false, true, true>);
    return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, false, false, true>);
  }
  if (h && d) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, true, true, false>);
  if (h) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, true, false, false>);
  if (d) return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, false, true, false>);
  return reinterpret_cast<const void*>(&k_stage_batched<R, RR, TY, false, false, false>);
}

// Kernel of an entry (kind 3: EMF, otherwise the stage kernel) for the ensemble's settings.
template <typename R> const void* ens_kernel(const Ensemble& e, int ty, int kind) {
  if (ty == 4) {
    if (e.rr == 1) return ens_kernel_t<R, 1, 4>(e, kind);
    if (e.rr == 2) return ens_kernel_t<R, 2, 4>(e, kind);
    return ens_kernel_t<R, 3, 4>(e, kind);
  }
  if (e.rr == 1) return ens_kernel_t<R, 1, 8>(e, kind);
  if (e.rr == 2) return ens_kernel_t<R, 2, 8>(e, kind);
  return ens_kernel_t<R, 3, 8>(e, kind);
}

template <typename R> hipError_t ens_launch_steps(Ensemble& e, hipStream_t s) {
  const FArgs<R>* base = (const FArgs<R>*)e.dargs;
  const size_t nm = e.m.size();
  size_t k = 0;
  for (int step = 0; step < e.steps_per_graph; ++step)
    for (const StageSpec& st : e.plan) {
      const FArgs<R>* a = base + k * nm;
      ++k;
      void* kargs[3] = {(void*)&a, (void*)&e.dtiles, (void*)&e.ntiles};
      const int grid = st.kind == 3 ? e.grid_emf : e.grid_stage;
      hipError_t r = hipLaunchKernel(ens_kernel<R>(e, e.ty, st.kind), dim3(grid), dim3(64, e.ty), kargs, 0, s);
      if (r != hipSuccess) return r;
    }
  return hipSuccess;
}

template <typename R> int ens_configure(Ensemble& e, int ty, int zc, int order, std::string& err) {
  e.ty = ty;
  e.zc = zc;
  e.order = order;
  std::vector<FArgs<R>> args = ensemble_args<R>(e);
  const size_t ab = args.size() * sizeof(FArgs<R>);
  if (e.args_bytes < ab) {
    if (e.dargs) (void)hipFree(e.dargs);
    e.dargs = nullptr;
    if (hipMalloc(&e.dargs, ab) != hipSuccess) return err = "hipMalloc failed", -4;
    e.args_bytes = ab;
  }
  if (hipMemcpyAsync(e.dargs, args.data(), ab, hipMemcpyHostToDevice, e.stream) != hipSuccess) return err = "copy failed", -4;
  std::vector<int4> tiles = ensemble_tiles(e, ty, zc, order);
  if (e.dtiles) (void)hipFree(e.dtiles);
  e.dtiles = nullptr;
  if (hipMalloc(&e.dtiles, tiles.size() * sizeof(int4)) != hipSuccess) return err = "hipMalloc failed", -4;
  if (hipMemcpyAsync(e.dtiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, e.stream) != hipSuccess)
    return err = "copy failed", -4;
  e.ntiles = (int)tiles.size();
  // persistent grids: as many blocks as can be resident at once (registers and LDS bound it)
  int occ_emf = 1, occ_stage = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_emf, ens_kernel<R>(e, ty, 3), 64 * ty, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_stage, ens_kernel<R>(e, ty, 0), 64 * ty, 0) != hipSuccess)
    return err = "occupancy query failed", -4;
  // multiples of 8 (one equal share of resident slots per XCD), no more than the tiles need
  const int need = (e.ntiles + 7) / 8 * 8;
  e.grid_emf = std::min(need, (e.ncu * std::max(1, occ_emf) + 7) / 8 * 8);
  e.grid_stage = std::min(need, (e.ncu * std::max(1, occ_stage) + 7) / 8 * 8);
  if (hipStreamSynchronize(e.stream) != hipSuccess) return err = "sync failed", -4;
  return 0;
}

template <typename R> int ens_capture(Ensemble& e, std::string& err) {
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  if (e.graph) (void)hipGraphDestroy(e.graph);
  e.exec = nullptr;
  e.graph = nullptr;
  if (hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return err = "capture begin failed", -4;
  hipError_t r = ens_launch_steps<R>(e, e.stream);
  if (hipStreamEndCapture(e.stream, &e.graph) != hipSuccess || r != hipSuccess) return err = "capture failed", -4;
  if (hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0) != hipSuccess) return err = "graph instantiate failed", -4;
  return 0;
}

size_t member_bytes(const Ensemble& e, const ieq_member& mb) {
  return (size_t)3 * ((((size_t)mb.nx * mb.ny * mb.nz) + 63) / 64 * 64) * e.real_bytes;
}

template <typename R> int ens_tune(Ensemble& e, std::string& err) {
  // save the registers the trial steps overwrite
  size_t total = 0;
  for (auto& mb : e.m) total += 3 * member_bytes(e, mb);
  char* save = nullptr;
  if (hipMalloc(&save, total) != hipSuccess) return err = "hipMalloc failed (tuning snapshot)", -4;
  size_t off = 0;
  for (auto& mb : e.m) {
    void* f[3] = {mb.b1, mb.b2, mb.b3};
    for (int j = 0; j < 3; ++j) {
      (void)hipMemcpyAsync(save + off, f[j], member_bytes(e, mb), hipMemcpyDeviceToDevice, e.stream);
      off += member_bytes(e, mb);
    }
  }
  float best = 1e30f;
  int bty = 8, bzc = 8, bord = 1;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  const int tys[2] = {4, 8}, zcs[3] = {4, 8, 16};
  for (int ty : tys)
    for (int zc : zcs)
      for (int ord = 0; ord < 2; ++ord) {
        int rc = ens_configure<R>(e, ty, zc, ord, err);
        if (rc) return rc;
        hipError_t r = ens_launch_steps<R>(e, e.stream);  // warm-up
        (void)hipEventRecord(t0, e.stream);
        if (r == hipSuccess) r = ens_launch_steps<R>(e, e.stream);
        if (r == hipSuccess) r = ens_launch_steps<R>(e, e.stream);
        (void)hipEventRecord(t1, e.stream);
        (void)hipEventSynchronize(t1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (r != hipSuccess) return err = std::string("tuning launch failed: ") + hipGetErrorString(r), -4;
        if (ms < best) {
          best = ms;
          bty = ty;
          bzc = zc;
          bord = ord;
        }
      }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  off = 0;
  for (auto& mb : e.m) {
    void* f[3] = {mb.b1, mb.b2, mb.b3};
    for (int j = 0; j < 3; ++j) {
      (void)hipMemcpyAsync(f[j], save + off, member_bytes(e, mb), hipMemcpyDeviceToDevice, e.stream);
      off += member_bytes(e, mb);
    }
  }
  (void)hipStreamSynchronize(e.stream);
  int rc = ens_configure<R>(e, bty, bzc, bord, err);
  if (rc) return rc;
  rc = ens_capture<R>(e, err);
  if (rc) return rc;
  {  // graph replay against direct launches of the same period (fields restored afterwards)
    hipEvent_t u0, u1, u2;
    (void)hipEventCreate(&u0);
    (void)hipEventCreate(&u1);
    (void)hipEventCreate(&u2);
    off = 0;
    (void)hipGraphLaunch(e.exec, e.stream);  // warm-up
    (void)hipEventRecord(u0, e.stream);
    for (int k = 0; k < 2; ++k) (void)hipGraphLaunch(e.exec, e.stream);
    (void)hipEventRecord(u1, e.stream);
    for (int k = 0; k < 2; ++k) (void)ens_launch_steps<R>(e, e.stream);
    (void)hipEventRecord(u2, e.stream);
    (void)hipEventSynchronize(u2);
    float tg = 0.f, td = 0.f;
    (void)hipEventElapsedTime(&tg, u0, u1);
    (void)hipEventElapsedTime(&td, u1, u2);
    e.use_graph = tg <= td;
    (void)hipEventDestroy(u0);
    (void)hipEventDestroy(u1);
    (void)hipEventDestroy(u2);
    for (auto& mb : e.m) {
      void* f[3] = {mb.b1, mb.b2, mb.b3};
      for (int j = 0; j < 3; ++j) {
        (void)hipMemcpyAsync(f[j], save + off, member_bytes(e, mb), hipMemcpyDeviceToDevice, e.stream);
        off += member_bytes(e, mb);
      }
    }
    (void)hipStreamSynchronize(e.stream);
  }
  (void)hipFree(save);
  e.tuned = true;
  return hipGetLastError() == hipSuccess ? 0 : (err = "ensemble tuning failed", -4);
}

}  // namespace

int ensemble_create(const Params& p, const ieq_program_info& info, const ieq_member* members, int n, const char* integrator,
                    Ensemble** out, std::string& err) {
  *out = nullptr;
  if (!p.periodic) return err = "ensembles need a periodic program", IEQ_ERR_UNSUPPORTED;
  if (p.diss != 0 && p.diss != 1) return err = "ensembles support high-order dissipation only", IEQ_ERR_UNSUPPORTED;
  std::string I(integrator);
  if (I != "SSPRK33" && I != "CarpenterKennedy2N54" && I != "ToulorgeDesmet2N84F")
    return err = "ensembles support SSPRK33, CarpenterKennedy2N54 and ToulorgeDesmet2N84F", IEQ_ERR_UNSUPPORTED;
  Ensemble* e = new Ensemble();
  e->real_bytes = info.real_bytes;
  if (!fused_plan(I, p.hall, e->plan, e->nf)) {
    delete e;
    return err = "unknown integrator", IEQ_ERR_ARGS;
  }
  e->rr = central_coefs<double>(info.opset, p.ho_factor, e->coef);
  if (e->rr < 1 || e->rr > 3) {
    delete e;
    return err = "ensembles need an operator whose central rows are at most 7 wide", IEQ_ERR_UNSUPPORTED;
  }
  e->hall = p.hall;
  e->hod = p.diss == 1;
  e->src_zero = p.form_source == 0;
  for (int i = 0; i < n; ++i) {
    const ieq_member& mb = members[i];
    if (mb.nx < 1 || mb.ny < 1 || mb.nz < 1 || !mb.b1 || !mb.b2 || !mb.b3 || !mb.emf || !(mb.rho != 0.0) || !(mb.dx > 0) ||
        !(mb.dy > 0) || !(mb.dz > 0)) {
      delete e;
      return err = "invalid member " + std::to_string(i), IEQ_ERR_ARGS;
    }
    if (!offsets_fit(mb.nx, mb.ny, mb.nz, info.real_bytes)) {
      delete e;
      return err = "member " + std::to_string(i) + " is too large (one component must stay below 2 GiB)", IEQ_ERR_UNSUPPORTED;
    }
    if (mb.ux != 0.0 || mb.uy != 0.0 || mb.uz != 0.0) e->conv = 1;
    e->m.push_back(mb);
  }
  e->steps_per_graph = (e->plan.size() > 0 && e->plan.back().kind == 4 && (e->plan.size() / (p.hall ? 2 : 1)) % 2) ? 2 : 1;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) e->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return err = "stream creation failed", IEQ_ERR_HIP;
  }
  *out = e;
  return IEQ_OK;
}

int ensemble_step(Ensemble* e, long long nsteps, double* seconds, std::string& err) {
  int rc = 0;
  if (!e->tuned) {
    rc = e->real_bytes == 8 ? ens_tune<double>(*e, err) : ens_tune<float>(*e, err);
    if (rc) return IEQ_ERR_HIP;
  }
  if (nsteps % e->steps_per_graph) return err = "this integrator advances the ensemble in pairs of steps", IEQ_ERR_ARGS;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  (void)hipEventRecord(t0, e->stream);
  for (long long k = 0; k < nsteps; k += e->steps_per_graph) {
    if (e->use_graph) {
      if (hipGraphLaunch(e->exec, e->stream) != hipSuccess) return err = "graph launch failed", IEQ_ERR_HIP;
    } else {
      hipError_t r = e->real_bytes == 8 ? ens_launch_steps<double>(*e, e->stream) : ens_launch_steps<float>(*e, e->stream);
      if (r != hipSuccess) return err = std::string("ensemble launch failed: ") + hipGetErrorString(r), IEQ_ERR_HIP;
    }
  }
  (void)hipEventRecord(t1, e->stream);
  if (hipEventSynchronize(t1) != hipSuccess) return err = "ensemble step failed", IEQ_ERR_HIP;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (seconds) *seconds = ms * 1e-3;
  return hipGetLastError() == hipSuccess ? IEQ_OK : (err = "ensemble kernel failed", IEQ_ERR_HIP);
}

void ensemble_config(const Ensemble* e, int* ty, int* zc, int* order) {
  *ty = e->ty;
  *zc = e->zc;
  *order = e->order + (e->use_graph ? 2 : 0);
}

void ensemble_release(Ensemble* e) {
  if (!e) return;
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->dargs) (void)hipFree(e->dargs);
  if (e->dtiles) (void)hipFree(e->dtiles);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int fused_step(const Params& p, const ieq_program_info& info, hipStream_t s, FusedWork** work, const char* integrator,
               long long nsteps, int variable_u, void* const* args, int* used, std::string& err) {
  *used = 0;
  if (!p.periodic || variable_u || (p.diss != 0 && p.diss != 1)) return IEQ_OK;
  if (info.real_bytes == 8) return fused_step_t<double>(p, info, s, work, integrator, nsteps, args, used, err);
  return fused_step_t<float>(p, info, s, work, integrator, nsteps, args, used, err);
}

}  // namespace ieq
