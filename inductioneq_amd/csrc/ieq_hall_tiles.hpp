// One kernel per RK stage for the periodic Hall problem: the EMF E = (curl_D B / rho) x B is
// formed inside the stage kernel instead of making a round trip through device memory.
//
// A workgroup owns a TX x TY tile of output columns and marches along z.  The stage needs E
// on the tile's x/y stencil halo (half-width RR) and on the z-stencil of its own columns, and
// E at a halo node needs curl_D B there, i.e. B up to 2*RR outside the tile.  Per z-plane:
//   * an LDS plane of B with a 2*RR halo (x/y derivatives of B for J, and the x/y stencils of
//     the dissipation and convection terms);
//   * an LDS plane of E with an RR halo (x/y derivatives of E);
//   * register windows along z: B of the own column (J's z-derivatives, dissipation and
//     convection z-terms, the RK update), B_x/B_y of one E-halo column per thread (J's
//     z-derivatives there), E_x/E_y of the own column (z-derivatives of E).
// Iteration `o` receives B plane q = o + 2*RR, evaluates E and every RHS term except the
// z-derivatives of E on plane pe = o + RR, and completes and writes output plane o.
// Loads follow the memory-counter discipline of the other z-marching kernels: per iteration
// the stores of the finished plane, then the staging loads of the next B plane's halo, then
// the column and pointwise loads PD planes ahead, unconditionally on every thread.
#pragma once

#include <type_traits>

#include "ieq_stencil.hpp"

namespace ieq {
namespace {

#ifndef IEQ_HALL_HSKIP
#define IEQ_HALL_HSKIP 0
#endif
// E at the E-halo columns only on the planes the chunk outputs (1), or on every plane (0).
#ifndef IEQ_HALL_ZSKIP
#define IEQ_HALL_ZSKIP 1
#endif
// Extra LDS row elements (bank spreading of the scattered halo accesses); 0 by default.
#ifndef IEQ_HALL_PADB
#define IEQ_HALL_PADB 0
#endif
#ifndef IEQ_HALL_PADE
#define IEQ_HALL_PADE 0
#endif
template <int RR, int TX, int TY> struct HallGeom {
  static constexpr int NT = TX * TY;
  static constexpr int BX = TX + 4 * RR + IEQ_HALL_PADB, BY = TY + 4 * RR;  // B plane with a 2*RR halo
  static constexpr int EX = TX + 2 * RR + IEQ_HALL_PADE, EY = TY + 2 * RR;  // E plane with an RR halo
  static constexpr int NXH = RR * TY, NYH = RR * TX;        // E-halo columns per x side / per y side
  static constexpr int NH = 2 * (NXH + NYH);                // E-halo columns
  // B nodes beyond the E region that the x/y derivatives of J at E-halo nodes reach: the four
  // RR x RR corners (all components), the outer x strips (B_y, B_z, differentiated along x)
  // and the outer y strips (B_x, B_z, differentiated along y)
  static constexpr int NCORNER = 4 * RR * RR, NXO = 2 * RR * TY, NYO = 2 * RR * TX;
  static constexpr int NITEMS = 3 * NCORNER + 2 * NXO + 2 * NYO;
  static constexpr int K = (NITEMS + NT - 1) / NT;  // staging loads per thread and plane
  static_assert(NH <= NT, "one E-halo column per thread");
};

// NBUF = 1: single planes, two barriers per z-step; NBUF = 2: double-buffered planes, one barrier.
template <typename R, int RR, int TX, int TY, int NBUF> struct HallLds {
  using G = HallGeom<RR, TX, TY>;
  R sB[NBUF][3][G::BY][G::BX];
  R sE[NBUF][3][G::EY][G::EX];
};

// E-halo column h of a tile: offset (lx, ly) from the tile origin.
template <int RR, int TX, int TY> __device__ inline void halo_col(int h, int& lx, int& ly) {
  using G = HallGeom<RR, TX, TY>;
  if (h < G::NXH) {
    lx = -RR + h % RR;
    ly = h / RR;
  } else if (h < 2 * G::NXH) {
    h -= G::NXH;
    lx = TX + h % RR;
    ly = h / RR;
  } else if (h < 2 * G::NXH + G::NYH) {
    h -= 2 * G::NXH;
    lx = h % TX;
    ly = -RR + h / TX;
  } else {
    h -= 2 * G::NXH + G::NYH;
    lx = h % TX;
    ly = TY + h / TX;
  }
}

// Staging item i: offset (lx, ly) from the tile origin and component c.
template <int RR, int TX, int TY> __device__ inline void stage_item(int i, int& lx, int& ly, int& c) {
  using G = HallGeom<RR, TX, TY>;
  if (i < 3 * G::NCORNER) {
    c = i / G::NCORNER;
    const int j = i % G::NCORNER, corner = j / (RR * RR), k = j % (RR * RR);
    lx = (corner & 1) ? TX + k % RR : -RR + k % RR;
    ly = (corner & 2) ? TY + k / RR : -RR + k / RR;
  } else if (i < 3 * G::NCORNER + 2 * G::NXO) {
    const int j = i - 3 * G::NCORNER;
    c = 1 + j / G::NXO;
    const int k = j % G::NXO, side = k / (RR * TY), k2 = k % (RR * TY);
    lx = side ? TX + RR + k2 % RR : -2 * RR + k2 % RR;
    ly = k2 / RR;
  } else {
    const int j = i - 3 * G::NCORNER - 2 * G::NXO;
    c = (j / G::NYO) ? 2 : 0;
    const int k = j % G::NYO, side = k / (RR * TX), k2 = k % (RR * TX);
    lx = k2 % TX;
    ly = side ? TY + RR + k2 / TX : -2 * RR + k2 / TX;
  }
}

// E = (curl_D B / rho) x B at LDS node (bx, by) of the B plane, with the undivided z-derivatives of
// B_x and B_y already formed (SBP_operator.cl:126-135, kernel_time_integrator.cl:147).  Each
// undivided derivative along d is scaled by a.jsc[d] = 1/(dx_d rho).
template <typename R, int RR, int BY, int BX>
__device__ __forceinline__ void emf_at(const FArgs<R>& a, R (*sB)[BY][BX], int bx, int by, R dzBx, R dzBy, R Bx, R By,
                                       R Bz, R& Ex, R& Ey, R& Ez) {
  constexpr int Q = 2 * RR + 1;
  R v[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = sB[1][by][bx - RR + k];
  const R dxBy = deriv<R, RR>(a.c, v);
#pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = sB[2][by][bx - RR + k];
  const R dxBz = deriv<R, RR>(a.c, v);
#pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = sB[0][by - RR + k][bx];
  const R dyBx = deriv<R, RR>(a.c, v);
#pragma unroll
  for (int k = 0; k < Q; ++k) v[k] = sB[2][by - RR + k][bx];
  const R dyBz = deriv<R, RR>(a.c, v);
  const R Jx = a.jsc[1] * dyBz - a.jsc[2] * dzBy, Jy = a.jsc[2] * dzBx - a.jsc[0] * dxBz,
          Jz = a.jsc[0] * dxBy - a.jsc[1] * dyBx;
  Ex = Jy * Bz - Jz * By;
  Ey = Jz * Bx - Jx * Bz;
  Ez = Jx * By - Jy * Bx;
}

// Stage kernel body.  MODE as in stage_tile: 0 out = cP*P + cQ*Q + cD*RHS;  1 (2N) b2 = cP*b2 +
// cD*RHS, out = Y + cQ*b2;  2 as 0 with Q either Y or absent (only P loaded);  3 as 0 with P
// either Y or absent and no Q.
// SK = 0: iteration o stages B plane pe = o + RR, forms E(pe), then (second barrier) the x/y
// derivatives of E on pe.  SK = 1: the LDS planes are double-buffered and iteration o stages
// pe = o + RR + 1 and forms E(pe) while it differentiates E(pe - 1) from the other buffer, so a
// z-step needs one barrier; the windows are one plane longer.
template <typename R, int RR, int TX, int TY, bool HOD, bool CONV, int MODE, int PD, int SK>
__device__ __forceinline__ void hall_stage_tile(const FArgs<R>& a, int x0, int y0, int z0, HallLds<R, RR, TX, TY, 1 + SK>& L) {
  using G = HallGeom<RR, TX, TY>;
  constexpr int Q = 2 * RR + 1;
  constexpr int QB = Q + SK;              // B window: planes o .. q = o + 2RR + SK
  constexpr int NRX = RR + 1 + SK;        // pending RHS: planes o .. pe
  constexpr int NPW = MODE == 0 ? 6 : MODE == 3 ? 0 : 3;  // pointwise loads: P (b2), [Q]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  const bool active = x0 + tx < a.nx && y0 + ty < a.ny;
  const long long S = a.S;
  const unsigned RB = sizeof(R);
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;  // plane stride in bytes
  const unsigned colb = (unsigned)(wrap(x0 + tx, a.nx) + a.nx * wrap(y0 + ty, a.ny)) * RB;
  const int zend = min(z0 + a.zc, a.nz);
  const Comps<R, 3> B = comps<R, 3>(a.Y, S);
  const R* Pp = MODE == 1 ? a.b2 : (a.P ? a.P : a.Y);
  const R* Qp = a.Q ? a.Q : a.Y;

  // E-halo column of this thread (threads past the last one repeat it and do not store)
  const bool has_h = tid < G::NH;
  int hlx, hly;
  halo_col<RR, TX, TY>(has_h ? tid : G::NH - 1, hlx, hly);
  const unsigned hcolb = (unsigned)(wrap(x0 + hlx, a.nx) + a.nx * wrap(y0 + hly, a.ny)) * RB;
  // staging items of this thread: LDS index into a B plane (flattened) and byte offset (component included)
  int sidx[G::K];
  unsigned soff[G::K];
  bool sown[G::K];
#pragma unroll
  for (int k = 0; k < G::K; ++k) {
    const int i = tid + k * G::NT;
    sown[k] = i < G::NITEMS;
    int lx, ly, c;
    stage_item<RR, TX, TY>(sown[k] ? i : G::NITEMS - 1, lx, ly, c);
    sidx[k] = (c * G::BY + ly + 2 * RR) * G::BX + lx + 2 * RR;
    soff[k] = (unsigned)(wrap(x0 + lx, a.nx) + a.nx * wrap(y0 + ly, a.ny)) * RB + (unsigned)(c * S) * RB;
  }

  R qb[3][QB];         // B of the own column: planes o .. q
  R qh[2][Q];          // B_x, B_y of the E-halo column: planes pe-RR .. pe+RR
  R qe[2][QB];         // E_x, E_y of the own column: planes o-RR .. pe
  R rx[3][NRX];        // RHS without the z-derivatives of E, planes o .. pe
  R cs[PD][5];         // column loads in flight (B of the own column, B_x/B_y of the halo column)
  R pw[PD][NPW > 0 ? NPW : 1];  // pointwise loads in flight (output plane)
  R st[G::K], shz;     // staging loads in flight (B halo of the next plane, B_z of the halo column)

  const int o0 = z0 - 2 * RR - SK;
  const int zlast = zend - 1 + 2 * RR;  // last B plane the chunk needs
  auto colld = [&](R* d, int pl) {
    const unsigned pz = (unsigned)wrap(min(pl, zlast), a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ldo(B.c[c], colb + pz);
    d[3] = ldo(B.c[0], hcolb + pz);
    d[4] = ldo(B.c[1], hcolb + pz);
  };
  auto ptld = [&](R* d, int pl) {
    if constexpr (NPW == 0) return;
    const unsigned o = colb + (unsigned)wrap(max(min(pl, zend - 1), z0), a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ldo(Pp + c * S, o);
    if constexpr (MODE == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[3 + c] = ldo(Qp + c * S, o);
    }
  };
  auto stld = [&](int pl) {
    const unsigned pz = (unsigned)wrap(min(pl, zlast - RR), a.nz) * plb;
#pragma unroll
    for (int k = 0; k < G::K; ++k) st[k] = ldo(B.c[0], soff[k] + pz);
    shz = ldo(B.c[2], hcolb + pz);
  };
  // prologue: the windows' planes before the first iteration's heads (plane o0 + 2RR + SK)
#pragma unroll
  for (int j = 0; j < QB - 1; ++j) {
    const unsigned pz = (unsigned)wrap(o0 + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][j] = ldo(B.c[c], colb + pz);
  }
#pragma unroll
  for (int j = 0; j < Q - 1; ++j) {
    const unsigned pz = (unsigned)wrap(o0 + SK + j, a.nz) * plb;
    qh[0][j] = ldo(B.c[0], hcolb + pz);
    qh[1][j] = ldo(B.c[1], hcolb + pz);
  }
  __builtin_amdgcn_sched_barrier(0);
  colld(cs[0], o0 + 2 * RR + SK);  // as issued by iteration o0 - PD
  ptld(pw[0], o0);
  __builtin_amdgcn_sched_barrier(0);
  stld(o0 + RR + SK);  // as issued by iteration o0 - 1
  if constexpr (PD == 2) {
    colld(cs[1], o0 + 2 * RR + SK + 1);
    ptld(pw[1], o0 + 1);
  }
  __builtin_amdgcn_sched_barrier(0);

  auto body = [&](auto slot, int o) {
    constexpr int UL = decltype(slot)::value;  // position in the two-iteration loop body
    constexpr int SL = UL % PD;                 // register slot of the loads in flight
    const int pe = o + RR + SK;
    const int buf = SK ? UL : 0;
    auto& sB = L.sB[buf];
    auto& sE = L.sE[buf];
    // window heads: plane q = pe + RR
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][QB - 1] = cs[SL][c];
    qh[0][Q - 1] = cs[SL][3];
    qh[1][Q - 1] = cs[SL][4];
    // B plane pe into LDS: own column, E-halo column, staged halo items
    const int cx = tx + 2 * RR, cy = ty + 2 * RR;
#pragma unroll
    for (int c = 0; c < 3; ++c) sB[c][cy][cx] = qb[c][RR + SK];
    if (has_h) {
      sB[0][hly + 2 * RR][hlx + 2 * RR] = qh[0][RR];
      sB[1][hly + 2 * RR][hlx + 2 * RR] = qh[1][RR];
      sB[2][hly + 2 * RR][hlx + 2 * RR] = shz;
    }
    R* sBf = &sB[0][0][0];
#pragma unroll
    for (int k = 0; k < G::K; ++k)
      if (sown[k]) sBf[sidx[k]] = st[k];
    lds_barrier();
    // E on plane pe at the own column, and the dissipation and convection terms of plane pe.  Each
    // x/y row of B is read from LDS once and feeds J (the derivatives curl B needs), the dissipation
    // sum and the convection derivative; the z rows come from the register window.
    const bool rhs_pe = pe >= z0 && pe < zend;
    R r[3] = {R(0), R(0), R(0)};
    {
      R dB[3][3];   // undivided derivative sums: component c along direction d
      R hsum[3];    // dissipation sums per component: w_0 B over the three directions, then w_|k| B(n + k e_d)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (HOD) hsum[c] = (a.c.w[0] * R(3)) * qb[c][RR + SK];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const bool for_j = c != d;  // curl B uses every derivative across components
          if (!(for_j || CONV || HOD)) continue;
          R v[Q];
#pragma unroll
          for (int k = 0; k < Q; ++k)
            v[k] = k == RR ? qb[c][RR + SK]
                           : d == 0 ? sB[c][cy][cx - RR + k] : d == 1 ? sB[c][cy - RR + k][cx] : qb[c][SK + k];
          if (for_j || CONV) dB[c][d] = deriv<R, RR>(a.c, v);
          if constexpr (HOD) {
            // the dissipation sums of the three directions in one chain (hodsum per direction, merged)
            if (rhs_pe) {
#pragma unroll
              for (int k = 1; k <= RR; ++k) {
                hsum[c] = hsum[c] + a.c.w[k] * v[RR - k];
                hsum[c] = hsum[c] + a.c.w[k] * v[RR + k];
              }
            }
          }
        }
      }
      // J = curl_D B / rho, E = J x B  (SBP_operator.cl:126-135, kernel_time_integrator.cl:147)
      const R Jx = a.jsc[1] * dB[2][1] - a.jsc[2] * dB[1][2];
      const R Jy = a.jsc[2] * dB[0][2] - a.jsc[0] * dB[2][0];
      const R Jz = a.jsc[0] * dB[1][0] - a.jsc[1] * dB[0][1];
      const R Bx = qb[0][RR + SK], By = qb[1][RR + SK], Bz = qb[2][RR + SK];
      const R Ex = Jy * Bz - Jz * By, Ey = Jz * Bx - Jx * Bz, Ez = Jx * By - Jy * Bx;
      sE[0][ty + RR][tx + RR] = Ex;
      sE[1][ty + RR][tx + RR] = Ey;
      sE[2][ty + RR][tx + RR] = Ez;
      qe[0][QB - 1] = Ex;
      qe[1][QB - 1] = Ey;
      if (rhs_pe) {
        if constexpr (CONV) {
          //   sum_d (2/dx_d) 0.5 ([ZERO] u (D_d B)_d - u_d (D_d B))
#pragma unroll
          for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const R f = a.two_h[d] * R(0.5);
              r[c] = r[c] - f * (a.u[d] * dB[c][d]);
              if (c == d && a.src_zero) {
#pragma unroll
                for (int e = 0; e < 3; ++e) r[e] = r[e] + f * (a.u[e] * dB[c][d]);
              }
            }
          }
        }
        if constexpr (HOD) {
#pragma unroll
          for (int c = 0; c < 3; ++c) r[c] = r[c] + hsum[c];
        }
      }
    }
    // E at the E-halo column: only planes whose x/y derivatives of E the chunk takes (the z lead-in
    // and tail planes feed the own column's z-derivatives alone); a wave whose threads all lie past
    // the last halo column may skip it too
#if IEQ_HALL_ZSKIP
    if (rhs_pe)
#endif
#if IEQ_HALL_HSKIP
    if (has_h)
#endif
    {
      R hx, hy, hz;
      emf_at<R, RR, G::BY, G::BX>(a, sB, hlx + 2 * RR, hly + 2 * RR, deriv<R, RR>(a.c, qh[0]), deriv<R, RR>(a.c, qh[1]),
                                  qh[0][RR], qh[1][RR], shz, hx, hy, hz);
      if (has_h) {
        sE[0][hly + RR][hlx + RR] = hx;
        sE[1][hly + RR][hlx + RR] = hy;
        sE[2][hly + RR][hlx + RR] = hz;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) rx[c][NRX - 1] = r[c];
    // -curl_D E, x/y terms of plane pe - SK (this plane after a second barrier, the previous plane
    // from the other buffer)
    if constexpr (SK == 0) lds_barrier();
    const int pd = pe - SK;
    if (pd >= z0 && pd < zend) {
      auto& sEd = L.sE[SK ? buf ^ 1 : 0];
      R v[Q];
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sEd[1][ty + RR][tx + k];
      const R dxEy = deriv<R, RR>(a.c, v);
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sEd[2][ty + RR][tx + k];
      const R dxEz = deriv<R, RR>(a.c, v);
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sEd[0][ty + k][tx + RR];
      const R dyEx = deriv<R, RR>(a.c, v);
#pragma unroll
      for (int k = 0; k < Q; ++k) v[k] = sEd[2][ty + k][tx + RR];
      const R dyEz = deriv<R, RR>(a.c, v);
      rx[0][RR] = rx[0][RR] + a.inv_h[1] * (-dyEz);
      rx[1][RR] = rx[1][RR] + a.inv_h[0] * dxEz;
      rx[2][RR] = rx[2][RR] + (a.inv_h[0] * (-dxEy) + a.inv_h[1] * dyEx);
    }
    // output plane o: z-derivatives of E, then the RK update
    if (active && o >= z0 && o < zend) {
      R rhs[3] = {rx[0][0], rx[1][0], rx[2][0]};
      const R dzEx = deriv<R, RR>(a.c, qe[0]);
      const R dzEy = deriv<R, RR>(a.c, qe[1]);
      rhs[0] = rhs[0] + a.inv_h[2] * dzEy;
      rhs[1] = rhs[1] + a.inv_h[2] * (-dzEx);
      const unsigned oc = colb + (unsigned)o * plb;
      if constexpr (MODE == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          R ov = a.cD * rhs[c];
          if (a.Q) ov = a.cQ * pw[SL][3 + c] + ov;
          if (a.P) ov = a.cP * pw[SL][c] + ov;
          sto(a.out + c * S, oc, ov);
        }
      } else if constexpr (MODE == 2) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          R ov = a.cD * rhs[c];
          if (a.Q) ov = a.cQ * qb[c][0] + ov;
          ov = a.cP * pw[SL][c] + ov;
          sto(a.out + c * S, oc, ov);
        }
      } else if constexpr (MODE == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          R ov = a.cD * rhs[c];
          if (a.P) ov = a.cP * qb[c][0] + ov;
          sto(a.out + c * S, oc, ov);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const R b2 = a.cP * pw[SL][c] + a.cD * rhs[c];
          sto(a.b2 + c * S, oc, b2);
          sto(a.out + c * S, oc, qb[c][0] + a.cQ * b2);
        }
      }
    }
    // next loads: staging of plane pe + 1, then the columns of plane q + PD and the pointwise
    // operands of output plane o + PD
    stld(pe + 1);
    colld(cs[SL], pe + RR + PD);
    ptld(pw[SL], o + PD);
#pragma unroll
    for (int j = 0; j < QB - 1; ++j) {
#pragma unroll
      for (int c = 0; c < 3; ++c) qb[c][j] = qb[c][j + 1];
#pragma unroll
      for (int c = 0; c < 2; ++c) qe[c][j] = qe[c][j + 1];
    }
#pragma unroll
    for (int j = 0; j < Q - 1; ++j) {
#pragma unroll
      for (int c = 0; c < 2; ++c) qh[c][j] = qh[c][j + 1];
    }
#pragma unroll
    for (int k = 0; k < NRX - 1; ++k) {
#pragma unroll
      for (int c = 0; c < 3; ++c) rx[c][k] = rx[c][k + 1];
    }
  };
  // two iterations per trip (the loop body is instantiated once per position, so the windows
  // shift by renaming inside the trip); a trip may run one iteration past zend, which writes nothing
  for (int o = o0; o < zend; o += 2) {
    body(std::integral_constant<int, 0>{}, o);
    body(std::integral_constant<int, 1>{}, o + 1);
  }
}

#ifndef IEQ_HALL_WAVES
#define IEQ_HALL_WAVES 2
#endif
// The batched FP32 form holds the tile loop's state besides the single-grid kernel's: at 4
// waves per SIMD (128 VGPRs) its 6th-order instantiations spill 14-20 VGPRs, so it is compiled
// for 3 (64-member ensemble: 0.227 s against 0.234 s on one box, where the tuner then prefers
// the two-kernel form).
#ifndef IEQ_HALL_WAVES32B
#define IEQ_HALL_WAVES32B 3
#endif
#ifndef IEQ_HALL_WAVES32
#define IEQ_HALL_WAVES32 4
#endif
// Planes of column loads in flight: one (measured faster than two once the loop body holds two
// iterations in both precisions).
#ifndef IEQ_HALL_PD64
#define IEQ_HALL_PD64 1
#endif
template <typename R> constexpr int hall_pd() { return sizeof(R) == 8 ? IEQ_HALL_PD64 : 1; }
// One barrier per z-step with double-buffered LDS planes (1) or two barriers and single planes (0).
#ifndef IEQ_HALL_SKEW
#define IEQ_HALL_SKEW 0
#endif

template <typename R, int RR, int TX, int TY, bool HOD, bool CONV, int MODE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? IEQ_HALL_WAVES32 : IEQ_HALL_WAVES, 8))) void k_hall_stage(
    FArgs<R> a) {
  __shared__ HallLds<R, RR, TX, TY, 1 + IEQ_HALL_SKEW> L;
  int x0, y0, z0;
  tile_origin<TY, TX>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  hall_stage_tile<R, RR, TX, TY, HOD, CONV, MODE, hall_pd<R>(), IEQ_HALL_SKEW>(a, x0, y0, z0, L);
}

// Persistent batched form over a tile list covering every ensemble member.
template <typename R, int RR, int TX, int TY, bool HOD, bool CONV, int MODE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? IEQ_HALL_WAVES32B : IEQ_HALL_WAVES, 8))) void k_hall_stage_batched(
    const FArgs<R>* __restrict__ args, const int4* __restrict__ tiles, int ntiles) {
  __shared__ HallLds<R, RR, TX, TY, 1 + IEQ_HALL_SKEW> L;
  const int chunk = (ntiles + kXcds - 1) / kXcds, slots = gridDim.x / kXcds;  // gridDim.x is a multiple of 8
  const int first = (blockIdx.x % kXcds) * chunk, last = min(first + chunk, ntiles);
  for (int t = first + (int)blockIdx.x / kXcds; t < last; t += slots) {
    const int4 d = tiles[t];
    hall_stage_tile<R, RR, TX, TY, HOD, CONV, MODE, hall_pd<R>(), IEQ_HALL_SKEW>(args[d.w], d.x, d.y, d.z, L);
    lds_barrier();
  }
}

}  // namespace
}  // namespace ieq
