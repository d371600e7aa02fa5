// Device side of the fused periodic path, shared by the single-grid stage launches
// (ieq_fused.hip) and the batched ensemble launches (ieq_ensemble.hip): the z-marching
// EMF and stage tile bodies and their kernels.  Templates with internal linkage; each
// translation unit instantiates the kernels it launches.
#pragma once

#include <type_traits>

#include "ieq_stencil.hpp"

namespace ieq {
namespace {

// Planes of column loads in flight per thread in the z-marching kernels: two for FP64 when the
// stencil half-width allows (the loads for plane p + 2 are issued while plane p is computed),
// else one.  Measured at 128^3: FP64 (2 waves per SIMD) gains 9 % from the second plane; FP32
// (4 waves per SIMD, batched or not) is faster with one.
template <typename R, int RR> constexpr int march_pd() { return (sizeof(R) == 8 && RR >= 2) ? 2 : 1; }

// Memory-counter discipline of the z-marching loops.  On gfx9-family GPUs one counter tracks
// loads and stores and retires in issue order, and the compiler waits for a register's load
// with a count of the accesses issued after it on every path.  The loops therefore
//   * issue per iteration: stores of the finished plane, then the halo of the next plane
//     (only on tile-edge threads, waited for one iteration later), then the column and
//     pointwise loads for PD planes ahead, which every thread issues on every path;
//   * keep loads in flight in register slots indexed by compile-time constants (the loop
//     body is instantiated once per slot), so no register with a pending load is copied;
//   * issue the prologue's loads in the same order as the loop (pinned by scheduling
//     barriers), so the wait at the loop head is the same on entry and on the back edge.
// The only waits in the loop are then for the loads issued PD iterations earlier.

// EMF kernel: E = (curl_D B / rho) x B.
//
// Lead-aligned z-march: iteration `it` stages plane p = it + RR (the newest plane of the
// z-queue) into LDS together with its x/y halo and evaluates the x/y-derivative terms of
// J for that plane; the z-derivative terms of plane `it` come from the register queue and
// complete J(it).  The column value and the halo of a plane are therefore read in the same
// iteration, which is also when the neighbouring tiles (dispatched together to the same
// XCD) read those rows as their own column values, so the halo rows are L2 hits.
template <typename R, int RR, int TY, int NBUF = 2, int TX = 64>
__device__ __forceinline__ void emf_tile(const FArgs<R>& a, int x0, int y0, int z0, R (*sB)[3][TY + 2 * RR][TX + 2 * RR]) {
  constexpr int Q = 2 * RR + 1, PD = march_pd<R, RR>();
  const int tx = threadIdx.x, ty = threadIdx.y;
  const bool active = x0 + tx < a.nx && y0 + ty < a.ny;
  const unsigned RB = sizeof(R);
  const unsigned colb = (unsigned)(wrap(x0 + tx, a.nx) + a.nx * wrap(y0 + ty, a.ny)) * RB;  // own column
  const unsigned plb = (unsigned)(a.nx * a.ny) * RB;                                       // plane stride
  const int zend = min(z0 + a.zc, a.nz);
  const Comps<R, 3> B = comps<R, 3>(a.Y, a.S);
  const R irho = R(1) / a.rho;
  PlaneStage<R, RR, TY, 3, 0b110, 0b101, TX> hs;  // x halo: B_y, B_z;  y halo: B_x, B_z
  hs.init(tx, ty, x0, y0, a.nx, a.ny);

  R q[3][Q];        // own column: planes it-RR .. it+RR
  R cs[PD][3];      // column loads in flight: plane p in slot (p - z0) % PD
  R jr[3][RR + 1];  // x/y parts of curl_D B for planes it .. it+RR: (d_y B_z, -d_x B_z, d_x B_y - d_y B_x)
  const int it0 = z0 - RR;
  // Loads issued for planes past the last one the chunk needs (the prefetch overshoot of the
  // final iterations) re-read that last plane, a cache hit, instead of fetching unused planes.
  const int zlast = zend - 1 + RR;
  auto col = [&](R* dst, int pl) {
    const unsigned o = colb + (unsigned)wrap(min(pl, zlast), a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[c] = ldo(B.c[c], o);
  };
#pragma unroll
  for (int j = RR; j < Q - 1; ++j) {  // planes z0-RR .. z0-1 (older slots are never read)
    const unsigned o = colb + (unsigned)wrap(it0 - RR + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c][j] = ldo(B.c[c], o);
  }
  __builtin_amdgcn_sched_barrier(0);
  col(cs[0], z0);  // as issued by iteration it0 - PD
  __builtin_amdgcn_sched_barrier(0);
  hs.load(B, (unsigned)z0 * plb);  // as issued by iteration it0 - 1
  if constexpr (PD == 2) col(cs[1], z0 + 1);
  __builtin_amdgcn_sched_barrier(0);
  auto body = [&](auto slot, int it) {
    constexpr int S = decltype(slot)::value;
    const int p = it + RR;  // plane staged in this iteration
    const int buf = NBUF == 2 ? (PD == 2 ? S : (it - it0) & 1) : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c][Q - 1] = cs[S][c];  // queue head: plane p
    if constexpr (NBUF == 1) {
      if (it > it0) lds_barrier();
    }
    if (p < zend) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sB[buf][c][ty + RR][tx + RR] = q[c][Q - 1];
      hs.store(sB[buf], tx, ty);
    }
    lds_barrier();
    if (active) {
      if (p < zend) {
        R vx[Q], vy[Q];
#pragma unroll
        for (int k = 0; k < Q; ++k) vx[k] = sB[buf][1][ty + RR][tx + k];
        const R dxBy = deriv<R, RR>(a.c, vx) * a.inv_h[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) vx[k] = sB[buf][2][ty + RR][tx + k];
        const R dxBz = deriv<R, RR>(a.c, vx) * a.inv_h[0];
#pragma unroll
        for (int k = 0; k < Q; ++k) vy[k] = sB[buf][0][ty + k][tx + RR];
        const R dyBx = deriv<R, RR>(a.c, vy) * a.inv_h[1];
#pragma unroll
        for (int k = 0; k < Q; ++k) vy[k] = sB[buf][2][ty + k][tx + RR];
        const R dyBz = deriv<R, RR>(a.c, vy) * a.inv_h[1];
        jr[0][RR] = dyBz;
        jr[1][RR] = -dxBz;
        jr[2][RR] = dxBy - dyBx;
      }
      if (it >= z0 && it < zend) {
        const R dzBx = deriv<R, RR>(a.c, q[0]) * a.inv_h[2];
        const R dzBy = deriv<R, RR>(a.c, q[1]) * a.inv_h[2];
        // J = curl B / rho  (SBP_operator.cl:126-135, kernel_time_integrator.cl:147)
        const R Jx = (jr[0][0] - dzBy) * irho, Jy = (dzBx + jr[1][0]) * irho, Jz = jr[2][0] * irho;
        const R Bx = q[0][RR], By = q[1][RR], Bz = q[2][RR];
        const unsigned o = colb + (unsigned)it * plb;
        sto(a.Eout, o, Jy * Bz - Jz * By);
        sto(a.Eout + a.S, o, Jz * Bx - Jx * Bz);
        sto(a.Eout + 2 * a.S, o, Jx * By - Jy * Bx);
      }
    }
    // next loads: halo of plane p+1 (edge threads), then the column of plane p+PD (all threads)
    if (p + 1 < zend) hs.load(B, (unsigned)(p + 1) * plb);
    col(cs[S], p + PD);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) q[c][j] = q[c][j + 1];
#pragma unroll
      for (int k = 0; k < RR; ++k) jr[c][k] = jr[c][k + 1];
    }
  };
  for (int it = it0; it < zend; it += PD) {
    body(std::integral_constant<int, 0>{}, it);
    if constexpr (PD == 2) body(std::integral_constant<int, 1>{}, it + 1);
  }
}

template <typename R, int RR, int TY, int NBUF, int TX = 64>
__global__ __launch_bounds__(TX * TY) void k_emf(FArgs<R> a) {
  __shared__ R sB[NBUF][3][TY + 2 * RR][TX + 2 * RR];
  int x0, y0, z0;
  tile_origin<TY, TX>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  emf_tile<R, RR, TY, NBUF, TX>(a, x0, y0, z0, sB);
}

// Persistent batched form: blocks stride over a tile list covering every member.
template <typename R, int RR, int TY, int TX = 64>
__global__ __launch_bounds__(TX * TY) void k_emf_batched(const FArgs<R>* __restrict__ args, const int4* __restrict__ tiles,
                                                         int ntiles) {
  __shared__ R sB[2][3][TY + 2 * RR][TX + 2 * RR];
  const int chunk = (ntiles + kXcds - 1) / kXcds, slots = gridDim.x / kXcds;  // gridDim.x is a multiple of 8
  const int first = (blockIdx.x % kXcds) * chunk, last = min(first + chunk, ntiles);
  for (int t = first + (int)blockIdx.x / kXcds; t < last; t += slots) {
    const int4 d = tiles[t];
    emf_tile<R, RR, TY, 2, TX>(args[d.w], d.x, d.y, d.z, sB);
    lds_barrier();
  }
}

// Stage kernel: RHS = convection + (-curl E) + dissipation, then the RK update.
// Same pipelining as k_emf, with B (3 components) and E (3 in LDS, x/y in the queue).
#ifndef IEQ_STAGE_WAVES
#define IEQ_STAGE_WAVES 2
#endif
#ifndef IEQ_STAGE_WAVES32
#define IEQ_STAGE_WAVES32 4
#endif
template <typename R, int RR, int TY, bool HALL, int NBUF = 2, int TX = 64> struct StageLds {
  static constexpr int LX = TX + 2 * RR, LY = TY + 2 * RR;
  static constexpr int NE = HALL ? 3 : 1, LYE = HALL ? LY : 1, LXE = HALL ? LX : 1;
  R sB[NBUF][3][LY][LX];
  R sE[NBUF][NE][LYE][LXE];
};

// Stage kernel body, lead-aligned like emf_tile: iteration `it` stages plane p = it + RR of B
// (and E) into LDS and evaluates the x/y-derivative terms of the RHS of plane p; the output
// plane `it` adds the z-derivative terms from the register queues and applies the RK update.
// MODE 0: out = cP*P + cQ*Q + cD*RHS (+ cY*Y in the adjoint);  MODE 1 (2N): b2 = cP*b2 + cD*RHS,
// out = Y + cQ*b2;  MODE 2: as 0 with Q either Y or absent (Y comes from the z-queue, only P
// is loaded);  MODE 3: as 0 with P either Y or absent and no Q (no pointwise loads).  The
// SSPRK stages are modes 2 and 3.  Loads follow the memory-counter discipline described above
// emf_tile.
template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV, bool ADJ, int NBUF, int MODE, int TX = 64>
__device__ __forceinline__ void stage_tile(const FArgs<R>& a, int x0, int y0, int z0, StageLds<R, RR, TY, HALL, NBUF, TX>& L) {
  constexpr int Q = 2 * RR + 1, PD = march_pd<R, RR>();
  constexpr int NC = HALL ? 6 : 3;                            // column loads: B, then E_x, E_y, E_z
  constexpr int NPW = (MODE == 0 ? 6 : MODE == 3 ? 0 : 3) + (ADJ ? 3 : 0);  // pointwise loads: P (b2), [Q], [G]
  static_assert(!ADJ || MODE == 0, "the transposed stages use mode 0");
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
  // pointwise sources; an absent operand is read from Y (valid memory) and never used.  Only
  // the 2N form (mode 1) reads b2; the three-register forms (modes 0, 2) read P.
  const R* Pp = MODE == 1 ? a.b2 : (a.P ? a.P : a.Y);
  const R* Qp = a.Q ? a.Q : a.Y;
  const R* Gp = a.G ? a.G : a.Y;
  PlaneStage<R, RR, TY, 3, 7, 7, TX> hsb;
  PlaneStage<R, RR, TY, 3, 0b110, 0b101, TX> hse;  // x halo: E_y, E_z;  y halo: E_x, E_z
  hsb.init(tx, ty, x0, y0, a.nx, a.ny);
  if constexpr (HALL) hse.init(tx, ty, x0, y0, a.nx, a.ny);

  R qb[3][Q];                  // B of the own column: planes it-RR .. it+RR
  R qe[2][HALL ? Q : 1];       // E_x, E_y of the own column (z-derivatives of curl E)
  R cs[PD][NC];                // column loads in flight: plane p in slot (p - z0) % PD
  R pw[PD][NPW > 0 ? NPW : 1];  // pointwise loads in flight: output plane o in slot (o - it0) % PD
  R rx[3][RR + 1];             // x/y terms of the RHS for planes it .. it+RR
  const int it0 = z0 - RR;
  // Loads past the last plane the chunk needs re-read that plane (a cache hit), as in emf_tile.
  const int zlast = zend - 1 + RR;
  auto colld = [&](R* d, int pl) {
    const unsigned o = colb + (unsigned)wrap(min(pl, zlast), a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ldo(B.c[c], o);
    if constexpr (HALL) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[3 + c] = ldo(E.c[c], o);
    }
  };
  auto ptld = [&](R* d, int pl) {
    if constexpr (NPW == 0) return;
    const unsigned o = colb + (unsigned)wrap(min(pl, zend - 1), a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) d[c] = ldo(Pp + c * S, o);
    if constexpr (MODE == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[3 + c] = ldo(Qp + c * S, o);
    }
    if constexpr (ADJ) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[NPW - 3 + c] = ldo(Gp + c * S, o);
    }
  };
#pragma unroll
  for (int j = RR; j < Q - 1; ++j) {  // planes z0-RR .. z0-1 (older slots are never read)
    const unsigned o = colb + (unsigned)wrap(it0 - RR + j, a.nz) * plb;
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][j] = ldo(B.c[c], o);
    if constexpr (HALL) {
      qe[0][j] = ldo(E.c[0], o);
      qe[1][j] = ldo(E.c[1], o);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  colld(cs[0], z0);  // as issued by iteration it0 - PD
  ptld(pw[0], it0);
  __builtin_amdgcn_sched_barrier(0);
  hsb.load(B, (unsigned)z0 * plb);  // as issued by iteration it0 - 1
  if constexpr (HALL) hse.load(E, (unsigned)z0 * plb);
  if constexpr (PD == 2) {
    colld(cs[1], z0 + 1);
    ptld(pw[1], it0 + 1);
  }
  __builtin_amdgcn_sched_barrier(0);
  auto body = [&](auto slot, int it) {
    constexpr int SL = decltype(slot)::value;
    const int p = it + RR;
    const int buf = NBUF == 2 ? (PD == 2 ? SL : (it - it0) & 1) : 0;
    // queue heads: plane p
#pragma unroll
    for (int c = 0; c < 3; ++c) qb[c][Q - 1] = cs[SL][c];
    if constexpr (HALL) {
      qe[0][Q - 1] = cs[SL][3];
      qe[1][Q - 1] = cs[SL][4];
    }
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
        sE[buf][2][ty + RR][tx + RR] = cs[SL][5];
        hse.store(sE[buf], tx, ty);
      }
    }
    lds_barrier();
    if (active) {
      if (p < zend) {  // x/y terms of plane p
        R r[3] = {R(0), R(0), R(0)};
        R v[Q];
        R hod[3];
        if constexpr (CONV || HOD) {
          // One pass over the B stencils: each row of B along x or y is read from LDS once and
          // feeds both the convection derivative and the dissipation sum, so only one row is
          // live at a time.
          //   convection: sum_d (2/dx_d) 0.5 ([ZERO] u (D_d B)_d - u_d (D_d B))
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            R h = R(0);
#pragma unroll
            for (int d = 0; d < 2; ++d) {
              // (the centre value is the queue head, so only the 2*RR neighbours come from LDS)
#pragma unroll
              for (int k = 0; k < Q; ++k)
                v[k] = k == RR ? qb[c][Q - 1] : d == 0 ? sB[buf][c][ty + RR][tx + k] : sB[buf][c][ty + k][tx + RR];
              if constexpr (CONV) {
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
              if constexpr (HOD) h = d == 0 ? hodsum<R, RR>(a.c, v) : h + hodsum<R, RR>(a.c, v);
            }
            if constexpr (HOD) hod[c] = h;
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
          for (int c = 0; c < 3; ++c) r[c] = r[c] + hod[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) rx[c][RR] = r[c];
      }
      if (it >= z0 && it < zend) {  // output plane it: z terms and the RK update
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
          if (a.G) {
#pragma unroll
            for (int c = 0; c < 3; ++c) rhs[c] = rhs[c] - pw[SL][NPW - 3 + c];
          }
        }
        const unsigned oc = colb + (unsigned)it * plb;
        if constexpr (MODE == 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            R o = a.cD * rhs[c];
            if constexpr (ADJ) o = a.cY * qb[c][RR] + o;
            if (a.Q) o = a.cQ * pw[SL][3 + c] + o;
            if (a.P) o = a.cP * pw[SL][c] + o;
            sto(a.out + c * S, oc, o);
          }
        } else if constexpr (MODE == 2) {  // same operation order as mode 0 with Q = Y
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            R o = a.cD * rhs[c];
            if (a.Q) o = a.cQ * qb[c][RR] + o;
            o = a.cP * pw[SL][c] + o;
            sto(a.out + c * S, oc, o);
          }
        } else if constexpr (MODE == 3) {  // same operation order as mode 0 with P = Y, no Q
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            R o = a.cD * rhs[c];
            if (a.P) o = a.cP * qb[c][RR] + o;
            sto(a.out + c * S, oc, o);
          }
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const R b2 = a.cP * pw[SL][c] + a.cD * rhs[c];
            sto(a.b2 + c * S, oc, b2);
            sto(a.out + c * S, oc, qb[c][RR] + a.cQ * b2);
          }
        }
      }
    }
    // next loads: halos of plane p+1 (edge threads), then the column of plane p+PD and the
    // pointwise operands of output plane it+PD (all threads)
    if (p + 1 < zend) {
      hsb.load(B, (unsigned)(p + 1) * plb);
      if constexpr (HALL) hse.load(E, (unsigned)(p + 1) * plb);
    }
    colld(cs[SL], p + PD);
    ptld(pw[SL], it + PD);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int j = 0; j < Q - 1; ++j) qb[c][j] = qb[c][j + 1];
#pragma unroll
      for (int k = 0; k < RR; ++k) rx[c][k] = rx[c][k + 1];
    }
    if constexpr (HALL) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int j = 0; j < Q - 1; ++j) qe[c][j] = qe[c][j + 1];
      }
    }
  };
  for (int it = it0; it < zend; it += PD) {
    body(std::integral_constant<int, 0>{}, it);
    if constexpr (PD == 2) body(std::integral_constant<int, 1>{}, it + 1);
  }
}

// FP64: 2 waves per SIMD (the stage state needs ~200 VGPRs); FP32 fits 4 waves without spills
template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV, bool ADJ, int NBUF, int MODE, int TX = 64>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? IEQ_STAGE_WAVES32 : IEQ_STAGE_WAVES, 8))) void k_stage(
    FArgs<R> a) {
  __shared__ StageLds<R, RR, TY, HALL, NBUF, TX> L;
  int x0, y0, z0;
  tile_origin<TY, TX>(xcd_tile(blockIdx.x, gridDim.x), a.nx, a.ny, a.zc, x0, y0, z0);
  stage_tile<R, RR, TY, HALL, HOD, CONV, ADJ, NBUF, MODE, TX>(a, x0, y0, z0, L);
}

template <typename R, int RR, int TY, bool HALL, bool HOD, bool CONV, int MODE, int TX = 64>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? IEQ_STAGE_WAVES32 : IEQ_STAGE_WAVES, 8))) void k_stage_batched(
    const FArgs<R>* __restrict__ args, const int4* __restrict__ tiles, int ntiles) {
  __shared__ StageLds<R, RR, TY, HALL, 2, TX> L;
  const int chunk = (ntiles + kXcds - 1) / kXcds, slots = gridDim.x / kXcds;  // gridDim.x is a multiple of 8
  const int first = (blockIdx.x % kXcds) * chunk, last = min(first + chunk, ntiles);
  for (int t = first + (int)blockIdx.x / kXcds; t < last; t += slots) {
    const int4 d = tiles[t];
    stage_tile<R, RR, TY, HALL, HOD, CONV, false, 2, MODE, TX>(args[d.w], d.x, d.y, d.z, L);
    lds_barrier();
  }
}

}  // namespace
}  // namespace ieq
