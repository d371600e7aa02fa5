// Stencil machinery shared by the fused stage kernels, ensembles and the adjoint:
// central-row coefficients, per-launch stage arguments, tile mapping, z-marching
// helpers, and the fused forward launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace ieq {

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
  R inv_h[3];  // (R)(1/dx_d): Hall_volume and curl scaling
  R two_h[3];  // (R)(2/dx_d): convec_volume scaling
  R rho;       // uniform density (J = curl B / rho)
  R jsc[3];    // (R)(1/(dx_d rho)): scaling of the undivided derivatives that form J
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
  // adjoint stage only (stage kernels instantiated with ADJ = true):
  const R* G;   // pointwise term subtracted from the RHS (null: none)
  R cY;         // out += cY * Y at the node
  int src_adj;  // source term in transposed form: rhs_d += (1/dx_d) sum_c u_c D_d Y_c
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8) and
// each XCD has its own L2.  Giving XCD k the contiguous tile range
// [start_k, start_k + count_k) keeps the halo rows a tile shares with its y- and z-neighbours
// inside one L2 while they are being read.  Bijective for any nb.
constexpr int kXcds = 8;
__device__ inline int xcd_tile(int b, int nb) {
  const int q = nb / kXcds, r = nb % kXcds, k = b % kXcds;
  return k * q + min(k, r) + b / kXcds;
}

// Linear tile index (x fastest, then y, then z-chunk) -> origin of a TX x TY tile.
template <int TY, int TX = 64> __device__ inline void tile_origin(int t, int nx, int ny, int zc, int& x0, int& y0, int& z0) {
  const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
  x0 = (t % ntx) * TX;
  t /= ntx;
  y0 = (t % nty) * TY;
  z0 = (t / nty) * zc;
}

// Periodic index into [0, n).  The z-marching kernels call it with indices at most a few
// stencil widths outside the grid, so one correction usually suffices; the loops keep it
// exact for any v without an integer division (which would cost tens of scalar
// instructions per call).
__device__ inline int wrap(int v, int n) {
  while (v < 0) v += n;
  while (v >= n) v -= n;
  return v;
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

// Barrier over LDS traffic only: the LDS stores of this plane are visible after it,
// while global prefetch loads of the next plane stay in flight across it.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Global accesses of the fused kernels use a uniform base pointer per field component
// (held in SGPRs) and a 32-bit per-lane byte offset, which maps onto the SGPR-base form of
// the global load/store instructions (no 64-bit address arithmetic per access).  The host
// only takes these paths when one component plane of a field is below 2^31 bytes.
template <typename R> __device__ __forceinline__ R ldo(const R* __restrict__ base, unsigned off) {
  return *(const R*)((const char*)base + off);
}
template <typename R> __device__ __forceinline__ void sto(R* __restrict__ base, unsigned off, R v) {
  *(R*)((char*)base + off) = v;
}
// Host check: one component plane of an nx*ny*nz field of `rb`-byte REALs is addressable with
// 32-bit byte offsets.
inline bool offsets_fit(long long nx, long long ny, long long nz, long long rb) {
  return nx * ny * nz * rb < (1LL << 31);
}

template <typename R, int NC> struct Comps {  // component base pointers of a field
  const R* c[NC];
};
template <typename R, int NC> __device__ __forceinline__ Comps<R, NC> comps(const R* f, long long S) {
  Comps<R, NC> r;
#pragma unroll
  for (int k = 0; k < NC; ++k) r.c[k] = f + k * S;
  return r;
}

// Per-thread staging of one xy-plane of an NC-component field into an LDS tile (TX x TY
// cells) with a halo of RR cells.  The thread owns its column (x0+tx, y0+ty) and, at the tile edges,
// up to one x-halo cell and up to two y-halo cells (all coordinates wrap periodically).
// XM / YM select the components whose x / y halo is staged (bit c = component c): a
// component that is only differentiated along y needs no x halo, and vice versa.
template <typename R, int RR, int TY, int NC, int XM = (1 << NC) - 1, int YM = (1 << NC) - 1, int TX = 64>
struct PlaneStage {
  static constexpr int LX = TX + 2 * RR, LY = TY + 2 * RR;
  // With TY >= 2*RR a thread stages at most one y-halo cell (top rows or bottom rows).
  static constexpr bool ONE_Y = TY >= 2 * RR;
  bool hx, ht, hb;
  int lcx, lry;                // LDS column of the x-halo cell; LDS row of the y-halo cell (ONE_Y)
  unsigned gx, gt, gb;         // in-plane byte offsets of the halo cells
  R vx[NC], vt[NC], vb[ONE_Y ? 1 : NC];  // prefetched halo values of the next plane to stage

  __device__ void init(int tx, int ty, int x0, int y0, int nx, int ny) {
    const int xw = wrap(x0 + tx, nx), yw = wrap(y0 + ty, ny);
    hx = tx < RR || tx >= TX - RR;
    ht = ty < RR;
    hb = ty >= TY - RR;
    lcx = tx < RR ? tx : tx + 2 * RR;
    gx = (unsigned)((tx < RR ? wrap(x0 - RR + tx, nx) : wrap(x0 + tx + RR, nx)) + nx * yw) * (unsigned)sizeof(R);
    gt = (unsigned)(xw + nx * wrap(y0 - RR + ty, ny)) * (unsigned)sizeof(R);
    gb = (unsigned)(xw + nx * wrap(y0 + ty + RR, ny)) * (unsigned)sizeof(R);
    if (ONE_Y) {
      lry = ht ? ty : ty + 2 * RR;
      if (hb) gt = gb;
      ht = ht || hb;
    }
  }
  // halo of the plane starting at byte offset pz
  __device__ void load(const Comps<R, NC>& f, unsigned pz) {
    if (hx)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if ((XM >> c) & 1) vx[c] = ldo(f.c[c], gx + pz);
    if (ht)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if ((YM >> c) & 1) vt[c] = ldo(f.c[c], gt + pz);
    if constexpr (!ONE_Y) {
      if (hb)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if ((YM >> c) & 1) vb[c] = ldo(f.c[c], gb + pz);
    }
  }
  __device__ void store(R (*sm)[LY][LX], int tx, int ty) {
    if (hx)
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if ((XM >> c) & 1) sm[c][ty + RR][lcx] = vx[c];
    if constexpr (ONE_Y) {
      if (ht)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if ((YM >> c) & 1) sm[c][lry][tx + RR] = vt[c];
    } else {
      if (ht)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if ((YM >> c) & 1) sm[c][ty][tx + RR] = vt[c];
      if (hb)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if ((YM >> c) & 1) sm[c][ty + 2 * RR][tx + RR] = vb[c];
    }
  }
};

// Stencil half-width RR of the central rows and their coefficients (derivative and
// HO dissipation, HO_DISSIPATION_FACTOR folded in); -1 when the central row is wider.
template <typename R> int central_coefs(int ops, double ho_factor, StencilCoef<R>& c);

struct StageSpec {
  int kind;  // 0: K_STAGE, 1: 2R update, 2: SSPRK104_06, 3: EMF, 4: 2N fused (a+b)
  int out, P, Q, Y;
  double cP, cQ, cD;  // K_STAGE: cD multiplies DT; 2R: (a, b); 2N: A (cP), B (cQ)
};

// Stage-kernel variant of a stage with operands P and Q (absent: null / -1) and differentiated
// field Y: 1 for the 2N update; 2 when only P must be loaded (Q is Y or absent); 3 when no
// operand must be loaded (P is Y or absent, no Q); 0 otherwise.
inline int stage_mode(bool two_n, const void* P, const void* Q, const void* Y) {
  if (two_n) return 1;
  if (P && P != Y && (!Q || Q == Y)) return 2;
  if ((!P || P == Y) && !Q) return 3;
  return 0;
}
inline int stage_mode(bool two_n, int P, int Q, int Y) {
  if (two_n) return 1;
  if (P >= 0 && P != Y && (Q < 0 || Q == Y)) return 2;
  if ((P < 0 || P == Y) && Q < 0) return 3;
  return 0;
}

// Fused plan of one step (same arithmetic as the reference schedule, time kernels dropped).
bool fused_plan(const std::string& I, int hall, std::vector<StageSpec>& v, int& nf);

// Fused forward launches (RR in {1, 2, 3}; tile shape code 4: 64 x 4, 8: 64 x 8, 32: 32 x 8).
template <typename R> void launch_emf(int rr, int ty, const FArgs<R>& a, hipStream_t s);
template <typename R> void launch_stage(int rr, int ty, const FArgs<R>& a, bool hall, bool hod, hipStream_t s);
// z-chunk that makes the stage kernel's tile grid one round of resident workgroups.
template <typename R> int pick_zc(int rr, int ty, const FArgs<R>& a, bool hall, bool hod);
template <typename R> int pick_zc_emf(int rr, int ty, const FArgs<R>& a);

}  // namespace ieq
