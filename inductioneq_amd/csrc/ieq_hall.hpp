// Single-kernel RK stage of the periodic Hall problem (EMF formed in the stage kernel).
#pragma once

#include <hip/hip_runtime.h>

#include "ieq_stencil.hpp"

namespace ieq {

// Stage kernel for an update form (stage_mode), dissipation and convection, single-grid or
// batched over a tile list (int4 {x0, y0, z0, member}, FArgs per member).  Two tiles: shape 0
// (FP64 32 x 8, FP32 64 x 8) and shape 1 (FP64 16 x 16, FP32 32 x 8).  FP64 single grids up to
// about 240^3 take shape 1 (hall_shape_for), ensembles tune it.
template <typename R> const void* hall_stage_kernel(int rr, int mode, bool hod, bool conv, bool batched, int shape = 0);
// Tile shape (threads per workgroup = tx * ty) of the kernels above.
template <typename R> void hall_tile_shape(int* tx, int* ty, int shape = 0);
template <typename R> void launch_hall_stage(int rr, const FArgs<R>& a, bool hod, hipStream_t s, int shape = 0);
// z-chunk that makes the tile grid one round of resident workgroups.
template <typename R> int pick_zc_hall(int rr, const FArgs<R>& a, bool hod, int shape = 0);
// Tile of single-grid launches for an nx x ny x nz grid of `real_bytes` REALs (IEQ_HALL_TILE=0/1
// overrides): in FP64 16 x 16 up to 1 GB for the three vector fields a 2N stage touches (128^3:
// 66.3 against 67.9 us per stage, 160^3-224^3: 1-2 % faster), 32 x 8 beyond (256^3: 436.7
// against 449.3 us); in FP32 64 x 8 where 64 divides nx, 32 x 8 otherwise (96^3: 35.2 against
// 39.0 us).
int hall_shape_for(long long nx, long long ny, long long nz, int real_bytes);

}  // namespace ieq
