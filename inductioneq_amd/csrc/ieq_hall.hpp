// Single-kernel RK stage of the periodic Hall problem (EMF formed in the stage kernel).
#pragma once

#include <hip/hip_runtime.h>

#include "ieq_stencil.hpp"

namespace ieq {

// Stage kernel for an update form (stage_mode), dissipation and convection, single-grid or
// batched over a tile list (int4 {x0, y0, z0, member}, FArgs per member).  `shape` 0 is the
// single-grid tile; shape 1 (batched launches only) is a second tile for ensembles of small
// or odd-sized members: FP32 32 x 8, FP64 16 x 16.
template <typename R> const void* hall_stage_kernel(int rr, int mode, bool hod, bool conv, bool batched, int shape = 0);
// Tile shape (threads per workgroup = tx * ty) of the kernels above.
template <typename R> void hall_tile_shape(int* tx, int* ty, int shape = 0);
template <typename R> void launch_hall_stage(int rr, const FArgs<R>& a, bool hod, hipStream_t s);
// z-chunk that makes the tile grid one round of resident workgroups.
template <typename R> int pick_zc_hall(int rr, const FArgs<R>& a, bool hod);

}  // namespace ieq
