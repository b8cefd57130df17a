// LDS staging of a shape's parameter block.  Kernels of families whose lanes
// index their tables with per-lane indices (polycone sections, polyhedron
// quadrilaterals) copy the block into LDS once per workgroup (or per tile),
// turning per-lane global gathers into LDS reads.
#pragma once
#include <hip/hip_runtime.h>

namespace vgm {

// The whole workgroup copies `nd` doubles; the caller places the barrier.
__device__ __forceinline__ void stage_block(double *dst, const double *__restrict__ src, int nd)
{
  for (int k = int(threadIdx.x); k < nd; k += int(blockDim.x)) dst[k] = src[k];
}

}  // namespace vgm
