// Batched transport: records shared by the host-side work-queue builder
// (batch.cpp) and the persistent kernels (transport.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgm {

enum TransportMode : int32_t { kTransportFP64 = 0, kTransportMixed = 1, kTransportClassify = 2 };

// One solid of the batch: its shape family and parameter block, the track
// range it owns and the mean physics step.
struct SolidRec {
  int32_t type;
  int32_t off;  // parameter block offset in the pool (doubles)
  int64_t begin, end;
  double lambda;  // mean physics step (mm)
  double scale;   // bounding-box diagonal (mm), sets the float guard band
  double wc[3];   // periodic world box: centre and half-widths (1.3 x bounding box)
  double wh[3];
};

// A tile of the work queue: consecutive tracks of one solid.
struct Tile {
  int32_t solid;
  int32_t pad;
  int64_t begin, end;
};

struct TransportArgs {
  const double *pool;
  const SolidRec *solids;
  const Tile *tiles;
  int32_t ntiles;
  int32_t nsteps;
  int32_t step0;  // index of the first step (RNG counter)
  int32_t pad;
  uint64_t seed;
  double *x, *y, *z;
  double *dx, *dy, *dz;
  int8_t *side;
  int32_t *crossings;
  int *counter;
  unsigned long long *fp64_steps;
};

struct LaunchShape {
  int block;        // threads per workgroup
  int grid;         // persistent workgroups
  int tile_tracks;  // tracks per tile
};

hipError_t launch_transport(hipStream_t stream, int mode, const TransportArgs &a, const LaunchShape &ls);

}  // namespace vgm
