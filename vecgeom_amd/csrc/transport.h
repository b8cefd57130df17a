// Batched transport: records shared by the host-side work-queue builder
// (batch.cpp) and the persistent kernels (transport.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mlp.h"

namespace vgm {

enum TransportMode : int32_t { kTransportFP64 = 0, kTransportMixed = 1, kTransportClassify = 2 };

// One solid of the batch: its shape family and parameter block, the track
// range it owns and the mean physics step.
struct SolidRec {
  int32_t type;
  int32_t off;  // parameter block offset in the pool (doubles)
  int32_t nd;   // parameter block size (doubles)
  int32_t mlp;  // learned-safety network (index into TransportArgs::mlps), -1 if none
  int64_t begin, end;
  double lambda;  // mean physics step (mm)
  double scale;   // bounding-box diagonal (mm), sets the float guard band
  double wc[3];   // periodic world box: centre and half-widths (1.3 x bounding box)
  double wh[3];
  double bc[3];   // bounding box of the solid (Extent): centre and half-widths
  double bh[3];
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
  // when set, the index of the first step is read here instead (graph
  // replays advance it without re-capturing)
  const int32_t *step0_dev;
  uint64_t seed;
  double *x, *y, *z;
  double *dx, *dy, *dz;
  int8_t *side;
  int32_t *crossings;
  int *counter;
  unsigned long long *fp64_steps;
  // learned safety (VGM_LEARNED): networks, and counters of bound
  // evaluations, steps moved on a bound, exact fallbacks and (audit runs)
  // bounds above the exact safety
  const MlpDesc *mlps;
  unsigned long long *learned;  // [4]
  int32_t audit;
  int32_t fallback_pct;  // a wave runs its fallbacks once this % of its live tracks wait
  int32_t pad3;
  int32_t fallback_grid;   // workgroups of the learned fallback kernel
  int32_t cull;            // learned launches: outside moves clear of the bounding box need no bound
  int32_t pad2;
  // learned families, between their per-step kernels: carried bound and
  // flags per track (indexed like x), and the queue of exact fallbacks
  // (entries solid << 40 | track) with its two counters (by step parity)
  double *tcarry;
  uint8_t *tflags;
  uint64_t *queue;
  int *qcount;
};

enum LearnedCounter : int { kLearnedQueries = 0, kLearnedServed = 1, kLearnedFallbacks = 2, kLearnedViolations = 3 };

// Upper bound of the workgroup size of the transport kernels.  It is also
// their launch bound: 256 threads let the compiler give a wave up to 256
// VGPRs, which the FP64 shape paths need to stay out of scratch.
constexpr int kTransportMaxBlock = 256;

struct LaunchShape {
  int block;        // threads per workgroup
  int grid;         // persistent workgroups
  int tile_tracks;  // tracks per tile
};

// One family's share of a transport launch: tiles [t0, t1) of the sorted
// queue.  counter != nullptr selects the persistent form with `grid`
// workgroups; otherwise one workgroup runs one tile.
struct FamilyLaunch {
  int type;
  int t0, t1;
  int grid;
  int *counter;
  int stage_nd;  // > 0: stage parameter blocks of up to this many doubles in LDS
  int mode;      // FamilyMode of this launch
  bool has_mlp;  // some solid of the family has a learned-safety network
  int64_t queue_off;  // this family's region of the fallback queue (its track count long)
};

// How a family's track-steps are computed.  All but kFamMixed take the FP64
// decisions of step_fp64 (bit-identical trajectories):
//   kFamFP64       safety, then the boundary distance when the step exceeds it
//   kFamMixed      float filter, FP64 redo of steps near a boundary
//   kFamDistFirst  for families whose safety never exceeds the boundary
//                  distance: the boundary distance first and the safety only
//                  where the two orders could differ (within
//                  kDistanceFirstWindow of it); other families take the
//                  kFamFP64 step
//   kFamLearned    learned safety bound, then kFamCulled as the fallback
//   kFamCulled     kFamDistFirst, and outside moves clear of the solid's
//                  bounding box skip the queries
enum FamilyMode : int { kFamFP64 = 0, kFamMixed = 1, kFamDistFirst = 2, kFamLearned = 3, kFamCulled = 4 };

// Writes the first step index of a graph replay (TransportArgs::step0_dev).
hipError_t launch_set_step0(hipStream_t stream, int32_t *dst, int32_t v);
hipError_t launch_reset(hipStream_t stream, int *counters, int n, unsigned long long *fp64_steps,
                        unsigned long long *learned);
// True when the family can serve learned bounds (its safety is a distance).
bool family_learnable(int type);
hipError_t launch_family(hipStream_t stream, const TransportArgs &a, const FamilyLaunch &f, int block);
// The launcher of one shape family S, instantiated in that family's own
// translation unit (transport_family.hip).
template <class S>
hipError_t launch_family_of(hipStream_t stream, const TransportArgs &a, const FamilyLaunch &f, int block);
// Initial inside flags (persistent over all tiles, uses a.counter).
hipError_t launch_classify(hipStream_t stream, const TransportArgs &a, int grid, int block);

}  // namespace vgm
