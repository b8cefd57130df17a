// Batched heterogeneous track-step transport: many solids of different
// families, uneven track counts.
//
// Work is cut into tiles of consecutive tracks of one solid; tiles are sorted
// by shape family, and each family's range of tiles is run by a kernel
// instantiated for that family alone, so its register allocation is that of
// its own code path (a box step needs a third of the registers of a
// polyhedron step).  The family kernels run concurrently on separate streams.
// Either one workgroup runs one tile, or the launch is persistent: a grid
// sized to the family's share of the machine pulls tiles from a device
// counter until its range is drained.  Each track keeps its state in
// registers for all steps and touches HBM once on entry and once on exit.
//
// Precision modes:
//   FP64   every operation in double, bit-identical to the CPU oracle.
//   MIXED  safety and boundary distance evaluated in float as a filter; any
//          step that could reach a boundary is redone with the FP64 kernels,
//          so positions and crossings follow the FP64 trajectory except where
//          float and double disagree by more than the guard band.
#include "vgm_internal.h"
#include "transport.h"
#include "stepping.hpp"
#include "shapes/all_shapes.hpp"
#include "staging.hpp"

namespace vgm {

namespace {

struct TrackState {
  V3<double> p, d;
  bool in;
  int32_t crossings;
  double carry;  // lower bound on the safety at p (0: none), see step_fp64
};

// Slack of the carried safety bound: rounding of the moves and the
// tolerance-level slack of the facet tests.
VGM_DEV double carry_margin(const V3<double> &p)
{
  return 1e-7 + 1e-12 * (vabs(p.x) + vabs(p.y) + vabs(p.z));
}

// One track-step in FP64 (see stepping.hpp for the definition).  Returns
// true when the step ended on a boundary crossing.
//
// Families whose safety is a distance to their surface (kSafetyIsDistance:
// the tessellated solid's minimum point-to-facet distance) can skip it: a
// safety s at p bounds the safety at any p' by s - |p' - p|, so after a free
// move of l the track carries s - l (less a margin) and the next step moves
// without any query while its length stays below that bound.  The decision
// is the one the exact safety would take, so trajectories are unchanged; a
// crossing or a wrap of the periodic world clears the bound.
template <class S>
VGM_DEV bool step_fp64(const typename S::Data &s, TrackState &t, double l)
{
  if constexpr (S::kSafetyIsDistance) {
    if (l <= t.carry) {
      const double m = carry_margin(t.p);
      t.p.x = t.p.x + l * t.d.x;
      t.p.y = t.p.y + l * t.d.y;
      t.p.z = t.p.z + l * t.d.z;
      t.carry = t.carry - l - m;
      return false;
    }
  }
  double saf;
  if constexpr (S::kSafetyIsDistance) {
    // SafetyToOut/In is the surface distance when the point is on the
    // track's side and -1 otherwise; the side check (a ray walk) only matters
    // when the distance would let the step move freely.
    saf = S::template surface_distance<double>(s, t.p, t.in);
    if (l <= saf && S::template contains<double>(s, t.p) != t.in) saf = -1.0;
  } else {
    saf = t.in ? S::template safety_to_out<double>(s, t.p) : S::template safety_to_in<double>(s, t.p);
  }
  if (l <= saf) {
    if constexpr (S::kSafetyIsDistance) t.carry = saf - l - carry_margin(t.p);
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  t.carry = 0.0;
  double db = t.in ? S::template distance_to_out<double>(s, t.p, t.d, kInfLength)
                   : S::template distance_to_in<double>(s, t.p, t.d, kInfLength);
  if (db < 0.0) db = 0.0;
  if (l < db) {
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  const double m = db + kBoundaryPush;
  t.p.x = t.p.x + m * t.d.x;
  t.p.y = t.p.y + m * t.d.y;
  t.p.z = t.p.z + m * t.d.z;
  t.in = !t.in;
  t.crossings++;
  return true;
}

// Absolute guard for float results: a few float ulps of the largest
// coordinate magnitude in play, plus a relative margin on the distance.
VGM_DEV float guard_band(const V3<float> &p, float scale)
{
  return 4.0e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + scale);
}

template <class S>
VGM_DEV bool step_mixed(const typename S::Data &s, TrackState &t, double l, float scale, uint32_t &fp64_steps)
{
  if constexpr (S::kSafetyIsDistance) {
    if (l <= t.carry) return step_fp64<S>(s, t, l);  // bound carried from an FP64 safety
  }
  const V3<float> pf{float(t.p.x), float(t.p.y), float(t.p.z)};
  const V3<float> df{float(t.d.x), float(t.d.y), float(t.d.z)};
  const float lf = float(l);
  const float g = guard_band(pf, scale);
  float saf;
  if constexpr (S::kSafetyIsDistance) {
    saf = S::template surface_distance<float>(s, pf, t.in);
    if (lf + g <= saf && S::template contains<float>(s, pf) != t.in) saf = -1.0f;
  } else {
    saf = t.in ? S::template safety_to_out<float>(s, pf) : S::template safety_to_in<float>(s, pf);
  }
  bool move_free = lf + g <= saf;
  const bool by_safety = move_free;
  if (!move_free) {
    const float db = t.in ? S::template distance_to_out<float>(s, pf, df, FLT_MAX)
                          : S::template distance_to_in<float>(s, pf, df, FLT_MAX);
    move_free = (db < FLT_MAX) ? (lf + g + 1.0e-3f * db < db) : (!t.in);
  }
  if (move_free) {
    // a guarded float safety bounds the safety as the FP64 one does, so a
    // surface-distance family carries it like an FP64 safety
    t.carry = 0.0;
    if constexpr (S::kSafetyIsDistance)
      if (by_safety) t.carry = (double(saf) - double(g)) - l - carry_margin(t.p);
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  fp64_steps++;
  return step_fp64<S>(s, t, l);
}

template <class S, int Mode>
VGM_DEV void run_tile(const TransportArgs &a, const SolidRec &sr, const typename S::Data &s, const Tile &tl,
                      uint32_t &fp64_steps)
{
  const float scale = float(sr.scale);
  for (int64_t i = tl.begin + threadIdx.x; i < tl.end; i += blockDim.x) {
    TrackState t;
    t.p = V3<double>{a.x[i], a.y[i], a.z[i]};
    t.d = V3<double>{a.dx[i], a.dy[i], a.dz[i]};
    t.in = a.side[i] != 0;
    t.crossings = 0;
    t.carry = 0.0;
    // after a step: a new direction unless the step crossed, and the
    // periodic world for tracks outside (a wrap invalidates the carried bound)
    auto finish = [&](uint32_t step, bool crossed) {
      if (!crossed) t.d = new_direction(a.seed, uint64_t(i), step, t.d);
      if (!t.in) {
        const V3<double> q = t.p;
        t.p.x = wrap(t.p.x, sr.wc[0], sr.wh[0]);
        t.p.y = wrap(t.p.y, sr.wc[1], sr.wh[1]);
        t.p.z = wrap(t.p.z, sr.wc[2], sr.wh[2]);
        if (t.p.x != q.x || t.p.y != q.y || t.p.z != q.z) t.carry = 0.0;
      }
    };
    int k = 0;
    while (k < a.nsteps) {
      uint32_t step = uint32_t(a.step0 + k);
      double l = physics_step(sr.lambda, a.seed, uint64_t(i), step);
      if constexpr (S::kSafetyIsDistance) {
        // Steps covered by the carried bound need no geometry query.  Each
        // lane runs through its own run of them before the next query, so
        // the lanes of a wave meet at the expensive step instead of idling
        // through each other's cheap ones.
        while (l <= t.carry) {
          const double m = carry_margin(t.p);
          t.p.x = t.p.x + l * t.d.x;
          t.p.y = t.p.y + l * t.d.y;
          t.p.z = t.p.z + l * t.d.z;
          t.carry = t.carry - l - m;
          finish(step, false);
          if (++k >= a.nsteps) break;
          step = uint32_t(a.step0 + k);
          l = physics_step(sr.lambda, a.seed, uint64_t(i), step);
        }
        if (k >= a.nsteps) break;
      }
      const bool crossed =
          (Mode == kTransportFP64) ? step_fp64<S>(s, t, l) : step_mixed<S>(s, t, l, scale, fp64_steps);
      finish(step, crossed);
      ++k;
    }
    a.x[i] = t.p.x;
    a.y[i] = t.p.y;
    a.z[i] = t.p.z;
    a.dx[i] = t.d.x;
    a.dy[i] = t.d.y;
    a.dz[i] = t.d.z;
    a.side[i] = t.in ? 1 : 0;
    a.crossings[i] += t.crossings;
  }
}

}  // namespace

// One shape family, one precision mode.  Tiles [t0, t1) of the type-sorted
// queue belong to this family.  With a counter the launch is persistent:
// the grid is sized to the family's share of the machine and workgroups pull
// tiles until the range is drained.  Without one, workgroup b runs tile t0+b.
// Every family has its own kernel, so each gets the register allocation of
// its own code path rather than the largest one of all families.
template <class S, int Mode, bool Staged>
__global__ __launch_bounds__(kTransportMaxBlock) void transport_family_kernel(TransportArgs a, int t0, int t1,
                                                                              int *counter)
{
  extern __shared__ double lds_block[];
  __shared__ int tile_s;
  uint32_t fp64_steps = 0;
  int staged_solid = -1;
  for (int k = 0;; ++k) {
    int t;
    if (!counter) {
      if (k > 0) break;
      t = t0 + int(blockIdx.x);
    } else {
      if (threadIdx.x == 0) tile_s = t0 + atomicAdd(counter, 1);
      __syncthreads();
      t = tile_s;
      __syncthreads();
    }
    if (t >= t1) break;
    const Tile tl = a.tiles[t];
    const SolidRec sr = a.solids[tl.solid];
    if constexpr (Staged) {
      if (tl.solid != staged_solid) {  // uniform across the workgroup
        __syncthreads();
        stage_block(lds_block, a.pool + sr.off, sr.nd);
        __syncthreads();
        staged_solid = tl.solid;
      }
      run_tile<S, Mode>(a, sr, *reinterpret_cast<const typename S::Data *>(lds_block), tl, fp64_steps);
    } else {
      run_tile<S, Mode>(a, sr, *reinterpret_cast<const typename S::Data *>(a.pool + sr.off), tl, fp64_steps);
    }
  }
  if (Mode != kTransportFP64 && a.fp64_steps && fp64_steps) atomicAdd(a.fp64_steps, (unsigned long long)fp64_steps);
}

namespace {
struct Classifier {
  const TransportArgs &a;
  const SolidRec &sr;
  const Tile &tl;
  template <class S>
  __device__ void operator()()
  {
    const typename S::Data &s = *reinterpret_cast<const typename S::Data *>(a.pool + sr.off);
    for (int64_t i = tl.begin + threadIdx.x; i < tl.end; i += blockDim.x)
      a.side[i] = S::template contains<double>(s, V3<double>{a.x[i], a.y[i], a.z[i]}) ? 1 : 0;
  }
};
}  // namespace

// Initial inside/outside flag of every track: Contains on its solid.
__global__ __launch_bounds__(kTransportMaxBlock) void classify_kernel(TransportArgs a)
{
  __shared__ int tile_s;
  for (;;) {
    if (threadIdx.x == 0) tile_s = atomicAdd(a.counter, 1);
    __syncthreads();
    const int t = tile_s;
    __syncthreads();
    if (t >= a.ntiles) break;
    const Tile tl = a.tiles[t];
    const SolidRec sr = a.solids[tl.solid];
    Classifier c{a, sr, tl};
    dispatch_shape_type(sr.type, c);
  }
}

__global__ void reset_counters_kernel(int *counters, int n, unsigned long long *fp64_steps)
{
  const int i = int(threadIdx.x);
  if (i < n) counters[i] = 0;
  if (i == 0 && fp64_steps) *fp64_steps = 0;
}

namespace {
struct FamilyLauncher {
  hipStream_t stream;
  int mode;
  const TransportArgs &a;
  const FamilyLaunch &f;
  int block;
  template <class S>
  void operator()()
  {
    const int grid = f.counter ? f.grid : (f.t1 - f.t0);
    if (grid <= 0) return;
    const bool staged = S::kStage && f.stage_nd > 0 && size_t(f.stage_nd) * sizeof(double) <= size_t(kMaxStageBytes);
    const size_t lds = staged ? size_t(f.stage_nd) * sizeof(double) : 0;
    if (mode == kTransportMixed) {
      if (staged)
        hipLaunchKernelGGL((transport_family_kernel<S, kTransportMixed, S::kStage>), dim3(grid), dim3(block), lds,
                           stream, a, f.t0, f.t1, f.counter);
      else
        hipLaunchKernelGGL((transport_family_kernel<S, kTransportMixed, false>), dim3(grid), dim3(block), 0, stream,
                           a, f.t0, f.t1, f.counter);
    } else {
      if (staged)
        hipLaunchKernelGGL((transport_family_kernel<S, kTransportFP64, S::kStage>), dim3(grid), dim3(block), lds,
                           stream, a, f.t0, f.t1, f.counter);
      else
        hipLaunchKernelGGL((transport_family_kernel<S, kTransportFP64, false>), dim3(grid), dim3(block), 0, stream,
                           a, f.t0, f.t1, f.counter);
    }
  }
};
}  // namespace

hipError_t launch_reset(hipStream_t stream, int *counters, int n, unsigned long long *fp64_steps)
{
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters, n, fp64_steps);
  return hipGetLastError();
}

hipError_t launch_family(hipStream_t stream, int mode, const TransportArgs &a, const FamilyLaunch &f, int block)
{
  FamilyLauncher l{stream, mode, a, f, block};
  if (!dispatch_shape_type(f.type, l)) return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_classify(hipStream_t stream, const TransportArgs &a, int grid, int block)
{
  hipLaunchKernelGGL(classify_kernel, dim3(grid), dim3(block), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vgm
