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
//   MIXED  a guarded float safety settles the steps it covers, and the
//          boundary distance comes from the mixed kernels (float selects the
//          surface, FP64 evaluates it): the same decisions and positions as
//          FP64, bit for bit.
//
// This header holds the kernel templates; each shape family instantiates
// them in a translation unit of its own (transport_family.hip, built once per
// family), so the families compile in parallel.  transport.hip holds the
// family-independent kernels and the dispatch on the run-time type.
#pragma once
#include "vgm_internal.h"
#include "transport.h"
#include "stepping.hpp"
#include "shapes/all_shapes.hpp"
#include "staging.hpp"
#include "mlp_serve.hpp"

namespace vgm {

namespace {

// Index of the first step of this launch (the RNG counter of step 0).
VGM_DEV int32_t first_step(const TransportArgs &a) { return a.step0_dev ? *a.step0_dev : a.step0; }

struct TrackState {
  V3<double> p, d;
  bool in;
  int32_t crossings;
  double carry;  // lower bound on the safety at p (0: none), see step_fp64
};

// DistanceToOut from inside, DistanceToIn from outside.  A family that
// provides boundary_distance runs both in one code path, so a wave whose
// lanes sit on both sides walks once instead of once per side.
template <class S, class = void>
struct HasBoundaryDistance : std::false_type {};
template <class S>
struct HasBoundaryDistance<S, std::void_t<decltype(&S::template boundary_distance<double>)>> : std::true_type {};

template <class S>
VGM_DEV double boundary_distance(const typename S::Data &s, const V3<double> &p, const V3<double> &d, bool in)
{
  if constexpr (HasBoundaryDistance<S>::value)
    return S::template boundary_distance<double>(s, p, d, in);
  else
    return in ? S::template distance_to_out<double>(s, p, d, kInfLength)
              : S::template distance_to_in<double>(s, p, d, kInfLength);
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
  double db = boundary_distance<S>(s, t.p, t.d, t.in);
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

// Absolute guard for float safeties: a few float ulps of the largest
// coordinate magnitude in play (a float safety is a short closed form or a
// minimum of such, so its error stays well inside).
VGM_DEV float guard_band(const V3<float> &p, float scale)
{
  return 4.0e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z) + scale);
}

// Mixed-precision track-step: the decisions and the position arithmetic of
// the FP64 step, with float doing the bulk of the work.
//  - kFloatSafety / kSafetyIsDistance families: the float safety less the
//    guard band is a lower bound of the FP64 safety, so when it covers the
//    step the FP64 step moves freely too, and so does this one (the side of
//    a surface-distance family is checked in FP64, as in the FP64 step);
//  - otherwise the boundary distance comes first (where distance_first_ok
//    holds), from the mixed kernels where the family has them (float
//    selects the surface, FP64 evaluates it: the FP64 distance bit for
//    bit), else in FP64;
//  - the remaining families take the FP64 step.
// Trajectories are the FP64 ones bit for bit.  fp64_steps counts the steps
// the float safety did not settle.
template <class S>
VGM_DEV bool step_distance_first(const typename S::Data &s, TrackState &t, double l, bool &verified,
                                 float scale = 0.0f, bool mixed_distance = false);

template <class S>
VGM_DEV bool step_mixed(const typename S::Data &s, TrackState &t, double l, float scale, uint32_t &fp64_steps)
{
  if constexpr (S::kSafetyIsDistance) {
    if (l <= t.carry) return step_fp64<S>(s, t, l);  // under the carried bound: moves without a query
  }
  if constexpr (S::kFloatSafety || S::kSafetyIsDistance) {
    const V3<float> pf{float(t.p.x), float(t.p.y), float(t.p.z)};
    const float g = guard_band(pf, scale);
    const float lf = float(l);  // g covers its rounding
    float saf;
    if constexpr (S::kSafetyIsDistance) {
      saf = S::template surface_distance<float>(s, pf, t.in);
      if (lf + g <= saf && S::template contains<double>(s, t.p) != t.in) saf = -1.0f;
    } else {
      saf = t.in ? S::template safety_to_out<float>(s, pf) : S::template safety_to_in<float>(s, pf);
    }
    if (lf + g <= saf) {
      t.carry = 0.0;
      if constexpr (S::kSafetyIsDistance) t.carry = (double(saf) - double(g)) - l - carry_margin(t.p);
      t.p.x = t.p.x + l * t.d.x;
      t.p.y = t.p.y + l * t.d.y;
      t.p.z = t.p.z + l * t.d.z;
      return false;
    }
  }
  fp64_steps++;
  if constexpr (kDistanceFirstAny<S>) {
    if (distance_first_ok<S>(s, t.p, t.in)) {
      bool verified = false;
      return step_distance_first<S>(s, t, l, verified, scale, S::kMixedDistance);
    }
  }
  return step_fp64<S>(s, t, l);
}

// Exact track-step, boundary distance first.  The FP64 step moves freely
// when l <= safety or l < d_b and crosses otherwise.  For families whose
// safety is a lower bound of the distance to their surface along any
// direction (kSafetyBelowDistance: SafetyToIn <= DistanceToIn and
// SafetyToOut <= DistanceToOut up to the tolerance of the surface tests),
// the step is free either way when l < d_b and crosses either way when
// l > d_b + kDistanceFirstWindow; only in between is the safety (with its
// side check) computed.  Same decisions and arithmetic as step_fp64; only
// the carried bound differs.  `verified` is set when the side check ran and
// agreed (see learned_bound_kernel).  Launches of other families ignore the
// distance-first request.
// With mixed_distance the boundary distance comes from the family's mixed
// kernels, which return the FP64 value.
template <class S>
VGM_DEV bool step_distance_first(const typename S::Data &s, TrackState &t, double l, bool &verified, float scale,
                                 bool mixed_distance)
{
  double db;
  if constexpr (S::kMixedDistance) {
    if (mixed_distance)
      db = t.in ? S::distance_to_out_mixed(s, t.p, t.d, scale) : S::distance_to_in_mixed(s, t.p, t.d, scale);
    else
      db = boundary_distance<S>(s, t.p, t.d, t.in);
  } else {
    db = boundary_distance<S>(s, t.p, t.d, t.in);
  }
  if (db < 0.0) db = 0.0;
  if (l < db) {
    t.carry = 0.0;
    t.p.x = t.p.x + l * t.d.x;
    t.p.y = t.p.y + l * t.d.y;
    t.p.z = t.p.z + l * t.d.z;
    return false;
  }
  if (l <= db + kDistanceFirstWindow) {
    if constexpr (S::kSafetyIsDistance) {
      const double saf = S::template surface_distance<double>(s, t.p, t.in);
      if (l <= saf && S::template contains<double>(s, t.p) == t.in) {
        verified = true;
        t.carry = saf - l - carry_margin(t.p);
        t.p.x = t.p.x + l * t.d.x;
        t.p.y = t.p.y + l * t.d.y;
        t.p.z = t.p.z + l * t.d.z;
        return false;
      }
    } else {
      const double saf =
          t.in ? S::template safety_to_out<double>(s, t.p) : S::template safety_to_in<double>(s, t.p);
      if (l <= saf) {
        t.p.x = t.p.x + l * t.d.x;
        t.p.y = t.p.y + l * t.d.y;
        t.p.z = t.p.z + l * t.d.z;
        return false;
      }
    }
  }
  t.carry = 0.0;
  const double m = db + kBoundaryPush;
  t.p.x = t.p.x + m * t.d.x;
  t.p.y = t.p.y + m * t.d.y;
  t.p.z = t.p.z + m * t.d.z;
  t.in = !t.in;
  t.crossings++;
  return true;
}

// Step order of an FP64 family kernel, fixed at compile time so that each
// kernel holds the registers of one order only: the FP64 step, distance
// first (where distance_first_ok holds; the FP64 step elsewhere), or
// distance first with the bounding-box cull.
enum StepOrder : int { kOrderFP64 = 0, kOrderDistFirst = 1, kOrderCulled = 2 };

// A step that moves by l without any shape query: under the carried bound of
// a surface-distance family, or (culled order) an outside move clear of the
// bounding box.
template <class S, int Order>
VGM_DEV bool cheap_step(const SolidRec &sr, const TrackState &t, double l)
{
  if constexpr (S::kSafetyIsDistance)
    if (l <= t.carry) return true;
  if constexpr (Order == kOrderCulled) return !t.in && move_clears_box(sr.bc, sr.bh, t.p, t.d, l);
  return false;
}

template <class S, int Mode, int Order, bool Staged>
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
      uint32_t step = uint32_t(first_step(a) + k);
      double l = physics_step(sr.lambda, a.seed, uint64_t(i), step);
      // Steps that need no geometry query (under the carried bound, or clear
      // of the bounding box).  Each lane runs through its own run of them
      // before the next query, so the lanes of a wave meet at the expensive
      // step instead of idling through each other's cheap ones.
      while (cheap_step<S, Order>(sr, t, l)) {
        const double m = carry_margin(t.p);
        t.p.x = t.p.x + l * t.d.x;
        t.p.y = t.p.y + l * t.d.y;
        t.p.z = t.p.z + l * t.d.z;
        if constexpr (S::kSafetyIsDistance) t.carry = t.carry - l - m;
        finish(step, false);
        if (++k >= a.nsteps) break;
        step = uint32_t(first_step(a) + k);
        l = physics_step(sr.lambda, a.seed, uint64_t(i), step);
      }
      if (k >= a.nsteps) break;
      bool crossed;
      // a block staged in LDS is re-read there each step rather than held in
      // registers across the track loop (which spills it to scratch)
      const typename S::Data &ss = Staged ? opaque_ref(s) : s;
      if constexpr (Mode == kTransportFP64) {
        bool verified = false;  // only the learned tiles read it
        if constexpr (Order != kOrderFP64 && S::kSafetyBelowDistance)
          crossed = step_distance_first<S>(ss, t, l, verified);
        else if constexpr (Order != kOrderFP64 && kDistanceFirstAny<S>)
          crossed = distance_first_ok<S>(ss, t.p, t.in) ? step_distance_first<S>(ss, t, l, verified)
                                                        : step_fp64<S>(ss, t, l);
        else
          crossed = step_fp64<S>(ss, t, l);
      } else {
        crossed = step_mixed<S>(ss, t, l, scale, fp64_steps);
      }
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
template <class S, int Mode, bool Staged, int Order>
__global__ __launch_bounds__(kTransportMaxBlock) __attribute__((amdgpu_waves_per_eu(S::kWavesPerEU))) void
transport_family_kernel(TransportArgs a, int t0, int t1, int *counter)
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
      // the same value in every lane: scalar registers, so the tile and solid
      // records below load into scalar registers too
      t = __builtin_amdgcn_readfirstlane(tile_s);
      __syncthreads();
    }
    if (t >= t1) break;
    const Tile tl = a.tiles[t];
    const SolidRec sr = a.solids[tl.solid];
    const typename S::Data *sp = reinterpret_cast<const typename S::Data *>(a.pool + sr.off);
    if constexpr (Staged) {
      if (tl.solid != staged_solid) {  // uniform across the workgroup
        __syncthreads();
        stage_block(lds_block, a.pool + sr.off, sr.nd);
        __syncthreads();
        staged_solid = tl.solid;
      }
      sp = reinterpret_cast<const typename S::Data *>(lds_block);
    }
    run_tile<S, Mode, Order, Staged>(a, sr, *sp, tl, fp64_steps);
  }
  if (Mode != kTransportFP64 && a.fp64_steps && fp64_steps) atomicAdd(a.fp64_steps, (unsigned long long)fp64_steps);
}

// ---- learned safety: one track-step per launch pair ------------------------
//
// A learned family advances its tracks one track-step at a time with two
// kernels per step:
//   learned_bound_kernel     every track of the family: a step under the
//                            carried bound moves at once; otherwise the
//                            network evaluates a bound for the whole wave on
//                            the matrix cores and a track whose step fits
//                            under it moves; the others are appended to a
//                            compacted queue
//   learned_fallback_kernel  the queued tracks: the exact distance-first step
//                            (or the Contains that verifies a bound's side)
// so the latency-bound facet walks run on full waves at the occupancy of
// their own kernel, and the network's kernel does no walks.  Between the two
// (and between steps) a track's carried bound and flags live in HBM.
//
// A bound at or below the exact safety moves where the FP64 step moves, so
// trajectories are those of the FP64 path.  The exact safety is -1 when the
// point is not on the side of the track's flag (after the periodic wrap puts
// a track down inside the solid, or right after a crossing); the FP64 step
// then takes the boundary-distance path, which restores the flag.  A bound is
// only a distance, so it moves a track only once its side is verified: by a
// Contains that agrees with the flag (in the fallback kernel); moves shorter
// than the distance to the surface keep it verified, a wrap or a crossing
// clears it.  Audit runs count any bound above the exact surface distance.

constexpr uint8_t kTrackVerified = 1;

struct LearnedCounts {
  uint32_t q = 0, served = 0, fallbacks = 0, violations = 0;
};

__device__ __forceinline__ void add_learned_counts(const TransportArgs &a, const LearnedCounts &cnt)
{
  uint32_t v[4] = {cnt.q, cnt.served, cnt.fallbacks, cnt.violations};
  for (int c = 0; c < 4; ++c) {
    uint32_t x = v[c];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63u) == 0 && x) atomicAdd(a.learned + c, (unsigned long long)x);
  }
}

// New direction unless the step crossed; the periodic world for tracks
// outside (a wrap clears the carried bound and the side verification).
__device__ __forceinline__ void finish_step(const TransportArgs &a, const SolidRec &sr, uint64_t i, uint32_t step,
                                            bool crossed, TrackState &t, bool &verified)
{
  if (!crossed) t.d = new_direction(a.seed, i, step, t.d);
  if (crossed) verified = false;
  if (!t.in) {
    const V3<double> q = t.p;
    t.p.x = wrap(t.p.x, sr.wc[0], sr.wh[0]);
    t.p.y = wrap(t.p.y, sr.wc[1], sr.wh[1]);
    t.p.z = wrap(t.p.z, sr.wc[2], sr.wh[2]);
    if (t.p.x != q.x || t.p.y != q.y || t.p.z != q.z) {
      t.carry = 0.0;
      verified = false;
    }
  }
}

// Step k of the tiles [t0, t1) of one family, one workgroup per tile.  The
// audit variant also computes the exact surface distance of every query.
template <class S, bool Audit>
__global__ __launch_bounds__(kTransportMaxBlock) void learned_bound_kernel(TransportArgs a, int t0, int t1, int k)
{
  __shared__ mlp_act_rows act_buf[kTransportMaxBlock / 64];
  const int t = t0 + int(blockIdx.x);
  if (t >= t1) return;
  const Tile tl = a.tiles[t];
  const SolidRec sr = a.solids[tl.solid];
  const typename S::Data &s = *reinterpret_cast<const typename S::Data *>(a.pool + sr.off);
  const bool net = sr.mlp >= 0;
  MlpDesc md{};
  if (net) md = a.mlps[sr.mlp];
  const uint32_t step = uint32_t(first_step(a) + k);
  const int lane = int(threadIdx.x & 63u);
  LearnedCounts cnt;
  for (int64_t base = tl.begin + int64_t(threadIdx.x & ~63u); base < tl.end; base += blockDim.x) {
    const int64_t i = base + lane;
    const bool act = i < tl.end;
    TrackState t;
    t.p = act ? V3<double>{a.x[i], a.y[i], a.z[i]} : V3<double>{sr.wc[0], sr.wc[1], sr.wc[2]};
    t.d = act ? V3<double>{a.dx[i], a.dy[i], a.dz[i]} : V3<double>{0.0, 0.0, 1.0};
    t.in = act && a.side[i] != 0;
    t.crossings = 0;
    t.carry = (act && k > 0) ? a.tcarry[i] : 0.0;
    bool verified = act && k > 0 && (a.tflags[i] & kTrackVerified);
    const double l = act ? physics_step(sr.lambda, a.seed, uint64_t(i), step) : 0.0;
    // no query under the carried bound or for an outside move clear of the box
    const bool cheap = act && (l <= t.carry || (a.cull && !t.in && move_clears_box(sr.bc, sr.bh, t.p, t.d, l)));
    const bool want = act && !cheap && net;
    double b = -1.0;
    if (__any(want)) {  // uniform: the whole wave evaluates the network
      const float y = mlp_eval_wave(*md.img, act_buf[threadIdx.x >> 6], mlp_coord(t.p.x, md.c[0], md.inv_hw[0]),
                                    mlp_coord(t.p.y, md.c[1], md.inv_hw[1]), mlp_coord(t.p.z, md.c[2], md.inv_hw[2]));
      b = double(y) * md.scale - md.margin;
    }
    bool moved = false, queued = false;
    if (act) {
      if (cheap) {
        moved = true;
      } else if (want) {
        ++cnt.q;
        if constexpr (Audit)
          if (b > S::template surface_distance<double>(s, t.p, t.in)) ++cnt.violations;
        if (verified) {
          if (b > t.carry) t.carry = b;
          moved = l <= t.carry;
          cnt.served += moved ? 1u : 0u;
        } else if (l <= b) {
          t.carry = b;  // a bound for the fallback kernel to use once the side is verified
        }
        queued = !moved;
      } else {
        queued = true;
      }
    }
    if (moved) {
      const double m = carry_margin(t.p);
      t.p.x = t.p.x + l * t.d.x;
      t.p.y = t.p.y + l * t.d.y;
      t.p.z = t.p.z + l * t.d.z;
      t.carry = t.carry - l - m;
      finish_step(a, sr, uint64_t(i), step, false, t, verified);
    }
    // append the queued tracks, one atomic per wave
    const uint64_t qm = __ballot(queued);
    if (qm) {
      int slot0 = 0;
      if (lane == 0) slot0 = atomicAdd(a.qcount + (k & 1), __popcll(qm));
      slot0 = __shfl(slot0, 0);
      if (queued) {
        const int slot = slot0 + __popcll(qm & ((1ull << lane) - 1ull));
        a.queue[slot] = (uint64_t(uint32_t(tl.solid)) << 40) | uint64_t(i);
      }
    }
    if (act) {
      if (moved) {
        a.x[i] = t.p.x;
        a.y[i] = t.p.y;
        a.z[i] = t.p.z;
        a.dx[i] = t.d.x;
        a.dy[i] = t.d.y;
        a.dz[i] = t.d.z;
      }
      a.tcarry[i] = t.carry;
      a.tflags[i] = verified ? kTrackVerified : 0;
    }
  }
  add_learned_counts(a, cnt);
}

// Step k of the queued tracks (grid-stride over the queue).  Also clears the
// queue counter of step k + 1, which the next bound kernel fills.
template <class S>
__global__ __launch_bounds__(kTransportMaxBlock) __attribute__((amdgpu_waves_per_eu(S::kWavesPerEU))) void
learned_fallback_kernel(TransportArgs a, int k)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) a.qcount[(k + 1) & 1] = 0;
  const int nq = a.qcount[k & 1];
  const uint32_t step = uint32_t(first_step(a) + k);
  LearnedCounts cnt;
  for (int j = int(blockIdx.x * blockDim.x + threadIdx.x); j < nq; j += int(gridDim.x * blockDim.x)) {
    const uint64_t e = a.queue[j];
    const int64_t i = int64_t(e & ((uint64_t(1) << 40) - 1));
    const SolidRec sr = a.solids[int(e >> 40)];
    const typename S::Data &s = *reinterpret_cast<const typename S::Data *>(a.pool + sr.off);
    TrackState t;
    t.p = V3<double>{a.x[i], a.y[i], a.z[i]};
    t.d = V3<double>{a.dx[i], a.dy[i], a.dz[i]};
    t.in = a.side[i] != 0;
    t.crossings = 0;
    t.carry = a.tcarry[i];
    bool verified = (a.tflags[i] & kTrackVerified) != 0;
    const double l = physics_step(sr.lambda, a.seed, uint64_t(i), step);
    bool crossed = false;
    if (!verified && l <= t.carry && S::template contains<double>(s, t.p) == t.in) {
      verified = true;  // the bound of the bound kernel now applies
      const double m = carry_margin(t.p);
      t.p.x = t.p.x + l * t.d.x;
      t.p.y = t.p.y + l * t.d.y;
      t.p.z = t.p.z + l * t.d.z;
      t.carry = t.carry - l - m;
      ++cnt.served;
    } else {
      if (!verified) t.carry = 0.0;
      crossed = step_distance_first<S>(s, t, l, verified);
      ++cnt.fallbacks;
    }
    finish_step(a, sr, uint64_t(i), step, crossed, t, verified);
    a.x[i] = t.p.x;
    a.y[i] = t.p.y;
    a.z[i] = t.p.z;
    a.dx[i] = t.d.x;
    a.dy[i] = t.d.y;
    a.dz[i] = t.d.z;
    a.side[i] = t.in ? 1 : 0;
    if (crossed) a.crossings[i] += 1;
    a.tcarry[i] = t.carry;
    a.tflags[i] = verified ? kTrackVerified : 0;
  }
  add_learned_counts(a, cnt);
}

namespace {
struct FamilyLauncher {
  hipStream_t stream;
  const TransportArgs &a;
  const FamilyLaunch &f;
  int block;
  hipError_t err = hipSuccess;

  template <class S, int Mode, int Order>
  void go(const TransportArgs &args)
  {
    const int grid = f.counter ? f.grid : (f.t1 - f.t0);
    if (grid <= 0) return;
    const bool staged = S::kStage && f.stage_nd > 0 && size_t(f.stage_nd) * sizeof(double) <= size_t(kMaxStageBytes);
    if (staged)
      hipLaunchKernelGGL((transport_family_kernel<S, Mode, S::kStage, Order>), dim3(grid), dim3(block),
                         size_t(f.stage_nd) * sizeof(double), stream, args, f.t0, f.t1, f.counter);
    else
      hipLaunchKernelGGL((transport_family_kernel<S, Mode, false, Order>), dim3(grid), dim3(block), 0, stream, args,
                         f.t0, f.t1, f.counter);
  }

  template <class S>
  void operator()()
  {
    TransportArgs args = a;
    const bool culled = f.mode == kFamCulled || f.mode == kFamLearned;
    args.cull = culled ? 1 : 0;
    if (f.mode == kFamMixed) {
      go<S, kTransportMixed, kOrderFP64>(args);
      return;
    }
    if constexpr (S::kSafetyIsDistance) {
      if (f.mode == kFamLearned && f.has_mlp) {
        args.queue = a.queue + f.queue_off;
        args.qcount = a.qcount + 2 * f.type;
        const unsigned tiles = unsigned(f.t1 - f.t0);
        for (int k = 0; k < a.nsteps && tiles > 0; ++k) {
          if (a.audit)
            hipLaunchKernelGGL((learned_bound_kernel<S, true>), dim3(tiles), dim3(block), 0, stream, args, f.t0, f.t1,
                               k);
          else
            hipLaunchKernelGGL((learned_bound_kernel<S, false>), dim3(tiles), dim3(block), 0, stream, args, f.t0,
                               f.t1, k);
          hipLaunchKernelGGL(learned_fallback_kernel<S>, dim3(unsigned(a.fallback_grid)), dim3(kTransportMaxBlock),
                             0, stream, args, k);
        }
        return;
      }
    }
    // distance first is the FP64 step for families whose safety may exceed
    // their distance everywhere
    constexpr int kDistFirst = kDistanceFirstAny<S> ? kOrderDistFirst : kOrderFP64;
    if (f.mode == kFamCulled)
      go<S, kTransportFP64, kOrderCulled>(args);
    else if (f.mode == kFamDistFirst)
      go<S, kTransportFP64, kDistFirst>(args);
    else
      go<S, kTransportFP64, kOrderFP64>(args);
  }
};

}  // namespace

// Launches one family's share of a transport launch (see launch_family).
template <class S>
hipError_t launch_family_of(hipStream_t stream, const TransportArgs &a, const FamilyLaunch &f, int block)
{
  FamilyLauncher l{stream, a, f, block};
  l.template operator()<S>();
  return hipGetLastError();
}

}  // namespace vgm
