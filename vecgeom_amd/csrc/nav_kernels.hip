// Navigation kernels: one lane per track.
//   locate     GlobalLocator::LocateGlobalPoint from the world (top = true)
//   step       one ComputeStepAndPropagatedState per track (with a safety
//              output: ComputeStepAndSafetyAndPropagatedState)
//   transport  nsteps of it with an unlimited physics step, moving the global
//              point by each step; a track that leaves the world stops
//              (persistent grid over a work queue of tracks)
//   safety     VSafetyEstimator::ComputeSafety (mother SafetyToOut, then the
//              daughters through the BVH)
// States and geometry tables live in HBM; a track's point, direction and
// state stay in registers for the whole transport launch.
#include "nav_kernels.h"
#include "navigation.hpp"

namespace vgm {

constexpr int kNavBlock = 128;

static_assert(kNavBlock <= kNavWalkThreads, "walk stack columns");

// Kernel variants by the set of shape families in the geometry: the barrel
// of tubes and boxes needs neither the polygon nor the facet code.
constexpr unsigned kMaskBoxTube = kShapeBit(kBox) | kShapeBit(kTube);
constexpr unsigned kMaskSimple = kShapeBit(kBox) | kShapeBit(kTube) | kShapeBit(kCone) | kShapeBit(kTrd) |
                                 kShapeBit(kTrapezoid) | kShapeBit(kOrb) |
                                 kShapeBit(kCutTube);

// Waves per SIMD the step kernels of the box-and-tube variant ask the
// compiler for: the BVH and shape walks hide their latency better at 5 waves
// than at the 3 the unconstrained register allocation gives (barrel transport
// 0.48 -> 0.62 G steps/s), even with the spills that brings.  The variants
// with more families call out-of-line shape code and are left unconstrained.
template <unsigned Mask>
constexpr int kNavWavesPerEU = Mask == kMaskBoxTube ? 5 : 1;

template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) void nav_locate_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  Navigator<Mask> nav{a.g};
  a.state[i] = nav.locate_global(V3<double>{a.x[i], a.y[i], a.z[i]});
  if (a.last) a.last[i] = 0;
  if (a.boundary) a.boundary[i] = 0;
}

template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kNavWavesPerEU<Mask>))) void nav_step_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  if (!a.state[i]) {
    a.step[i] = 0.0;
    if (a.safety) a.safety[i] = 0.0;
    return;
  }
  Navigator<Mask> nav{a.g};
  const V3<double> p{a.x[i], a.y[i], a.z[i]};
  // ComputeStepAndSafetyAndPropagatedState: the safety of the start point first
  if (a.safety) a.safety[i] = (!a.calcsafety || a.calcsafety[i]) ? nav.safety(p, a.state[i]) : 0.0;
  const NavState in{a.state[i], a.last[i], a.boundary[i] != 0};
  NavState out;
  a.step[i] = nav.step(p, V3<double>{a.dx[i], a.dy[i], a.dz[i]}, a.step_limit ? a.step_limit[i] : kInfLength, in, out,
                       !a.no_relocate);
  a.state[i] = out.ind;
  a.last[i] = out.last;
  a.boundary[i] = out.boundary ? 1 : 0;
}

// Transports run as persistent grids over a work queue of tracks: a lane
// whose track has taken its steps (or left the world) stores it and takes
// the next one, so a wave stays full until the queue drains instead of
// waiting for its longest track (in the barrel, tracks leave the world after
// about a quarter of their 200 steps).  Each track's steps are the same as
// one lane per track would take, so the results are identical.
//
// The next track index for every lane of the wave that asks: one atomic per
// wave on the queue counter.
__device__ __forceinline__ long long pull_track(unsigned long long *queue, bool need)
{
  const unsigned long long mask = __ballot(need ? 1 : 0);
  if (!mask) return -1;
  const int lane = int(__lane_id());
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(queue, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  const unsigned long long below = mask & ((1ull << lane) - 1ull);
  return need ? (long long)(base + (unsigned long long)__popcll(below)) : -1;
}

__global__ void nav_queue_reset_kernel(unsigned long long *queue) { *queue = 0ull; }

// LevelLocate in the logical volume of state[i] for a point (and, when dx is
// given, a direction) in that volume's frame, skipping the daughter whose
// state is last[i] (0: none): the daughter's state and the point in its
// frame, or 0 and the point itself.
template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) void nav_level_locate_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  const V3<double> p{a.x[i], a.y[i], a.z[i]};
  const unsigned node = a.state[i];
  uint32_t child = 0u;
  V3<double> q = p;
  if (node) {
    Navigator<Mask> nav{a.g};
    const NavNode &N = a.g.nodes[node];
    const int lv = a.g.pvs[N.pv].lv;
    const unsigned ex = a.last ? a.last[i] : 0u;
    const int excl_pv = ex ? a.g.nodes[ex].pv : -1;
    V3<double> dp;
    const int k = a.dx ? nav.level_locate_dir(lv, excl_pv, p, V3<double>{a.dx[i], a.dy[i], a.dz[i]}, dp)
                       : nav.level_locate(lv, excl_pv, p, dp);
    if (k >= 0) {
      child = uint32_t(a.g.children[N.child0 + k]);
      q = dp;
    }
  }
  a.child[i] = child;
  a.ox[i] = q.x;
  a.oy[i] = q.y;
  a.oz[i] = q.z;
}

template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kNavWavesPerEU<Mask>))) void nav_transport_kernel(NavArgs a)
{
  Navigator<Mask> nav{a.g};
  const long long n = (long long)a.n;
  long long i = -1;
  NavState s{0u, 0u, false};
  V3<double> p{0.0, 0.0, 0.0}, d{0.0, 0.0, 0.0};
  int k = 0;
  for (;;) {
    const bool need = i < 0;
    const long long next = pull_track(a.queue, need);
    if (need) {
      if (next >= n) break;  // the queue is drained
      i = next;
      s = NavState{a.state[i], a.last[i], a.boundary[i] != 0};
      p = V3<double>{a.x[i], a.y[i], a.z[i]};
      d = V3<double>{a.dx[i], a.dy[i], a.dz[i]};
      k = 0;
    }
    if (k < a.nsteps && s.ind) {
      NavState out;
      const double st = nav.step(p, d, kInfLength, s, out);
      p.x = p.x + st * d.x;
      p.y = p.y + st * d.y;
      p.z = p.z + st * d.z;
      s = out;
      ++k;
    }
    if (k >= a.nsteps || !s.ind) {
      a.x[i] = p.x;
      a.y[i] = p.y;
      a.z[i] = p.z;
      a.state[i] = s.ind;
      a.last[i] = s.last;
      a.boundary[i] = s.boundary ? 1 : 0;
      a.steps[i] += k;
      i = -1;
    }
  }
}

// One step of the straight-line transport for every track still in the
// world: the body of the step loop as a kernel of its own, for the
// graph-captured multi-kernel loop (one launch per step; the state goes
// through HBM between steps).
template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kNavWavesPerEU<Mask>))) void nav_step_move_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n || !a.state[i]) return;
  Navigator<Mask> nav{a.g};
  const NavState s{a.state[i], a.last[i], a.boundary[i] != 0};
  V3<double> p{a.x[i], a.y[i], a.z[i]};
  const V3<double> d{a.dx[i], a.dy[i], a.dz[i]};
  NavState out;
  const double st = nav.step(p, d, kInfLength, s, out);
  p.x = p.x + st * d.x;
  p.y = p.y + st * d.y;
  p.z = p.z + st * d.z;
  a.x[i] = p.x;
  a.y[i] = p.y;
  a.z[i] = p.z;
  a.state[i] = out.ind;
  a.last[i] = out.last;
  a.boundary[i] = out.boundary ? 1 : 0;
  a.steps[i] += 1;
}

// Transport under a physics step limit L per track: each step is
// ComputeStepAndSafetyAndPropagatedState with calcsafety
// (VecGeom/navigation/VNavigator.h:543-592): the safety first, and when it
// covers L the track moves L with no boundary distance computed
// (`safetydone`); otherwise the step limited to L.  With reuse_safety the
// safety left after a safety-only move (safety - L, still a lower bound in
// the same volume) is carried, and the safety is recomputed only when that
// no longer covers L or after a geometry step.
template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) __attribute__((amdgpu_waves_per_eu(kNavWavesPerEU<Mask>))) void nav_transport_limited_kernel(NavArgs a)
{
  Navigator<Mask> nav{a.g};
  const long long n = (long long)a.n;
  long long i = -1;
  NavState s{0u, 0u, false};
  V3<double> p{0.0, 0.0, 0.0}, d{0.0, 0.0, 0.0};
  double L = 0.0, saf = -1.0;
  int k = 0, geo = 0;
  for (;;) {
    const bool need = i < 0;
    const long long next = pull_track(a.queue, need);
    if (need) {
      if (next >= n) break;  // the queue is drained
      i = next;
      s = NavState{a.state[i], a.last[i], a.boundary[i] != 0};
      p = V3<double>{a.x[i], a.y[i], a.z[i]};
      d = V3<double>{a.dx[i], a.dy[i], a.dz[i]};
      L = a.step_limit[i];
      saf = -1.0;
      k = 0;
      geo = 0;
    }
    if (k < a.nsteps && s.ind) {
      if (!(a.reuse_safety && saf >= L)) saf = nav.safety(p, s.ind);
      double st;
      if (saf >= L) {
        st = L;
        s.boundary = false;
        saf = saf - L;
      } else {
        NavState out;
        st = nav.step(p, d, L, s, out);
        s = out;
        saf = -1.0;
        ++geo;
      }
      p.x = p.x + st * d.x;
      p.y = p.y + st * d.y;
      p.z = p.z + st * d.z;
      ++k;
    }
    if (k >= a.nsteps || !s.ind) {
      a.x[i] = p.x;
      a.y[i] = p.y;
      a.z[i] = p.z;
      a.state[i] = s.ind;
      a.last[i] = s.last;
      a.boundary[i] = s.boundary ? 1 : 0;
      a.steps[i] += k;
      a.geo_steps[i] += geo;
      i = -1;
    }
  }
}

template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) void nav_approach_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n || !a.state[i]) return;
  Navigator<Mask> nav{a.g};
  const NavState s{a.state[i], a.last[i], false};
  a.step[i] = nav.approach(V3<double>{a.x[i], a.y[i], a.z[i]}, V3<double>{a.dx[i], a.dy[i], a.dz[i]}, a.step[i], s);
}

template <unsigned Mask>
__global__ __launch_bounds__(kNavBlock) void nav_safety_kernel(NavArgs a)
{
  const size_t i = size_t(blockIdx.x) * kNavBlock + threadIdx.x;
  if (i >= a.n) return;
  Navigator<Mask> nav{a.g};
  a.step[i] = nav.safety(V3<double>{a.x[i], a.y[i], a.z[i]}, a.state[i]);
}

namespace {
// Workgroups of a persistent transport: enough to keep every SIMD of the
// device at the kernels' occupancy, and no more than the tracks need.
unsigned persistent_blocks(const NavArgs &a, unsigned blocks)
{
  const unsigned full = unsigned(a.num_cus > 0 ? a.num_cus : 256) * 16u;
  return blocks < full ? blocks : full;
}

template <unsigned Mask>
void launch_nav_mask(hipStream_t stream, int kind, const NavArgs &a, unsigned blocks, bool &ok)
{
  ok = true;
  switch (kind) {
  case kNavLocate: hipLaunchKernelGGL(nav_locate_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a); break;
  case kNavStep: hipLaunchKernelGGL(nav_step_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a); break;
  case kNavTransport:
    hipLaunchKernelGGL(nav_queue_reset_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    hipLaunchKernelGGL(nav_transport_kernel<Mask>, dim3(persistent_blocks(a, blocks)), dim3(kNavBlock), 0, stream, a);
    break;
  case kNavSafety: hipLaunchKernelGGL(nav_safety_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a); break;
  case kNavApproach:
    hipLaunchKernelGGL(nav_approach_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a);
    break;
  case kNavStepMove:
    hipLaunchKernelGGL(nav_step_move_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a);
    break;
  case kNavLevelLocate:
    hipLaunchKernelGGL(nav_level_locate_kernel<Mask>, dim3(blocks), dim3(kNavBlock), 0, stream, a);
    break;
  case kNavTransportLimited:
    hipLaunchKernelGGL(nav_queue_reset_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    hipLaunchKernelGGL(nav_transport_limited_kernel<Mask>, dim3(persistent_blocks(a, blocks)), dim3(kNavBlock), 0,
                       stream, a);
    break;
  default: ok = false;
  }
}
}  // namespace

hipError_t launch_nav(hipStream_t stream, int kind, const NavArgs &a)
{
  if (a.n == 0) return hipSuccess;
  const unsigned blocks = unsigned((a.n + kNavBlock - 1) / kNavBlock);
  bool ok = false;
  if ((a.shape_mask & ~kMaskBoxTube) == 0)
    launch_nav_mask<kMaskBoxTube>(stream, kind, a, blocks, ok);
  else if ((a.shape_mask & ~kMaskSimple) == 0)
    launch_nav_mask<kMaskSimple>(stream, kind, a, blocks, ok);
  else
    launch_nav_mask<kAllShapes>(stream, kind, a, blocks, ok);
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace vgm
