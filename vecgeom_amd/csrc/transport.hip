// Batched heterogeneous track-step transport: the family-independent
// kernels (initial classification, counters) and the dispatch of a family's
// launch to its kernels, which transport_kernels.hpp defines and
// transport_family.hip instantiates once per shape family.
#include "vgm_internal.h"
#include "transport.h"
#include "stepping.hpp"
#include "shapes/all_shapes.hpp"

namespace vgm {

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

__global__ void reset_counters_kernel(int *counters, int n, unsigned long long *fp64_steps,
                                      unsigned long long *learned)
{
  const int i = int(threadIdx.x);
  if (i < n) counters[i] = 0;
  if (i == 0 && fp64_steps) *fp64_steps = 0;
  if (i < 4 && learned) learned[i] = 0;
}

namespace {
// Calls the family's launcher, compiled in its own translation unit.
struct FamilyDispatch {
  hipStream_t stream;
  const TransportArgs &a;
  const FamilyLaunch &f;
  int block;
  hipError_t err = hipSuccess;
  template <class S>
  void operator()()
  {
    err = launch_family_of<S>(stream, a, f, block);
  }
};
}  // namespace


namespace {
struct LearnableQuery {
  bool r = false;
  template <class S>
  void operator()()
  {
    r = S::kSafetyIsDistance;
  }
};
}  // namespace

bool family_learnable(int type)
{
  LearnableQuery q;
  return dispatch_shape_type(type, q) && q.r;
}

__global__ void set_step0_kernel(int32_t *dst, int32_t v) { *dst = v; }

hipError_t launch_set_step0(hipStream_t stream, int32_t *dst, int32_t v)
{
  hipLaunchKernelGGL(set_step0_kernel, dim3(1), dim3(1), 0, stream, dst, v);
  return hipGetLastError();
}

hipError_t launch_reset(hipStream_t stream, int *counters, int n, unsigned long long *fp64_steps,
                        unsigned long long *learned)
{
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters, n, fp64_steps, learned);
  return hipGetLastError();
}

hipError_t launch_family(hipStream_t stream, const TransportArgs &a, const FamilyLaunch &f, int block)
{
  if (f.mode == kFamLearned && f.has_mlp &&
      (!a.mlps || !a.learned || !a.queue || !a.qcount || !a.tcarry || !a.tflags || a.fallback_grid <= 0 ||
       block > kTransportMaxBlock || block % 64))
    return hipErrorInvalidValue;
  FamilyDispatch d{stream, a, f, block};
  if (!dispatch_shape_type(f.type, d)) return hipErrorInvalidValue;
  return d.err;
}

hipError_t launch_classify(hipStream_t stream, const TransportArgs &a, int grid, int block)
{
  hipLaunchKernelGGL(classify_kernel, dim3(grid), dim3(block), 0, stream, a);
  return hipGetLastError();
}

}  // namespace vgm
