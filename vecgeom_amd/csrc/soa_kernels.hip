// SOA3D batch kernels: one thread per point, coalesced FP64 streams.
// Replaces the per-point loops of SIMDSpecializedVolImplHelper
// (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:184-257) and
// UnplacedVolumeImplHelper (VecGeom/volumes/UnplacedVolumeImplHelper.h:34-67)
// and the CUDA benchmark kernels (test/VecGeomBenchmark/Benchmarker.cu:13-61).
//
// The kernels are templated on shape family and method, so nothing is
// dispatched per point; the shape parameters are wave-uniform and come in
// through scalar loads.
#include "vgm_internal.h"
#include "transform.hpp"
#include "shapes/all_shapes.hpp"
#include "staging.hpp"

namespace vgm {

constexpr int kSoaBlock = 256;

template <class S, int M, bool Staged>
__global__ __launch_bounds__(kSoaBlock) void soa_fp64_kernel(const typename S::Data *__restrict__ sd, int nd,
                                                             Transform tr, int use_tr, SoaArgs a)
{
  extern __shared__ double lds_block[];
  const typename S::Data *sp = sd;
  if constexpr (Staged) {
    stage_block(lds_block, reinterpret_cast<const double *>(sd), nd);
    __syncthreads();
    sp = reinterpret_cast<const typename S::Data *>(lds_block);
  }
  const size_t i = size_t(blockIdx.x) * kSoaBlock + threadIdx.x;
  if (i >= a.n) return;
  const typename S::Data &s = *sp;
  V3<double> p{a.px[i], a.py[i], a.pz[i]};
  if (use_tr) p = to_local_point(tr, p);
  if constexpr (M == kContains) {
    static_cast<uint8_t *>(a.out)[i] = S::template contains<double>(s, p) ? 1 : 0;
  } else if constexpr (M == kInside) {
    static_cast<int32_t *>(a.out)[i] = S::template inside<double>(s, p);
  } else if constexpr (M == kSafetyToIn) {
    static_cast<double *>(a.out)[i] = S::template safety_to_in<double>(s, p);
  } else if constexpr (M == kSafetyToOut) {
    static_cast<double *>(a.out)[i] = S::template safety_to_out<double>(s, p);
  } else {
    V3<double> d{a.dx[i], a.dy[i], a.dz[i]};
    if (use_tr) d = to_local_dir(tr, d);
    const double step = a.step ? a.step[i] : kInfLength;
    double r;
    if constexpr (M == kDistanceToIn)
      r = S::template distance_to_in<double>(s, p, d, step);
    else
      r = S::template distance_to_out<double>(s, p, d, step);
    static_cast<double *>(a.out)[i] = r;
  }
}

// Mixed precision (VGM_MIXED): float does the bulk of the work and FP64
// decides every result, so each result equals the FP64 kernel's.
//   Contains, Inside  the float classification stands when the float safety
//                     puts the point farther than the guard band from the
//                     surface; otherwise (and for float "surface") the point
//                     is classified in FP64 (kFloatClassify families)
//   DistanceToIn/Out  float selects the surface, FP64 evaluates it with the
//                     FP64 kernel's expressions; uncertain selections take the
//                     FP64 kernel (kMixedDistance families: tube, cone); FP64
//                     for the others
//   SafetyToIn/Out    FP64 (a single closed form or a walk whose value is the
//                     result: a float value would carry float error)
// The guard band is a few float ulps of the largest coordinate in play.
template <class S, int M>
__global__ __launch_bounds__(kSoaBlock) void soa_mixed_kernel(const typename S::Data *__restrict__ sd, Transform tr,
                                                              int use_tr, float scale, SoaArgs a)
{
  const size_t i = size_t(blockIdx.x) * kSoaBlock + threadIdx.x;
  if (i >= a.n) return;
  const typename S::Data &s = *sd;
  V3<double> p{a.px[i], a.py[i], a.pz[i]};
  if (use_tr) p = to_local_point(tr, p);
  if constexpr (M == kContains || M == kInside) {
    if constexpr (S::kFloatClassify) {
      const V3<float> pf{float(p.x), float(p.y), float(p.z)};
      const float g = 4.0e-6f * (fabsf(pf.x) + fabsf(pf.y) + fabsf(pf.z) + scale);
      const int r = S::template inside<float>(s, pf);
      float sf = 0.0f;
      if (r == kInsideV) sf = S::template safety_to_out<float>(s, pf);
      if (r == kOutsideV) sf = S::template safety_to_in<float>(s, pf);
      const bool firm = r != kSurfaceV && sf > g;
      if constexpr (M == kContains)
        static_cast<uint8_t *>(a.out)[i] = (firm ? r == kInsideV : S::template contains<double>(s, p)) ? 1 : 0;
      else
        static_cast<int32_t *>(a.out)[i] = firm ? r : S::template inside<double>(s, p);
    } else {
      if constexpr (M == kContains) static_cast<uint8_t *>(a.out)[i] = S::template contains<double>(s, p) ? 1 : 0;
      else static_cast<int32_t *>(a.out)[i] = S::template inside<double>(s, p);
    }
  } else if constexpr (M == kSafetyToIn) {
    static_cast<double *>(a.out)[i] = S::template safety_to_in<double>(s, p);
  } else if constexpr (M == kSafetyToOut) {
    static_cast<double *>(a.out)[i] = S::template safety_to_out<double>(s, p);
  } else {
    V3<double> d{a.dx[i], a.dy[i], a.dz[i]};
    if (use_tr) d = to_local_dir(tr, d);
    double r;
    if constexpr (S::kMixedDistance) {
      r = (M == kDistanceToIn) ? S::distance_to_in_mixed(s, p, d, scale) : S::distance_to_out_mixed(s, p, d, scale);
    } else {
      const double step = a.step ? a.step[i] : kInfLength;
      r = (M == kDistanceToIn) ? S::template distance_to_in<double>(s, p, d, step)
                               : S::template distance_to_out<double>(s, p, d, step);
    }
    static_cast<double *>(a.out)[i] = r;
  }
}

// VPlacedVolume::Normal (source/PlacedVolume.cpp:126-137) over a batch:
// the master point into the local frame, the family's unplaced Normal, the
// normal back to the master frame.  Outputs: the three components and
// whether the point is on the surface.
template <class S, bool Staged>
__global__ __launch_bounds__(kSoaBlock) void soa_normal_kernel(const typename S::Data *__restrict__ sd, int nd,
                                                               Transform tr, NormalArgs a)
{
  extern __shared__ double lds_block[];
  const typename S::Data *sp = sd;
  if constexpr (Staged) {
    stage_block(lds_block, reinterpret_cast<const double *>(sd), nd);
    __syncthreads();
    sp = reinterpret_cast<const typename S::Data *>(lds_block);
  }
  const size_t i = size_t(blockIdx.x) * kSoaBlock + threadIdx.x;
  if (i >= a.n) return;
  V3<double> ln{0.0, 0.0, 0.0};
  const bool valid = S::normal(*sp, to_local_point(tr, V3<double>{a.px[i], a.py[i], a.pz[i]}), ln);
  const V3<double> m = to_master_dir(tr, ln);
  a.nx[i] = m.x;
  a.ny[i] = m.y;
  a.nz[i] = m.z;
  a.valid[i] = valid ? 1 : 0;
}

namespace {
template <class S, class = void>
struct HasNormal : std::false_type {};
template <class S>
struct HasNormal<S, std::void_t<decltype(&S::normal)>> : std::true_type {};

struct NormalLaunch {
  hipStream_t stream;
  int nd;
  const void *params;
  const Transform &tr;
  const NormalArgs &a;
  hipError_t err;
  template <class S>
  void operator()()
  {
    if constexpr (HasNormal<S>::value) {
      const unsigned blocks = unsigned((a.n + kSoaBlock - 1) / kSoaBlock);
      const auto *p = static_cast<const typename S::Data *>(params);
      if (S::kStage && nd > 0 && size_t(nd) * sizeof(double) <= size_t(kMaxStageBytes))
        hipLaunchKernelGGL((soa_normal_kernel<S, S::kStage>), dim3(blocks), dim3(kSoaBlock),
                           size_t(nd) * sizeof(double), stream, p, nd, tr, a);
      else
        hipLaunchKernelGGL((soa_normal_kernel<S, false>), dim3(blocks), dim3(kSoaBlock), 0, stream, p, nd, tr, a);
      err = hipGetLastError();
    } else {
      err = hipErrorNotSupported;
    }
  }
};
struct NormalQuery {
  bool r = false;
  template <class S>
  void operator()()
  {
    r = HasNormal<S>::value;
  }
};
}  // namespace

bool family_has_normal(int type)
{
  NormalQuery q;
  return dispatch_shape_type(type, q) && q.r;
}

hipError_t launch_normal(hipStream_t stream, const ShapeRec &rec, const void *params, const Transform &tr,
                         const NormalArgs &a)
{
  if (a.n == 0) return hipSuccess;
  NormalLaunch L{stream, rec.nd, params, tr, a, hipSuccess};
  if (!dispatch_shape_type(rec.type, L)) return hipErrorInvalidValue;
  return L.err;
}

namespace {
struct SoaLaunch {
  hipStream_t stream;
  int method;
  int precision;
  float scale;
  int nd;
  const void *params;
  const Transform &tr;
  int use_tr;
  const SoaArgs &a;
  hipError_t err;

  template <class S, int M>
  void go()
  {
    const unsigned blocks = unsigned((a.n + kSoaBlock - 1) / kSoaBlock);
    const auto *p = static_cast<const typename S::Data *>(params);
    if (precision == kMixed) {
      hipLaunchKernelGGL((soa_mixed_kernel<S, M>), dim3(blocks), dim3(kSoaBlock), 0, stream, p, tr, use_tr, scale, a);
      err = hipGetLastError();
      return;
    }
    if (S::kStage && nd > 0 && size_t(nd) * sizeof(double) <= size_t(kMaxStageBytes))
      hipLaunchKernelGGL((soa_fp64_kernel<S, M, S::kStage>), dim3(blocks), dim3(kSoaBlock),
                         size_t(nd) * sizeof(double), stream, p, nd, tr, use_tr, a);
    else
      hipLaunchKernelGGL((soa_fp64_kernel<S, M, false>), dim3(blocks), dim3(kSoaBlock), 0, stream, p, nd, tr,
                         use_tr, a);
    err = hipGetLastError();
  }

  template <class S>
  void operator()()
  {
    switch (method) {
    case kContains: go<S, kContains>(); break;
    case kInside: go<S, kInside>(); break;
    case kDistanceToIn: go<S, kDistanceToIn>(); break;
    case kDistanceToOut: go<S, kDistanceToOut>(); break;
    case kSafetyToIn: go<S, kSafetyToIn>(); break;
    case kSafetyToOut: go<S, kSafetyToOut>(); break;
    default: err = hipErrorInvalidValue;
    }
  }
};
}  // namespace

hipError_t launch_soa(hipStream_t stream, int method, int precision, const ShapeRec &rec, const void *params,
                      const Transform &tr, bool use_transform, const SoaArgs &a)
{
  if (a.n == 0) return hipSuccess;
  double d2 = 0;
  for (int k = 0; k < 3; ++k) d2 += (rec.bmax[k] - rec.bmin[k]) * (rec.bmax[k] - rec.bmin[k]);
  SoaLaunch L{stream, method, precision, float(__builtin_sqrt(d2)), rec.nd, params, tr, use_transform ? 1 : 0, a,
              hipSuccess};
  if (!dispatch_shape_type(rec.type, L)) return hipErrorInvalidValue;
  return L.err;
}

}  // namespace vgm
