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

// Mixed precision (VGM_MIXED): every query is evaluated in float; FP64 is
// kept for the decisions float cannot be trusted with.
//   Contains, Inside  the float classification stands when the float safety
//                     puts the point farther than the guard band from the
//                     surface; otherwise (and for float "surface") the point
//                     is classified in FP64 - so classifications equal FP64's
//   DistanceToIn/Out  float; FP64 when the float safety or the float distance
//                     is below the guard band (the point may be on the surface
//                     or the wrong side of it), or when the float hit point is
//                     not on the surface (float safety there above 16 guards);
//                     FP64 throughout for families whose float DistanceToOut
//                     can end on another surface (kFloatDistanceToOut false)
//                     or whose float DistanceToIn can miss where FP64 hits
//                     (kFloatDistanceToIn false)
//   SafetyToIn/Out    float; FP64 when the float safety is below the guard band
// The guard band is a few float ulps of the largest coordinate in play.  The
// error of the float results against FP64 is measured by the tests and the
// benchmark, per query.
template <class S, int M>
__global__ __launch_bounds__(kSoaBlock) void soa_mixed_kernel(const typename S::Data *__restrict__ sd, Transform tr,
                                                              int use_tr, float scale, SoaArgs a)
{
  const size_t i = size_t(blockIdx.x) * kSoaBlock + threadIdx.x;
  if (i >= a.n) return;
  const typename S::Data &s = *sd;
  V3<double> p{a.px[i], a.py[i], a.pz[i]};
  if (use_tr) p = to_local_point(tr, p);
  if constexpr (!S::kFloatQueries) {  // every query in FP64
    V3<double> d{0.0, 0.0, 0.0};
    if constexpr (M == kDistanceToIn || M == kDistanceToOut) {
      d = V3<double>{a.dx[i], a.dy[i], a.dz[i]};
      if (use_tr) d = to_local_dir(tr, d);
    }
    const double step = (M == kDistanceToIn || M == kDistanceToOut) && a.step ? a.step[i] : kInfLength;
    if constexpr (M == kContains) static_cast<uint8_t *>(a.out)[i] = S::template contains<double>(s, p) ? 1 : 0;
    else if constexpr (M == kInside) static_cast<int32_t *>(a.out)[i] = S::template inside<double>(s, p);
    else if constexpr (M == kSafetyToIn) static_cast<double *>(a.out)[i] = S::template safety_to_in<double>(s, p);
    else if constexpr (M == kSafetyToOut) static_cast<double *>(a.out)[i] = S::template safety_to_out<double>(s, p);
    else if constexpr (M == kDistanceToIn) static_cast<double *>(a.out)[i] = S::template distance_to_in<double>(s, p, d, step);
    else static_cast<double *>(a.out)[i] = S::template distance_to_out<double>(s, p, d, step);
    return;
  }
  const V3<float> pf{float(p.x), float(p.y), float(p.z)};
  const float g = 4.0e-6f * (fabsf(pf.x) + fabsf(pf.y) + fabsf(pf.z) + scale);
  if constexpr (M == kContains || M == kInside) {
    const int r = S::template inside<float>(s, pf);
    float sf = 0.0f;
    if (r == kInsideV) sf = S::template safety_to_out<float>(s, pf);
    if (r == kOutsideV) sf = S::template safety_to_in<float>(s, pf);
    const bool firm = r != kSurfaceV && sf > g;
    if constexpr (M == kContains)
      static_cast<uint8_t *>(a.out)[i] = (firm ? r == kInsideV : S::template contains<double>(s, p)) ? 1 : 0;
    else
      static_cast<int32_t *>(a.out)[i] = firm ? r : S::template inside<double>(s, p);
  } else if constexpr (M == kSafetyToIn || M == kSafetyToOut) {
    float sf = (M == kSafetyToIn) ? S::template safety_to_in<float>(s, pf) : S::template safety_to_out<float>(s, pf);
    double r = double(sf);
    if (!(sf > g))
      r = (M == kSafetyToIn) ? S::template safety_to_in<double>(s, p) : S::template safety_to_out<double>(s, p);
    static_cast<double *>(a.out)[i] = r;
  } else {
    V3<double> d{a.dx[i], a.dy[i], a.dz[i]};
    if (use_tr) d = to_local_dir(tr, d);
    if constexpr ((M == kDistanceToOut && !S::kFloatDistanceToOut) || (M == kDistanceToIn && !S::kFloatDistanceToIn)) {
      const double step = a.step ? a.step[i] : kInfLength;
      static_cast<double *>(a.out)[i] = (M == kDistanceToOut) ? S::template distance_to_out<double>(s, p, d, step)
                                                              : S::template distance_to_in<double>(s, p, d, step);
      return;
    }
    const V3<float> df{float(d.x), float(d.y), float(d.z)};
    const double step = a.step ? a.step[i] : kInfLength;
    const float stepf = step >= double(FLT_MAX) ? FLT_MAX : float(step);
    // near the surface the float point may sit on the other side of it:
    // FP64 there, and where the float distance itself is below the guard
    const float sf = (M == kDistanceToIn) ? S::template safety_to_in<float>(s, pf)
                                          : S::template safety_to_out<float>(s, pf);
    double r;
    float rf = 0.0f;
    if (sf > g)
      rf = (M == kDistanceToIn) ? S::template distance_to_in<float>(s, pf, df, stepf)
                                : S::template distance_to_out<float>(s, pf, df, stepf);
    // a float hit must land on the surface: the float safety there, on the
    // side the query ends on, must vanish within the guard band (a walk that
    // stopped at an internal section junction lands inside the solid)
    bool ok = sf > g && rf > g;
    if (ok && rf < FLT_MAX) {
      const V3<float> q{pf.x + rf * df.x, pf.y + rf * df.y, pf.z + rf * df.z};
      const float gq = 4.0e-6f * (fabsf(q.x) + fabsf(q.y) + fabsf(q.z) + scale);
      const float sq = (M == kDistanceToIn) ? S::template safety_to_in<float>(s, q) : S::template safety_to_out<float>(s, q);
      ok = fabsf(sq) <= 16.0f * gq;
    }
    if (ok)
      r = rf >= FLT_MAX ? kInfLength : double(rf);
    else
      r = (M == kDistanceToIn) ? S::template distance_to_in<double>(s, p, d, step)
                               : S::template distance_to_out<double>(s, p, d, step);
    static_cast<double *>(a.out)[i] = r;
  }
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
