// Device-side scalar helpers for the shape kernels.  Every kernel that must
// bit-match the FP64 oracle is compiled with -ffp-contract=off, so the order of
// operations written here is the order executed.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>

namespace vgm {

// Tolerances and sentinels (VecGeom/base/Math.h:27-53, FP64 build).
constexpr double kTolerance = 1e-9;
constexpr double kHalfTolerance = 0.5 * 1e-9;
constexpr double kConeTolerance = 1e-7;
constexpr double kAngTolerance = 1e-9;
constexpr double kPushTolerance = 1e-6;
constexpr double kInfLength = DBL_MAX;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2. * kPi;
constexpr double kHalfPi = 0.5 * kPi;
constexpr double kTiny = 1e-30;

#define VGM_DEV __host__ __device__ __forceinline__

template <typename R>
struct V3 {
  R x, y, z;
};

template <typename R>
VGM_DEV R vmax(R a, R b) { return a > b ? a : b; }
template <typename R>
VGM_DEV R vmin(R a, R b) { return a < b ? a : b; }
template <typename R>
VGM_DEV R vmax3(R a, R b, R c) { return vmax(vmax(a, b), c); }
template <typename R>
VGM_DEV R vmin3(R a, R b, R c) { return vmin(vmin(a, b), c); }
VGM_DEV double vabs(double a) { return __builtin_fabs(a); }
VGM_DEV float vabs(float a) { return __builtin_fabsf(a); }
VGM_DEV double vsqrt(double a) { return __builtin_sqrt(a); }
VGM_DEV float vsqrt(float a) { return __builtin_sqrtf(a); }
// copysign(1, x): VecCore's Sign for scalar doubles.
VGM_DEV double vsign(double x) { return __builtin_copysign(1.0, x); }
VGM_DEV float vsign(float x) { return __builtin_copysignf(1.0f, x); }
VGM_DEV double vcopysign(double a, double b) { return __builtin_copysign(a, b); }
VGM_DEV float vcopysign(float a, float b) { return __builtin_copysignf(a, b); }
// NonZero(x) = x + copysign(1e-30, x)  (VecGeom/base/Global.h:170-177)
// The same reference, but opaque to the optimiser: loads through it cannot be
// hoisted above this point or merged with loads through the original.  A
// kernel calls it once per track-step on a parameter block in LDS, so the
// header fields are read from LDS inside the step instead of being held in
// registers across the whole track loop (where they spill to scratch).  The
// offset stays wave-uniform (a scalar register), so loads from global memory
// remain scalar loads.
template <class T>
VGM_DEV const T &opaque_ref(const T &r)
{
#if defined(__HIP_DEVICE_COMPILE__)
  int z = 0;
  asm volatile("" : "+s"(z));
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(&r) + z);
#else
  return r;
#endif
}

// Vector3D::Unit (VecGeom/base/Vector3D.h:284-290): v / sqrt(|v|^2 + DBL_MIN)
VGM_DEV V3<double> v3_unit(const V3<double> &v)
{
  const double m = __builtin_sqrt(v.x * v.x + v.y * v.y + v.z * v.z + DBL_MIN);
  return V3<double>{v.x / m, v.y / m, v.z / m};
}
// Vector3D::Normalize (Vector3D.h:198): v * (1 / |v|)
VGM_DEV V3<double> v3_normalized(const V3<double> &v)
{
  const double s = 1.0 / __builtin_sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  return V3<double>{v.x * s, v.y * s, v.z * s};
}
VGM_DEV double v3_mag2(const V3<double> &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

VGM_DEV double nonzero(double x) { return x + __builtin_copysign(1.0e-30, x); }
VGM_DEV float nonzero(float x) { return x + __builtin_copysignf(1.0e-30f, x); }
// NonZeroAbs(x) = |x| + 1e-30 (VecGeom/base/Global.h:162-167)
VGM_DEV double nonzero_abs(double x) { return __builtin_fabs(x) + 1.0e-30; }
VGM_DEV float nonzero_abs(float x) { return __builtin_fabsf(x) + 1.0e-30f; }

template <typename R>
VGM_DEV R inf_length();
template <>
VGM_DEV double inf_length<double>() { return DBL_MAX; }
template <>
VGM_DEV float inf_length<float>() { return FLT_MAX; }

}  // namespace vgm
