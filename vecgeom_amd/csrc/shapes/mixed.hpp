// Mixed precision: float selects, FP64 decides the value.
//
// A boundary distance of a tube or cone is the smallest of a few surface
// candidates (z planes, the outer and inner curved surfaces, two phi
// planes), each accepted or rejected by tests on its hit point.  The mixed
// kernels evaluate every candidate in float and certify each outcome with an
// error bound: a test counts only if it passes or fails by more than the
// float error of its operands (`Cert`), and two accepted candidates count as
// ordered only if their float distances differ by more than both errors.
// When every outcome is certain, the FP64 kernel would take the same branches
// and return the same candidate, so that candidate alone is evaluated in FP64
// by the FP64 kernel's own expressions: the result is the FP64 result, bit
// for bit.  Any uncertain outcome (grazing rays, hits near an edge, points
// within the float error of a surface, near ties) sends the query to the full
// FP64 kernel.  The error bounds are generous multiples of the float unit
// roundoff (kMixEps), so a float rounding cannot certify a wrong outcome.
#pragma once
#include "../vgm_math.hpp"

namespace vgm {

// float unit roundoff x 64: bound on the relative error of a short float
// expression (a few operations on operands rounded from FP64)
constexpr float kMixEps = 64.0f * 5.9604645e-8f;
// absolute length margin per unit of |p|_1 + solid diagonal: covers the
// rounding of the coordinates and of the solid's parameters to float
constexpr float kMixLen = 2.0e-5f;
// a root whose relative error bound exceeds this is not trusted
constexpr float kMixMaxRel = 1.0e-3f;

// Float square root and quotient of the selection: the hardware's 1-ulp
// instructions (v_sqrt_f32, v_rcp_f32) rather than the correctly rounded
// sequences; kMixEps covers their error.  Device code only.
#define VGM_MIX __device__ __forceinline__
VGM_MIX float mix_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
VGM_MIX float mix_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

enum Cert : int { kCertNo = 0, kCertYes = 1, kCertUnsure = 2 };

// a >= b (or a > b: the same at a margin) with both known to within m
VGM_DEV int cert_ge(float a, float b, float m) { return a > b + m ? kCertYes : (a < b - m ? kCertNo : kCertUnsure); }
VGM_DEV int cert_le(float a, float b, float m) { return cert_ge(b, a, m); }
VGM_DEV int cert_and(int a, int b)
{
  if (a == kCertNo || b == kCertNo) return kCertNo;
  return (a == kCertYes && b == kCertYes) ? kCertYes : kCertUnsure;
}
VGM_DEV int cert_or(int a, int b)
{
  if (a == kCertYes || b == kCertYes) return kCertYes;
  return (a == kCertNo && b == kCertNo) ? kCertNo : kCertUnsure;
}

// One float candidate: certified acceptance, distance and its error bound.
struct FCand {
  int acc;
  float d, err;
};

VGM_DEV FCand cand_no() { return FCand{kCertNo, FLT_MAX, 0.0f}; }
VGM_DEV FCand cand_unsure() { return FCand{kCertUnsure, FLT_MAX, 0.0f}; }

// Smallest accepted candidate, in the FP64 kernel's order; `unsure` once any
// candidate's acceptance or its order against the current best is uncertain.
struct FSelect {
  int best = -1;
  float d = FLT_MAX, err = 0.0f;
  bool unsure = false;

  VGM_DEV void add(int id, const FCand &c)
  {
    if (c.acc == kCertUnsure) unsure = true;
    if (c.acc != kCertYes) return;
    if (best >= 0) {
      if (vabs(c.d - d) <= c.err + err) {
        unsure = true;
        return;
      }
      if (!(c.d < d)) return;
    }
    best = id;
    d = c.d;
    err = c.err;
  }
};

// Exit through the z plane at height zt (dir.z != 0 in FP64): always taken,
// at (zt - pz) / dz; a direction with |dz| below float range is far away.
VGM_MIX FCand z_plane_cand(float zt, float pz, float dz, float m)
{
  if (vabs(dz) < 1.0e-20f) return FCand{kCertYes, 1.0e30f, 0.0f};
  const float r = __builtin_amdgcn_rcpf(dz);
  const float d = (zt - pz) * r;
  return FCand{kCertYes, d, (kMixEps * (vabs(zt) + vabs(pz)) + m) * vabs(r) + kMixEps * vabs(d)};
}

}  // namespace vgm
