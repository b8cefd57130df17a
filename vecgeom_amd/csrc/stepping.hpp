// Track-step loop shared by the batched transport kernels.
//
// One track-step (the unit of the throughput metric; the reference has no
// such loop, this is the project's definition):
//   s  = SafetyToOut(p) if inside else SafetyToIn(p)
//   l  = physics step ~ Exp(mean = lambda_solid), counter-based RNG
//   if l <= s          : p += l * d, then a new isotropic direction
//   else d_b = DistanceToOut/In(p, d), clamped at 0
//        if l < d_b    : p += l * d, then a new isotropic direction
//        else          : p += (d_b + 1e-6) * d, flip inside/outside
//   outside the world box (1.3 x the solid's bounding box) a coordinate is
//   wrapped to the opposite face (periodic world), keeping the population
//   around the solid.
// Everything here is deterministic and uses only + - * / so the device and
// the CPU oracle draw identical steps.
#pragma once
#include "vgm_math.hpp"

namespace vgm {

constexpr double kBoundaryPush = 1e-6;  // push past a crossed boundary (MovePointAfterBoundary)

VGM_DEV uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform in (0, 1) for (seed, track, step).
VGM_DEV double step_uniform(uint64_t seed, uint64_t track, uint32_t step)
{
  const uint64_t h = mix64(seed ^ mix64(track * 0x9E3779B97F4A7C15ull + step));
  return (double(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// -ln(u) for u in (0, 1): exponent split from the bits, then
// ln(m) = 2 atanh((m-1)/(m+1)) by a fixed odd series.  Not correctly rounded,
// but deterministic: the same operations on every platform.
VGM_DEV double neg_log(double u)
{
  const uint64_t bits = __builtin_bit_cast(uint64_t, u);
  const int e = int((bits >> 52) & 0x7ff) - 1022;
  const double m = __builtin_bit_cast(double, (bits & 0x800FFFFFFFFFFFFFull) | (1022ull << 52));  // [0.5, 1)
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double poly = 1.0 / 15.0;
  poly = poly * s2 + 1.0 / 13.0;
  poly = poly * s2 + 1.0 / 11.0;
  poly = poly * s2 + 1.0 / 9.0;
  poly = poly * s2 + 1.0 / 7.0;
  poly = poly * s2 + 1.0 / 5.0;
  poly = poly * s2 + 1.0 / 3.0;
  poly = poly * s2 + 1.0;
  const double lnm = 2.0 * s * poly;
  return -(double(e) * 0.6931471805599453 + lnm);
}

// Random numbers of step k: 4k (length), 4k+1..4k+3 (direction).
VGM_DEV double physics_step(double lambda, uint64_t seed, uint64_t track, uint32_t step)
{
  return lambda * neg_log(step_uniform(seed, track, 4u * step));
}

// Isotropic direction as volumeUtilities::SampleDirection: normalise
// (1-2u1, 1-2u2, 1-2u3).
VGM_DEV V3<double> new_direction(uint64_t seed, uint64_t track, uint32_t step, V3<double> old)
{
  const double vx = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 1u);
  const double vy = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 2u);
  const double vz = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 3u);
  const double m2 = vx * vx + vy * vy + vz * vz;
  if (!(m2 > 0.0)) return old;
  const double inv = 1.0 / __builtin_sqrt(m2);
  return V3<double>{vx * inv, vy * inv, vz * inv};
}

// Periodic world box [c - h, c + h] per axis.
VGM_DEV double wrap(double p, double c, double h)
{
  if (p > c + h) return p - 2.0 * h;
  if (p < c - h) return p + 2.0 * h;
  return p;
}

// Exact reorderings of the track-step (same decisions, same arithmetic on
// the position):
//
// Carried bound.  For a family whose safety is the distance to its surface,
// a safety s at p bounds the safety at p' by s - |p' - p|; a move of l keeps
// s - l - carry_margin(p) as the bound at the new point, and a later step
// of length up to that bound moves without any query.
template <typename R>
VGM_DEV R carry_margin(const V3<R> &p)
{
  return R(1e-7) + R(1e-12) * (vabs(p.x) + vabs(p.y) + vabs(p.z));
}

// Distance first.  For a family whose safety never exceeds its boundary
// distance, a step moves freely when l < d_b and crosses when l exceeds d_b
// by more than this window; only in between is the safety needed.
constexpr double kDistanceFirstWindow = 1e-6;

// Bounding-box cull.  A track outside its solid whose move [p, p + l d]
// stays clear of the solid's bounding box (centre bc, half-widths bh) cannot
// reach the surface: every point DistanceToIn can return lies on the surface
// (within the tolerance of the surface tests) and the surface lies in the
// box, so DistanceToIn > l and the step moves the track by l without a
// crossing whichever order it takes.  Separating-axis test of the segment,
// lengthened by twice the distance-first window, against the box grown by
// the window and the carried-bound margin (which cover the surface tolerance
// and the rounding of the test): the three box axes, then the cross products
// of the segment with them.  Rounding only errs towards "no", after which
// the ordinary step runs.
template <typename R>
VGM_DEV bool move_clears_box(const double *bc, const double *bh, const V3<R> &p, const V3<R> &d, R l)
{
  const R hl = R(0.5) * (l + R(2.0 * kDistanceFirstWindow));
  const R m = R(kDistanceFirstWindow) + carry_margin(p);
  const R wx = hl * d.x, wy = hl * d.y, wz = hl * d.z;  // half segment
  const R cx = (p.x + wx) - R(bc[0]), cy = (p.y + wy) - R(bc[1]), cz = (p.z + wz) - R(bc[2]);
  const R hx = R(bh[0]) + m, hy = R(bh[1]) + m, hz = R(bh[2]) + m;
  const R ex = vabs(wx), ey = vabs(wy), ez = vabs(wz);
  if (vabs(cx) > hx + ex || vabs(cy) > hy + ey || vabs(cz) > hz + ez) return true;
  return vabs(cy * wz - cz * wy) > hy * ez + hz * ey || vabs(cz * wx - cx * wz) > hx * ez + hz * ex ||
         vabs(cx * wy - cy * wx) > hx * ey + hy * ex;
}

}  // namespace vgm
