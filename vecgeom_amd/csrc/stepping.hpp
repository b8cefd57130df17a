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

}  // namespace vgm
