// Oracle: batched track-step transport (TEST INFRASTRUCTURE ONLY).
//
// The reference has no track-step loop; this is the project's definition,
// restated independently of the HIP kernels so that their results can be
// checked bit for bit and timed on the host as the CPU baseline:
//   s = SafetyToOut(p) inside / SafetyToIn(p) outside
//   l = mean_step * (-ln u), u from a SplitMix64-style counter hash of
//       (seed, track, step), -ln evaluated by exponent split + atanh series
//   l <= s          : p += l d, new isotropic direction
//   else b = max(DistanceToOut/In(p, d), 0)
//        l < b      : p += l d, new isotropic direction
//        otherwise  : p += (b + 1e-6) d, flip the side (crossing)
//   outside the solid each coordinate wraps periodically within 1.3x the
//   solid's bounding box.  Random numbers 4k (length) and 4k+1..4k+3
//   (direction: normalised (1-2u1, 1-2u2, 1-2u3)) belong to step k.
#include <omp.h>
#include <cstring>
#include "geom.h"
#include "oracle.h"

using namespace orc;

namespace {

uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double step_uniform(uint64_t seed, uint64_t track, uint32_t step)
{
  const uint64_t h = mix64(seed ^ mix64(track * 0x9E3779B97F4A7C15ull + step));
  return (double(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

double neg_log(double u)
{
  uint64_t bits;
  std::memcpy(&bits, &u, 8);
  const int e = int((bits >> 52) & 0x7ff) - 1022;
  const uint64_t mb = (bits & 0x800FFFFFFFFFFFFFull) | (1022ull << 52);
  double m;
  std::memcpy(&m, &mb, 8);
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

Vec3 new_direction(uint64_t seed, uint64_t track, uint32_t step, const Vec3 &old)
{
  const double vx = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 1u);
  const double vy = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 2u);
  const double vz = 1.0 - 2.0 * step_uniform(seed, track, 4u * step + 3u);
  const double m2 = vx * vx + vy * vy + vz * vz;
  if (!(m2 > 0.0)) return old;
  const double inv = 1.0 / std::sqrt(m2);
  return Vec3(vx * inv, vy * inv, vz * inv);
}

double wrap(double p, double c, double h)
{
  if (p > c + h) return p - 2.0 * h;
  if (p < c - h) return p + 2.0 * h;
  return p;
}

void advance(Vec3 &p, const Vec3 &d, double t)
{
  p.x = p.x + t * d.x;
  p.y = p.y + t * d.y;
  p.z = p.z + t * d.z;
}

}  // namespace

extern "C" {

double orc_physics_step(double mean, uint64_t seed, uint64_t track, uint32_t step)
{
  return mean * neg_log(step_uniform(seed, track, 4u * step));
}

int orc_batch_transport(orc_geom *g, int nsolids, const int *shapes, const int64_t *offsets, const double *mean_step,
                        double *x, double *y, double *z, const double *dx, const double *dy, const double *dz,
                        int8_t *side, int32_t *crossings, int nsteps, int step0, uint64_t seed, int classify,
                        int nthreads)
{
  auto *G = reinterpret_cast<Geometry *>(g);
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  for (int s = 0; s < nsolids; ++s) {
    const Shape &S = *G->shapes[shapes[s]];
    const double lambda = mean_step[s];
    Vec3 mn, mx;
    S.Extent(mn, mx);
    double wc[3], wh[3];
    for (int k = 0; k < 3; ++k) {
      wc[k] = 0.5 * (mn[k] + mx[k]);
      wh[k] = 1.3 * (0.5 * (mx[k] - mn[k]));
    }
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads)
    for (long long i = offsets[s]; i < offsets[s + 1]; ++i) {
      Vec3 p(x[i], y[i], z[i]);
      if (classify) side[i] = S.Contains(p) ? 1 : 0;
      Vec3 d(dx[i], dy[i], dz[i]);
      bool in = side[i] != 0;
      int32_t cross = 0;
      for (int k = 0; k < nsteps; ++k) {
        const uint32_t step = uint32_t(step0 + k);
        const double l = lambda * neg_log(step_uniform(seed, uint64_t(i), 4u * step));
        bool crossed = false;
        const double saf = in ? S.SafetyToOut(p) : S.SafetyToIn(p);
        if (l <= saf) {
          advance(p, d, l);
        } else {
          double b = in ? S.DistanceToOut(p, d, kInfLength) : S.DistanceToIn(p, d, kInfLength);
          if (b < 0.0) b = 0.0;
          if (l < b) {
            advance(p, d, l);
          } else {
            advance(p, d, b + kPushTolerance);
            in = !in;
            cross++;
            crossed = true;
          }
        }
        if (!crossed) d = new_direction(seed, uint64_t(i), step, d);
        if (!in) {
          p.x = wrap(p.x, wc[0], wh[0]);
          p.y = wrap(p.y, wc[1], wh[1]);
          p.z = wrap(p.z, wc[2], wh[2]);
        }
      }
      x[i] = p.x;
      y[i] = p.y;
      z[i] = p.z;
      const_cast<double *>(dx)[i] = d.x;
      const_cast<double *>(dy)[i] = d.y;
      const_cast<double *>(dz)[i] = d.z;
      side[i] = in ? 1 : 0;
      crossings[i] += cross;
    }
  }
  return 0;
}

}  // extern "C"
