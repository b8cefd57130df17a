// Oracle: polyhedron (TEST INFRASTRUCTURE ONLY), scalar backend.
//   construction: PolyhedronStruct(phiStart, phiDelta, sideCount, zPlaneCount,
//     zPlanes, rMin, rMax) and Initialize (VecGeom/volumes/PolyhedronStruct.h:116-440)
//   quadrilaterals: VecGeom/volumes/Quadrilaterals.h, source/Quadrilaterals.cpp
//     (Set, FlipSign), VecGeom/volumes/Planes.h and source/Planes.cpp
//   kernels: VecGeom/volumes/kernel/PolyhedronImplementation.h:300-1200
//     (generic EInnerRadii / EPhiCutout treatment; the specialisations the
//     factory picks, source/UnplacedPolyhedron.cpp:75-101, compute the same)
//   NormalizeAngle: VecGeom/volumes/kernel/GenericKernels.h:143-146
//   Vector3D equality is within kTolerance per component (Vector3D.h:387-390).
#include <vector>
#include <memory>
#include "tube.h"

namespace orc {

namespace {

double NormalizeAngle(double a) { return a + kTwoPi * (int(a < 0) - int(a / kTwoPi)); }

Vec3 FixZeroes(Vec3 v)
{
  for (int i = 0; i < 3; ++i)
    if (Abs(v[i]) < kTolerance) v[i] = 0.0;
  return v;
}

Vec3 Normalized(const Vec3 &v) { return v * (1. / v.Mag()); }

bool TolEqual(const Vec3 &a, const Vec3 &b)
{
  return Abs(a[0] - b[0]) < kTolerance && Abs(a[1] - b[1]) < kTolerance && Abs(a[2] - b[2]) < kTolerance;
}

Vec3 FromCylindrical(double r, double phi, double z) { return Vec3(r * std::cos(phi), r * std::sin(phi), z); }

// DistanceToLineSegmentSquared1 (GenericKernels.h:190-215)
double SegmentDistSq(const Vec3 &c0, const Vec3 &c1, const Vec3 &p)
{
  const Vec3 line = c1 - c0;
  const Vec3 projection = p - c0;
  const double dot0 = Dot(projection, line);
  if (dot0 <= 0) return (p - c0).Mag2();
  const double dot1 = line.Mag2();
  if (dot1 <= dot0) return (p - c1).Mag2();
  const Vec3 q = c0 + line * (dot0 / dot1);
  return (q - p).Mag2();
}

struct Quads {
  int n = 0;
  std::vector<Vec3> normal;
  std::vector<double> dist;
  std::vector<Vec3> sideN[4];
  std::vector<double> sideD[4];
  std::vector<Vec3> corner[4];

  void Resize(int size)
  {
    n = size;
    normal.assign(size, Vec3());
    dist.assign(size, 0.);
    for (int j = 0; j < 4; ++j) {
      sideN[j].assign(size, Vec3());
      sideD[j].assign(size, 0.);
      corner[j].assign(size, Vec3());
    }
  }

  // Planes::Set(index, normal, x0) (source/Planes.cpp:30-38)
  static void SetSide(Vec3 &nOut, double &dOut, const Vec3 &nrm, const Vec3 &x0)
  {
    Vec3 fixed = FixZeroes(nrm);
    double inv = 1. / fixed.Mag();
    nOut = Vec3(inv * fixed.x, inv * fixed.y, inv * fixed.z);
    dOut = inv * -Dot(fixed, x0);
  }

  void Set(int index, const Vec3 &c0, const Vec3 &c1, const Vec3 &c2, const Vec3 &c3)
  {
    corner[0][index] = c0;
    corner[1][index] = c1;
    corner[2][index] = c2;
    corner[3][index] = c3;
    Vec3 ch[4];
    ch[0] = c0;
    int assigned = 1, tested = 1;
    const Vec3 cs[4] = {c0, c1, c2, c3};
    while (tested < 4) {
      bool chosen = true;
      for (int j = 0; j < assigned; ++j)
        if (TolEqual(ch[j], cs[tested])) {
          chosen = false;
          break;
        }
      if (chosen) ch[assigned++] = cs[tested];
      ++tested;
    }
    double a = ch[0][1] * (ch[1][2] - ch[2][2]) + ch[1][1] * (ch[2][2] - ch[0][2]) + ch[2][1] * (ch[0][2] - ch[1][2]);
    double b = ch[0][2] * (ch[1][0] - ch[2][0]) + ch[1][2] * (ch[2][0] - ch[0][0]) + ch[2][2] * (ch[0][0] - ch[1][0]);
    double c = ch[0][0] * (ch[1][1] - ch[2][1]) + ch[1][0] * (ch[2][1] - ch[0][1]) + ch[2][0] * (ch[0][1] - ch[1][1]);
    double d = -ch[0][0] * (ch[1][1] * ch[2][2] - ch[2][1] * ch[1][2]) -
               ch[1][0] * (ch[2][1] * ch[0][2] - ch[0][1] * ch[2][2]) -
               ch[2][0] * (ch[0][1] * ch[1][2] - ch[1][1] * ch[0][2]);
    Vec3 nrm(a, b, c);
    double inverseLength = 1. / nrm.Mag();
    nrm = nrm * inverseLength;
    d *= inverseLength;
    nrm = FixZeroes(nrm);
    normal[index] = nrm;
    dist[index] = d;
    auto side = [&](int j, const Vec3 &a0, const Vec3 &a1) {
      if (!TolEqual(a1, a0)) {
        Vec3 sv = Normalized(Cross(nrm, a1 - a0));
        SetSide(sideN[j][index], sideD[j][index], sv, a0);
      } else {
        sideN[j][index] = Vec3(0, 0, 0);
        sideD[j][index] = 0;
      }
    };
    side(0, c0, c1);
    side(1, c1, c2);
    side(2, c2, c3);
    side(3, c3, c0);
  }

  void FlipSign(int i)
  {
    normal[i] = -normal[i];
    dist[i] = -dist[i];
  }

  bool Contains(const Vec3 &p) const  // Planes::Contains
  {
    bool r = true;
    for (int i = 0; i < n; ++i) r &= Dot(p, normal[i]) + dist[i] < 0;
    return r;
  }

  int InsideOne(const Vec3 &p, int i) const  // Planes::Inside(point, i)
  {
    double dr = normal[i].x * p[0] + normal[i].y * p[1] + normal[i].z * p[2] + dist[i];
    int r = kInside;
    if (dr > kHalfTolerance) r = kOutside;
    if (r == kInside && dr > -kHalfTolerance) r = kSurface;
    return r;
  }

  int Inside(const Vec3 &p) const  // Planes::Inside(point)
  {
    int r = kInside;
    for (int i = 0; i < n; ++i) {
      double dr = normal[i].x * p[0] + normal[i].y * p[1] + normal[i].z * p[2] + dist[i];
      if (dr > kHalfTolerance) r = kOutside;
      if (r == kInside && dr > -kHalfTolerance) r = kSurface;
      if (r == kOutside) break;
    }
    return r;
  }

  bool RayHits(int i, const Vec3 &x) const
  {
    bool valid = true;
    for (int j = 0; j < 4; ++j) {
      valid &= Dot(x, sideN[j][i]) + sideD[j][i] >= -kTolerance;
      if (!valid) break;
    }
    return valid;
  }

  // Quadrilaterals::DistanceToIn<behindPlanesT> (Quadrilaterals.h:320-350)
  double DistanceToIn(const Vec3 &p, const Vec3 &d, bool behind) const
  {
    double best = kInfLength;
    for (int i = 0; i < n; ++i) {
      double distance = Dot(p, normal[i]) + dist[i];
      bool valid = (behind ? -distance : distance) > -kTolerance;
      if (!valid) continue;
      double proj = Dot(d, normal[i]);
      valid &= (behind ? proj : -proj) > 0;
      if (!valid) continue;
      distance /= -(proj + std::copysign(1e-20, proj));
      Vec3 x = p + d * distance;
      for (int j = 0; j < 4; ++j) {
        valid &= Dot(x, sideN[j][i]) + sideD[j][i] >= -kTolerance;
        if (!valid) break;
      }
      if (valid) best = distance;
      if (best < kInfLength) break;
    }
    return Max(0., best);
  }

  // Quadrilaterals::DistanceToOut(point, direction, zMin, zMax) (:420-455)
  double DistanceToOut(const Vec3 &p, const Vec3 &d, double zMin, double zMax) const
  {
    double best = kInfLength;
    for (int i = 0; i < n; ++i) {
      double dt = Dot(p, normal[i]) + dist[i];
      bool valid = dt < kTolerance;
      if (!valid) continue;
      double proj = Dot(d, normal[i]);
      valid &= proj > 0;
      if (!valid) continue;
      dt /= -proj;
      valid &= dt < best;
      if (!valid) continue;
      if (zMin == zMax) {
        Vec3 x = p + d * dt;
        valid = RayHits(i, x);
      } else {
        double zp = p[2] + dt * d[2];
        valid &= (zp >= zMin - kTolerance) && (zp < zMax + kTolerance);
      }
      if (!valid) continue;
      best = dt;
    }
    if (best > -kTolerance) best = Max(best, 0.);
    return best;
  }

  // Quadrilaterals::ScalarDistanceSquared (:460-490)
  double DistanceSquared(int i, const Vec3 &p) const
  {
    const Vec3 &pn = normal[i];
    double distance = Dot(p, pn) + dist[i];
    Vec3 x = p - pn * distance;
    bool within[4];
    for (int j = 0; j < 4; ++j)
      within[j] = x[0] * sideN[j][i].x + x[1] * sideN[j][i].y + x[2] * sideN[j][i].z + sideD[j][i] >= 0;
    if (within[0] && within[1] && within[2] && within[3]) return distance * distance;
    double dsq = kInfLength;
    for (int j = 0; j < 4; ++j) {
      if (!within[j]) {
        distance = SegmentDistSq(corner[j][i], corner[(j + 1) % 4][i], p);
        if (distance < dsq) dsq = distance;
      }
    }
    return dsq;
  }
};

struct Seg {
  Quads outer, phi, inner;
  bool hasInner() const { return inner.n > 0; }
};

}  // namespace

struct Polyhedron final : Shape {
  int fSideCount;
  bool fHasInnerRadii = false, fHasPhiCutout, fHasLargePhiCutout;
  double fPhiStart, fPhiDelta;
  Wedge fPhiWedge;
  std::vector<Seg> seg;
  std::vector<double> fZPlanes, fRMin, fRMax;
  std::vector<bool> fSameZ;
  std::vector<Vec3> fPhiSections;
  std::unique_ptr<Tube> fBoundingTube;
  double fBoundingTubeOffset = 0;

  Polyhedron(double phiStart, double phiDelta, int sideCount, int nz, const double *zPlanes, const double *rMin,
             const double *rMax)
      : fSideCount(sideCount), fHasPhiCutout(phiDelta < kTwoPi), fHasLargePhiCutout(phiDelta < kPi),
        fPhiStart(NormalizeAngle(phiStart)), fPhiDelta((phiDelta > kTwoPi) ? kTwoPi : phiDelta),
        fPhiWedge(fPhiDelta, fPhiStart)
  {
    fZPlanes.assign(zPlanes, zPlanes + nz);
    fRMin.assign(rMin, rMin + nz);
    fRMax.assign(rMax, rMax + nz);
    fSameZ.assign(nz, false);
    for (int i = 0; i < nz; i++)
      if (i > 0 && i < nz - 1 && fZPlanes[i] == fZPlanes[i + 1]) fSameZ[i] = true;
    seg.resize(nz - 1);
    for (int i = 0; i < nz - 1; ++i) {
      bool hasInnerRadius = rMin[i] > 0 || rMin[i + 1] > 0;
      int mult = (zPlanes[i] == zPlanes[i + 1] && rMax[i] == rMax[i + 1]) ? 0 : 1;
      seg[i].outer.Resize(sideCount * mult);
      if (fHasPhiCutout) {
        mult = (zPlanes[i] == zPlanes[i + 1]) ? 0 : 1;
        seg[i].phi.Resize(2 * mult);
      }
      mult = (zPlanes[i] == zPlanes[i + 1] && rMin[i] == rMin[i + 1]) ? 0 : 1;
      if (hasInnerRadius && mult > 0) {
        seg[i].inner.Resize(sideCount * mult);
        fHasInnerRadii = true;
      }
    }
    phiStart = NormalizeAngle(phiStart);
    if (phiDelta > kTwoPi) phiDelta = kTwoPi;
    const double sidePhi = phiDelta / sideCount;
    std::vector<double> vertixPhi(sideCount + 1);
    fPhiSections.resize(sideCount + 1);
    for (int i = 0; i < sideCount + 1; ++i) {
      vertixPhi[i] = NormalizeAngle(phiStart + i * sidePhi);
      Vec3 cornerVector = FixZeroes(Normalized(FromCylindrical(1., vertixPhi[i], 0)));
      fPhiSections[i] = Cross(Normalized(cornerVector), Vec3(0, 0, -1));
    }
    if (!fHasPhiCutout) vertixPhi[sideCount] = vertixPhi[0];
    const double cosHalfDeltaPhi = std::cos(0.5 * sidePhi);
    double innerRadius = kInfLength, outerRadius = -kInfLength;
    for (int i = 0; i < nz; ++i) {
      if (rMin[i] < innerRadius) innerRadius = rMin[i];
      if (rMax[i] > outerRadius) outerRadius = rMax[i];
    }
    outerRadius /= cosHalfDeltaPhi;
    const double boundingTubeZ = 0.5 * (zPlanes[nz - 1] - zPlanes[0]) + kTolerance;
    const double kPhiTolerance = 100 * kTolerance;
    const double boundsPhiStart = !fHasPhiCutout ? 0 : phiStart - kPhiTolerance;
    const double boundsPhiDelta = !fHasPhiCutout ? kTwoPi : phiDelta + 2 * kPhiTolerance;
    fBoundingTube.reset(new Tube(innerRadius - kHalfTolerance, outerRadius + kHalfTolerance, boundingTubeZ,
                                 boundsPhiStart, boundsPhiDelta));
    fBoundingTubeOffset = 0.5 * (zPlanes[0] + zPlanes[nz - 1]);
    auto VI = [&](int plane, int corner) { return plane * (sideCount + 1) + corner; };
    const int nV = nz * (sideCount + 1);
    std::vector<Vec3> ov(nV), iv(nV);
    for (int i = 0; i < nz; ++i) {
      for (int j = 0, jMax = sideCount + (fHasPhiCutout ? 1 : 0); j < jMax; ++j) {
        ov[VI(i, j)] = FixZeroes(FromCylindrical(rMax[i] / cosHalfDeltaPhi, vertixPhi[j], zPlanes[i]));
        iv[VI(i, j)] = FixZeroes(FromCylindrical(rMin[i] / cosHalfDeltaPhi, vertixPhi[j], zPlanes[i]));
      }
      if (!fHasPhiCutout) {
        ov[VI(i, sideCount)] = ov[VI(i, 0)];
        iv[VI(i, sideCount)] = iv[VI(i, 0)];
      }
    }
    auto wrong = [](const Vec3 &nrm, const Vec3 &c) { return nrm[0] * c[0] + nrm[1] * c[1] < 0; };
    for (int ip = 0; ip < nz - 1; ++ip) {
      for (int is = 0; is < seg[ip].outer.n; ++is) {
        seg[ip].outer.Set(is, ov[VI(ip, is)], ov[VI(ip, is + 1)], ov[VI(ip + 1, is + 1)], ov[VI(ip + 1, is)]);
        if (wrong(seg[ip].outer.normal[is], ov[VI(ip, is)])) seg[ip].outer.FlipSign(is);
      }
      for (int is = 0; is < seg[ip].inner.n; ++is) {
        seg[ip].inner.Set(is, iv[VI(ip, is)], iv[VI(ip, is + 1)], iv[VI(ip + 1, is + 1)], iv[VI(ip + 1, is)]);
        if (wrong(seg[ip].inner.normal[is], iv[VI(ip, is)])) seg[ip].inner.FlipSign(is);
      }
      if (fHasPhiCutout && seg[ip].phi.n == 2) {
        seg[ip].phi.Set(0, iv[VI(ip, 0)], iv[VI(ip + 1, 0)], ov[VI(ip + 1, 0)], ov[VI(ip, 0)]);
        if (Dot(seg[ip].phi.normal[0], fPhiSections[0]) > 0) seg[ip].phi.FlipSign(0);
        seg[ip].phi.Set(1, ov[VI(ip, sideCount)], ov[VI(ip + 1, sideCount)], iv[VI(ip + 1, sideCount)],
                        iv[VI(ip, sideCount)]);
        if (Dot(seg[ip].phi.normal[1], fPhiSections[fSideCount]) < 0) seg[ip].phi.FlipSign(1);
      }
    }
  }

  int NSeg() const { return int(seg.size()); }

  // FindZSegmentKernel<Precision> (PolyhedronImplementation.h:313-327)
  int FindZSegment(double pz) const
  {
    const double *begin = fZPlanes.data(), *end = fZPlanes.data() + fZPlanes.size();
    int index = -1;
    while (begin < end - 1 && pz - kTolerance > *begin) {
      ++index;
      ++begin;
    }
    if (pz + kTolerance > *begin) return (index + 1);
    return index;
  }

  int FindPhiSegment(const Vec3 &p) const
  {
    int index = -1;
    double first = p[0] * fPhiSections[0].x + p[1] * fPhiSections[0].y + p[2] * fPhiSections[0].z;
    for (int i = 1; i < fSideCount + 1; ++i) {
      double second = p[0] * fPhiSections[i].x + p[1] * fPhiSections[i].y + p[2] * fPhiSections[i].z;
      if (first > -kTolerance && second < kTolerance) index = i - 1;
      if (index >= 0) break;
      first = second;
    }
    return index;
  }

  bool BoundsContain(const Vec3 &p) const
  {
    return fBoundingTube->Contains(Vec3(p[0], p[1], p[2] - fBoundingTubeOffset));
  }

  double DistanceToInZSegment(int si, const Vec3 &p, const Vec3 &d) const
  {
    const Seg &s = seg[si];
    double distance = s.outer.DistanceToIn(p, d, false);
    bool done = distance < kInfLength;
    if (done) return distance;
    if (fHasPhiCutout) distance = s.phi.DistanceToIn(p, d, false);
    done |= distance < kInfLength;
    if (done) return distance;
    if (s.hasInner()) distance = s.inner.DistanceToIn(p, d, true);
    return distance;
  }

  double DistanceToOutZSegment(int si, double zMin, double zMax, const Vec3 &p, const Vec3 &d) const
  {
    const Seg &s = seg[si];
    bool done = false;
    double distance = kInfLength;
    if (s.hasInner()) {
      distance = s.inner.DistanceToIn(p, d, false);
      done = distance < kInfLength;
      if (done) return distance;
    }
    if (fHasPhiCutout) {
      double distphi = s.phi.DistanceToIn(p, d, true);
      if (!done && distance > -kTolerance) distance = distphi;
      done = distance > -kTolerance && distance < kInfLength;
      if (done) return distance;
    }
    double distout = s.outer.DistanceToOut(p, d, zMin, zMax);
    if (!done) distance = distout;
    return distance;
  }

  // A zero-height end segment has no phi quads: the reference would read
  // past the shell there (undefined); the restatement says "not in the wedge".
  bool InPhiCutoutWedge(const Seg &s, const Vec3 &p) const
  {
    if (s.phi.n < 2) return false;
    bool s0 = Dot(p, s.phi.normal[0]) + s.phi.dist[0] >= 0;
    bool s1 = Dot(p, s.phi.normal[1]) + s.phi.dist[1] >= 0;
    if (fHasLargePhiCutout) return s0 || s1;
    return s0 && s1;
  }

  double SafetyToZSegmentSquared(int si, int &phiIndex, const Vec3 &p, bool pt_inside, int &iSurf) const
  {
    const Seg &s = seg[si];
    bool in_cutout = phiIndex < 0;
    double safetySquared = kInfLength;
    if (fHasPhiCutout && s.phi.n == 2) {
      if (pt_inside || in_cutout) {
        iSurf = 0;
        safetySquared = s.phi.DistanceSquared(0, p);
        double saf = s.phi.DistanceSquared(1, p);
        if (saf < safetySquared) {
          safetySquared = saf;
          iSurf = 1;
        }
        if (in_cutout) {
          if (fHasLargePhiCutout && s.hasInner()) {
            if (s.inner.n > 0) {
              double sq = s.inner.DistanceSquared(0, p);
              if (sq < safetySquared) {
                iSurf = 2;
                phiIndex = 0;
                safetySquared = sq;
              }
              if (s.inner.n > 1) {
                sq = s.inner.DistanceSquared(s.inner.n - 1, p);
                if (sq < safetySquared) {
                  iSurf = 2;
                  phiIndex = s.inner.n - 1;
                  safetySquared = sq;
                }
              }
            }
          }
          return safetySquared;
        }
      }
    }
    if (in_cutout && si > 0 && si < NSeg() - 1 && fZPlanes[si] == fZPlanes[si + 1]) {
      if (s.outer.n > 0) {
        double so = s.outer.DistanceSquared(0, p);
        if (so < safetySquared) {
          iSurf = 3;
          phiIndex = 0;
          safetySquared = so;
        }
        if (s.outer.n > 1) {
          so = s.outer.DistanceSquared(s.outer.n - 1, p);
          if (so < safetySquared) {
            iSurf = 3;
            phiIndex = s.outer.n - 1;
            safetySquared = so;
          }
        }
      }
      if (s.hasInner()) {
        if (s.inner.n > 0) {
          double sq = s.inner.DistanceSquared(0, p);
          if (sq < safetySquared) {
            iSurf = 2;
            phiIndex = 0;
            safetySquared = sq;
          }
          if (s.inner.n > 1) {
            sq = s.inner.DistanceSquared(s.inner.n - 1, p);
            if (sq < safetySquared) {
              iSurf = 2;
              phiIndex = s.inner.n - 1;
              safetySquared = sq;
            }
          }
        }
      }
      return safetySquared;
    }
    // A point in the phi cutout (phiIndex < 0) reaches this point only for a
    // zero-height first or last segment, which has no phi quadrilaterals; the
    // reference then indexes the shells with -1
    // (PolyhedronImplementation.h:541-546), an out-of-bounds read.  The
    // restatement skips those lookups instead; such shapes are kept out of
    // the parity sets.
    if (phiIndex < 0) return safetySquared;
    double so = kInfLength;
    if (s.outer.n > 0) so = s.outer.DistanceSquared(phiIndex, p);
    double sq = kInfLength;
    if (s.hasInner() && s.inner.n > 0) sq = s.inner.DistanceSquared(phiIndex, p);
    if (sq < safetySquared) {
      iSurf = 2;
      safetySquared = sq;
    }
    if (so < safetySquared) {
      iSurf = 3;
      safetySquared = so;
    }
    return safetySquared;
  }

  void DistanceToEndcaps(bool pointInside, const Vec3 &p, const Vec3 &d, double &distance) const
  {
    const Seg *s;
    double zPlane;
    const int last = NSeg() - 1;
    if (pointInside ? d[2] < 0 : !(d[2] < 0)) {
      s = &seg[0];
      zPlane = fZPlanes[0];
    } else {
      s = &seg[last];
      zPlane = fZPlanes[NSeg()];
    }
    double dt = (zPlane - p[2]) / NonZero(d[2]);
    if (dt < -kTolerance || dt >= distance) return;
    Vec3 x = p + d * dt;
    if (!s->outer.Contains(x)) return;
    if (s->hasInner() && s->inner.Contains(x)) return;
    if (fHasPhiCutout && InPhiCutoutWedge(*s, x)) return;
    distance = dt;
  }

  void SafetyToEndcapsSquared(const Vec3 &p, double &distanceSquared) const
  {
    double firstDistance = fZPlanes[0] - p[2];
    double lastDistance = fZPlanes[NSeg()] - p[2];
    bool isFirst = Abs(firstDistance) < Abs(lastDistance);
    const Seg &s = isFirst ? seg[0] : seg[NSeg() - 1];
    double dt = isFirst ? firstDistance : lastDistance;
    double dt2 = dt * dt;
    if (dt2 >= distanceSquared) return;
    Vec3 x(p[0], p[1], p[2] + dt);
    if (!s.outer.Contains(x)) return;
    if (s.hasInner() && s.inner.Contains(x)) return;
    if (fHasPhiCutout && InPhiCutoutWedge(s, x)) return;
    distanceSquared = dt2;
  }

  int WedgeInside(const Vec3 &p) const
  {
    bool ci = false, co = false;
    fPhiWedge.GenericKernel(p, true, ci, co);
    int r = kSurface;
    if (co) r = kOutside;
    if (ci) r = kInside;
    return r;
  }

  int InsideSegBorder(const Vec3 &p, int zIndex) const
  {
    const Seg &s = seg[zIndex];
    int phiIndex = FindPhiSegment(p);
    if (phiIndex < 0) return kOutside;
    if (!s.outer.n && !s.inner.n) return kOutside;  // no shell to take a radius from
    const Vec3 &vout = s.outer.n ? s.outer.sideN[0][phiIndex] : s.inner.sideN[0][phiIndex];
    double rdv = Abs(Dot(p, vout));
    bool in1 = (rdv > fRMin[zIndex] - kTolerance) && (rdv < fRMax[zIndex] + kTolerance);
    bool in2 = (rdv > fRMin[zIndex + 1] - kTolerance) && (rdv < fRMax[zIndex + 1] + kTolerance);
    if (in1 && in2) {
      if ((rdv < fRMin[zIndex] + kTolerance) || (rdv > fRMax[zIndex] - kTolerance) ||
          (rdv < fRMin[zIndex + 1] + kTolerance) || (rdv > fRMax[zIndex + 1] - kTolerance))
        return kSurface;
      if (fHasPhiCutout) return WedgeInside(p);
      return kInside;
    }
    if (!in1 && !in2) return kOutside;
    return kSurface;
  }

  int InsideSegPhi(const Vec3 &p, int zIndex, int phiIndex) const
  {
    if (phiIndex < 0) return kOutside;
    double dz = Abs(p[2] - fBoundingTubeOffset) - 0.5 * (fZPlanes[NSeg()] - fZPlanes[0]);
    if (dz > kHalfTolerance) return kOutside;
    if (fSameZ[zIndex]) return InsideSegBorder(p, zIndex);
    const Seg &s = seg[zIndex];
    // (a zero-height end segment may have no outer shell: skipped)
    if (s.outer.n > 0) {
      int io = s.outer.InsideOne(p, phiIndex);
      if (io != kInside) return io;
    }
    if (s.hasInner()) {
      int ii = s.inner.InsideOne(p, phiIndex);
      if (ii == kInside) return kOutside;
      if (ii == kSurface) return kSurface;
    }
    if (fHasPhiCutout) {
      int ip = WedgeInside(p);
      if (ip != kInside) return ip;
    }
    if (Abs(dz) < kHalfTolerance) return kSurface;
    return kInside;
  }

  bool Contains(const Vec3 &p) const override
  {
    if (!BoundsContain(p)) return false;
    int zIndex = FindZSegment(p[2]);
    if (!((zIndex >= 0) && (zIndex < NSeg()))) return false;
    const Seg &s = seg[zIndex];
    if (fSameZ[zIndex]) {
      int phiIndex = FindPhiSegment(p);
      if (phiIndex < 0) return false;
      if (!s.outer.n && !s.inner.n) return false;  // no shell to take a radius from
      const Vec3 &vout = s.outer.n ? s.outer.sideN[0][phiIndex] : s.inner.sideN[0][phiIndex];
      double rdv = Abs(Dot(p, vout));
      bool in1 = (rdv >= fRMin[zIndex]) && (rdv <= fRMax[zIndex]);
      bool in2 = (rdv >= fRMin[zIndex + 1]) && (rdv <= fRMax[zIndex + 1]);
      return in1 | in2;
    }
    if (!s.outer.Contains(p)) return false;
    if (s.hasInner() && s.inner.Contains(p)) return false;
    if (fHasPhiCutout && !s.phi.Contains(p)) return false;
    return true;
  }

  int Inside(const Vec3 &p) const override
  {
    if (!BoundsContain(p)) return kOutside;
    int zIndex = FindZSegment(p[2]);
    if (zIndex > NSeg() - 1) zIndex = NSeg() - 1;
    if (zIndex < 0) zIndex = 0;
    const Seg &s = seg[zIndex];
    if (fSameZ[zIndex]) return InsideSegBorder(p, zIndex);
    int io = s.outer.Inside(p);
    if (io != kInside) return io;
    if (s.hasInner()) {
      int ii = s.inner.Inside(p);
      if (ii == kInside) return kOutside;
      if (ii == kSurface) return kSurface;
    }
    if (fHasPhiCutout) {
      int ip = s.phi.Inside(p);
      if (ip != kInside) return ip;
    }
    double dz = Abs(Abs(p[2] - fBoundingTubeOffset) - 0.5 * (fZPlanes[NSeg()] - fZPlanes[0]));
    if (dz < kTolerance) return kSurface;
    return kInside;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &d, double stepMax) const override
  {
    if ((p[2] < fZPlanes[0] + kTolerance) && d[2] <= 0) return kInfLength;
    if ((p[2] > fZPlanes[NSeg()] - kTolerance) && d[2] >= 0) return kInfLength;
    if (Inside(p) == kInside) return -1.;
    double tubeDistance = 0.;
    {
      Vec3 bp(p[0], p[1], p[2] - fBoundingTubeOffset);
      if (!fBoundingTube->Contains(bp)) {
        tubeDistance = fBoundingTube->DistanceToIn(bp, d, stepMax);
        if (tubeDistance == kInfLength) return kInfLength;
      }
    }
    int zIndex = FindZSegment(p[2]);
    const int zMax = NSeg();
    zIndex = zIndex < 0 ? 0 : (zIndex >= zMax ? zMax - 1 : zIndex);
    bool goingRight = d[2] >= 0;
    double distance = kInfLength;
    if (goingRight) {
      for (; zIndex < zMax; ++zIndex) {
        distance = DistanceToInZSegment(zIndex, p, d);
        if (distance >= 0 && distance < kInfLength) break;
      }
    } else {
      for (; zIndex >= 0; --zIndex) {
        distance = DistanceToInZSegment(zIndex, p, d);
        if (distance >= 0 && distance < kInfLength) break;
      }
    }
    DistanceToEndcaps(false, p, d, distance);
    return (distance >= tubeDistance - 1E-6) ? distance : kInfLength;
  }

  double Safety(const Vec3 &p, bool pt_inside) const
  {
    double safety = kInfLength;
    int iSurf = 0;
    const int zMax = NSeg();
    int zIndex = FindZSegment(p[2]);
    zIndex = zIndex < 0 ? 0 : (zIndex >= zMax ? zMax - 1 : zIndex);
    int phiIndex = FindPhiSegment(p);
    int inside = InsideSegPhi(p, zIndex, phiIndex);
    if (inside == kSurface) return 0.;
    bool contains = (inside == kInside);
    if (contains ^ pt_inside) return -1.;
    for (int z = zIndex; z < zMax;) {
      safety = Min(safety, SafetyToZSegmentSquared(z, phiIndex, p, pt_inside, iSurf));
      ++z;
      double dz = fZPlanes[z] - p[2];
      if (dz * dz > safety) break;
    }
    for (int z = zIndex - 1; z >= 0; --z) {
      safety = Min(safety, SafetyToZSegmentSquared(z, phiIndex, p, pt_inside, iSurf));
      double dz = p[2] - fZPlanes[z];
      if (dz * dz > safety) break;
    }
    SafetyToEndcapsSquared(p, safety);
    return Sqrt(safety);
  }

  double SafetyToIn(const Vec3 &p) const override { return Safety(p, false); }
  double SafetyToOut(const Vec3 &p) const override { return Safety(p, true); }

  double DistanceToOut(const Vec3 &p, const Vec3 &d, double) const override
  {
    const int zMax = NSeg();
    if ((p[2] < fZPlanes[0] - kTolerance) || (p[2] > fZPlanes[zMax] + kTolerance)) return -1.;
    if (Inside(p) == kOutside) return -1.;
    int zIndex = FindZSegment(p[2]);
    zIndex = zIndex < 0 ? 0 : (zIndex >= zMax ? zMax - 1 : zIndex);
    bool goingRight = d[2] >= 0;
    double distance = kInfLength;
    if (goingRight) {
      for (; zIndex < zMax; ++zIndex) {
        distance = DistanceToOutZSegment(zIndex, fZPlanes[zIndex], fZPlanes[zIndex + 1], p, d);
        if (distance >= 0 && distance < kInfLength) break;
        if (fZPlanes[zIndex] - p[2] > distance) break;
      }
    } else {
      for (; zIndex >= 0; --zIndex) {
        distance = DistanceToOutZSegment(zIndex, fZPlanes[zIndex], fZPlanes[zIndex + 1], p, d);
        if (distance >= 0 && distance < kInfLength) break;
        if (p[2] - fZPlanes[zIndex] > distance) break;
      }
    }
    DistanceToEndcaps(true, p, d, distance);
    if (distance >= kInfLength) distance = 0.;
    return distance;
  }

  // UnplacedPolyhedron.cpp:273-303: bounding box over every corner of every
  // z plane (radii are side distances, corners sit 1/cos(sidePhi/2) further).
  // ScalarSafetyToEndcapsSquared with the endcap index (:620-650): iz = -1
  // or +1 when the endcap is the nearest surface so far
  void SafetyToEndcapsSquaredIz(const Vec3 &p, double &distanceSquared, int &iz) const
  {
    const double firstDistance = fZPlanes[0] - p[2];
    const double lastDistance = fZPlanes[NSeg()] - p[2];
    const bool isFirst = Abs(firstDistance) < Abs(lastDistance);
    iz = 0;
    const Seg &s = isFirst ? seg[0] : seg[NSeg() - 1];
    const double dt = isFirst ? firstDistance : lastDistance;
    const double dt2 = dt * dt;
    if (dt2 >= distanceSquared) return;
    const Vec3 x(p[0], p[1], p[2] + dt);
    if (!s.outer.Contains(x)) return;
    if (s.hasInner() && s.inner.Contains(x)) return;
    if (fHasPhiCutout && InPhiCutoutWedge(s, x)) return;
    iz = isFirst ? -1 : 1;
    distanceSquared = dt2;
  }

  // ScalarNormalKernel (:999-1068): the normal of the nearest quadrilateral
  // or endcap.  The reference indexes the outer shell with phiIndex even when
  // no surface was recorded (iSurf == -1, phiIndex possibly -1); such a point
  // gets the zero vector here.
  bool Normal(const Vec3 &point, Vec3 &normal) const override
  {
    double safety = kInfLength;
    const int zMax = NSeg();
    const int zIndex = FindZSegment(point[2]);
    if (zIndex < 0) {
      normal = Vec3(0, 0, -1);
      return true;
    }
    if (zIndex >= zMax) {
      normal = Vec3(0, 0, 1);
      return true;
    }
    int iSeg = zIndex, iSurf = -1, iz = 0;
    int phiIndex = FindPhiSegment(point);
    for (int z = zIndex; z < zMax;) {
      int iSurfCrt = -1;
      const double safetySeg = SafetyToZSegmentSquared(z, phiIndex, point, true, iSurfCrt);
      if (safetySeg < safety) {
        safety = safetySeg;
        iSeg = z;
        iSurf = iSurfCrt;
      }
      ++z;
      const double dz = fZPlanes[z] - point[2];
      if (dz * dz > safety) break;
    }
    for (int z = zIndex - 1; z >= 0; --z) {
      int iSurfCrt = -1;
      const double safetySeg = SafetyToZSegmentSquared(z, phiIndex, point, true, iSurfCrt);
      if (safetySeg < safety) {
        safety = safetySeg;
        iSeg = z;
        iSurf = iSurfCrt;
      }
      const double dz = point[2] - fZPlanes[z];
      if (dz * dz > safety) break;
    }
    SafetyToEndcapsSquaredIz(point, safety, iz);
    if (iz != 0) {
      normal = Vec3(0, 0, iz);
      return true;
    }
    const Seg &segment = seg[iSeg];
    if (iSurf >= 0 && iSurf < 2) {
      normal = segment.phi.normal[iSurf];
    } else if (iSurf == 2) {
      const Vec3 n = segment.inner.normal[phiIndex];
      normal = Vec3(-1. * n.x, -1. * n.y, -1. * n.z);
    } else if (phiIndex >= 0 && phiIndex < segment.outer.n) {
      normal = segment.outer.normal[phiIndex];
    } else {
      normal = Vec3();
    }
    return true;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = Vec3(kInfLength, kInfLength, kInfLength);
    mx = Vec3(-kInfLength, -kInfLength, -kInfLength);
    const double sidePhi = fPhiDelta / fSideCount;
    const double conv = 1. / std::cos(0.5 * sidePhi);
    for (int iphi = 0; iphi <= fSideCount; ++iphi) {
      const double phi = fPhiStart + iphi * sidePhi;
      const double cx = conv * std::cos(phi), cy = conv * std::sin(phi);
      for (size_t k = 0; k < fZPlanes.size(); ++k) {
        const double rr[2] = {fRMin[k], fRMax[k]};
        for (double r : rr) {
          const double c[3] = {r * cx, r * cy, fZPlanes[k]};
          for (int i = 0; i < 3; ++i) {
            mn[i] = c[i] < mn[i] ? c[i] : mn[i];
            mx[i] = mx[i] > c[i] ? mx[i] : c[i];
          }
        }
      }
    }
  }
};

Shape *make_polyhedron(double phiStart, double phiDelta, int sideCount, int nz, const double *z, const double *rmin,
                       const double *rmax)
{
  if (nz < 2 || sideCount < 1 || !(phiDelta > 0)) return nullptr;
  for (int i = 0; i + 1 < nz; ++i)
    if (z[i] > z[i + 1]) return nullptr;
  return new Polyhedron(phiStart, phiDelta, sideCount, nz, z, rmin, rmax);
}

}  // namespace orc
