// Oracle: polycone (PolyconeStruct::Init, VecGeom/volumes/PolyconeStruct.h;
// kernels VecGeom/volumes/kernel/PolyconeImplementation.h:66-515; sections are
// UniversalCone ConeStructs shifted along z, ConeTypes used as polycone types,
// source/UnplacedPolycone.cpp:48).  GetSectionIndex is the reference's linear
// search over fZs.  Scalar backend.
#include <memory>
#include <vector>
#include "cone.h"

namespace orc {

struct Polycone final : Shape {
  struct Section {
    std::unique_ptr<Cone> solid;
    double shift;
  };
  std::vector<Section> secs;
  std::vector<double> zs;
  double rmaxExtent = 0, zmin = 0, zmax = 0;

  Polycone(double phiStart, double phiTotal, int nz, const double *zPlane, const double *rInner, const double *rOuter)
  {
    std::vector<double> z(nz), rmin(nz), rmax(nz);
    for (int i = 0; i < nz; ++i) {
      z[i] = zPlane[i];
      rmin[i] = rInner[i];
      rmax[i] = rOuter[i];
    }
    if (zPlane[0] > zPlane[nz - 1]) {
      for (int i = 0; i < nz; ++i) {
        z[i] = zPlane[nz - 1 - i];
        rmin[i] = rInner[nz - 1 - i];
        rmax[i] = rOuter[nz - 1 - i];
      }
    }
    double RMaxextent = rmax[0];
    for (int j = 1; j < nz; ++j)
      if (rmax[j] > RMaxextent) RMaxextent = rmax[j];
    rmaxExtent = RMaxextent;
    double prevZ = z[0], prevRmax = 0, prevRmin = 0;
    int dirZ = 1;
    if (z[1] < z[0]) dirZ = -1;
    for (int i = 0; i < nz; ++i) {
      double rMin = rmin[i], rMax = rmax[i], zi = z[i];
      if (i > 0) {
        if (((zi > prevZ + kTolerance) && (dirZ > 0)) || ((zi < prevZ - kTolerance) && (dirZ < 0))) {
          double dz = (zi - prevZ) / 2;
          Section s;
          s.solid.reset(new Cone(prevRmin, prevRmax, rMin, rMax, dz, phiStart, phiTotal));
          zs.push_back(zi);
          int k = int(zs.size()) - 1;
          s.shift = zs[k - 1] + 0.5 * (zs[k] - zs[k - 1]);
          secs.push_back(std::move(s));
        }
      } else {
        zs.push_back(zi);
      }
      prevZ = zi;
      prevRmin = rMin;
      prevRmax = rMax;
    }
    zmin = zs.front();
    zmax = zs.back();
  }

  int NSec() const { return int(secs.size()); }

  // PolyconeStruct::GetSectionIndex (linear search)
  int SectionIndex(double zp) const
  {
    if (zp < zs[0]) return -1;
    for (int i = 0; i < NSec(); ++i)
      if (zp >= zs[i] && zp <= zs[i + 1]) return i;
    return -2;
  }

  static Vec3 Shifted(const Vec3 &p, double s) { return Vec3(p.x - 0., p.y - 0., p.z - s); }

  void SectionGeneric(int isect, const Vec3 &p, bool forInside, bool &in, bool &out) const
  {
    if (isect < 0) {
      in = false;
      out = true;
      return;
    }
    const Section &s = secs[isect];
    s.solid->Generic(Shifted(p, s.shift), forInside, in, out);
  }

  // GenericKernelForContainsAndInside (:126-187)
  void Generic(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    int lo = SectionIndex(p.z - kTolerance);
    int hi = SectionIndex(p.z + kTolerance);
    if (lo < 0 && hi < 0) {
      cout = true;
      return;
    }
    if (lo < 0 && hi == 0) {
      SectionGeneric(0, p, forInside, cin, cout);
      return;
    }
    if (hi < 0 && lo == NSec() - 1) {
      SectionGeneric(NSec() - 1, p, forInside, cin, cout);
      return;
    }
    if (lo >= 0 && hi >= 0) {
      if (lo == hi) {
        SectionGeneric(lo, p, forInside, cin, cout);
        return;
      }
      bool inLo = false, outLo = false, inHi = false, outHi = false;
      SectionGeneric(lo, p, forInside, inLo, outLo);
      SectionGeneric(hi, p, forInside, inHi, outHi);
      bool surfLo = !inLo && !outLo;
      bool surfHi = !inHi && !outHi;
      if (surfLo && surfHi) {
        cin = true;
        return;
      }
      if (outLo && outHi) {
        cout = true;
        return;
      }
    }
  }

  bool Contains(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    Generic(p, false, ci, co);
    return !co;
  }

  int Inside(const Vec3 &p) const override
  {
    bool ci = false, co = false;
    Generic(p, true, ci, co);
    int r = kSurface;
    if (co) r = kOutside;
    if (ci) r = kInside;
    return r;
  }

  double DistanceToIn(const Vec3 &p, const Vec3 &v, double stepMax) const override  // :192-237
  {
    double distance = kInfLength;
    int increment = (v.z > 0) ? 1 : -1;
    if (std::fabs(v.z) < kTolerance) increment = 0;
    int index = SectionIndex(p.z);
    if (index == -1) index = 0;
    if (index == -2) index = NSec() - 1;
    do {
      const Section &s = secs[index];
      distance = s.solid->DistanceToIn(Shifted(p, s.shift), v, stepMax);
      if (distance < kInfLength || !increment) break;
      index += increment;
    } while (index >= 0 && index < NSec());
    return distance;
  }

  double DistanceToOut(const Vec3 &point, const Vec3 &dir, double stepMax) const override  // :242-379
  {
    double distance = kInfLength;
    Vec3 pn = point;
    if (NSec() == 1) {
      const Section &s = secs[0];
      return s.solid->DistanceToOut(Shifted(point, s.shift), dir, stepMax);
    }
    int lo = SectionIndex(point.z - kTolerance);
    int hi = SectionIndex(point.z + kTolerance);
    int index = 0;
    if (lo < 0 && hi < 0) {
      return -1;
    } else if (lo < 0 && hi >= 0) {
      index = hi;
      const Section &s = secs[index];
      if (s.solid->Inside(Shifted(point, s.shift)) == kOutside) return -1;
    } else if (lo != hi && (lo >= 0)) {
      const Section &s = secs[lo];
      if (s.solid->Inside(Shifted(point, s.shift)) == kOutside)
        index = hi;
      else
        index = lo;
    } else {
      index = lo;
      if (index < 0) index = SectionIndex(point.z);
    }
    if (index < 0) {
      return 0.;
    } else {
      const Section &s = secs[index];
      if (s.solid->Inside(Shifted(point, s.shift)) == kOutside) return -1;
    }
    double totalDistance = 0.;
    double dist;
    int increment = (dir.z > 0) ? 1 : -1;
    if (std::fabs(dir.z) < kTolerance) increment = 0;
    int istep = 0;
    do {
      const Section &s = secs[index];
      if ((totalDistance != 0) || (istep < 2)) {
        pn = point + dir * totalDistance;
        pn.z -= s.shift;
        if (s.solid->Inside(pn) == kOutside) break;
      } else {
        pn.z -= s.shift;
      }
      istep++;
      dist = s.solid->DistanceToOut(pn, dir, stepMax);
      if (dist == -1) return distance;
      if (std::fabs(dist) < 0.5 * kTolerance) {
        int index1 = index;
        if ((index > 0) && (index < NSec() - 1)) {
          index1 += increment;
        } else {
          if ((index == 0) && (increment > 0)) index1 += increment;
          if ((index == NSec() - 1) && (increment < 0)) index1 += increment;
        }
        Vec3 pte = point + dir * (totalDistance + dist);
        const Section &s1 = secs[index1];
        pte.z -= s1.shift;
        int inside22 = s1.solid->Inside(pte);
        if (inside22 == 3 || (increment == 0)) break;
      }
      totalDistance += dist;
      index += increment;
    } while (increment != 0 && index >= 0 && index < NSec());
    distance = totalDistance;
    return distance;
  }

  double SafetyToIn(const Vec3 &p) const override  // :384-452
  {
    int index = SectionIndex(p.z);
    bool needZ = false;
    if (index < 0) {
      needZ = true;
      if (index == -1) index = 0;
      if (index == -2) index = NSec() - 1;
    }
    const Section &sec = secs[index];
    double safety = sec.solid->SafetyToIn(Shifted(p, sec.shift));
    if (needZ) return safety;
    if (safety < kTolerance) return safety;
    double minSafety = safety;
    double zbase = zs[index + 1];
    for (int i = index + 1; i < NSec(); ++i) {
      double dz = zs[i] - zbase;
      if (dz >= minSafety) break;
      safety = secs[i].solid->SafetyToIn(Shifted(p, secs[i].shift));
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = zs[index - 1];
      for (int i = index - 1; i >= 0; --i) {
        double dz = zbase - zs[i];
        if (dz >= minSafety) break;
        safety = secs[i].solid->SafetyToIn(Shifted(p, secs[i].shift));
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }

  double SafetyToOut(const Vec3 &point) const override  // :455-515
  {
    bool ci = false, co = false;
    Generic(point, true, ci, co);
    if (co) return -1;
    int index = SectionIndex(point.z);
    if (index < 0) return -1;
    const Section &sec = secs[index];
    double safety = sec.solid->SafetyToOut(Shifted(point, sec.shift));
    double minSafety = safety;
    if (minSafety == kInfLength) return 0.;
    if (minSafety < kTolerance) return 0.;
    double zbase = zs[index + 1];
    for (int i = index + 1; i < NSec(); ++i) {
      double dz = zs[i] - zbase;
      if (dz >= minSafety) break;
      safety = secs[i].solid->SafetyToIn(Shifted(point, secs[i].shift));
      if (safety < minSafety) minSafety = safety;
    }
    if (index > 0) {
      zbase = zs[index - 1];
      for (int i = index - 1; i >= 0; --i) {
        double dz = zbase - zs[i];
        if (dz >= minSafety) break;
        safety = secs[i].solid->SafetyToIn(Shifted(point, secs[i].shift));
        if (safety < minSafety) minSafety = safety;
      }
    }
    return minSafety;
  }

  // UnplacedPolycone::Normal (source/UnplacedPolycone.cpp:782-832): the
  // section's cone normal; on the plane between two sections the next
  // section's too
  bool Normal(const Vec3 &point, Vec3 &norm) const override
  {
    bool valid = true;
    const int index = SectionIndex(point.z - kTolerance);
    if (index < 0) {
      valid = false;
      if (index == -1) norm = Vec3(0., 0., -1.);
      if (index == -2) norm = Vec3(0., 0., 1.);
      return valid;
    }
    const Section &sec = secs[index];
    valid = sec.solid->Normal(Shifted(point, sec.shift), norm);
    if (index + 1 < NSec() && std::abs(point.z - zs[index + 1]) < kTolerance) {
      const Section &sec2 = secs[index + 1];
      Vec3 norm2;
      const bool valid2 = sec2.solid->Normal(Shifted(point, sec2.shift), norm2);
      if (!valid && valid2) {
        norm = norm2;
        valid = valid2;
      }
      if (valid && valid2) {
        const double s = kTolerance * 10;
        const Vec3 q = point + Vec3(s * norm2.x, s * norm2.y, s * norm2.z);
        const bool c2 = sec2.solid->Contains(Shifted(q, sec2.shift));
        if (c2) {
          norm = norm2;
        } else {
          norm = norm + norm2;
          if (norm.Mag2() < kTolerance) norm = Vec3(point.x, point.y, 0);
        }
      }
    }
    if (valid) norm = norm / norm.Mag();
    return valid;
  }
  bool HasNormal() const override { return true; }

  void Extent(Vec3 &mn, Vec3 &mx) const override
  {
    mn = Vec3(-rmaxExtent, -rmaxExtent, zmin);
    mx = Vec3(rmaxExtent, rmaxExtent, zmax);
  }
};

Shape *make_polycone(double sphi, double dphi, int nz, const double *z, const double *rmin, const double *rmax)
{
  if (nz < 2) return nullptr;
  Polycone *p = new Polycone(sphi, dphi, nz, z, rmin, rmax);
  if (p->NSec() < 1) {
    delete p;
    return nullptr;
  }
  return p;
}

}  // namespace orc
