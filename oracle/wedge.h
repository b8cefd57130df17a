// Oracle: phi wedge (VecGeom/volumes/Wedge_Evolution.h:43-420,
// source/Wedge_Evolution.cpp for the along/normal vectors).  TEST ONLY.
#pragma once
#include "core.h"

namespace orc {

struct Wedge {
  double fSPhi = 0, fDPhi = 0;
  Vec3 fAlong1, fAlong2, fNormal1, fNormal2;

  Wedge() = default;
  Wedge(double angle, double zeroangle) : fSPhi(zeroangle), fDPhi(angle)
  {
    fAlong1 = Vec3(std::cos(zeroangle), std::sin(zeroangle), 0);
    fAlong2 = Vec3(std::cos(zeroangle + angle), std::sin(zeroangle + angle), 0);
    fNormal1 = Vec3(-std::sin(zeroangle), std::cos(zeroangle), 0);
    fNormal2 = Vec3(std::sin(zeroangle + angle), -std::cos(zeroangle + angle), 0);
  }

  Vec3 GetNormal1() const { return fNormal1; }
  Vec3 GetNormal2() const { return fNormal2; }

  static bool IsOnSurfaceGeneric(const Vec3 &along, const Vec3 &normal, const Vec3 &p)
  {
    bool c1 = along.x * p.x + along.y * p.y >= 0.;
    if (!c1) return false;
    return Abs(normal.x * p.x + normal.y * p.y) < kTolerance;
  }

  void GenericKernel(const Vec3 &p, bool forInside, bool &cin, bool &cout) const
  {
    double startCheck = (-p.x * fAlong1.y + p.y * fAlong1.x);
    double endCheck = (-fAlong2.x * p.y + fAlong2.y * p.x);
    cout = startCheck < 0.;
    if (fDPhi < kPi)
      cout |= endCheck < 0.;
    else
      cout &= endCheck < 0.;
    if (forInside) {
      bool onSurface = IsOnSurfaceGeneric(fAlong1, fNormal1, p) || IsOnSurfaceGeneric(fAlong2, fNormal2, p);
      cout &= !onSurface;
      cin = !onSurface && !cout;
    }
  }

  bool Contains(const Vec3 &p) const
  {
    bool ci = false, co = false;
    GenericKernel(p, false, ci, co);
    return !co;
  }

  double SafetyToOut(const Vec3 &p) const
  {
    double d1 = p.x * fNormal1.x + p.y * fNormal1.y;
    double d2 = p.x * fNormal2.x + p.y * fNormal2.y;
    return (fDPhi < kPi) ? Min(d1, d2) : Max(d1, d2);
  }

  double SafetyToIn(const Vec3 &p) const
  {
    double d1 = p.x * fNormal1.x + p.y * fNormal1.y;
    double d2 = p.x * fNormal2.x + p.y * fNormal2.y;
    return (fDPhi < kPi) ? Max(-1 * d1, -1 * d2) : Min(-1 * d1, -1 * d2);
  }

  void DistanceToIn(const Vec3 &p, const Vec3 &d, double &w1, double &w2) const
  {
    w1 = kInfLength;
    w2 = kInfLength;
    double c1 = d.x * fNormal1.x + d.y * fNormal1.y;
    double c2 = d.x * fNormal2.x + d.y * fNormal2.y;
    if (c1 > 0.) {
      double t = -(p.x * fNormal1.x + p.y * fNormal1.y) / c1;
      if (t > 0.) w1 = t;
    }
    if (c2 > 0.) {
      double t = -(p.x * fNormal2.x + p.y * fNormal2.y) / c2;
      if (t > 0.) w2 = t;
    }
  }

  void DistanceToOut(const Vec3 &p, const Vec3 &d, double &w1, double &w2) const
  {
    double c1 = d.x * fNormal1.x + d.y * fNormal1.y;
    double c2 = d.x * fNormal2.x + d.y * fNormal2.y;
    w1 = kInfLength;
    w2 = kInfLength;
    if (c1 < 0.) {
      double t = -(p.x * fNormal1.x + p.y * fNormal1.y) / c1;
      if (t > 0.) w1 = t;
    }
    if (c2 < 0.) {
      double t = -(p.x * fNormal2.x + p.y * fNormal2.y) / c2;
      if (t > 0.) w2 = t;
    }
  }
};

}  // namespace orc
