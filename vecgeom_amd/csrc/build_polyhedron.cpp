// Polyhedron builder: PolyhedronStruct(phiStart, phiDelta, sideCount,
// zPlaneCount, zPlanes, rMin, rMax) and Initialize
// (VecGeom/volumes/PolyhedronStruct.h:116-440), with the quadrilateral plane
// and side-plane derivation of Quadrilaterals::Set (source/Quadrilaterals.cpp)
// and Planes::Set (source/Planes.cpp:30-38).  Radii are distances to the
// sides; corners sit 1/cos(sidePhi/2) further out.  The derived numbers must
// equal the reference's bit for bit, so every expression keeps its order.
#include <cmath>
#include <vector>
#include "geometry.h"
#include "shapes/polyhedron_data.h"

namespace vgm {

namespace {

constexpr double kTol = 1e-9;
constexpr double kHalfTol = 0.5e-9;
constexpr double kPiH = 3.14159265358979323846;
constexpr double kTwoPiH = 2. * kPiH;

struct P3 {
  double x = 0, y = 0, z = 0;
  P3() = default;
  P3(double a, double b, double c) : x(a), y(b), z(c) {}
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  P3 operator-(const P3 &o) const { return P3(x - o.x, y - o.y, z - o.z); }
  P3 operator*(double s) const { return P3(x * s, y * s, z * s); }
  double mag() const { return std::sqrt(x * x + y * y + z * z); }
};

double dot(const P3 &a, const P3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 cross(const P3 &a, const P3 &b) { return P3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
P3 normalized(const P3 &v) { return v * (1. / v.mag()); }
// NormalizeAngle (VecGeom/volumes/kernel/GenericKernels.h:143-146)
double normalize_angle(double a) { return a + kTwoPiH * (int(a < 0) - int(a / kTwoPiH)); }
// Vector3D::FixZeroes
P3 fix_zeroes(P3 v)
{
  for (int i = 0; i < 3; ++i)
    if (std::fabs(v[i]) < kTol) v[i] = 0.0;
  return v;
}
// Vector3D operator== compares each component within kTolerance.
bool tol_equal(const P3 &a, const P3 &b)
{
  return std::fabs(a[0] - b[0]) < kTol && std::fabs(a[1] - b[1]) < kTol && std::fabs(a[2] - b[2]) < kTol;
}
P3 from_cylindrical(double r, double phi, double z) { return P3(r * std::cos(phi), r * std::sin(phi), z); }

// Planes::Set(index, normal, origin)
void set_plane(double n_out[3], double &d_out, const P3 &nrm, const P3 &x0)
{
  const P3 fixed = fix_zeroes(nrm);
  const double inv = 1. / fixed.mag();
  n_out[0] = inv * fixed.x;
  n_out[1] = inv * fixed.y;
  n_out[2] = inv * fixed.z;
  d_out = inv * -dot(fixed, x0);
}

// Quadrilaterals::Set(index, c0, c1, c2, c3): the plane through the first
// three distinct corners, then one inward side plane per non-degenerate edge.
void set_quad(PolyhedronQuad &q, const P3 &c0, const P3 &c1, const P3 &c2, const P3 &c3)
{
  const P3 cs[4] = {c0, c1, c2, c3};
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) q.c[j][k] = cs[j][k];
  P3 ch[4];
  ch[0] = c0;
  int assigned = 1;
  for (int tested = 1; tested < 4; ++tested) {
    bool chosen = true;
    for (int j = 0; j < assigned; ++j)
      if (tol_equal(ch[j], cs[tested])) {
        chosen = false;
        break;
      }
    if (chosen) ch[assigned++] = cs[tested];
  }
  const double a = ch[0][1] * (ch[1][2] - ch[2][2]) + ch[1][1] * (ch[2][2] - ch[0][2]) + ch[2][1] * (ch[0][2] - ch[1][2]);
  const double b = ch[0][2] * (ch[1][0] - ch[2][0]) + ch[1][2] * (ch[2][0] - ch[0][0]) + ch[2][2] * (ch[0][0] - ch[1][0]);
  const double c = ch[0][0] * (ch[1][1] - ch[2][1]) + ch[1][0] * (ch[2][1] - ch[0][1]) + ch[2][0] * (ch[0][1] - ch[1][1]);
  double d = -ch[0][0] * (ch[1][1] * ch[2][2] - ch[2][1] * ch[1][2]) -
             ch[1][0] * (ch[2][1] * ch[0][2] - ch[0][1] * ch[2][2]) -
             ch[2][0] * (ch[0][1] * ch[1][2] - ch[1][1] * ch[0][2]);
  P3 nrm(a, b, c);
  const double inverseLength = 1. / nrm.mag();
  nrm = nrm * inverseLength;
  d *= inverseLength;
  nrm = fix_zeroes(nrm);
  q.n[0] = nrm.x;
  q.n[1] = nrm.y;
  q.n[2] = nrm.z;
  q.d = d;
  for (int j = 0; j < 4; ++j) {
    const P3 &a0 = cs[j], &a1 = cs[(j + 1) % 4];
    if (!tol_equal(a1, a0)) {
      set_plane(q.sn[j], q.sd[j], normalized(cross(nrm, a1 - a0)), a0);
    } else {
      q.sn[j][0] = q.sn[j][1] = q.sn[j][2] = 0;
      q.sd[j] = 0;
    }
  }
}

void flip_sign(PolyhedronQuad &q)
{
  for (int k = 0; k < 3; ++k) q.n[k] = -q.n[k];
  q.d = -q.d;
}

}  // namespace

int build_polyhedron(HostGeometry &g, double phiStart, double phiDelta, int sideCount, int nz, const double *zPlanes,
                     const double *rMin, const double *rMax, std::string &err)
{
  if (nz < 2 || sideCount < 1 || !zPlanes || !rMin || !rMax || !(phiDelta > 0)) {
    err = "polyhedron needs at least two z planes, one side and phiDelta > 0";
    return -1;
  }
  for (int i = 0; i + 1 < nz; ++i)
    if (!(zPlanes[i] <= zPlanes[i + 1])) {
      err = "polyhedron z planes must be increasing";
      return -1;
    }
  for (int i = 0; i < nz; ++i)
    if (!(rMin[i] >= 0 && rMax[i] >= rMin[i])) {
      err = "polyhedron needs 0 <= rmin <= rmax";
      return -1;
    }
  const int nseg = nz - 1;
  const bool hasPhiCutout = phiDelta < kTwoPiH;
  const bool hasLargePhiCutout = phiDelta < kPiH;
  const double fPhiStart = normalize_angle(phiStart);
  const double fPhiDelta = (phiDelta > kTwoPiH) ? kTwoPiH : phiDelta;

  // Shell sizes per segment: zero-height segments lose the shells that do
  // not change radius there, and all phi quads.
  std::vector<PolyhedronSeg> segs(nseg);
  int nquads = 0;
  for (int i = 0; i < nseg; ++i) {
    PolyhedronSeg &s = segs[i];
    s = PolyhedronSeg{};
    const bool hasInnerRadius = rMin[i] > 0 || rMin[i + 1] > 0;
    int mult = (zPlanes[i] == zPlanes[i + 1] && rMax[i] == rMax[i + 1]) ? 0 : 1;
    s.outer = nquads;
    s.n_outer = sideCount * mult;
    nquads += s.n_outer;
    s.phi = nquads;
    if (hasPhiCutout) {
      mult = (zPlanes[i] == zPlanes[i + 1]) ? 0 : 1;
      s.n_phi = 2 * mult;
      nquads += s.n_phi;
    }
    mult = (zPlanes[i] == zPlanes[i + 1] && rMin[i] == rMin[i + 1]) ? 0 : 1;
    s.inner = nquads;
    if (hasInnerRadius && mult > 0) {
      s.n_inner = sideCount * mult;
      nquads += s.n_inner;
    }
  }

  const double sidePhi = fPhiDelta / sideCount;
  std::vector<double> vertixPhi(sideCount + 1);
  std::vector<P3> phiSections(sideCount + 1);
  for (int i = 0; i < sideCount + 1; ++i) {
    vertixPhi[i] = normalize_angle(fPhiStart + i * sidePhi);
    const P3 cornerVector = fix_zeroes(normalized(from_cylindrical(1., vertixPhi[i], 0)));
    phiSections[i] = cross(normalized(cornerVector), P3(0, 0, -1));
  }
  if (!hasPhiCutout) vertixPhi[sideCount] = vertixPhi[0];
  const double cosHalfDeltaPhi = std::cos(0.5 * sidePhi);

  double innerRadius = 1.7976931348623157e308, outerRadius = -1.7976931348623157e308;
  for (int i = 0; i < nz; ++i) {
    if (rMin[i] < innerRadius) innerRadius = rMin[i];
    if (rMax[i] > outerRadius) outerRadius = rMax[i];
  }
  outerRadius /= cosHalfDeltaPhi;
  const double boundingTubeZ = 0.5 * (zPlanes[nz - 1] - zPlanes[0]) + kTol;
  const double kPhiTolerance = 100 * kTol;
  const double boundsPhiStart = !hasPhiCutout ? 0 : fPhiStart - kPhiTolerance;
  const double boundsPhiDelta = !hasPhiCutout ? kTwoPiH : fPhiDelta + 2 * kPhiTolerance;

  const auto VI = [&](int plane, int corner) { return plane * (sideCount + 1) + corner; };
  std::vector<P3> ov(nz * (sideCount + 1)), iv(nz * (sideCount + 1));
  for (int i = 0; i < nz; ++i) {
    for (int j = 0, jMax = sideCount + (hasPhiCutout ? 1 : 0); j < jMax; ++j) {
      ov[VI(i, j)] = fix_zeroes(from_cylindrical(rMax[i] / cosHalfDeltaPhi, vertixPhi[j], zPlanes[i]));
      iv[VI(i, j)] = fix_zeroes(from_cylindrical(rMin[i] / cosHalfDeltaPhi, vertixPhi[j], zPlanes[i]));
    }
    if (!hasPhiCutout) {
      ov[VI(i, sideCount)] = ov[VI(i, 0)];
      iv[VI(i, sideCount)] = iv[VI(i, 0)];
    }
  }

  std::vector<PolyhedronQuad> quads(nquads);
  const auto wrong = [](const PolyhedronQuad &q, const P3 &c) { return q.n[0] * c[0] + q.n[1] * c[1] < 0; };
  for (int ip = 0; ip < nseg; ++ip) {
    const PolyhedronSeg &s = segs[ip];
    for (int is = 0; is < s.n_outer; ++is) {
      PolyhedronQuad &q = quads[s.outer + is];
      set_quad(q, ov[VI(ip, is)], ov[VI(ip, is + 1)], ov[VI(ip + 1, is + 1)], ov[VI(ip + 1, is)]);
      if (wrong(q, ov[VI(ip, is)])) flip_sign(q);
    }
    for (int is = 0; is < s.n_inner; ++is) {
      PolyhedronQuad &q = quads[s.inner + is];
      set_quad(q, iv[VI(ip, is)], iv[VI(ip, is + 1)], iv[VI(ip + 1, is + 1)], iv[VI(ip + 1, is)]);
      if (wrong(q, iv[VI(ip, is)])) flip_sign(q);
    }
    if (hasPhiCutout && s.n_phi == 2) {
      PolyhedronQuad &q0 = quads[s.phi];
      set_quad(q0, iv[VI(ip, 0)], iv[VI(ip + 1, 0)], ov[VI(ip + 1, 0)], ov[VI(ip, 0)]);
      if (q0.n[0] * phiSections[0].x + q0.n[1] * phiSections[0].y + q0.n[2] * phiSections[0].z > 0) flip_sign(q0);
      PolyhedronQuad &q1 = quads[s.phi + 1];
      set_quad(q1, ov[VI(ip, sideCount)], ov[VI(ip + 1, sideCount)], iv[VI(ip + 1, sideCount)],
               iv[VI(ip, sideCount)]);
      const P3 &ps = phiSections[sideCount];
      if (q1.n[0] * ps.x + q1.n[1] * ps.y + q1.n[2] * ps.z < 0) flip_sign(q1);
    }
  }

  // Assemble the block.
  std::vector<double> block(PolyhedronData::bytes(nseg, sideCount, nquads) / sizeof(double), 0.0);
  PolyhedronData *h = reinterpret_cast<PolyhedronData *>(block.data());
  h->nseg = nseg;
  h->nsides = sideCount;
  h->phi_cutout = hasPhiCutout;
  h->large_cutout = hasLargePhiCutout;
  h->phi_start = fPhiStart;
  h->phi_delta = fPhiDelta;
  h->bt_offset = 0.5 * (zPlanes[0] + zPlanes[nz - 1]);
  h->half_len = 0.5 * (zPlanes[nseg] - zPlanes[0]);
  // Wedge(fPhiDelta, fPhiStart) (source/Wedge_Evolution.cpp)
  h->w_a1x = std::cos(fPhiStart);
  h->w_a1y = std::sin(fPhiStart);
  h->w_a2x = std::cos(fPhiStart + fPhiDelta);
  h->w_a2y = std::sin(fPhiStart + fPhiDelta);
  h->w_n1x = -std::sin(fPhiStart);
  h->w_n1y = std::cos(fPhiStart);
  h->w_n2x = std::sin(fPhiStart + fPhiDelta);
  h->w_n2y = -std::cos(fPhiStart + fPhiDelta);
  h->w_lt_pi = fPhiDelta < kPiH;
  h->nquads = nquads;
  make_tube_data(h->btube, innerRadius - kHalfTol, outerRadius + kHalfTol, boundingTubeZ, boundsPhiStart,
                 boundsPhiDelta);
  double *zp = const_cast<double *>(h->zplanes());
  double *rn = const_cast<double *>(h->rmin());
  double *rx = const_cast<double *>(h->rmax());
  double *sz = const_cast<double *>(h->samez());
  for (int i = 0; i < nz; ++i) {
    zp[i] = zPlanes[i];
    rn[i] = rMin[i];
    rx[i] = rMax[i];
    sz[i] = (i > 0 && i < nz - 1 && zPlanes[i] == zPlanes[i + 1]) ? 1.0 : 0.0;
  }
  double *ps = const_cast<double *>(h->phisec());
  for (int i = 0; i < sideCount + 1; ++i) {
    ps[3 * i + 0] = phiSections[i].x;
    ps[3 * i + 1] = phiSections[i].y;
    ps[3 * i + 2] = phiSections[i].z;
  }
  PolyhedronSeg *sg = const_cast<PolyhedronSeg *>(h->segs());
  for (int i = 0; i < nseg; ++i) sg[i] = segs[i];
  PolyhedronQuad *qd = const_cast<PolyhedronQuad *>(h->quads());
  for (int i = 0; i < nquads; ++i) qd[i] = quads[i];

  // UnplacedPolyhedron::Extent (source/UnplacedPolyhedron.cpp:273-303)
  double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  const double conv = 1. / std::cos(0.5 * sidePhi);
  for (int iphi = 0; iphi <= sideCount; ++iphi) {
    const double phi = fPhiStart + iphi * sidePhi;
    const double cx = conv * std::cos(phi), cy = conv * std::sin(phi);
    for (int k = 0; k < nz; ++k) {
      const double rr[2] = {rMin[k], rMax[k]};
      for (double r : rr) {
        const double c[3] = {r * cx, r * cy, zPlanes[k]};
        for (int i = 0; i < 3; ++i) {
          lo[i] = c[i] < lo[i] ? c[i] : lo[i];
          hi[i] = hi[i] > c[i] ? hi[i] : c[i];
        }
      }
    }
  }
  return g.add_shape(kPolyhedron, block.data(), block.size(), nseg, sideCount, lo, hi);
}

}  // namespace vgm
