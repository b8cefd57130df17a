// Polyhedron kernels (VecGeom/volumes/kernel/PolyhedronImplementation.h:300-1200,
// the generic inner-radius / phi-cutout treatment that every factory
// specialisation computes identically) over the quadrilateral shells of
// Quadrilaterals.h / Planes.h.  Scalar-backend semantics, FP64 or FP32 by R.
//
// Each z segment holds up to three shells of quads (outer, inner, two phi
// cut planes).  A lane walks segments in z order and scans one shell at a
// time; the quads of one solid sit together in the pool, so the scans hit
// L1/L2 after the first lane of a wave touches them.
#pragma once
#include "tube.hpp"
#include "wedge.hpp"
#include "polyhedron_data.h"

namespace vgm {

struct PolyhedronShape {
  using Data = PolyhedronData;
  static constexpr bool kStage = true;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for: the latency
  // of the quad walks hides better at 6 waves (80 VGPRs) than at the 2 its
  // unconstrained register allocation allows, although a few values then
  // spill to scratch (50 steps of 6.7M tracks, distance first: 87.2 ms at 4
  // waves, 79.6 at 5, 77.5 at 6, 78.7 at 8)
  static constexpr int kWavesPerEU = 6;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = false;
  // SafetyToIn/Out never exceed DistanceToIn/Out beyond the surface tolerance:
  // the distance-first step takes the FP64 step's decisions
  static constexpr bool kSafetyBelowDistance = true;
  // mixed precision: Inside/Contains classified in float away from the surface
  static constexpr bool kFloatClassify = true;
  // mixed precision: the float safety less the guard band bounds the FP64 safety
  static constexpr bool kFloatSafety = false;
  // mixed precision: distance_to_in_mixed / distance_to_out_mixed (float selects
  // the surface, FP64 evaluates it)
  static constexpr bool kMixedDistance = false;
  using Quad = PolyhedronQuad;

  template <typename R>
  VGM_DEV static R pdot(const double *n, V3<R> p)
  {
    return p.x * R(n[0]) + p.y * R(n[1]) + p.z * R(n[2]);
  }
  template <typename R>
  VGM_DEV static R plane(const Quad &q, V3<R> p)
  {
    return pdot<R>(q.n, p) + R(q.d);
  }
  template <typename R>
  VGM_DEV static bool side_ok(const Quad &q, int j, V3<R> x, R tol)
  {
    return pdot<R>(q.sn[j], x) + R(q.sd[j]) >= tol;
  }

  // Planes::Contains: strictly behind every plane of the shell.
  template <typename R>
  VGM_DEV static bool shell_contains(const Quad *q, int n, V3<R> p)
  {
    for (int i = 0; i < n; ++i)
      if (!(plane<R>(q[i], p) < R(0))) return false;
    return true;
  }

  // Planes::Inside(point, i)
  template <typename R>
  VGM_DEV static int quad_inside(const Quad &q, V3<R> p)
  {
    const R dr = R(q.n[0]) * p.x + R(q.n[1]) * p.y + R(q.n[2]) * p.z + R(q.d);
    int r = kInsideV;
    if (dr > R(kHalfTolerance)) r = kOutsideV;
    if (r == kInsideV && dr > R(-kHalfTolerance)) r = kSurfaceV;
    return r;
  }

  // Planes::Inside(point): the first plane that puts the point outside ends the scan.
  template <typename R>
  VGM_DEV static int shell_inside(const Quad *q, int n, V3<R> p)
  {
    int r = kInsideV;
    for (int i = 0; i < n; ++i) {
      const R dr = R(q[i].n[0]) * p.x + R(q[i].n[1]) * p.y + R(q[i].n[2]) * p.z + R(q[i].d);
      if (dr > R(kHalfTolerance)) r = kOutsideV;
      if (r == kInsideV && dr > R(-kHalfTolerance)) r = kSurfaceV;
      if (r == kOutsideV) break;
    }
    return r;
  }

  // Quadrilaterals::DistanceToIn<behindPlanesT> (Quadrilaterals.h:320-350):
  // the first quad hit from its front (or, behind, back) side wins.
  template <typename R>
  VGM_DEV static R shell_distance_to_in(const Quad *q, int n, V3<R> p, V3<R> d, bool behind)
  {
    R best = inf_length<R>();
    for (int i = 0; i < n; ++i) {
      R distance = plane<R>(q[i], p);
      if (!((behind ? -distance : distance) > R(-kTolerance))) continue;
      const R proj = pdot<R>(q[i].n, d);
      if (!((behind ? proj : -proj) > R(0))) continue;
      distance /= -(proj + vcopysign(R(1e-20), proj));
      const V3<R> x{p.x + d.x * distance, p.y + d.y * distance, p.z + d.z * distance};
      bool valid = true;
      for (int j = 0; j < 4; ++j) {
        valid &= side_ok<R>(q[i], j, x, R(-kTolerance));
        if (!valid) break;
      }
      if (valid) best = distance;
      if (best < inf_length<R>()) break;
    }
    return vmax(R(0), best);
  }

  // Quadrilaterals::DistanceToOut(point, direction, zMin, zMax) (:420-455)
  template <typename R>
  VGM_DEV static R shell_distance_to_out(const Quad *q, int n, V3<R> p, V3<R> d, R zMin, R zMax)
  {
    R best = inf_length<R>();
    for (int i = 0; i < n; ++i) {
      R dt = plane<R>(q[i], p);
      if (!(dt < R(kTolerance))) continue;
      const R proj = pdot<R>(q[i].n, d);
      if (!(proj > R(0))) continue;
      dt /= -proj;
      if (!(dt < best)) continue;
      bool valid;
      if (zMin == zMax) {
        const V3<R> x{p.x + d.x * dt, p.y + d.y * dt, p.z + d.z * dt};
        valid = true;
        for (int j = 0; j < 4; ++j) {
          valid &= side_ok<R>(q[i], j, x, R(-kTolerance));
          if (!valid) break;
        }
      } else {
        const R zp = p.z + dt * d.z;
        valid = (zp >= zMin - R(kTolerance)) && (zp < zMax + R(kTolerance));
      }
      if (!valid) continue;
      best = dt;
    }
    if (best > R(-kTolerance)) best = vmax(best, R(0));
    return best;
  }

  // DistanceToLineSegmentSquared1 (GenericKernels.h:190-215)
  template <typename R>
  VGM_DEV static R segment_dist2(const double *c0, const double *c1, V3<R> p)
  {
    const R lx = R(c1[0]) - R(c0[0]), ly = R(c1[1]) - R(c0[1]), lz = R(c1[2]) - R(c0[2]);
    const R px = p.x - R(c0[0]), py = p.y - R(c0[1]), pz = p.z - R(c0[2]);
    const R dot0 = px * lx + py * ly + pz * lz;
    if (dot0 <= R(0)) return px * px + py * py + pz * pz;
    const R dot1 = lx * lx + ly * ly + lz * lz;
    if (dot1 <= dot0) {
      const R qx = p.x - R(c1[0]), qy = p.y - R(c1[1]), qz = p.z - R(c1[2]);
      return qx * qx + qy * qy + qz * qz;
    }
    const R t = dot0 / dot1;
    const R qx = R(c0[0]) + lx * t, qy = R(c0[1]) + ly * t, qz = R(c0[2]) + lz * t;
    const R ex = qx - p.x, ey = qy - p.y, ez = qz - p.z;
    return ex * ex + ey * ey + ez * ez;
  }

  // Quadrilaterals::ScalarDistanceSquared (:460-490): exact squared
  // distance to the quad, through its plane or its nearest edge.
  template <typename R>
  VGM_DEV static R quad_dist2(const Quad &q, V3<R> p)
  {
    const R distance = plane<R>(q, p);
    const V3<R> x{p.x - R(q.n[0]) * distance, p.y - R(q.n[1]) * distance, p.z - R(q.n[2]) * distance};
    unsigned outside = 0;  // bit j: the foot point is outside side plane j
    for (int j = 0; j < 4; ++j) outside |= side_ok<R>(q, j, x, R(0)) ? 0u : (1u << j);
    if (!outside) return distance * distance;
    R dsq = inf_length<R>();
    // one call site for the edge distance keeps the kernel small
#pragma unroll 1
    for (int j = 0; j < 4; ++j) {
      if (outside & (1u << j)) {
        const R e = segment_dist2<R>(q.c[j], q.c[(j + 1) & 3], p);
        if (e < dsq) dsq = e;
      }
    }
    return dsq;
  }

  // FindZSegmentKernel (PolyhedronImplementation.h:313-327)
  template <typename R>
  VGM_DEV static int find_z_segment(const Data &h, R pz)
  {
    const double *zp = h.zplanes();
    const int nz = h.nseg + 1;
    int index = -1, k = 0;
    while (k < nz - 1 && pz - R(kTolerance) > R(zp[k])) {
      ++index;
      ++k;
    }
    if (pz + R(kTolerance) > R(zp[k])) return index + 1;
    return index;
  }

  // FindPhiSegment (:344-364)
  template <typename R>
  VGM_DEV static int find_phi_segment(const Data &h, V3<R> p)
  {
    const double *ps = h.phisec();
    R first = p.x * R(ps[0]) + p.y * R(ps[1]) + p.z * R(ps[2]);
    for (int i = 1; i < h.nsides + 1; ++i) {
      const R second = p.x * R(ps[3 * i]) + p.y * R(ps[3 * i + 1]) + p.z * R(ps[3 * i + 2]);
      if (first > R(-kTolerance) && second < R(kTolerance)) return i - 1;
      first = second;
    }
    return -1;
  }

  template <typename R>
  VGM_DEV static bool bounds_contain(const Data &h, V3<R> p)
  {
    return TubeShape::contains<R>(h.btube, V3<R>{p.x, p.y, p.z - R(h.bt_offset)});
  }

  // DistanceToInZSegment (:368-400) and DistanceToOutZSegment (:403-445) as
  // one sequence with the side a run-time argument, so that the lanes of a
  // wave on both sides share the first two shell scans:
  //   outside: outer shell (front), phi shell (front), inner shell (behind)
  //   inside:  inner shell (front), phi shell (behind), outer shell exit
  // each returning the first shell's hit.  (The inside sequence keeps a phi
  // hit only when no inner hit came first, and a shell hit is never negative,
  // so it reduces to the same first-hit rule.)
  template <typename R>
  VGM_DEV static R zseg_distance(const Data &h, int si, R zMin, R zMax, V3<R> p, V3<R> d, bool in)
  {
    const PolyhedronSeg &s = h.segs()[si];
    const Quad *q = h.quads();
    if (!in || s.n_inner > 0) {
      const R a = shell_distance_to_in<R>(q + (in ? s.inner : s.outer), in ? s.n_inner : s.n_outer, p, d, false);
      if (a < inf_length<R>()) return a;
    }
    if (h.phi_cutout) {
      const R b = shell_distance_to_in<R>(q + s.phi, s.n_phi, p, d, in);
      if (b < inf_length<R>()) return b;
    }
    if (in) return shell_distance_to_out<R>(q + s.outer, s.n_outer, p, d, zMin, zMax);
    return s.n_inner > 0 ? shell_distance_to_in<R>(q + s.inner, s.n_inner, p, d, true) : inf_length<R>();
  }

  // InPhiCutoutWedge (:659-671).  A zero-height end segment has no phi
  // quads; the reference then reads past the shell (undefined), here the
  // point counts as outside the cutout.
  template <typename R>
  VGM_DEV static bool in_phi_cutout_wedge(const Data &h, const PolyhedronSeg &s, V3<R> p)
  {
    if (s.n_phi < 2) return false;
    const Quad *q = h.quads() + s.phi;
    const bool s0 = pdot<R>(q[0].n, p) + R(q[0].d) >= R(0);
    const bool s1 = pdot<R>(q[1].n, p) + R(q[1].d) >= R(0);
    if (h.large_cutout) return s0 || s1;
    return s0 && s1;
  }

  // ScalarSafetyToZSegmentSquared (:448-560).  The reference's branches pick
  // at most four quads and keep the smallest squared distance (some of them
  // also record the phi index of the winner, and the surface it lies on:
  // 0/1 the phi planes, 2 the inner shell, 3 the outer).  Here the branches
  // only fill a list of (quad, phi index to record, surface) and one loop
  // evaluates it, so the quad distance is compiled once; the comparisons and
  // their order are the reference's.  `iSurf` (the Normal kernel's) receives
  // the winner's surface; the safeties pass none.
  template <typename R>
  VGM_DEV static R safety_zseg2(const Data &h, int si, int &phiIndex, V3<R> p, bool pt_inside,
                                int *iSurf = nullptr)
  {
    const PolyhedronSeg &s = h.segs()[si];
    const bool in_cutout = phiIndex < 0;
    uint64_t cq = 0;   // quad indices, 16 bits each
    uint32_t cp = 0;   // phi index to record + 1, 8 bits each (0: none)
    uint32_t cs = 0;   // surface, 2 bits each
    int nc = 0;
    bool first_assigns = false;  // the first candidate is assigned, not compared
    auto add = [&](int quad, int phi_rec, int surf) {
      cq |= uint64_t(uint32_t(quad) & 0xffffu) << (16 * nc);
      cp |= uint32_t(phi_rec + 1) << (8 * nc);
      cs |= uint32_t(surf) << (2 * nc);
      ++nc;
    };
    bool done = false;
    if (h.phi_cutout && s.n_phi == 2 && (pt_inside || in_cutout)) {
      first_assigns = true;
      add(s.phi, -1, 0);
      add(s.phi + 1, -1, 1);
      if (in_cutout) {
        if (h.large_cutout && s.n_inner > 0) {
          add(s.inner, 0, 2);
          if (s.n_inner > 1) add(s.inner + s.n_inner - 1, s.n_inner - 1, 2);
        }
        done = true;
      }
    }
    if (!done) {
      if (in_cutout && si > 0 && si < h.nseg - 1 && h.zplanes()[si] == h.zplanes()[si + 1]) {
        if (s.n_outer > 0) {
          add(s.outer, 0, 3);
          if (s.n_outer > 1) add(s.outer + s.n_outer - 1, s.n_outer - 1, 3);
        }
        if (s.n_inner > 0) {
          add(s.inner, 0, 2);
          if (s.n_inner > 1) add(s.inner + s.n_inner - 1, s.n_inner - 1, 2);
        }
      } else if (phiIndex >= 0) {
        // A phi-cutout point reaches the shells only for a zero-height end
        // segment, where the reference indexes them with -1 (out of bounds);
        // it is skipped here.  Inner is compared before outer.
        if (s.n_inner > 0) add(s.inner + phiIndex, -1, 2);
        if (s.n_outer > 0) add(s.outer + phiIndex, -1, 3);
      }
    }
    const Quad *q = h.quads();
    R safetySquared = inf_length<R>();
#pragma unroll 1
    for (int c = 0; c < nc; ++c) {
      const R v = quad_dist2<R>(q[int((cq >> (16 * c)) & 0xffffu)], p);
      if ((c == 0 && first_assigns) || v < safetySquared) {
        safetySquared = v;
        const int rec = int((cp >> (8 * c)) & 0xffu) - 1;
        if (rec >= 0) phiIndex = rec;
        if (iSurf) *iSurf = int((cs >> (2 * c)) & 3u);
      }
    }
    return safetySquared;
  }

  // ScalarDistanceToEndcaps<pointInsideT> (:565-615)
  template <typename R>
  VGM_DEV static void distance_to_endcaps(const Data &h, bool pointInside, V3<R> p, V3<R> d, R &distance)
  {
    const int last = h.nseg - 1;
    const bool first = pointInside ? (d.z < R(0)) : !(d.z < R(0));
    const PolyhedronSeg &s = h.segs()[first ? 0 : last];
    const R zPlane = R(h.zplanes()[first ? 0 : h.nseg]);
    const R dt = (zPlane - p.z) / nonzero(d.z);
    if (dt < R(-kTolerance) || dt >= distance) return;
    const V3<R> x{p.x + d.x * dt, p.y + d.y * dt, p.z + d.z * dt};
    const Quad *q = h.quads();
    if (!shell_contains<R>(q + s.outer, s.n_outer, x)) return;
    if (s.n_inner > 0 && shell_contains<R>(q + s.inner, s.n_inner, x)) return;
    if (h.phi_cutout && in_phi_cutout_wedge<R>(h, s, x)) return;
    distance = dt;
  }

  // ScalarSafetyToEndcapsSquared (:620-650); `iz` (the Normal kernel's)
  // receives -1 or +1 when an endcap is the nearer surface
  template <typename R>
  VGM_DEV static void safety_to_endcaps2(const Data &h, V3<R> p, R &distanceSquared, int *iz = nullptr)
  {
    const R firstDistance = R(h.zplanes()[0]) - p.z;
    const R lastDistance = R(h.zplanes()[h.nseg]) - p.z;
    const bool isFirst = vabs(firstDistance) < vabs(lastDistance);
    const PolyhedronSeg &s = h.segs()[isFirst ? 0 : h.nseg - 1];
    const R dt = isFirst ? firstDistance : lastDistance;
    const R dt2 = dt * dt;
    if (dt2 >= distanceSquared) return;
    const V3<R> x{p.x, p.y, p.z + dt};
    const Quad *q = h.quads();
    if (!shell_contains<R>(q + s.outer, s.n_outer, x)) return;
    if (s.n_inner > 0 && shell_contains<R>(q + s.inner, s.n_inner, x)) return;
    if (h.phi_cutout && in_phi_cutout_wedge<R>(h, s, x)) return;
    if (iz) *iz = isFirst ? -1 : 1;
    distanceSquared = dt2;
  }

  template <typename R>
  VGM_DEV static int wedge_inside(const Data &h, V3<R> p)
  {
    bool ci = false, co = false;
    wedge_generic<R, true>(h.w_a1x, h.w_a1y, h.w_a2x, h.w_a2y, h.w_n1x, h.w_n1y, h.w_n2x, h.w_n2y, h.w_lt_pi != 0,
                           p, ci, co);
    int r = kSurfaceV;
    if (co) r = kOutsideV;
    if (ci) r = kInsideV;
    return r;
  }

  // Side-plane normal 0 of the segment's outer (else inner) quad at phiIndex:
  // the radial direction used for the "radius" of a point on a z border.
  template <typename R>
  VGM_DEV static bool border_radius(const Data &h, const PolyhedronSeg &s, int phiIndex, V3<R> p, R &rdv)
  {
    const Quad *q = h.quads();
    const Quad *qq = s.n_outer ? &q[s.outer + phiIndex] : (s.n_inner ? &q[s.inner + phiIndex] : nullptr);
    if (!qq) return false;
    rdv = vabs(pdot<R>(qq->sn[0], p));
    return true;
  }

  // ScalarInsideSegBorder (:797-834)
  template <typename R>
  VGM_DEV static int inside_seg_border(const Data &h, V3<R> p, int zIndex)
  {
    const PolyhedronSeg &s = h.segs()[zIndex];
    const int phiIndex = find_phi_segment<R>(h, p);
    if (phiIndex < 0) return kOutsideV;
    R rdv;
    if (!border_radius<R>(h, s, phiIndex, p, rdv)) return kOutsideV;
    const double *rn = h.rmin(), *rx = h.rmax();
    const bool in1 = (rdv > R(rn[zIndex]) - R(kTolerance)) && (rdv < R(rx[zIndex]) + R(kTolerance));
    const bool in2 = (rdv > R(rn[zIndex + 1]) - R(kTolerance)) && (rdv < R(rx[zIndex + 1]) + R(kTolerance));
    if (in1 && in2) {
      if ((rdv < R(rn[zIndex]) + R(kTolerance)) || (rdv > R(rx[zIndex]) - R(kTolerance)) ||
          (rdv < R(rn[zIndex + 1]) + R(kTolerance)) || (rdv > R(rx[zIndex + 1]) - R(kTolerance)))
        return kSurfaceV;
      if (h.phi_cutout) return wedge_inside<R>(h, p);
      return kInsideV;
    }
    if (!in1 && !in2) return kOutsideV;
    return kSurfaceV;
  }

  // ScalarInsideSegPhi (:837-885)
  template <typename R>
  VGM_DEV static int inside_seg_phi(const Data &h, V3<R> p, int zIndex, int phiIndex)
  {
    if (phiIndex < 0) return kOutsideV;
    const R dz = vabs(p.z - R(h.bt_offset)) - R(h.half_len);
    if (dz > R(kHalfTolerance)) return kOutsideV;
    if (h.samez()[zIndex] != 0.0) return inside_seg_border<R>(h, p, zIndex);
    const PolyhedronSeg &s = h.segs()[zIndex];
    const Quad *q = h.quads();
    if (s.n_outer > 0) {
      const int io = quad_inside<R>(q[s.outer + phiIndex], p);
      if (io != kInsideV) return io;
    }
    if (s.n_inner > 0) {
      const int ii = quad_inside<R>(q[s.inner + phiIndex], p);
      if (ii == kInsideV) return kOutsideV;
      if (ii == kSurfaceV) return kSurfaceV;
    }
    if (h.phi_cutout) {
      const int ip = wedge_inside<R>(h, p);
      if (ip != kInsideV) return ip;
    }
    if (vabs(dz) < R(kHalfTolerance)) return kSurfaceV;
    return kInsideV;
  }

  // ScalarContainsKernel (:675-733)
  template <typename R>
  VGM_DEV static bool contains(const Data &h, V3<R> p)
  {
    if (!bounds_contain<R>(h, p)) return false;
    const int zIndex = find_z_segment<R>(h, p.z);
    if (!((zIndex >= 0) && (zIndex < h.nseg))) return false;
    const PolyhedronSeg &s = h.segs()[zIndex];
    const Quad *q = h.quads();
    if (h.samez()[zIndex] != 0.0) {
      const int phiIndex = find_phi_segment<R>(h, p);
      if (phiIndex < 0) return false;
      R rdv;
      if (!border_radius<R>(h, s, phiIndex, p, rdv)) return false;
      const double *rn = h.rmin(), *rx = h.rmax();
      const bool in1 = (rdv >= R(rn[zIndex])) && (rdv <= R(rx[zIndex]));
      const bool in2 = (rdv >= R(rn[zIndex + 1])) && (rdv <= R(rx[zIndex + 1]));
      return in1 | in2;
    }
    if (!shell_contains<R>(q + s.outer, s.n_outer, p)) return false;
    if (s.n_inner > 0 && shell_contains<R>(q + s.inner, s.n_inner, p)) return false;
    if (h.phi_cutout && !shell_contains<R>(q + s.phi, s.n_phi, p)) return false;
    return true;
  }

  // ScalarInsideKernel (:737-793)
  template <typename R>
  VGM_DEV static int inside(const Data &h, V3<R> p)
  {
    if (!bounds_contain<R>(h, p)) return kOutsideV;
    int zIndex = find_z_segment<R>(h, p.z);
    if (zIndex > h.nseg - 1) zIndex = h.nseg - 1;
    if (zIndex < 0) zIndex = 0;
    if (h.samez()[zIndex] != 0.0) return inside_seg_border<R>(h, p, zIndex);
    const PolyhedronSeg &s = h.segs()[zIndex];
    const Quad *q = h.quads();
    const int io = shell_inside<R>(q + s.outer, s.n_outer, p);
    if (io != kInsideV) return io;
    if (s.n_inner > 0) {
      const int ii = shell_inside<R>(q + s.inner, s.n_inner, p);
      if (ii == kInsideV) return kOutsideV;
      if (ii == kSurfaceV) return kSurfaceV;
    }
    if (h.phi_cutout) {
      const int ip = shell_inside<R>(q + s.phi, s.n_phi, p);
      if (ip != kInsideV) return ip;
    }
    const R dz = vabs(vabs(p.z - R(h.bt_offset)) - R(h.half_len));
    if (dz < R(kTolerance)) return kSurfaceV;
    return kInsideV;
  }

  // DistanceToIn (:895-960) from outside, DistanceToOut (:1080-1140) from
  // inside, as one function of the side: the Inside check, the z-segment
  // walk and the endcaps are shared by the lanes of a wave on both sides.
  // (The reference's step_max reaches only the bounding tube's DistanceToIn,
  // which ignores it.)
  template <typename R>
  VGM_DEV static R boundary_distance(const Data &h, V3<R> p, V3<R> d, bool in)
  {
    const double *zp = h.zplanes();
    const int zMax = h.nseg;
    if (in) {
      if ((p.z < R(zp[0]) - R(kTolerance)) || (p.z > R(zp[zMax]) + R(kTolerance))) return R(-1);
    } else {
      if ((p.z < R(zp[0]) + R(kTolerance)) && d.z <= R(0)) return inf_length<R>();
      if ((p.z > R(zp[zMax]) - R(kTolerance)) && d.z >= R(0)) return inf_length<R>();
    }
    if (inside<R>(h, p) == (in ? kOutsideV : kInsideV)) return R(-1);
    R tubeDistance = R(0);
    if (!in) {
      const V3<R> bp{p.x, p.y, p.z - R(h.bt_offset)};
      if (!TubeShape::contains<R>(h.btube, bp)) {
        tubeDistance = TubeShape::distance_to_in<R>(h.btube, bp, d, inf_length<R>());
        if (tubeDistance == inf_length<R>()) return inf_length<R>();
      }
    }
    int zIndex = find_z_segment<R>(h, p.z);
    zIndex = zIndex < 0 ? 0 : (zIndex >= zMax ? zMax - 1 : zIndex);
    R distance = inf_length<R>();
    const bool up = d.z >= R(0);  // walk segments along the ray's z
    for (; zIndex >= 0 && zIndex < zMax; zIndex += up ? 1 : -1) {
      distance = zseg_distance<R>(h, zIndex, R(zp[zIndex]), R(zp[zIndex + 1]), p, d, in);
      if (distance >= R(0) && distance < inf_length<R>()) break;
      if (in && (up ? R(zp[zIndex]) - p.z : p.z - R(zp[zIndex])) > distance) break;
    }
    distance_to_endcaps<R>(h, in, p, d, distance);
    if (in) return (distance >= inf_length<R>()) ? R(0) : distance;
    return (distance >= tubeDistance - R(1E-6)) ? distance : inf_length<R>();
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &h, V3<R> p, V3<R> d, R /*step_max*/)
  {
    return boundary_distance<R>(h, p, d, false);
  }

  // SafetyToIn / SafetyToOut (:965-1075): walk z segments outwards from the
  // point's segment while the z gap is below the best squared distance.
  template <typename R>
  VGM_DEV static R safety(const Data &h, V3<R> p, bool pt_inside)
  {
    R s2 = inf_length<R>();
    const int zMax = h.nseg;
    int zIndex = find_z_segment<R>(h, p.z);
    zIndex = zIndex < 0 ? 0 : (zIndex >= zMax ? zMax - 1 : zIndex);
    int phiIndex = find_phi_segment<R>(h, p);
    const int in = inside_seg_phi<R>(h, p, zIndex, phiIndex);
    if (in == kSurfaceV) return R(0);
    const bool contains = (in == kInsideV);
    if (contains ^ pt_inside) return R(-1);
    const double *zp = h.zplanes();
    // upwards from the point's segment while the next plane is nearer than
    // the best distance, then downwards from the segment below it
    int z = zIndex;
    bool up = true;
    while (z >= 0) {
      if (up && z >= zMax) {
        up = false;
        z = zIndex - 1;
        continue;
      }
      s2 = vmin(s2, safety_zseg2<R>(h, z, phiIndex, p, pt_inside));
      if (up) {
        ++z;
        const R dz = R(zp[z]) - p.z;
        if (dz * dz > s2) {
          up = false;
          z = zIndex - 1;
        }
      } else {
        const R dz = p.z - R(zp[z]);
        if (dz * dz > s2) break;
        --z;
      }
    }
    safety_to_endcaps2<R>(h, p, s2);
    return vsqrt(s2);
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &h, V3<R> p)
  {
    return safety<R>(h, p, false);
  }
  template <typename R>
  VGM_DEV static R safety_to_out(const Data &h, V3<R> p)
  {
    return safety<R>(h, p, true);
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &h, V3<R> p, V3<R> d, R /*step_max*/)
  {
    return boundary_distance<R>(h, p, d, true);
  }

  // ScalarNormalKernel (:999-1068): the normal of the nearest quadrilateral
  // or endcap.  The reference indexes the outer shell with phiIndex even when
  // no surface was recorded (phiIndex possibly -1); such a point gets the
  // zero vector here.
  VGM_DEV static bool normal(const Data &h, V3<double> p, V3<double> &n)
  {
    double safety = kInfLength;
    const int zMax = h.nseg;
    const int zIndex = find_z_segment<double>(h, p.z);
    if (zIndex < 0) {
      n = V3<double>{0, 0, -1};
      return true;
    }
    if (zIndex >= zMax) {
      n = V3<double>{0, 0, 1};
      return true;
    }
    const double *zp = h.zplanes();
    int iSeg = zIndex, iSurf = -1, iz = 0;
    int phiIndex = find_phi_segment<double>(h, p);
    for (int z = zIndex; z < zMax;) {
      int iSurfCrt = -1;
      const double safetySeg = safety_zseg2<double>(h, z, phiIndex, p, true, &iSurfCrt);
      if (safetySeg < safety) {
        safety = safetySeg;
        iSeg = z;
        iSurf = iSurfCrt;
      }
      ++z;
      const double dz = zp[z] - p.z;
      if (dz * dz > safety) break;
    }
    for (int z = zIndex - 1; z >= 0; --z) {
      int iSurfCrt = -1;
      const double safetySeg = safety_zseg2<double>(h, z, phiIndex, p, true, &iSurfCrt);
      if (safetySeg < safety) {
        safety = safetySeg;
        iSeg = z;
        iSurf = iSurfCrt;
      }
      const double dz = p.z - zp[z];
      if (dz * dz > safety) break;
    }
    safety_to_endcaps2<double>(h, p, safety, &iz);
    if (iz != 0) {
      n = V3<double>{0, 0, double(iz)};
      return true;
    }
    const PolyhedronSeg &s = h.segs()[iSeg];
    const Quad *q = h.quads();
    if (iSurf >= 0 && iSurf < 2) {
      const Quad &f = q[s.phi + iSurf];
      n = V3<double>{f.n[0], f.n[1], f.n[2]};
    } else if (iSurf == 2) {
      const Quad &f = q[s.inner + phiIndex];
      n = V3<double>{-1. * f.n[0], -1. * f.n[1], -1. * f.n[2]};
    } else if (phiIndex >= 0 && phiIndex < s.n_outer) {
      const Quad &f = q[s.outer + phiIndex];
      n = V3<double>{f.n[0], f.n[1], f.n[2]};
    } else {
      n = V3<double>{0., 0., 0.};
    }
    return true;
  }
};

}  // namespace vgm
