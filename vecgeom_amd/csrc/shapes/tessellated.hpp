// Tessellated kernels (VecGeom/volumes/kernel/TessellatedImplementation.h:62-318)
// with the per-facet arithmetic of TessellatedCluster (DistanceToCluster,
// InsideCluster, SafetySq; VecGeom/volumes/TessellatedCluster.h:228-545) and
// the facet setup of Tile<3> (VecGeom/volumes/Tile.h:75-150).
//
// The reference walks float boxes of facet clusters in distance order and
// keeps the strict minimum; its results are the minimum over all facets.
// Here a four-wide bounding-volume tree over the facets (boxes enlarged by
// a margin, so pruning is conservative) is walked with a per-lane stack;
// equal distances go to the facet defined first.  Facet data stay in global
// memory: a 20k-facet solid is 3.8 MB, served from L2 / infinity cache.
#pragma once
#include "../vgm_math.hpp"
#include "tessellated_data.h"

namespace vgm {

// Tree-walk stack.  An entry is a node index and a lower bound of the
// node's key (the entry distance of its box along the ray, or its squared
// distance from the point) in the upper half of a float, so a popped node
// whose key exceeds the current bound is dropped without reading it.  On the
// device the stack lives in LDS, one column per thread of the workgroup (a
// private array would be spilled to scratch memory, whose traffic dominated
// the walks); on the host it is an ordinary local array.  The tree builder
// keeps node indices below 65536 and (kTessWidth - 1) x depth + 1 within
// kTessWalkDepth, so a nearest-first walk (at most kTessWidth - 1 deferred
// siblings per level) never overflows.
constexpr int kTessWalkDepth = 32;
constexpr int kTessWalkThreads = 256;  // largest workgroup of any kernel that walks
#if defined(__HIP_DEVICE_COMPILE__)
static __shared__ uint32_t tess_walk_lds[kTessWalkDepth * kTessWalkThreads];
struct TessWalkStack {
  uint32_t *col;
  __device__ TessWalkStack() : col(tess_walk_lds + threadIdx.x) {}
  __device__ void set(int i, uint32_t v) { col[i * kTessWalkThreads] = v; }
  __device__ uint32_t get(int i) const { return col[i * kTessWalkThreads]; }
};
#else
struct TessWalkStack {
  uint32_t a[kTessWalkDepth];
  void set(int i, uint32_t v) { a[i] = v; }
  uint32_t get(int i) const { return a[i]; }
};
#endif

// Stack entry of node `node` with key x: the 16 upper bits of a float at or
// below x (a negative x, which never prunes, is stored as -infinity).
template <typename R>
VGM_DEV uint32_t tess_walk_entry(int node, R x)
{
  uint32_t key = 0xff80u;  // -inf
  if (x >= R(0)) {
    const float f = float(x);
    uint32_t bits = __builtin_bit_cast(uint32_t, f);
    if (R(f) > x) bits -= 1u;  // the next float below
    key = bits >> 16;          // truncation rounds a positive float down
  }
  return (key << 16) | uint32_t(node);
}
template <typename R>
VGM_DEV R tess_walk_key(uint32_t e)
{
  return R(__builtin_bit_cast(float, e & 0xffff0000u));
}

struct TessellatedShape {
  using Data = TessData;
  static constexpr bool kStage = false;  // parameter block staged in LDS
  // waves per SIMD the transport kernel asks the compiler for: the latency
  // of the tree walks hides better at 4 waves than at the 2 its
  // unconstrained register allocation allows, although some values then
  // spill to scratch (measured: 214 -> 164 ms for 50 steps of 3M tracks)
  static constexpr int kWavesPerEU = 4;
  // safety is the distance to the surface (1-Lipschitz): carried bounds are valid
  static constexpr bool kSafetyIsDistance = true;
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

  template <typename R>
  VGM_DEV static bool facet_contains(const TessFacet &f, R px, R py, R pz)
  {
    bool inside = true;
    for (int i = 0; i < 3; ++i)
      inside &= (px - R(f.v[i][0])) * R(f.sv[i][0]) + (py - R(f.v[i][1])) * R(f.sv[i][1]) +
                    (pz - R(f.v[i][2])) * R(f.sv[i][2]) >
                R(-kTolerance);
    return inside;
  }

  // DistanceToLineSegmentSquared2, one lane with mask = true
  template <typename R>
  VGM_DEV static R segment_dist2(const double *c0, const double *c1, V3<R> p)
  {
    const R lx = R(c1[0]) - R(c0[0]), ly = R(c1[1]) - R(c0[1]), lz = R(c1[2]) - R(c0[2]);
    const R px = p.x - R(c0[0]), py = p.y - R(c0[1]), pz = p.z - R(c0[2]);
    const R dot0 = px * lx + py * ly + pz * lz;
    if (dot0 <= R(0)) return px * px + py * py + pz * pz;
    const R dot1 = lx * lx + ly * ly + lz * lz;
    R result = inf_length<R>();
    if (dot1 <= dot0) {
      const R qx = p.x - R(c1[0]), qy = p.y - R(c1[1]), qz = p.z - R(c1[2]);
      result = qx * qx + qy * qy + qz * qz;
    }
    if (result < inf_length<R>()) return result;
    const R t = dot0 / dot1;
    const R ex = (R(c0[0]) + lx * t) - p.x, ey = (R(c0[1]) + ly * t) - p.y, ez = (R(c0[2]) + lz * t) - p.z;
    return ex * ex + ey * ey + ez * ez;
  }

  // TessellatedCluster::SafetySq for one facet.  The side is a run-time
  // argument so that lanes on both sides run one copy of the walk.
  template <typename R>
  VGM_DEV static R facet_safety2(const TessFacet &f, V3<R> p, bool to_in)
  {
    R s = p.x * R(f.n[0]) + p.y * R(f.n[1]) + p.z * R(f.n[2]) + R(f.d);
    const R xx = p.x - R(f.n[0]) * s, xy = p.y - R(f.n[1]) * s, xz = p.z - R(f.n[2]) * s;
    bool within = facet_contains<R>(f, xx, xy, xz);
    // The side test's tolerance is kTolerance in FP64.  A float evaluation
    // widens it by its rounding error: a point within that error of the facet
    // plane takes the plane distance, the smaller of the two values, so the
    // float minimum never exceeds the FP64 one by more than rounding (the
    // edge distance would jump away from the plane's).
    R side_tol = R(kTolerance);
    if constexpr (sizeof(R) == sizeof(float))
      side_tol = side_tol + R(4.0e-6) * (vabs(p.x) + vabs(p.y) + vabs(p.z) + vabs(R(f.d)));
    if (to_in)
      within &= s > -side_tol;
    else
      within &= s < side_tol;
    s *= s;
    if (within) return s;
    R e = inf_length<R>();
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
      const R ei = segment_dist2<R>(f.v[i], f.v[i == 2 ? 0 : i + 1], p);
      e = (i == 0) ? ei : vmin(e, ei);
    }
    return e;
  }

  // Entry distance of the ray into child box i (-: the box contains p),
  // infinity when the ray misses it.
  template <typename R>
  VGM_DEV static R box_entry(const TessNode &nd, int i, V3<R> p, V3<R> inv)
  {
    const double *lo = nd.lo[i], *hi = nd.hi[i];
    R a = (R(lo[0]) - p.x) * inv.x, b = (R(hi[0]) - p.x) * inv.x;
    R t0 = vmin(a, b), t1 = vmax(a, b);
    a = (R(lo[1]) - p.y) * inv.y;
    b = (R(hi[1]) - p.y) * inv.y;
    t0 = vmax(t0, vmin(a, b));
    t1 = vmin(t1, vmax(a, b));
    a = (R(lo[2]) - p.z) * inv.z;
    b = (R(hi[2]) - p.z) * inv.z;
    t0 = vmax(t0, vmin(a, b));
    t1 = vmin(t1, vmax(a, b));
    return (t0 <= t1) ? t0 : inf_length<R>();
  }

  // Squared distance from p to child box i.
  template <typename R>
  VGM_DEV static R box_dist2(const TessNode &nd, int i, V3<R> p)
  {
    const double *lo = nd.lo[i], *hi = nd.hi[i];
    const R ex = vmax(vmax(R(lo[0]) - p.x, p.x - R(hi[0])), R(0));
    const R ey = vmax(vmax(R(lo[1]) - p.y, p.y - R(hi[1])), R(0));
    const R ez = vmax(vmax(R(lo[2]) - p.z, p.z - R(hi[2])), R(0));
    return ex * ex + ey * ey + ez * ez;
  }

  // A node's children in increasing key order.  A child's sort entry holds
  // the upper 16 bits of a float at or below its key, mapped so that the
  // keys order as unsigned integers (a key below 0, a box holding the point,
  // as -infinity; a box the ray misses last), and the child's slot in the
  // low bits: a sorting network of unsigned min/max orders them in four
  // registers.  The decoded key bounds the exact one from below, so it
  // prunes conservatively.
  template <typename R>
  struct SlotOrder {
    uint32_t e0, e1, e2, e3;
    VGM_DEV static uint32_t pack(R x, int slot)
    {
      uint32_t m;
      if (!(x < inf_length<R>())) {
        m = 0xffffu;  // the ray misses the box
      } else if (x < R(0)) {
        m = 0x007fu;  // -infinity
      } else {
        const float f = float(x);
        uint32_t bits = __builtin_bit_cast(uint32_t, f);
        if (R(f) > x) bits -= 1u;  // the next float below
        m = 0x8000u | (bits >> 16);
      }
      return (m << 16) | uint32_t(slot);
    }
    VGM_DEV static void cswap(uint32_t &a, uint32_t &b)
    {
      const uint32_t lo = a < b ? a : b;
      b = a < b ? b : a;
      a = lo;
    }
    VGM_DEV SlotOrder(R a, R b, R c, R d) : e0(pack(a, 0)), e1(pack(b, 1)), e2(pack(c, 2)), e3(pack(d, 3))
    {
      cswap(e0, e1);
      cswap(e2, e3);
      cswap(e0, e2);
      cswap(e1, e3);
      cswap(e1, e2);
    }
    VGM_DEV uint32_t entry(int o) const { return o == 0 ? e0 : (o == 1 ? e1 : (o == 2 ? e2 : e3)); }
    VGM_DEV static int slot(uint32_t e) { return int(e & 3u); }
    VGM_DEV static bool miss(uint32_t e) { return (e >> 16) == 0xffffu; }
    VGM_DEV static R key(uint32_t e)
    {
      const uint32_t m = e >> 16;
      if (!(m & 0x8000u)) return -inf_length<R>();
      return R(__builtin_bit_cast(float, (m & 0x7fffu) << 16));
    }
  };
  static_assert(kTessWidth == 4, "SlotOrder sorts four children");

  // Minimum entering and exiting facet hits (DistanceToCluster over all
  // facets).  The walk is bounded by the minimum of the kind selected by
  // to_in; the other kind is only needed below that bound.
  template <typename R>
  struct Hits {
    R dIn, dOut;
    int iIn, iOut;      // storage slot of the facet, -1 if none
    int oIn, oOut;      // definition index (tie-break)
  };

  template <typename R>
  VGM_DEV static void ray_hits(const Data &t, V3<R> p, V3<R> d, bool to_in, Hits<R> &h)
  {
    h.dIn = h.dOut = inf_length<R>();
    h.iIn = h.iOut = h.oIn = h.oOut = -1;
    const TessFacet *fa = t.facets();
    const TessNode *nodes = t.nodes();
    const V3<R> inv{R(1) / nonzero(d.x), R(1) / nonzero(d.y), R(1) / nonzero(d.z)};
    // nearest-first walk: at each node the leaf children are tested (nearer
    // first), then the internal ones pushed (nearer last, so it is popped
    // first); a popped node whose key exceeds the shrinking bound is dropped
    auto bound = [&]() {
      return to_in ? (h.iIn >= 0 ? h.dIn : inf_length<R>()) : (h.iOut >= 0 ? h.dOut : inf_length<R>());
    };
    TessWalkStack stack;
    int sp = 0;
    stack.set(sp++, tess_walk_entry<R>(0, R(-1)));
    while (sp) {
      const uint32_t e = stack.get(--sp);
      if (tess_walk_key<R>(e) > bound()) continue;
      const TessNode &nd = nodes[e & 0xffffu];
      const SlotOrder<R> ord(box_entry<R>(nd, 0, p, inv), box_entry<R>(nd, 1, p, inv), box_entry<R>(nd, 2, p, inv),
                             box_entry<R>(nd, 3, p, inv));
#pragma unroll 1
      for (int o = 0; o < kTessWidth; ++o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<R>::slot(se);
        const R te = SlotOrder<R>::key(se);
        if (SlotOrder<R>::miss(se) || te > bound()) break;  // so are the children after it
        if (nd.child[i] >= 0) continue;
        const int f0 = nd.first[i], f1 = f0 + nd.count[i];
        for (int k = f0; k < f1; ++k) {
          const TessFacet &f = fa[k];
          const R ndd = nonzero(d.x * R(f.n[0]) + d.y * R(f.n[1]) + d.z * R(f.n[2]));
          const R saf = p.x * R(f.n[0]) + p.y * R(f.n[1]) + p.z * R(f.n[2]) + R(f.d);
          const bool vIn = ndd < R(0) && saf > R(-kTolerance);
          const bool vOut = ndd > R(0) && saf < R(kTolerance);
          if (!vIn && !vOut) continue;
          const R dist = -saf / ndd;
          if (!facet_contains<R>(f, p.x + d.x * dist, p.y + d.y * dist, p.z + d.z * dist)) continue;
          if (vIn) {
            if (dist < h.dIn || (dist == h.dIn && f.orig < h.oIn)) {
              h.dIn = dist;
              h.iIn = k;
              h.oIn = f.orig;
            }
          } else if (dist < h.dOut || (dist == h.dOut && f.orig < h.oOut)) {
            h.dOut = dist;
            h.iOut = k;
            h.oOut = f.orig;
          }
        }
      }
#pragma unroll 1
      for (int o = kTessWidth - 1; o >= 0; --o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<R>::slot(se);
        const R te = SlotOrder<R>::key(se);
        if (nd.child[i] < 0 || SlotOrder<R>::miss(se) || te > bound()) continue;
        stack.set(sp++, tess_walk_entry<R>(nd.child[i], te));
      }
    }
  }

  template <typename R>
  VGM_DEV static R safety2(const Data &t, V3<R> p, bool to_in)
  {
    const TessFacet *fa = t.facets();
    const TessNode *nodes = t.nodes();
    R best = inf_length<R>();
    TessWalkStack stack;
    int sp = 0;
    stack.set(sp++, tess_walk_entry<R>(0, R(-1)));
    while (sp) {
      const uint32_t e = stack.get(--sp);
      if (tess_walk_key<R>(e) > best) continue;
      const TessNode &nd = nodes[e & 0xffffu];
      const SlotOrder<R> ord(box_dist2<R>(nd, 0, p), box_dist2<R>(nd, 1, p), box_dist2<R>(nd, 2, p),
                             box_dist2<R>(nd, 3, p));
#pragma unroll 1
      for (int o = 0; o < kTessWidth; ++o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<R>::slot(se);
        if (SlotOrder<R>::key(se) > best) break;  // so are the children after it
        if (nd.child[i] >= 0) continue;
        const int f0 = nd.first[i], f1 = f0 + nd.count[i];
        for (int k = f0; k < f1; ++k) {
          const R s = facet_safety2<R>(fa[k], p, to_in);
          if (s < best) best = s;
        }
      }
#pragma unroll 1
      for (int o = kTessWidth - 1; o >= 0; --o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<R>::slot(se);
        const R di = SlotOrder<R>::key(se);
        if (nd.child[i] < 0 || di > best) continue;
        stack.set(sp++, tess_walk_entry<R>(nd.child[i], di));
      }
    }
    return best;
  }

  // Storage slot of the facet nearest to p by TessellatedCluster::SafetySq
  // with ToIn = false; equal distances go to the facet defined first.  The
  // same walk as safety2, which prunes only boxes strictly beyond the best
  // distance, so every facet that ties with it is visited.
  VGM_DEV static int nearest_facet(const Data &t, V3<double> p)
  {
    const TessFacet *fa = t.facets();
    const TessNode *nodes = t.nodes();
    double best = kInfLength;
    int slot = -1, orig = 0;
    TessWalkStack stack;
    int sp = 0;
    stack.set(sp++, tess_walk_entry<double>(0, -1.0));
    while (sp) {
      const uint32_t e = stack.get(--sp);
      if (tess_walk_key<double>(e) > best) continue;
      const TessNode &nd = nodes[e & 0xffffu];
      const SlotOrder<double> ord(box_dist2<double>(nd, 0, p), box_dist2<double>(nd, 1, p),
                                  box_dist2<double>(nd, 2, p), box_dist2<double>(nd, 3, p));
#pragma unroll 1
      for (int o = 0; o < kTessWidth; ++o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<double>::slot(se);
        if (SlotOrder<double>::key(se) > best) break;  // so are the children after it
        if (nd.child[i] >= 0) continue;
        const int f0 = nd.first[i], f1 = f0 + nd.count[i];
        for (int k = f0; k < f1; ++k) {
          const double s = facet_safety2<double>(fa[k], p, false);
          if (s < best || (s == best && fa[k].orig < orig)) {
            best = s;
            slot = k;
            orig = fa[k].orig;
          }
        }
      }
#pragma unroll 1
      for (int o = kTessWidth - 1; o >= 0; --o) {
        const uint32_t se = ord.entry(o);
        const int i = SlotOrder<double>::slot(se);
        const double di = SlotOrder<double>::key(se);
        if (nd.child[i] < 0 || di > best) continue;
        stack.set(sp++, tess_walk_entry<double>(nd.child[i], di));
      }
    }
    return slot;
  }

  // UnplacedTessellated::Normal (source/UnplacedTessellated.cpp:83-89,
  // TessellatedImplementation.h:163-172): the normal of the nearest facet,
  // valid everywhere.
  VGM_DEV static bool normal(const Data &t, V3<double> p, V3<double> &n)
  {
    const int k = nearest_facet(t, p);
    const TessFacet &f = t.facets()[k];
    n = V3<double>{f.n[0], f.n[1], f.n[2]};
    return true;
  }

  // BoxImplementation::IntersectCachedKernel2 over the extent, t0 = -kTolerance,
  // t1 = kInfLength (BoxImplementation.h:326-368)
  template <typename R>
  VGM_DEV static R bbox_intersect(const Data &t, V3<R> p, V3<R> d)
  {
    const R ix = R(1.0) / nonzero(d.x), iy = R(1.0) / nonzero(d.y), iz = R(1.0) / nonzero(d.z);
    const bool sx = ix < R(0), sy = iy < R(0), sz = iz < R(0);
    R tmin = (R(sx ? t.mx[0] : t.mn[0]) - p.x) * ix;
    const R tymax = (R(sy ? t.mn[1] : t.mx[1]) - p.y) * iy;
    if (tmin > tymax) return inf_length<R>();
    R tmax = (R(sx ? t.mn[0] : t.mx[0]) - p.x) * ix;
    const R tymin = (R(sy ? t.mx[1] : t.mn[1]) - p.y) * iy;
    if (tymin > tmax) return inf_length<R>();
    tmin = vmax(tmin, tymin);
    tmax = vmin(tmax, tymax);
    const R tzmin = (R(sz ? t.mx[2] : t.mn[2]) - p.z) * iz;
    const R tzmax = (R(sz ? t.mn[2] : t.mx[2]) - p.z) * iz;
    if ((tmin > tzmax) || (tzmin > tmax)) return inf_length<R>();
    tmin = vmax(tmin, tzmin);
    tmax = vmin(tmax, tzmax);
    if (!((tmin <= inf_length<R>() + R(kTolerance)) && (tmax > R(-kTolerance) - R(kTolerance))))
      return inf_length<R>();
    return tmin;
  }

  // DistanceToSolid<ToIn> (TessellatedImplementation.h:176-279), the side a
  // run-time argument
  template <typename R>
  VGM_DEV static R distance_to_solid(const Data &t, V3<R> p, V3<R> d, R step_max, bool ToIn, Hits<R> &h,
                                     int &isurf)
  {
    isurf = -1;
    R distance;
    if (ToIn) {
      distance = bbox_intersect<R>(t, p, d);
      if (distance >= step_max) {
        h.iIn = h.iOut = -1;
        return distance;
      }
    }
    ray_hits<R>(t, p, d, ToIn, h);
    const TessFacet *fa = t.facets();
    distance = inf_length<R>();
    if (ToIn) {
      if (h.iIn >= 0) {
        isurf = h.iIn;
        distance = h.dIn;
      }
      if (h.iIn < 0) {
        if (h.iOut >= 0 && h.dOut * (d.x * R(fa[h.iOut].n[0]) + d.y * R(fa[h.iOut].n[1]) + d.z * R(fa[h.iOut].n[2])) >
                               R(kTolerance))
          distance = R(-1);
      } else {
        if (h.iOut >= 0 && h.dOut > R(kTolerance) && h.dOut < h.dIn) distance = R(-1);
      }
    } else {
      if (h.iOut >= 0) {
        isurf = h.iOut;
        distance = h.dOut;
      }
      if (h.iOut < 0) {
        distance = R(-1);
      } else if (h.iIn >= 0 && h.dIn < h.dOut &&
                 h.dIn * (d.x * R(fa[h.iIn].n[0]) + d.y * R(fa[h.iIn].n[1]) + d.z * R(fa[h.iIn].n[2])) <
                     R(-kTolerance)) {
        distance = R(-1);
        isurf = -1;
      }
    }
    return distance;
  }

  template <typename R>
  VGM_DEV static V3<R> test_dir(const Data &t)
  {
    return V3<R>{R(t.test_dir[0]), R(t.test_dir[1]), R(t.test_dir[2])};
  }

  template <typename R>
  VGM_DEV static bool contains(const Data &t, V3<R> p)
  {
    Hits<R> h;
    int isurf;
    distance_to_solid<R>(t, p, test_dir<R>(t), inf_length<R>(), false, h, isurf);
    return isurf >= 0;
  }

  template <typename R>
  VGM_DEV static int inside(const Data &t, V3<R> p)
  {
    Hits<R> h;
    int iOut;
    const V3<R> td = test_dir<R>(t);
    const R dOut = distance_to_solid<R>(t, p, td, inf_length<R>(), false, h, iOut);
    if (iOut < 0) return kOutsideV;
    const TessFacet *fa = t.facets();
    if (dOut < R(0) || dOut * (td.x * R(fa[iOut].n[0]) + td.y * R(fa[iOut].n[1]) + td.z * R(fa[iOut].n[2])) <
                           R(kTolerance))
      return kSurfaceV;
    if (h.iIn < 0 || dOut < h.dIn) return kInsideV;
    if (h.dIn < R(0) ||
        h.dIn * (td.x * R(fa[h.iIn].n[0]) + td.y * R(fa[h.iIn].n[1]) + td.z * R(fa[h.iIn].n[2])) > R(-kTolerance))
      return kSurfaceV;
    return kOutsideV;
  }

  template <typename R>
  VGM_DEV static R distance_to_in(const Data &t, V3<R> p, V3<R> d, R step_max)
  {
    Hits<R> h;
    int isurf;
    return distance_to_solid<R>(t, p, d, step_max, true, h, isurf);
  }

  template <typename R>
  VGM_DEV static R distance_to_out(const Data &t, V3<R> p, V3<R> d, R step_max)
  {
    Hits<R> h;
    int isurf;
    return distance_to_solid<R>(t, p, d, step_max, false, h, isurf);
  }

  // DistanceToOut (inside) or DistanceToIn (outside) in one walk, so the
  // lanes of a wave on both sides of the surface share it.
  template <typename R>
  VGM_DEV static R boundary_distance(const Data &t, V3<R> p, V3<R> d, bool inside)
  {
    Hits<R> h;
    int isurf;
    return distance_to_solid<R>(t, p, d, inf_length<R>(), !inside, h, isurf);
  }

  // Distance from p to the surface, without the side check of
  // SafetyToIn/Out: SafetyToOut(p) = inside ? surface_distance(p, true) : -1,
  // SafetyToIn(p) = inside ? -1 : surface_distance(p, false).
  template <typename R>
  VGM_DEV static R surface_distance(const Data &t, V3<R> p, bool from_inside)
  {
    return vsqrt(safety2<R>(t, p, !from_inside));
  }

  template <typename R>
  VGM_DEV static R safety_to_in(const Data &t, V3<R> p)
  {
    if (contains<R>(t, p)) return R(-1);
    return vsqrt(safety2<R>(t, p, true));
  }

  template <typename R>
  VGM_DEV static R safety_to_out(const Data &t, V3<R> p)
  {
    if (!contains<R>(t, p)) return R(-1);
    return vsqrt(safety2<R>(t, p, false));
  }
};

}  // namespace vgm
