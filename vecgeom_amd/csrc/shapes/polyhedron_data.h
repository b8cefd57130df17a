// Polyhedron parameter block (PolyhedronStruct, VecGeom/volumes/PolyhedronStruct.h:55-110).
// A header followed in the pool by
//   zplanes[nz], rmin[nz], rmax[nz], samez[nz] (0/1 as double),
//   phisec[(nsides + 1) * 3]        phi-section plane normals (fPhiSections),
//   segs[nseg]                      per z segment: first quad and count of the
//                                   outer, inner and phi shells,
//   quads[]                         32 doubles each (plane, four side planes,
//                                   four corners), one cache line pair.
// All members are 8-byte aligned; the block is a whole number of doubles.
#pragma once
#include <stdint.h>
#include "tube_data.h"

namespace vgm {

struct PolyhedronQuad {
  double n[3], d;     // plane: n.p + d, normal points out of the shell
  double sn[4][3];    // side planes through each edge, pointing inwards
  double sd[4];
  double c[4][3];     // corners (for the exact point-to-quad distance)
};
static_assert(sizeof(PolyhedronQuad) == 32 * sizeof(double), "quad is 32 doubles");

struct PolyhedronSeg {
  int32_t outer, n_outer;
  int32_t inner, n_inner;
  int32_t phi, n_phi;
  int32_t pad0, pad1;
};
static_assert(sizeof(PolyhedronSeg) == 4 * sizeof(double), "segment record is 4 doubles");

struct PolyhedronData {
  int32_t nseg, nsides;
  int32_t phi_cutout;    // fHasPhiCutout: phiDelta < 2 pi
  int32_t large_cutout;  // fHasLargePhiCutout: phiDelta < pi
  double phi_start, phi_delta;
  double bt_offset;      // fBoundingTubeOffset
  double half_len;       // 0.5 * (zplanes[nseg] - zplanes[0])
  // fPhiWedge(phiDelta, phiStart)
  double w_a1x, w_a1y, w_a2x, w_a2y, w_n1x, w_n1y, w_n2x, w_n2y;
  int32_t w_lt_pi, nquads;
  TubeData btube;        // fBoundingTube

  __host__ __device__ const double *zplanes() const { return reinterpret_cast<const double *>(this + 1); }
  __host__ __device__ const double *rmin() const { return zplanes() + nseg + 1; }
  __host__ __device__ const double *rmax() const { return rmin() + nseg + 1; }
  __host__ __device__ const double *samez() const { return rmax() + nseg + 1; }
  __host__ __device__ const double *phisec() const { return samez() + nseg + 1; }
  __host__ __device__ const PolyhedronSeg *segs() const
  {
    return reinterpret_cast<const PolyhedronSeg *>(phisec() + 3 * (nsides + 1));
  }
  __host__ __device__ const PolyhedronQuad *quads() const
  {
    return reinterpret_cast<const PolyhedronQuad *>(segs() + nseg);
  }
  __host__ __device__ static size_t bytes(int nseg, int nsides, int nquads)
  {
    return sizeof(PolyhedronData) + sizeof(double) * (4 * (nseg + 1) + 3 * (nsides + 1)) +
           sizeof(PolyhedronSeg) * nseg + sizeof(PolyhedronQuad) * nquads;
  }
};
static_assert(sizeof(PolyhedronData) % sizeof(double) == 0, "header is whole doubles");

}  // namespace vgm
