// Polycone parameter block (PolyconeStruct, VecGeom/volumes/PolyconeStruct.h):
// a header followed in the pool by zs[nsec+1], shift[nsec] and nsec cone
// sections (ConeData), all 8-byte aligned.
#pragma once
#include <stdint.h>
#include "cone_data.h"

namespace vgm {

struct PolyconeData {
  int32_t nsec;
  int32_t pad;

  __host__ __device__ const double *zs() const { return reinterpret_cast<const double *>(this + 1); }
  __host__ __device__ const double *shift() const { return zs() + nsec + 1; }
  __host__ __device__ const ConeData *sec() const { return reinterpret_cast<const ConeData *>(shift() + nsec); }
  __host__ __device__ static size_t bytes(int nsec)
  {
    return sizeof(PolyconeData) + sizeof(double) * (2 * nsec + 1) + sizeof(ConeData) * nsec;
  }
};

}  // namespace vgm
