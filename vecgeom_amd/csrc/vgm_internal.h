// Internal interfaces between the C-ABI layer (capi.cpp) and the kernel
// launchers.  Not installed; callers use include/vgm/vgm.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string>
#include "vgm_types.h"

namespace vgm {

// One SOA3D batch (SOA3D<double>, VecGeom/base/SOA3D.h:36-41) plus outputs.
struct SoaArgs {
  const double *px, *py, *pz;
  const double *dx, *dy, *dz;
  const double *step;
  size_t n;
  void *out;
};

enum Precision : int32_t { kFP64 = 0, kMixed = 1 };

// Launch one placed-volume SOA query on `stream`.  `rec` describes the shape,
// `params` is the device address of its parameter block, `use_transform`
// selects master-frame input (Contains/Inside/DistanceToIn/SafetyToIn and
// PlacedDistanceToOut) versus local-frame input.
hipError_t launch_soa(hipStream_t stream, int method, int precision, const ShapeRec &rec, const void *params,
                      const Transform &tr, bool use_transform, const SoaArgs &a);

// VPlacedVolume::Normal over a batch of master-frame points.
struct NormalArgs {
  const double *px, *py, *pz;
  size_t n;
  double *nx, *ny, *nz;
  uint8_t *valid;
};
// True when the family has a Normal kernel (every family but the tessellated solid).
bool family_has_normal(int type);
hipError_t launch_normal(hipStream_t stream, const ShapeRec &rec, const void *params, const Transform &tr,
                         const NormalArgs &a);

void set_error(const std::string &msg);

}  // namespace vgm
