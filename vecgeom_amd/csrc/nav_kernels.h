// Interface between the navigation C-ABI (nav_capi.cpp) and the navigation
// kernels (nav_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "navigation.hpp"

namespace vgm {

enum NavKind : int32_t {
  kNavLocate = 0,
  kNavStep = 1,
  kNavTransport = 2,
  kNavSafety = 3,
  kNavApproach = 4,           // ApproachNextDaughter per track
  kNavTransportLimited = 5,   // transport under a physics step, safety first
  kNavStepMove = 6,           // one straight-line transport step (graph-captured loops)
  kNavLevelLocate = 7,        // LevelLocate in the volume of a state (with a direction: the entering rule)
};

struct NavArgs {
  GeomView g;
  double *x, *y, *z;
  const double *dx, *dy, *dz;
  const double *step_limit;  // per track, may be null (unlimited)
  uint32_t *state, *last;
  uint8_t *boundary;
  double *step;    // step: output step; safety: output safety
  const uint8_t *calcsafety;  // step with safeties: per track, may be null (every track)
  double *safety;             // step with safeties: output safety (null: plain step)
  int32_t *steps;  // transport: steps taken (accumulated)
  int32_t *geo_steps;  // limited transport: steps that needed the boundary distance (accumulated)
  uint32_t *child;              // level locate: the daughter's state, 0 if none (excluded one: `last`)
  double *ox, *oy, *oz;         // level locate: the point in the daughter's frame
  int32_t reuse_safety;  // limited transport: carry the safety left after a safety-only move
  int32_t no_relocate;   // step: ComputeStep, the out state without relocation
  unsigned long long *queue;  // transports: work-queue counter (device, reset by the launch)
  int32_t num_cus;            // transports: compute units of the device (persistent grid size)
  size_t n;
  int32_t nsteps;
  uint32_t shape_mask;  // families used by the navigation tree (bit per ShapeType)
};

hipError_t launch_nav(hipStream_t stream, int kind, const NavArgs &a);

}  // namespace vgm
