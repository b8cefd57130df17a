// One shape family's transport kernels: built once per family with
// -DVGM_FAMILY_SHAPE=<shape struct> (see the Makefile), so the families'
// kernels compile in parallel.
#include "transport_kernels.hpp"

#ifndef VGM_FAMILY_SHAPE
#error "VGM_FAMILY_SHAPE must name the shape family"
#endif

namespace vgm {
template hipError_t launch_family_of<VGM_FAMILY_SHAPE>(hipStream_t, const TransportArgs &, const FamilyLaunch &, int);
}  // namespace vgm
