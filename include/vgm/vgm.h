/* vgm — MI355X (gfx950) geometry hot path behind a C ABI.
 *
 * Every entry point returns 0 on success and a negative status otherwise;
 * vgm_last_error() returns a thread-local description of the last failure.
 * No C++ exceptions and no library types cross this boundary: plain pointers,
 * sizes and integer handles only.
 *
 * Units are millimetres.  Result conventions (sentinels, frames) are those of
 * VecGeom: a miss is kInfLength = DBL_MAX, a query from the wrong side returns
 * a negative value, Inside returns 1 (inside), 2 (surface) or 3 (outside)
 * (VecGeom/base/Global.h:71-88).
 */
#ifndef VGM_VGM_H
#define VGM_VGM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vgm_ctx vgm_ctx;

/* Query flags. */
#define VGM_HOST_PTRS 0u   /* inputs/outputs are host memory; call is synchronous */
#define VGM_DEVICE_PTRS 1u /* inputs/outputs are device memory on the context's device */
#define VGM_ASYNC 2u       /* with VGM_DEVICE_PTRS: return after enqueueing on the context stream */
#define VGM_MIXED 4u       /* FP32 bulk arithmetic, FP64 refinement and tolerance decisions */
#define VGM_TIMING 8u      /* vgm_batch_transport: time each shape family's kernel (vgm_batch_family_ms) */
#define VGM_LEARNED 16u    /* vgm_batch_transport: learned safety bound before the exact safety */
#define VGM_AUDIT 32u      /* with VGM_LEARNED: check every served bound against the exact safety */
#define VGM_DISTANCE_FIRST 64u /* transport: boundary distance before the safety (same decisions as FP64) */
/* Shape types, as indices of the per-family arrays of the batch calls:
 * 0 box, 1 tube, 2 cone, 3 trd, 4 trapezoid, 5 polycone, 6 polyhedron,
 * 7 tessellated, 8 orb, 9 cut tube, 10 boolean, 11 sphere, 12 torus. */
#define VGM_NUM_SHAPE_TYPES 13
#define VGM_TUNED 128u     /* transport: each family in the mode vgm_batch_autotune chose for it */
#define VGM_CULL 256u /* transport: distance first, and outside moves clear of the solid's bounding box skip \
                          the queries (same decisions as FP64) */
#define VGM_GRAPH 512u /* transports, with VGM_DEVICE_PTRS: the step loop as a captured HIP graph of one-step \
                          kernels, replayed while the arguments stay the same */
#define VGM_STEP_KERNELS 1024u /* vgm_nav_transport: the step loop as one-step kernels launched one by one \
                                  (what VGM_GRAPH captures) */

/* Status codes. */
#define VGM_OK 0
#define VGM_EINVAL -1
#define VGM_EHIP -2
#define VGM_ESTATE -3
#define VGM_ENOMEM -4

const char *vgm_last_error(void);

/* ---- context ---------------------------------------------------------------
 * A context owns one geometry and one HIP stream on one device.  `stream` may
 * be NULL (the context creates its own) or an existing hipStream_t, in which
 * case all work is enqueued there.  Replaces the process-wide singletons
 * GeoManager::Instance() / CudaManager::Instance()
 * (VecGeom/management/GeoManager.h, VecGeom/management/CudaManager.h). */
int vgm_ctx_create(int device, void *stream, vgm_ctx **out);
int vgm_ctx_destroy(vgm_ctx *ctx);
int vgm_ctx_synchronize(vgm_ctx *ctx);
/* Returns the hipStream_t the context launches on. */
void *vgm_ctx_stream(vgm_ctx *ctx);

/* ---- unplaced shapes ------------------------------------------------------
 * Constructor arguments are those of the reference's Unplaced* classes; the
 * library derives the cached quantities itself.  On success *shape is the id. */

/* UnplacedBox(dx, dy, dz): half-lengths (VecGeom/volumes/UnplacedBox.h). */
int vgm_box_create(vgm_ctx *ctx, double dx, double dy, double dz, int *shape);
/* UnplacedOrb(r) (VecGeom/volumes/UnplacedOrb.h; kernels
 * VecGeom/volumes/kernel/OrbImplementation.h:64-225): a full sphere. */
int vgm_orb_create(vgm_ctx *ctx, double r, int *shape);
/* UnplacedSphere(rmin, rmax, sphi, dphi, stheta, dtheta)
 * (VecGeom/volumes/UnplacedSphere.h:40-41, SphereStruct.h; kernels
 * VecGeom/volumes/kernel/SphereImplementation.h with the theta cone of
 * ThetaCone_Evolution.h): a spherical shell cut in phi and theta. */
int vgm_sphere_create(vgm_ctx *ctx, double rmin, double rmax, double sphi, double dphi, double stheta, double dtheta,
                      int *shape);
/* UnplacedTorus2(rmin, rmax, rtor, sphi, dphi) (VecGeom/volumes/UnplacedTorus2.h:64-75,
 * TorusStruct2.h; kernels VecGeom/volumes/kernel/TorusImplementation2.h,
 * Normal source/UnplacedTorus2.cpp:201-285): a torus of tube radii
 * rmin..rmax around a circle of radius rtor, cut in phi.  FP64; its
 * distances solve a quartic (cbrt, pow, cos, atan2, then Newton) and agree
 * with a host evaluation to the refinement's tolerance. */
int vgm_torus_create(vgm_ctx *ctx, double rmin, double rmax, double rtor, double sphi, double dphi, int *shape);
/* UnplacedBooleanVolume<op>(left placed, right placed)
 * (VecGeom/volumes/UnplacedBooleanVolume.h, BooleanStruct.h:8-31; kernels
 * VecGeom/volumes/kernel/BooleanImplementation.h:39-290,
 * BooleanIntersectionImplementation.h and BooleanUnionImplementation.h):
 * op 0 = kUnion, 1 = kIntersection, 2 = kSubtraction (left minus right),
 * the reference's BooleanOperation values.  The operands are
 * existing primitive shapes (not booleans, tessellated solids or tori:
 * VGM_EINVAL) placed by
 * left_tr / right_tr (12 doubles: translation, row-major rotation; NULL =
 * identity).  The operand parameters are copied at creation. */
int vgm_boolean_create(vgm_ctx *ctx, int op, int left, const double *left_tr, int right, const double *right_tr,
                       int *shape);
/* UnplacedCutTube(rmin, rmax, z, sphi, dphi, bx, by, bz, tx, ty, tz)
 * (VecGeom/volumes/UnplacedCutTube.h:54-70; kernels
 * VecGeom/volumes/kernel/CutTubeImplementation.h:106-321): a tube of
 * half-length z cut by planes through (0,0,-z) and (0,0,z) with outward
 * normals (bx,by,bz), bz < 0, and (tx,ty,tz), tz > 0. */
int vgm_cut_tube_create(vgm_ctx *ctx, double rmin, double rmax, double z, double sphi, double dphi, double bx,
                        double by, double bz, double tx, double ty, double tz, int *shape);

/* UnplacedTube(rmin, rmax, dz, sphi, dphi): radii, half-length, start and
 * delta phi in radians (VecGeom/volumes/UnplacedTube.h:29-35). */
int vgm_tube_create(vgm_ctx *ctx, double rmin, double rmax, double dz, double sphi, double dphi, int *shape);

/* UnplacedCone(rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi): radii at -dz and
 * +dz (VecGeom/volumes/UnplacedCone.h:51-57). */
int vgm_cone_create(vgm_ctx *ctx, double rmin1, double rmax1, double rmin2, double rmax2, double dz, double sphi,
                    double dphi, int *shape);

/* UnplacedTrd(x1, x2, y1, y2, z): half-lengths in x and y at -z and +z
 * (VecGeom/volumes/UnplacedTrd.h). */
int vgm_trd_create(vgm_ctx *ctx, double x1, double x2, double y1, double y2, double z, int *shape);

/* UnplacedTrapezoid(dz, theta, phi, dy1, dx1, dx2, alpha1, dy2, dx3, dx4,
 * alpha2), angles in radians (VecGeom/volumes/UnplacedTrapezoid.h:56-66). */
int vgm_trapezoid_create(vgm_ctx *ctx, double dz, double theta, double phi, double dy1, double dx1, double dx2,
                         double alpha1, double dy2, double dx3, double dx4, double alpha2, int *shape);

/* UnplacedPolycone(phistart, deltaphi, Nz, z, rmin, rmax): nz planes with
 * their inner and outer radii (VecGeom/volumes/UnplacedPolycone.h:79-89). */
int vgm_polycone_create(vgm_ctx *ctx, double sphi, double dphi, int nz, const double *z, const double *rmin,
                        const double *rmax, int *shape);

/* UnplacedPolyhedron(phiStart, phiDelta, sideCount, zPlaneCount, zPlanes,
 * rMin, rMax): radii are distances to the sides, not to the corners; z planes
 * must be non-decreasing (VecGeom/volumes/UnplacedPolyhedron.h:119-121). */
int vgm_polyhedron_create(vgm_ctx *ctx, double sphi, double dphi, int nsides, int nz, const double *z,
                          const double *rmin, const double *rmax, int *shape);

/* UnplacedTessellated with triangular facets only
 * (TessellatedStruct::AddTriangularFacet, VecGeom/volumes/TessellatedStruct.h:596-610,
 * then Close()): triangles holds nfacets x 3 vertices x 3 coordinates, each
 * facet anti-clockwise seen from outside; degenerate facets are dropped.
 * test_dir is the direction of the Contains/Inside test ray (fTestDir, drawn
 * at random by the reference); NULL selects a fixed default. */
int vgm_tessellated_create(vgm_ctx *ctx, int nfacets, const double *triangles, const double *test_dir, int *shape);

/* Axis-aligned local bounding box of a shape (VUnplacedVolume::Extent,
 * VecGeom/volumes/UnplacedVolume.h). */
int vgm_shape_extent(vgm_ctx *ctx, int shape, double *min3, double *max3);

/* ---- volumes --------------------------------------------------------------
 * LogicalVolume(shape) (VecGeom/volumes/LogicalVolume.h) and placed volumes
 * carrying a Transformation3D given as 12 doubles: translation[3] followed by
 * the rotation matrix[9] in the reference's storage order
 * (VecGeom/base/Transformation3D.h:54-63).  `transform` may be NULL for the
 * identity.  Adding a daughter corresponds to LogicalVolume::PlaceDaughter. */
int vgm_lv_create(vgm_ctx *ctx, int shape, int *lv);
int vgm_pv_create(vgm_ctx *ctx, int lv, const double *transform, int *pv);
int vgm_lv_add_daughter(vgm_ctx *ctx, int mother_lv, int pv);

/* ---- SOA3D queries on one placed volume -----------------------------------
 * Each replaces the matching VPlacedVolume SOA3D virtual
 * (VecGeom/volumes/PlacedVolume.h:273,299,332,372,405,421).  Points and
 * directions are three separate coordinate arrays of length n (SOA3D<double>,
 * VecGeom/base/SOA3D.h:36-41); directions must be unit vectors.
 *
 * Frames follow the reference: Contains, Inside, DistanceToIn and SafetyToIn
 * take master-frame points and apply the placement
 * (VecGeom/volumes/SpecializedPlacedVolImplHelper.h:186-257); DistanceToOut
 * and SafetyToOut take local-frame points
 * (VecGeom/volumes/PlacedVolImplHelper.h:76-124); PlacedDistanceToOut takes
 * master-frame input (SpecializedPlacedVolImplHelper.h:148-164).
 * `step_max` may be NULL (kInfLength everywhere). */
int vgm_contains(vgm_ctx *ctx, int pv, const double *x, const double *y, const double *z, size_t n, uint8_t *out,
                 unsigned flags);
int vgm_inside(vgm_ctx *ctx, int pv, const double *x, const double *y, const double *z, size_t n, int32_t *out,
               unsigned flags);
int vgm_distance_to_in(vgm_ctx *ctx, int pv, const double *px, const double *py, const double *pz,
                       const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                       double *out, unsigned flags);
int vgm_distance_to_out(vgm_ctx *ctx, int pv, const double *px, const double *py, const double *pz,
                        const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                        double *out, unsigned flags);
int vgm_placed_distance_to_out(vgm_ctx *ctx, int pv, const double *px, const double *py, const double *pz,
                               const double *dx, const double *dy, const double *dz, const double *step_max,
                               size_t n, double *out, unsigned flags);
int vgm_safety_to_in(vgm_ctx *ctx, int pv, const double *x, const double *y, const double *z, size_t n,
                     double *out, unsigned flags);
int vgm_safety_to_out(vgm_ctx *ctx, int pv, const double *x, const double *y, const double *z, size_t n,
                      double *out, unsigned flags);

/* VPlacedVolume::Normal (VecGeom/volumes/PlacedVolume.h:437-439,
 * source/PlacedVolume.cpp:126-137) over n master-frame points: the unit
 * normal of the surface the point is on, or of the nearest one (master
 * frame), and valid[i] = 1 when the point is on the surface (the families'
 * own rules: the tessellated solid's nearest-facet normal is always valid,
 * a torus point off every surface gets a zero normal).  The reference has
 * only the scalar form.  Every family; FP64 only (VGM_MIXED is refused). */
int vgm_normal(vgm_ctx *ctx, int pv, const double *x, const double *y, const double *z, size_t n, double *nx,
               double *ny, double *nz, uint8_t *valid, unsigned flags);

/* ---- allocation-free SOA3D query --------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  It is the host-pointer form of the seven queries above
 * (method = one of VGM_Q_*; the same frames, sentinels and results as the
 * matching vgm_* call with VGM_HOST_PTRS, VGM_MIXED allowed), but it stages
 * the inputs and outputs through a fixed device array of
 * VGM_NOALLOC_SCRATCH_BYTES that the library holds per device until the
 * process ends, and calls no hipMalloc/hipFree.  Each coordinate, direction
 * and step_max array and the output take n elements rounded up to 256 bytes;
 * a call that needs more than the array fails with VGM_EINVAL before any
 * work.  The geometry must already be on the device: after creating or
 * changing volumes, run one of the allocating calls first (otherwise
 * VGM_ESTATE).  Calls on one device share the array and overlapping calls
 * race: serialise them (one host thread, or a lock, per device).
 * VGM_DEVICE_PTRS is rejected: device arrays need no staging. */
#define VGM_Q_CONTAINS 0
#define VGM_Q_INSIDE 1
#define VGM_Q_DISTANCE_TO_IN 2
#define VGM_Q_DISTANCE_TO_OUT 3
#define VGM_Q_PLACED_DISTANCE_TO_OUT 4
#define VGM_Q_SAFETY_TO_IN 5
#define VGM_Q_SAFETY_TO_OUT 6
#define VGM_NOALLOC_SCRATCH_BYTES (32u << 20)
int vgm_query_noalloc(vgm_ctx *ctx, int pv, int method, const double *px, const double *py, const double *pz,
                      const double *dx, const double *dy, const double *dz, const double *step_max, size_t n,
                      void *out, unsigned flags);

/* ---- navigation -------------------------------------------------------------
 * The drop-in for VNavigator::ComputeStepAndPropagatedState
 * (VecGeom/navigation/VNavigator.h:51-55, BVHNavigator over every logical
 * volume with daughters), GlobalLocator::LocateGlobalPoint
 * (VecGeom/navigation/GlobalLocator.h:35) and VSafetyEstimator::ComputeSafety
 * (VecGeom/navigation/VSafetyEstimator.h:45) on SOA3D batches of global
 * points and directions.
 *
 * A navigation state is a uint32 index (NavStateIndex): the node of the
 * placement tree reached by a depth-first walk from the world, 1 = world,
 * 0 = outside.  A track carries (state, last_exited, boundary) like
 * NavStateIndex's fNavInd / fLastExited / fOnBoundary.  vgm_nav_node decodes
 * a state into its placed volume, level and mother state.
 *
 * vgm_nav_close builds the tables for the tree under world_pv (ABBoxes with
 * a 1e-3 mm margin, the reference's surface-area BVH per logical volume,
 * cached global matrices); call it again after changing the geometry. */
int vgm_nav_close(vgm_ctx *ctx, int world_pv, int *nnodes);
int vgm_nav_node(vgm_ctx *ctx, uint32_t state, int *pv, int *level, int *mother);
/* NavStatePath <-> state index (NavIndexTable::ValidateState,
 * source/NavIndexTable.cpp:86-140, and NavStatePath::At).  A path is the
 * placed-volume ids from the world (pvs[0]) down to the top volume
 * (pvs[depth - 1]); each must be a daughter of the previous one's logical
 * volume.  vgm_nav_path_to_state fails (VGM_EINVAL) on a path that is not in
 * the placement tree, where ValidateState throws.  vgm_nav_state_to_path
 * writes depth = level + 1 entries (at most cap). */
int vgm_nav_path_to_state(vgm_ctx *ctx, const int *pvs, int depth, uint32_t *state);
int vgm_nav_state_to_path(vgm_ctx *ctx, uint32_t state, int *pvs, int cap, int *depth);
/* state[i] = located state of global point i (0 if outside the world). */
int vgm_nav_locate(vgm_ctx *ctx, const double *x, const double *y, const double *z, size_t n, uint32_t *state,
                   unsigned flags);
/* One ComputeStepAndPropagatedState per track: step[i] = geometric step
 * (bounded by step_limit[i]; step_limit may be NULL for no limit); state,
 * last and boundary are replaced by the propagated state. */
int vgm_nav_step(vgm_ctx *ctx, const double *x, const double *y, const double *z, const double *dx, const double *dy,
                 const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                 uint8_t *boundary, double *step, unsigned flags);
/* The same call under the name the north-star transport codes use for it
 * ("ComputeStepAndNextVolume": the step and, in state, the next volume);
 * identical to vgm_nav_step. */
int vgm_nav_compute_step_and_next_volume(vgm_ctx *ctx, const double *x, const double *y, const double *z,
                                         const double *dx, const double *dy, const double *dz,
                                         const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                                         uint8_t *boundary, double *step, unsigned flags);
/* VNavigator::ComputeStep (VecGeom/navigation/VNavigator.h:494-540): the
 * step of vgm_nav_step without the relocation after the boundary.  The out
 * state is PrepareOutState's (:192-235): on a daughter hit the daughter's
 * node with boundary = 1 and last = the current state; on leaving the
 * mother the current state with boundary = 1 (the caller relocates); a
 * physics-limited step keeps the state with boundary = 0; a point whose
 * mother DistanceToOut is kInfLength goes to the mother's mother with step
 * kTolerance. */
int vgm_nav_compute_step(vgm_ctx *ctx, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, const double *step_limit, size_t n, uint32_t *state,
                         uint32_t *last, uint8_t *boundary, double *step, unsigned flags);
/* ComputeStepsAndSafetiesAndPropagatedStates (VecGeom/navigation/
 * VNavigator.h:136-148; per track ComputeStepAndSafetyAndPropagatedState,
 * :436-492 with the safety first): as vgm_nav_step, and before the step
 * safety[i] = the safety of the point in its state where calcsafety[i] != 0
 * (calcsafety may be NULL: every track), else 0; 0 for state 0. */
int vgm_nav_step_safety(vgm_ctx *ctx, const double *x, const double *y, const double *z, const double *dx,
                        const double *dy, const double *dz, const double *step_limit, size_t n, uint32_t *state,
                        uint32_t *last, uint8_t *boundary, double *step, const uint8_t *calcsafety,
                        double *safety, unsigned flags);
/* nsteps straight-line steps per track (no physics limit): each step moves
 * the global point by the geometric step and takes the propagated state; a
 * track that leaves the world stops.  steps[i] is incremented by the number
 * of steps taken.  By default one persistent kernel runs all steps over a
 * work queue of tracks; with VGM_GRAPH (device pointers only) the loop is
 * nsteps launches of a one-step kernel, captured once into a HIP graph and
 * replayed by later calls with the same arrays, sizes and geometry;
 * VGM_STEP_KERNELS launches those kernels one by one.  All three give the
 * same results. */
int vgm_nav_transport(vgm_ctx *ctx, double *x, double *y, double *z, const double *dx, const double *dy,
                      const double *dz, size_t n, uint32_t *state, uint32_t *last, uint8_t *boundary, int nsteps,
                      int32_t *steps, unsigned flags);
/* nsteps steps per track under a physics step limit step_limit[i] > 0: each
 * step is ComputeStepAndSafetyAndPropagatedState with calcsafety
 * (VecGeom/navigation/VNavigator.h:543-592): the safety first, and when it
 * covers the limit the track moves by it with no boundary distance computed;
 * otherwise the geometric step limited to it.  With reuse_safety the safety
 * left after such a move is carried to the next step.  steps[i] and
 * geo_steps[i] (steps that computed a boundary distance) are incremented. */
int vgm_nav_transport_limited(vgm_ctx *ctx, double *x, double *y, double *z, const double *dx, const double *dy,
                              const double *dz, const double *step_limit, size_t n, uint32_t *state, uint32_t *last,
                              uint8_t *boundary, int nsteps, int reuse_safety, int32_t *steps, int32_t *geo_steps,
                              unsigned flags);
/* BVH::ApproachNextDaughter (source/BVH.cpp:395-449, VecGeom/base/BVH.h:158)
 * per track: step[i] (in: the current step, out: the approached step) is
 * shortened to the distance along the ray to the nearest daughter bounding
 * box of the volume of state[i] (negative when the point is inside one),
 * skipping the volume last[i] from its surface.  Tracks outside the world
 * keep their step. */
int vgm_nav_approach(vgm_ctx *ctx, const double *x, const double *y, const double *z, const double *dx,
                     const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *last,
                     double *step, unsigned flags);
/* VLevelLocator::LevelLocateExclVol through the BVH
 * (VecGeom/navigation/BVHLevelLocator.h:73-95, VecGeom/base/BVH.h:219-344):
 * for point i in the frame of the volume of state[i], the first daughter in
 * BVH order, other than the daughter whose state is excl[i] (excl NULL or
 * excl[i] = 0: none), that holds it: by its Contains without a direction
 * (dx, dy, dz NULL), or with one, by its Inside being kInside, or kSurface
 * with the direction entering it (its Normal against the direction < 0).
 * child[i] = that daughter's state and (qx, qy, qz)[i] the point in its
 * frame; child[i] = 0 and the point unchanged when none holds it or
 * state[i] = 0. */
int vgm_nav_level_locate(vgm_ctx *ctx, const double *x, const double *y, const double *z, const double *dx,
                         const double *dy, const double *dz, size_t n, const uint32_t *state, const uint32_t *excl,
                         uint32_t *child, double *qx, double *qy, double *qz, unsigned flags);
/* out[i] = safety of global point i in its state (mother SafetyToOut, then
 * the daughters through the BVH); 0 for state 0. */
int vgm_nav_safety(vgm_ctx *ctx, const double *x, const double *y, const double *z, size_t n,
                   const uint32_t *state, double *out, unsigned flags);

/* ---- batched heterogeneous transport ---------------------------------------
 * No counterpart call exists in the reference: it replaces a caller's loop of
 * per-solid SOA3D calls (VPlacedVolume::SafetyToIn/Out, DistanceToIn/Out,
 * VecGeom/volumes/PlacedVolume.h:332-421) over many solids.  Solid i owns the
 * tracks [offsets[i], offsets[i+1]) of one set of SOA3D arrays, given in the
 * solid's local frame.  One track-step is: safety on the current side; a
 * physics step l ~ Exp(mean_step[i]) from a counter-based generator keyed by
 * (seed, track, step0 + k); if l exceeds the safety, the boundary distance
 * decides between moving l or crossing (boundary + 1e-6 mm push, side flips).
 * A step that does not cross ends in an interaction with a new isotropic
 * direction; outside the solid, coordinates wrap periodically within 1.3x
 * the solid's bounding box.  Positions and directions are updated in place.
 * All pointers are device memory.  Tracks are cut into tiles sorted by shape
 * family; each family runs as its own kernel, the families concurrently.
 * vgm_batch_autotune picks block, grid and tile size. */
/* Launch shape: block = threads per workgroup (multiple of 64, at most 256);
 * grid > 0 = persistent launch with that many workgroups in total, shared
 * between families by cost, pulling tiles from a queue; grid <= 0 = one
 * workgroup per tile; tile_tracks = tracks per tile.
 * VGM_MIXED selects float filtering with FP64 boundary steps. */
typedef struct vgm_batch vgm_batch;
int vgm_batch_create(vgm_ctx *ctx, int nsolids, const int *shapes, const int64_t *offsets, const double *mean_step,
                     vgm_batch **out);
int vgm_batch_destroy(vgm_batch *batch);
/* side[i] = Contains(point i) of its solid. */
int vgm_batch_classify(vgm_batch *batch, const double *x, const double *y, const double *z, int8_t *side);
/* Advances every track by nsteps track-steps; crossings[i] is
 * incremented per boundary crossed.  fp64_steps (host, may be NULL) receives the
 * number of steps of a VGM_MIXED run that the float safety did not settle.
 * With VGM_GRAPH the whole call (for learned families, two kernels per
 * track-step) is captured into a HIP graph and replayed by later calls with
 * the same arrays, step count, seed, flags and launch settings; step0 may
 * change between replays.  Not with VGM_TIMING. */
int vgm_batch_transport(vgm_batch *batch, double *x, double *y, double *z, double *dx, double *dy, double *dz,
                        int8_t *side, int32_t *crossings, int nsteps, int step0, uint64_t seed,
                        unsigned flags, unsigned long long *fp64_steps);
/* Kernel time in ms of each shape family (index = shape type: box, tube, cone,
 * trd, trapezoid, polycone, polyhedron, tessellated, orb, cut tube, boolean,
 * sphere, torus; VGM_NUM_SHAPE_TYPES
 * entries) in the last
 * vgm_batch_transport called with VGM_TIMING, measured with events on the
 * stream that family ran on; -1 for families not run.  Waits for them. */
int vgm_batch_family_ms(vgm_batch *batch, float *ms);
/* Counters of the last VGM_LEARNED transport: [0] bound evaluations, [1]
 * steps moved on a learned bound, [2] exact fallback steps, [3] (VGM_AUDIT)
 * bounds above the exact safety. */
int vgm_batch_learned_stats(vgm_batch *batch, unsigned long long *out4);
/* Mode per shape family chosen by the last vgm_batch_autotune (index = shape
 * type; 0 FP64, 1 mixed, 2 distance-first, 3 learned, 4 culled; -1 not
 * tuned).  The autotuner tries FP64 and each mode its flags allow (VGM_MIXED,
 * VGM_DISTANCE_FIRST, VGM_LEARNED, VGM_CULL), each family alone, and keeps the
 * fastest; VGM_TUNED transports use the choice. */
int vgm_batch_get_modes(vgm_batch *batch, int *modes);
/* Relative cost of each family (index = shape type): after
 * vgm_batch_autotune, the ms of that family's kernel alone in its mode;
 * before, its track count.  It splits persistent grids and orders launches. */
int vgm_batch_get_weights(vgm_batch *batch, double *ms);
/* Sets the modes VGM_TUNED transports use (distance-first, and the distance
 * order of culled, fall back to FP64 for families whose safety can exceed
 * their distance). */
int vgm_batch_set_modes(vgm_batch *batch, const int *modes);
/* A wave of the learned kernel runs its exact fallbacks once this percentage
 * of its live tracks wait for one (default 50). */
int vgm_batch_set_fallback(vgm_batch *batch, int pct);
int vgm_batch_set_launch(vgm_batch *batch, int block, int grid, int tile_tracks);
/* The families of a transport run concurrently on the batch's 4 family
 * streams (4, the default) or back to back on the context stream (1); the
 * autotuner tries both. */
int vgm_batch_set_streams(vgm_batch *batch, int streams);
int vgm_batch_get_streams(vgm_batch *batch, int *streams);
int vgm_batch_get_launch(vgm_batch *batch, int *shape3);
int vgm_batch_autotune(vgm_batch *batch, const double *x, const double *y, const double *z, const double *dx,
                       const double *dy, const double *dz, const int8_t *side, int nsteps, unsigned flags,
                       int *best3);


/* ---- Learned safety --------------------------------------------------------
 * A per-shape network (51 positional-encoding features of the point, four
 * 128-unit SiLU layers, one output; BF16 matrix products, FP32 accumulation)
 * trained on the GPU on exact safeties the shape kernels generate: SafetyToOut
 * for points inside, SafetyToIn outside, points uniform in `domain` x the
 * bounding box.  After training, the largest overestimate on `validation`
 * fresh points sets the margin; the served bound is max(0, prediction -
 * margin).  It has no counterpart in the reference: it is a fast lower bound
 * on VSafetyEstimator::ComputeSafetyForLocalPoint
 * (VecGeom/navigation/VSafetyEstimator.h:45-80) that the batched transport
 * uses in place of the exact safety when VGM_LEARNED is set, with the exact
 * kernels as fallback. */
typedef struct vgm_mlp_config {
  int64_t samples;      /* training samples (default 6250000) */
  int64_t validation;   /* margin validation samples (default 1000000) */
  int batch;            /* samples per step, a multiple of 256 (default 8192) */
  int epochs;           /* passes over the samples (default 4) */
  float lr;             /* Adam learning rate, cosine decay to 5% (default 2e-3) */
  float over_weight;    /* loss weight of overestimates (default 4) */
  double margin_factor; /* margin = factor * largest overestimate + ... (default 1.25) */
  double margin_abs;    /* ... + margin_abs * bounding-box diagonal (default 1e-4) */
  double domain;        /* sampling box / bounding box, per axis (default 1.3) */
  uint64_t seed;        /* default 42 */
} vgm_mlp_config;

typedef struct vgm_mlp_report {
  double margin;       /* mm */
  double max_over;     /* largest prediction - exact on the validation points, mm */
  double mean_abs_err; /* mm */
  double rms_err;      /* mm */
  double final_loss;   /* mean loss of the last steps (targets in units of scale) */
  double train_ms;     /* sample generation + training, GPU time */
  double gflop;        /* algorithmic training work */
  double scale;        /* bounding-box diagonal, mm */
  int64_t steps, samples;
} vgm_mlp_report;

void vgm_mlp_config_default(vgm_mlp_config *cfg);
int vgm_mlp_train(vgm_ctx *ctx, int shape, const vgm_mlp_config *cfg, vgm_mlp_report *report);
/* Prediction (mm) and served bound max(0, prediction - margin) at points in
 * the shape's frame. */
int vgm_mlp_predict(vgm_ctx *ctx, int shape, const double *x, const double *y, const double *z, size_t n,
                    double *out, unsigned flags);
int vgm_mlp_bound(vgm_ctx *ctx, int shape, const double *x, const double *y, const double *z, size_t n,
                  double *out, unsigned flags);
/* Save / load: 56321 FP32 parameters in the order W1[128][51], b1[128],
 * W2[128][128], b2, W3, b3, W4, b4, w5[128], b5, and 8 doubles: domain
 * centre (3), domain half-widths (3), scale, margin. */
int vgm_mlp_get(vgm_ctx *ctx, int shape, float *params, double *geom8);
int vgm_mlp_set(vgm_ctx *ctx, int shape, const float *params, const double *geom8);
/* Gradient of the training loss for one batch (host arrays): n points given
 * by normalised coordinates u (n x 3), targets = safety / scale, n a multiple
 * of 64.  Runs the training kernels without an update. */
int vgm_mlp_gradient(vgm_ctx *ctx, const float *params, const float *u, const float *target, size_t n,
                     float over_weight, float *grad, float *loss);

#ifdef __cplusplus
}
#endif

#endif /* VGM_VGM_H */
