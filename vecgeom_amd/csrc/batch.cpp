// Batched heterogeneous transport: work-queue construction, launch-shape
// autotuner and the C-ABI entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "context.h"
#include "transport.h"
#include "mlp_train.h"

using namespace vgm;

// Families run concurrently on this many streams of their own, forked from
// and joined back into the context stream.
constexpr int kFamilyStreams = 4;
// Device counters: one tile counter per family, one for classify, then two
// fallback-queue counters per family.
constexpr int kNumCounters = kNumShapeTypes + 1 + 2 * kNumShapeTypes;
static_assert(kNumCounters * sizeof(int) <= 256, "counters fit their 256-byte slot");

struct vgm_batch {
  vgm_ctx *ctx = nullptr;
  std::vector<SolidRec> solids;
  std::vector<Tile> tiles;
  std::vector<FamilyLaunch> fams;  // one per family present, in tile order
  double weight[kNumShapeTypes] = {};  // relative cost of each family's tiles
  int64_t ntracks = 0;
  LaunchShape shape{256, 0, 2048};
  int streams = kFamilyStreams;  // 1: the families back to back on the context stream
  bool tiles_dirty = true;
  char *dbuf = nullptr;  // solids | tiles | counters | fp64 step counter
  size_t dcap = 0;
  SolidRec *d_solids = nullptr;
  Tile *d_tiles = nullptr;
  int *d_counters = nullptr;  // one per family, then one for classify
  unsigned long long *d_fp64 = nullptr;
  hipStream_t fam_stream[kFamilyStreams] = {};
  hipEvent_t fork = nullptr, join[kFamilyStreams] = {};
  // VGM_TIMING: events around each family's kernel, on that family's stream
  hipEvent_t fam_t0[kNumShapeTypes] = {}, fam_t1[kNumShapeTypes] = {};
  bool fam_timed[kNumShapeTypes] = {};
  // learned safety: shape of each solid, network descriptors of the solids
  // that have one (SolidRec::mlp indexes them), device counters
  std::vector<int> shape_ids;
  std::vector<MlpDesc> descs;
  MlpDesc *d_mlps = nullptr;
  unsigned long long *d_learned = nullptr;
  int fallback_pct = 50;
  // learned families: per-track carried bound and flags, fallback queue
  char *lbuf = nullptr;
  double *d_tcarry = nullptr;
  uint8_t *d_tflags = nullptr;
  uint64_t *d_queue = nullptr;
  int fam_mode[kNumShapeTypes];  // FamilyMode chosen by the autotuner (VGM_TUNED), -1 if none
  // VGM_GRAPH: the captured transport of the last call and what it was
  // captured with; the first step index lives on the device
  int32_t *d_step0 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<char> graph_key;
  vgm_batch()
  {
    for (int &m : fam_mode) m = -1;
  }
};

namespace {

// Tiles of `tile_tracks` consecutive tracks, ordered by shape family and then
// by solid so that neighbouring workgroups share code and parameters.
void build_tiles(vgm_batch *b)
{
  std::vector<int> order(b->solids.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b->solids[a].type < b->solids[c].type; });
  b->tiles.clear();
  b->fams.clear();
  const int64_t tt = b->shape.tile_tracks;
  for (int s : order) {
    const SolidRec &r = b->solids[s];
    if (r.end > r.begin && (b->fams.empty() || b->fams.back().type != r.type))
      b->fams.push_back(
          FamilyLaunch{r.type, int(b->tiles.size()), int(b->tiles.size()), 0, nullptr, 0, kFamFP64, false, 0});
    if (r.end > r.begin && r.mlp >= 0) b->fams.back().has_mlp = true;
    if (r.end > r.begin) b->fams.back().stage_nd = std::max(b->fams.back().stage_nd, r.nd);
    for (int64_t t = r.begin; t < r.end; t += tt) {
      Tile tl;
      tl.solid = s;
      tl.pad = 0;
      tl.begin = t;
      tl.end = std::min<int64_t>(t + tt, r.end);
      b->tiles.push_back(tl);
    }
    if (!b->fams.empty()) b->fams.back().t1 = int(b->tiles.size());
  }
  int64_t off = 0;  // queue regions: each family's track count, in family order
  for (FamilyLaunch &f : b->fams) {
    f.queue_off = off;
    for (int t = f.t0; t < f.t1; ++t) off += b->tiles[size_t(t)].end - b->tiles[size_t(t)].begin;
  }
}

int ensure_streams(vgm_batch *b)
{
  if (b->fork) return VGM_OK;
  for (int i = 0; i < kFamilyStreams; ++i) {
    VGM_HIP(hipStreamCreateWithFlags(&b->fam_stream[i], hipStreamNonBlocking));
    VGM_HIP(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
  }
  for (int t = 0; t < kNumShapeTypes; ++t) {
    VGM_HIP(hipEventCreate(&b->fam_t0[t]));
    VGM_HIP(hipEventCreate(&b->fam_t1[t]));
  }
  VGM_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  return VGM_OK;
}

// Networks of the solids for a learned run (none otherwise); a change
// re-uploads the solid records.
void assign_networks(vgm_batch *b, bool learned)
{
  vgm_ctx *c = b->ctx;
  std::vector<MlpDesc> descs;
  for (size_t i = 0; i < b->solids.size(); ++i) {
    SolidRec &r = b->solids[i];
    const size_t sh = size_t(b->shape_ids[i]);
    int m = -1;
    if (learned && family_learnable(r.type) && sh < c->mlps.size() && c->mlps[sh].ready) {
      m = int(descs.size());
      descs.push_back(mlp_desc(c->mlps[sh]));
    }
    if (r.mlp != m) {
      r.mlp = m;
      b->tiles_dirty = true;
    }
  }
  if (descs.size() != b->descs.size() ||
      (!descs.empty() && std::memcmp(descs.data(), b->descs.data(), descs.size() * sizeof(MlpDesc)) != 0))
    b->tiles_dirty = true;
  b->descs = descs;
}

int upload(vgm_batch *b, bool learned = false)
{
  vgm_ctx *c = b->ctx;
  int rc = sync_geometry(c);
  if (rc) return rc;
  assign_networks(b, learned);
  if (learned && !b->lbuf && !b->descs.empty()) {
    const size_t n = size_t(std::max<int64_t>(1, b->ntracks));
    const size_t s1 = align256(n * sizeof(double)), s2 = align256(n), s3 = align256(n * sizeof(uint64_t));
    VGM_HIP(hipMalloc(&b->lbuf, s1 + s2 + s3));
    b->d_tcarry = reinterpret_cast<double *>(b->lbuf);
    b->d_tflags = reinterpret_cast<uint8_t *>(b->lbuf + s1);
    b->d_queue = reinterpret_cast<uint64_t *>(b->lbuf + s1 + s2);
  }
  if (!b->tiles_dirty && b->dbuf) return VGM_OK;
  build_tiles(b);
  const size_t s1 = align256(b->solids.size() * sizeof(SolidRec) + 1);
  const size_t s2 = align256(b->tiles.size() * sizeof(Tile) + 1);
  const size_t s3 = align256(b->solids.size() * sizeof(MlpDesc) + 1);
  const size_t total = s1 + s2 + 512 + s3;
  if (b->dcap < total) {
    if (b->dbuf) {
      VGM_HIP(hipStreamSynchronize(c->stream));
      VGM_HIP(hipFree(b->dbuf));
    }
    VGM_HIP(hipMalloc(&b->dbuf, total));
    b->dcap = total;
  }
  b->d_solids = reinterpret_cast<SolidRec *>(b->dbuf);
  b->d_tiles = reinterpret_cast<Tile *>(b->dbuf + s1);
  b->d_counters = reinterpret_cast<int *>(b->dbuf + s1 + s2);
  b->d_fp64 = reinterpret_cast<unsigned long long *>(b->dbuf + s1 + s2 + 256);
  b->d_learned = b->d_fp64 + 1;
  b->d_mlps = reinterpret_cast<MlpDesc *>(b->dbuf + s1 + s2 + 512);
  if (!b->descs.empty())
    VGM_HIP(hipMemcpyAsync(b->d_mlps, b->descs.data(), b->descs.size() * sizeof(MlpDesc), hipMemcpyHostToDevice,
                           c->stream));
  VGM_HIP(hipMemcpyAsync(b->d_solids, b->solids.data(), b->solids.size() * sizeof(SolidRec), hipMemcpyHostToDevice,
                         c->stream));
  VGM_HIP(hipMemcpyAsync(b->d_tiles, b->tiles.data(), b->tiles.size() * sizeof(Tile), hipMemcpyHostToDevice,
                         c->stream));
  VGM_HIP(hipStreamSynchronize(c->stream));
  b->tiles_dirty = false;
  return VGM_OK;
}

TransportArgs make_args(vgm_batch *b, double *x, double *y, double *z, double *dx, double *dy, double *dz,
                        int8_t *side, int32_t *crossings, int nsteps, int step0, uint64_t seed, unsigned flags = 0)
{
  TransportArgs a{};
  a.pool = b->ctx->dev.pool;
  a.solids = b->d_solids;
  a.tiles = b->d_tiles;
  a.ntiles = int32_t(b->tiles.size());
  a.nsteps = nsteps;
  a.step0 = step0;
  a.seed = seed;
  a.x = x, a.y = y, a.z = z;
  a.dx = dx, a.dy = dy, a.dz = dz;
  a.side = side;
  a.crossings = crossings;
  a.counter = b->d_counters + kNumShapeTypes;
  a.fp64_steps = b->d_fp64;
  a.mlps = b->d_mlps;
  a.learned = b->d_learned;
  a.audit = (flags & VGM_AUDIT) ? 1 : 0;
  a.fallback_pct = b->fallback_pct;
  a.fallback_grid = 8 * std::max(1, b->ctx->num_cus);
  a.tcarry = b->d_tcarry;
  a.tflags = b->d_tflags;
  a.queue = b->d_queue;
  a.qcount = b->d_counters + kNumShapeTypes + 1;
  return a;
}

LaunchShape resolved(vgm_batch *b)
{
  LaunchShape ls = b->shape;
  // persistent: never more workgroups than tiles, the surplus would only spin
  if (ls.grid > 0) ls.grid = int(std::min<size_t>(size_t(ls.grid), std::max<size_t>(1, b->tiles.size())));
  return ls;
}

// Mode of one family in a transport call: the autotuner's choice under
// VGM_TUNED, else from the flags (learned where a network exists, then mixed,
// then distance-first, else FP64).
int family_mode(const vgm_batch *b, const FamilyLaunch &f, unsigned flags)
{
  int m = kFamFP64;
  if ((flags & VGM_TUNED) && b->fam_mode[f.type] >= 0)
    m = b->fam_mode[f.type];
  else if ((flags & VGM_LEARNED) && f.has_mlp)
    m = kFamLearned;
  else if (flags & VGM_MIXED)
    m = kFamMixed;
  else if (flags & VGM_CULL)
    m = kFamCulled;
  else if (flags & VGM_DISTANCE_FIRST)
    m = kFamDistFirst;
  if (m == kFamLearned && !f.has_mlp) m = kFamCulled;
  return m;
}

// All families of one transport call: counters reset on the context stream,
// then each family's kernel on one of the family streams (heaviest families
// first, each to the least loaded stream), joined back into the context
// stream.  A single family runs on the context stream directly.
hipError_t launch_families(vgm_batch *b, unsigned flags, const TransportArgs &a, const LaunchShape &ls,
                           bool timing = false)
{
  hipStream_t main = b->ctx->stream;
  for (bool &t : b->fam_timed) t = false;
  // one family's kernel on `st`, bracketed by timing events when asked
  auto one = [&](hipStream_t st, const FamilyLaunch &f) -> hipError_t {
    hipError_t r;
    if (timing && (r = hipEventRecord(b->fam_t0[f.type], st)) != hipSuccess) return r;
    if ((r = launch_family(st, a, f, ls.block)) != hipSuccess) return r;
    if (timing) {
      if ((r = hipEventRecord(b->fam_t1[f.type], st)) != hipSuccess) return r;
      b->fam_timed[f.type] = true;
    }
    return hipSuccess;
  };
  hipError_t e = launch_reset(main, b->d_counters, kNumCounters, a.fp64_steps, a.learned);
  if (e != hipSuccess) return e;
  double wsum = 0;
  for (const FamilyLaunch &f : b->fams) wsum += b->weight[f.type];
  std::vector<FamilyLaunch> fl = b->fams;
  for (FamilyLaunch &f : fl) {
    const int nt = f.t1 - f.t0;
    f.mode = family_mode(b, f, flags);
    if (ls.grid > 0) {
      const double share = wsum > 0 ? b->weight[f.type] / wsum : 1.0 / double(fl.size());
      f.grid = std::max(1, std::min(nt, int(std::lround(share * ls.grid))));
      f.counter = b->d_counters + f.type;
    } else {
      f.grid = nt;
      f.counter = nullptr;
    }
  }
  std::sort(fl.begin(), fl.end(),
            [&](const FamilyLaunch &x, const FamilyLaunch &y) { return b->weight[x.type] > b->weight[y.type]; });
  if (fl.size() == 1 || b->streams == 1) {
    for (const FamilyLaunch &f : fl)
      if ((e = one(main, f)) != hipSuccess) return e;
    return hipSuccess;
  }
  double load[kFamilyStreams] = {};
  bool used[kFamilyStreams] = {};
  if ((e = hipEventRecord(b->fork, main)) != hipSuccess) return e;
  for (const FamilyLaunch &f : fl) {
    int s = 0;
    for (int i = 1; i < kFamilyStreams; ++i)
      if (load[i] < load[s]) s = i;
    if (!used[s]) {
      if ((e = hipStreamWaitEvent(b->fam_stream[s], b->fork, 0)) != hipSuccess) return e;
      used[s] = true;
    }
    load[s] += b->weight[f.type];
    if ((e = one(b->fam_stream[s], f)) != hipSuccess) return e;
  }
  for (int s = 0; s < kFamilyStreams; ++s) {
    if (!used[s]) continue;
    if ((e = hipEventRecord(b->join[s], b->fam_stream[s])) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(main, b->join[s], 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int vgm_batch_create(vgm_ctx *c, int nsolids, const int *shapes, const int64_t *offsets, const double *mean_step,
                     vgm_batch **out)
{
  if (!c || !shapes || !offsets || !mean_step || !out || nsolids <= 0) return fail(VGM_EINVAL, "null argument");
  vgm_batch *b = new vgm_batch();
  b->ctx = c;
  for (int i = 0; i < nsolids; ++i) {
    if (shapes[i] < 0 || size_t(shapes[i]) >= c->host.shapes.size() || offsets[i + 1] < offsets[i] ||
        !(mean_step[i] > 0)) {
      delete b;
      return fail(VGM_EINVAL, "invalid shape id, track range or mean step");
    }
    const ShapeRec &r = c->host.shapes[shapes[i]];
    SolidRec s{};
    s.type = r.type;
    s.off = r.off;
    s.nd = r.nd;
    s.mlp = -1;
    s.begin = offsets[i];
    s.end = offsets[i + 1];
    s.lambda = mean_step[i];
    double d2 = 0;
    for (int k = 0; k < 3; ++k) d2 += (r.bmax[k] - r.bmin[k]) * (r.bmax[k] - r.bmin[k]);
    s.scale = std::sqrt(d2);
    for (int k = 0; k < 3; ++k) {
      s.wc[k] = 0.5 * (r.bmin[k] + r.bmax[k]);
      s.wh[k] = 1.3 * (0.5 * (r.bmax[k] - r.bmin[k]));
      s.bc[k] = 0.5 * (r.bmin[k] + r.bmax[k]);
      s.bh[k] = 0.5 * (r.bmax[k] - r.bmin[k]);
    }
    b->solids.push_back(s);
    b->shape_ids.push_back(shapes[i]);
  }
  b->ntracks = offsets[nsolids];
  for (const SolidRec &s : b->solids) b->weight[s.type] += double(s.end - s.begin);
  *out = b;
  return VGM_OK;
}

int vgm_batch_destroy(vgm_batch *b)
{
  if (!b) return VGM_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_step0) (void)hipFree(b->d_step0);
  if (b->dbuf) (void)hipFree(b->dbuf);
  if (b->lbuf) (void)hipFree(b->lbuf);
  for (int i = 0; i < kFamilyStreams; ++i) {
    if (b->fam_stream[i]) (void)hipStreamDestroy(b->fam_stream[i]);
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (int t = 0; t < kNumShapeTypes; ++t) {
    if (b->fam_t0[t]) (void)hipEventDestroy(b->fam_t0[t]);
    if (b->fam_t1[t]) (void)hipEventDestroy(b->fam_t1[t]);
  }
  delete b;
  return VGM_OK;
}

int vgm_batch_learned_stats(vgm_batch *b, unsigned long long *out4)
{
  if (!b || !out4) return fail(VGM_EINVAL, "null argument");
  if (!b->d_learned) {
    for (int k = 0; k < 4; ++k) out4[k] = 0;
    return VGM_OK;
  }
  VGM_HIP(hipMemcpyAsync(out4, b->d_learned, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, b->ctx->stream));
  VGM_HIP(hipStreamSynchronize(b->ctx->stream));
  return VGM_OK;
}

int vgm_batch_get_weights(vgm_batch *b, double *ms)
{
  if (!b || !ms) return fail(VGM_EINVAL, "null argument");
  for (int t = 0; t < kNumShapeTypes; ++t) ms[t] = b->weight[t];
  return VGM_OK;
}

int vgm_batch_set_modes(vgm_batch *b, const int *modes)
{
  if (!b || !modes) return fail(VGM_EINVAL, "null argument");
  for (int t = 0; t < kNumShapeTypes; ++t)
    if (modes[t] < -1 || modes[t] > kFamCulled) return fail(VGM_EINVAL, "invalid family mode");
  for (int t = 0; t < kNumShapeTypes; ++t) b->fam_mode[t] = modes[t];
  return VGM_OK;
}

int vgm_batch_get_modes(vgm_batch *b, int *modes)
{
  if (!b || !modes) return fail(VGM_EINVAL, "null argument");
  for (int t = 0; t < kNumShapeTypes; ++t) modes[t] = b->fam_mode[t];
  return VGM_OK;
}

int vgm_batch_set_fallback(vgm_batch *b, int pct)
{
  if (!b || pct < 1 || pct > 100) return fail(VGM_EINVAL, "fallback percentage must be in [1, 100]");
  b->fallback_pct = pct;
  return VGM_OK;
}

int vgm_batch_family_ms(vgm_batch *b, float *ms)
{
  if (!b || !ms) return fail(VGM_EINVAL, "null argument");
  for (int t = 0; t < kNumShapeTypes; ++t) {
    ms[t] = -1.0f;
    if (!b->fam_timed[t]) continue;
    VGM_HIP(hipEventSynchronize(b->fam_t1[t]));
    VGM_HIP(hipEventElapsedTime(&ms[t], b->fam_t0[t], b->fam_t1[t]));
  }
  return VGM_OK;
}

int vgm_batch_set_launch(vgm_batch *b, int block, int grid, int tile_tracks)
{
  if (!b || block <= 0 || block > kTransportMaxBlock || block % 64 || tile_tracks <= 0)
    return fail(VGM_EINVAL, "block must be a multiple of 64 up to 256, tile_tracks > 0");
  if (tile_tracks != b->shape.tile_tracks) b->tiles_dirty = true;
  b->shape = LaunchShape{block, grid, tile_tracks};
  return VGM_OK;
}

int vgm_batch_set_streams(vgm_batch *b, int streams)
{
  if (!b || (streams != 1 && streams != kFamilyStreams))
    return fail(VGM_EINVAL, "streams must be 1 or " + std::to_string(kFamilyStreams));
  b->streams = streams;
  return VGM_OK;
}

int vgm_batch_get_streams(vgm_batch *b, int *streams)
{
  if (!b || !streams) return fail(VGM_EINVAL, "null argument");
  *streams = b->streams;
  return VGM_OK;
}

int vgm_batch_get_launch(vgm_batch *b, int *shape3)
{
  if (!b || !shape3) return fail(VGM_EINVAL, "null argument");
  const LaunchShape ls = resolved(b);
  shape3[0] = ls.block;
  shape3[1] = ls.grid;
  shape3[2] = ls.tile_tracks;
  return VGM_OK;
}

int vgm_batch_classify(vgm_batch *b, const double *x, const double *y, const double *z, int8_t *side)
{
  if (!b || !x || !y || !z || !side) return fail(VGM_EINVAL, "null argument");
  VGM_HIP(hipSetDevice(b->ctx->device));
  int rc = upload(b);
  if (rc) return rc;
  TransportArgs a = make_args(b, const_cast<double *>(x), const_cast<double *>(y), const_cast<double *>(z), nullptr,
                              nullptr, nullptr, side, nullptr, 0, 0, 0);
  VGM_HIP(launch_reset(b->ctx->stream, b->d_counters, kNumCounters, nullptr, nullptr));
  const int grid = int(std::min<size_t>(b->tiles.size(), size_t(4 * std::max(1, b->ctx->num_cus))));
  VGM_HIP(launch_classify(b->ctx->stream, a, std::max(1, grid), 256));
  VGM_HIP(hipStreamSynchronize(b->ctx->stream));
  return VGM_OK;
}

namespace {
// VGM_GRAPH: the whole transport call (counter reset, the families' kernels
// on their streams, and for learned families the per-step kernel pairs)
// captured once into a HIP graph on the context stream and replayed while
// the arrays, step count, seed, flags, launch shape, streams, modes and
// uploaded tables stay the same.  The first step index is the one argument
// that changes from call to call: the kernels read it from the device, and a
// one-thread kernel writes it before each replay.
int graph_transport(vgm_batch *b, unsigned flags, TransportArgs a, int step0)
{
  hipStream_t main = b->ctx->stream;
  if (!b->d_step0) VGM_HIP(hipMalloc(&b->d_step0, 256));
  a.step0_dev = b->d_step0;
  a.step0 = 0;
  const LaunchShape ls = resolved(b);
  // everything the captured launches depend on, byte for byte
  std::vector<char> key;
  auto put = [&key](const void *p, size_t n) {
    const char *c = static_cast<const char *>(p);
    key.insert(key.end(), c, c + n);
  };
  put(&a, sizeof a);
  const unsigned f = flags & ~(VGM_GRAPH | VGM_ASYNC);
  put(&f, sizeof f);
  put(&ls, sizeof ls);
  put(&b->streams, sizeof b->streams);
  put(b->fam_mode, sizeof b->fam_mode);
  put(b->weight, sizeof b->weight);
  VGM_HIP(launch_set_step0(main, b->d_step0, step0));
  if (!b->graph_exec || key != b->graph_key) {
    if (b->graph_exec) VGM_HIP(hipGraphExecDestroy(b->graph_exec));
    if (b->graph) VGM_HIP(hipGraphDestroy(b->graph));
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
    VGM_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    const hipError_t e = launch_families(b, flags, a, ls, false);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(main, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e != hipSuccess ? e : e2, "capturing the batched transport");
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      (void)hipGraphDestroy(g);
      return hip_fail(e3, "hipGraphInstantiate");
    }
    b->graph = g;
    b->graph_exec = x;
    b->graph_key = key;
  }
  VGM_HIP(hipGraphLaunch(b->graph_exec, main));
  return VGM_OK;
}
}  // namespace

int vgm_batch_transport(vgm_batch *b, double *x, double *y, double *z, double *dx, double *dy, double *dz,
                        int8_t *side, int32_t *crossings, int nsteps, int step0, uint64_t seed,
                        unsigned flags, unsigned long long *fp64_steps)
{
  if (!b || !x || !y || !z || !dx || !dy || !dz || !side || !crossings || nsteps < 0)
    return fail(VGM_EINVAL, "null argument or negative step count");
  VGM_HIP(hipSetDevice(b->ctx->device));
  int rc = upload(b, (flags & (VGM_LEARNED | VGM_TUNED)) != 0);
  if (rc) return rc;
  TransportArgs a = make_args(b, x, y, z, dx, dy, dz, side, crossings, nsteps, step0, seed, flags);
  if ((rc = ensure_streams(b))) return rc;
  if (flags & VGM_GRAPH) {
    if (flags & VGM_TIMING) return fail(VGM_EINVAL, "VGM_GRAPH does not time the family kernels (VGM_TIMING)");
    if ((rc = graph_transport(b, flags, a, step0))) return rc;
  } else {
    VGM_HIP(launch_families(b, flags, a, resolved(b), (flags & VGM_TIMING) != 0));
  }
  if (fp64_steps) {
    VGM_HIP(hipMemcpyAsync(fp64_steps, b->d_fp64, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           b->ctx->stream));
    VGM_HIP(hipStreamSynchronize(b->ctx->stream));
  } else if (!(flags & VGM_ASYNC)) {
    VGM_HIP(hipStreamSynchronize(b->ctx->stream));
  }
  return VGM_OK;
}

// Autotuner.  Works on scratch copies of the track state (the caller's
// buffers are left untouched).  First each family is timed alone, which
// gives the weights that split the persistent grid and balance the family
// streams; then every candidate launch shape is timed on the whole batch and
// the fastest is kept.  Candidates: workgroup size (64, 128, 256), tile size
// (512, 2048 tracks), and either one workgroup per tile or a persistent grid
// of 4 workgroups per CU.
int vgm_batch_autotune(vgm_batch *b, const double *x, const double *y, const double *z, const double *dx,
                       const double *dy, const double *dz, const int8_t *side, int nsteps, unsigned flags,
                       int *best3)
{
  if (!b || !x || !y || !z || !dx || !dy || !dz || !side || nsteps <= 0) return fail(VGM_EINVAL, "null argument");
  vgm_ctx *c = b->ctx;
  VGM_HIP(hipSetDevice(c->device));
  int rc = ensure_streams(b);
  if (rc) return rc;
  const size_t n = size_t(b->ntracks);
  char *work = nullptr;
  const size_t bytes = align256(6 * n * sizeof(double)) + align256(n) + align256(n * sizeof(int32_t));
  VGM_HIP(hipMalloc(&work, bytes));
  double *wx = reinterpret_cast<double *>(work);
  double *wy = wx + n, *wz = wy + n, *wdx = wz + n, *wdy = wdx + n, *wdz = wdy + n;
  int8_t *ws = reinterpret_cast<int8_t *>(work + align256(6 * n * sizeof(double)));
  int32_t *wc = reinterpret_cast<int32_t *>(work + align256(6 * n * sizeof(double)) + align256(n));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const LaunchShape keep = b->shape;
  const int keep_streams = b->streams;
  int keep_modes[kNumShapeTypes];
  std::memcpy(keep_modes, b->fam_mode, sizeof keep_modes);
  const int cus = std::max(1, c->num_cus);
  float best = 1e30f;
  LaunchShape best_shape = keep;
  int best_streams = keep_streams;
  double weight[kNumShapeTypes] = {};

  auto stage = [&]() -> bool {
    return !(hipMemcpyAsync(wx, x, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(wy, y, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(wz, z, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(wdx, dx, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(wdy, dy, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(wdz, dz, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream) ||
             hipMemcpyAsync(ws, side, n, hipMemcpyDeviceToDevice, c->stream) ||
             hipMemsetAsync(wc, 0, n * sizeof(int32_t), c->stream));
  };
  // time `launch` on freshly staged state, second of two repetitions
  auto timed = [&](auto &&launch, float &ms) -> int {
    for (int rep = 0; rep < 2; ++rep) {
      if (!stage()) return fail(VGM_EHIP, "autotune staging copy failed");
      (void)hipEventRecord(e0, c->stream);
      if (launch() != hipSuccess) return fail(VGM_EHIP, "autotune launch failed");
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&ms, e0, e1);
    }
    return VGM_OK;
  };

  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(VGM_EHIP, "hipEventCreate failed");
    goto done;
  }
  // 1. each family alone, one workgroup of 256 per 2048-track tile, in every
  // mode the flags allow (FP64 always; mixed, distance-first and learned on
  // request): the fastest mode is kept, its time is the family's weight
  vgm_batch_set_launch(b, 256, 0, 2048);
  if ((rc = upload(b, (flags & VGM_LEARNED) != 0))) goto done;
  for (const FamilyLaunch &f0 : b->fams) {
    std::vector<int> cand{kFamFP64};
    if (flags & VGM_MIXED) cand.push_back(kFamMixed);
    if (flags & VGM_DISTANCE_FIRST) cand.push_back(kFamDistFirst);
    if (flags & VGM_CULL) cand.push_back(kFamCulled);
    if ((flags & VGM_LEARNED) && f0.has_mlp) cand.push_back(kFamLearned);
    float fbest = 1e30f;
    for (int m : cand) {
      FamilyLaunch f = f0;
      f.grid = f.t1 - f.t0;
      f.counter = nullptr;
      f.mode = m;
      TransportArgs a = make_args(b, wx, wy, wz, wdx, wdy, wdz, ws, wc, nsteps, 0, 1, flags);
      float ms = 0;
      auto one = [&]() {
        hipError_t e = launch_reset(c->stream, b->d_counters, kNumCounters, nullptr, a.learned);
        return e != hipSuccess ? e : launch_family(c->stream, a, f, 256);
      };
      if ((rc = timed(one, ms))) goto done;
      if (ms < fbest) {
        fbest = ms;
        b->fam_mode[f.type] = m;
      }
    }
    weight[f0.type] = std::max(1e-3, double(fbest));
  }
  for (int t = 0; t < kNumShapeTypes; ++t)
    if (weight[t] > 0) b->weight[t] = weight[t];
  // 2. launch shapes on the whole batch, the families concurrently on the
  // family streams or back to back on one
  {
    const int blocks[] = {64, 128, 256};
    const int tiles[] = {256, 512, 2048};
    const int per_cu[] = {0, 4};
    const int nstreams[] = {kFamilyStreams, 1};
    for (int ns : nstreams)
      for (int bl : blocks)
        for (int tt : tiles)
          for (int pc : per_cu) {
            if (tt < bl) continue;
            vgm_batch_set_launch(b, bl, pc * cus, tt);
            b->streams = ns;
            if ((rc = upload(b, (flags & VGM_LEARNED) != 0))) goto done;
            TransportArgs a = make_args(b, wx, wy, wz, wdx, wdy, wdz, ws, wc, nsteps, 0, 1, flags);
            const LaunchShape ls = resolved(b);
            float ms = 0;
            if ((rc = timed([&]() { return launch_families(b, flags | VGM_TUNED, a, ls); }, ms))) goto done;
            if (ms < best) {
              best = ms;
              best_shape = LaunchShape{bl, pc * cus, tt};
              best_streams = ns;
            }
          }
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(work);
  if (rc) {
    b->shape = keep;
    b->streams = keep_streams;
    std::memcpy(b->fam_mode, keep_modes, sizeof keep_modes);
    b->tiles_dirty = true;
    return rc;
  }
  vgm_batch_set_launch(b, best_shape.block, best_shape.grid, best_shape.tile_tracks);
  b->streams = best_streams;
  if (best3) {
    best3[0] = best_shape.block;
    best3[1] = best_shape.grid;
    best3[2] = best_shape.tile_tracks;
  }
  return VGM_OK;
}

}  // extern "C"
