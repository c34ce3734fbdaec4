// Host driver of the Hausdorff dilation engine: the C ABI of include/hdm.h.
//
// One call of the single-pair entry points enqueues, on one stream:
//   reset -> pack -> totals -> tile metadata -> persistent dilation, whose
//   workgroup 0 writes the result into pinned host memory as it leaves.
// There are no host round trips between the kernels (the reference reads its
// counters back twice mid-pipeline to size allocations, hello.cu:952, 1012):
// the workspace is sized up front from the dims, grows only, and is reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/hdm.h"
#include "../../include/hdm_h5.h"
#include "hdm_host.h"

using namespace hdm;

namespace hdm {

int map_hip(hipError_t e) {
  if (e == hipSuccess) return HDM_OK;
  if (e == hipErrorOutOfMemory) return HDM_E_NOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return HDM_E_NODEVICE;
  return HDM_E_HIP;
}

int DevMem::ensure(size_t n) {
  if (n <= cap) return HDM_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = std::max(n, cap * 5 / 4);
  if (hipMalloc(&p, want) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return HDM_E_NOMEM;
  }
  cap = want;
  return HDM_OK;
}

void DevMem::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

static std::mutex g_ctx_mu;
static std::map<int, Ctx *> g_ctx;

int get_ctx(Ctx **out) {
  int dev = -1;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return HDM_E_NODEVICE;
  }
  if (hipGetDevice(&dev) != hipSuccess) return HDM_E_NODEVICE;
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  auto it = g_ctx.find(dev);
  if (it != g_ctx.end()) {
    *out = it->second;
    return HDM_OK;
  }
  Ctx *c = new Ctx();
  c->dev = dev;
  int rc = HDM_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = HDM_E_HIP;
  if (rc == HDM_OK &&
      hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    rc = HDM_E_HIP;
  int khz = 0;
  if (rc == HDM_OK &&
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
    rc = HDM_E_HIP;
  c->clock_khz = khz > 0 ? khz : 100000;
  for (int k = 1; rc == HDM_OK && k < kLaunchIds; ++k) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hdm_tile_launch(k)->dilate_fn,
                                                     kWgThreads, 0) != hipSuccess ||
        nb < 1)
      rc = HDM_E_HIP;
    c->occ[k] = nb;
    if (rc == HDM_OK &&  // the one-pair kernel has its own resident grid
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hdm_tile_launch(k)->dilate_single_fn,
                                                      kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_single[k] = nb;
    if (rc == HDM_OK &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hdm_tile_launch(k)->pack_fn,
                                                      kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_pack[k] = nb;
    if (rc == HDM_OK &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, hdm_tile_launch(k)->fused_fn,
                                                      kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_fused[k] = nb;
    const void *ds_fn = hdm_tile_launch(k)->fused_ds_fn;
    if (rc == HDM_OK && ds_fn &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ds_fn, kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_fused_ds[k] = ds_fn ? nb : 0;
    const void *tb_fn = hdm_tile_launch(k)->fused_tb_fn;
    if (rc == HDM_OK && tb_fn &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tb_fn, kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_fused_tb[k] = tb_fn ? nb : 0;
    const void *near_fn = hdm_tile_launch(k)->fused_near_fn;
    if (rc == HDM_OK && near_fn &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, near_fn, kWgThreads, 0) != hipSuccess ||
         nb < 1))
      rc = HDM_E_HIP;
    c->occ_fused_near[k] = near_fn ? nb : 0;
  }
  // the dilation kernel stores the result and the engine head here directly:
  // coherent (fine-grained) host memory, so the stores are seen by the host
  // once the kernel has completed
  if (rc == HDM_OK && hipHostMalloc(&c->host_out, 4096, hipHostMallocCoherent) != hipSuccess)
    rc = HDM_E_HIP;
  // pinned memory can come back from an earlier allocation with its contents:
  // no ready marker left there may match this context's first call
  if (rc == HDM_OK) std::memset(c->host_out, 0, 4096);
  for (int i = 0; rc == HDM_OK && i < 5; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = HDM_E_HIP;
  if (rc != HDM_OK) {
    (void)hipGetLastError();
    delete c;
    return rc;
  }
  g_ctx[dev] = c;
  *out = c;
  return HDM_OK;
}

// npairs selects the dilation kernel the grid is for (launch_dilate runs the
// one-pair kernel for npairs == 1)
int resident_grid(const Ctx *c, int tile, int requested, int npairs) {
  const int occ = npairs == 1 ? c->occ_single[tile] : c->occ[tile];
  const int maxres = std::min(occ * c->cus, kMaxGrid);
  if (requested > 0) return std::min(requested, maxres);
  // every resident workgroup: dense iterations want many waves, and the
  // barrier polls arrivals in batches, so a sparse queue pays little for them
  return maxres;
}

// Bound of every grid-barrier wait: 2 s of wall-clock ticks.  HDM_SPIN_TICKS
// (read at each call or batch / server creation) replaces it, so tests can
// make the waits expire and check the timeout path; every workgroup then
// leaves its loop and the call returns HDM_E_TIMEOUT.
uint64_t spin_ticks(const Ctx *c) {
  const char *e = std::getenv("HDM_SPIN_TICKS");
  if (e && *e) return std::strtoull(e, nullptr, 10);
  return (uint64_t)(c->clock_khz * 1000.0 * 2.0);
}

// Grid barrier: 0 (default) flag array with agent-coherent data accesses and
// no cache maintenance; 1 the same plus L2 write-back / invalidate at every
// barrier (HDM_BARRIER_FENCE=1, a conservative setting for comparison).
// A single pair runs as one fused launch unless HDM_FUSED=0 asks for the
// four-launch path (reset, pack, seed, dilation), kept for batches and for
// comparison.
bool fused_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("HDM_FUSED");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// HDM_COOP=1 launches the fused single-pair kernel with
// hipLaunchCooperativeKernel: the runtime then guarantees that the whole grid
// is resident at once (or refuses the launch), where a plain launch relies on
// the grid being sized within this process's occupancy.
bool coop_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("HDM_COOP");
    return e && std::atoi(e) == 1;
  }();
  return on;
}

// The dilation runs two steps per grid barrier (64x8x8 tiles, 6 element) --
// the fused single-pair launch, batches and the server -- unless HDM_DS=0
// asks for one.
bool ds_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("HDM_DS");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// HDM_TB: 0 runs every single pair in the two-step fused launch; 1 (the
// default) queues the region launch behind it, and the two-step launch hands
// over the pairs whose HD lower bound reaches 16 steps; 2 hands over every pair.
static int tb_setting() {
  static const int v = [] {
    const char *e = std::getenv("HDM_TB");
    if (!e || !*e) return 1;
    const int x = std::atoi(e);
    return x == 0 ? 0 : x == 2 ? 2 : 1;
  }();
  return v;
}
bool tb_enabled() { return tb_setting() != 0; }
int tb_mode() { return tb_setting(); }
// HDM_SKIP_ROWS=0: a batch with a near pass writes every row in its pack (no
// rows skipped, no zero fill afterwards)
bool skip_rows() {
  static const bool on = [] {
    const char *e = std::getenv("HDM_SKIP_ROWS");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}
int pack_pools() {
  static const int v = [] {
    const char *e = std::getenv("HDM_PACK_POOLS");
    return e && std::atoi(e) == 0 ? 0 : 8;
  }();
  return v;
}

bool near_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("HDM_NEAR");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// Default grid of the fused launch: one workgroup per two tiles, at least
// 64, at most every resident one.  A small volume's iterations queue few
// entries, and every workgroup costs the grid barriers an arrival to poll and
// a release to read (a 64^3 pair measured 57 -> 54 us on 64 workgroups
// instead of 512; pairs of 2048 tiles or more run fastest on all of them).
int fused_grid_of(int occ, int cus, int requested, uint64_t ntiles) {
  const int maxres = std::min(std::max(1, occ) * cus, kMaxGrid);
  if (requested > 0) return std::min(requested, maxres);
  return (int)std::min<uint64_t>((uint64_t)maxres, std::max<uint64_t>(64, ntiles / 2));
}
// Default grid of the near launch: the fused rule over at most two workgroups
// per CU.  Its pack streams as fast on two as on three, and every workgroup
// more costs the pack barrier, the share count and the arrival count (CT
// pair 52.8 -> 49.3 us with 512 instead of 768 on 256 CUs; the 512^3 cube,
// with more near-pass entries than workgroups, 74.0 -> 74.4 us).  A requested
// grid keeps the fused rule's bound.
int near_grid_of(int occ, int cus, int requested, uint64_t ntiles) {
  return fused_grid_of(requested > 0 ? occ : std::min(occ, 2), cus, requested, ntiles);
}
// Grid of the region launch: the fused launch's rule over tiles (the launch
// may run the two-step body), at most four workgroups per region and at most
// 2 x kWgThreads in all (its seed period publishes one count per workgroup
// into a table of that size and gives each a segment of the queue area sized
// by regions per workgroup)
int tb_grid(int occ, int cus, int requested, uint64_t ntiles, uint32_t nregions) {
  int g = fused_grid_of(occ, cus, requested, ntiles);
  g = (int)std::min<uint64_t>((uint64_t)g, 4ull * std::max<uint32_t>(1u, nregions));
  return std::min(g, 2 * kWgThreads);
}
int fused_grid(const Ctx *c, int tile, int requested, bool ds, uint64_t ntiles) {
  return fused_grid_of(ds ? c->occ_fused_ds[tile] : c->occ_fused[tile], c->cus, requested,
                       ntiles);
}

int barrier_mode() {
  static int mode = [] {
    const char *e = std::getenv("HDM_BARRIER_FENCE");
    return (e && std::atoi(e) == 1) ? 1 : 0;
  }();
  return mode;
}

int norm_tile(int tile) { return (tile <= 0 || tile >= HDM_TILE_COUNT) ? HDM_TILE_64x8x8 : tile; }

// ds: the 64x8x8 tile with the 6 element takes the launch id whose seed and
// dilation kernels run two steps per engine iteration
int select_launch(int tile_opt, int conn_opt, int *conn, bool ds) {
  if (tile_opt < 0 || tile_opt >= HDM_TILE_COUNT) return -1;
  const int cn = conn_opt == 0 ? 6 : conn_opt;
  if (cn != 6 && cn != 18 && cn != 26) return -1;
  // the 18 and 26 elements run on 64x8x8 tiles (one tile per wavefront)
  if (cn != 6 && tile_opt != HDM_TILE_AUTO && tile_opt != HDM_TILE_64x8x8) return -1;
  *conn = cn;
  return launch_id(norm_tile(tile_opt), cn, ds);
}

// Launch shapes.  Pack: at most the resident waves, each walking one
// contiguous range of tiles, two wave steps each when the grid allows.  Seed:
// one thread per tile.  Dilation: occupancy-bounded.
LaunchCfg launch_cfg(const Ctx *c, int tile, uint64_t total_rw, uint64_t total_tiles,
                     int grid_req, int chunk, int npairs) {
  LaunchCfg L;
  (void)total_rw;
  const uint64_t rows = (uint64_t)tile_shape(tile).rows();
  const uint64_t tpw = (uint64_t)kWave / rows;  // tiles per wave step of the pack
  const uint64_t pack_waves = (total_tiles + 2 * tpw - 1) / (2 * tpw);  // two steps per wave
  const uint64_t pack_want = (pack_waves + kWavesPerWg - 1) / kWavesPerWg;
  const uint64_t pack_res = (uint64_t)std::max(1, c->occ_pack[tile]) * c->cus;
  L.pack_blocks = (unsigned)std::max<uint64_t>(1, std::min(pack_want, pack_res));
  const uint64_t seed_want = (total_tiles + kWgThreads - 1) / kWgThreads;
  L.seed_blocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(seed_want, 8192));
  L.grid = resident_grid(c, tile, grid_req, npairs);
  L.chunk = std::max(1, chunk);
  L.spin = spin_ticks(c);
  return L;
}

hipError_t enqueue_group(hipStream_t s, const TileLaunch *tl, const GroupPtrs &P, int npairs,
                         const PairDesc *single, const PackIndex *single_idx,
                         uint64_t total_rw, uint64_t total_tiles, uint32_t qcap,
                         const LaunchCfg &cfg, int qstats_len) {
  auto mark = [&](int i) {
    // timing 3 brackets the dilation only: its two records are queued while
    // the pack still runs, so they add no idle time to the call
    if (!cfg.ev || (cfg.ev_mode == 3 && i < 2)) return hipSuccess;
    return hipEventRecord(cfg.ev[i], s);
  };
  hipError_t e = hdm_launch_reset(s, P.descs, single, single_idx, P.idx, P.ctrs, 0, npairs,
                                  P.engine, 1, qcap);
  if (e != hipSuccess || (e = mark(0)) != hipSuccess) return e;
  (void)total_rw;
  // several pairs: the last quarter of the steps from the engine's per-XCD
  // counters, which the reset above zeroed.  A batch whose near pass runs
  // writes no rows of a mask that is empty in a tile (the near pass reads
  // such rows as zero from the tile flags) and zero-fills them afterwards
  // for the pairs the near pass leaves to the seed and the dilation
  const bool skip = cfg.near_list && npairs > 1 && tl->pack_skip && tl->fill && skip_rows();
  e = (skip ? tl->pack_skip : tl->pack)(dim3(cfg.pack_blocks), s, P.descs, P.idx, P.ctrs, npairs,
                                        total_tiles,
                                        npairs > 1 && pack_pools() ? &P.engine->pack_grab[0] : nullptr);
  if (e != hipSuccess) return e;
  if (npairs > 1 && (e = hdm_launch_totals(s, P.ctrs, 0, npairs)) != hipSuccess) return e;
  if (cfg.near_list && npairs > 1 &&
      (e = hdm_launch_near_batch(s, P.descs, P.idx, P.ctrs, npairs, cfg.near_words,
                                 &P.engine->near_count,
                                 cfg.near_list, cfg.near_cap, cfg.near_grid)) != hipSuccess)
    return e;
  if (skip && (e = tl->fill(dim3(cfg.pack_blocks), s, P.descs, P.ctrs, npairs)) != hipSuccess)
    return e;
  if ((e = mark(1)) != hipSuccess) return e;
  e = tl->seed(dim3(cfg.seed_blocks), s, P.descs, P.idx, P.ctrs, npairs, P.engine, P.q, qcap,
               total_tiles, 0);
  if (e != hipSuccess || (e = mark(2)) != hipSuccess) return e;
  DilateArgs a{};
  a.descs = P.descs;
  a.ctrs = P.ctrs;
  a.npairs = npairs;
  a.max_iters = INT_MAX;
  a.es = P.engine;
  a.q = P.q;
  a.qcap = qcap;
  a.chunk = cfg.chunk;
  a.spin = cfg.spin;
  a.qstats = P.qstats;
  a.qstats_len = qstats_len;
  a.flags = barrier_mode();
  a.trace = cfg.trace;
  a.outs = cfg.outs ? cfg.outs : P.outs;
  a.head = cfg.head;
  a.ready = cfg.ready;
  a.ready_seq = cfg.ready_seq;
  e = tl->dilate(dim3(cfg.grid), s, a);
  if (e != hipSuccess) return e;
  return mark(3);
}

GroupPtrs carve_common(char *base, const CommonLayout &L) {
  GroupPtrs P;
  P.descs = (PairDesc *)(base + L.descs);
  P.idx = (PackIndex *)(base + L.idx);
  P.ctrs = (PairCtr *)(base + L.ctrs);
  P.outs = (PairOut *)(base + L.outs);
  P.engine = (Engine *)(base + L.engine);
  P.q = (uint64_t *)(base + L.q);
  P.qstats = (uint32_t *)(base + L.qstats);
  return P;
}

void fill_result(hdm_result *out, const PairOut &o, const EngineHead *es, uint32_t tiles,
                 int iterations) {
  std::memset(out, 0, sizeof(*out));
  out->hd = o.hd, out->hd_fp = o.hd_fp, out->hd_fn = o.hd_fn;
  out->iterations = iterations;
  out->sum_fp = o.sum_fp, out->cnt_fp = o.cnt_fp, out->sum_fn = o.sum_fn, out->cnt_fn = o.cnt_fn;
  out->sum = o.sum_fp + o.sum_fn;
  out->count = o.cnt_fp + o.cnt_fn;
  out->n_gold = o.n_gold, out->n_segm = o.n_segm, out->n_fp = o.n_fp, out->n_fn = o.n_fn;
  // the one floating-point step, on the host (hello.cu never forms the average)
  out->avg_hd = out->count ? (double)out->sum / (double)out->count : 0.0;
  if (es) {
    out->queue_total = es->qtotal;
    out->queue_max = es->qmax;
  }
  out->tiles = tiles;
}

}  // namespace hdm

// ---------------------------------------------------------------- misc ABI
extern "C" int hdm_abi_version(void) { return HDM_ABI_VERSION; }

extern "C" const char *hdm_strerror(int code) {
  switch (code) {
    case HDM_OK: return "ok";
    case HDM_E_INVALID: return "invalid argument";
    case HDM_E_EMPTY_MASK: return "one mask is empty while the other has voxels";
    case HDM_E_HIP: return "HIP runtime error";
    case HDM_E_NOMEM: return "device memory allocation failed";
    case HDM_E_TIMEOUT: return "device grid barrier timed out";
    case HDM_E_BUDGET: return "request exceeds the HBM budget";
    case HDM_E_NODEVICE: return "no HIP device";
    case HDM_E_STUCK: return "work queue drained with targets uncovered";
    case HDM_E_CAPACITY: return "capacity exceeded";
    case HDM_E_IO: return "file cannot be opened or read";
    case HDM_E_UNSUPPORTED: return "HDF5 feature not supported by the reader";
    case HDM_E_FORMAT: return "not an HDF5 file, or a corrupt structure";
    case HDM_E_NOTFOUND: return "no dataset at that path";
    default: return "unknown error";
  }
}

extern "C" void hdm_opts_init(hdm_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->robustness = 1.0;
  o->ptr_kind = HDM_PTR_DEVICE;
  o->tile = HDM_TILE_AUTO;
  o->crop = 1;
}

extern "C" int hdm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" uint64_t hdm_workspace_bytes(int nx, int ny, int nz, int tile) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  const TileShape ts = tile_shape(norm_tile(tile));
  const PairGeom g = pair_geom(nx, ny, nz, ts);
  const CommonLayout L = common_layout(1, 2ull * g.ntiles, g.hist_len);
  return L.total + g.total;
}

// Waits until the kernel's ready marker holds seq; the stream is queried
// every 200 us so a launch that fails or ends without the marker is reported.
int hdm::wait_ready(hipStream_t s, volatile uint32_t *ready, uint32_t seq) {
  using clk = std::chrono::steady_clock;
  auto next = clk::now() + std::chrono::microseconds(200);
  for (uint32_t k = 0;; ++k) {
    if (*ready == seq) break;
    if ((k & 63) == 63 && clk::now() >= next) {
      const hipError_t q = hipStreamQuery(s);
      if (q == hipSuccess) {
        if (*ready == seq) break;
        return HDM_E_HIP;
      }
      if (q != hipErrorNotReady) return map_hip(q);
      next = clk::now() + std::chrono::microseconds(200);
    }
    __builtin_ia32_pause();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return HDM_OK;
}

// ---------------------------------------------------------------- single pair
static int hausdorff_core(const void *gold, const void *segm, int dtype, int label, int nx,
                          int ny, int nz, const hdm_opts *opts_in, uint16_t *map_out,
                          hdm_result *out) {
  hdm_opts opts;
  hdm_opts_init(&opts);
  if (opts_in) opts = *opts_in;
  if (!gold || !segm || !out || nx < 1 || ny < 1 || nz < 1 || nx > 65535 || ny > 65535 ||
      nz > 65535)
    return HDM_E_INVALID;
  if (!(opts.robustness >= 0.0 && opts.robustness <= 1.0)) return HDM_E_INVALID;
  if (opts.ptr_kind != HDM_PTR_DEVICE && opts.ptr_kind != HDM_PTR_HOST) return HDM_E_INVALID;
  int conn = 6;
  const int tile = select_launch(opts.tile, opts.connectivity, &conn);
  if (tile < 0) return HDM_E_INVALID;
  if (!map_out && opts.map_out) map_out = opts.map_out;
  // the map holds d(v) as uint16; the largest first hit in an nx x ny x nz
  // box is nx + ny + nz - 3 steps (6 element; fewer for 18 and 26), so dims
  // whose sum could exceed that are refused instead of wrapping
  if (map_out && (int64_t)nx + ny + nz - 3 > 65535) return HDM_E_INVALID;
  Ctx *c = nullptr;
  int rc = get_ctx(&c);
  if (rc != HDM_OK) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  // NULL = the legacy default stream, so the call is ordered after work the
  // caller queued on it or on any blocking stream (torch's default stream)
  hipStream_t s = (hipStream_t)opts.stream;
  // the previous call's launch may still be leaving on its stream: wait for
  // it before work goes to another stream, or a buffer it uses is replaced
  if (c->has_pending && c->pending != s && (rc = c->settle()) != HDM_OK) return rc;
  const TileLaunch *tl = hdm_tile_launch(tile);
  const TileShape ts = tile_shape(tile);
  const PairGeom g = pair_geom(nx, ny, nz, ts);
  if ((uint64_t)g.ntiles * 2 > 0xFFFFFFFFull) return HDM_E_INVALID;
  const uint32_t qcap = 2u * g.ntiles;
  const CommonLayout L = common_layout(1, qcap, g.hist_len);
  const uint64_t nvox = (uint64_t)nx * ny * nz;
  const size_t esz = dtype == HDM_DTYPE_U8 ? 1 : 4;
  // the near launch (section of the fused path below) and its region marks
  const bool near_path = fused_enabled() && conn == 6 && tl->fused_near && ds_enabled() &&
                         tb_enabled() && near_enabled() && !map_out &&
                         g.nregions <= kNearMaxRegions;
  // two bitmap sets of 2 * regions bits each
  const size_t marks_bytes = near_path ? (size_t)2 * 4 * ((2 * (size_t)g.nregions + 31) / 32) : 0;
  const bool grow = L.total + g.total > c->ws.cap || marks_bytes > c->marks.cap ||
                    (opts.ptr_kind == HDM_PTR_HOST && 2 * align256(nvox * esz) > c->in.cap) ||
                    (map_out && opts.ptr_kind == HDM_PTR_HOST && nvox * 2 > c->map.cap) ||
                    (opts.timing == 2 && (size_t)g.hist_len * 64 > c->trace.cap);
  if (grow && (rc = c->settle()) != HDM_OK) return rc;
  if ((rc = c->ws.ensure(L.total + g.total)) != HDM_OK) return rc;
  if (marks_bytes > c->marks.cap) {
    if ((rc = c->marks.ensure(marks_bytes)) != HDM_OK) return rc;
    // fresh memory: both sets start empty
    if (hipMemsetAsync(c->marks.p, 0, c->marks.cap, s) != hipSuccess) return HDM_E_HIP;
    c->marks_set = 0;
  }
  char *base = (char *)c->ws.p;
  GroupPtrs P = carve_common(base, L);
  PairDesc d{};
  const void *gd = gold, *sd = segm;
  if (opts.ptr_kind == HDM_PTR_HOST) {
    if ((rc = c->in.ensure(2 * align256(nvox * esz))) != HDM_OK) return rc;
    char *ib = (char *)c->in.p;
    if (hipMemcpyAsync(ib, gold, nvox * esz, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(ib + align256(nvox * esz), segm, nvox * esz, hipMemcpyHostToDevice, s) !=
            hipSuccess)
      return HDM_E_HIP;
    gd = ib, sd = ib + align256(nvox * esz);
  }
  fill_desc(d, g, gd, sd, dtype, label, opts.robustness, opts.crop ? 1 : 0, tile, 0, conn);
  carve_pair(d, g, base + L.total);
  if (map_out) {
    if (opts.ptr_kind == HDM_PTR_DEVICE) {
      d.map = map_out;
    } else {
      if ((rc = c->map.ensure(nvox * 2)) != HDM_OK) return rc;
      d.map = (uint16_t *)c->map.p;
    }
  }
  const PackIndex pi{0, 0};
  LaunchCfg cfg = launch_cfg(c, tile, g.row_words, g.ntiles, opts.grid_ctas, opts.chunk, 1);
  hipEvent_t ring_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int pair = -1;
  if (opts.timing == 3) {  // a fresh pair of the ring brackets the dominant launch
    if (c->ring.empty()) {
      c->ring.assign(2 * Ctx::kRing, nullptr);
      for (auto &ev : c->ring)
        if (hipEventCreate(&ev) != hipSuccess) {
          (void)hipGetLastError();
          for (auto &x : c->ring)
            if (x) (void)hipEventDestroy(x);
          c->ring.clear();
          return HDM_E_HIP;
        }
    }
    pair = (int)(c->ring_next++ % Ctx::kRing);
    ring_ev[2] = c->ring[2 * pair], ring_ev[3] = c->ring[2 * pair + 1];
    cfg.ev = ring_ev, cfg.ev_mode = 3;
  } else if (opts.timing) {
    cfg.ev = c->ev, cfg.ev_mode = opts.timing;
  }
  const size_t trace_bytes = (size_t)g.hist_len * 8 * sizeof(uint64_t);
  if (opts.timing == 2) {  // per-iteration device trace (diagnostic, not graph-safe)
    if ((rc = c->trace.ensure(trace_bytes)) != HDM_OK) return rc;
    if (hipMemsetAsync(c->trace.p, 0, trace_bytes, s) != hipSuccess) return HDM_E_HIP;
    cfg.trace = (uint64_t *)c->trace.p;
  }
  // the dilation kernel writes the result and the engine head straight into
  // the pinned host block; a status the kernel never writes marks "not run"
  char *ho = (char *)c->host_out;
  PairOut *hout = (PairOut *)ho;
  EngineHead *hhead = (EngineHead *)(ho + 256);
  hout->status = 1;
  cfg.outs = hout;
  cfg.head = hhead;
  // the result can be taken from the ready marker unless the call also
  // copies a map or histograms out, or times itself with events
  volatile uint32_t *ready = (volatile uint32_t *)(ho + 1024);
  const bool by_marker = !(map_out && opts.ptr_kind == HDM_PTR_HOST) &&
                         !(opts.hist_out && opts.hist_len > 0) &&
                         (opts.timing == 0 || opts.timing == 3);
  if (++c->seq == 0) ++c->seq;
  if (by_marker) cfg.ready = (uint32_t *)ready, cfg.ready_seq = c->seq;
  hipError_t e = hipSuccess;
  const bool fused = fused_enabled();
  const uint32_t ebase = c->ebase;
  bool ds = false;  // two dilation steps per barrier
  bool tb = false;  // then the region launch (eight steps per barrier)
  bool handed = false;  // the first launch handed the pair to the region launch
  if (fused) {
    // one launch: reset, pack, seed and dilation.  Its barrier tags use
    // epochs above every earlier launch's on this workspace; a new workspace
    // or a restarting base gets its engine cleared first.
    if (c->ebase_ws != c->ws.p || ebase > (1u << 23)) {
      if (hipMemsetAsync(P.engine, 0, sizeof(Engine), s) != hipSuccess) return HDM_E_HIP;
      if (c->marks.p && hipMemsetAsync(c->marks.p, 0, c->marks.cap, s) != hipSuccess)
        return HDM_E_HIP;
      c->ebase = 0;
      c->ebase_ws = c->ws.p;
    }
    if (d.map && hipMemsetAsync(d.map, 0, nvox * 2, s) != hipSuccess) return HDM_E_HIP;
    DilateArgs a{};
    a.descs = P.descs;
    a.ctrs = P.ctrs;
    a.npairs = 1;
    a.max_iters = INT_MAX;
    a.es = P.engine;
    a.q = P.q;
    a.qcap = qcap;
    a.chunk = cfg.chunk;
    a.spin = cfg.spin;
    a.qstats = P.qstats;
    a.qstats_len = g.hist_len;
    a.flags = barrier_mode();
    a.trace = cfg.trace;
    a.outs = hout;
    a.head = hhead;
    a.desc = d;
    a.total_tiles = g.ntiles;
    a.ebase = c->ebase;
    a.ready = cfg.ready;
    a.ready_seq = cfg.ready_seq;
    a.acc_set = c->acc_set;
    a.coop = coop_enabled() ? 1 : 0;
    a.tb_active = c->cus;
    ds = tl->fused_ds && conn == 6 && ds_enabled();
    tb = ds && tl->fused_tb && tb_enabled();
    a.tb_mode = tb ? tb_mode() : 0;
    a.pack_pools = pack_pools();
    if (near_path) {  // fills one set, zeroes the other for the next near launch
      const uint32_t set_words = (uint32_t)(c->marks.cap / 4 / 2);
      a.marks = (uint32_t *)c->marks.p + (size_t)c->marks_set * set_words;
      a.marks_zero = (uint32_t *)c->marks.p + (size_t)(c->marks_set ^ 1) * set_words;
      a.marks_zero_words = set_words;
      a.desc.marks = a.marks;  // the pack marks through the descriptor
    }
    if (cfg.ev) e = hipEventRecord(cfg.ev_mode == 3 ? cfg.ev[2] : cfg.ev[0], s);
    if (e == hipSuccess) {
      if (near_path) {
        e = tl->fused_near(dim3(near_grid_of(c->occ_fused_near[tile], c->cus, opts.grid_ctas,
                                              g.ntiles)),
                           s, a);
        if (e == hipSuccess) c->marks_set ^= 1;
      } else
        e = (ds ? tl->fused_ds : tl->fused)(
            dim3(fused_grid(c, tile, opts.grid_ctas, ds, g.ntiles)), s, a);
      if (e == hipSuccess) c->acc_set ^= 1;  // the launch zeroes the other set first
    }
    if (e == hipSuccess && cfg.ev) e = hipEventRecord(cfg.ev[3], s);
    if (e == hipSuccess && tb) {
      // the near (or two-step) launch either computes the pair or, after its
      // pack (or its near pass), hands it over (status 2): then the region
      // launch follows on the stream, and the end event is recorded again
      // behind it
      if (by_marker) {
        if ((rc = wait_ready(s, ready, c->seq)) != HDM_OK) return rc;
      } else if ((e = hipStreamSynchronize(s)) != hipSuccess) {
        return map_hip(e);
      }
      if (__atomic_load_n(&hout->status, __ATOMIC_ACQUIRE) == 2) {
        hout->status = 1;
        handed = true;
        if (by_marker) {
          if (++c->seq == 0) ++c->seq;
          a.ready_seq = c->seq;
        }
        // a traced call: the region launch's per-period records start clear
        // (the first launch wrote per-workgroup records into the same buffer)
        if (cfg.trace) e = hipMemsetAsync(cfg.trace, 0, trace_bytes, s);
        if (e == hipSuccess)
            e = tl->fused_tb(
              dim3(tb_grid(c->occ_fused_tb[tile], c->cus, opts.grid_ctas, g.ntiles, g.nregions)),
              s, a);
        if (e == hipSuccess && cfg.ev) e = hipEventRecord(cfg.ev[3], s);
      }
    }
    // until the result is read, assume the launch used every epoch it could
    c->ebase = a.ebase + 8 + (uint32_t)g.hist_len;
  } else {
    c->ebase_ws = nullptr;  // the reset kernel clears the tagged words
    e = enqueue_group(s, tl, P, 1, &d, &pi, g.row_words, g.ntiles, qcap, cfg, g.hist_len);
  }
  if (e != hipSuccess) return map_hip(e);
  if (map_out && opts.ptr_kind == HDM_PTR_HOST &&
      hipMemcpyAsync(map_out, d.map, nvox * 2, hipMemcpyDeviceToHost, s) != hipSuccess)
    return HDM_E_HIP;
  if (opts.hist_out && opts.hist_len > 0) {
    const int n = std::min(opts.hist_len, g.hist_len);
    if (hipMemcpyAsync(opts.hist_out, d.hist[0], n * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(opts.hist_out + opts.hist_len, d.hist[1], n * 4, hipMemcpyDeviceToHost,
                       s) != hipSuccess)
      return HDM_E_HIP;
  }
  if (by_marker) {
    if ((rc = wait_ready(s, ready, c->seq)) != HDM_OK) return rc;
    c->pending = s, c->has_pending = true;
  } else {
    c->has_pending = false;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return map_hip(e);
  }
  PairOut o;
  EngineHead es{};
  std::memcpy(&o, hout, sizeof(o));
  std::memcpy(&es, hhead, sizeof(es));
  if (o.status == 1) return HDM_E_HIP;
  // the body that ran (eight steps: the pair was handed to the region launch)
  c->last_ms[6] = (float)std::max(1, o.steps);
  if (fused) {
    c->ebase = c->ebase - 8 - (uint32_t)g.hist_len + 4 + es.t;  // the epochs it used
    c->last_ms[5] = handed ? 2.f : 1.f;  // launches of the fused path
    if (opts.timing && opts.timing != 3) {
      float all = -1.f;
      if (hipEventElapsedTime(&all, c->ev[0], c->ev[3]) != hipSuccess) all = -1.f;
      (void)hipGetLastError();
      {  // phases from the launch's own clock stamps
        const double tick_ms = 1.0 / c->clock_khz;
        c->last_ms[0] = (float)((es.phase[1] - es.phase[0]) * tick_ms);
        c->last_ms[1] = (float)((es.phase[2] - es.phase[1]) * tick_ms);
        c->last_ms[2] = (float)((es.phase[3] - es.phase[2]) * tick_ms);
        c->last_ms[3] = 0.f;
        c->last_ms[4] = all;
      }
    }
  } else if (opts.timing == 3) {
    c->last_ms[5] = 0.f;
  } else if (opts.timing) {
    // phases: pack (+ reset), metadata, dilation (+ finalize); total
    float ms[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
      if (hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]) != hipSuccess) ms[i] = -1.f;
    float all = 0.f;
    if (hipEventElapsedTime(&all, c->ev[0], c->ev[3]) != hipSuccess) all = -1.f;
    for (int i = 0; i < 3; ++i) c->last_ms[i] = ms[i];
    c->last_ms[3] = 0.f;
    c->last_ms[4] = all;
    c->last_ms[5] = 0.f;
    (void)hipGetLastError();
  }
  if (opts.timing == 3) {  // read when asked: the dilation, or the whole fused launch
    for (int i = 0; i < 5; ++i) c->last_ms[i] = -1.f;
    c->last_pair = pair;
    c->last_pair_fused = fused;
    if (c->log.size() >= Ctx::kRing) c->log.erase(c->log.begin());
    c->log.push_back((uint32_t)pair);
  }
  if (opts.timing == 2) {
    const size_t nrec = std::min<size_t>(es.t, (size_t)g.hist_len);
    c->last_trace.assign(nrec * 8, 0);
    if (nrec && hipMemcpy(c->last_trace.data(), c->trace.p, nrec * 8 * sizeof(uint64_t),
                          hipMemcpyDeviceToHost) != hipSuccess)
      return HDM_E_HIP;
  }
  fill_result(out, o, &es, g.ntiles, o.iters);
  if (o.status == HDM_OK && !o.done) return HDM_E_STUCK;  // queue drained first
  return o.status;
}

extern "C" int hdm_hausdorff_u8(const uint8_t *gold, const uint8_t *segm, int nx, int ny, int nz,
                                uint8_t label, const hdm_opts *opts, hdm_result *out) {
  return hausdorff_core(gold, segm, HDM_DTYPE_U8, label, nx, ny, nz, opts, nullptr, out);
}

extern "C" int hdm_hausdorff_i32(const int32_t *gold, const int32_t *segm, int nx, int ny,
                                 int nz, int32_t label, const hdm_opts *opts, hdm_result *out) {
  return hausdorff_core(gold, segm, HDM_DTYPE_I32, label, nx, ny, nz, opts, nullptr, out);
}

extern "C" int hdm_hausdorff_bool(const uint8_t *img1, const uint8_t *img2, int nx, int ny,
                                  int nz, const hdm_opts *opts, hdm_result *out) {
  // bool volumes hold 0 or 1 (forBench/Volume.h:35-38 allocates them with calloc)
  return hausdorff_core(img1, img2, HDM_DTYPE_U8, 1, nx, ny, nz, opts, nullptr, out);
}

extern "C" int hdm_first_hit_map_u8(const uint8_t *gold, const uint8_t *segm, int nx, int ny,
                                    int nz, uint8_t label, const hdm_opts *opts, uint16_t *map,
                                    hdm_result *out) {
  if (!map) return HDM_E_INVALID;
  return hausdorff_core(gold, segm, HDM_DTYPE_U8, label, nx, ny, nz, opts, map, out);
}

// ---------------------------------------------------------------- batch
struct hdm_batch {
  Ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  int n = 0, tile = 0, grid = 0;
  const TileLaunch *tl = nullptr;
  TileShape ts{};
  LaunchCfg cfg{};
  uint64_t total_rw = 0, total_tiles = 0;
  uint32_t qcap = 0;
  int qstats_len = 0;
  void *ws = nullptr;
  GroupPtrs P{};
  std::vector<PairGeom> geoms;
  bool graph = false;
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void *host = nullptr;
  float last_ms = 0.f;
  uint32_t near_entries = 0;  // of the last synced run
};

static int batch_enqueue(hdm_batch *b) {
  return map_hip(enqueue_group(b->s, b->tl, b->P, b->n, nullptr, nullptr, b->total_rw,
                               b->total_tiles, b->qcap, b->cfg, b->qstats_len));
}

extern "C" void hdm_batch_destroy(hdm_batch *b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->g) (void)hipGraphDestroy(b->g);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->ws) (void)hipFree(b->ws);
  if (b->host) (void)hipHostFree(b->host);
  if (b->s) (void)hipStreamDestroy(b->s);
  (void)hipGetLastError();
  delete b;
}

extern "C" int hdm_batch_create(const hdm_pair *pairs, int n, const hdm_opts *opts_in,
                                int use_graph, hdm_batch **out) {
  if (!pairs || n < 1 || n > 65535 || !out) return HDM_E_INVALID;
  *out = nullptr;
  hdm_opts opts;
  hdm_opts_init(&opts);
  if (opts_in) opts = *opts_in;
  int conn = 6;
  const int lid = select_launch(opts.tile, opts.connectivity, &conn, ds_enabled());
  if (!(opts.robustness >= 0.0 && opts.robustness <= 1.0) || lid < 0 ||
      opts.ptr_kind != HDM_PTR_DEVICE)
    return HDM_E_INVALID;
  for (int i = 0; i < n; ++i) {
    const hdm_pair &p = pairs[i];
    if (!p.gold || !p.segm || p.nx < 1 || p.ny < 1 || p.nz < 1 || p.nx > 65535 ||
        p.ny > 65535 || p.nz > 65535 || (p.dtype != HDM_DTYPE_U8 && p.dtype != HDM_DTYPE_I32))
      return HDM_E_INVALID;
  }
  Ctx *c = nullptr;
  int rc = get_ctx(&c);
  if (rc != HDM_OK) return rc;
  hdm_batch *b = new hdm_batch();
  b->ctx = c;
  b->n = n;
  b->tile = lid;
  b->tl = hdm_tile_launch(b->tile);
  b->ts = tile_shape(b->tile);
  b->grid = resident_grid(c, b->tile, opts.grid_ctas, n);
  std::vector<PairDesc> descs(n);
  std::vector<PackIndex> idx(n);
  uint64_t pair_bytes = 0, qcap = 0;
  int maxh = 0;
  // the near pass for the 64x8x8 tile with the 6 element at two steps per
  // iteration (HDM_NEAR=0 turns it off): each pair's region bitmap, and a list
  // with room for every bit of every bitmap
  const bool near_batch = b->tl->fused_near && conn == 6 && near_enabled();
  uint64_t mark_words = 0, mark_bits = 0;
  for (int i = 0; i < n; ++i) {
    const hdm_pair &p = pairs[i];
    const PairGeom g = pair_geom(p.nx, p.ny, p.nz, b->ts);
    b->geoms.push_back(g);
    idx[i] = PackIndex{b->total_rw, b->total_tiles, mark_words};
    b->total_rw += g.row_words;
    b->total_tiles += g.ntiles;
    qcap += 2ull * g.ntiles;
    pair_bytes += g.total;
    maxh = std::max(maxh, g.hist_len);
    if (near_batch) {
      mark_words += near_mark_words(g.nregions);
      mark_bits += 2ull * g.nregions;
    }
  }
  if (mark_bits > 0xFFFFFFFFull) {
    delete b;
    return HDM_E_INVALID;
  }
  const uint64_t near_bytes = near_batch ? align256(4 * mark_words) + align256(8 * mark_bits) : 0;
  if (qcap > 0xFFFFFFFFull || b->total_tiles > 0xFFFFFFFFull) {
    delete b;
    return HDM_E_INVALID;
  }
  b->qcap = (uint32_t)qcap;
  b->qstats_len = maxh;
  b->cfg = launch_cfg(c, b->tile, b->total_rw, b->total_tiles, opts.grid_ctas, opts.chunk, n);
  b->grid = b->cfg.grid;
  const CommonLayout L = common_layout(n, qcap, maxh);
  if (hipMalloc(&b->ws, L.total + pair_bytes + near_bytes) != hipSuccess) {
    (void)hipGetLastError();
    b->ws = nullptr;
    hdm_batch_destroy(b);
    return HDM_E_NOMEM;
  }
  b->P = carve_common((char *)b->ws, L);
  char *pb = (char *)b->ws + L.total;
  for (int i = 0; i < n; ++i) {
    const hdm_pair &p = pairs[i];
    fill_desc(descs[i], b->geoms[i], p.gold, p.segm, p.dtype, p.label, opts.robustness,
              opts.crop ? 1 : 0, b->tile, 0, conn);
    carve_pair(descs[i], b->geoms[i], pb);
    pb += b->geoms[i].total;
  }
  if (near_batch) {  // [bitmaps of every pair][entry list]
    uint32_t *mk = (uint32_t *)pb;
    for (int i = 0; i < n; ++i) descs[i].marks = mk + idx[i].mark_base;
    b->cfg.near_list = (uint64_t *)(pb + align256(4 * mark_words));
    b->cfg.near_words = mark_words;
    b->cfg.near_cap = (uint32_t)mark_bits;
    b->cfg.near_grid = std::min(kMaxGrid, std::max(1, c->occ_fused_near[b->tile]) * c->cus);
  }
  // a blocking stream: a run is ordered after work on the legacy default stream
  // (where callers such as torch usually produced the input volumes)
  if (hipStreamCreate(&b->s) != hipSuccess ||
      hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess ||
      hipHostMalloc(&b->host, sizeof(PairOut) * n + 256, hipHostMallocDefault) != hipSuccess ||
      hipMemcpy(b->P.descs, descs.data(), sizeof(PairDesc) * n, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->P.idx, idx.data(), sizeof(PackIndex) * n, hipMemcpyHostToDevice) !=
          hipSuccess) {
    hdm_batch_destroy(b);
    return HDM_E_HIP;
  }
  if (use_graph) {
    if (hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      hdm_batch_destroy(b);
      return HDM_E_HIP;
    }
    rc = batch_enqueue(b);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->s, &g);
    b->g = g;
    if (rc != HDM_OK || ee != hipSuccess ||
        hipGraphInstantiate(&b->exec, b->g, nullptr, nullptr, 0) != hipSuccess) {
      hdm_batch_destroy(b);
      return rc != HDM_OK ? rc : HDM_E_HIP;
    }
    b->graph = true;
  }
  *out = b;
  return HDM_OK;
}

extern "C" int hdm_batch_run(hdm_batch *b) {
  if (!b) return HDM_E_INVALID;
  if (hipEventRecord(b->e0, b->s) != hipSuccess) return HDM_E_HIP;
  int rc = HDM_OK;
  if (b->graph)
    rc = map_hip(hipGraphLaunch(b->exec, b->s));
  else
    rc = batch_enqueue(b);
  if (rc != HDM_OK) return rc;
  if (hipEventRecord(b->e1, b->s) != hipSuccess) return HDM_E_HIP;
  return HDM_OK;
}

extern "C" int hdm_batch_sync(hdm_batch *b, hdm_result *out) {
  if (!b) return HDM_E_INVALID;
  hipError_t e = hipStreamSynchronize(b->s);
  if (e != hipSuccess) return map_hip(e);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, b->e0, b->e1) == hipSuccess) b->last_ms = ms;
  (void)hipGetLastError();
  b->near_entries = 0;
  if (b->cfg.near_list &&
      hipMemcpy(&b->near_entries, &b->P.engine->near_count, sizeof(uint32_t),
                hipMemcpyDeviceToHost) != hipSuccess)
    return HDM_E_HIP;
  if (!out) return HDM_OK;
  if (hipMemcpy(b->host, b->P.outs, sizeof(PairOut) * b->n, hipMemcpyDeviceToHost) !=
          hipSuccess ||
      hipMemcpy((char *)b->host + sizeof(PairOut) * b->n, b->P.engine, kEngineHostBytes,
                hipMemcpyDeviceToHost) != hipSuccess)
    return HDM_E_HIP;
  EngineHead es{};
  std::memcpy(&es, (char *)b->host + sizeof(PairOut) * b->n, kEngineHostBytes);
  int rc = HDM_OK;
  for (int i = 0; i < b->n; ++i) {
    PairOut o;
    std::memcpy(&o, (char *)b->host + sizeof(PairOut) * i, sizeof(o));
    fill_result(&out[i], o, &es, b->geoms[i].ntiles, std::max(o.hd_fp, o.hd_fn));
    const int st = (o.status == HDM_OK && !o.done) ? HDM_E_STUCK : o.status;
    if (st != HDM_OK && rc == HDM_OK) rc = st;
  }
  return rc;
}

extern "C" float hdm_batch_last_ms(hdm_batch *b) { return b ? b->last_ms : 0.f; }
extern "C" uint32_t hdm_batch_near_entries(hdm_batch *b) { return b ? b->near_entries : 0u; }

extern "C" int hdm_hausdorff_batch(const hdm_pair *pairs, int n, const hdm_opts *opts,
                                   hdm_result *out) {
  hdm_batch *b = nullptr;
  int rc = hdm_batch_create(pairs, n, opts, 0, &b);
  if (rc != HDM_OK) return rc;
  rc = hdm_batch_run(b);
  if (rc == HDM_OK) rc = hdm_batch_sync(b, out);
  hdm_batch_destroy(b);
  return rc;
}

// ---------------------------------------------------------------- autotuner
extern "C" int hdm_autotune(const hdm_pair *pairs, int n, const int32_t *tiles, int n_tiles,
                            const int32_t *chunks, int n_chunks, const int32_t *grids,
                            int n_grids, int reps, hdm_tune_entry *out, int out_cap,
                            int *n_out) {
  if (!pairs || n < 1 || !out || out_cap < 1 || !n_out || reps < 1) return HDM_E_INVALID;
  static const int32_t def_tiles[] = {HDM_TILE_64x8x8, HDM_TILE_64x4x4, HDM_TILE_64x8x4,
                                      HDM_TILE_32x8x8};
  static const int32_t def_chunks[] = {1, 2, 4};
  static const int32_t def_grids[] = {0};
  if (!tiles || n_tiles < 1) tiles = def_tiles, n_tiles = 4;
  if (!chunks || n_chunks < 1) chunks = def_chunks, n_chunks = 3;
  if (!grids || n_grids < 1) grids = def_grids, n_grids = 1;
  std::vector<hdm_result> ref(n), cur(n);
  bool have_ref = false;
  int k = 0;
  for (int ti = 0; ti < n_tiles; ++ti)
    for (int ci = 0; ci < n_chunks; ++ci)
      for (int gi = 0; gi < n_grids; ++gi) {
        if (k >= out_cap) break;
        hdm_opts o;
        hdm_opts_init(&o);
        o.tile = tiles[ti];
        o.chunk = chunks[ci];
        o.grid_ctas = grids[gi];
        hdm_tune_entry &E = out[k++];
        E.tile = tiles[ti], E.chunk = chunks[ci], E.grid_ctas = grids[gi];
        E.ms = 0.f;
        hdm_batch *b = nullptr;
        int rc = hdm_batch_create(pairs, n, &o, 1, &b);
        if (rc == HDM_OK) rc = hdm_batch_run(b);
        if (rc == HDM_OK) rc = hdm_batch_sync(b, cur.data());
        std::vector<float> ms;
        for (int r = 0; rc == HDM_OK && r < reps; ++r) {
          rc = hdm_batch_run(b);
          if (rc == HDM_OK) rc = hdm_batch_sync(b, nullptr);
          ms.push_back(hdm_batch_last_ms(b));
        }
        if (b) E.grid_ctas = b->grid;
        hdm_batch_destroy(b);
        if (rc == HDM_OK) {
          if (!have_ref) {
            ref = cur;
            have_ref = true;
          } else {
            for (int i = 0; i < n; ++i)
              if (cur[i].hd != ref[i].hd || cur[i].sum != ref[i].sum ||
                  cur[i].count != ref[i].count)
                rc = HDM_E_STUCK;  // candidates must agree bit for bit
          }
          std::sort(ms.begin(), ms.end());
          E.ms = ms.empty() ? 0.f : ms[ms.size() / 2];
        }
        E.status = rc;
      }
  std::stable_sort(out, out + k, [](const hdm_tune_entry &a, const hdm_tune_entry &b) {
    if ((a.status == HDM_OK) != (b.status == HDM_OK)) return a.status == HDM_OK;
    return a.ms < b.ms;
  });
  *n_out = k;
  return HDM_OK;
}

extern "C" int hdm_last_trace(uint64_t *out, int max_iters, double *tick_ns) {
  Ctx *c = nullptr;
  if (get_ctx(&c) != HDM_OK) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  if (tick_ns) *tick_ns = 1e6 / c->clock_khz;
  const int n = std::min<int>(max_iters, (int)(c->last_trace.size() / 8));
  if (out && n > 0) std::memcpy(out, c->last_trace.data(), (size_t)n * 8 * sizeof(uint64_t));
  return n;
}

extern "C" int hdm_last_timings(float *ms, int n) {
  if (!ms || n < 1) return 0;
  Ctx *c = nullptr;
  if (get_ctx(&c) != HDM_OK) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->last_pair >= 0) {  // a timing-3 call: wait for its pair now
    float t = -1.f;
    hipEvent_t *p = &c->ring[2 * c->last_pair];
    if (hipEventSynchronize(p[1]) != hipSuccess || hipEventElapsedTime(&t, p[0], p[1]) != hipSuccess)
      t = -1.f;
    (void)hipGetLastError();
    c->last_ms[c->last_pair_fused ? 4 : 2] = t;
    c->last_pair = -1;
  }
  const int k = std::min(n, 7);
  for (int i = 0; i < k; ++i) ms[i] = c->last_ms[i];
  return k;
}

extern "C" int hdm_timing_log(float *ms, int n) {
  if (!ms || n < 1) return 0;
  Ctx *c = nullptr;
  if (get_ctx(&c) != HDM_OK) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  int k = 0;
  for (uint32_t pr : c->log) {
    if (k >= n) break;
    float t = -1.f;
    hipEvent_t *p = &c->ring[2 * pr];
    if (hipEventSynchronize(p[1]) != hipSuccess || hipEventElapsedTime(&t, p[0], p[1]) != hipSuccess)
      t = -1.f;
    (void)hipGetLastError();
    ms[k++] = t;
  }
  c->log.clear();
  return k;
}
