// hdm_hausdorff_u8_noalloc: the single-pair entry point for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).  It runs the same launches as
// hdm_hausdorff_u8, but its whole device workspace -- descriptor, counters,
// engine state, queue, packed masks, dilation buffers, region marks, results -- lives in
// `scratch`, a file-scope device array of kScratchBytes, and the results come
// back with a stream-ordered copy into the caller's hdm_result.  No call here
// allocates: no hipMalloc / hipFree, no pinned host memory, no events, no
// library context.
//
// The first call on a device (with that device current) takes the array's
// device address with hipGetSymbolAddress and the launch shapes from the
// occupancy API, once; every later call uses that address.  The array's name
// is a host-side stand-in and is never passed to a kernel or a copy.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents (the engine words are tagged with epochs that grow from call to
// call, so nothing has to be cleared between calls).  Calls from the threads
// of one process are serialized per device (a mutex held until the call's
// stream is done); two processes each hold their own array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/hdm.h"
#include "hdm_host.h"

using namespace hdm;

namespace {

constexpr uint64_t kScratchBytes = 64ull << 20;  // HDM_NOALLOC_MAX_BYTES
constexpr int kMaxDevices = 64;
// the near launch's two region-mark bitmaps, at a fixed place (the end of the
// array), so the set one call zeroes is the set the next call fills whatever
// the pair sizes: 2 x kNearMaxRegions bits per set
constexpr uint64_t kMarkSetWords = 2ull * kNearMaxRegions / 32;
constexpr uint64_t kMarkBytes = 2 * kMarkSetWords * 4;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NoAllocDev {
  std::mutex mu;
  bool ready = false;
  unsigned char *ptr = nullptr;  // this device's copy of scratch
  int cus = 0;
  int occ_fused[kLaunchIds] = {};
  int occ_fused_ds[kLaunchIds] = {};
  int occ_fused_tb[kLaunchIds] = {};
  int occ_fused_near[kLaunchIds] = {};
  int marks_set = 0;   // the mark bitmap the next near launch fills
  double clock_khz = 0;
  uint32_t ebase = 0;  // epoch base of the next launch's barrier tags
  int acc_set = 0;     // the pack-partial set the next launch packs into
  bool cleared = false;
};

NoAllocDev g_dev[kMaxDevices];

// Set-up on the current device, once: the array's address and launch shapes.
int setup(NoAllocDev &d, int dev) {
  if (d.ready) return HDM_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return HDM_E_HIP;
  }
  int cus = 0, khz = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) {
    (void)hipGetLastError();
    return HDM_E_HIP;
  }
  for (int k = 1; k < kLaunchIds; ++k) {
    const TileLaunch *tl = hdm_tile_launch(k);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tl->fused_fn, kWgThreads, 0) !=
            hipSuccess ||
        nb < 1)
      return HDM_E_HIP;
    d.occ_fused[k] = nb;
    if (tl->fused_ds_fn) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tl->fused_ds_fn, kWgThreads, 0) !=
              hipSuccess ||
          nb < 1)
        return HDM_E_HIP;
      d.occ_fused_ds[k] = nb;
    }
    if (tl->fused_tb_fn) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tl->fused_tb_fn, kWgThreads, 0) !=
              hipSuccess ||
          nb < 1)
        return HDM_E_HIP;
      d.occ_fused_tb[k] = nb;
    }
    if (tl->fused_near_fn) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tl->fused_near_fn, kWgThreads, 0) !=
              hipSuccess ||
          nb < 1)
        return HDM_E_HIP;
      d.occ_fused_near[k] = nb;
    }
  }
  d.ptr = (unsigned char *)p;
  d.cus = cus;
  d.clock_khz = khz > 0 ? khz : 100000;
  d.ready = true;
  return HDM_OK;
}

}  // namespace

extern "C" uint64_t hdm_noalloc_max_bytes(void) { return kScratchBytes; }

extern "C" int hdm_hausdorff_u8_noalloc(const uint8_t *gold, const uint8_t *segm, int nx, int ny,
                                        int nz, uint8_t label, const hdm_opts *opts_in,
                                        hdm_result *out) {
  hdm_opts opts;
  hdm_opts_init(&opts);
  if (opts_in) opts = *opts_in;
  if (!gold || !segm || !out || nx < 1 || ny < 1 || nz < 1 || nx > 65535 || ny > 65535 ||
      nz > 65535)
    return HDM_E_INVALID;
  if (!(opts.robustness >= 0.0 && opts.robustness <= 1.0)) return HDM_E_INVALID;
  // device inputs only (a host input needs a staging copy), no event timing
  if (opts.ptr_kind != HDM_PTR_DEVICE || opts.timing != 0) return HDM_E_INVALID;
  if (opts.map_out && (int64_t)nx + ny + nz - 3 > 65535) return HDM_E_INVALID;
  int conn = 6;
  const int tile = select_launch(opts.tile, opts.connectivity, &conn);
  if (tile < 0) return HDM_E_INVALID;
  const TileLaunch *tl = hdm_tile_launch(tile);
  const TileShape ts = tile_shape(tile);
  const PairGeom g = pair_geom(nx, ny, nz, ts);
  if ((uint64_t)g.ntiles * 2 > 0xFFFFFFFFull) return HDM_E_INVALID;
  const uint32_t qcap = 2u * g.ntiles;
  const CommonLayout L = common_layout(1, qcap, g.hist_len);
  // [common][pair][PairOut][EngineHead] ... [mark bitmaps at the end]
  const uint64_t off_out = align256(L.total + g.total);
  const uint64_t off_head = off_out + align256(sizeof(PairOut));
  const uint64_t off_marks = kScratchBytes - kMarkBytes;
  if (off_head + align256(sizeof(EngineHead)) > off_marks) return HDM_E_CAPACITY;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return HDM_E_NODEVICE;
  }
  NoAllocDev &D = g_dev[dev];
  std::lock_guard<std::mutex> lk(D.mu);
  int rc = setup(D, dev);
  if (rc != HDM_OK) return rc;
  hipStream_t s = (hipStream_t)opts.stream;
  unsigned char *base = D.ptr;
  GroupPtrs P = carve_common((char *)base, L);
  PairDesc d{};
  fill_desc(d, g, gold, segm, HDM_DTYPE_U8, label, opts.robustness, opts.crop ? 1 : 0, tile, 0,
            conn);
  carve_pair(d, g, (char *)base + L.total);
  d.map = opts.map_out;
  PairOut *dout = (PairOut *)(base + off_out);
  EngineHead *dhead = (EngineHead *)(base + off_head);
  // the engine (at a fixed offset: the common layout of one pair starts with
  // fixed-size regions) is cleared on the first call and when the epoch base
  // restarts; tagged words of earlier calls never match later epochs
  uint32_t *marks = (uint32_t *)(base + off_marks);
  if (!D.cleared || D.ebase > (1u << 23)) {
    if (hipMemsetAsync(P.engine, 0, sizeof(Engine), s) != hipSuccess) return HDM_E_HIP;
    if (hipMemsetAsync(marks, 0, kMarkBytes, s) != hipSuccess) return HDM_E_HIP;
    D.ebase = 0;
    D.acc_set = 0;
    D.marks_set = 0;
    D.cleared = true;
  }
  const uint64_t nvox = (uint64_t)nx * ny * nz;
  if (d.map && hipMemsetAsync(d.map, 0, nvox * 2, s) != hipSuccess) return HDM_E_HIP;
  // a head the kernel never wrote reads nwg = 0
  if (hipMemsetAsync(dhead, 0, sizeof(EngineHead), s) != hipSuccess) return HDM_E_HIP;
  DilateArgs a{};
  a.descs = P.descs;
  a.ctrs = P.ctrs;
  a.npairs = 1;
  a.max_iters = INT_MAX;
  a.es = P.engine;
  a.q = P.q;
  a.qcap = qcap;
  a.chunk = std::max(1, opts.chunk);
  a.spin = (uint64_t)(D.clock_khz * 1000.0 * 2.0);
  {
    const char *e = std::getenv("HDM_SPIN_TICKS");
    if (e && *e) a.spin = std::strtoull(e, nullptr, 10);
  }
  a.qstats = P.qstats;
  a.qstats_len = g.hist_len;
  a.flags = barrier_mode();
  a.outs = dout;
  a.head = dhead;
  a.desc = d;
  a.total_tiles = g.ntiles;
  a.ebase = D.ebase;
  a.acc_set = D.acc_set;
  a.coop = coop_enabled() ? 1 : 0;
  a.tb_active = D.cus;
  // the launches hdm_hausdorff_u8 makes: the near launch (pack and eight
  // steps around the regions with voxels to cover) or two steps per barrier
  // (else one), then the region launch, to which the first hands long pairs
  const bool ds = tl->fused_ds && conn == 6 && ds_enabled();
  const bool tb = ds && tl->fused_tb && tb_enabled();
  const bool near_path = tb && tl->fused_near && near_enabled() && !d.map &&
                         g.nregions <= kNearMaxRegions;
  a.tb_mode = tb ? tb_mode() : 0;
  a.pack_pools = pack_pools();
  if (near_path) {
    a.marks = marks + (size_t)D.marks_set * kMarkSetWords;
    a.marks_zero = marks + (size_t)(D.marks_set ^ 1) * kMarkSetWords;
    a.marks_zero_words = (uint32_t)kMarkSetWords;
    a.desc.marks = a.marks;  // the pack marks through the descriptor
  }
  const int grid = near_path ? near_grid_of(D.occ_fused_near[tile], D.cus, opts.grid_ctas, g.ntiles)
                             : fused_grid_of(ds ? D.occ_fused_ds[tile] : D.occ_fused[tile], D.cus,
                                             opts.grid_ctas, g.ntiles);
  hipError_t e = (near_path ? tl->fused_near : ds ? tl->fused_ds : tl->fused)(dim3(grid), s, a);
  if (e == hipSuccess) {
    D.acc_set ^= 1;  // the launch zeroes the other set for the next one
    if (near_path) D.marks_set ^= 1;
  }
  // until the head is read, assume the launches used every epoch they could
  D.ebase = a.ebase + 8 + (uint32_t)g.hist_len;
  if (e != hipSuccess) return map_hip(e);
  PairOut o;
  EngineHead h;
  if (tb) {
    // the near (or two-step) launch computes the pair or hands it over
    // (status 2): then the region launch follows
    if (hipMemcpyAsync(&o, dout, sizeof(o), hipMemcpyDeviceToHost, s) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return e != hipSuccess ? map_hip(e) : HDM_E_HIP;
    if (o.status == 2) {
      e = tl->fused_tb(
          dim3(tb_grid(D.occ_fused_tb[tile], D.cus, opts.grid_ctas, g.ntiles, g.nregions)), s,
          a);
      if (e != hipSuccess) return map_hip(e);
    }
  }
  if (hipMemcpyAsync(&o, dout, sizeof(o), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&h, dhead, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
    return HDM_E_HIP;
  if (opts.hist_out && opts.hist_len > 0) {
    const int n = std::min(opts.hist_len, g.hist_len);
    if (hipMemcpyAsync(opts.hist_out, d.hist[0], n * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(opts.hist_out + opts.hist_len, d.hist[1], n * 4, hipMemcpyDeviceToHost,
                       s) != hipSuccess)
      return HDM_E_HIP;
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return map_hip(e);
  if (h.nwg == 0) return HDM_E_HIP;  // the launch did not complete
  D.ebase = a.ebase + 4 + h.t;      // the epochs it used
  fill_result(out, o, &h, g.ntiles, o.iters);
  if (o.status == HDM_OK && !o.done) return HDM_E_STUCK;
  return o.status;
}
