// Serving engine: continuous batching over the persistent dilation kernel.
//
// No counterpart in the reference (one pair per blocking call, hello.cu:2025).
// Requests wait in a FIFO, are admitted into one of max_slots slots when a
// slot is free and the HBM arena (the budget) has room for their workspace,
// join the running engine at an iteration boundary (their descriptor carries
// the engine iteration at which they start, so every pair keeps its own
// iteration count), and retire as soon as both of their passes are done.
//
// Admission has two halves.  Staging -- the inputs' copy (host volumes), the
// slot's reset and the pack (or, for registered masks, the combine pass) --
// runs on the admission stream; it touches only the new pair's slot and
// arena block, and the slot's descriptor keeps an unused start iteration.
// Joining -- the start iteration and the seed into the engine's queue -- runs
// on the engine stream after the staging's event.  One step = stage and join
// what waits, launch one slice of slice_iters dilation steps, retire what it
// finished.  Staging never runs beside a slice: a pack on the admission
// stream during a slice held part of the slice's grid off the GPU, and its
// grid barrier waited out its 2 s bound (measured: slices of two steps at
// 5000 requests/s), so the kernels of both halves run between slices.
// A lone request (nothing in flight, nothing else waiting) skips the
// slots: it runs as the library's one-pair call on the engine stream.
//
// Registered masks: a volume packed once (hdm_server_register) stays in the
// arena as packed bits, face words and tile flags; requests naming two of
// them read those instead of re-packing their volumes.
//
// Admission groups (64x8x8 tiles, 6 element, near pass on): the waiting
// requests are staged together as one group on the admission stream -- one
// reset, one pack over the group's volume pairs (a combine per registered
// pair), the totals, the near pass and an outcome kernel, whose ready marker
// the step waits for.  A pair whose cuts the near pass reached within eight
// steps (most evaluation pairs) completes there and never takes a slot; the
// others move into free slots and join the engine as staged pairs do.  The
// group runs between slices, so no kernel of it runs beside a slice's grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/hdm.h"
#include "hdm_host.h"

using namespace hdm;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// first-fit allocator over [0, size) with 256-byte granules
struct Arena {
  uint64_t size = 0, used = 0;
  std::map<uint64_t, uint64_t> free;  // offset -> length
  void init(uint64_t n) {
    size = n;
    used = 0;
    free.clear();
    free[0] = n;
  }
  bool alloc(uint64_t n, uint64_t *off) {
    n = align256(n);
    for (auto it = free.begin(); it != free.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const uint64_t rest = it->second - n, at = it->first + n;
        free.erase(it);
        if (rest) free[at] = rest;
        used += n;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = align256(n);
    used -= n;
    auto it = free.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
};

constexpr uint32_t kUnused = 0xFFFFFFFFu;  // start iteration of a slot outside the engine

// A registered mask: [packed bits][face words][tile flags][descriptor][index]
// [counters] in one arena block; the last three serve its own pack only.
struct MaskReg {
  int nx = 0, ny = 0, nz = 0;
  uint64_t off = 0, bytes = 0;
  void *orig = nullptr;
  uint64_t *xf = nullptr;
  uint32_t *tflags = nullptr;
  int refs = 0;          // queued, staged and in-flight requests naming it
  bool released = false; // unregistered: freed when refs reaches 0
};

struct Request {
  uint64_t id = 0;
  const void *gold = nullptr, *segm = nullptr;
  uint64_t mg = 0, ms = 0;  // registered masks (0: the volumes above)
  int nx = 0, ny = 0, nz = 0, dtype = 0, label = 1, ptr_kind = HDM_PTR_DEVICE;
  double t_submit = 0, t_admit = 0;
  int slot = -1;
  bool active = false;  // joined the engine (else staged)
  uint64_t off = 0, bytes = 0;  // arena block: [inputs if copied][workspace]
  PairGeom geom{};
  hipEvent_t staged = nullptr;  // the staging's end on the admission stream
};

}  // namespace

struct hdm_server {
  hdm_server_opts o{};
  Ctx *ctx = nullptr;
  hipStream_t s = nullptr;      // the engine: joins and slices
  hipStream_t s_adm = nullptr;  // admissions: copies, resets, packs
  int tile = HDM_TILE_64x8x8;  // launch id
  int conn = 6;
  const TileLaunch *tl = nullptr;
  TileShape ts{};
  LaunchCfg cfg{};
  uint32_t qcap = 0;
  void *common = nullptr;
  GroupPtrs P{};
  void *arena_mem = nullptr;
  Arena arena;
  std::deque<Request> pending;
  std::vector<Request> slots;  // slot index -> request (id 0 = free)
  std::vector<uint32_t> slot_gen;
  // A slot freed after slice k is reused only after slice k+1 has run: that
  // slice drops the queue entries its last owner left (stale generation), so
  // one slot never holds a pair's entries and a predecessor's at once and
  // the queue stays within 2 * tiles(max dims) per slot.
  std::vector<uint64_t> freed_after;
  uint64_t slices = 0;
  std::vector<hipEvent_t> event_pool;
  void *host = nullptr;  // pinned: outs + engine head + ready marker (zeroed at creation)
  uint32_t seq = 0;      // ready marker value of the last slice
  // lone requests run as one-pair calls (off with HDM_SERVER_LONE=0, or for
  // good once that call found no device memory for its workspace)
  bool lone = true;
  uint64_t next_id = 1;
  uint32_t t_engine = 0;
  // completions of a failed slice beyond the caller's capacity, reported
  // first by the next steps
  std::deque<hdm_completion> backlog;
  std::map<uint64_t, MaskReg> masks;
  uint64_t next_mask = 1;
  hdm_server_stats st{};
  // admission groups (near pass on): device descriptors, pack index,
  // counters, entry count and entry list of up to max_slots pairs; pinned
  // staging of the descriptors and index, the outcomes and their marker
  bool near = false;
  int near_grid = 1;
  void *gmem = nullptr;
  PairDesc *gdescs = nullptr;
  PackIndex *gidx = nullptr;
  PairCtr *gctrs = nullptr;
  uint32_t *gcount = nullptr;
  uint64_t *glist = nullptr;
  uint32_t glist_cap = 0;
  void *ghost = nullptr;
  PairDesc *hdescs = nullptr;
  PackIndex *hidx = nullptr;
  PairOut *gouts = nullptr;
  volatile uint32_t *gready = nullptr;
  uint32_t gseq = 0;
  // registered pairs of a group: their masks' tile flags (device, and the
  // pinned staging)
  const uint32_t **gflags = nullptr, **hflags = nullptr;
};

static hipEvent_t take_event(hdm_server *s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

extern "C" void hdm_server_opts_init(hdm_server_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->hbm_budget = 64ull << 30;
  o->max_slots = 64;
  o->slice_iters = 8;
  o->tile = HDM_TILE_64x8x8;
  o->robustness = 1.0;
  o->max_nx = o->max_ny = o->max_nz = 512;
}

extern "C" double hdm_server_now(void) { return now_s(); }

extern "C" void hdm_server_destroy(hdm_server *s) {
  if (!s) return;
  if (s->ctx && s->s) {
    // a lone request ran as a one-pair call on this stream and may have left
    // it as the context's pending stream: settle it before the stream goes,
    // so no later call synchronizes a destroyed (or reused) handle
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    if (s->ctx->has_pending && s->ctx->pending == s->s) (void)s->ctx->settle();
  }
  if (s->s) (void)hipStreamSynchronize(s->s);
  if (s->s_adm) (void)hipStreamSynchronize(s->s_adm);
  for (auto &r : s->slots)
    if (r.staged) (void)hipEventDestroy(r.staged);
  for (auto e : s->event_pool) (void)hipEventDestroy(e);
  if (s->common) (void)hipFree(s->common);
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  if (s->host) (void)hipHostFree(s->host);
  if (s->gmem) (void)hipFree(s->gmem);
  if (s->ghost) (void)hipHostFree(s->ghost);
  if (s->s) (void)hipStreamDestroy(s->s);
  if (s->s_adm) (void)hipStreamDestroy(s->s_adm);
  (void)hipGetLastError();
  delete s;
}

extern "C" int hdm_server_create(const hdm_server_opts *oin, hdm_server **out) {
  if (!out) return HDM_E_INVALID;
  *out = nullptr;
  hdm_server_opts o;
  hdm_server_opts_init(&o);
  if (oin) o = *oin;
  int conn = 6;
  const int lid = select_launch(o.tile, o.connectivity, &conn, ds_enabled());
  if (o.max_slots < 2 || o.max_slots > 4096 || o.slice_iters < 1 || lid < 0 ||
      !(o.robustness >= 0.0 && o.robustness <= 1.0) ||
      o.max_nx < 1 || o.max_ny < 1 || o.max_nz < 1 || o.hbm_budget < (1u << 20))
    return HDM_E_INVALID;
  Ctx *c = nullptr;
  int rc = get_ctx(&c);
  if (rc != HDM_OK) return rc;
  hdm_server *s = new hdm_server();
  s->o = o;
  s->ctx = c;
  {
    const char *e = std::getenv("HDM_SERVER_LONE");
    s->lone = !(e && std::atoi(e) == 0);
  }
  s->tile = lid;
  s->conn = conn;
  s->tl = hdm_tile_launch(s->tile);
  s->ts = tile_shape(s->tile);
  const PairGeom gmax = pair_geom(o.max_nx, o.max_ny, o.max_nz, s->ts);
  const uint64_t qcap = 2ull * gmax.ntiles * (uint64_t)o.max_slots;
  if (qcap > 0xFFFFFFFFull) {
    delete s;
    return HDM_E_INVALID;
  }
  s->qcap = (uint32_t)qcap;
  const CommonLayout L = common_layout(o.max_slots, qcap, 1);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  uint64_t arena = o.hbm_budget;
  if (free_b && arena + L.total > (uint64_t)(free_b * 0.9)) arena = (uint64_t)(free_b * 0.9) - L.total;
  if (hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->s_adm, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->common, L.total) != hipSuccess || hipMalloc(&s->arena_mem, arena) != hipSuccess ||
      hipHostMalloc(&s->host, align256(sizeof(PairOut) * o.max_slots) + 256,
                    hipHostMallocCoherent) !=
          hipSuccess) {
    (void)hipGetLastError();
    hdm_server_destroy(s);
    return HDM_E_NOMEM;
  }
  // pinned memory comes back from earlier allocations with their contents: a
  // marker left by an earlier server must not match this one's first slice
  std::memset(s->host, 0, align256(sizeof(PairOut) * o.max_slots) + 256);
  s->arena.init(arena);
  s->P = carve_common((char *)s->common, L);
  s->slots.resize(o.max_slots);
  s->slot_gen.assign(o.max_slots, 0);
  s->freed_after.assign(o.max_slots, 0);
  s->slices = 1;
  s->cfg = launch_cfg(c, s->tile, 1, 1, o.grid_ctas, 1, o.max_slots);  // several pairs
  s->near = s->tl->fused_near && conn == 6 && near_enabled();
  if (s->near) {
    const uint64_t G = (uint64_t)o.max_slots;
    const uint64_t lcap = G * 2 * (uint64_t)gmax.nregions;
    s->glist_cap = (uint32_t)std::min<uint64_t>(lcap, 0xFFFFFFFFull);
    const uint64_t bd = align256(G * sizeof(PairDesc)), bi = align256(G * sizeof(PackIndex)),
                   bc = align256(G * sizeof(PairCtr));
    const uint64_t bf = align256(2 * G * sizeof(void *));
    const uint64_t hb = bd + bi + align256(G * sizeof(PairOut)) + 256 + bf;
    if (hipMalloc(&s->gmem, bd + bi + bc + 256 + bf + 8 * (uint64_t)s->glist_cap) != hipSuccess ||
        hipHostMalloc(&s->ghost, hb, hipHostMallocCoherent) != hipSuccess) {
      (void)hipGetLastError();
      hdm_server_destroy(s);
      return HDM_E_NOMEM;
    }
    char *g = (char *)s->gmem;
    s->gdescs = (PairDesc *)g;
    s->gidx = (PackIndex *)(g + bd);
    s->gctrs = (PairCtr *)(g + bd + bi);
    s->gcount = (uint32_t *)(g + bd + bi + bc);
    s->gflags = (const uint32_t **)(g + bd + bi + bc + 256);
    s->glist = (uint64_t *)(g + bd + bi + bc + 256 + bf);
    std::memset(s->ghost, 0, hb);  // no marker left by an earlier allocation matches
    char *h = (char *)s->ghost;
    s->hdescs = (PairDesc *)h;
    s->hidx = (PackIndex *)(h + bd);
    s->gouts = (PairOut *)(h + bd + bi);
    s->gready = (volatile uint32_t *)(h + bd + bi + align256(G * sizeof(PairOut)));
    s->hflags = (const uint32_t **)(h + bd + bi + align256(G * sizeof(PairOut)) + 256);
    s->near_grid = std::max(1, std::min(kMaxGrid, c->occ_fused_near[s->tile] * c->cus));
  }
  // every slot starts unused: a start beyond any engine iteration keeps it
  // out of the per-iteration bookkeeping
  std::vector<PairDesc> d(o.max_slots);
  for (auto &x : d) {
    std::memset(&x, 0, sizeof(x));
    x.t_start = kUnused;
  }
  std::vector<PairCtr> cz(o.max_slots);
  std::memset(cz.data(), 0, sizeof(PairCtr) * o.max_slots);
  for (auto &x : cz) x.done_from[0] = x.done_from[1] = 0;
  if (hipMemcpy(s->P.descs, d.data(), sizeof(PairDesc) * o.max_slots, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(s->P.ctrs, cz.data(), sizeof(PairCtr) * o.max_slots, hipMemcpyHostToDevice) !=
          hipSuccess) {
    hdm_server_destroy(s);
    return HDM_E_HIP;
  }
  // engine state: t = 0, empty queue
  hipError_t e = hdm_launch_reset(s->s, s->P.descs, nullptr, nullptr, s->P.idx, s->P.ctrs, 0, 0,
                                  s->P.engine, 1, s->qcap);
  if (e != hipSuccess || hipStreamSynchronize(s->s) != hipSuccess) {
    hdm_server_destroy(s);
    return HDM_E_HIP;
  }
  *out = s;
  return HDM_OK;
}

static uint64_t request_bytes(const hdm_server *s, const Request &r, uint64_t *in_bytes) {
  const uint64_t nvox = (uint64_t)r.nx * r.ny * r.nz;
  const uint64_t esz = r.dtype == HDM_DTYPE_U8 ? 1 : 4;
  *in_bytes = r.ptr_kind == HDM_PTR_HOST && !r.mg ? 2 * align256(nvox * esz) : 0;
  // with admission groups, the pair's near-pass bitmap after its workspace
  const uint64_t marks = s->near ? align256(4ull * near_mark_words(r.geom.nregions)) : 0;
  return *in_bytes + r.geom.total + marks;
}

extern "C" int hdm_server_submit(hdm_server *s, const void *gold, const void *segm, int nx,
                                 int ny, int nz, int dtype, int label, int ptr_kind,
                                 uint64_t *id_out) {
  if (!s || !gold || !segm || nx < 1 || ny < 1 || nz < 1 ||
      (dtype != HDM_DTYPE_U8 && dtype != HDM_DTYPE_I32) ||
      (ptr_kind != HDM_PTR_DEVICE && ptr_kind != HDM_PTR_HOST))
    return HDM_E_INVALID;
  if (nx > s->o.max_nx || ny > s->o.max_ny || nz > s->o.max_nz) return HDM_E_CAPACITY;
  Request r;
  r.id = s->next_id++;
  r.gold = gold, r.segm = segm;
  r.nx = nx, r.ny = ny, r.nz = nz, r.dtype = dtype, r.label = label, r.ptr_kind = ptr_kind;
  r.geom = pair_geom(nx, ny, nz, s->ts);
  uint64_t inb = 0;
  r.bytes = request_bytes(s, r, &inb);
  if (align256(r.bytes) > s->arena.size) return HDM_E_BUDGET;
  r.t_submit = now_s();
  s->pending.push_back(r);
  if (id_out) *id_out = r.id;
  return HDM_OK;
}

// Registration: the volume packed as both masks of a one-pair pack into the
// mask's own block (both masks' stores go to the same words with the same
// values), on the admission stream, waited for before returning.
extern "C" int hdm_server_register(hdm_server *s, const void *vol, int nx, int ny, int nz,
                                   int dtype, int label, int ptr_kind, uint64_t *handle_out) {
  if (!s || !vol || !handle_out || nx < 1 || ny < 1 || nz < 1 ||
      (dtype != HDM_DTYPE_U8 && dtype != HDM_DTYPE_I32) ||
      (ptr_kind != HDM_PTR_DEVICE && ptr_kind != HDM_PTR_HOST))
    return HDM_E_INVALID;
  if (dtype == HDM_DTYPE_U8 && (label < 0 || label > 255)) return HDM_E_INVALID;
  if (nx > s->o.max_nx || ny > s->o.max_ny || nz > s->o.max_nz) return HDM_E_CAPACITY;
  const PairGeom g = pair_geom(nx, ny, nz, s->ts);
  const uint64_t nvox = (uint64_t)nx * ny * nz;
  const uint64_t esz = dtype == HDM_DTYPE_U8 ? 1 : 4;
  const uint64_t b_desc = align256(sizeof(PairDesc)), b_idx = align256(sizeof(PackIndex)),
                 b_ctr = align256(sizeof(PairCtr));
  const uint64_t keep = g.buf_bytes + g.xf_bytes + g.tf_bytes;
  const uint64_t bytes = keep + b_desc + b_idx + b_ctr;
  const uint64_t in_bytes = ptr_kind == HDM_PTR_HOST ? align256(nvox * esz) : 0;
  MaskReg m;
  m.nx = nx, m.ny = ny, m.nz = nz;
  if (!s->arena.alloc(bytes, &m.off)) return HDM_E_BUDGET;
  m.bytes = bytes;
  uint64_t in_off = 0;
  if (in_bytes && !s->arena.alloc(in_bytes, &in_off)) {
    s->arena.release(m.off, m.bytes);
    return HDM_E_BUDGET;
  }
  char *base = (char *)s->arena_mem + m.off;
  m.orig = base;
  m.xf = (uint64_t *)(base + g.buf_bytes);
  m.tflags = (uint32_t *)(base + g.buf_bytes + g.xf_bytes);
  PairDesc *ddev = (PairDesc *)(base + keep);
  PackIndex *idev = (PackIndex *)(base + keep + b_desc);
  PairCtr *cdev = (PairCtr *)(base + keep + b_desc + b_idx);
  const void *src = vol;
  hipError_t e = hipSuccess;
  if (in_bytes) {
    src = (char *)s->arena_mem + in_off;
    e = hipMemcpyAsync((void *)src, vol, nvox * esz, hipMemcpyHostToDevice, s->s_adm);
  }
  PairDesc d{};
  fill_desc(d, g, src, src, dtype, label, s->o.robustness, 1, s->tile, kUnused, s->conn);
  d.orig[0] = d.orig[1] = m.orig;
  d.xf_orig[0] = d.xf_orig[1] = m.xf;
  d.tflags = m.tflags;
  d.hist_len = 0;  // the reset below clears no histogram
  const PackIndex pi{0, 0};
  if (e == hipSuccess)
    e = hdm_launch_reset(s->s_adm, ddev, &d, &pi, idev, cdev, 0, 1, nullptr, 0, 0);
  if (e == hipSuccess) {
    const LaunchCfg lc = launch_cfg(s->ctx, s->tile, g.row_words, g.ntiles, 0, 1, 1);
    e = s->tl->pack(dim3(lc.pack_blocks), s->s_adm, ddev, idev, cdev, 1, g.ntiles, nullptr);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->s_adm);
  if (in_bytes) s->arena.release(in_off, in_bytes);
  if (e != hipSuccess) {
    s->arena.release(m.off, m.bytes);
    return map_hip(e);
  }
  const uint64_t h = s->next_mask++;
  s->masks[h] = m;
  s->st.registered_bytes += align256(m.bytes);
  *handle_out = h;
  return HDM_OK;
}

static void mask_unref(hdm_server *s, uint64_t h) {
  auto it = s->masks.find(h);
  if (it == s->masks.end()) return;
  MaskReg &m = it->second;
  if (--m.refs <= 0 && m.released) {
    s->arena.release(m.off, m.bytes);
    s->st.registered_bytes -= align256(m.bytes);
    s->masks.erase(it);
  }
}

extern "C" int hdm_server_unregister(hdm_server *s, uint64_t handle) {
  if (!s) return HDM_E_INVALID;
  auto it = s->masks.find(handle);
  if (it == s->masks.end() || it->second.released) return HDM_E_INVALID;
  it->second.released = true;
  ++it->second.refs;  // mask_unref below drops it now when nothing names it
  mask_unref(s, handle);
  return HDM_OK;
}

extern "C" int hdm_server_submit_masks(hdm_server *s, uint64_t gold, uint64_t segm,
                                       uint64_t *id_out) {
  if (!s) return HDM_E_INVALID;
  auto ig = s->masks.find(gold), is = s->masks.find(segm);
  if (ig == s->masks.end() || is == s->masks.end() || ig->second.released ||
      is->second.released)
    return HDM_E_INVALID;
  const MaskReg &g = ig->second, &m = is->second;
  if (g.nx != m.nx || g.ny != m.ny || g.nz != m.nz) return HDM_E_INVALID;
  Request r;
  r.id = s->next_id++;
  r.mg = gold, r.ms = segm;
  r.nx = g.nx, r.ny = g.ny, r.nz = g.nz, r.dtype = HDM_DTYPE_U8;
  r.geom = pair_geom(r.nx, r.ny, r.nz, s->ts);
  uint64_t inb = 0;
  r.bytes = request_bytes(s, r, &inb);
  if (align256(r.bytes) > s->arena.size) return HDM_E_BUDGET;
  ++ig->second.refs, ++is->second.refs;
  r.t_submit = now_s();
  s->pending.push_back(r);
  if (id_out) *id_out = r.id;
  return HDM_OK;
}

// Staging of one request into slot p, on the admission stream: inputs
// (copied, resident, or registered masks), descriptor with an unused start,
// counters, pack or combine.  The running slice ignores the slot (its start
// is unused), and nothing here touches the engine.
static int stage(hdm_server *s, Request &r, int p) {
  uint64_t inb = 0;
  request_bytes(s, r, &inb);
  char *base = (char *)s->arena_mem + r.off;
  const void *gd = r.gold, *sd = r.segm;
  hipError_t e = hipSuccess;
  if (r.ptr_kind == HDM_PTR_HOST && !r.mg) {
    const uint64_t nb = (uint64_t)r.nx * r.ny * r.nz * (r.dtype == HDM_DTYPE_U8 ? 1 : 4);
    if (hipMemcpyAsync(base, r.gold, nb, hipMemcpyHostToDevice, s->s_adm) != hipSuccess ||
        hipMemcpyAsync(base + inb / 2, r.segm, nb, hipMemcpyHostToDevice, s->s_adm) !=
            hipSuccess)
      return HDM_E_HIP;
    gd = base, sd = base + inb / 2;
  }
  PairDesc d{};
  fill_desc(d, r.geom, gd, sd, r.dtype, r.label, s->o.robustness, 1, s->tile, kUnused, s->conn);
  carve_pair(d, r.geom, base + inb);
  const MaskReg *mg = nullptr, *ms = nullptr;
  if (r.mg) {
    mg = &s->masks.at(r.mg), ms = &s->masks.at(r.ms);
    d.gold = d.segm = nullptr;
    d.orig[0] = mg->orig, d.orig[1] = ms->orig;
    d.xf_orig[0] = mg->xf, d.xf_orig[1] = ms->xf;
  }
  d.gen = (++s->slot_gen[p]) & kGenMask;  // drops queue entries of the slot's last owner
  const PackIndex pi{0, 0};
  e = hdm_launch_reset(s->s_adm, s->P.descs, &d, &pi, s->P.idx, s->P.ctrs, p, 1, nullptr, 0,
                       s->qcap);
  if (e != hipSuccess) return map_hip(e);
  const LaunchCfg lc = launch_cfg(s->ctx, s->tile, r.geom.row_words, r.geom.ntiles, 0, 1, 1);
  if (r.mg)
    e = s->tl->combine(dim3(lc.pack_blocks), s->s_adm, s->P.descs + p, s->P.ctrs + p,
                       mg->tflags, ms->tflags, r.geom.ntiles);
  else
    e = s->tl->pack(dim3(lc.pack_blocks), s->s_adm, s->P.descs + p, s->P.idx + p,
                    s->P.ctrs + p, 1, r.geom.ntiles, nullptr);
  if (e != hipSuccess) return map_hip(e);
  r.staged = take_event(s);
  if (!r.staged || hipEventRecord(r.staged, s->s_adm) != hipSuccess) return HDM_E_HIP;
  r.slot = p;
  r.active = false;
  r.t_admit = now_s();
  return HDM_OK;
}

// Joining the staged requests, on the engine stream between slices: after
// their staging events, their start is the engine's current iteration (one
// launch for all), and each one's seed (which folds its totals) puts its
// first tiles in the current iteration's queue.  Returns the number joined.
static int join_staged(hdm_server *s, int *joined) {
  std::vector<int> list;
  for (auto &r : s->slots) {
    if (r.id == 0 || r.active) continue;
    if (hipStreamWaitEvent(s->s, r.staged, 0) != hipSuccess) return HDM_E_HIP;
    s->event_pool.push_back(r.staged);  // the wait holds the record it saw
    r.staged = nullptr;
    list.push_back(r.slot);
  }
  *joined = (int)list.size();
  if (list.empty()) return HDM_OK;
  hipError_t e = hdm_launch_set_start(s->s, s->P.descs, list.data(), (int)list.size(), s->t_engine);
  for (int p : list) {
    if (e != hipSuccess) break;
    Request &r = s->slots[p];
    const LaunchCfg lc = launch_cfg(s->ctx, s->tile, r.geom.row_words, r.geom.ntiles, 0, 1, 1);
    e = s->tl->seed(dim3(lc.seed_blocks), s->s, s->P.descs + p, s->P.idx + p, s->P.ctrs + p, 1,
                    s->P.engine, s->P.q, s->qcap, r.geom.ntiles, p);
    r.active = true;
  }
  return e == hipSuccess ? HDM_OK : map_hip(e);
}

// Stages pending requests in arrival order while a slot and arena room exist.
static int stage_pending(hdm_server *s, bool *quarantined) {
  while (!s->pending.empty()) {
    int p = -1;
    for (int i = 0; i < (int)s->slots.size(); ++i)
      if (s->slots[i].id == 0) {
        if (s->freed_after[i] < s->slices) {
          p = i;
          break;
        }
        *quarantined = true;
      }
    if (p < 0) break;
    Request &r = s->pending.front();
    if (!s->arena.alloc(r.bytes, &r.off)) break;
    const int rc = stage(s, r, p);
    if (rc != HDM_OK) {
      s->arena.release(r.off, r.bytes);
      return rc;
    }
    ++s->st.admitted;
    s->slots[p] = r;
    s->pending.pop_front();
  }
  return HDM_OK;
}

// One admission group (servers with the near pass): the waiting requests, in
// arrival order, while a free slot for each (any of them may need one), arena
// room and room among the caller's completions remain.  Volume pairs come
// first (one pack over them), registered pairs after (a combine each).  The
// pairs the near pass finishes complete here; the others move into slots as
// staged pairs, which the join then starts.  Returns an error only for a
// failure of the step itself (the group's requests then complete with it).
static int run_group(hdm_server *s, hdm_completion *done, int cap, int *n_done,
                     bool *quarantined) {
  std::vector<int> free_slots;
  for (int i = 0; i < (int)s->slots.size(); ++i)
    if (s->slots[i].id == 0) {
      if (s->freed_after[i] < s->slices)
        free_slots.push_back(i);
      else
        *quarantined = true;
    }
  const int room = std::min<int>((int)free_slots.size(), cap - *n_done);
  std::vector<Request> grp;
  while (!s->pending.empty() && (int)grp.size() < room) {
    Request &r = s->pending.front();
    if (!s->arena.alloc(r.bytes, &r.off)) break;
    grp.push_back(r);
    s->pending.pop_front();
  }
  if (grp.empty()) return HDM_OK;
  ++s->st.groups;
  s->st.group_pairs += grp.size();
  std::stable_partition(grp.begin(), grp.end(), [](const Request &r) { return r.mg == 0; });
  const int n = (int)grp.size();
  int nv = 0;
  uint64_t tiles = 0, rw = 0, words = 0, rtiles = 0;
  hipError_t e = hipSuccess;
  const double t_admit = now_s();
  for (int k = 0; k < n && e == hipSuccess; ++k) {
    Request &r = grp[k];
    uint64_t inb = 0;
    request_bytes(s, r, &inb);
    char *base = (char *)s->arena_mem + r.off;
    const void *gd = r.gold, *sd = r.segm;
    if (r.ptr_kind == HDM_PTR_HOST && !r.mg) {
      const uint64_t nb = (uint64_t)r.nx * r.ny * r.nz * (r.dtype == HDM_DTYPE_U8 ? 1 : 4);
      e = hipMemcpyAsync(base, r.gold, nb, hipMemcpyHostToDevice, s->s_adm);
      if (e == hipSuccess)
        e = hipMemcpyAsync(base + inb / 2, r.segm, nb, hipMemcpyHostToDevice, s->s_adm);
      gd = base, sd = base + inb / 2;
    }
    PairDesc d{};
    fill_desc(d, r.geom, gd, sd, r.dtype, r.label, s->o.robustness, 1, s->tile, kUnused, s->conn);
    carve_pair(d, r.geom, base + inb);
    d.marks = (uint32_t *)(base + inb + r.geom.total);
    const uint64_t G = (uint64_t)s->o.max_slots;
    if (r.mg) {
      const MaskReg &mg = s->masks.at(r.mg), &ms = s->masks.at(r.ms);
      d.gold = d.segm = nullptr;
      d.orig[0] = mg.orig, d.orig[1] = ms.orig;
      d.xf_orig[0] = mg.xf, d.xf_orig[1] = ms.xf;
      s->hflags[k - nv] = mg.tflags, s->hflags[G + k - nv] = ms.tflags;
    } else {
      ++nv;
    }
    s->hdescs[k] = d;
    // tile bases: volume pairs among themselves (the pack), registered pairs
    // among themselves (the combine)
    s->hidx[k] = PackIndex{rw, r.mg ? rtiles : tiles, words};
    if (!r.mg) tiles += r.geom.ntiles, rw += r.geom.row_words;
    else rtiles += r.geom.ntiles;
    words += near_mark_words(r.geom.nregions);
    r.t_admit = t_admit;
  }
  const uint32_t seq = ++s->gseq ? s->gseq : ++s->gseq;
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->gdescs, s->hdescs, sizeof(PairDesc) * n, hipMemcpyHostToDevice, s->s_adm);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->gidx, s->hidx, sizeof(PackIndex) * n, hipMemcpyHostToDevice, s->s_adm);
  if (e == hipSuccess && nv < n)
    e = hipMemcpyAsync(s->gflags, s->hflags, 2 * sizeof(void *) * s->o.max_slots,
                       hipMemcpyHostToDevice, s->s_adm);
  if (e == hipSuccess) e = hipMemsetAsync(s->gcount, 0, sizeof(uint32_t), s->s_adm);
  if (e == hipSuccess)
    e = hdm_launch_reset(s->s_adm, s->gdescs, nullptr, nullptr, s->gidx, s->gctrs, 0, n, nullptr,
                         0, 0);
  if (e == hipSuccess && nv > 0) {
    const LaunchCfg lc = launch_cfg(s->ctx, s->tile, rw, tiles, 0, 1, nv);
    e = s->tl->pack(dim3(lc.pack_blocks), s->s_adm, s->gdescs, s->gidx, s->gctrs, nv, tiles,
                    nullptr);
  }
  if (e == hipSuccess && nv < n) {
    const LaunchCfg lc = launch_cfg(s->ctx, s->tile, 0, rtiles, 0, 1, n - nv);
    e = hdm_launch_combine_group(s->s_adm, s->gdescs + nv, s->gidx + nv, s->gctrs + nv, n - nv,
                                 rtiles, s->gflags, s->gflags + s->o.max_slots,
                                 (int)lc.pack_blocks);
  }
  if (e == hipSuccess) e = hdm_launch_totals(s->s_adm, s->gctrs, 0, n);
  if (e == hipSuccess)
    e = hdm_launch_near_batch(s->s_adm, s->gdescs, s->gidx, s->gctrs, n, words, s->gcount,
                              s->glist, s->glist_cap, s->near_grid);
  if (e == hipSuccess)
    e = hdm_launch_group_out(s->s_adm, s->gdescs, s->gctrs, n, s->gouts, (uint32_t *)s->gready,
                             seq);
  int rc = e == hipSuccess ? wait_ready(s->s_adm, s->gready, seq) : map_hip(e);
  const double t_now = now_s();
  std::vector<int> src, dst;
  std::vector<uint32_t> gen;
  size_t next_slot = 0;
  for (int k = 0; k < n; ++k) {
    Request &r = grp[k];
    PairOut o;
    std::memcpy(&o, s->gouts + k, sizeof(o));
    if (rc == HDM_OK && o.status == 2) {  // to the engine
      const int p = free_slots[next_slot++];
      src.push_back(k), dst.push_back(p);
      gen.push_back((++s->slot_gen[p]) & kGenMask);
      r.slot = p;
      r.active = false;
      s->slots[p] = r;
      ++s->st.admitted;
      continue;
    }
    hdm_completion &c = done[(*n_done)++];
    std::memset(&c, 0, sizeof(c));
    c.id = r.id;
    c.status = rc == HDM_OK ? o.status : rc;
    if (rc == HDM_OK) fill_result(&c.result, o, nullptr, r.geom.ntiles, std::max(o.hd_fp, o.hd_fn));
    c.t_submit = r.t_submit;
    c.t_admit = r.t_admit;
    c.t_done = t_now;
    s->arena.release(r.off, r.bytes);
    if (r.mg) mask_unref(s, r.mg), mask_unref(s, r.ms);
    if (rc == HDM_OK) ++s->st.near_finished;
  }
  if (rc != HDM_OK) return HDM_OK;  // the group's requests completed with the error
  if (!src.empty()) {
    e = hdm_launch_move_pairs(s->s_adm, s->gdescs, s->gctrs, s->P.descs, s->P.ctrs, s->P.idx,
                              src.data(), dst.data(), gen.data(), (int)src.size());
    for (int p : dst) {
      Request &r = s->slots[p];
      r.staged = take_event(s);
      if (e == hipSuccess && (!r.staged || hipEventRecord(r.staged, s->s_adm) != hipSuccess))
        e = hipErrorUnknown;
    }
    if (e != hipSuccess) return map_hip(e);
  }
  return HDM_OK;
}

// The lone request at the head of the FIFO as a one-pair call (hdm_hausdorff_*)
// on the engine stream: the same results as the pair computed alone, which
// every request's result equals.  Its workspace is the library context's (the
// one the one-pair calls use, sized by the request), not the arena.  A
// per-request outcome (done, empty mask, timeout, stuck) completes the
// request; any other error is the step's and the request stays queued.
static int run_lone(hdm_server *s, hdm_completion *done, int *n_done) {
  Request r = s->pending.front();
  hdm_opts o;
  hdm_opts_init(&o);
  o.robustness = s->o.robustness;
  o.ptr_kind = r.ptr_kind;
  o.tile = s->o.tile;
  o.connectivity = s->o.connectivity;
  o.stream = s->s;
  hdm_result res;
  std::memset(&res, 0, sizeof(res));
  r.t_admit = now_s();
  const int rc =
      r.dtype == HDM_DTYPE_U8
          ? hdm_hausdorff_u8((const uint8_t *)r.gold, (const uint8_t *)r.segm, r.nx, r.ny, r.nz,
                             (uint8_t)r.label, &o, &res)
          : hdm_hausdorff_i32((const int32_t *)r.gold, (const int32_t *)r.segm, r.nx, r.ny,
                              r.nz, (int32_t)r.label, &o, &res);
  if (rc != HDM_OK && rc != HDM_E_EMPTY_MASK && rc != HDM_E_TIMEOUT && rc != HDM_E_STUCK)
    return rc;
  s->pending.pop_front();
  ++s->st.lone;
  hdm_completion &c = done[(*n_done)++];
  std::memset(&c, 0, sizeof(c));
  c.id = r.id;
  c.status = rc;
  c.result = res;
  c.t_submit = r.t_submit;
  c.t_admit = r.t_admit;
  c.t_done = now_s();
  return HDM_OK;
}

extern "C" int hdm_server_step(hdm_server *s, hdm_completion *done, int cap, int *n_done) {
  if (!s || (!done && cap > 0) || !n_done) return HDM_E_INVALID;
  *n_done = 0;
  while (!s->backlog.empty() && *n_done < cap) {  // a failed slice's leftovers
    done[(*n_done)++] = s->backlog.front();
    s->backlog.pop_front();
  }
  if (!s->backlog.empty()) return HDM_OK;
  bool idle = true;  // no request staged or in flight
  for (auto &r : s->slots) idle = idle && r.id == 0;
  if (idle && s->lone && s->pending.size() == 1 && cap > *n_done && !s->pending.front().mg &&
      (s->pending.front().dtype != HDM_DTYPE_U8 ||
       (s->pending.front().label >= 0 && s->pending.front().label <= 255))) {
    const int rc = run_lone(s, done, n_done);
    if (rc != HDM_E_NOMEM) return rc;
    // no device memory for the one-pair workspace (the arena may hold most
    // of it): this request and every later one go through the slots
    s->lone = false;
    (void)hipGetLastError();
  }
  // join what the last step staged; with nothing in flight, stage and join
  // now (there is no slice to overlap)
  // stage what waits, then join it: between slices, never beside one
  bool quarantined = false;
  int active = 0, joined = 0;
  int rc = s->near ? run_group(s, done, cap, n_done, &quarantined) : stage_pending(s, &quarantined);
  if (rc != HDM_OK) return rc;
  if ((rc = join_staged(s, &joined)) != HDM_OK) return rc;
  for (auto &r : s->slots) active += r.id != 0 && r.active;
  // with nothing in flight a slice still runs when requests wait on
  // quarantined slots: it drops the stale entries and frees them
  if (active == 0 && !(quarantined && !s->pending.empty())) return HDM_OK;
  // one slice of the engine over every slot
  const int nslots = (int)s->slots.size();
  ++s->slices;
  ++s->st.slices;
  s->st.pair_slices += (uint64_t)active;
  s->st.max_in_flight = std::max(s->st.max_in_flight, active);
  // the slice's workgroup 0 writes every slot's result and the engine head
  // into the pinned host block as the launch leaves
  char *h = (char *)s->host;
  EngineHead *hhead = (EngineHead *)(h + align256(sizeof(PairOut) * nslots));
  hhead->nwg = 0;
  DilateArgs a{};
  a.descs = s->P.descs;
  a.ctrs = s->P.ctrs;
  a.npairs = nslots;
  // engine iterations of slice_iters dilation steps
  a.max_iters = (s->o.slice_iters + s->tl->steps - 1) / s->tl->steps;
  a.es = s->P.engine;
  a.q = s->P.q;
  a.qcap = s->qcap;
  a.chunk = 1;
  a.spin = spin_ticks(s->ctx);  // read per slice (HDM_SPIN_TICKS)
  a.flags = barrier_mode();
  a.outs = (PairOut *)h;
  a.head = hhead;
  // the slice stores its sequence number after the results and the head;
  // the step takes them from there instead of waiting for the stream (the
  // launch's workgroups touch no pair data after its last barrier, so the
  // slots and arena blocks freed below are free for the next admissions,
  // whose work the streams order after the launch)
  volatile uint32_t *ready =
      (volatile uint32_t *)(h + align256(sizeof(PairOut) * nslots) + 192);
  if (++s->seq == 0) ++s->seq;
  a.ready = (uint32_t *)ready;
  a.ready_seq = s->seq;
  hipError_t e = s->tl->dilate(dim3(s->cfg.grid), s->s, a);
  if (e != hipSuccess) return map_hip(e);
  const int wr = wait_ready(s->s, ready, s->seq);
  if (wr != HDM_OK) return wr;
  EngineHead head{};
  std::memcpy(&head, hhead, sizeof(head));
  if (head.nwg == 0) return HDM_E_HIP;  // the launch did not complete
  s->t_engine = head.t;
  const double t_now = now_s();
  std::vector<int> retired;
  for (int p = 0; p < nslots; ++p) {
    Request &r = s->slots[p];
    if (r.id == 0 || !r.active) continue;  // free, or staged during this slice
    PairOut o;
    std::memcpy(&o, h + sizeof(PairOut) * p, sizeof(o));
    int st = o.status;
    bool finished = o.done || st != HDM_OK;
    if (!finished && head.qnext == 0) {  // the queue drained with this pair unfinished
      st = HDM_E_STUCK;
      finished = true;
    }
    if (head.error) {
      st = HDM_E_TIMEOUT;
      finished = true;
    }
    if (!finished) continue;
    hdm_completion c;
    std::memset(&c, 0, sizeof(c));
    c.id = r.id;
    c.status = st;
    fill_result(&c.result, o, nullptr, r.geom.ntiles, std::max(o.hd_fp, o.hd_fn));
    c.t_submit = r.t_submit;
    c.t_admit = r.t_admit;
    c.t_done = t_now;
    if (*n_done < cap)
      done[(*n_done)++] = c;
    else if (head.error)
      s->backlog.push_back(c);  // the engine restarts below: report it later
    else
      continue;  // report it at the next step
    retired.push_back(p);
    s->arena.release(r.off, r.bytes);
    if (r.mg) mask_unref(s, r.mg), mask_unref(s, r.ms);
    r = Request();
    s->freed_after[p] = s->slices;
  }
  // the retired pairs leave the engine's bookkeeping for good: with their
  // start beyond every iteration, no later fold reads their counters or
  // writes their (released) histograms, and a later staging of the slot,
  // which waits for the next slice, finds it unused
  if (!retired.empty()) {
    e = hdm_launch_set_start(s->s, s->P.descs, retired.data(), (int)retired.size(), kUnused);
    if (e != hipSuccess) return map_hip(e);
  }
  if (head.error) {
    // a barrier wait expired and every request in flight failed with it: the
    // engine restarts at iteration 0 with an empty queue and a clear error
    // flag (otherwise every later slice would fail at its first barrier)
    const hipError_t e2 = hdm_launch_reset(s->s, s->P.descs, nullptr, nullptr, s->P.idx,
                                           s->P.ctrs, 0, 0, s->P.engine, 1, s->qcap);
    if (e2 != hipSuccess) return map_hip(e2);
    s->t_engine = 0;
  }
  return HDM_OK;
}

extern "C" int hdm_server_load(hdm_server *s, int *queued, int *in_flight, uint64_t *bytes) {
  if (!s) return HDM_E_INVALID;
  int busy = 0;
  for (auto &r : s->slots) busy += r.id != 0;
  if (queued) *queued = (int)s->pending.size();
  if (in_flight) *in_flight = busy;
  if (bytes) *bytes = s->arena.used;
  return HDM_OK;
}

extern "C" int hdm_server_stats_get(hdm_server *s, hdm_server_stats *out) {
  if (!s || !out) return HDM_E_INVALID;
  *out = s->st;
  return HDM_OK;
}
