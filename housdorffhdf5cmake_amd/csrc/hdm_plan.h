// Host-side planning: tile geometry, workspace layout and launch shapes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/hdm.h"
#include "hdm_engine.h"

namespace hdm {

struct PackIndex {
  uint64_t rw_base;
  uint64_t tile_base;
  uint64_t mark_base;  // first word of the pair's near-pass bitmap in the batch's
};

// Arguments of one persistent dilation launch.  When the launch leaves, its
// workgroup 0 writes every pair's result to outs and the engine head to head
// (either may be host-visible pinned memory; head may be null).
struct DilateArgs {
  const PairDesc *descs;
  PairCtr *ctrs;
  int npairs;
  int max_iters;  // iterations this launch may run (a serving slice)
  Engine *es;
  uint64_t *q;
  uint32_t qcap;
  int chunk;       // queue entries per wave step, in tiles per wave
  uint64_t spin;   // bound of every barrier wait, in wall-clock ticks
  uint32_t *qstats;
  int qstats_len;
  int flags;       // 1: fence the barrier (release/acquire at agent scope)
  uint64_t *trace; // per-iteration timestamps (diagnostic) or null
  PairOut *outs;
  EngineHead *head;
  // the fused single-pair launch (reset, pack, seed and dilation in one
  // kernel): the pair's descriptor, its tile count, and the epoch base of its
  // barrier tags (every launch's epochs exceed the previous launch's, so
  // tagged barrier words never need clearing)
  PairDesc desc;
  uint64_t total_tiles;
  uint32_t ebase;
  // when the launch leaves, workgroup 0 stores ready_seq to *ready (host
  // memory) after the results and the head, behind a system-scope fence, so
  // the host can take the results without waiting for the stream (or null)
  uint32_t ready_seq;
  uint32_t *ready;
  int32_t acc_set;  // fused launch: the set of Engine::acc it packs into (0 or 1)
  // fused launch: launch it cooperatively (hipLaunchCooperativeKernel), so the
  // runtime refuses a grid that cannot be resident at once instead of letting
  // the grid barrier wait for workgroups that never start (host choice only)
  int32_t coop;
  // the region launch: workgroups that stay for the periods after the first
  // (the others leave once the seed period is over); 0 = all
  int32_t tb_active;
  // the two-step fused launch: 0 computes every pair itself; 1 hands a pair
  // whose HD lower bound reaches kTbMinSteps over after its pack (result
  // status 2; the host then queues the region launch); 2 hands every pair over
  int32_t tb_mode;
  // the near launch: a bitmap of 2 * regions bits (pass 0 then pass 1), bit
  // set for a region holding voxels to be covered, which the launch fills,
  // and the other set of the two used by turns, which it zeroes for the next
  // near launch (marks_zero_words words)
  uint32_t *marks;
  uint32_t *marks_zero;
  // fused launch: the pack's last quarter of steps taken one at a time from
  // counters (0: every wave packs an even static share; 1: one counter; 8:
  // one counter per XCD, each over its own eighth of those steps; one counter
  // for the whole chip serializes its atomics and runs the pack at half speed)
  int32_t pack_pools;
  uint32_t marks_zero_words;
};

struct TileLaunch {
  // the last argument: the pack's step counters (Engine::pack_grab, zeroed) or null
  hipError_t (*pack)(dim3, hipStream_t, const PairDesc *, const PackIndex *, PairCtr *, int,
                     uint64_t, uint32_t *);
  hipError_t (*seed)(dim3, hipStream_t, const PairDesc *, const PackIndex *, PairCtr *, int,
                     Engine *, uint64_t *, uint32_t, uint64_t, int);
  hipError_t (*dilate)(dim3, hipStream_t, const DilateArgs &);
  hipError_t (*fused)(dim3, hipStream_t, const DilateArgs &);  // one pair, one launch
  // the fused launch with two dilation steps per barrier (null: not built)
  hipError_t (*fused_ds)(dim3, hipStream_t, const DilateArgs &);
  // the fused launch over regions of 1 x 2 x 2 tiles staged in LDS, eight
  // dilation steps per barrier (null: not built)
  hipError_t (*fused_tb)(dim3, hipStream_t, const DilateArgs &);
  // the fused launch that packs and runs the near pass (eight steps around the
  // regions holding voxels to be covered), handing longer pairs to fused_tb
  hipError_t (*fused_near)(dim3, hipStream_t, const DilateArgs &);
  const void *dilate_fn;         // several pairs
  const void *dilate_single_fn;  // one pair
  const void *fused_fn;          // one pair, reset + pack + seed + dilation
  const void *fused_ds_fn;
  const void *fused_tb_fn;
  const void *fused_near_fn;
  const void *pack_fn;
  // the pack of a pair of masks packed before (combine_kernel): counts, box
  // and the pair's tile flags from the masks' own flags
  hipError_t (*combine)(dim3, hipStream_t, const PairDesc *, PairCtr *, const uint32_t *,
                        const uint32_t *, uint64_t);
  // a batch whose near pass runs (64x8x8, 6 element; null otherwise): the
  // pack that writes no rows of a mask empty in a tile, and after the near
  // pass the zero rows of the pairs it did not finish
  hipError_t (*pack_skip)(dim3, hipStream_t, const PairDesc *, const PackIndex *, PairCtr *, int,
                          uint64_t, uint32_t *);
  hipError_t (*fill)(dim3, hipStream_t, const PairDesc *, const PairCtr *, int);
  int txb, ty, tz;
  int conn;   // structuring element of the seed and dilation kernels
  int steps;  // dilation steps per engine iteration of the seed and dilation kernels
};

// launch ids (hdm_tile_launch): the HDM_TILE_* kinds with the 6 element, then
// the 64x8x8 tile with the 18 and the 26 element, then the 64x8x8 tile with the
// 6 element at two dilation steps per engine iteration (ds)
constexpr int kLaunchIds = 8;
constexpr int kLaunchDs = 7;
inline int launch_id(int tile_kind, int conn, bool ds = false) {
  if (ds && conn == 6 && tile_kind == HDM_TILE_64x8x8) return kLaunchDs;
  return conn == 18 ? 5 : conn == 26 ? 6 : tile_kind;
}

struct TileShape {
  int txb, ty, tz;
  int rows() const { return ty * tz; }
};

inline TileShape tile_shape(int kind) {
  switch (kind) {
    case HDM_TILE_64x4x4: return {64, 4, 4};
    case HDM_TILE_64x8x4: return {64, 8, 4};
    case HDM_TILE_32x8x8: return {32, 8, 8};
    default: return {64, 8, 8};
  }
}

inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// Geometry and byte offsets (relative to the pair's block) of one pair.
struct PairGeom {
  int nx, ny, nz, wxn, gx, gy, gz;
  uint32_t ntiles;
  uint64_t words;      // ntiles * rows
  uint64_t row_words;  // wxn * ny * nz, the pack grid
  int hist_len;
  // regions of 1 x 2 x 2 tiles (the eight-step single-pair launch)
  uint32_t nregions;
  uint64_t buf_bytes, xf_bytes, st_bytes, rs_bytes, tf_bytes, hist_bytes, total;
};

inline PairGeom pair_geom(int nx, int ny, int nz, const TileShape &ts) {
  PairGeom g{};
  g.nx = nx, g.ny = ny, g.nz = nz;
  g.wxn = (nx + ts.txb - 1) / ts.txb;
  g.gx = g.wxn;
  g.gy = (ny + ts.ty - 1) / ts.ty;
  g.gz = (nz + ts.tz - 1) / ts.tz;
  g.ntiles = (uint32_t)((uint64_t)g.gx * g.gy * g.gz);
  g.words = (uint64_t)g.ntiles * ts.rows();
  g.row_words = (uint64_t)g.wxn * ny * nz;
  g.hist_len = nx + ny + nz + 2;
  g.buf_bytes = align256(g.words * (ts.txb / 8));
  g.xf_bytes = align256((uint64_t)g.ntiles * 8 * xf_words(ts.txb, ts.ty, ts.tz));
  g.st_bytes = align256((uint64_t)g.ntiles * 4);
  g.tf_bytes = align256((uint64_t)g.ntiles * 4);
  g.nregions = (uint32_t)((uint64_t)g.gx * ((g.gy + 1) / 2) * ((g.gz + 1) / 2));
  g.rs_bytes = align256((uint64_t)g.nregions * 4);
  g.hist_bytes = align256((uint64_t)g.hist_len * 4);
  g.total = 6 * g.buf_bytes + 6 * g.xf_bytes + 2 * g.st_bytes + 2 * g.rs_bytes + g.tf_bytes +
            2 * g.hist_bytes;
  return g;
}

// Fill the device pointers of a descriptor from the pair's block base.
inline void carve_pair(PairDesc &d, const PairGeom &g, char *base) {
  char *p = base;
  auto take = [&](uint64_t n) {
    char *q = p;
    p += n;
    return q;
  };
  d.orig[0] = take(g.buf_bytes);
  d.orig[1] = take(g.buf_bytes);
  for (int m = 0; m < 2; ++m)
    for (int s = 0; s < 2; ++s) d.buf[m][s] = take(g.buf_bytes);
  d.xf_orig[0] = (uint64_t *)take(g.xf_bytes);
  d.xf_orig[1] = (uint64_t *)take(g.xf_bytes);
  for (int m = 0; m < 2; ++m)
    for (int s = 0; s < 2; ++s) d.xf_buf[m][s] = (uint64_t *)take(g.xf_bytes);
  d.state[0] = (uint32_t *)take(g.st_bytes);
  d.state[1] = (uint32_t *)take(g.st_bytes);
  d.rstate[0] = (uint32_t *)take(g.rs_bytes);
  d.rstate[1] = (uint32_t *)take(g.rs_bytes);
  d.tflags = (uint32_t *)take(g.tf_bytes);
  d.hist[0] = (uint32_t *)take(g.hist_bytes);
  d.hist[1] = (uint32_t *)take(g.hist_bytes);
}

inline void fill_desc(PairDesc &d, const PairGeom &g, const void *gold, const void *segm,
                      int dtype, int label, double robust, int crop, int tile_kind,
                      uint32_t t_start, int conn) {
  d.gold = gold, d.segm = segm;
  d.nx = g.nx, d.ny = g.ny, d.nz = g.nz, d.wxn = g.wxn;
  d.gx = g.gx, d.gy = g.gy, d.gz = g.gz;
  d.ntiles = g.ntiles;
  d.dtype = dtype, d.label = label;
  d.hist_len = g.hist_len;
  d.t_start = t_start;
  d.crop = crop;
  d.tile_kind = tile_kind;
  d.robust = robust;
  d.words = g.words;
  d.map = nullptr;
  d.conn = conn;
}

// Shared (per launch group) region: descriptors, indices, counters, outputs,
// engine state and the two queue arrays.
// The queue is 2 (parity) x kSubQ sub-queues of qcap entries each: a (tile,
// mask) is queued at most once per iteration, so any one sub-queue can take the
// whole iteration.
struct CommonLayout {
  uint64_t descs, idx, ctrs, outs, engine, q, qstats, total;
};

inline CommonLayout common_layout(int npairs, uint64_t qcap, int qstats_len) {
  CommonLayout L{};
  uint64_t o = 0;
  auto take = [&](uint64_t n) {
    const uint64_t r = o;
    o += align256(n);
    return r;
  };
  L.descs = take(sizeof(PairDesc) * npairs);
  L.idx = take(sizeof(PackIndex) * npairs);
  L.ctrs = take(sizeof(PairCtr) * npairs);
  L.outs = take(sizeof(PairOut) * npairs);
  L.engine = take(sizeof(Engine));
  L.q = take(8 * qcap * 2 * kSubQ);
  L.qstats = take(4 * (uint64_t)qstats_len);
  L.total = o;
  return L;
}

}  // namespace hdm

extern "C" const hdm::TileLaunch *hdm_tile_launch(int kind);
extern "C" hipError_t hdm_launch_reset(hipStream_t s, hdm::PairDesc *descs,
                                       const hdm::PairDesc *single,
                                       const hdm::PackIndex *single_idx, hdm::PackIndex *idx,
                                       hdm::PairCtr *ctrs, int p0, int npairs, hdm::Engine *es,
                                       int reset_engine, uint32_t qcap);
extern "C" hipError_t hdm_launch_totals(hipStream_t s, hdm::PairCtr *ctrs, int p0, int npairs);
// the near pass over npairs pairs (bitmaps of total_words words in all): list
// (count, zero on entry, receives the entries; at most cap), steps, fold
extern "C" hipError_t hdm_launch_near_batch(hipStream_t s, const hdm::PairDesc *descs,
                                            const hdm::PackIndex *idx, hdm::PairCtr *ctrs,
                                            int npairs, uint64_t total_words, uint32_t *count,
                                            uint64_t *list, uint32_t cap, int grid);
// server admission group: outs[k] = the results of pair k if its near pass
// finished it (or a mask is empty), else status 2; then seq into *ready
// (pinned host memory) behind a system-scope fence
extern "C" hipError_t hdm_launch_group_out(hipStream_t s, const hdm::PairDesc *descs,
                                           const hdm::PairCtr *ctrs, int n, hdm::PairOut *outs,
                                           uint32_t *ready, uint32_t seq);
// server admission group, registered masks (64x8x8 tiles): the combine pass
// of npairs pairs at once (pair p: tiles from idx[p].tile_base, its masks'
// tile flags gfs[p] and sfs[p])
extern "C" hipError_t hdm_launch_combine_group(hipStream_t s, const hdm::PairDesc *descs,
                                               const hdm::PackIndex *idx, hdm::PairCtr *ctrs,
                                               int npairs, uint64_t total_tiles,
                                               const uint32_t *const *gfs,
                                               const uint32_t *const *sfs, int grid);
// server admission group: pair src[i] of the group into engine slot dst[i]
// (counters, and the descriptor with generation gen[i] and no start yet)
extern "C" hipError_t hdm_launch_move_pairs(hipStream_t s, const hdm::PairDesc *gdescs,
                                            const hdm::PairCtr *gctrs, hdm::PairDesc *descs,
                                            hdm::PairCtr *ctrs, hdm::PackIndex *idx,
                                            const int *src, const int *dst, const uint32_t *gen,
                                            int n);
// descs[slots[i]].t_start = t_start for i < n
extern "C" hipError_t hdm_launch_set_start(hipStream_t s, hdm::PairDesc *descs, const int *slots,
                                           int n, uint32_t t_start);
