// Internal host-side helpers shared by the single-pair, batch and serving drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/hdm.h"
#include "hdm_plan.h"

namespace hdm {

// grow-only device buffer
struct DevMem {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t n);
  void release();
};

// per-device state of the single-pair entry points
struct Ctx {
  int dev = -1;
  hipStream_t stream = nullptr;
  int cus = 0;
  int occ[kLaunchIds] = {};        // resident dilate workgroups per CU, several pairs
  int occ_single[kLaunchIds] = {}; // the same for the one-pair dilation kernel
  int occ_pack[kLaunchIds] = {};  // resident pack workgroups per CU
  int occ_fused[kLaunchIds] = {};  // resident fused single-pair workgroups per CU
  int occ_fused_ds[kLaunchIds] = {};  // the same, two dilation steps per barrier
  int occ_fused_tb[kLaunchIds] = {};  // the same, regions of eight steps per barrier
  int occ_fused_near[kLaunchIds] = {};  // the same, pack and near pass
  double clock_khz = 0;               // wall_clock64() rate
  DevMem ws, in, map;
  // region mark bitmaps of the near launch, two sets used by turns
  // (marks_set: the one the next near launch fills; it zeroes the other);
  // cleared whenever the buffer is allocated
  DevMem marks;
  int marks_set = 0;
  void *host_out = nullptr;  // pinned, 4 KiB
  hipEvent_t ev[5] = {};     // phase boundaries of a timed call
  // pack, seed, dilation, 0, all; 1 when the call was one fused launch;
  // dilation steps per grid barrier (1 or 2)
  float last_ms[7] = {};
  // epoch base of the next fused launch's barrier tags; the tagged words of
  // the workspace's engine are cleared whenever the base restarts at 0
  uint32_t ebase = 0;
  void *ebase_ws = nullptr;  // the workspace the base belongs to
  // the set of Engine::acc the next fused launch packs into; each launch
  // zeroes the other set, and clearing the engine zeroes both
  int acc_set = 0;
  std::vector<uint64_t> last_trace;  // 8 words per iteration of a traced call
  DevMem trace;
  std::mutex mu;
  // A call that took its result from the kernel's ready marker returns while
  // the launch may still be leaving: its stream, waited for before the next
  // call uses another stream or reallocates a buffer (settle)
  uint32_t seq = 0;
  // timing 3: one event pair per call from a ring; the durations are read
  // when asked (hdm_last_timings, hdm_timing_log), so the call need not wait
  static constexpr uint32_t kRing = 1024;
  std::vector<hipEvent_t> ring;  // 2 * kRing, created on first use
  uint32_t ring_next = 0;
  std::vector<uint32_t> log;     // pairs recorded since the log was read, oldest first
  int last_pair = -1;            // pair of the last timing-3 call, not yet read
  bool last_pair_fused = false;
  hipStream_t pending = nullptr;
  bool has_pending = false;
  int settle() {
    if (!has_pending) return HDM_OK;
    has_pending = false;
    return hipStreamSynchronize(pending) == hipSuccess ? HDM_OK : HDM_E_HIP;
  }
};

// launch shapes and optional phase events of one enqueue
struct LaunchCfg {
  unsigned pack_blocks = 1, seed_blocks = 1;
  int grid = 1, chunk = 1;
  uint64_t spin = 0;
  const hipEvent_t *ev = nullptr;  // 4 events (phase bounds) or null
  int ev_mode = 0;                 // opts.timing: 1 or 2 every phase, 3 the dilation only
  uint64_t *trace = nullptr;       // per-iteration device trace or null
  PairOut *outs = nullptr;         // where the results go (null: the group's outs)
  EngineHead *head = nullptr;      // where the engine head goes (null: nowhere)
  uint32_t *ready = nullptr;       // the dilation's ready marker (host) or null
  uint32_t ready_seq = 0;
  // batches: the near pass between the totals and the seed (pairs' bitmaps of
  // near_words words in all, an entry list of near_cap entries), or off
  uint64_t *near_list = nullptr;
  uint64_t near_words = 0;
  uint32_t near_cap = 0;
  int near_grid = 0;
};

struct GroupPtrs {
  PairDesc *descs = nullptr;
  PackIndex *idx = nullptr;
  PairCtr *ctrs = nullptr;
  PairOut *outs = nullptr;
  Engine *engine = nullptr;
  uint64_t *q = nullptr;
  uint32_t *qstats = nullptr;
};

int map_hip(hipError_t e);
int get_ctx(Ctx **out);
int resident_grid(const Ctx *c, int tile, int requested, int npairs);
uint64_t spin_ticks(const Ctx *c);
int norm_tile(int tile);
// launch id (hdm_tile_launch) of a tile option and a connectivity option (0 =
// 6), or -1 when they are invalid; *conn receives the element
int select_launch(int tile_opt, int conn_opt, int *conn, bool ds = false);
int barrier_mode();
// Spins until the launch on s has stored seq into *ready (pinned coherent
// host memory), checking the stream now and then so a launch that ended
// without it is an error, not a hang.
int wait_ready(hipStream_t s, volatile uint32_t *ready, uint32_t seq);
bool fused_enabled();
// the fused single-pair launch is cooperative (HDM_COOP=1) or plain
bool coop_enabled();
bool ds_enabled();
// the two-step fused single-pair launch (64x8x8 tiles, 6 element) is
// followed by the region launch, to which it hands long pairs, unless HDM_TB=0
bool tb_enabled();
// a single pair on 64x8x8 tiles with the 6 element and no first-hit map runs
// the near launch first (pack, then eight steps around the regions holding
// voxels to be covered), which hands pairs with HD > 8 to the region launch,
// unless HDM_NEAR=0 (then the two-step launch runs as before)
bool near_enabled();
// DilateArgs::pack_pools of the fused launches: 8 (a counter per XCD for the
// pack's last quarter), or 0 with HDM_PACK_POOLS=0 (even static shares)
int pack_pools();
// regions beyond which a pair skips the near launch: every workgroup of it
// reads the mark bits of every region (16 words per thread at most)
constexpr uint32_t kNearMaxRegions = 1u << 16;
// DilateArgs::tb_mode of the two-step launch: 0, 1 (by the HD bound) or 2
// (every pair), from HDM_TB
int tb_mode();
int fused_grid(const Ctx *c, int tile, int requested, bool ds, uint64_t ntiles);
// the same from the kernel's resident workgroups per CU and the CU count
int fused_grid_of(int occ, int cus, int requested, uint64_t ntiles);
int near_grid_of(int occ, int cus, int requested, uint64_t ntiles);
// the grid of the region launch (regions of 1 x 2 x 2 tiles)
int tb_grid(int occ, int cus, int requested, uint64_t ntiles, uint32_t nregions);
GroupPtrs carve_common(char *base, const CommonLayout &L);
hipError_t enqueue_group(hipStream_t s, const TileLaunch *tl, const GroupPtrs &P, int npairs,
                         const PairDesc *single, const PackIndex *single_idx,
                         uint64_t total_rw, uint64_t total_tiles, uint32_t qcap,
                         const LaunchCfg &cfg, int qstats_len);
LaunchCfg launch_cfg(const Ctx *c, int tile, uint64_t total_rw, uint64_t total_tiles,
                     int grid_req, int chunk, int npairs);
void fill_result(hdm_result *out, const PairOut &o, const EngineHead *es, uint32_t tiles,
                 int iterations);

}  // namespace hdm
