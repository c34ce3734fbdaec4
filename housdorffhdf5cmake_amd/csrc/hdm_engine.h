// Device-side data structures shared by the kernels and the host driver.
//
// Layout in HBM, per pair (tile = TXB(x) * TY(y) * TZ(z) voxels, R = TY*TZ rows
// of TXB bits; one wavefront lane owns one row):
//   orig[m]          packed mask m (0 gold, 1 segm), tile-major: word
//                    tile*R + (y%TY) + TY*(z%TZ) holds voxels x = TXB*tx .. +TXB-1
//   buf[m][s]        two ping-pong slots for the dilated mask m
//   xf_*[m]          per tile two R-bit words: bit r = bit 0 / bit TXB-1 of row r
//                    (the tile's -x / +x faces), so a neighbour reads one word
//                    instead of the whole tile
//   state[m][tile]   0 idle; bit 31 full; else 1 + the pair-local iteration at
//                    which the tile was first queued
//   hist[m][t]       number of voxels of the other mask first covered at t
// Where a neighbour's state at iteration t lives: full -> all ones; queued
// before t -> buf[m][(t-1)&1]; otherwise (idle, or first queued at t) -> orig.
//
// Inside the persistent kernel every word that one workgroup writes and
// another reads (tile rows, face words, states, queue entries, counters) is
// accessed with agent-scope atomic loads and stores, so the grid barrier
// needs no L2 write-back or invalidate.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hdm {

constexpr int kWave = 64;
constexpr int kWgThreads = 256;
constexpr int kWavesPerWg = kWgThreads / kWave;
constexpr int kLocalQueue = 2048;  // entries staged in LDS per workgroup
constexpr int kSubQ = 16;          // sub-queues: workgroup b flushes into b % kSubQ
constexpr int kHitSlots = 16;      // per (pair, mask) per-iteration hit counters
constexpr int kSumSlots = 64;      // per (pair, field) partial voxel counts
constexpr int kMaxGrid = 4096;     // persistent workgroups (arrival flags)
constexpr int kSlotListMax = 64;   // server slots per start-iteration update launch
constexpr int kLineWords = 32;     // one 128-byte line
// x-face words per tile: bit y + TY z of word [side] is the voxel of row (y, z)
// in column 0 (side 0) or TXB-1 (side 1); 64x8x8 tiles also keep the second
// column from each face in words [2 + side] (columns 1 and TXB-2), the part
// of a neighbour that two dilation steps reach
constexpr int xf_words(int txb, int ty, int tz) {
  return (txb == 64 && ty == 8 && tz == 8) ? 4 : 2;
}
constexpr int kReleaseCopies = 32; // replicas of the barrier release line
// dilation steps per grid barrier of the single-pair region kernel, and the
// words of a workgroup's barrier arrival record (hits per pass and step,
// entries pushed and kept)
constexpr int kTbSteps = 8;
constexpr int kArrWords = 2 + 2 * kTbSteps;

constexpr uint32_t kFull = 0x80000000u;
constexpr uint32_t kNoTile = 0xFFFFFFFFu;

// queue entry: tile (32) | pair (16) | mask (1) | first (1) | check (1) |
// generation of the pair slot (13): a serving slot reused by a new request
// gets a new generation, so entries left by its previous owner are dropped
constexpr uint32_t kGenMask = 0x1FFF;
__host__ __device__ inline uint64_t make_entry(uint32_t tile, uint32_t pair, uint32_t m,
                                               uint32_t first, uint32_t check, uint32_t gen) {
  return (uint64_t)tile | ((uint64_t)pair << 32) | ((uint64_t)m << 48) |
         ((uint64_t)first << 49) | ((uint64_t)check << 50) | ((uint64_t)(gen & kGenMask) << 51);
}
__host__ __device__ inline uint32_t entry_gen(uint64_t e) { return (uint32_t)(e >> 51) & kGenMask; }

struct PairDesc {
  const void *gold;
  const void *segm;
  int32_t nx, ny, nz, wxn;  // wxn: TXB-bit words per x-row
  int32_t gx, gy, gz;       // tile grid
  uint32_t ntiles;
  int32_t dtype, label;
  int32_t hist_len;
  uint32_t t_start;  // engine iteration at which the pair entered
  int32_t crop;
  int32_t tile_kind;
  uint32_t gen;  // slot generation carried by this pair's queue entries
  int32_t conn;  // structuring element: 6 (cross), 18 (faces, xy diagonals, corners), 26 (cube)
  double robust;
  uint64_t words;  // ntiles * R, words per packed mask
  void *orig[2];
  void *buf[2][2];
  uint64_t *xf_orig[2];
  uint64_t *xf_buf[2][2];
  uint32_t *state[2];
  // per region of 1 x 2 x 2 tiles (the eight-step launch): 0 idle (data =
  // orig), kFull, else 1 + the period at which it was first queued
  uint32_t *rstate[2];
  // per tile, low 16 bits gold / high 16 bits segm: bit 0 non-empty, 1 full,
  // 2..7 voxels on the -x, +x, -y, +y, -z, +z face, 8..13 voxels in the two
  // outermost layers toward -x .. +z, 14 / 15 voxels in the eight outermost
  // columns toward -x / +x (written by the pack pass; 8..15 on 64x8x8 tiles)
  uint32_t *tflags;
  uint32_t *hist[2];
  uint16_t *map;
  // near pass (64x8x8 tiles, 6 element): the pair's bitmap of 2 * regions bits
  // (regions of 1 x 2 x 2 tiles; bit m * regions + rg: region rg holds voxels
  // pass m covers), zero before the pack sets it; null: no marks
  uint32_t *marks;
};
// regions of 1 x 2 x 2 tiles of a pair's 64x8x8 tile grid
__host__ __device__ inline uint32_t pair_regions(const PairDesc &d) {
  return (uint32_t)d.gx * (uint32_t)((d.gy + 1) / 2) * (uint32_t)((d.gz + 1) / 2);
}
// words of a pair's near-pass bitmap
__host__ __device__ constexpr uint32_t near_mark_words(uint32_t nreg) { return (2 * nreg + 31) / 32; }

struct PairCtr {
  unsigned long long n_gold, n_segm, n_fp, n_fn;  // totals (written before the dilation)
  unsigned long long cov[2];                      // covered so far; barrier master only
  // the part of cov covered by the second of two dilation steps per engine
  // iteration (barrier master only; zero with one step per iteration)
  unsigned long long cov2[2];
  // Sum of first-hit iterations of the covered voxels and the last iteration
  // with a hit, per pass; the master keeps them with cov while it folds, so
  // they cover exactly the histogram bins up to the pass's stop.
  unsigned long long sum_d[2];
  int32_t lo[3], hi[3];  // tile bounding box of non-empty tiles (inclusive)
  int32_t status;
  // First pair-local iteration at which pass m has nothing left to do; written
  // by the barrier master between iterations, so every workgroup reads the
  // same value during an iteration.
  int32_t done_from[2];
  uint32_t hd_last[2];
  // 1 when the batched near pass finished the pair (both passes' cuts reached
  // within eight steps): its counters hold the results, and the seed, the
  // dilation and its fold leave it alone
  int32_t near_done;
  // cumulative hits, spread over slots: [m] pass m, or with two dilation
  // steps per engine iteration [2k + m] pass m in step k of each iteration
  unsigned long long hits[4][kHitSlots];
  unsigned long long sums[4][kSumSlots];        // pack partials: gold, segm, fp, fn
  int32_t box[6][kSumSlots];  // pack partials of lo[0..2] and hi[0..2] (box_enc_*)
  // batched near pass: new hits per (pass m, step k) at index m + 2k, spread
  // over four slots
  unsigned long long near_hits[2 * kTbSteps][4];
};

struct PairOut {
  int32_t hd, hd_fp, hd_fn, status;
  unsigned long long sum_fp, cnt_fp, sum_fn, cnt_fn;
  unsigned long long n_gold, n_segm, n_fp, n_fn;
  int32_t done;  // both passes finished at the engine's current iteration
  int32_t iters; // dilation iterations: the last one a pass needed, or all run
  int32_t steps; // dilation steps per engine iteration of the body that ran
  int32_t pad;
};

// The engine fields the host reads after a launch; the dilation kernel writes
// a copy next to the pair results when it leaves.
// Pack partials with zero as the empty value, so zeroed memory is a valid
// start: voxel counts (gold, segm, FP, FN) and the box of non-empty tiles,
// lo[k] as box_enc_lo (0x7FFFFFFF - lo) and hi[k] as box_enc_hi (hi + 1), both
// combined by max.
__host__ __device__ inline int32_t box_enc_lo(int32_t v) { return 0x7FFFFFFF - v; }
__host__ __device__ inline int32_t box_enc_hi(int32_t v) { return v + 1; }
__host__ __device__ inline int32_t box_dec_lo(int32_t e) { return 0x7FFFFFFF - e; }
__host__ __device__ inline int32_t box_dec_hi(int32_t e) { return e - 1; }
struct PackSums {
  unsigned long long sums[4][kSumSlots];
  // rows 0..5 the box of the tiles with voxels of either mask (lo x, y, z,
  // hi x, y, z), 6..11 of gold, 12..17 of segm
  int32_t box[18][kSumSlots];
  // pack steps handed out one at a time past the waves' static shares: one
  // counter per pool ([32 p], each in a cache line of its own, away from the
  // partials' atomics)
  alignas(128) uint32_t grab[8 * 32];
  // near pass of the single-pair launch: hits per (pass, step), index m + 2k,
  // spread over kHitSlots by workgroup, and the workgroups' arrivals, so the
  // last workgroup to arrive folds the counts.  Zeroed with the set, by the
  // launch before, so a workgroup that arrives early (after a barrier that
  // timed out) still counts from zero.
  unsigned long long near_hits[2 * kTbSteps][kHitSlots];
  uint32_t near_done;
};

struct EngineHead {
  uint32_t t;      // completed engine iterations
  uint32_t error;  // 1: a barrier wait exceeded its limit
  uint32_t qmax;
  uint32_t nwg;
  unsigned long long qtotal;
  uint32_t qcap;
  uint32_t qnext;  // entries queued for the next iteration when the kernel left
  // fused single-pair launch: wall-clock ticks at its start and at the end of
  // its pack, seed and dilation phases (0 otherwise)
  unsigned long long phase[4];
};

struct Engine {
  // head (EngineHead layout)
  uint32_t t;      // completed engine iterations
  uint32_t error;  // 1: a barrier wait exceeded its limit
  uint32_t qmax;
  uint32_t nwg;
  unsigned long long qtotal;
  uint32_t qcap;
  uint32_t qnext;  // entries queued for the next iteration when the kernel left
  uint32_t pad1[8];
  // device only
  // [t % 4][sub-queue]: iteration t reads set t%4 and fills (t+1)%4; the
  // barrier master zeroes (t+2)%4 at the end of t, a full barrier before use
  uint32_t qlen[4][kLineWords];
  uint32_t arrive[kMaxGrid];           // last epoch each workgroup arrived at
  // Release lines, kReleaseCopies replicas so polling spreads over channels:
  // word 0 the epoch, words 1..kSubQ+1 the prefix offsets of the sub-queues of
  // the iteration the epoch starts.  Workgroup b polls replica b % kReleaseCopies.
  uint32_t rel[kReleaseCopies][kLineWords];
  // Single-pair barrier: every word carries its epoch in the top 24 bits
  // (tag()), so a reader that sees the right tag has the value that goes with
  // it and no ordering between words is needed.  arr[j][wg], word j of
  // workgroup wg's record: this iteration's hits of pass 0 and 1 in its first
  // dilation step, the entries the workgroup pushed and kept, then the hits
  // of pass 0 and 1 in each further step of the iteration (two or eight steps
  // per barrier).  Word-major, so the master's poll of word j over 64
  // consecutive workgroups is one load of four cache lines, not 64 requests.
  // rel2[copy]: word 0 flags (bit m: pass m done, bit 2: error), words
  // 1..kSubQ+1 prefix offsets of the next queue.
  unsigned long long arr[kArrWords][kMaxGrid];
  unsigned long long rel2[kReleaseCopies][kLineWords];
  // tag(epoch, first global entry) of each workgroup's free wave slots: the
  // master's prefix over workgroups of (slots - entries the workgroup kept)
  unsigned long long fslot[kMaxGrid];
  // pack partials of the fused single-pair launch, two sets used by turns:
  // a launch packs into one and zeroes the other for the next launch, so no
  // barrier has to separate the zeroing from the pack
  PackSums acc[2];
  // the two-step launch hands a packed pair to the region launch queued
  // behind it: word 0 tag(epoch base, 2), words 1 and 2 its clock stamps of
  // the start and the end of the pack
  unsigned long long handoff[3];
  // near pass of the single-pair launch: workgroup 0's clock stamps of its
  // start, the end of its pack and the pack barrier, for the engine head
  unsigned long long near_phase[3];
  // batched near pass: entries written to the list (zeroed by the reset)
  uint32_t near_count;
  // the multi-pair pack's step counters, one per XCD in a cache line of its
  // own (zeroed by the reset; see PackSums::grab)
  alignas(128) uint32_t pack_grab[8 * 32];
};
constexpr int kTagShift = 40;
constexpr unsigned long long kTagValue = (1ull << kTagShift) - 1;
__host__ __device__ inline unsigned long long tag(uint32_t epoch, unsigned long long v) {
  return ((unsigned long long)(epoch & 0xFFFFFFu) << kTagShift) | (v & kTagValue);
}
__host__ __device__ inline bool tagged(unsigned long long w, uint32_t epoch) {
  return (uint32_t)(w >> kTagShift) == (epoch & 0xFFFFFFu);
}
constexpr size_t kEngineHostBytes = sizeof(EngineHead);
static_assert(sizeof(EngineHead) == 64, "engine head is one 64-byte block");

// ceil(r * n) clamped to [0, n]; the robust cut (old/testMainPassesKernels.cu:81-82)
__host__ __device__ inline unsigned long long robust_cut(double r, unsigned long long n) {
  if (n == 0) return 0;
  double c = ceil(r * (double)n);
  if (c < 0.0) c = 0.0;
  if (c > (double)n) c = (double)n;
  return (unsigned long long)c;
}

}  // namespace hdm
