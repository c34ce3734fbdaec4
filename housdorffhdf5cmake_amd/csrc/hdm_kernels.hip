// HIP kernels of the Hausdorff dilation engine, written for gfx950 (CDNA4,
// wave64).  Integer bit logic only; the path is bound by memory traffic and
// by the latency of the per-iteration grid barrier.
//
//   reset_kernel     clears per-pair counters, engine state, histograms
//   pack_kernel      dense labels -> packed tile-major bits, x-face words and
//                    tile flags, partial voxel counts and the tile bounding
//                    box, in one read of the input (replaces getMinMaxes and
//                    boolPrepareKernel, hello.cu:396-526, 537-748)
//   totals_kernel    folds the partial counts into |gold|, |segm|, |FP|, |FN|
//                    (batches; a single pair's seed pass does it itself)
//   seed_kernel      tile states and the first work queue from the tile flags
//                    (firstMetaPrepareKernel, hello.cu:762-877)
//   dilate_kernel    persistent kernel: one grid barrier per iteration, a
//                    queue of the tiles still changing (mainPassKernel,
//                    hello.cu:1066-1896); when it leaves, its workgroup 0
//                    writes every pair's HD, sums and counts (finalize_pair).
//                    The one-pair variant can run the reset, pack and seed
//                    itself first (FUSED: one launch per call); on 64x8x8
//                    tiles with the 6 element an iteration is two dilation
//                    steps (DS; dilate_ds_kernel is the fused one-pair form)
//
// Same-address atomics serialize (tens of ns each on MI355X), so every hot
// counter is spread: partial voxel counts over kSumSlots, per-iteration hits
// over kHitSlots after aggregation in LDS, queue flushes over kSubQ
// sub-queues, and the grid barrier uses one arrival flag per workgroup.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "hdm_plan.h"

namespace hdm {

#define AGENT __HIP_MEMORY_SCOPE_AGENT

template <int TXB>
using RowT = std::conditional_t<TXB == 64, uint64_t, uint32_t>;

// Every pointer these helpers take is global memory.  Pointers read out of a
// descriptor are generic to the compiler, which would emit flat accesses
// (waited for on both the vector and the LDS counters); the casts make them
// global_load / global_store / global_atomic.
#define GLB __attribute__((address_space(1)))

// agent-coherent load / store (global_load / global_store with sc1)
template <typename T>
__device__ __forceinline__ T ld_sc(const T *p) {
  return __hip_atomic_load((GLB T *)p, __ATOMIC_RELAXED, AGENT);
}
template <typename T>
__device__ __forceinline__ void st_sc(T *p, T v) {
  __hip_atomic_store((GLB T *)p, v, __ATOMIC_RELAXED, AGENT);
}
// plain global load (data no launch in flight writes: packed originals)
template <typename T>
__device__ __forceinline__ T ldg(const T *p) {
  return *(const GLB T *)p;
}
// compare-and-swap at agent scope; returns the value found
__device__ __forceinline__ uint32_t cas_g(uint32_t *p, uint32_t expect, uint32_t v) {
  __hip_atomic_compare_exchange_strong((GLB uint32_t *)p, &expect, v, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, AGENT);
  return expect;
}
__device__ __forceinline__ void or_g(uint32_t *p, uint32_t v) {
  __hip_atomic_fetch_or((GLB uint32_t *)p, v, __ATOMIC_RELAXED, AGENT);
}

__device__ __forceinline__ uint32_t shfl32(uint32_t v, int src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = shfl32((uint32_t)v, src), hi = shfl32((uint32_t)(v >> 32), src);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
template <typename Row>
__device__ __forceinline__ Row shfl_row(Row v, int src) {
  if constexpr (sizeof(Row) == 8)
    return shfl64(v, src);
  else
    return shfl32(v, src);
}
template <typename Row>
__device__ __forceinline__ int popc_row(Row v) {
  if constexpr (sizeof(Row) == 8)
    return __popcll(v);
  else
    return __popc(v);
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-wide reductions through DPP moves (VALU latency, where a butterfly of
// ds_bpermute shuffles waits on the LDS pipe at every level): an inclusive
// scan inside each row of 16 lanes (row_shr 1, 2, 4, 8), then the total of
// row 0 into row 1 and of row 2 into row 3 (row_bcast 15), then rows 0-1
// into rows 2-3 (row_bcast 31); lane 63 holds the result, which every lane
// reads back (readlane).  A source lane outside the row, or a row outside the
// step's row mask, contributes the identity (bound_ctrl off: the move keeps
// its old value).  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp32(int v, int identity) {
  return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  const uint32_t lo = (uint32_t)dpp32<CTRL, ROWS>((int)(uint32_t)v, 0);
  const uint32_t hi = (uint32_t)dpp32<CTRL, ROWS>((int)(uint32_t)(v >> 32), 0);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
template <typename Op>
__device__ __forceinline__ int wave_reduce32(int v, int identity, Op op) {
  v = op(v, dpp32<0x111, 0xF>(v, identity));  // row_shr:1
  v = op(v, dpp32<0x112, 0xF>(v, identity));  // row_shr:2
  v = op(v, dpp32<0x114, 0xF>(v, identity));  // row_shr:4
  v = op(v, dpp32<0x118, 0xF>(v, identity));  // row_shr:8
  v = op(v, dpp32<0x142, 0xA>(v, identity));  // row_bcast:15 into rows 1, 3
  v = op(v, dpp32<0x143, 0xC>(v, identity));  // row_bcast:31 into rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}
// inclusive prefix sum over the 64 lanes (the same DPP steps, no read-back)
__device__ __forceinline__ uint32_t wave_scan32(uint32_t x) {
  int v = (int)x;
  v += dpp32<0x111, 0xF>(v, 0);
  v += dpp32<0x112, 0xF>(v, 0);
  v += dpp32<0x114, 0xF>(v, 0);
  v += dpp32<0x118, 0xF>(v, 0);
  v += dpp32<0x142, 0xA>(v, 0);
  v += dpp32<0x143, 0xC>(v, 0);
  return (uint32_t)v;
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  v += dpp64<0x111, 0xF>(v);
  v += dpp64<0x112, 0xF>(v);
  v += dpp64<0x114, 0xF>(v);
  v += dpp64<0x118, 0xF>(v);
  v += dpp64<0x142, 0xA>(v);
  v += dpp64<0x143, 0xC>(v);
  return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63) << 32);
}

// A 64-bit row from lane - 1, + 1, - 2 or + 2 inside the same row of 16
// lanes (DPP row_shr / row_shl; a source outside the row reads 0): for the y
// neighbours of a 64x8x8 tile's rows, whose 8-lane y groups never straddle a
// row of 16, where a ds_bpermute shuffle would wait on the LDS pipe.
template <int CTRL>
__device__ __forceinline__ uint64_t row_dpp64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, true);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, true);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
constexpr int kFromLeft1 = 0x111, kFromLeft2 = 0x112;    // row_shr: lane - 1, lane - 2
constexpr int kFromRight1 = 0x101, kFromRight2 = 0x102;  // row_shl: lane + 1, lane + 2

// sum over the aligned group of R lanes (R a power of two <= 64), in every
// lane of the group; a whole wave (R = 64) must be active
template <int R>
__device__ __forceinline__ uint32_t group_sum(uint32_t v, int lane) {
  if constexpr (R == 64) {
    return (uint32_t)wave_reduce32((int)v, 0,
                                   [](int a, int b) { return (int)((uint32_t)a + (uint32_t)b); });
  } else {
#pragma unroll
    for (int off = R / 2; off >= 1; off >>= 1) v += shfl32(v, lane ^ off);
    return v;
  }
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v, int lane) { return group_sum<64>(v, lane); }
__device__ __forceinline__ int wave_min(int v, int) {
  return wave_reduce32(v, INT_MAX, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max(int v, int) {
  return wave_reduce32(v, INT_MIN, [](int a, int b) { return max(a, b); });
}

// bits of the R-lane group holding `lane` out of a wave ballot, shifted to bit 0
template <int R>
__device__ __forceinline__ uint64_t group_bits(uint64_t ballot, int lane) {
  if constexpr (R == 64)
    return ballot;
  else
    return (ballot >> (lane & ~(R - 1))) & ((1ull << R) - 1);
}

template <typename Row>
__device__ __forceinline__ Row valid_bits(int tx, int nx) {
  constexpr int TXB = sizeof(Row) * 8;
  const int rem = nx - tx * TXB;
  if (rem >= TXB) return ~Row(0);
  return (Row(1) << rem) - 1;
}

// pair index of a global position by binary search over ascending bases
template <typename F>
__device__ __forceinline__ int find_pair(int npairs, uint64_t g, F base_of) {
  int lo = 0, hi = npairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base_of(mid) <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint64_t *sub_queue(uint64_t *q, uint32_t qcap, uint32_t parity,
                                               uint32_t k) {
  return q + ((uint64_t)(parity * kSubQ + k)) * qcap;
}

// ---------------------------------------------------------------- reset
// Blocks stride over the pairs: each zeroes a pair's counters and histograms
// with all its threads.  Block 0 also resets the engine, and stores the
// descriptor of a single-pair call (passed by value, so this launch reads the
// histogram pointers from the argument, not from the descriptor it writes).
__global__ __launch_bounds__(kWgThreads) void reset_kernel(
    PairDesc *descs, PairDesc single, PackIndex single_idx, PackIndex *idx, int write_single,
    PairCtr *ctrs, int p0, int npairs, Engine *es, int reset_engine, uint32_t qcap) {
  const int tid = threadIdx.x;
  if (write_single && blockIdx.x == 0 && tid == 0) {
    descs[p0] = single;
    idx[p0] = single_idx;
  }
  if (reset_engine) {  // spread over the grid's blocks
    const int g = blockIdx.x * blockDim.x + tid, gs = gridDim.x * blockDim.x;
    if (g == 0) {
      es->t = 0;
      es->error = 0;
      es->qmax = 0;
      es->qtotal = 0;
      es->qcap = qcap;
      es->qnext = 0;
      es->near_count = 0;
    }
    if (g < 8) es->pack_grab[32 * g] = 0;
    for (int i = g; i < 4 * kLineWords; i += gs) es->qlen[0][i] = 0;
    for (int i = g; i < kMaxGrid; i += gs) es->arrive[i] = 0;
    for (int i = g; i < kReleaseCopies * kLineWords; i += gs) es->rel[0][i] = 0;
    constexpr int kArrWords = (int)(sizeof(es->arr) / sizeof(es->arr[0][0]));
    for (int i = g; i < kArrWords; i += gs) (&es->arr[0][0])[i] = 0;
    for (int i = g; i < kMaxGrid; i += gs) es->fslot[i] = 0;
    for (int i = g; i < kReleaseCopies * kLineWords; i += gs) es->rel2[0][i] = 0;
  }
  for (int p = p0 + blockIdx.x; p < p0 + npairs; p += gridDim.x) {
    PairCtr &c = ctrs[p];
    if (tid == 0) {
      c.n_gold = c.n_segm = c.n_fp = c.n_fn = 0;
      c.cov[0] = c.cov[1] = 0;
      c.cov2[0] = c.cov2[1] = 0;
      c.sum_d[0] = c.sum_d[1] = 0;
      c.lo[0] = c.lo[1] = c.lo[2] = 0x7FFFFFFF;
      c.hi[0] = c.hi[1] = c.hi[2] = -1;
      c.status = 0;
      c.done_from[0] = c.done_from[1] = 0x7FFFFFFF;
      c.hd_last[0] = c.hd_last[1] = 0;
    }
    for (int i = tid; i < 4 * kHitSlots; i += blockDim.x) (&c.hits[0][0])[i] = 0;
    for (int i = tid; i < 4 * kSumSlots; i += blockDim.x) (&c.sums[0][0])[i] = 0;
    if (tid == 0) c.near_done = 0;
    for (int i = tid; i < 2 * kTbSteps * 4; i += blockDim.x) (&c.near_hits[0][0])[i] = 0;
    {  // the pair's near-pass bitmap, which the pack then sets
      const PairDesc &dd = write_single ? single : descs[p];
      if (dd.marks) {
        const uint32_t nw = near_mark_words(pair_regions(dd));
        for (uint32_t i = tid; i < nw; i += blockDim.x) dd.marks[i] = 0;
      }
    }
    for (int i = tid; i < 6 * kSumSlots; i += blockDim.x)
      (&c.box[0][0])[i] = 0;
    uint32_t *h0 = write_single ? single.hist[0] : descs[p].hist[0];
    uint32_t *h1 = write_single ? single.hist[1] : descs[p].hist[1];
    const int len = write_single ? single.hist_len : descs[p].hist_len;
    for (int i = tid; i < len; i += blockDim.x) h0[i] = 0, h1[i] = 0;
  }
}

// Pass m of a pair has nothing left to do at pair-local iteration t: its cut
// is zero, or the cut was reached by the hits of an iteration before t.
__device__ __forceinline__ bool pass_done(const PairDesc &d, const PairCtr &c, int m, int t) {
  if (robust_cut(d.robust, m == 0 ? c.n_fp : c.n_fn) == 0) return true;
  return ld_sc(&c.done_from[m]) <= t;
}

// ---------------------------------------------------------------- pack
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// 16 labels (bytes) -> 16 bits: bit k set iff byte k == label.
__device__ __forceinline__ uint32_t eq_bits_u8x16(u32x4 v, uint32_t lab4) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  uint32_t out = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t x = w[j] ^ lab4;                           // zero byte where equal
    const uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;  // byte msb = byte != 0
    const uint32_t eq = (~t & 0x80808080u) >> 7;               // bit 8k = byte k equal
    out |= ((eq * 0x10204080u) >> 28) << (4 * j);
  }
  return out;
}

// The label compare of hello.cu:453 / 615 for TXB consecutive voxels of a row.
template <typename Row>
__device__ __forceinline__ Row load_row_bits(const PairDesc &d, const void *vol, uint64_t base,
                                             int count) {
  constexpr int TXB = sizeof(Row) * 8;
  Row bits = 0;
  if (d.dtype == 0) {
    const uint8_t *p = (const uint8_t *)vol + base;
    const uint32_t lab = (uint8_t)d.label;
    if (count == TXB && (((uintptr_t)p) & 15) == 0) {
      const u32x4 *q = (const u32x4 *)p;  // read once: nontemporal, as the pack's pairs
#pragma unroll
      for (int k = 0; k < TXB / 16; ++k)
        bits |= (Row)eq_bits_u8x16(__builtin_nontemporal_load(q + k), lab * 0x01010101u)
                << (16 * k);
    } else {
      for (int k = 0; k < count; ++k) bits |= (Row)(p[k] == lab) << k;
    }
  } else {
    const int32_t *p = (const int32_t *)vol + base;
    const int32_t lab = d.label;
    if (count == TXB && (((uintptr_t)p) & 15) == 0) {
      const i32x4 *q = (const i32x4 *)p;
#pragma unroll
      for (int k = 0; k < TXB / 4; ++k) {
        const i32x4 v = __builtin_nontemporal_load(q + k);
        bits |= (Row)(v.x == lab) << (4 * k);
        bits |= (Row)(v.y == lab) << (4 * k + 1);
        bits |= (Row)(v.z == lab) << (4 * k + 2);
        bits |= (Row)(v.w == lab) << (4 * k + 3);
      }
    } else {
      for (int k = 0; k < count; ++k) bits |= (Row)(p[k] == lab) << k;
    }
  }
  return bits;
}

// Per-wave accumulator of the pack statistics of one pair.
struct PackAcc {
  uint32_t c[4];  // |gold|, |segm|, |FP|, |FN| bits seen by this lane
  // tile box of the tiles with voxels of either mask (0..2), of gold (3..5)
  // and of segm (6..8); the per-mask boxes only where the partials have room
  // for them (PackSums, the fused single-pair launch)
  int mn[9], mx[9];
};
__device__ __forceinline__ void acc_clear(PackAcc &a) {
  a.c[0] = a.c[1] = a.c[2] = a.c[3] = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) a.mn[k] = 0x7FFFFFFF, a.mx[k] = -1;
}
template <typename S>
constexpr int box_rows() {
  return (int)(sizeof(S::box) / sizeof(S::box[0]));
}
// into the pack partials (sums, box) of a PairCtr or a PackSums
template <typename S>
__device__ __forceinline__ void acc_publish(S &c, const PackAcc &a, int slot) {
  constexpr int kBoxes = box_rows<S>() / 6;  // 1 (either mask) or 3 (also per mask)
#pragma unroll
  for (int f = 0; f < 4; ++f)
    if (a.c[f]) atomicAdd(&c.sums[f][slot], (unsigned long long)a.c[f]);
#pragma unroll
  for (int bx = 0; bx < kBoxes; ++bx) {
    if (a.mx[3 * bx] >= 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        atomicMax(&c.box[6 * bx + k][slot], box_enc_lo(a.mn[3 * bx + k]));
        atomicMax(&c.box[6 * bx + 3 + k][slot], box_enc_hi(a.mx[3 * bx + k]));
      }
    }
  }
}
// wave-reduce the accumulator, lane 0 publishes
template <typename S>
__device__ __forceinline__ void acc_flush(S &c, PackAcc &a, int lane, int slot) {
  constexpr int kVals = box_rows<S>() / 2;  // 3 or 9 box coordinates
  PackAcc r;
#pragma unroll
  for (int f = 0; f < 4; ++f) r.c[f] = wave_sum(a.c[f], lane);
#pragma unroll
  for (int k = 0; k < kVals; ++k) r.mn[k] = wave_min(a.mn[k], lane), r.mx[k] = wave_max(a.mx[k], lane);
  if (lane == 0) acc_publish(c, r, slot);
  acc_clear(a);
}

// One wave per tile step (R lanes per tile, one lane per row): waves walk
// contiguous tile ranges so the pair index changes rarely.  A lane compares
// its row's TXB labels with the label (hello.cu:453 / 615), stores the packed
// word (the R words of a tile are contiguous, so the store is coalesced), and
// the tile's ballots give its x-face words and its flags for the seed pass:
// non-empty, full, and which of its six faces hold voxels.  Voxel counts and
// the box of non-empty tiles (getMinMaxes, hello.cu:396-526) accumulate per
// lane and are published once per (wave, pair) into spread slots.

// The near-pass mark of a 64x8x8 tile (tx, ty, tz) that holds FP (fp) or FN
// (fn) voxels: bit m * regions + rg of the pair's bitmap, for its region rg
// of 1 x 2 x 2 tiles (pass 0: FP, pass 1: FN), by an atomic OR.
__device__ __forceinline__ void mark_tile(const PairDesc &d, int tx, int ty, int tz, bool fp,
                                          bool fn) {
  const uint32_t rgx = (uint32_t)d.gx, rgy = (uint32_t)((d.gy + 1) / 2);
  const uint32_t nreg = pair_regions(d);
  const uint32_t rg = (uint32_t)tx + rgx * ((uint32_t)(ty >> 1) + rgy * (uint32_t)(tz >> 1));
  if (fp) or_g(d.marks + (rg >> 5), 1u << (rg & 31));
  if (fn) or_g(d.marks + ((nreg + rg) >> 5), 1u << ((nreg + rg) & 31));
}

// store that other workgroups of the same launch read later (COH: agent-
// coherent, written through) or that only later launches read (plain)
template <bool COH, typename T>
__device__ __forceinline__ void put(T *p, T v) {
  if constexpr (COH)
    st_sc(p, v);
  else
    *p = v;
}

// The pack pass over all tiles of the launch's pairs, spread over the grid's
// waves.  COH: the same launch reads the results in later phases (the fused
// single-pair launch), so every store is written through; the first-hit map
// is then cleared by the host instead.  MARK (64x8x8 tiles): a tile with FP or
// FN voxels of a pair whose descriptor carries a near-pass bitmap (marks)
// sets, by an atomic OR, the bit m * regions + rg of its region rg of
// 1 x 2 x 2 tiles, for pass m = 0 (FP = segm and not gold, the voxels gold's
// dilation covers) and m = 1 (FN).  grab (zeroed counters, 32 words apart,
// or null) and pools (1 or 8): the waves split three quarters of the steps
// evenly and take the rest one step at a time from a counter (with 8 pools,
// the counter of their workgroup's XCD, over its eighth of those steps), so
// CUs that stream faster pack more and the pack ends at about the same time
// on every CU.  LEAN (the near launch, whose only later reader is the region
// launch): no x-face words, and of the tile flags only bits 0, 1, 14 and 15
// per mask -- a quarter of the ballots per tile (512^3 pack about 66 -> 57 us
// measured alone: the waves compute between their steps' loads).  SKIP
// (LEAN's default; a batch with a near pass): no rows of a mask that is
// empty in a tile -- its readers take them as zero from the tile flags, or
// fill_kernel writes them for the pairs a batch's near pass leaves.
template <int TXB, int TY, int TZ, bool COH, bool MARK = false, bool LEAN = false,
          bool SKIP = LEAN>
__device__ __forceinline__ void pack_tiles(const PairDesc *__restrict__ descs,
                                           const PackIndex *__restrict__ idx, PairCtr *ctrs,
                                           int npairs, uint64_t total_tiles,
                                           uint16_t (*s_pieces)[1024], PackSums *alt = nullptr,
                                           uint32_t *grab = nullptr, int pools = 1) {
  static_assert(!MARK || (TXB == 64 && TY == 8 && TZ == 8), "near marks: 64x8x8 tiles");
  static_assert(!LEAN || MARK, "the lean pack is the near launch's");
  static_assert(!SKIP || MARK, "rows of empty masks are skipped on 64x8x8 tiles with marks");
  using Row = RowT<TXB>;
  constexpr int R = TY * TZ;
  constexpr int TPW = kWave / R;  // tiles per wave step
  constexpr int kXf = xf_words(TXB, TY, TZ);
  constexpr int U = 2;            // wave steps whose loads are all issued first
                                  // (the 128-byte path below pairs steps 0 and 1)
  const int lane = threadIdx.x & 63;
  const int sub = lane / R, r = lane % R;
  const int yl = r % TY, zl = r / TY;
  const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerWg +
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t nw = (uint64_t)gridDim.x * kWavesPerWg;
  constexpr uint64_t kStep = (uint64_t)TPW * U;
  // the wave steps split evenly over every wave of the grid (ranges differ by
  // at most one step), so each CU packs the same share whatever waves it
  // holds; ranges start on a step, so the 128-byte path's tile pairs stay
  // x-adjacent
  const uint64_t nsteps = (total_tiles + kStep - 1) / kStep;
  // with grab: a static share of sfix steps per wave, then steps dyn0 + the
  // counter's values; the next one is asked for while the wave packs its last
  // static step or its current dynamic one, so the atomic's round trip hides
  // behind that step's loads
  const uint64_t sfix = grab ? nsteps * 3 / 4 / nw : 0;
  uint64_t dyn0 = sfix * nw, dyn1 = nsteps;  // the wave's pool of steps
  // one pair: the static steps strided (gw, gw + nw, ...) and the pools'
  // dynamic steps interleaved, so all waves sweep the volume together in
  // tile order (the cube's launch 86.9 -> 85.4 us against contiguous ranges);
  // several pairs: contiguous ranges, which keep the pair walk short
  const bool strided = grab && npairs == 1;
  uint32_t np = 1, pool = 0;
  if (grab && pools > 1) {
    // pool = blockIdx % pools, which is the workgroup's XCD when the GPU runs
    // as one partition (workgroups go to the XCDs round robin); every pool
    // has workgroups as long as the grid has at least as many as pools, and
    // a grid with fewer takes one pool (a pool without waves would leave its
    // steps unpacked)
    np = min((uint32_t)pools, gridDim.x);
    pool = blockIdx.x % np;
    if (!strided) {
      const uint64_t nd = nsteps - dyn0, b = dyn0;
      dyn0 = b + nd * pool / np;
      dyn1 = b + nd * (pool + 1) / np;
    }
    grab += 32 * pool;
  }
  uint64_t si = 0;  // strided: the static steps begun
  const uint64_t w0 = strided ? (sfix ? min(total_tiles, gw * kStep) : 0)
                      : grab  ? min(total_tiles, gw * sfix * kStep)
                              : min(total_tiles, gw * nsteps / nw * kStep);
  uint64_t w1 = strided ? (sfix ? min(total_tiles, gw * kStep + kStep) : 0)
                : grab  ? min(total_tiles, (gw + 1) * sfix * kStep)
                        : min(total_tiles, (gw + 1) * nsteps / nw * kStep);
  uint32_t gv = 0;     // lane 0: the counter's value for the wave's next step
  bool asked = false;  // an atomic for it is in flight
  // the counter's address as a per-lane value (an opaque zero offset): the
  // compiler then keeps one plain returning atomic for lane 0 instead of
  // combining the wave's lanes, which reads the result back at once; this
  // way the wave waits for it only where the value is used
  uint32_t gz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(gz));
  auto ask = [&]() {
    if (lane == 0) gv = atomicAdd(grab + gz, 1u);
    asked = true;
  };
  const int slot = (int)(gw % kSumSlots);
  PackAcc acc;
  acc_clear(acc);
  int ap = -1;  // pair the accumulator belongs to (wave-uniform)
  // the pair of the wave's running position (wave-uniform): one search at the
  // start of its range, then a step past each pair boundary as the range is
  // walked in order; only a lane beyond the running pair's end searches
  int cp = 0;
  uint64_t ce = total_tiles;  // end of pair cp
  if (npairs > 1 && w0 < w1) {
    cp = find_pair(npairs, w0, [&](int i) { return idx[i].tile_base; });
    ce = cp + 1 < npairs ? idx[cp + 1].tile_base : total_tiles;
  }
  for (uint64_t base = w0;;) {
    if (strided && base >= w1 && si + 1 < sfix) {  // the next strided static step
      ++si;
      base = (gw + si * nw) * kStep;
      w1 = min(total_tiles, base + kStep);
    }
    // one call site for the atomic: at the last static step, at an empty
    // static share, and at every dynamic step (which is its range's last)
    if (grab && !asked && (!strided || si + 1 >= sfix) && base + kStep >= w1) ask();
    if (base >= w1) {
      if (!grab) break;
      const uint64_t gval = (uint32_t)__builtin_amdgcn_readfirstlane((int)gv);
      const uint64_t st = strided ? dyn0 + gval * np + pool : dyn0 + gval;
      asked = false;
      if (st >= dyn1) break;
      base = st * kStep;
      w1 = min(total_tiles, base + kStep);
      if (npairs > 1) {  // a jump past any number of pairs: find the pair again
        cp = find_pair(npairs, base, [&](int i) { return idx[i].tile_base; });
        ce = cp + 1 < npairs ? idx[cp + 1].tile_base : total_tiles;
      }
      continue;  // asks for the next one, then packs this one
    }
    bool inU[U], okU[U];
    int pU[U], txU[U], tyU[U], tzU[U], cntU[U];
    uint32_t tileU[U];
    uint64_t vbU[U];
    Row aU[U], bU[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t g = base + (uint64_t)u * TPW + sub;
      inU[u] = g < w1;
      pU[u] = 0;
      if (npairs > 1) {
        const uint64_t g0 = base + (uint64_t)u * TPW;  // the step's first tile
        while (g0 < w1 && g0 >= ce) {
          ++cp;
          ce = cp + 1 < npairs ? idx[cp + 1].tile_base : total_tiles;
        }
        pU[u] = cp;
        if (inU[u] && g >= ce) pU[u] = find_pair(npairs, g, [&](int i) { return idx[i].tile_base; });
      }
      aU[u] = bU[u] = 0;
      okU[u] = false;
      tileU[u] = 0;
      txU[u] = tyU[u] = tzU[u] = 0;
      cntU[u] = 0;
      vbU[u] = 0;
      if (inU[u]) {
        const PairDesc &d = descs[pU[u]];
        const uint32_t tile = (uint32_t)(g - (npairs > 1 ? idx[pU[u]].tile_base : 0));
        tileU[u] = tile;
        txU[u] = (int)(tile % (uint32_t)d.gx);
        tyU[u] = (int)((tile / (uint32_t)d.gx) % (uint32_t)d.gy);
        tzU[u] = (int)(tile / ((uint32_t)d.gx * d.gy));
        const int y = tyU[u] * TY + yl, z = tzU[u] * TZ + zl;
        okU[u] = y < d.ny && z < d.nz;
        cntU[u] = min(TXB, d.nx - txU[u] * TXB);
        vbU[u] = (uint64_t)txU[u] * TXB + (uint64_t)d.nx * ((uint64_t)y + (uint64_t)d.ny * z);
      }
    }
    // Two x-adjacent interior 64x8x8 tiles of a uint8 pair: their rows are
    // 128 contiguous bytes, so each load instruction reads 8 whole 128-byte
    // lines (lane L: row y+L/8 of z-slice k, bytes 16*(L%8) ..), and the 16-bit
    // compare pieces go through LDS to the lane that owns the row.
    bool fast = false;
    if constexpr (TXB == 64 && TY == 8 && TZ == 8) {
      const PairDesc &d = descs[pU[0]];
      fast = inU[0] && inU[1] && pU[0] == pU[1] && d.dtype == 0 && txU[1] == txU[0] + 1 &&
             tyU[1] == tyU[0] && tzU[1] == tzU[0] && (txU[1] + 1) * 64 <= d.nx &&
             (tyU[0] + 1) * 8 <= d.ny && (tzU[0] + 1) * 8 <= d.nz && (d.nx & 15) == 0 &&
             (((uintptr_t)d.gold | (uintptr_t)d.segm) & 15) == 0;
      fast = __builtin_amdgcn_readfirstlane((int)fast) != 0;  // one tile pair per wave
      if (fast) {
        const uint64_t nx = (uint64_t)d.nx, nxy = nx * (uint64_t)d.ny;
        const uint64_t b0 = (uint64_t)txU[0] * 64 + nx * ((uint64_t)tyU[0] * 8) +
                            nxy * ((uint64_t)tzU[0] * 8) + nx * (uint64_t)(lane >> 3) +
                            16u * (uint64_t)(lane & 7);
        const uint8_t *gp = (const uint8_t *)d.gold + b0, *sp = (const uint8_t *)d.segm + b0;
        u32x4 vg[8], vs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          vg[k] = __builtin_nontemporal_load((const u32x4 *)(gp + nxy * k));
          vs[k] = __builtin_nontemporal_load((const u32x4 *)(sp + nxy * k));
        }
        // all sixteen loads in flight before the first compare: one round
        // trip per step (the scheduler would otherwise start on the first
        // loads and issue the rest after their wait).  The volumes are read
        // once: nontemporal loads keep them from evicting the packed masks,
        // which the same launch reads back, from the caches (512^3 pair:
        // 94 -> 88 us per launch)
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t lab4 = (uint32_t)(uint8_t)d.label * 0x01010101u;
        uint16_t *pc = s_pieces[threadIdx.x >> 6];
#pragma unroll
        for (int k = 0; k < 8; ++k) {  // row 8k + L/8, piece L%8
          pc[(8 * k + (lane >> 3)) * 8 + (lane & 7)] = (uint16_t)eq_bits_u8x16(vg[k], lab4);
          pc[512 + (8 * k + (lane >> 3)) * 8 + (lane & 7)] = (uint16_t)eq_bits_u8x16(vs[k], lab4);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint64_t *row = (const uint64_t *)(pc + lane * 8);
        const uint64_t *rows = (const uint64_t *)(pc + 512 + lane * 8);
        aU[0] = (Row)row[0], aU[1] = (Row)row[1];
        bU[0] = (Row)rows[0], bU[1] = (Row)rows[1];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // the pieces are read before the next step writes
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (!fast) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (inU[u] && okU[u]) {
          const PairDesc &d = descs[pU[u]];
          aU[u] = load_row_bits<Row>(d, d.gold, vbU[u], cntU[u]);
          bU[u] = load_row_bits<Row>(d, d.segm, vbU[u], cntU[u]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = pU[u];
      const int lead = __builtin_amdgcn_readfirstlane(p);
      if (ap != lead) {
        if (ap >= 0) {
          if (alt)
            acc_flush(*alt, acc, lane, slot);
          else
            acc_flush(ctrs[ap], acc, lane, slot);
        }
        ap = lead;
      }
      const bool in = inU[u], rowok = okU[u];
      const PairDesc &d = descs[p];
      const uint32_t tile = tileU[u];
      const Row a = aU[u], b = bU[u];
      const uint64_t o = (uint64_t)tile * R + r;
      if constexpr (SKIP) {
        // a mask with no voxel in the tile writes no rows there: the readers
        // of the near launch's pack (near_window, the region launch) take
        // zero rows wherever the tile's flag says the mask is empty
        // (512^3 pack measured alone, a fifth of the tiles non-empty: 58 ->
        // 53 us)
        if (rowok && __ballot(a != 0) != 0) put<COH>((Row *)d.orig[0] + o, a);
        if (rowok && __ballot(b != 0) != 0) put<COH>((Row *)d.orig[1] + o, b);
      } else if (rowok) {
        put<COH>((Row *)d.orig[0] + o, a);
        put<COH>((Row *)d.orig[1] + o, b);
      }
      const Row vb = valid_bits<Row>(txU[u], d.nx);
      uint32_t flags = 0;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const Row w = m == 0 ? a : b;
        const bool any = group_bits<R>(__ballot(w != 0), lane) != 0;
        const bool full = group_bits<R>(__ballot(rowok && w != vb), lane) == 0;
        constexpr bool kDeep = TXB == 64 && TY == 8 && TZ == 8;
        bool sxl = false, sxh = false;  // voxels in the eight columns toward -x / +x
        if constexpr (kDeep) {
          sxl = group_bits<R>(__ballot(rowok && (w & 0xFFu) != 0), lane) != 0;
          sxh = group_bits<R>(__ballot(rowok && (w >> (TXB - 8)) != 0), lane) != 0;
        }
        flags |= ((uint32_t)any | (uint32_t)full << 1 | (uint32_t)sxl << 14 | (uint32_t)sxh << 15)
                 << (16 * m);
        if constexpr (LEAN) continue;
        const uint64_t fxl = group_bits<R>(__ballot(rowok && (w & 1)), lane);
        const uint64_t fxh = group_bits<R>(__ballot(rowok && ((w >> (TXB - 1)) & 1)), lane);
        const bool fyl = group_bits<R>(__ballot(yl == 0 && w != 0), lane) != 0;
        const bool fyh = group_bits<R>(__ballot(yl == TY - 1 && w != 0), lane) != 0;
        const bool fzl = group_bits<R>(__ballot(zl == 0 && w != 0), lane) != 0;
        const bool fzh = group_bits<R>(__ballot(zl == TZ - 1 && w != 0), lane) != 0;
        // the two outermost layers toward each face (reach within two steps;
        // read only by the two-steps-per-barrier seed and steps, which run on
        // 64x8x8): the second column from each x face as face words, and flags
        bool gxl = false, gxh = false, gyl = false, gyh = false, gzl = false, gzh = false;
        uint64_t fx2l = 0, fx2h = 0;
        if constexpr (kDeep) {
          fx2l = group_bits<R>(__ballot(rowok && ((w >> 1) & 1)), lane);
          fx2h = group_bits<R>(__ballot(rowok && ((w >> (TXB - 2)) & 1)), lane);
          gxl = (fxl | fx2l) != 0;
          gxh = (fxh | fx2h) != 0;
          gyl = group_bits<R>(__ballot(yl <= 1 && w != 0), lane) != 0;
          gyh = group_bits<R>(__ballot(yl >= TY - 2 && w != 0), lane) != 0;
          gzl = group_bits<R>(__ballot(zl <= 1 && w != 0), lane) != 0;
          gzh = group_bits<R>(__ballot(zl >= TZ - 2 && w != 0), lane) != 0;
        }
        if (in && r == 0) {
          put<COH>(d.xf_orig[m] + kXf * (uint64_t)tile, fxl);
          put<COH>(d.xf_orig[m] + kXf * (uint64_t)tile + 1, fxh);
          if constexpr (kDeep) {
            put<COH>(d.xf_orig[m] + kXf * (uint64_t)tile + 2, fx2l);
            put<COH>(d.xf_orig[m] + kXf * (uint64_t)tile + 3, fx2h);
          }
        }
        flags |= ((uint32_t)(fxl != 0) << 2 | (uint32_t)(fxh != 0) << 3 | (uint32_t)fyl << 4 |
                  (uint32_t)fyh << 5 | (uint32_t)fzl << 6 | (uint32_t)fzh << 7 |
                  (uint32_t)gxl << 8 | (uint32_t)gxh << 9 | (uint32_t)gyl << 10 |
                  (uint32_t)gyh << 11 | (uint32_t)gzl << 12 | (uint32_t)gzh << 13)
                 << (16 * m);
      }
      if (in && r == 0) put<COH>(d.tflags + tile, flags);
      if (!COH && d.map && rowok) {
        uint16_t *mp = d.map + vbU[u];
        for (int k = 0; k < cntU[u]; ++k) mp[k] = 0;
      }
      // counts of hello.cu:624-625 plus the non-empty tile box
      const uint32_t c0 = popc_row(a), c1 = popc_row(b);
      const uint32_t c2 = popc_row((Row)(b & ~a)), c3 = popc_row((Row)(a & ~b));
      if constexpr (MARK) {  // one tile per wave: the ballots cover the tile
        const bool fp = __ballot(c2 != 0) != 0, fn = __ballot(c3 != 0) != 0;
        if (in && r == 0 && (fp || fn) && d.marks) mark_tile(d, txU[u], tyU[u], tzU[u], fp, fn);
      }
      const bool ne = (a | b) != 0;
      if (p == ap) {
        acc.c[0] += c0, acc.c[1] += c1, acc.c[2] += c2, acc.c[3] += c3;
        if (ne) {
          acc.mn[0] = min(acc.mn[0], txU[u]), acc.mx[0] = max(acc.mx[0], txU[u]);
          acc.mn[1] = min(acc.mn[1], tyU[u]), acc.mx[1] = max(acc.mx[1], tyU[u]);
          acc.mn[2] = min(acc.mn[2], tzU[u]), acc.mx[2] = max(acc.mx[2], tzU[u]);
        }
        if constexpr (COH) {  // per-mask boxes (the fused launch's PackSums)
          const int tc[3] = {txU[u], tyU[u], tzU[u]};
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            if (a != 0) acc.mn[3 + k] = min(acc.mn[3 + k], tc[k]), acc.mx[3 + k] = max(acc.mx[3 + k], tc[k]);
            if (b != 0) acc.mn[6 + k] = min(acc.mn[6 + k], tc[k]), acc.mx[6 + k] = max(acc.mx[6 + k], tc[k]);
          }
        }
      } else if (in) {  // a lane past a pair boundary publishes its own row
        PairCtr &cp = ctrs[p];
        if (c0) atomicAdd(&cp.sums[0][slot], (unsigned long long)c0);
        if (c1) atomicAdd(&cp.sums[1][slot], (unsigned long long)c1);
        if (c2) atomicAdd(&cp.sums[2][slot], (unsigned long long)c2);
        if (c3) atomicAdd(&cp.sums[3][slot], (unsigned long long)c3);
        if (ne) {
          atomicMax(&cp.box[0][slot], box_enc_lo(txU[u]));
          atomicMax(&cp.box[3][slot], box_enc_hi(txU[u]));
          atomicMax(&cp.box[1][slot], box_enc_lo(tyU[u]));
          atomicMax(&cp.box[4][slot], box_enc_hi(tyU[u]));
          atomicMax(&cp.box[2][slot], box_enc_lo(tzU[u]));
          atomicMax(&cp.box[5][slot], box_enc_hi(tzU[u]));
        }
      }
    }
    base += kStep;
  }
  if (ap >= 0) {
    if (alt)
      acc_flush(*alt, acc, lane, slot);
    else
      acc_flush(ctrs[ap], acc, lane, slot);
  }
}

template <int TXB, int TY, int TZ, bool SKIP = false>
__global__ __launch_bounds__(kWgThreads) void pack_kernel(const PairDesc *__restrict__ descs,
                                                          const PackIndex *__restrict__ idx,
                                                          PairCtr *ctrs, int npairs,
                                                          uint64_t total_tiles, uint32_t *grab) {
  // per wave: the 16-bit compare pieces of 64 rows x 8 chunks, gold then segm
  __shared__ __align__(16) uint16_t s_pieces[kWavesPerWg][1024];
  constexpr bool kMark = TXB == 64 && TY == 8 && TZ == 8;  // pairs with a bitmap get marks
  pack_tiles<TXB, TY, TZ, false, kMark, false, SKIP>(descs, idx, ctrs, npairs, total_tiles,
                                                     s_pieces, nullptr, grab, 8);
}

// Zero rows for the masks a skipping batch pack (pack_kernel<64, 8, 8, true>)
// left unwritten -- a mask empty in a tile, flag bit 0 / 16 clear -- of the
// pairs the batched near pass did not finish: their seed and dilation read
// every row, while the near pass takes such rows as zero from the flags.  One
// wave per tile, lane r row r; a finished pair (near_done) is skipped.
__global__ __launch_bounds__(kWgThreads) void fill_kernel(const PairDesc *__restrict__ descs,
                                                          const PairCtr *__restrict__ ctrs,
                                                          int npairs) {
  constexpr int R = 64;
  const int lane = threadIdx.x & 63;
  const uint32_t gw = blockIdx.x * kWavesPerWg +
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t nw = gridDim.x * kWavesPerWg;
  // lane i checks pair p0 + i (the counters lie kilobytes apart: one load
  // per pair in turn would cost a round trip each); the wave walks the pairs
  // left open
  for (int p0 = 0; p0 < npairs; p0 += kWave) {
    const int pl = p0 + lane;
    uint64_t left = __ballot(pl < npairs && ctrs[pl].near_done == 0);
    while (left) {
      const int p = p0 + __builtin_ctzll(left);
      left &= left - 1;
      const PairDesc &d = descs[p];
      uint64_t *o0 = (uint64_t *)d.orig[0], *o1 = (uint64_t *)d.orig[1];
      for (uint32_t tile = gw; tile < d.ntiles; tile += nw) {
        const uint32_t f = (uint32_t)__builtin_amdgcn_readfirstlane((int)d.tflags[tile]);
        const uint64_t o = (uint64_t)tile * R + lane;
        if (!(f & 1u)) o0[o] = 0ull;
        if (!(f & 0x10000u)) o1[o] = 0ull;
      }
    }
  }
}

// The pack pass of a pair whose two masks were packed before, each on its own
// (the server's resident masks): d.orig and d.xf_orig point at them, and
// gflags / sflags are their tile flags (the low 16 bits of each).  One wave
// per tile step as in the pack: a lane reads its row of both masks (2 bits
// per voxel instead of the pack's 2 bytes), counts |gold|, |segm|, |FP|, |FN|
// and the non-empty tile box into the pair's partials, and writes the pair's
// tile flags (gold low, segm high).  Tiles empty in both masks read nothing.
template <int TXB, int TY, int TZ>
__global__ __launch_bounds__(kWgThreads) void combine_kernel(const PairDesc *__restrict__ dp,
                                                             PairCtr *ctr,
                                                             const uint32_t *__restrict__ gflags,
                                                             const uint32_t *__restrict__ sflags,
                                                             uint64_t total_tiles) {
  using Row = RowT<TXB>;
  constexpr int R = TY * TZ;
  constexpr int TPW = kWave / R;
  const PairDesc &d = *dp;
  const int lane = threadIdx.x & 63;
  const int sub = lane / R, r = lane % R;
  const int yl = r % TY, zl = r / TY;
  const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerWg + (threadIdx.x >> 6);
  const uint64_t nw = (uint64_t)gridDim.x * kWavesPerWg;
  const int slot = (int)(gw % kSumSlots);
  PackAcc acc;
  acc_clear(acc);
  for (uint64_t base = gw * TPW; base < total_tiles; base += nw * TPW) {
    const uint64_t g = base + sub;
    const bool in = g < total_tiles;
    const uint32_t tile = in ? (uint32_t)g : 0u;
    const uint32_t gf = in ? gflags[tile] & 0xFFFFu : 0u, sf = in ? sflags[tile] & 0xFFFFu : 0u;
    if (in && r == 0) d.tflags[tile] = gf | (sf << 16);
    const int tx = (int)(tile % (uint32_t)d.gx);
    const int ty = (int)((tile / (uint32_t)d.gx) % (uint32_t)d.gy);
    const int tz = (int)(tile / ((uint32_t)d.gx * d.gy));
    // rows beyond the volume were never written by the masks' pack
    const bool rowok = in && ty * TY + yl < d.ny && tz * TZ + zl < d.nz;
    const uint64_t o = (uint64_t)tile * R + r;
    const Row a = rowok && (gf & 1u) ? ((const Row *)d.orig[0])[o] : (Row)0;
    const Row b = rowok && (sf & 1u) ? ((const Row *)d.orig[1])[o] : (Row)0;
    acc.c[0] += popc_row(a), acc.c[1] += popc_row(b);
    acc.c[2] += popc_row((Row)(b & ~a)), acc.c[3] += popc_row((Row)(a & ~b));
    if constexpr (TXB == 64 && TY == 8 && TZ == 8) {  // one tile per wave: near-pass marks
      const bool fp = __ballot((b & ~a) != 0) != 0, fn = __ballot((a & ~b) != 0) != 0;
      if (in && r == 0 && (fp || fn) && d.marks) mark_tile(d, tx, ty, tz, fp, fn);
    }
    if ((a | b) != 0) {
      acc.mn[0] = min(acc.mn[0], tx), acc.mx[0] = max(acc.mx[0], tx);
      acc.mn[1] = min(acc.mn[1], ty), acc.mx[1] = max(acc.mx[1], ty);
      acc.mn[2] = min(acc.mn[2], tz), acc.mx[2] = max(acc.mx[2], tz);
    }
  }
  acc_flush(*ctr, acc, lane, slot);
}

// The combine pass of several pairs of registered masks (a server admission
// group; 64x8x8 tiles): pair p's tiles are [idx[p].tile_base, + its tiles),
// its masks' tile flags gf[p] and sf[p]; each wave walks a contiguous range
// of tiles and publishes its counts and box once per pair it crossed.
__global__ __launch_bounds__(kWgThreads) void combine_group_kernel(
    const PairDesc *__restrict__ descs, const PackIndex *__restrict__ idx, PairCtr *ctrs,
    int npairs, uint64_t total_tiles, const uint32_t *const *__restrict__ gfs,
    const uint32_t *const *__restrict__ sfs) {
  using Row = uint64_t;
  constexpr int R = 64;
  const int lane = threadIdx.x & 63;
  const int yl = lane % 8, zl = lane / 8;
  const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerWg + (threadIdx.x >> 6);
  const uint64_t nw = (uint64_t)gridDim.x * kWavesPerWg;
  const uint64_t t0 = min(total_tiles, gw * total_tiles / nw);
  const uint64_t t1 = min(total_tiles, (gw + 1) * total_tiles / nw);
  const int slot = (int)(gw % kSumSlots);
  PackAcc acc;
  acc_clear(acc);
  int ap = -1;
  uint64_t ce = 0;  // end of pair ap's tiles
  for (uint64_t g = t0; g < t1; ++g) {
    if (ap < 0 || g >= ce) {  // the next pair (wave-uniform)
      if (ap >= 0) acc_flush(ctrs[ap], acc, lane, slot);
      ap = find_pair(npairs, g, [&](int i) { return idx[i].tile_base; });
      ce = ap + 1 < npairs ? idx[ap + 1].tile_base : total_tiles;
    }
    const PairDesc &d = descs[ap];
    const uint32_t tile = (uint32_t)(g - idx[ap].tile_base);
    const uint32_t gf = gfs[ap][tile] & 0xFFFFu, sf = sfs[ap][tile] & 0xFFFFu;
    if (lane == 0) d.tflags[tile] = gf | (sf << 16);
    const int tx = (int)(tile % (uint32_t)d.gx);
    const int ty = (int)((tile / (uint32_t)d.gx) % (uint32_t)d.gy);
    const int tz = (int)(tile / ((uint32_t)d.gx * d.gy));
    // rows beyond the volume were never written by the masks' pack
    const bool rowok = ty * 8 + yl < d.ny && tz * 8 + zl < d.nz;
    const uint64_t o = (uint64_t)tile * R + lane;
    const Row a = rowok && (gf & 1u) ? ((const Row *)d.orig[0])[o] : (Row)0;
    const Row b = rowok && (sf & 1u) ? ((const Row *)d.orig[1])[o] : (Row)0;
    acc.c[0] += popc_row(a), acc.c[1] += popc_row(b);
    acc.c[2] += popc_row((Row)(b & ~a)), acc.c[3] += popc_row((Row)(a & ~b));
    const bool fp = __ballot((b & ~a) != 0) != 0, fn = __ballot((a & ~b) != 0) != 0;
    if (lane == 0 && (fp || fn) && d.marks) mark_tile(d, tx, ty, tz, fp, fn);
    if ((a | b) != 0) {
      acc.mn[0] = min(acc.mn[0], tx), acc.mx[0] = max(acc.mx[0], tx);
      acc.mn[1] = min(acc.mn[1], ty), acc.mx[1] = max(acc.mx[1], ty);
      acc.mn[2] = min(acc.mn[2], tz), acc.mx[2] = max(acc.mx[2], tz);
    }
  }
  if (ap >= 0) acc_flush(ctrs[ap], acc, lane, slot);
}

// The engine iteration at which the pairs of some server slots start (a
// staged pair joining, or kUnused for pairs that retire), written between two
// slices on the engine's stream.
struct SlotList {
  int n;
  uint16_t slot[kSlotListMax];
};
__global__ void set_start_kernel(PairDesc *descs, SlotList l, uint32_t t_start) {
  if ((int)threadIdx.x < l.n) descs[l.slot[threadIdx.x]].t_start = t_start;
}

// |gold|, |segm|, |FP|, |FN| per pair from the spread partials (one wave per
// pair); the single-pair seed pass folds them itself
__global__ void totals_kernel(PairCtr *ctrs, int p0, int npairs) {
  const int p = p0 + blockIdx.x;
  if (p >= p0 + npairs) return;
  const int lane = threadIdx.x;
  PairCtr &c = ctrs[p];
  unsigned long long v[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    v[f] = c.sums[f][lane];
    v[f] = wave_sum64(v[f]);
  }
  if (lane == 0) c.n_gold = v[0], c.n_segm = v[1], c.n_fp = v[2], c.n_fn = v[3];
  int bx[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    bx[k] = wave_max(c.box[k][lane], lane);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) c.lo[k] = box_dec_lo(bx[k]), c.hi[k] = box_dec_hi(bx[3 + k]);
  }
}

// ---------------------------------------------------------------- helpers
struct Box {
  int lo[3], hi[3];
};
__device__ __forceinline__ Box pair_box(const PairDesc &d, const int *lo, const int *hi) {
  Box b;
  if (d.crop) {
    b.lo[0] = lo[0], b.lo[1] = lo[1], b.lo[2] = lo[2];
    b.hi[0] = hi[0], b.hi[1] = hi[1], b.hi[2] = hi[2];
    if (d.conn == 18 && b.hi[0] >= 0) {
      // a shortest path of the 18 element can step one voxel outside the
      // voxels' box in x or y (offset (2,0,2) is the corners (1,1,1) and
      // (1,-1,1)): one more tile on each side there
      b.lo[0] = max(b.lo[0] - 1, 0), b.hi[0] = min(b.hi[0] + 1, d.gx - 1);
      b.lo[1] = max(b.lo[1] - 1, 0), b.hi[1] = min(b.hi[1] + 1, d.gy - 1);
    }
  } else {
    b.lo[0] = b.lo[1] = b.lo[2] = 0;
    b.hi[0] = d.gx - 1, b.hi[1] = d.gy - 1, b.hi[2] = d.gz - 1;
  }
  return b;
}
__device__ __forceinline__ Box pair_box(const PairDesc &d, const PairCtr &c) {
  return pair_box(d, c.lo, c.hi);
}

// What a dilation step reads of its pair, with the box applied: for one pair
// the launch keeps it in registers, for several a step loads it in one batch.
struct PairView {
  const void *orig[2];
  void *buf[2][2];
  const uint64_t *xf_orig[2];
  uint64_t *xf_buf[2][2];
  uint32_t *state[2];
  uint32_t *rstate[2];
  uint32_t *hist0, *hist1;
  uint16_t *map;
  int nx, ny, nz, gx, gy, hist_len;
  uint32_t t_start, gen;
  Box b;
  unsigned long long cut[2];  // robust cut per pass; 0: nothing to do
};
// from the descriptor and explicit box and FP / FN counts
__device__ __forceinline__ PairView make_view(const PairDesc &d, const int *lo, const int *hi,
                                              unsigned long long n_fp, unsigned long long n_fn) {
  PairView v;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    v.orig[m] = d.orig[m];
    v.buf[m][0] = d.buf[m][0], v.buf[m][1] = d.buf[m][1];
    v.xf_orig[m] = d.xf_orig[m];
    v.xf_buf[m][0] = d.xf_buf[m][0], v.xf_buf[m][1] = d.xf_buf[m][1];
    v.state[m] = d.state[m];
    v.rstate[m] = d.rstate[m];
  }
  v.hist0 = d.hist[0], v.hist1 = d.hist[1];
  v.map = d.map;
  v.nx = d.nx, v.ny = d.ny, v.nz = d.nz, v.gx = d.gx, v.gy = d.gy, v.hist_len = d.hist_len;
  v.t_start = d.t_start, v.gen = d.gen;
  v.b = pair_box(d, lo, hi);
  v.cut[0] = robust_cut(d.robust, n_fp);
  v.cut[1] = robust_cut(d.robust, n_fn);
  return v;
}
__device__ __forceinline__ PairView make_view(const PairDesc &d, const PairCtr &c) {
  return make_view(d, c.lo, c.hi, c.n_fp, c.n_fn);
}

// neighbour tile of `tile` in direction dir (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z),
// or kNoTile when it is outside the (cropped) grid; G has the grid dims gx, gy
template <typename G>
__device__ __forceinline__ uint32_t neighbour(const G &d, const Box &b, uint32_t tile, int tx,
                                              int ty, int tz, int dir) {
  switch (dir) {
    case 0: return tx > b.lo[0] ? tile - 1 : kNoTile;
    case 1: return tx < b.hi[0] ? tile + 1 : kNoTile;
    case 2: return ty > b.lo[1] ? tile - (uint32_t)d.gx : kNoTile;
    case 3: return ty < b.hi[1] ? tile + (uint32_t)d.gx : kNoTile;
    case 4: return tz > b.lo[2] ? tile - (uint32_t)d.gx * d.gy : kNoTile;
    default: return tz < b.hi[2] ? tile + (uint32_t)d.gx * d.gy : kNoTile;
  }
}

// Workgroup-local staging of queue pushes: LDS, flushed once per iteration
// into sub-queue blockIdx % kSubQ; straight to that sub-queue when LDS is full.
struct LocalQueue {
  uint64_t e[kLocalQueue];
  uint32_t n, base;
};

// pushed (LDS, may be null) counts the entries that went straight to gq
__device__ __forceinline__ void push_entries(bool want, uint64_t ent, LocalQueue &lq,
                                             uint64_t *gq, uint32_t *gq_len, int lane,
                                             uint32_t *pushed = nullptr) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const uint32_t n = (uint32_t)__popcll(m);
  uint32_t base = 0;
  bool global = false;
  if (lane == 0) {  // reserve only if the whole run fits, so LDS stays hole-free
    uint32_t cur = lq.n;
    for (;;) {
      if (cur + n > (uint32_t)kLocalQueue) {
        base = atomicAdd(gq_len, n);
        if (pushed) atomicAdd(pushed, n);
        global = true;
        break;
      }
      const uint32_t prev = atomicCAS(&lq.n, cur, cur + n);
      if (prev == cur) {
        base = cur;
        break;
      }
      cur = prev;
    }
  }
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);  // lane 0's, as a scalar
  global = __builtin_amdgcn_readlane((int)global, 0) != 0;
  if (want) {
    const uint32_t pos = base + lanes_below(m);
    if (global)
      st_sc(&gq[pos], ent);
    else
      lq.e[pos] = ent;
  }
}

// Pushes of a single-pair launch: kept in the workgroup's own LDS queue st
// (count *stn, capacity cap) for its next iteration, so no global queue
// round trip; when it is full they go to the staging queue lq, which the
// workgroup flushes into the global queue once at the end of the iteration.
__device__ __forceinline__ void push_sticky(bool want, uint64_t ent, uint64_t *st, uint32_t *stn,
                                            uint32_t cap, LocalQueue &lq, uint64_t *gq,
                                            uint32_t *gq_len, int lane, uint32_t *pushed) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const uint32_t n = (uint32_t)__popcll(m);
  uint32_t base = 0;
  bool kept = false;
  if (lane == 0) {  // reserve only if the whole run fits
    uint32_t cur = *stn;
    while (cur + n <= cap) {
      const uint32_t prev = atomicCAS(stn, cur, cur + n);
      if (prev == cur) {
        base = cur;
        kept = true;
        break;
      }
      cur = prev;
    }
  }
  kept = __builtin_amdgcn_readlane((int)kept, 0) != 0;  // lane 0's, as a scalar
  if (!kept) {  // staged; straight to gq (counted in *pushed) when lq is full too
    push_entries(want, ent, lq, gq, gq_len, lane, pushed);
    return;
  }
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
  if (want) st[base + lanes_below(m)] = ent;
}

// returns the number of entries flushed (the same in every thread)
__device__ __forceinline__ uint32_t flush_local_queue(LocalQueue &lq, uint64_t *gq,
                                                      uint32_t *gq_len) {
  __syncthreads();
  const uint32_t n = lq.n;
  if (n != 0) {
    if (threadIdx.x == 0) lq.base = atomicAdd(gq_len, n);
    __syncthreads();
    const uint32_t b = lq.base;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) st_sc(&gq[b + i], lq.e[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) lq.n = 0;
  return n;
}

// ---------------------------------------------------------------- seed
// Bit m of the result: an empty tile is reached in one step of the element
// from a non-empty neighbour of mask m, judged from the neighbours' tile
// flags (flags(j, nb) gives those of neighbour nb at offset index
// j = (ox+1) + 3(oy+1) + 9(oz+1)).  CONN 6: a face neighbour has voxels on its
// face toward the tile (hello.cu:1649-1663).  CONN 18 / 26: a neighbour at
// offset o among the 26 has voxels on each of its faces that o crosses (the
// per-face flags make this a superset of the tiles reached; a tile queued
// early only runs steps that leave it as it is).
// DS (two dilation steps per barrier, CONN 6): reached within two steps -- a
// face neighbour with voxels in its two layers toward the tile, or an edge
// neighbour with voxels on both faces toward it (a superset of its edge line).
template <int CONN, bool DS = false, typename G, typename F>
__device__ __forceinline__ uint32_t seed_reach(const G &d, const Box &b, uint32_t tile, int tx,
                                               int ty, int tz, F flags) {
  uint32_t act = 0;
  if constexpr (DS) {
    static_assert(CONN == 6, "two steps per barrier with the 6 element");
#pragma unroll
    for (int j = 0; j < 27; ++j) {
      const int ox = j % 3 - 1, oy = (j / 3) % 3 - 1, oz = j / 9 - 1;
      const int nzc = (ox != 0) + (oy != 0) + (oz != 0);
      if (nzc == 0 || nzc == 3) continue;
      const int ux = tx + ox, uy = ty + oy, uz = tz + oz;
      if (ux < b.lo[0] || ux > b.hi[0] || uy < b.lo[1] || uy > b.hi[1] || uz < b.lo[2] ||
          uz > b.hi[2])
        continue;
      const uint32_t f = flags(j, (uint32_t)((int)tile + ox + oy * d.gx + oz * d.gx * d.gy));
      // the neighbour's faces toward this tile: one layer (bits 2..7) or two (8..13)
      const int base = nzc == 1 ? 8 : 2;
      const uint32_t need = (ox ? 1u << (base + (ox > 0 ? 0 : 1)) : 0u) |
                            (oy ? 1u << (base + (oy > 0 ? 2 : 3)) : 0u) |
                            (oz ? 1u << (base + (oz > 0 ? 4 : 5)) : 0u);
#pragma unroll
      for (int m = 0; m < 2; ++m) act |= (uint32_t)((((f >> (16 * m)) & need) == need)) << m;
    }
  } else if constexpr (CONN == 6) {
    // bit of the neighbour's flags that marks its face toward this tile, per
    // direction (0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z)
    constexpr int kFacing[6] = {3, 2, 5, 4, 7, 6};
    constexpr int kJ[6] = {12, 14, 10, 16, 4, 22};  // offset index of each direction
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      const uint32_t nb = neighbour(d, b, tile, tx, ty, tz, dir);
      if (nb == kNoTile) continue;
      const uint32_t f = flags(kJ[dir], nb);
#pragma unroll
      for (int m = 0; m < 2; ++m) act |= ((f >> (16 * m + kFacing[dir])) & 1u) << m;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 27; ++j) {
      if (j == 13) continue;
      const int ox = j % 3 - 1, oy = (j / 3) % 3 - 1, oz = j / 9 - 1;
      const int ux = tx + ox, uy = ty + oy, uz = tz + oz;
      if (ux < b.lo[0] || ux > b.hi[0] || uy < b.lo[1] || uy > b.hi[1] || uz < b.lo[2] ||
          uz > b.hi[2])
        continue;
      const uint32_t nb = (uint32_t)((int)tile + ox + oy * d.gx + oz * d.gx * d.gy);
      const uint32_t f = flags(j, nb);
      // the neighbour's faces toward this tile: -x (bit 2) when it lies at +x
      const uint32_t need = (ox ? 1u << (ox > 0 ? 2 : 3) : 0u) | (oy ? 1u << (oy > 0 ? 4 : 5) : 0u) |
                            (oz ? 1u << (oz > 0 ? 6 : 7) : 0u);
#pragma unroll
      for (int m = 0; m < 2; ++m) act |= (uint32_t)((((f >> (16 * m)) & need) == need)) << m;
    }
  }
  return act;
}

// One thread per tile decides, for both masks, the tile's state and whether
// it enters the first queue, from its own flags and those of its six
// neighbours (no atomics): a non-empty full tile is kFull and stays out; a
// non-empty tile that is not full is queued; an empty tile is queued when a
// non-empty neighbour has voxels on the face toward it (hello.cu:1649-1663).
// Only tiles inside the pair's box of non-empty tiles take part
// (hello.cu:940-973), and a pass whose cut is zero queues nothing.  Every
// entry is first = 1 (input from orig) and check = 1 (count first hits).
// With one pair the launch also folds the pack partials into the totals.
// DS: the dilation runs two steps per engine iteration, so an empty tile is
// queued when a neighbour's voxels reach it within two steps.
template <int TXB, int TY, int TZ, int CONN, bool DS = false>
__global__ __launch_bounds__(kWgThreads) void seed_kernel(const PairDesc *__restrict__ descs,
                                                          const PackIndex *__restrict__ idx,
                                                          PairCtr *ctrs, int npairs, Engine *es,
                                                          uint64_t *q, uint32_t qcap,
                                                          uint64_t total_tiles, int p0) {
  __shared__ LocalQueue lq;
  __shared__ unsigned long long s_tot[4];
  __shared__ int s_box[6];
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) lq.n = 0;
  if (npairs == 1 && threadIdx.x < 64) {
    PairCtr &c = ctrs[0];
    unsigned long long v[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      v[f] = c.sums[f][lane];
      v[f] = wave_sum64(v[f]);
    }
    int bx[6];
#pragma unroll
    for (int k = 0; k < 6; ++k)
      bx[k] = k < 3 ? box_dec_lo(wave_max(c.box[k][lane], lane))
                    : box_dec_hi(wave_max(c.box[k][lane], lane));
    if (lane == 0) {
#pragma unroll
      for (int f = 0; f < 4; ++f) s_tot[f] = v[f];
#pragma unroll
      for (int k = 0; k < 6; ++k) s_box[k] = bx[k];
      if (blockIdx.x == 0) {
        c.n_gold = v[0], c.n_segm = v[1], c.n_fp = v[2], c.n_fn = v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) c.lo[k] = bx[k], c.hi[k] = bx[3 + k];
      }
    }
  }
  __syncthreads();
  const uint32_t t0 = es->t;
  const uint32_t kq = blockIdx.x % kSubQ;
  uint64_t *gq = sub_queue(q, qcap, t0 & 1, kq);
  uint32_t *gq_len = &es->qlen[t0 % 4][kq];
  for (uint64_t tb = (uint64_t)blockIdx.x * kWgThreads; tb < total_tiles;
       tb += (uint64_t)gridDim.x * kWgThreads) {
    const uint64_t g = tb + threadIdx.x;
    const bool in = g < total_tiles;
    int p = 0;
    if (npairs > 1)
      p = find_pair(npairs, in ? g : total_tiles - 1, [&](int i) { return idx[i].tile_base; });
    const PairDesc &d = descs[p];
    const PairCtr &c = ctrs[p];
    const uint32_t tile = in ? (uint32_t)(g - idx[p].tile_base) : 0;
    const int tx = (int)(tile % (uint32_t)d.gx);
    const int ty = (int)((tile / (uint32_t)d.gx) % (uint32_t)d.gy);
    const int tz = (int)(tile / ((uint32_t)d.gx * d.gy));
    const Box b = npairs == 1 ? pair_box(d, s_box, s_box + 3) : pair_box(d, c);
    const bool in_box = in && tx >= b.lo[0] && tx <= b.hi[0] && ty >= b.lo[1] &&
                        ty <= b.hi[1] && tz >= b.lo[2] && tz <= b.hi[2];
    uint32_t own = 0, reach = 0;
    if (in_box) {
      own = d.tflags[tile];
      reach = seed_reach<CONN, DS>(d, b, tile, tx, ty, tz, [&](int, uint32_t nb) -> uint32_t {
        return d.tflags[nb];
      });
    }
    const unsigned long long nfp = npairs == 1 ? s_tot[2] : c.n_fp;
    const unsigned long long nfn = npairs == 1 ? s_tot[3] : c.n_fn;
    bool want[2];
    uint64_t ent[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const uint32_t f = (own >> (16 * m)) & 0xFFFF;
      const bool any = f & 1, full = (f >> 1) & 1;
      const bool act = ((reach >> m) & 1u) != 0;
      // a pass whose cut is zero, or whose pair the batched near pass
      // finished, queues nothing
      const bool needed = robust_cut(d.robust, m == 0 ? nfp : nfn) != 0 && !c.near_done;
      const uint32_t st = !in_box ? 0u : (any && full) ? kFull : (any || (act && needed)) ? 1u : 0u;
      if (in) d.state[m][tile] = st;
      want[m] = in_box && needed && (any ? !full : act);
      ent[m] = make_entry(tile, (uint32_t)(p0 + p), (uint32_t)m, 1, 1, d.gen);
    }
    push_entries(want[0], ent[0], lq, gq, gq_len, lane);
    push_entries(want[1], ent[1], lq, gq, gq_len, lane);
  }
  flush_local_queue(lq, gq, gq_len);
}

// ---------------------------------------------------------------- barrier
// Grid barrier on an occupancy-bounded grid (every workgroup resident).
// Arrival: each wave drains its own memory operations (s_waitcnt 0: its
// stores, all agent-coherent, are then performed), then the workgroup stores
// its epoch into its own flag.  Workgroup 0 waits for every flag, folds the
// iteration's hit counters (fold), then publishes the epoch; the others wait
// for it.  Waits are bounded by wall-clock ticks: on expiry the engine error
// flag is set and the caller leaves its loop, so no wave waits forever.
// With fence set, the arrival and release also write back / invalidate L2.
// The master also publishes, with the epoch, the prefix offsets of the
// sub-queues of the next iteration (next_qlen), so one read of its release
// replica gives every workgroup both; s_pre receives them.
template <typename Fold>
__device__ __forceinline__ bool grid_barrier(Engine *es, uint32_t epoch, uint64_t limit,
                                             bool fence, int *s_ok, Fold fold, uint64_t *tr,
                                             const uint32_t *next_qlen, uint32_t *s_pre) {
  if (fence) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    *s_ok = 1;
    st_sc(&es->arrive[blockIdx.x], epoch);
  }
  if (blockIdx.x == 0) {
    __syncthreads();
    int ok = 1;
    const uint64_t t0 = wall_clock64();
    constexpr int kPoll = 4;  // flags polled together per thread (one round trip each poll)
    for (uint32_t i0 = threadIdx.x; i0 < gridDim.x && ok; i0 += kPoll * blockDim.x) {
      uint32_t pending = 0;
#pragma unroll
      for (int k = 0; k < kPoll; ++k)
        if (i0 + k * blockDim.x < gridDim.x) pending |= 1u << k;
      while (pending) {
        uint32_t f[kPoll];
#pragma unroll
        for (int k = 0; k < kPoll; ++k)
          f[k] = ((pending >> k) & 1) ? ld_sc(&es->arrive[i0 + k * blockDim.x]) : epoch;
#pragma unroll
        for (int k = 0; k < kPoll; ++k)
          if (f[k] == epoch) pending &= ~(1u << k);
        if (!pending) break;
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > limit) {
          ok = 0;
          break;
        }
      }
    }
    if (!ok) {
      atomicOr(&es->error, 1u);
      *s_ok = 0;
    }
    __syncthreads();
    if (tr && threadIdx.x == 0) tr[3] = wall_clock64();
    if (*s_ok) fold();
    if (threadIdx.x < kSubQ) s_pre[threadIdx.x + 1] = ld_sc(&next_qlen[threadIdx.x]);
    __syncthreads();
    if (threadIdx.x == 0) {
      s_pre[0] = 0;
      for (int k = 1; k <= kSubQ; ++k) s_pre[k] += s_pre[k - 1];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kReleaseCopies * (kSubQ + 1); j += blockDim.x)
      st_sc(&es->rel[j / (kSubQ + 1)][1 + j % (kSubQ + 1)], s_pre[j % (kSubQ + 1)]);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tr && threadIdx.x == 0) tr[4] = wall_clock64();
    if (threadIdx.x < kReleaseCopies) {
      if (fence) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      st_sc(&es->rel[threadIdx.x][0], epoch);
    }
  } else if (threadIdx.x == 0) {
    const uint32_t *line = es->rel[blockIdx.x % kReleaseCopies];
    const uint64_t t0 = wall_clock64();
    while (ld_sc(&line[0]) != epoch) {
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > limit) {
        atomicOr(&es->error, 1u);
        *s_ok = 0;
        break;
      }
    }
    for (int k = 0; k <= kSubQ; ++k) s_pre[k] = ld_sc(&line[1 + k]);
  }
  // one thread reads the error flag, so the whole workgroup takes one decision
  if (threadIdx.x == 0 && ld_sc(&es->error)) *s_ok = 0;
  __syncthreads();
  if (fence) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return *s_ok != 0;
}

// Bookkeeping of the single-pair barrier master, kept in registers of thread 0
// of workgroup 0 for the whole launch (loaded from and stored back to PairCtr).
struct PassState {
  unsigned long long cut[2], cov[2], sum[2];
  uint32_t hd[2];
  int done_from[2];
  uint32_t cur_total;   // entries of the iteration running (global + kept)
  uint32_t next_total;  // entries queued for the next one
  // the engine's queue totals: kept here between barriers and stored when
  // the launch leaves, so no barrier waits on them
  unsigned long long qtotal;
  uint32_t qmax;
};

// The fold of one barrier period's hit counts, one bin per lane of a wave
// (all 64 lanes call it, the state arrays wave-uniform; lane l < 2 KS takes
// bin l: pass m = l % 2, step it0 + l / 2).  A pass whose cut is reached at
// step s takes the bins up to s and stops (hello.cu:1748-1759), and each
// taken bin goes to the histogram (bins at hist_len and beyond are dropped):
// the loop of that rule as prefix sums and ballots, with no branch per bin.
template <int KS, typename Bin>
__device__ __forceinline__ void fold_bins(const unsigned long long *cut, unsigned long long *cov,
                                          unsigned long long *sum, uint32_t *hd, int *done_from,
                                          int it0, Bin bin, uint32_t *hist0, uint32_t *hist1,
                                          int hist_len) {
  static_assert(2 * KS <= 64, "one bin per lane");
  const int lane = threadIdx.x & 63;
  const int m = lane & 1, k = lane >> 1;
  const bool mine = lane < 2 * KS;
  const int it = it0 + k;
  // a pass folds unless its cut is zero or an earlier period reached it
  const bool act0 = cut[0] != 0 && done_from[0] >= it0;
  const bool act1 = cut[1] != 0 && done_from[1] >= it0;
  const bool active = mine && (m ? act1 : act0);
  const unsigned long long h = mine ? bin(lane) : 0ull;
  unsigned long long pre = m ? cov[1] : cov[0];  // coverage through this bin
#pragma unroll
  for (int j = 0; j < KS; ++j)
    if (j <= k) pre += bin(m + 2 * j);
  const uint64_t R = __ballot(active && pre >= (m ? cut[1] : cut[0]));
  constexpr uint64_t kPass0 = ((1ull << (2 * KS)) - 1) & 0x5555555555555555ull;
  const uint64_t Rm = R & (m ? kPass0 << 1 : kPass0);
  const int kstar = Rm ? (int)(__builtin_ctzll(Rm) >> 1) : KS;  // the bin reaching the cut
  const bool taken = active && k <= kstar && h != 0;
  const uint64_t T = __ballot(taken);
  if (taken && it < hist_len) st_sc((m ? hist1 : hist0) + it, (uint32_t)h);
  const unsigned long long term = taken ? h * (unsigned long long)it : 0ull;
  const unsigned long long s0 = wave_sum64(m == 0 ? term : 0ull);
  const unsigned long long s1 = wave_sum64(m == 1 ? term : 0ull);
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    if (!(p ? act1 : act0)) continue;
    const uint64_t Rp = R & (kPass0 << p), Tp = T & (kPass0 << p);
    const int ks = Rp ? (int)(__builtin_ctzll(Rp) >> 1) : KS - 1;
    const int src = p + 2 * ks;
    cov[p] = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pre, src) |
             ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pre >> 32), src) << 32);
    sum[p] += p ? s1 : s0;
    if (Tp) hd[p] = (uint32_t)(it0 + (int)((63 - __builtin_clzll(Tp)) >> 1));
    if (Rp) done_from[p] = it0 + ks;
  }
}

// Grid barrier of a single-pair launch.  It has the same structure as
// grid_barrier, but every word that crosses workgroups carries its epoch
// (tag()), so neither side needs a second round trip:
//   arrival  after s_waitcnt 0 and a workgroup barrier, thread 0 stores
//            tag(epoch, x) for x = this iteration's hits of pass 0 and 1, the
//            entries it pushed to the global queue, and the entries it kept
//            in its own LDS queue for the next iteration;
//   master   polls the records (thread i watches workgroups i, i+256, ...,
//            which all flush into sub-queue i % kSubQ), sums them, folds the
//            hits into the histogram and its PassState, decides the passes'
//            stop (hello.cu:1748-1759), zeroes the queue counters of the
//            iteration after next, and stores tag(epoch, flags) and
//            tag(epoch, prefix offsets of the next global queue) into every
//            release replica;
//   others   poll their replica until all words carry the epoch.
// Flags: bit m pass m done, bit 2 error, bit 3 entries remain (global or kept).
// The master also turns the kept counts into each workgroup's first global
// entry: a prefix over workgroups of their free wave slots (slots minus kept
// entries), stored tagged in es->fslot[wg] and read by each workgroup in the
// same poll as its release replica; s_fs[0] receives it, s_fs[1] the total of
// free slots, so global entries go first to waves that hold no kept entry.
// tdone is the iteration that ran (~0u before the first), qz the set of queue
// counters to zero.  Waits are bounded as in grid_barrier.  s_cnt[0..3] holds the four counts
// and is zeroed here, as is *s_zero (the kept count of the queue that becomes
// the next iteration's push target).  Returns false on a timeout.
// KS: dilation steps per engine iteration (1, 2 or 8).  With KS > 1 the
// arrival carries the hits of every step (s_cnt[m] for the first, s_cnt[2 +
// 2k + m] for step k >= 1), and the fold records each step as its own
// histogram bin, stopping a pass at the bin that reaches its cut.
template <int KS = 1>
__device__ __forceinline__ bool grid_barrier_single(
    Engine *es, uint32_t epoch, uint64_t limit, int *s_ok, uint32_t *s_cnt, uint32_t *s_zero,
    PassState &ps, const PairView &d, uint32_t tdone, uint32_t *qstats, int qstats_len,
    uint64_t *tr, uint32_t *s_pre, uint32_t *s_flags, unsigned long long *s_red, uint32_t slots,
    uint8_t *s_kept, uint32_t *s_fs, int qz, uint32_t narr = 0, uint32_t nnext = 0) {
  static_assert(KS == 1 || KS == 2 || KS == kTbSteps, "steps per barrier");
  // narr: the workgroups that arrive (0: the grid); nnext: those that take
  // part in the next iteration, among which the free slots are numbered (0:
  // narr).  Workgroups 0 .. narr-1 arrive; workgroup 0 is always one of them.
  if (narr == 0) narr = gridDim.x;
  if (nnext == 0) nnext = narr;
  constexpr int kArr = KS == 1 ? 4 : 2 + 2 * KS;  // arrival words
  constexpr int kH = 2 * KS;                      // hit words, index m + 2k
  static_assert(kArr <= kArrWords, "arrival record");
  // arrival word of hit index i (pass i % 2, step i / 2)
  auto hw = [](int i) { return i < 2 ? i : i + 2; };
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    *s_ok = 1;
#pragma unroll
    for (int j = 0; j < kArr; ++j) st_sc(&es->arr[j][blockIdx.x], tag(epoch, s_cnt[j]));
#pragma unroll
    for (int j = 0; j < kArr; ++j) s_cnt[j] = 0;  // the next iteration's counts start here
    *s_zero = 0;
  }
  constexpr int kWords = kSubQ + 3;  // flags, kSubQ + 1 prefix offsets, free slots
  __shared__ unsigned long long s_hw[kH][kWavesPerWg];  // hits per (pass, step) per wave
  if (blockIdx.x == 0) {
    unsigned long long h[kH], np = 0, nl = 0;
#pragma unroll
    for (int i = 0; i < kH; ++i) h[i] = 0;
    int ok = 1;
    const uint64_t t0 = wall_clock64();
    // up to kPoll records per thread are polled together: all their loads are
    // in flight at once, so a grid of kPoll * 256 workgroups costs one round
    // trip per poll, not one per workgroup
    constexpr int kPoll = KS > 2 ? 2 : 4;
    for (uint32_t i0 = threadIdx.x; i0 < narr && ok; i0 += kPoll * blockDim.x) {
      uint32_t pending = 0;
#pragma unroll
      for (int k = 0; k < kPoll; ++k)
        if (i0 + k * blockDim.x < narr) pending |= 1u << k;
      while (pending) {
        unsigned long long w[kPoll][kArr];
#pragma unroll
        for (int k = 0; k < kPoll; ++k) {
          if ((pending >> k) & 1) {
#pragma unroll
            for (int j = 0; j < kArr; ++j) w[k][j] = ld_sc(&es->arr[j][i0 + k * blockDim.x]);
          }
        }
#pragma unroll
        for (int k = 0; k < kPoll; ++k) {
          bool all = ((pending >> k) & 1) != 0;
#pragma unroll
          for (int j = 0; j < kArr; ++j) all = all && tagged(w[k][j], epoch);
          if (all) {
#pragma unroll
            for (int i = 0; i < kH; ++i) h[i] += w[k][hw(i)] & kTagValue;
            np += w[k][2] & kTagValue, nl += w[k][3] & kTagValue;
            s_kept[i0 + k * blockDim.x] = (uint8_t)(w[k][3] & kTagValue);
            pending &= ~(1u << k);
          }
        }
        if (!pending) break;
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > limit) {
          ok = 0;
          break;
        }
      }
    }
    // sums: hits per pass and step and kept entries over all workgroups,
    // global pushes per sub-queue
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < kSubQ) s_red[12 + threadIdx.x] = 0;
#pragma unroll
    for (int i = 0; i < kH; ++i) h[i] = wave_sum64(h[i]);
    nl = wave_sum64(nl);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < kH; ++i) s_hw[i][wave] = h[i];
      s_red[8 + wave] = nl;
    }
    if (!ok) *s_ok = 0;
    atomicAdd(&s_red[12 + threadIdx.x % kSubQ], np);
    __syncthreads();
    if (tr && threadIdx.x == 0) tr[3] = wall_clock64();
    const bool good = *s_ok != 0;
    {  // free slots: thread i owns workgroups [i*c, i*c + c); wave scan, then waves
      const uint32_t nwg = nnext, c = (nwg + blockDim.x - 1) / blockDim.x;
      const uint32_t wb = threadIdx.x * c, we = min(nwg, wb + c);
      uint32_t part = 0;
      for (uint32_t w = wb; w < we; ++w) part += slots - min(slots, (uint32_t)s_kept[w]);
      const uint32_t x = wave_scan32(part);  // inclusive
      __shared__ uint32_t s_wsum[kWavesPerWg];
      if (lane == 63) s_wsum[wave] = x;
      __syncthreads();
      uint32_t run = x - part, total = 0;
#pragma unroll
      for (int k = 0; k < kWavesPerWg; ++k) {
        if (k < wave) run += s_wsum[k];
        total += s_wsum[k];
      }
      for (uint32_t w = wb; w < we; ++w) {
        st_sc(&es->fslot[w], tag(epoch, run));
        run += slots - min(slots, (uint32_t)s_kept[w]);
      }
      if (threadIdx.x == 0) s_fs[0] = 0, s_fs[1] = total;  // workgroup 0's range starts at 0
    }
    // the fold by wave 0 (whose lanes all hold the pass state): this engine
    // iteration covered dilation steps KS tp + 1 .. KS tp + KS
    if (good && wave == 0)
      fold_bins<KS>(ps.cut, ps.cov, ps.sum, ps.hd, ps.done_from,
                    KS * (int)(tdone - d.t_start) + 1,
                    [&](int i) { return s_hw[i][0] + s_hw[i][1] + s_hw[i][2] + s_hw[i][3]; },
                    d.hist0, d.hist1, d.hist_len);
    if (threadIdx.x == 0) {
      uint32_t flags = 0;
      uint32_t acc = 0;
      for (int k = 0; k < kSubQ; ++k) {
        s_pre[k] = acc;
        acc += (uint32_t)s_red[12 + k];
      }
      s_pre[kSubQ] = acc;
      if (good) {
        const uint32_t kept = (uint32_t)(s_red[8] + s_red[9] + s_red[10] + s_red[11]);
        const int tp = (int)(tdone - d.t_start);
#pragma unroll
        for (int m = 0; m < 2; ++m)
          if (ps.cut[m] == 0 || ps.done_from[m] <= KS * tp + KS) flags |= 1u << m;
        ps.qtotal += ps.cur_total;
        ps.qmax = max(ps.qmax, ps.cur_total);
        if (qstats && tdone < (uint32_t)qstats_len) qstats[tdone] = ps.cur_total;
        ps.next_total = acc + kept;
        ps.cur_total = ps.next_total;
        if (ps.next_total) flags |= 8u;
      } else {
        flags = 4u;
        atomicOr(&es->error, 1u);
      }
      s_flags[0] = flags;
    }
    // the counters set qz (the one the iteration after next fills)
    if (threadIdx.x < kSubQ) st_sc(&es->qlen[qz][threadIdx.x], 0u);
    __syncthreads();
    if (tr && threadIdx.x == 0) tr[4] = wall_clock64();
    for (int j = threadIdx.x; j < kReleaseCopies * kWords; j += blockDim.x) {
      const int w = j % kWords;
      const uint32_t val = w == 0 ? s_flags[0] : w <= kSubQ + 1 ? s_pre[w - 1] : s_fs[1];
      st_sc(&es->rel2[j / kWords][w], tag(epoch, val));
    }
  } else {
    __shared__ unsigned long long s_rel[kWords + 1];
    // a workgroup that takes no part in the next iteration has no first slot
    if (threadIdx.x == kWords && blockIdx.x >= nnext) s_rel[kWords] = 0;
    if (threadIdx.x < kWords || (threadIdx.x == kWords && blockIdx.x < nnext)) {
      // the replica's words, then this workgroup's first slot
      const unsigned long long *src = threadIdx.x < kWords
                                          ? &es->rel2[blockIdx.x % kReleaseCopies][threadIdx.x]
                                          : &es->fslot[blockIdx.x];
      const uint64_t t0 = wall_clock64();
      unsigned long long w;
      for (;;) {
        w = ld_sc(src);
        if (tagged(w, epoch)) break;
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > limit) {
          atomicOr(&es->error, 1u);
          w = 4u;  // flags word or not, the workgroup leaves
          *s_ok = 0;
          break;
        }
      }
      s_rel[threadIdx.x] = w & kTagValue;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_flags[0] = (uint32_t)s_rel[0];
      for (int k = 0; k <= kSubQ; ++k) s_pre[k] = (uint32_t)s_rel[1 + k];
      s_fs[1] = (uint32_t)s_rel[kSubQ + 2];
      s_fs[0] = (uint32_t)s_rel[kWords];
      if (s_flags[0] & 4u) *s_ok = 0;
    }
  }
  __syncthreads();
  return *s_ok != 0;
}

// Grid barrier between the phases of the fused single-pair launch, carrying no
// data: each workgroup drains its stores and stores tag(epoch, 0) into word 0
// of its record; workgroup 0 waits for all of them and stores tag(epoch,
// flags) into word 0 of every release replica (flags bit 2: a wait timed out).
// Bounded like the other barriers; returns false on a timeout.
__device__ __forceinline__ bool grid_sync_tagged(Engine *es, uint32_t epoch, uint64_t limit,
                                                 int *s_ok) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    *s_ok = 1;
    st_sc(&es->arr[0][blockIdx.x], tag(epoch, 0));
  }
  if (blockIdx.x == 0) {
    int ok = 1;
    const uint64_t t0 = wall_clock64();
    constexpr int kPoll = 4;
    for (uint32_t i0 = threadIdx.x; i0 < gridDim.x && ok; i0 += kPoll * blockDim.x) {
      uint32_t pending = 0;
#pragma unroll
      for (int k = 0; k < kPoll; ++k)
        if (i0 + k * blockDim.x < gridDim.x) pending |= 1u << k;
      while (pending) {
        unsigned long long w[kPoll];
#pragma unroll
        for (int k = 0; k < kPoll; ++k)
          w[k] = ((pending >> k) & 1) ? ld_sc(&es->arr[0][i0 + k * blockDim.x]) : 0ull;
#pragma unroll
        for (int k = 0; k < kPoll; ++k)
          if (((pending >> k) & 1) && tagged(w[k], epoch)) pending &= ~(1u << k);
        if (!pending) break;
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > limit) {
          ok = 0;
          break;
        }
      }
    }
    __syncthreads();
    if (!ok) *s_ok = 0;
    __syncthreads();
    if (threadIdx.x == 0 && !*s_ok) atomicOr(&es->error, 1u);
    if (threadIdx.x < kReleaseCopies)
      st_sc(&es->rel2[threadIdx.x][0], tag(epoch, *s_ok ? 0u : 4u));
  } else if (threadIdx.x == 0) {
    const unsigned long long *w0 = &es->rel2[blockIdx.x % kReleaseCopies][0];
    const uint64_t t0 = wall_clock64();
    for (;;) {
      const unsigned long long w = ld_sc(w0);
      if (tagged(w, epoch)) {
        if (w & 4u) *s_ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > limit) {
        atomicOr(&es->error, 1u);
        *s_ok = 0;
        break;
      }
    }
  }
  __syncthreads();
  return *s_ok != 0;
}

// LDS table of per-(pair, step, mask) hit counts of the running iteration
constexpr int kHitTable = 64;
struct HitTable {
  uint32_t key[kHitTable];  // pair * 4 + step * 2 + mask, or ~0u
  uint32_t cnt[kHitTable];
};
__device__ __forceinline__ void hit_add(HitTable &h, PairCtr *ctrs, uint32_t key, uint32_t v) {
  uint32_t slot = key % kHitTable;
  for (int probe = 0; probe < 8; ++probe, slot = (slot + 1) % kHitTable) {
    const uint32_t k = atomicCAS(&h.key[slot], ~0u, key);
    if (k == ~0u || k == key) {
      atomicAdd(&h.cnt[slot], v);
      return;
    }
  }
  atomicAdd(&ctrs[key >> 2].hits[key & 3][blockIdx.x % kHitSlots],
            (unsigned long long)v);  // table full
}
__device__ __forceinline__ void hit_flush(HitTable &h, PairCtr *ctrs) {
  __syncthreads();
  if (threadIdx.x < kHitTable) {
    const uint32_t k = h.key[threadIdx.x];
    if (k != ~0u) {
      atomicAdd(&ctrs[k >> 2].hits[k & 3][blockIdx.x % kHitSlots],
                (unsigned long long)h.cnt[threadIdx.x]);
      h.key[threadIdx.x] = ~0u;
      h.cnt[threadIdx.x] = 0;
    }
  }
}

// ---------------------------------------------------------------- finalize
// Result of pair d at engine iteration t_engine.  The barrier master kept, per
// pass, the covered count, the sum of first-hit iterations and the last
// iteration with a hit while it folded the histogram bins in order, and it
// stops folding a pass at the bin where the covered count reaches the cut
// ceil(r * total) (hello.cu:1748-1759, 1887-1895); so HD, sum and count are
// read, not rescanned.
// steps: dilation iterations per engine iteration (2 when each barrier
// covers two steps)
// The same from explicit values: totals, and per pass the covered count, the
// sum of first hits, the last hit and the stop iteration.
__device__ PairOut finalize_from(const PairDesc &d, unsigned long long n_gold,
                                 unsigned long long n_segm, unsigned long long n_fp,
                                 unsigned long long n_fn, const unsigned long long *cov,
                                 const unsigned long long *sum_d, const uint32_t *hd_last,
                                 const int *done_from, int status, uint32_t t_engine, bool error,
                                 int steps) {
  PairOut o;
  const int tl = (int)(t_engine - d.t_start) * steps;  // pair-local iterations so far
  bool done = true, reached = true;
  unsigned long long sum[2], cnt[2];
  int hd[2], last = 0;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const unsigned long long cut = robust_cut(d.robust, m == 0 ? n_fp : n_fn);
    sum[m] = cut ? sum_d[m] : 0ull;
    cnt[m] = cut ? cov[m] : 0ull;
    hd[m] = cut ? (int)hd_last[m] : 0;
    const int df = done_from[m];
    done = done && (cut == 0 || df <= tl);
    if (cut) last = max(last, df);
    reached = reached && cnt[m] >= cut;
  }
  o.iters = done ? last : tl;
  o.steps = steps;
  o.pad = 0;
  o.hd_fp = hd[0], o.hd_fn = hd[1];
  o.hd = max(hd[0], hd[1]);
  o.sum_fp = sum[0], o.cnt_fp = cnt[0], o.sum_fn = sum[1], o.cnt_fn = cnt[1];
  o.n_gold = n_gold, o.n_segm = n_segm, o.n_fp = n_fp, o.n_fn = n_fn;
  o.done = done ? 1 : 0;
  int st = status;
  if ((n_fp && !n_gold) || (n_fn && !n_segm))
    st = -2;  // HDM_E_EMPTY_MASK
  else if (error)
    st = -5;  // HDM_E_TIMEOUT
  else if (st == 0 && done && !reached)
    st = -8;  // HDM_E_STUCK
  o.status = st;
  return o;
}
__device__ PairOut finalize_pair(const PairDesc &d, const PairCtr &c, uint32_t t_engine,
                                 bool error, int steps = 1) {
  unsigned long long cov[2], sum[2];
  uint32_t hd[2];
  int df[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    cov[m] = ld_sc(&c.cov[m]);
    sum[m] = ld_sc(&c.sum_d[m]);
    hd[m] = ld_sc(&c.hd_last[m]);
    df[m] = ld_sc(&c.done_from[m]);
  }
  // a pair the batched near pass finished is done at its stop steps (at most
  // eight), however many iterations the engine ran
  const bool nd = ld_sc(&c.near_done) != 0;
  return finalize_from(d, ld_sc(&c.n_gold), ld_sc(&c.n_segm), ld_sc(&c.n_fp), ld_sc(&c.n_fn), cov,
                       sum, hd, df, ld_sc(&c.status), nd ? d.t_start + kTbSteps : t_engine, error,
                       nd ? 1 : steps);
}

// A lower bound of the HD in dilation steps (6 element, 64x8x8 tiles) from
// the tile boxes of gold (g: lo x, y, z, hi x, y, z) and segm (s): a segm
// voxel outside gold's box is an FP voxel, at least as far from gold as from
// gold's box, and a gold voxel outside segm's box is an FN voxel.  A box that
// reaches d tiles beyond the other's along an axis holds a voxel at least
// (d - 1) * tile size + 1 voxels beyond it.  0 when a mask is empty.
__device__ __forceinline__ int hd_lower_bound(const int *g, const int *s) {
  if (g[3] < 0 || s[3] < 0) return 0;
  constexpr int kT[3] = {64, 8, 8};
  int lb = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int d[4] = {s[3 + k] - g[3 + k], g[k] - s[k], g[3 + k] - s[3 + k], s[k] - g[k]};
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (d[j] >= 1) lb = max(lb, (d[j] - 1) * kT[k] + 1);
  }
  return lb;
}
// Pairs whose HD bound reaches this many steps go to the region launch
// (eight steps per barrier); the others stay with the two-step launch, which
// keeps more waves busy in the dense first iterations of a short run.
constexpr int kTbMinSteps = 16;

// ---------------------------------------------------------------- dilate
// SINGLE: the launch holds one pair, so its descriptor and its per-iteration
// pass state are loop invariants instead of per-entry loads.
// FUSED (one pair): the launch first runs the reset, pack and seed passes
// itself, separated by grid barriers, so one call is one launch; every store
// that a later phase reads in another workgroup is written through, and those
// reads are agent-coherent.
template <int TXB, int TY, int TZ, bool SINGLE, bool FUSED = false, int CONN = 6,
          bool DS = false>
__device__ __forceinline__ void dilate_body(const DilateArgs &a) {
  static_assert(!DS || (TXB == 64 && TY == 8 && TZ == 8 && CONN == 6),
                "two steps per barrier: 64x8x8 tiles, 6 element");
  static_assert(!FUSED || SINGLE, "the fused launch holds one pair");
  const PairDesc *__restrict__ descs = a.descs;
  PairCtr *ctrs = a.ctrs;
  const int npairs = a.npairs;
  Engine *es = a.es;
  uint64_t *q = a.q;
  const uint32_t qcap = a.qcap;
  const int max_iters = a.max_iters, chunk = a.chunk;
  const uint64_t spin_limit = a.spin;
  uint32_t *qstats = a.qstats;
  const int qstats_len = a.qstats_len, flags = a.flags;
  uint64_t *trace = a.trace;
  using Row = RowT<TXB>;
  constexpr int R = TY * TZ;
  constexpr int TPW = kWave / R;
  constexpr int kXf = xf_words(TXB, TY, TZ);
  __shared__ LocalQueue lq;
  __shared__ HitTable ht;
  __shared__ uint32_t s_pre[kSubQ + 1];
  __shared__ int s_ok;
  // single pair: this iteration's hits per pass, entries pushed to the global
  // queue and entries kept; the released flags; the master's reduction space;
  // the workgroup's own queue of the running and of the next iteration
  __shared__ uint32_t s_cnt[6];
  __shared__ uint32_t s_flags[1];
  __shared__ unsigned long long s_red[12 + kSubQ];
  // kept entries: at most one per wave slot (kSlots), taken by slots 0..;
  // the other slots take global entries from the workgroup's range of the
  // master's free-slot prefix (s_fs), so while the queue fits, no wave runs
  // two steps in a row
  constexpr uint32_t kSlots = kWavesPerWg * TPW;
  constexpr uint32_t kStick = kSlots;
  __shared__ uint64_t s_st[2][kStick];
  __shared__ uint32_t s_stn[2];
  __shared__ uint8_t s_kept[kMaxGrid];
  __shared__ uint32_t s_fs[2];
  const bool fence = (flags & 1) != 0;
  constexpr bool single = SINGLE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int sub = lane / R;
  const int r = lane % R;
  const int yl = r % TY, zl = r / TY;
  const int gbase = lane - r;  // first lane of this lane's group
  const uint32_t gwave = blockIdx.x * kWavesPerWg + wave;
  const uint32_t nwave = gridDim.x * kWavesPerWg;
  const uint32_t step = (uint32_t)chunk * TPW;
  const uint32_t kq = blockIdx.x % kSubQ;
  if (threadIdx.x == 0) lq.n = 0;
  if (threadIdx.x < kHitTable) ht.key[threadIdx.x] = ~0u, ht.cnt[threadIdx.x] = 0;
  if (threadIdx.x < 6) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < 2) s_stn[threadIdx.x] = 0;
  if (threadIdx.x == 0) {  // no kept entries yet: every slot is free
    s_fs[0] = blockIdx.x * kSlots;
    s_fs[1] = gridDim.x * kSlots;
  }
  PairView v0{};  // one pair: its fields for the whole launch
  PassState ps{};
  uint32_t t = 0;
  bool run = true;
  // epoch of a dilation barrier = eoff + the iteration that follows it; the
  // fused launch's prologue uses ebase + 2 (end of the pack) and ebase + 3
  const uint32_t eoff = FUSED ? a.ebase + 3 : 0u;
  __shared__ unsigned long long s_phase[3];
  if constexpr (FUSED) {
    const PairDesc &d = a.desc;
    PairCtr &c0 = ctrs[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) s_phase[0] = wall_clock64();
    // R: workgroup 0 zeroes the pair's counters and histograms, the engine's
    // state (barrier words are tagged with ever larger epochs) and the pack
    // partials the next launch packs into.  Nothing here is read before the
    // barrier that ends the pack, except by workgroup 0 itself, and this
    // launch packs into the set the previous launch zeroed, so the pack
    // starts at once.
    PackSums *acc = &es->acc[a.acc_set & 1];
    if (blockIdx.x == 0) {
      const int tid = threadIdx.x;
      if (tid == 0) {
        st_sc(&c0.n_gold, 0ull), st_sc(&c0.n_segm, 0ull);
        st_sc(&c0.n_fp, 0ull), st_sc(&c0.n_fn, 0ull);
        for (int m = 0; m < 2; ++m) {
          st_sc(&c0.cov[m], 0ull), st_sc(&c0.sum_d[m], 0ull);
          st_sc(&c0.done_from[m], 0x7FFFFFFF), st_sc(&c0.hd_last[m], 0u);
        }
        st_sc(&c0.status, 0), st_sc(&c0.near_done, 0);
        st_sc(&es->t, 0u), st_sc(&es->error, 0u), st_sc(&es->qmax, 0u);
        st_sc(&es->qtotal, 0ull), st_sc(&es->qcap, qcap), st_sc(&es->qnext, 0u);
      }
      PackSums *next = &es->acc[(a.acc_set & 1) ^ 1];
      for (int i = tid; i < 4 * kSumSlots; i += blockDim.x) st_sc(&next->sums[0][0] + i, 0ull);
      for (int i = tid; i < 18 * kSumSlots; i += blockDim.x) st_sc(&next->box[0][0] + i, 0);
      if (tid < 8) st_sc(&next->grab[32 * tid], 0u);
      for (int i = tid; i < 2 * kTbSteps * kHitSlots; i += blockDim.x)
        st_sc(&next->near_hits[0][0] + i, 0ull);
      if (tid == 0) st_sc(&next->near_done, 0u);
      for (int i = tid; i < d.hist_len; i += blockDim.x)
        st_sc(d.hist[0] + i, 0u), st_sc(d.hist[1] + i, 0u);
      for (int i = tid; i < 4 * kLineWords; i += blockDim.x) st_sc(&es->qlen[0][0] + i, 0u);
    }
    // P: pack, every store written through
    {
      __shared__ __align__(16) uint16_t s_pieces[kWavesPerWg][1024];
      pack_tiles<TXB, TY, TZ, true>(&d, nullptr, ctrs, 1, a.total_tiles, s_pieces, acc,
                                    a.pack_pools ? &acc->grab[0] : nullptr, a.pack_pools);
      if (blockIdx.x == 0 && threadIdx.x == 0) s_phase[1] = wall_clock64();
      run = grid_sync_tagged(es, a.ebase + 2, spin_limit, &s_ok);
    }
    // S: totals and box from the pack partials, then tile states and the
    // first queue, whose entries go to this workgroup's own queue first
    if (run) {
      __shared__ unsigned long long s_tot[4];
      __shared__ int s_box[6];
      __shared__ int s_hand;  // 1: the pair goes to the region launch
      // the flags of this thread's first tile and of its neighbours in the
      // tile grid load together with the pack partials, one round trip for
      // both; the box, known once the partials are summed, selects among them
      constexpr bool kDiag = DS || CONN != 6;  // neighbours beyond the six faces
      uint32_t pf_own = 0, pf[27];
      {
        const uint64_t g = (uint64_t)blockIdx.x * kWgThreads + threadIdx.x;
        const bool in = g < a.total_tiles;
        const uint32_t tile = in ? (uint32_t)g : 0u;
        const int tx = (int)(tile % (uint32_t)d.gx);
        const int ty = (int)((tile / (uint32_t)d.gx) % (uint32_t)d.gy);
        const int tz = (int)(tile / ((uint32_t)d.gx * d.gy));
        if (in) pf_own = ld_sc(d.tflags + tile);
#pragma unroll
        for (int j = 0; j < 27; ++j) {
          pf[j] = 0;
          const int ox = j % 3 - 1, oy = (j / 3) % 3 - 1, oz = j / 9 - 1;
          const int nzc = (ox != 0) + (oy != 0) + (oz != 0);
          if (nzc == 0 || (!kDiag && nzc > 1) || (DS && nzc == 3)) continue;
          const int ux = tx + ox, uy = ty + oy, uz = tz + oz;
          if (in && ux >= 0 && ux < d.gx && uy >= 0 && uy < d.gy && uz >= 0 && uz < d.gz)
            pf[j] = ld_sc(d.tflags + (uint32_t)((int)tile + ox + oy * d.gx + oz * d.gx * d.gy));
        }
      }
      if (threadIdx.x < 64) {
        unsigned long long v[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          v[f] = ld_sc(&acc->sums[f][lane]);
          v[f] = wave_sum64(v[f]);
        }
        int bx[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const int x = wave_max(ld_sc(&acc->box[k][lane]), lane);
          bx[k] = k < 3 ? box_dec_lo(x) : box_dec_hi(x);
        }
        // hand the pair to the region launch (two-step launch only): always
        // (tb_mode 2), or when the HD bound from the gold and segm boxes
        // (rows 6..17) reaches kTbMinSteps (tb_mode 1)
        bool hand = false;
        if constexpr (DS) {
          if (a.tb_mode == 2) {
            hand = true;
          } else if (a.tb_mode == 1) {
            int gs[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) {
              const int x = wave_max(ld_sc(&acc->box[6 + k][lane]), lane);
              gs[k] = k % 6 < 3 ? box_dec_lo(x) : box_dec_hi(x);
            }
            hand = hd_lower_bound(gs, gs + 6) >= kTbMinSteps;
          }
        }
        if (lane == 0) {
#pragma unroll
          for (int f = 0; f < 4; ++f) s_tot[f] = v[f];
#pragma unroll
          for (int k = 0; k < 6; ++k) s_box[k] = bx[k];
          s_hand = hand ? 1 : 0;
          if (blockIdx.x == 0) {
            st_sc(&c0.n_gold, v[0]), st_sc(&c0.n_segm, v[1]);
            st_sc(&c0.n_fp, v[2]), st_sc(&c0.n_fn, v[3]);
#pragma unroll
            for (int k = 0; k < 3; ++k) st_sc(&c0.lo[k], bx[k]), st_sc(&c0.hi[k], bx[3 + k]);
          }
        }
      }
      __syncthreads();
      if (DS && s_hand) {
        // the region launch takes over from the packed pair: its region
        // states start idle, and workgroup 0 leaves the tagged hand-off word
        // (with the clock stamps of the start and the pack) and, in place of
        // the results, status 2 and the ready marker: the host then queues
        // the region launch, which starts once every workgroup here is done
        const uint32_t nreg = (uint32_t)d.gx * (uint32_t)((d.gy + 1) / 2) *
                              (uint32_t)((d.gz + 1) / 2);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nreg;
             i += gridDim.x * blockDim.x)
          st_sc(d.rstate[0] + i, 0u), st_sc(d.rstate[1] + i, 0u);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
          st_sc(&es->handoff[1], s_phase[0]);
          st_sc(&es->handoff[2], s_phase[1]);
          st_sc(&es->handoff[0], tag(a.ebase, 2ull));
          __hip_atomic_store(&a.outs[0].status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          if (a.ready) {
            __threadfence_system();
            __hip_atomic_store(a.ready, a.ready_seq, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
          }
        }
        return;
      }
      v0 = make_view(d, s_box, s_box + 3, s_tot[2], s_tot[3]);
      if (blockIdx.x == 0 && threadIdx.x < 64) {  // wave 0 holds the pass state
        for (int m = 0; m < 2; ++m) {
          ps.cut[m] = v0.cut[m];
          ps.cov[m] = ps.sum[m] = 0;
          ps.hd[m] = 0;
          ps.done_from[m] = 0x7FFFFFFF;
        }
      }
      // the rule of seed_kernel, with agent-coherent flag reads
      uint64_t *Q0 = sub_queue(q, qcap, 0, kq);
      uint32_t *Q0len = &es->qlen[0][kq];
      for (uint64_t tb = (uint64_t)blockIdx.x * kWgThreads; tb < a.total_tiles;
           tb += (uint64_t)gridDim.x * kWgThreads) {
        const uint64_t g = tb + threadIdx.x;
        const bool in = g < a.total_tiles;
        const uint32_t tile = in ? (uint32_t)g : 0u;
        const int tx = (int)(tile % (uint32_t)v0.gx);
        const int ty = (int)((tile / (uint32_t)v0.gx) % (uint32_t)v0.gy);
        const int tz = (int)(tile / ((uint32_t)v0.gx * v0.gy));
        const Box &b = v0.b;
        const bool in_box = in && tx >= b.lo[0] && tx <= b.hi[0] && ty >= b.lo[1] &&
                            ty <= b.hi[1] && tz >= b.lo[2] && tz <= b.hi[2];
        const bool first = tb == (uint64_t)blockIdx.x * kWgThreads;  // prefetched
        uint32_t own = 0, reach = 0;
        if (in_box) {
          own = first ? pf_own : ld_sc(d.tflags + tile);
          reach = seed_reach<CONN, DS>(v0, b, tile, tx, ty, tz,
                                       [&](int j, uint32_t nb) -> uint32_t {
                                         return first ? pf[j] : ld_sc(d.tflags + nb);
                                       });
        }
        bool want[2];
        uint64_t ent[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          const uint32_t f = (own >> (16 * m)) & 0xFFFF;
          const bool any = f & 1, full = (f >> 1) & 1;
          const bool act = ((reach >> m) & 1u) != 0;
          const bool needed = v0.cut[m] != 0;
          const uint32_t st = !in_box ? 0u : (any && full) ? kFull : (any || (act && needed)) ? 1u : 0u;
          if (in) st_sc((m ? v0.state[1] : v0.state[0]) + tile, st);
          want[m] = in_box && needed && (any ? !full : act);
          ent[m] = make_entry(tile, 0u, (uint32_t)m, 1, 1, v0.gen);
        }
        push_sticky(want[0], ent[0], s_st[0], &s_stn[0], kStick, lq, Q0, Q0len, lane, &s_cnt[2]);
        push_sticky(want[1], ent[1], s_st[0], &s_stn[0], kStick, lq, Q0, Q0len, lane, &s_cnt[2]);
      }
      const uint32_t nflushed = flush_local_queue(lq, Q0, Q0len);
      if (threadIdx.x == 0) s_cnt[2] += nflushed, s_cnt[3] = s_stn[0];
      if (blockIdx.x == 0 && threadIdx.x == 0) s_phase[2] = wall_clock64();
      // hands iteration 0 its flags, global prefix and free slots
      run = grid_barrier_single<DS ? 2 : 1>(es, a.ebase + 3, spin_limit, &s_ok, s_cnt, &s_stn[1], ps, v0,
                                ~0u, qstats, qstats_len, nullptr, s_pre, s_flags, s_red, kSlots,
                                s_kept, s_fs, 2);
    }
  } else {
    if constexpr (SINGLE) v0 = make_view(descs[0], ctrs[0]);
    if (SINGLE && blockIdx.x == 0 && threadIdx.x < 64) {  // the master's registers (wave 0)
      const PairCtr &c0 = ctrs[0];
      for (int m = 0; m < 2; ++m) {
        ps.cut[m] = robust_cut(descs[0].robust, ld_sc(m == 0 ? &c0.n_fp : &c0.n_fn));
        ps.cov[m] = ld_sc(&c0.cov[m]);
        ps.sum[m] = ld_sc(&c0.sum_d[m]);
        ps.hd[m] = ld_sc(&c0.hd_last[m]);
        ps.done_from[m] = ld_sc(&c0.done_from[m]);
      }
      ps.qtotal = ld_sc(&es->qtotal);
      ps.qmax = ld_sc(&es->qmax);
    }
    __syncthreads();
    t = ld_sc(&es->t);
  }
  for (int it = 0; run && it < max_iters; ++it) {
    // queue of iteration t: sub-queue lengths -> prefix offsets (after the
    // first iteration the barrier delivers them with the release)
    if (it == 0 && !FUSED) {
      if (threadIdx.x < kSubQ) s_pre[threadIdx.x + 1] = ld_sc(&es->qlen[t % 4][threadIdx.x]);
      __syncthreads();
      if (threadIdx.x == 0) {
        s_pre[0] = 0;
        for (int k = 1; k <= kSubQ; ++k) s_pre[k] += s_pre[k - 1];
      }
      __syncthreads();
    }
    const uint32_t L = s_pre[kSubQ];  // entries in the global queue
    // one pair: entries may also sit in the workgroups' own queues; the
    // release says whether any remain anywhere (uniform decision)
    // (the fused launch's prologue barrier released the first iteration too)
    const bool first_read = it == 0 && !FUSED;
    if (SINGLE ? (first_read ? L == 0 : (s_flags[0] & 8u) == 0) : L == 0) break;
    if (SINGLE && first_read && blockIdx.x == 0 && threadIdx.x == 0) ps.cur_total = L;
    bool pdone[2] = {false, false};
    if (single) {  // stop as soon as both passes are done (uniform decision)
      if (first_read) {
        const int tp0 = (int)(t - descs[0].t_start) * (DS ? 2 : 1);  // dilation steps
        pdone[0] = pass_done(descs[0], ctrs[0], 0, tp0);
        pdone[1] = pass_done(descs[0], ctrs[0], 1, tp0);
      } else {  // delivered with the release
        pdone[0] = (s_flags[0] & 1u) != 0;
        pdone[1] = (s_flags[0] & 2u) != 0;
      }
      if (pdone[0] && pdone[1]) break;
    }
    // per-iteration trace (diagnostic): [0] start, [1] last workgroup done,
    // [2] workgroup 0 done, [3] all arrived, [4] folded, [5] queue length
    uint64_t *tr = (trace && (int)t < qstats_len) ? trace + 8 * (uint64_t)t : nullptr;
    if (tr && blockIdx.x == 0 && threadIdx.x == 0) {
      tr[0] = wall_clock64();
      tr[5] = SINGLE ? ps.cur_total : L;
    }
    uint64_t *Qout = sub_queue(q, qcap, (t + 1) & 1, kq);
    uint32_t *qout_len = &es->qlen[(t + 1) % 4][kq];
    // a step's claims of idle neighbours (compare-and-swap on their state)
    // are pushed with the next step's entries, or after the wave's last step,
    // so the swap's round trip overlaps that step's loads or the stores the
    // barrier waits for anyway; claim_old: the state the swap found (0: this
    // lane won the tile), claim_ent: the entry (0: no claim)
    uint32_t claim_old = 0;
    uint64_t claim_ent = 0;
    // one step of queue entry ent (live: a real entry) by the R lanes of its
    // group; pushes the entries it produces
    auto step_tile = [&](uint64_t ent, bool live) {
        if constexpr (TPW == 1 && !SINGLE) {
          // one tile per wave: every lane holds the same entry, so the tile's
          // coordinates and its pair's fields live in scalar registers and
          // the descriptor is read with scalar loads (one pair keeps its
          // fields in registers anyway, and measured no faster with it)
          ent = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ent >> 32))
                 << 32) |
                (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ent);
        }
        const uint32_t tile = (uint32_t)ent;
        const int p = SINGLE ? 0 : (int)((ent >> 32) & 0xFFFF);
        const int m = (int)((ent >> 48) & 1);
        const bool first = (ent >> 49) & 1;
        const bool check = (ent >> 50) & 1;
        // the pair's fields: registers for the launch (one pair) or one batch
        // of loads together with the pass's stop state (several pairs)
        PairView v;
        int dfrom = 0x7FFFFFFF;
        if constexpr (SINGLE) {
          v = v0;
        } else {
          v = make_view(descs[p], ctrs[p]);
          dfrom = ld_sc(&ctrs[p].done_from[m]);
        }
        const uint32_t tp = t - v.t_start;  // pair-local iteration
        if (!SINGLE && live && entry_gen(ent) != (v.gen & kGenMask)) live = false;  // stale
        // a pass whose cut was reached stops (hello.cu:1748-1759)
        const unsigned long long cut_m = m ? v.cut[1] : v.cut[0];
        if (live && (SINGLE ? pdone[m] : (cut_m == 0 || dfrom <= (DS ? 2 : 1) * (int)tp)))
          live = false;
        if (live && (DS ? 2 * tp + 2 : tp + 1) >= (uint32_t)v.hist_len) {
          st_sc(&ctrs[p].status, -8);
          live = false;
        }
        const int tx = (int)(tile % (uint32_t)v.gx);
        const int ty = (int)((tile / (uint32_t)v.gx) % (uint32_t)v.gy);
        const int tz = (int)(tile / ((uint32_t)v.gx * v.gy));
        const int y = ty * TY + yl, z = tz * TZ + zl;
        const bool exists = y < v.ny && z < v.nz;
        const bool rowok = live && exists;
        const Row vb = valid_bits<Row>(tx, v.nx);
        const int rd = (tp + 1) & 1;  // slot holding iteration tp's input
        const int wr = tp & 1;        // slot receiving iteration tp's output
        // pointer selects, not indexing: indexing the view with run-time
        // values would put it in scratch memory
        const Row *orig_m = (const Row *)(m ? v.orig[1] : v.orig[0]);
        const Row *orig_o = (const Row *)(m ? v.orig[0] : v.orig[1]);
        Row *buf0 = (Row *)(m ? v.buf[1][0] : v.buf[0][0]);
        Row *buf1 = (Row *)(m ? v.buf[1][1] : v.buf[0][1]);
        const Row *bufr = rd ? buf1 : buf0;
        Row *bufw = wr ? buf1 : buf0;
        const uint64_t *xfo = m ? v.xf_orig[1] : v.xf_orig[0];
        uint64_t *xf0 = m ? v.xf_buf[1][0] : v.xf_buf[0][0];
        uint64_t *xf1 = m ? v.xf_buf[1][1] : v.xf_buf[0][1];
        const uint64_t *xfr = rd ? xf1 : xf0;
        uint64_t *xfw = wr ? xf1 : xf0;
        uint32_t *st_m = m ? v.state[1] : v.state[0];
        const uint64_t ro = (uint64_t)tile * R + r;
        // -- issue every load of this step before using any of them
        Row own = 0;
        if (rowok) own = first ? ldg(orig_m + ro) : ld_sc(bufr + ro);
        const Row other = (rowok && check) ? ldg(orig_o + ro) : Row(0);
        uint32_t nb_dir, nst;  // the neighbour this lane activates and its state
        Row dl, d1 = 0;
        if constexpr (DS) {
          // Two 6-connectivity steps in one (64x8x8 tiles, lane = row): the
          // second step's output is the union over the L1 ball of radius 2,
          //   D2 = x2(own) | x1(rows y+-1, z+-1) | rows y+-2, z+-2 | rows (y+-1, z+-1)
          // and the first step's D1 = x1(own) | rows y+-1, z+-1 gives the hits
          // of the first step.  Beyond the tile: rows at depth 1 and 2 of the y
          // and z neighbours (face lanes), the yz-diagonal row (corner lanes),
          // the x-face words of the x-side tiles at (y, z), (y+-1), (z+-1) for
          // the one-bit x halo of the rows y+-1, z+-1 (8x8 bit planes), and the
          // second-column face words of the x neighbours at (y, z) for the
          // two-bit x halo of the row itself.
          // Lane j < 27 reads the state of neighbour j = (ox+1) + 3(oy+1) +
          // 9(oz+1) of the 18 that two steps can reach (no corners).
          static_assert(TXB == 64 && TY == 8 && TZ == 8 && CONN == 6, "64x8x8, 6 element");
          const int j = lane;
          const int jox = j % 3 - 1, joy = (j / 3) % 3 - 1, joz = j / 9 - 1;
          const int jn = (jox != 0) + (joy != 0) + (joz != 0);
          nb_dir = kNoTile;
          if (live && j < 27 && jn >= 1 && jn <= 2) {
            const int ux = tx + jox, uy = ty + joy, uz = tz + joz;
            if (ux >= v.b.lo[0] && ux <= v.b.hi[0] && uy >= v.b.lo[1] && uy <= v.b.hi[1] &&
                uz >= v.b.lo[2] && uz <= v.b.hi[2])
              nb_dir = (uint32_t)((int)tile + jox + joy * v.gx + joz * v.gx * v.gy);
          }
          nst = nb_dir != kNoTile ? ld_sc(st_m + nb_dir) : 0u;
          // face words of x-side neighbour j toward this tile: the outermost
          // column, and for the neighbours at (y, z) (lanes 12, 14) the second
          uint64_t fwo = 0, fwb = 0, fw2o = 0, fw2b = 0;
          const bool x_only = joy == 0 && joz == 0;
          if (nb_dir != kNoTile && jox != 0) {
            const uint64_t o = kXf * (uint64_t)nb_dir + (jox < 0 ? 1 : 0);
            fwo = ldg(xfo + o);
            fwb = ld_sc(xfr + o);
            if (x_only) {
              fw2o = ldg(xfo + o + 2);
              fw2b = ld_sc(xfr + o + 2);
            }
          }
          // rows beyond the tile in y and z, depth 1 and 2, and the yz diagonal
          const int oy = yl == 0 ? -1 : yl == TY - 1 ? 1 : 0;
          const int oz = zl == 0 ? -1 : zl == TZ - 1 ? 1 : 0;
          auto nb_of = [&](int oyy, int ozz) -> uint32_t {
            const int uy = ty + oyy, uz = tz + ozz;
            if (uy < v.b.lo[1] || uy > v.b.hi[1] || uz < v.b.lo[2] || uz > v.b.hi[2]) return kNoTile;
            return (uint32_t)((int)tile + oyy * v.gx + ozz * v.gx * v.gy);
          };
          const uint32_t nby = (rowok && oy) ? nb_of(oy, 0) : kNoTile;
          const uint32_t nbz = (rowok && oz) ? nb_of(0, oz) : kNoTile;
          const uint32_t nbd = (rowok && oy && oz) ? nb_of(oy, oz) : kNoTile;
          // a neighbour above is a tail tile: its second layer may not exist
          const bool y2 = nby != kNoTile && (oy < 0 || (ty + 1) * TY + 1 < v.ny);
          const bool z2 = nbz != kNoTile && (oz < 0 || (tz + 1) * TZ + 1 < v.nz);
          const uint64_t ry1 = (uint64_t)nby * R + (oy < 0 ? TY - 1 : 0) + TY * zl;
          const uint64_t ry2 = (uint64_t)nby * R + (oy < 0 ? TY - 2 : 1) + TY * zl;
          const uint64_t rz1 = (uint64_t)nbz * R + yl + TY * (oz < 0 ? TZ - 1 : 0);
          const uint64_t rz2 = (uint64_t)nbz * R + yl + TY * (oz < 0 ? TZ - 2 : 1);
          const uint64_t rdd = (uint64_t)nbd * R + (oy < 0 ? TY - 1 : 0) + TY * (oz < 0 ? TZ - 1 : 0);
          Row y1o = 0, y1b = 0, y2o = 0, y2b = 0, z1o = 0, z1b = 0, z2o = 0, z2b = 0, edo = 0, edb = 0;
          if (nby != kNoTile) y1o = ldg(orig_m + ry1), y1b = ld_sc(bufr + ry1);
          if (y2) y2o = ldg(orig_m + ry2), y2b = ld_sc(bufr + ry2);
          if (nbz != kNoTile) z1o = ldg(orig_m + rz1), z1b = ld_sc(bufr + rz1);
          if (z2) z2o = ldg(orig_m + rz2), z2b = ld_sc(bufr + rz2);
          if (nbd != kNoTile) edo = ldg(orig_m + rdd), edb = ld_sc(bufr + rdd);
          // -- states -> the candidate that holds iteration tp's input
          const Row vfull = valid_bits<Row>(tx, v.nx);
          auto pick = [&](uint32_t nb, uint32_t st, Row o, Row bsel, Row fullv) -> Row {
            if (nb == kNoTile) return 0;
            if (st & kFull) return fullv;
            if (st != 0 && st - 1 < tp) return bsel;
            return o;
          };
          const uint32_t sty = shfl32(nst, 13 + 3 * oy), stz = shfl32(nst, 13 + 9 * oz);
          const uint32_t std_ = shfl32(nst, 13 + 3 * oy + 9 * oz);
          const Row ey1 = pick(nby, sty, y1o, y1b, vfull);
          const Row ey2 = y2 ? pick(nby, sty, y2o, y2b, vfull) : Row(0);
          const Row ez1 = pick(nbz, stz, z1o, z1b, vfull);
          const Row ez2 = z2 ? pick(nbz, stz, z2o, z2b, vfull) : Row(0);
          const Row ed = pick(nbd, std_, edo, edb, vfull);
          // face words; a full tile's face has a bit for every row it has, and
          // so has its second column unless it is a +x tail tile one column wide
          uint64_t fw = 0, fw2 = 0;
          if (nb_dir != kNoTile && jox != 0) {
            if (nst & kFull) {
              const int yy = v.ny - (ty + joy) * TY, zz = v.nz - (tz + joz) * TZ;
              const uint64_t ym = yy >= TY ? 0xFFull : ((1ull << yy) - 1);
              uint64_t rows = 0;
#pragma unroll
              for (int k = 0; k < TZ; ++k) rows |= k < zz ? ym << (TY * k) : 0ull;
              fw = rows;
              fw2 = (x_only && (jox < 0 || v.nx - (tx + 1) * TXB >= 2)) ? rows : 0ull;
            } else {
              const bool from_buf = nst != 0 && nst - 1 < tp;
              fw = from_buf ? fwb : fwo;
              fw2 = from_buf ? fw2b : fw2o;
            }
          }
          // -- the rows around (y, z): shuffles inside the tile, loaded rows
          // beyond; each is folded in as it comes, so few stay live.  f1: the
          // rows y+-1, z+-1; f2: rows y+-2, z+-2 and the four (y+-1, z+-1)
          Row f1, f2;
          {
            // y neighbours (lane +- 1, +- 2) through row DPP: each is used
            // only where its source lane is in the same y group
            Row t = row_dpp64<kFromLeft1>(own);
            f1 = yl > 0 ? t : ey1;
            t = row_dpp64<kFromRight1>(own);
            f1 |= yl < TY - 1 ? t : ey1;
            t = shfl_row<Row>(own, lane - TY);
            f1 |= zl > 0 ? t : ez1;
            t = shfl_row<Row>(own, lane + TY);
            f1 |= zl < TZ - 1 ? t : ez1;
            Row e = row_dpp64<kFromLeft1>(ey1);
            t = row_dpp64<kFromLeft2>(own);
            f2 = yl > 1 ? t : yl == 1 ? e : ey2;
            e = row_dpp64<kFromRight1>(ey1);
            t = row_dpp64<kFromRight2>(own);
            f2 |= yl < TY - 2 ? t : yl == TY - 2 ? e : ey2;
            e = shfl_row<Row>(ez1, lane - TY);
            t = shfl_row<Row>(own, lane - 2 * TY);
            f2 |= zl > 1 ? t : zl == 1 ? e : ez2;
            e = shfl_row<Row>(ez1, lane + TY);
            t = shfl_row<Row>(own, lane + 2 * TY);
            f2 |= zl < TZ - 2 ? t : zl == TZ - 2 ? e : ez2;
            // (y-1, z-1), (y+1, z-1), (y-1, z+1), (y+1, z+1): inside the tile,
            // or an edge lane's neighbour in ey1 / ez1, or the corner's ed
            const Row eym8 = shfl_row<Row>(ey1, lane - TY), eyp8 = shfl_row<Row>(ey1, lane + TY);
            const Row ezm1 = row_dpp64<kFromLeft1>(ez1), ezp1 = row_dpp64<kFromRight1>(ez1);
            t = shfl_row<Row>(own, lane - TY - 1);
            f2 |= yl > 0 ? (zl > 0 ? t : ezm1) : (zl > 0 ? eym8 : ed);
            t = shfl_row<Row>(own, lane - TY + 1);
            f2 |= yl < TY - 1 ? (zl > 0 ? t : ezp1) : (zl > 0 ? eym8 : ed);
            t = shfl_row<Row>(own, lane + TY - 1);
            f2 |= yl > 0 ? (zl < TZ - 1 ? t : ezm1) : (zl < TZ - 1 ? eyp8 : ed);
            t = shfl_row<Row>(own, lane + TY + 1);
            f2 |= yl < TY - 1 ? (zl < TZ - 1 ? t : ezp1) : (zl < TZ - 1 ? eyp8 : ed);
          }
          // -- x halo of the rows y+-1, z+-1: the x-side face planes (lanes
          // 12, 9, 15, 3, 21 at x-1; 14, 11, 17, 5, 23 at x+1), shifted by one
          // in y or z
          auto lane_word = [&](uint64_t w, int jj) -> uint64_t {
            return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)w, jj)) |
                   ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(w >> 32), jj)
                    << 32);
          };
          auto plane_of = [&](int jj) -> uint64_t { return lane_word(fw, jj); };
          auto cross = [&](int side) -> uint64_t {  // side 0: x-1, 2: x+1
            constexpr uint64_t kY0 = 0x0101010101010101ull, kY7 = 0x8080808080808080ull;
            const uint64_t c = plane_of(side + 12), ym = plane_of(side + 9), yp = plane_of(side + 15);
            const uint64_t zm = plane_of(side + 3), zp = plane_of(side + 21);
            return ((c >> 1) & ~kY7) | ((yp & kY0) << 7) | ((c << 1) & ~kY0) | ((ym & kY7) >> 7) |
                   (c >> 8) | (zp << 56) | (c << 8) | (zm >> 56);
          };
          const uint64_t clo = cross(0), chi = cross(2);
          // the row's x halo: columns 63 (bl1) and 62 (bl2) of the x-1
          // neighbour, 0 (bh1) and 1 (bh2) of the x+1 neighbour
          const Row bl1 = (Row)((plane_of(12) >> r) & 1), bh1 = (Row)((plane_of(14) >> r) & 1);
          const Row bl2 = (Row)((lane_word(fw2, 12) >> r) & 1);
          const Row bh2 = (Row)((lane_word(fw2, 14) >> r) & 1);
          const Row x1own = own | (Row)(own << 1) | (Row)(own >> 1) | bl1 | (Row)(bh1 << 63);
          const Row x2own = x1own | (Row)(own << 2) | (Row)(own >> 2) | bl2 | (Row)(bl1 << 1) |
                            (Row)(bh2 << 63) | (Row)(bh1 << 62);
          const Row x1f = f1 | (Row)(f1 << 1) | (Row)(f1 >> 1) | (Row)((clo >> r) & 1) |
                          (Row)(((chi >> r) & 1) << 63);
          d1 = (x1own | f1) & vb;
          dl = (x2own | x1f | f2) & vb;
          if (!rowok) d1 = 0, dl = 0;
        } else if constexpr (CONN == 6) {
          // neighbour tiles: lanes r < 6 also read the state of neighbour r
          nb_dir = (live && r < 6) ? neighbour(v, v.b, tile, tx, ty, tz, r) : kNoTile;
          const uint32_t nb_xl = live ? neighbour(v, v.b, tile, tx, ty, tz, 0) : kNoTile;
          const uint32_t nb_xh = live ? neighbour(v, v.b, tile, tx, ty, tz, 1) : kNoTile;
          const int on_face[4] = {yl == 0, yl == TY - 1, zl == 0, zl == TZ - 1};
          const int src_row[4] = {(TY - 1) + TY * zl, TY * zl, yl + TY * (TZ - 1), yl};
          uint32_t nb_yz[4];
  #pragma unroll
          for (int f = 0; f < 4; ++f)
            nb_yz[f] = (rowok && on_face[f]) ? neighbour(v, v.b, tile, tx, ty, tz, 2 + f) : kNoTile;
          nst = nb_dir != kNoTile ? ld_sc(st_m + nb_dir) : 0u;
          uint64_t fxl_o = 0, fxl_b = 0, fxh_o = 0, fxh_b = 0;
          if (nb_xl != kNoTile) {
            fxl_o = ldg(xfo + kXf * (uint64_t)nb_xl + 1);
            fxl_b = ld_sc(xfr + kXf * (uint64_t)nb_xl + 1);
          }
          if (nb_xh != kNoTile) {
            fxh_o = ldg(xfo + kXf * (uint64_t)nb_xh);
            fxh_b = ld_sc(xfr + kXf * (uint64_t)nb_xh);
          }
          // y / z face rows from the neighbour tiles, both candidate sources
          Row fo[4] = {0, 0, 0, 0}, fb[4] = {0, 0, 0, 0};
  #pragma unroll
          for (int f = 0; f < 4; ++f) {
            if (nb_yz[f] == kNoTile) continue;
            const uint64_t o = (uint64_t)nb_yz[f] * R + src_row[f];
            fo[f] = ldg(orig_m + o);
            fb[f] = ld_sc(bufr + o);
          }
          // -- neighbour states -> which candidate holds iteration tp's input
          const uint32_t st_xl = shfl32(nst, gbase + 0), st_xh = shfl32(nst, gbase + 1);
          const uint32_t st_f[4] = {shfl32(nst, gbase + 2), shfl32(nst, gbase + 3),
                                    shfl32(nst, gbase + 4), shfl32(nst, gbase + 5)};
          auto pick64 = [&](uint32_t nb, uint32_t st, uint64_t o, uint64_t bsel) -> uint64_t {
            if (nb == kNoTile) return 0;
            if (st & kFull) return ~0ull;
            if (st != 0 && st - 1 < tp) return bsel;
            return o;
          };
          const uint64_t wxl = pick64(nb_xl, st_xl, fxl_o, fxl_b);
          const uint64_t wxh = pick64(nb_xh, st_xh, fxh_o, fxh_b);
          Row face[4];
  #pragma unroll
          for (int f = 0; f < 4; ++f) {
            const uint32_t nb = nb_yz[f];
            const uint32_t st = st_f[f];
            face[f] = nb == kNoTile ? Row(0)
                      : (st & kFull) ? ~Row(0)
                      : (st != 0 && st - 1 < tp) ? fb[f] : fo[f];
          }
          // -- 6-connectivity dilation of the row (bitDilatate + faces,
          //    hello.cu:373-376, 1310-1542)
          const Row up = shfl_row<Row>(own, lane - 1);
          const Row dn = shfl_row<Row>(own, lane + 1);
          const Row bk = shfl_row<Row>(own, lane - TY);
          const Row fw = shfl_row<Row>(own, lane + TY);
          dl = own | (Row)(own << 1) | (Row)(own >> 1);
          dl |= (Row)((wxl >> r) & 1);
          dl |= (Row)((wxh >> r) & 1) << (TXB - 1);
          dl |= yl > 0 ? up : face[0];
          dl |= yl < TY - 1 ? dn : face[1];
          dl |= zl > 0 ? bk : face[2];
          dl |= zl < TZ - 1 ? fw : face[3];
          dl &= vb;
          if (!rowok) dl = 0;
        } else {
          // 18 / 26 elements (one 64x8x8 tile per wave, lane = row): lane j
          // < 27 reads the state of neighbour j = (ox+1) + 3(oy+1) + 9(oz+1)
          // and, when ox != 0, both candidates of that tile's face word
          // toward this one.  A row's (y, z) neighbours are other lanes'
          // rows (shuffles) or, on the tile's faces, rows of the y, z and
          // yz-diagonal neighbour tiles (ey, ez, ed: at most one of each
          // per lane); the x neighbours across the tile's x faces come from
          // the face words, dilated in (y, z) as 8x8 bit planes.
          static_assert(CONN == 6 || (TXB == 64 && TY == 8 && TZ == 8), "64x8x8 tiles only");
          const int j = lane;
          const int jox = j % 3 - 1, joy = (j / 3) % 3 - 1, joz = j / 9 - 1;
          nb_dir = kNoTile;
          if (live && j < 27 && j != 13) {
            const int ux = tx + jox, uy = ty + joy, uz = tz + joz;
            if (ux >= v.b.lo[0] && ux <= v.b.hi[0] && uy >= v.b.lo[1] && uy <= v.b.hi[1] &&
                uz >= v.b.lo[2] && uz <= v.b.hi[2])
              nb_dir = (uint32_t)((int)tile + jox + joy * v.gx + joz * v.gx * v.gy);
          }
          nst = nb_dir != kNoTile ? ld_sc(st_m + nb_dir) : 0u;
          uint64_t fwo = 0, fwb = 0;  // face word of neighbour j toward this tile
          if (nb_dir != kNoTile && jox != 0) {
            const uint64_t o = kXf * (uint64_t)nb_dir + (jox < 0 ? 1 : 0);
            fwo = ldg(xfo + o);
            fwb = ld_sc(xfr + o);
          }
          // rows beyond the tile: ey (y face), ez (z face), ed (yz corner)
          const int oy = yl == 0 ? -1 : yl == TY - 1 ? 1 : 0;
          const int oz = zl == 0 ? -1 : zl == TZ - 1 ? 1 : 0;
          const int jy = 13 + 3 * oy, jz = 13 + 9 * oz, jd = 13 + 3 * oy + 9 * oz;
          auto nb_of = [&](int oyy, int ozz) -> uint32_t {
            const int uy = ty + oyy, uz = tz + ozz;
            if (uy < v.b.lo[1] || uy > v.b.hi[1] || uz < v.b.lo[2] || uz > v.b.hi[2]) return kNoTile;
            return (uint32_t)((int)tile + oyy * v.gx + ozz * v.gx * v.gy);
          };
          const uint32_t nby = (rowok && oy) ? nb_of(oy, 0) : kNoTile;
          const uint32_t nbz = (rowok && oz) ? nb_of(0, oz) : kNoTile;
          const uint32_t nbd = (rowok && oy && oz) ? nb_of(oy, oz) : kNoTile;
          const int ry = (oy < 0 ? TY - 1 : 0) + TY * zl;     // the row in the y neighbour
          const int rz = yl + TY * (oz < 0 ? TZ - 1 : 0);     // in the z neighbour
          const int rd = (oy < 0 ? TY - 1 : 0) + TY * (oz < 0 ? TZ - 1 : 0);
          Row eyo = 0, eyb = 0, ezo = 0, ezb = 0, edo = 0, edb = 0;
          if (nby != kNoTile) {
            eyo = ldg(orig_m + (uint64_t)nby * R + ry);
            eyb = ld_sc(bufr + (uint64_t)nby * R + ry);
          }
          if (nbz != kNoTile) {
            ezo = ldg(orig_m + (uint64_t)nbz * R + rz);
            ezb = ld_sc(bufr + (uint64_t)nbz * R + rz);
          }
          if (nbd != kNoTile) {
            edo = ldg(orig_m + (uint64_t)nbd * R + rd);
            edb = ld_sc(bufr + (uint64_t)nbd * R + rd);
          }
          // -- states -> the candidate that holds iteration tp's input; a
          // full tile holds every valid bit of its rows
          const Row vfull = valid_bits<Row>(tx, v.nx);
          auto pick = [&](uint32_t nb, uint32_t st, Row o, Row bsel) -> Row {
            if (nb == kNoTile) return 0;
            if (st & kFull) return vfull;
            if (st != 0 && st - 1 < tp) return bsel;
            return o;
          };
          const Row ey = pick(nby, shfl32(nst, jy), eyo, eyb);
          const Row ez = pick(nbz, shfl32(nst, jz), ezo, ezb);
          const Row ed = pick(nbd, shfl32(nst, jd), edo, edb);
          // face word: a full tile's face has a bit for every row it has
          uint64_t fw = 0;
          if (nb_dir != kNoTile && jox != 0) {
            if (nst & kFull) {
              const int yy = v.ny - (ty + joy) * TY, zz = v.nz - (tz + joz) * TZ;
              const uint64_t ym = yy >= TY ? 0xFFull : ((1ull << yy) - 1);
              uint64_t rows = 0;
#pragma unroll
              for (int k = 0; k < TZ; ++k) rows |= k < zz ? ym << (TY * k) : 0ull;
              fw = rows;
            } else {
              fw = (nst != 0 && nst - 1 < tp) ? fwb : fwo;
            }
          }
          // -- (y, z) neighbourhood of the row
          const Row up = shfl_row<Row>(own, lane - 1), dn = shfl_row<Row>(own, lane + 1);
          const Row ysh = (yl > 0 ? up : ey) | (yl < TY - 1 ? dn : ey);  // rows y-1, y+1
          const Row zup = shfl_row<Row>(ez, lane - 1), zdn = shfl_row<Row>(ez, lane + 1);
          const Row zsh = (yl > 0 ? zup : ed) | (yl < TY - 1 ? zdn : ed);  // beyond z: y-1, y+1
          const Row yc = own | ysh;  // rows y-1 .. y+1 at z
          Row a_row, s_row;          // a_row: rows whose voxels also spread in x; s_row: x only
          if constexpr (CONN == 26) {
            const Row ycz = ez | zsh;
            const Row bk = shfl_row<Row>(yc, lane - TY), fw2 = shfl_row<Row>(yc, lane + TY);
            a_row = yc | (zl > 0 ? bk : ycz) | (zl < TZ - 1 ? fw2 : ycz);
            s_row = a_row;
          } else {  // 18: faces, xy diagonals, corners
            const Row bk = shfl_row<Row>(own, lane - TY), fw2 = shfl_row<Row>(own, lane + TY);
            const Row sbk = shfl_row<Row>(ysh, lane - TY), sfw = shfl_row<Row>(ysh, lane + TY);
            const Row zrow = (zl > 0 ? bk : ez) | (zl < TZ - 1 ? fw2 : ez);      // (0, +-1)
            const Row corner = (zl > 0 ? sbk : zsh) | (zl < TZ - 1 ? sfw : zsh);  // (+-1, +-1)
            a_row = yc | zrow;
            s_row = yc | corner;
          }
          // -- the x halo: the face planes of the 9 tiles at x-1 (lanes j =
          // 3k) and x+1 (j = 3k+2), k = (oy+1) + 3(oz+1), dilated in (y, z)
          // like the rows (bit yl + 8 zl of a plane is row (yl, zl))
          auto plane_of = [&](int jj) -> uint64_t {
            return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)fw, jj)) |
                   ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(fw >> 32), jj)
                    << 32);
          };
          auto halo = [&](int side) -> uint64_t {  // side 0: x-1, 2: x+1
            constexpr uint64_t kY0 = 0x0101010101010101ull, kY7 = 0x8080808080808080ull;
            uint64_t y3[3], ys[3];
#pragma unroll
            for (int kz = 0; kz < 3; ++kz) {
              const uint64_t c = plane_of(side + 3 * (1 + 3 * kz));
              const uint64_t m = plane_of(side + 3 * (0 + 3 * kz));
              const uint64_t pp = plane_of(side + 3 * (2 + 3 * kz));
              ys[kz] = ((c >> 1) & ~kY7) | ((pp & kY0) << 7) | ((c << 1) & ~kY0) | ((m & kY7) >> 7);
              y3[kz] = c | ys[kz];
            }
            const uint64_t *zs = CONN == 26 ? y3 : ys;
            return y3[1] | (zs[1] >> 8) | (zs[2] << 56) | (zs[1] << 8) | (zs[0] >> 56);
          };
          const uint64_t hlo = halo(0), hhi = halo(2);
          dl = a_row | (Row)(s_row << 1) | (Row)(s_row >> 1);
          dl |= (Row)((hlo >> r) & 1);
          dl |= (Row)((hhi >> r) & 1) << (TXB - 1);
          dl &= vb;
          if (!rowok) dl = 0;
        }
        // -- input full -> the tile leaves the queue (hello.cu:1319, 1634-1643)
        const bool full = group_bits<R>(__ballot(live && exists && own != vb), lane) == 0;
        // -- newly covered voxels of the other mask: res & ~src & ref (hello.cu:1590-1626)
        //    (DS: d1 holds the first step, hits at 2tp+1; dl the second, 2tp+2)
        const Row nw = DS ? (Row)(d1 & ~own & other) : (Row)(dl & ~own & other);
        const Row nw2 = DS ? (Row)(dl & ~d1 & other) : Row(0);
        const uint32_t cnt = group_sum<R>((uint32_t)popc_row(nw), lane);
        const uint32_t cnt2 = DS ? group_sum<R>((uint32_t)popc_row(nw2), lane) : 0u;
        const uint32_t rem = group_sum<R>((uint32_t)popc_row((Row)(other & ~dl)), lane);
        if ((nw | nw2) && v.map) {
          const uint64_t vrow =
              (uint64_t)tx * TXB + (uint64_t)v.nx * ((uint64_t)y + (uint64_t)v.ny * z);
#pragma unroll
          for (int k = 0; k < (DS ? 2 : 1); ++k) {
            Row bits = k ? nw2 : nw;
            const uint16_t it = (uint16_t)(DS ? 2 * tp + 1 + k : tp + 1);
            while (bits) {
              const int kb = (int)__builtin_ctzll((uint64_t)bits);
              bits &= bits - 1;
              v.map[vrow + kb] = it;
            }
          }
        }
        if (live && r == 0 && cnt) {
          if (SINGLE)
            atomicAdd(&s_cnt[m], cnt);
          else
            hit_add(ht, ctrs, (uint32_t)p * 4 + m, cnt);
        }
        if (DS && live && r == 0 && cnt2) {
          if (SINGLE)
            atomicAdd(&s_cnt[4 + m], cnt2);
          else
            hit_add(ht, ctrs, (uint32_t)p * 4 + 2 + m, cnt2);
        }
        // -- outputs
        const uint64_t oxl = group_bits<R>(__ballot(dl & 1), lane);
        const uint64_t oxh = group_bits<R>(__ballot((dl >> (TXB - 1)) & 1), lane);
        const uint64_t oany = group_bits<R>(__ballot(dl != 0), lane);
        // DS: the second column from each x face, and either of the two outermost
        const uint64_t o2l = DS ? group_bits<R>(__ballot((dl >> 1) & 1), lane) : 0ull;
        const uint64_t o2h = DS ? group_bits<R>(__ballot((dl >> (TXB - 2)) & 1), lane) : 0ull;
        const uint64_t ox2l = oxl | o2l, ox2h = oxh | o2h;
        const bool oyl = group_bits<R>(__ballot(yl == 0 && dl != 0), lane) != 0;
        const bool oyh = group_bits<R>(__ballot(yl == TY - 1 && dl != 0), lane) != 0;
        const bool ozl = group_bits<R>(__ballot(zl == 0 && dl != 0), lane) != 0;
        const bool ozh = group_bits<R>(__ballot(zl == TZ - 1 && dl != 0), lane) != 0;
        bool want = false;
        uint64_t pent = 0;
        uint32_t new_old = 0;  // this step's claim
        uint64_t new_ent = 0;
        if (live) {
          if (full) {
            if (r == 0) or_g(st_m + tile, kFull);
          } else {
            st_sc(bufw + ro, dl);
            if (r == 0) {
              st_sc(xfw + kXf * (uint64_t)tile, oxl);
              st_sc(xfw + kXf * (uint64_t)tile + 1, oxh);
              if constexpr (DS) {
                st_sc(xfw + kXf * (uint64_t)tile + 2, o2l);
                st_sc(xfw + kXf * (uint64_t)tile + 3, o2h);
              }
            }
            // lane kSelf re-queues the tile while it is not full; lanes
            // below it activate idle neighbours the output reaches
            constexpr int kSelf = (CONN == 6 && !DS) ? 6 : 27;
            if (r == kSelf) {
              want = true;
              pent = make_entry(tile, (uint32_t)p, (uint32_t)m, 0, rem != 0, v.gen);
            }
            if (r < kSelf && nb_dir != kNoTile && nst == 0) {
              bool reach;
              if constexpr (DS) {  // within two steps: two layers of a face, or an edge line
                const int jox = r % 3 - 1, joy = (r / 3) % 3 - 1, joz = r / 9 - 1;
                const bool face = (jox != 0) + (joy != 0) + (joz != 0) == 1;
                const uint64_t xs = jox < 0 ? (face ? ox2l : oxl) : jox > 0 ? (face ? ox2h : oxh) : oany;
                const uint64_t ym = joy < 0 ? (face ? 0x0303030303030303ull : 0x0101010101010101ull)
                                    : joy > 0 ? (face ? 0xC0C0C0C0C0C0C0C0ull : 0x8080808080808080ull)
                                              : ~0ull;
                const uint64_t zm = joz < 0 ? (face ? 0xFFFFull : 0xFFull)
                                    : joz > 0 ? (face ? 0xFFFFull << 48 : 0xFFull << 56) : ~0ull;
                reach = (xs & ym & zm) != 0;
              } else if constexpr (CONN == 6) {
                reach = r == 0 ? oxl != 0 : r == 1 ? oxh != 0 : r == 2 ? oyl
                        : r == 3 ? oyh : r == 4 ? ozl : ozh;
              } else {  // a voxel on every face that the offset of neighbour r crosses
                const int jox = r % 3 - 1, joy = (r / 3) % 3 - 1, joz = r / 9 - 1;
                const uint64_t xs = jox < 0 ? oxl : jox > 0 ? oxh : oany;
                const uint64_t ym = joy < 0 ? 0x0101010101010101ull
                                    : joy > 0 ? 0x8080808080808080ull : ~0ull;
                const uint64_t zm = joz < 0 ? 0xFFull : joz > 0 ? 0xFFull << 56 : ~0ull;
                reach = (xs & ym & zm) != 0;
              }
              if (reach) {
                new_old = cas_g(st_m + nb_dir, 0u, tp + 2);
                new_ent = make_entry(nb_dir, (uint32_t)p, (uint32_t)m, 1, 1, v.gen);
              }
            }
          }
        }
        // lane kSelf's own entry, lanes below it the claims won by the step before
        const bool won = claim_ent != 0 && claim_old == 0u;
        if constexpr (SINGLE)
          push_sticky(want || won, want ? pent : claim_ent, s_st[(t + 1) & 1],
                      &s_stn[(t + 1) & 1], kStick, lq, Qout, qout_len, lane, &s_cnt[2]);
        else
          push_entries(want || won, want ? pent : claim_ent, lq, Qout, qout_len, lane);
        claim_old = new_old;
        claim_ent = new_ent;
    };
    // global entry ei of the queue of iteration t
    // lane k holds the start of sub-queue k + 1 of this iteration's queue, read
    // once per iteration (sub-queue 0 starts at 0)
    const uint32_t pre_l = s_pre[1 + (lane & (kSubQ - 1))];
    auto global_entry = [&](uint32_t ei) {
      int sq = 0;
      uint32_t start = 0;
      if constexpr (TPW == 1) {  // ei is wave-uniform: lane k < kSubQ - 1 compares start k + 1
        sq = __popcll(__ballot(lane < kSubQ - 1 && pre_l <= ei));
        start = sq ? (uint32_t)__builtin_amdgcn_readlane((int)pre_l, sq - 1) : 0u;
      } else {
        while (sq + 1 < kSubQ && s_pre[sq + 1] <= ei) ++sq;
        start = s_pre[sq];
      }
      return ld_sc(sub_queue(q, qcap, t & 1, sq) + (ei - start));
    };
    uint32_t gfirst = 0;  // global entries below this were taken in the first pass
    if constexpr (SINGLE) {
      // first pass over the wave slots: slot s takes kept entry s, or else global entry
      // s_fs[0] + (s - kept) of this workgroup's free range
      const uint32_t nloc = s_stn[t & 1];
      const uint32_t s = (uint32_t)wave * TPW + sub;
      const uint32_t gi = s_fs[0] + (s - nloc);
      const bool glive = s >= nloc && gi < L && gi < s_fs[1];
      uint64_t ent = 0;
      if (s < nloc) ent = s_st[t & 1][s];
      if (glive) ent = global_entry(gi);
      if (__ballot(s < nloc || glive)) step_tile(ent, s < nloc || glive);
      gfirst = min(L, s_fs[1]);
    }
    // global entries beyond the free slots (one pair) or all of them
    // (several pairs), round robin over the waves: chunks of step entries
    // at base, base + nwave * step, ...
    if constexpr (SINGLE) {
      for (uint32_t base = gfirst + gwave * step; base < L; base += nwave * step) {
        for (uint32_t k = 0; k < step && base + k < L; k += TPW) {
          const uint32_t ei = base + k + sub;
          const bool live = ei < L;
          step_tile(live ? global_entry(ei) : 0ull, live);
        }
      }
      // the claims of the wave's last step
      push_sticky(claim_ent != 0 && claim_old == 0u, claim_ent, s_st[(t + 1) & 1],
                  &s_stn[(t + 1) & 1], kStick, lq, Qout, qout_len, lane, &s_cnt[2]);
      claim_ent = 0;
    } else {  // the same order, each entry loaded while the one before steps
      uint32_t base = gfirst + gwave * step, k = 0;
      uint64_t nxt = (base < L && base + sub < L) ? global_entry(base + sub) : 0ull;
      while (base < L) {
        const bool live = base + k + sub < L;
        const uint64_t cur = nxt;
        k += TPW;
        if (k >= step || base + k >= L) base += nwave * step, k = 0;
        if (base < L) {
          const uint32_t e2 = base + k + sub;
          nxt = e2 < L ? global_entry(e2) : 0ull;
        }
        step_tile(cur, live);
      }
      // the claims of the wave's last step
      push_entries(claim_ent != 0 && claim_old == 0u, claim_ent, lq, Qout, qout_len, lane);
      claim_ent = 0;
    }
    if (tr) {
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint64_t now = wall_clock64();
        atomicMax((unsigned long long *)&tr[1], (unsigned long long)now);
        if (blockIdx.x == 0) tr[2] = now;
      }
    }
    if constexpr (SINGLE) {  // staged overflow to the global queue; counts for the barrier
      const uint32_t nflushed = flush_local_queue(lq, Qout, qout_len);
      if (threadIdx.x == 0) s_cnt[2] += nflushed, s_cnt[3] = s_stn[(t + 1) & 1];
    } else {
      flush_local_queue(lq, Qout, qout_len);
      hit_flush(ht, ctrs);
    }
    const uint32_t tdone = t;  // iteration that just ran
    ++t;
    // master bookkeeping between iterations: fold the spread hit counters into
    // the histogram and the covered counts, decide each pass's stop
    // (hello.cu:1748-1759), reset the sub-queue counters that iteration t+1 fills
    auto fold = [&]() {
      // per lanes per (pair, mask): each sums kHitSlots / per cumulative hit
      // slots, the group sums the lanes; the increase over the last fold is
      // this iteration's histogram bin.  per shrinks with the pair count, so
      // up to 64 pairs fold in one pass, and every load of an item (its pair's
      // start, stop state and counters, and the hit slots) is issued at once:
      // one round trip per pass.  Two steps per iteration: the counters of each
      // step give the bins 2tp+1 and 2tp+2 (cov holds both steps, cov2 the
      // second's part), folded in order up to the one that reaches the cut.
      constexpr int kSteps = DS ? 2 : 1;
      int per = kHitSlots;
      while (per > 2 && per * 2 * npairs > kWgThreads) per >>= 1;
      const int s = threadIdx.x % per;
      for (int base = 0; base < 2 * npairs; base += kWgThreads / per) {
        const int i = base + threadIdx.x / per;
        const int pp = i >> 1, mm = i & 1;
        const bool in = i < 2 * npairs;
        unsigned long long v = 0, v2 = 0, cov = 0, c2 = 0, sum = 0, nm = 0;
        uint32_t ts = 0, hd = 0;
        int df0 = 0, nd = 0;
        if (in) {
          for (int k = s; k < kHitSlots; k += per) {
            v += ld_sc(&ctrs[pp].hits[mm][k]);
            if constexpr (DS) v2 += ld_sc(&ctrs[pp].hits[2 + mm][k]);
          }
          if (s == 0) {
            ts = descs[pp].t_start;
            nd = ctrs[pp].near_done;  // written before this launch
            df0 = ld_sc(&ctrs[pp].done_from[mm]);
            cov = ld_sc(&ctrs[pp].cov[mm]);
            sum = ld_sc(&ctrs[pp].sum_d[mm]);
            if constexpr (DS) {
              c2 = ld_sc(&ctrs[pp].cov2[mm]);
              hd = ld_sc(&ctrs[pp].hd_last[mm]);
            }
            nm = mm == 0 ? ctrs[pp].n_fp : ctrs[pp].n_fn;
          }
        }
        for (int off = per / 2; off >= 1; off >>= 1) {
          v += shfl64(v, (threadIdx.x & 63) ^ off);
          if constexpr (DS) v2 += shfl64(v2, (threadIdx.x & 63) ^ off);
        }
        if (!in || s != 0 || tdone < ts || nd) continue;  // finished by the near pass
        const int tpp = (int)(tdone - ts);
        if (df0 <= kSteps * tpp) continue;  // the pass stopped at an earlier bin
        PairCtr &cc = ctrs[pp];
        const PairDesc &dd = descs[pp];
        const unsigned long long cut = robust_cut(dd.robust, nm);
        if constexpr (DS) {
          const unsigned long long bin[2] = {v - (cov - c2), v2 - c2};
          if (bin[0] == 0 && bin[1] == 0) continue;
          int df = 0;
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            const int it = 2 * tpp + 1 + k;
            if (bin[k] == 0 || df) continue;
            if (it < dd.hist_len) dd.hist[mm][it] = (uint32_t)bin[k];
            const unsigned long long before = cov;
            cov += bin[k];
            sum += bin[k] * (unsigned long long)it;
            hd = (uint32_t)it;
            if (before < cut && cov >= cut) df = it;
          }
          st_sc(&cc.cov[mm], cov);
          st_sc(&cc.cov2[mm], v2);
          st_sc(&cc.sum_d[mm], sum);
          st_sc(&cc.hd_last[mm], hd);
          if (df) st_sc(&cc.done_from[mm], df);
        } else {
          const unsigned long long before = cov;
          if (v == before) continue;
          if (tpp + 1 < dd.hist_len) dd.hist[mm][tpp + 1] = (uint32_t)(v - before);
          st_sc(&cc.cov[mm], v);
          st_sc(&cc.sum_d[mm], sum + (v - before) * (unsigned long long)(tpp + 1));
          st_sc(&cc.hd_last[mm], (uint32_t)(tpp + 1));
          if (before < cut && v >= cut) st_sc(&cc.done_from[mm], tpp + 1);
        }
      }
      if (threadIdx.x < kSubQ) st_sc(&es->qlen[(t + 2) % 4][threadIdx.x], 0u);
      if (threadIdx.x == 0) {
        es->qtotal += L;  // several pairs: the fold reads pair state from memory anyway
        es->qmax = max(es->qmax, L);
        if (qstats && (int)tdone < qstats_len) qstats[tdone] = L;
      }
    };
    if (SINGLE) {
      if (!grid_barrier_single<DS ? 2 : 1>(es, eoff + t, spin_limit, &s_ok, s_cnt, &s_stn[tdone & 1], ps,
                               v0, tdone, qstats, qstats_len, tr, s_pre, s_flags, s_red, kSlots,
                               s_kept, s_fs, (t + 2) % 4))
        break;
    } else if (!grid_barrier(es, t, spin_limit, fence, &s_ok, fold, tr, es->qlen[t % 4], s_pre)) {
      break;
    }
  }
  if (SINGLE && blockIdx.x == 0 && threadIdx.x == 0) {  // the master's state, for finalize
    PairCtr &c0 = ctrs[0];
    for (int m = 0; m < 2; ++m) {
      st_sc(&c0.cov[m], ps.cov[m]);
      st_sc(&c0.sum_d[m], ps.sum[m]);
      st_sc(&c0.hd_last[m], ps.hd[m]);
      st_sc(&c0.done_from[m], ps.done_from[m]);
    }
  }
  // workgroup 0 holds every fold (it is the barrier master): it writes the
  // results of all pairs and the engine head
  if (blockIdx.x == 0) {
    const bool err = ld_sc(&es->error) != 0;
    if constexpr (FUSED) {
      if (threadIdx.x == 0) a.outs[0] = finalize_pair(a.desc, ctrs[0], t, err, DS ? 2 : 1);
    } else {
      for (int p = threadIdx.x; p < npairs; p += blockDim.x)
        a.outs[p] = finalize_pair(descs[p], ctrs[p], t, err, DS ? 2 : 1);
    }
    // with a ready marker, every thread's results reach the host before
    // thread 0 stores it
    if (a.ready) __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
      // entries left for a next iteration (one pair: the master's count,
      // which includes the workgroups' own queues)
      const uint32_t qn = SINGLE ? ps.next_total : s_pre[kSubQ];
      st_sc(&es->t, t);
      st_sc(&es->qnext, qn);
      if (SINGLE) st_sc(&es->qtotal, ps.qtotal), st_sc(&es->qmax, ps.qmax);
      if (a.head) {
        EngineHead h;
        h.t = t;
        h.error = err ? 1u : 0u;
        h.qmax = SINGLE ? ps.qmax : es->qmax;
        h.nwg = gridDim.x;
        h.qtotal = SINGLE ? ps.qtotal : es->qtotal;
        h.qcap = es->qcap;
        h.qnext = qn;
        h.phase[0] = FUSED ? s_phase[0] : 0ull;
        h.phase[1] = FUSED ? s_phase[1] : 0ull;
        h.phase[2] = FUSED ? s_phase[2] : 0ull;
        h.phase[3] = FUSED ? wall_clock64() : 0ull;
        *a.head = h;
      }
      if (a.ready) {  // the head too reaches the host before the marker
        __threadfence_system();
        __hip_atomic_store(a.ready, a.ready_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

// The multi-pair two-step kernel is held to four waves per SIMD (128 VGPRs,
// a few spilled): a 256-pair batch's dilation measured 4% faster than with
// three waves and no spill.
template <int TXB, int TY, int TZ, bool SINGLE, bool FUSED = false, int CONN = 6,
          bool DS = false>
__global__ __launch_bounds__(kWgThreads)
__attribute__((amdgpu_waves_per_eu(!SINGLE && DS ? 4 : 1))) void dilate_kernel(DilateArgs a) {
  dilate_body<TXB, TY, TZ, SINGLE, FUSED, CONN, DS>(a);
}

// The fused launch with two dilation steps per barrier: about 185 VGPRs, two
// resident workgroups per CU.  Held to three it spills in the step; with the
// barrier master's state moved to LDS it fits three without spilling, but
// every barrier period then grows with the 768 workgroups to poll and
// release, and it measured no faster (the long pair slower), so it runs
// with two.
__global__ __launch_bounds__(kWgThreads) void dilate_ds_kernel(DilateArgs a) {
  dilate_body<64, 8, 8, true, true, 6, true>(a);
}

// ---------------------------------------------------------------- regions
// The single-pair launch with eight dilation steps per grid barrier
// (dilate_tb_kernel; 64x8x8 tiles, 6 element).  A region is 1 x 2 x 2 tiles
// (64 x 16 x 16 voxels), region (rx, ry, rz) = tiles (rx, 2ry + {0,1},
// 2rz + {0,1}); the queue holds (region, mask) entries and one workgroup
// steps one entry at a time.  It holds the region's rows plus an eight-voxel
// halo -- a window of 32 x 32 rows in (y, z), each row 80 bits in x (eight
// columns of the x neighbours on each side) -- in registers, four rows of
// consecutive z per thread, and runs the eight 6-neighbour dilation steps
// there: x within the row, y from the lanes beside (DPP), z from the
// thread's own rows or, for its first and last row, from the neighbouring
// threads' rows through LDS, one workgroup barrier per step
// (hello.cu:373-376, 1310-1542 per step).  After step k the window is exact
// wherever the rows and columns it did not hold are more than k voxels away,
// so after eight steps the region itself is exact: the overlapped-tiling form
// of the reference's shared-memory block plus halo (hello.cu:1086, 1310-1542),
// taken eight steps deep.  New hits of the other mask in the region count per
// step (bins 8tp+1 .. 8tp+8, hello.cu:1590-1626); the barrier master folds
// them bin by bin.  Region states (rstate): 0 idle (data = orig), kFull, or
// 1 + the period at which the region was first queued (its data then sits in
// buf[(tp + 1) & 1] at period tp).  A region stays queued while its input is
// not full, and claims (compare-and-swap 0 -> tp + 2) every idle neighbour
// region whose core its output reaches within eight steps (L1 distance).
constexpr int kTbWin = 32;  // window rows per axis: 16 + 2 * 8

// A window row: o = the region's 64 voxels of the row; h = its x halo, bits
// 0..7 the voxels x = -8 .. -1 (the left tile's last eight columns), bits
// 16..23 the voxels x = 64 .. 71 (the right tile's first eight); bits 8..15
// stay empty, so a shift of h by one never carries one side into the other.
struct TbRow {
  uint64_t o;
  uint32_t h;
};
constexpr uint32_t kTbHalo = 0x00FF00FFu;
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));  // one LDS row
__device__ __forceinline__ TbRow tb_make(uint64_t left, uint64_t own, uint64_t right) {
  return TbRow{own, (uint32_t)(left >> 56) | ((uint32_t)(right & 0xFFull) << 16)};
}
// one 6-neighbour step of a row in x (bitDilatate, hello.cu:373-376), across
// the halo on both sides
__device__ __forceinline__ TbRow tb_xdil(const TbRow &r) {
  const uint64_t o = r.o | (r.o << 1) | (r.o >> 1) | (uint64_t)((r.h >> 7) & 1u) |
                     ((uint64_t)((r.h >> 16) & 1u) << 63);
  const uint32_t h = ((r.h | (r.h << 1) | (r.h >> 1)) & kTbHalo) | ((uint32_t)(r.o & 1u) << 7) |
                     ((uint32_t)(r.o >> 63) << 16);
  return TbRow{o, h};
}
// the row of lane - 1 (FROM_LOWER) or lane + 1 in the wave, 0 at the wave's ends
// (DPP wave_shr:1 / wave_shl:1)
template <bool FROM_LOWER>
__device__ __forceinline__ TbRow tb_lane_nb(const TbRow &r) {
  constexpr int kCtrl = FROM_LOWER ? 0x138 : 0x130;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)r.o, kCtrl, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(r.o >> 32), kCtrl, 0xF, 0xF, true);
  const uint32_t h = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r.h, kCtrl, 0xF, 0xF, true);
  return TbRow{(uint64_t)lo | ((uint64_t)hi << 32), h};
}

__device__ __forceinline__ bool tb_region_in_box(const Box &b, int rx, int ry, int rz) {
  return rx >= b.lo[0] && rx <= b.hi[0] && 2 * ry + 1 >= b.lo[1] && 2 * ry <= b.hi[1] &&
         2 * rz + 1 >= b.lo[2] && 2 * rz <= b.hi[2];
}
__device__ __forceinline__ bool tb_tile_in_box(const Box &b, int tx, int ty, int tz) {
  return tx >= b.lo[0] && tx <= b.hi[0] && ty >= b.lo[1] && ty <= b.hi[1] && tz >= b.lo[2] &&
         tz <= b.hi[2];
}

// The single-pair dilation of a pair the two-step launch handed over after
// its pack: region seed merged with the first eight steps, then periods of
// eight steps between grid barriers, then the results.  One workgroup per
// CU: with two, the 256 VGPRs a wave could hold spilled to scratch memory
// (the master's fold and barrier among them); with one, the overflow goes to
// AGPRs, and a long pair's periods run about 5% faster -- a sparse front has
// fewer regions than CUs anyway.
__global__ __launch_bounds__(kWgThreads)
__attribute__((amdgpu_waves_per_eu(1))) void dilate_tb_kernel(DilateArgs a) {
  constexpr int KS = kTbSteps;
  constexpr int R = 64;  // rows per tile
  PairCtr *ctrs = a.ctrs;
  Engine *es = a.es;
  uint64_t *q = a.q;
  const uint32_t qcap = a.qcap;
  const uint64_t spin_limit = a.spin;
  uint32_t *qstats = a.qstats;
  const int qstats_len = a.qstats_len;
  uint64_t *trace = a.trace;
  const PairDesc &d = a.desc;

  // the first and last of each thread's four rows, per step parity (the
  // rows other threads read in z), y fastest so that the lanes of a wave
  // touch consecutive 16-byte words (no bank conflicts)
  __shared__ u64x2 s_win[2][8][2][kTbWin];
  __shared__ LocalQueue lq;
  __shared__ uint32_t s_pre[kSubQ + 1];
  __shared__ int s_ok;
  __shared__ uint32_t s_cnt[kArrWords];
  __shared__ uint32_t s_flags[1];
  __shared__ unsigned long long s_red[12 + kSubQ];
  __shared__ uint64_t s_st[2][1];  // the entry a workgroup keeps for the next period
  __shared__ uint32_t s_stn[2];
  __shared__ uint8_t s_kept[kMaxGrid];
  __shared__ uint32_t s_fs[2];
  __shared__ uint32_t s_src[27];    // per neighbour region: 0 none, 1 orig, 2 buf, 3 full
  __shared__ uint32_t s_or[3];      // per entry: input not full, reach bits, other left
  __shared__ uint64_t s_ent;        // entry taken from the global queue
  __shared__ uint32_t s_seed[2 * kWgThreads];
  __shared__ uint32_t s_nseed;
  __shared__ unsigned long long s_phase[3];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t kq = blockIdx.x % kSubQ;
  // the step mapping: thread t holds rows (yy, zz0 .. zz0 + 3) of the window
  const int yy = threadIdx.x & 31;
  const int seg = threadIdx.x >> 5;
  const int zz0 = 4 * seg;
  const int tyi = yy / 8 - 1, tzi = seg / 2 - 1;  // tile offsets in the region, -1 .. 2
  const bool core = yy >= 8 && yy < 24 && seg >= 2 && seg < 6;
  const int cy = yy - 8;  // core coordinates of the thread's rows (core threads)

  if (threadIdx.x == 0) lq.n = 0;
  if (threadIdx.x < kArrWords) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x < 2) s_stn[threadIdx.x] = 0;
  if (threadIdx.x == 0) {  // no kept entries yet: every slot is free
    s_fs[0] = blockIdx.x;
    s_fs[1] = gridDim.x;
  }
  PassState ps{};
  uint32_t t = 0;
  bool run = true;
  const uint32_t eoff = a.ebase + 3;  // the barrier closing period tp has epoch eoff + tp + 1
  PackSums *acc = &es->acc[a.acc_set & 1];
  PairCtr &c0 = ctrs[0];
  // The two-step launch before this one on the stream reset and packed the
  // pair and zeroed its region states; it hands the pair over with a word
  // tagged with this launch's epoch base.  Without that word (a launch the
  // host queued by mistake) every workgroup leaves at once.
  {
    const unsigned long long h = ld_sc(&es->handoff[0]);
    if (!(tagged(h, a.ebase) && (h & kTagValue) == 2ull)) return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s_phase[0] = ld_sc(&es->handoff[1]);
    s_phase[1] = ld_sc(&es->handoff[2]);
  }
  const int rgx = d.gx, rgy = (d.gy + 1) / 2, rgz = (d.gz + 1) / 2;
  const uint32_t nreg = (uint32_t)rgx * (uint32_t)rgy * (uint32_t)rgz;
  // totals and box from the pack partials
  __shared__ unsigned long long s_tot[4];
  __shared__ int s_box[6];
  if (threadIdx.x < 64) {
    unsigned long long v[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) v[f] = wave_sum64(ld_sc(&acc->sums[f][lane]));
    int bx[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int x = wave_max(ld_sc(&acc->box[k][lane]), lane);
      bx[k] = k < 3 ? box_dec_lo(x) : box_dec_hi(x);
    }
    if (lane == 0) {
#pragma unroll
      for (int f = 0; f < 4; ++f) s_tot[f] = v[f];
#pragma unroll
      for (int k = 0; k < 6; ++k) s_box[k] = bx[k];
    }
  }
  __syncthreads();
  const PairView v0 = make_view(d, s_box, s_box + 3, s_tot[2], s_tot[3]);
  const Box &bb = v0.b;
  if (blockIdx.x == 0 && threadIdx.x < 64) {  // wave 0 holds the pass state
    for (int m = 0; m < 2; ++m) {
      ps.cut[m] = v0.cut[m];
      ps.cov[m] = ps.sum[m] = 0;
      ps.hd[m] = 0;
      ps.done_from[m] = 0x7FFFFFFF;
    }
  }
  bool pdone[2] = {v0.cut[0] == 0, v0.cut[1] == 0};
  // The seed period runs on every workgroup (its queue is as dense as the
  // volume's surface); the later periods, whose queue holds the moving fronts
  // only, on the first `active` (one per CU by default), because every
  // workgroup adds an arrival to the barrier.  The others keep no entry for
  // the next period and leave after the first barrier.
  const uint32_t active =
      a.tb_active > 0 ? min(gridDim.x, (uint32_t)a.tb_active) : gridDim.x;
  const bool leaving = blockIdx.x >= active;
  const uint32_t kcap = leaving ? 0u : 1u;  // entries kept for the next period

  // One queued (region, mask) by the whole workgroup at period tp; requeue:
  // whether it may queue itself for the next period (not when another
  // workgroup's claim already did).
  uint64_t *trp = nullptr;  // trace record of the running period (timing 2) or null
  auto step_region = [&](uint64_t ent, uint32_t tp, bool requeue) {
    const uint64_t ts1 = trp ? wall_clock64() : 0ull;
    const uint32_t rg = (uint32_t)ent;
    const int m = (int)((ent >> 48) & 1);
    const bool check = ((ent >> 50) & 1) != 0;
    const int rx = (int)(rg % (uint32_t)rgx);
    const int ry = (int)((rg / (uint32_t)rgx) % (uint32_t)rgy);
    const int rz = (int)(rg / ((uint32_t)rgx * (uint32_t)rgy));
    const int tx = rx, ty0 = 2 * ry, tz0 = 2 * rz;
    uint32_t *rst = m ? v0.rstate[1] : v0.rstate[0];
    const uint64_t *orig_m = (const uint64_t *)(m ? v0.orig[1] : v0.orig[0]);
    const uint64_t *orig_o = (const uint64_t *)(m ? v0.orig[0] : v0.orig[1]);
    // pointer selects, not indexing: indexing the view with run-time values
    // would put it in scratch memory
    uint64_t *buf0 = (uint64_t *)(m ? v0.buf[1][0] : v0.buf[0][0]);
    uint64_t *buf1 = (uint64_t *)(m ? v0.buf[1][1] : v0.buf[0][1]);
    const uint64_t *bufr = (tp & 1) ? buf0 : buf1;  // slot (tp + 1) & 1
    uint64_t *bufw = (tp & 1) ? buf1 : buf0;        // slot tp & 1
    // 1. in one round trip: the state of every region of the 3 x 3 x 3
    // neighbourhood (threads 0..26), and both candidates (orig, buf) of the
    // thread's four rows in the tiles at x-1, x, x+1; the states then pick
    if (threadIdx.x < 27) {
      const int j = threadIdx.x;
      const int nx_ = rx + j % 3 - 1, ny_ = ry + (j / 3) % 3 - 1, nz_ = rz + j / 9 - 1;
      uint32_t src = 0;
      if (tb_region_in_box(bb, nx_, ny_, nz_)) {
        const uint32_t st = ld_sc(rst + (uint32_t)nx_ + (uint32_t)rgx * ((uint32_t)ny_ + (uint32_t)rgy * (uint32_t)nz_));
        src = (st & kFull) ? 3u : (st != 0 && st - 1 < tp) ? 2u : 1u;
      }
      s_src[j] = src;
    }
    if (threadIdx.x < 3) s_or[threadIdx.x] = 0;
    const int ty = ty0 + tyi, tz = tz0 + tzi;
    const int yl = yy & 7;
    const bool yok = ty * 8 + yl < v0.ny;
    bool tin[3];
    uint64_t vb[3];
    uint32_t tix[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int ux = tx + c - 1;
      tin[c] = tb_tile_in_box(bb, ux, ty, tz);
      vb[c] = tin[c] ? valid_bits<uint64_t>(ux, v0.nx) : 0ull;
      tix[c] = tin[c] ? (uint32_t)ux + (uint32_t)v0.gx * ((uint32_t)ty + (uint32_t)v0.gy * (uint32_t)tz) : 0u;
    }
    bool rok[4];
    uint64_t wo[4][3], wb[4][3], oth[4];
    // the three tiles' flags with the rows: a mask empty in a tile has zero
    // original rows there, which the near launch's pack does not write
    uint32_t tfl[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) tfl[c] = ldg(d.tflags + tix[c]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int zl = (zz0 + i) & 7;
      rok[i] = yok && tz * 8 + zl < v0.nz;
      const uint32_t r = (uint32_t)(yl + 8 * zl);
      // every load is issued, at a valid address (tile 0 for a tile outside
      // the box), and dropped afterwards where it does not count: no branch
      // splits them into several round trips
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const uint64_t o = (uint64_t)tix[c] * R + r;
        wo[i][c] = ldg(orig_m + o);
        wb[i][c] = ld_sc(bufr + o);
      }
      oth[i] = ldg(orig_o + (uint64_t)tix[1] * R + r);
    }
    const bool oth_any = (tfl[1] >> (m ? 0 : 16)) & 1u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!(core && check && rok[i] && tin[1] && oth_any)) oth[i] = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!(rok[i] && tin[c])) wo[i][c] = wb[i][c] = 0;
        if (!((tfl[c] >> (m ? 16 : 0)) & 1u)) wo[i][c] = 0;
      }
    }
    __syncthreads();
    const int oyr = tyi < 0 ? -1 : tyi > 1 ? 1 : 0, ozr = tzi < 0 ? -1 : tzi > 1 ? 1 : 0;
    uint32_t srcx[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) srcx[c] = tin[c] ? s_src[c + 3 * (oyr + 1) + 9 * (ozr + 1)] : 0u;
    TbRow X[4], V[4];
    bool notfull = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      uint64_t w[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        w[c] = srcx[c] == 1 ? wo[i][c] : srcx[c] == 2 ? wb[i][c] : srcx[c] == 3 ? vb[c] : 0ull;
      V[i] = rok[i] ? tb_make(vb[0], vb[1], vb[2]) : TbRow{0, 0};
      X[i] = tb_make(w[0], w[1], w[2]);
      X[i].o &= V[i].o, X[i].h &= V[i].h;
      if (core && X[i].o != V[i].o) notfull = true;
    }
    if (__ballot(notfull) != 0 && lane == 0) s_or[0] = 1;
    __syncthreads();
    if (s_or[0] == 0) {  // input full: the region leaves the queue (hello.cu:1319, 1634-1643)
      if (threadIdx.x == 0) or_g(rst + rg, kFull);
      __syncthreads();
      return;
    }
    // 2. eight steps in the window
    // this thread's new hits per step (at most 4 x 64): step k < 7 in bits
    // 9k .. 9k + 8 of cntp, step 7 in cnt7 -- no array indexed by the step,
    // so the loop need not be unrolled
    static_assert(KS == 8, "eight steps: seven packed counts and one more");
    uint64_t cntp = 0;
    uint32_t cnt7 = 0;
#pragma unroll 1
    for (int k = 0; k < KS; ++k) {
      // rows z-1 and z+1 beyond the thread's four: the last row of the
      // segment below and the first of the segment above, through LDS.  At the
      // window's edges (z = 0 and 31 here, y = 0 and 31 below) the neighbour
      // read is some other row: a wrong voxel there moves inward one row per
      // step, so after eight steps it is still one row short of the region's
      // own rows (8 rows in) -- the edges need no masking
      s_win[k & 1][seg][0][yy] = u64x2{X[0].o, X[0].h};
      s_win[k & 1][seg][1][yy] = u64x2{X[3].o, X[3].h};
      __syncthreads();
      const u64x2 ezm = s_win[k & 1][seg > 0 ? seg - 1 : 0][1][yy];
      const u64x2 ezp = s_win[k & 1][seg < 7 ? seg + 1 : 7][0][yy];
      TbRow N[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // rows y-1 and y+1 are the same row of the lanes beside (lane = y in
        // each half wave)
        const TbRow ym = tb_lane_nb<true>(X[i]), yp = tb_lane_nb<false>(X[i]);
        const TbRow zm = i > 0 ? X[i - 1] : TbRow{ezm.x, (uint32_t)ezm.y};
        const TbRow zp = i < 3 ? X[i + 1] : TbRow{ezp.x, (uint32_t)ezp.y};
        const TbRow xd = tb_xdil(X[i]);
        N[i] = TbRow{(xd.o | ym.o | yp.o | zm.o | zp.o) & V[i].o,
                     (xd.h | ym.h | yp.h | zm.h | zp.h) & V[i].h};
      }
      // new hits: res & ~src & ref (hello.cu:1590-1626), step 8tp + k + 1
      // (oth is zero outside the core)
      uint64_t nw[4];
      uint32_t c = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        nw[i] = N[i].o & ~X[i].o & oth[i];
        c += (uint32_t)__popcll(nw[i]);
      }
      if (k < KS - 1)
        cntp += (uint64_t)c << (9 * k);
      else
        cnt7 = c;
      if (v0.map) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          uint64_t bits = nw[i];
          const int zl = (zz0 + i) & 7;
          const uint64_t vrow = (uint64_t)tx * 64 + (uint64_t)v0.nx *
                                ((uint64_t)(ty * 8 + yl) + (uint64_t)v0.ny * (uint64_t)(tz * 8 + zl));
          const uint16_t it = (uint16_t)(KS * tp + k + 1);
          while (bits) {
            const int kb = (int)__builtin_ctzll(bits);
            bits &= bits - 1;
            v0.map[vrow + kb] = it;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) X[i] = N[i];
    }
    // 3. the neighbour regions the output reaches within eight steps: a
    // voxel at (cx, cy, cz) of the region's 64 x 16 x 16 core is dy + dz away
    // in y and z from the neighbour at (ox, oy, oz) and, with ox != 0, cx + 1
    // or 64 - cx in x; the lowest and highest voxel of a row bound the x part
    uint32_t reach = 0;
    bool left = false;
    if (core) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t y = X[i].o;
        if (oth[i] & ~y) left = true;
        if (y == 0) continue;
        const int cz = zz0 + i - 8;
        const int xlo = __builtin_ctzll(y) + 1, xhi = __builtin_clzll(y) + 1;  // x distances
#pragma unroll
        for (int j = 0; j < 27; ++j) {
          if (j == 13) continue;
          const int ox = j % 3 - 1, oy = (j / 3) % 3 - 1, oz = j / 9 - 1;
          const int dy = oy < 0 ? cy + 1 : oy > 0 ? 16 - cy : 0;
          const int dz = oz < 0 ? cz + 1 : oz > 0 ? 16 - cz : 0;
          const int dx = ox < 0 ? xlo : ox > 0 ? xhi : 0;
          if (dx + dy + dz <= KS) reach |= 1u << j;
        }
      }
    }
    reach = (uint32_t)wave_reduce32((int)reach, 0, [](int x, int y) { return x | y; });
    const bool anyleft = __ballot(left) != 0;
    if (lane == 0) {
      if (reach) atomicOr(&s_or[1], reach);
      if (anyleft) s_or[2] = 1;
    }
    __syncthreads();
    // 4. claims first (their compare-and-swap travels while the output is
    // written and the hits are summed), then the pushes (wave 0): the region
    // itself for the next period, then the idle neighbours it won
    bool won = false;
    uint32_t nrg = 0;
    if (wave == 0 && lane < 27 && lane != 13 && ((s_or[1] >> lane) & 1) && s_src[lane] == 1) {
      const int j = lane;
      const int nx_ = rx + j % 3 - 1, ny_ = ry + (j / 3) % 3 - 1, nz_ = rz + j / 9 - 1;
      nrg = (uint32_t)nx_ + (uint32_t)rgx * ((uint32_t)ny_ + (uint32_t)rgy * (uint32_t)nz_);
      won = cas_g(rst + nrg, 0u, tp + 2) == 0u;
    }
    if (core && srcx[1] != 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int zl = (zz0 + i) & 7;
        if (rok[i]) st_sc(bufw + (uint64_t)tix[1] * R + (uint32_t)(yl + 8 * zl), X[i].o);
      }
    }
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const uint32_t ck = k < KS - 1 ? (uint32_t)(cntp >> (9 * k)) & 511u : cnt7;
      const uint32_t sum = (uint32_t)wave_reduce32((int)ck, 0, [](int x, int y) {
        return (int)((uint32_t)x + (uint32_t)y);
      });
      if (lane == 0 && sum) atomicAdd(&s_cnt[k == 0 ? m : 2 + 2 * k + m], sum);
    }
    if (wave == 0) {
      uint64_t *Qout = sub_queue(q, qcap, (tp + 1) & 1, kq);
      uint32_t *qout_len = &es->qlen[(tp + 1) % 4][kq];
      const uint32_t par = (tp + 1) & 1;
      push_sticky(requeue && lane == 0, make_entry(rg, 0u, (uint32_t)m, 0, s_or[2] != 0, v0.gen),
                  s_st[par], &s_stn[par], kcap, lq, Qout, qout_len, lane, &s_cnt[2]);
      push_sticky(won, make_entry(nrg, 0u, (uint32_t)m, 1, 1, v0.gen), s_st[par], &s_stn[par],
                  kcap, lq, Qout, qout_len, lane, &s_cnt[2]);
    }
    __syncthreads();
    // trace (timing 2): the period's entries stepped and the sum of their
    // durations in ticks
    if (trp && threadIdx.x == 0) {
      atomicAdd((unsigned long long *)&trp[6], 1ull);
      atomicAdd((unsigned long long *)&trp[7], (unsigned long long)(wall_clock64() - ts1));
    }
  };

  // Periods.  Period 0 is the seed: each workgroup scans its regions (region
  // i = blockIdx.x + gridDim.x * i, 256 at a time), queues as seed_kernel's
  // rule does -- a region that has voxels and is not full, or an empty one
  // whose neighbours hold voxels within eight steps of it (hello.cu:762-877,
  // 1649-1663) -- and steps the regions it queued itself at once; a full
  // region becomes kFull.  Claims of period-0 steps (0 -> 2) can reach a
  // region before its own scan: the scan then takes it over (2 -> 1) and steps
  // it without queueing it again.  Later periods take the workgroup's kept
  // entry or the first of its free range of the global queue, then the global
  // queue round robin.  Every entry goes through the one call of step_region.
  const uint32_t per = gridDim.x * blockDim.x;
  // the seed queue: parity 0 of the queue array, free during the first period
  uint64_t *seedq = sub_queue(q, qcap, 0, 0);
  const uint32_t seed_epoch = a.ebase + 3;
  // a seed word carries this launch's epoch and bit 39 (kSeedMark), which no
  // queue entry of a single pair has (its pair field is 0), so a slot still
  // holding an older entry never reads as a seed
  constexpr unsigned long long kSeedMark = 1ull << 39;
  auto is_seed = [&](uint64_t w) { return tagged(w, seed_epoch) && (w & kSeedMark) != 0; };
  bool scanned = false;
  uint32_t sj = 0, se = 0;  // this workgroup's share of the seeds: [sj, se)
  uint32_t chunk_end = nreg;
  // each workgroup's segment of the seed area: two seeds per region it scans
  const uint32_t seg_cap = 2 * ((nreg + gridDim.x - 1) / gridDim.x);
  while (run) {
    const uint32_t tp = t;
    if (tp > 0) {
      if ((s_flags[0] & 8u) == 0) break;  // no entries anywhere
      pdone[0] = (s_flags[0] & 1u) != 0;
      pdone[1] = (s_flags[0] & 2u) != 0;
      if (pdone[0] && pdone[1]) break;
    }
    uint64_t *tr = (trace && (int)t < qstats_len) ? trace + 8 * (uint64_t)t : nullptr;
    if (tr && blockIdx.x == 0 && threadIdx.x == 0) {
      tr[0] = wall_clock64();
      if (tp > 0) tr[5] = ps.cur_total;  // period 0: the seeds, once counted
    }
    trp = tr;
    const uint32_t L = tp > 0 ? s_pre[kSubQ] : 0u;
    const uint32_t nloc = tp > 0 ? s_stn[t & 1] : 0u;
    int phase = 0;             // later periods: 0 the first slot, 1 round robin, 2 done
    uint32_t ei = min(L, s_fs[1]) + blockIdx.x;  // next round-robin global entry
    for (;;) {
      uint64_t ent = 0;
      bool have = false, requeue = true, global = false;
      uint32_t gi = 0;
      if (tp == 0) {
        if (!scanned) {
          // this workgroup's regions, 256 at a time: seeds go to the
          // workgroup's own segment of the seed area as tagged words
          if (threadIdx.x == 0) s_nseed = 0;
          __syncthreads();
          // the segments fit the parity-0 queue area (16 sub-queues of qcap =
          // 2 x tiles entries) when the grid has at most one workgroup per
          // region and 2 x kWgThreads in all (the host sizes it so); a grid
          // that does not is refused here rather than written past
          if ((uint64_t)gridDim.x * seg_cap > (uint64_t)kSubQ * qcap ||
              gridDim.x > 2u * kWgThreads) {
            if (threadIdx.x == 0) atomicOr(&es->error, 1u);
            chunk_end = 0;
          }
          for (uint32_t chunk = 0; chunk < chunk_end; chunk += per) {
            const uint32_t rg = blockIdx.x + gridDim.x * (chunk / gridDim.x + threadIdx.x);
            if (rg >= nreg) continue;
            const int rx = (int)(rg % (uint32_t)rgx);
            const int ry = (int)((rg / (uint32_t)rgx) % (uint32_t)rgy);
            const int rz = (int)(rg / ((uint32_t)rgx * (uint32_t)rgy));
            if (!tb_region_in_box(bb, rx, ry, rz)) continue;
            // the flags of the region's tiles and of the ring of tiles
            // around them (x +-1, y and z +-1 tile), all loads at once
            uint32_t fl[3][4][4];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  const int ux = rx + c - 1, ty = 2 * ry + b - 1, tz = 2 * rz + e - 1;
                  const bool in = tb_tile_in_box(bb, ux, ty, tz);
                  // tile 0 stands in for a tile outside the box: every
                  // load issued at once, the stand-ins dropped below
                  const uint32_t ti = in ? (uint32_t)ux + (uint32_t)v0.gx * ((uint32_t)ty + (uint32_t)v0.gy * (uint32_t)tz) : 0u;
                  fl[c][b][e] = ld_sc(d.tflags + ti);
                }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                  if (!tb_tile_in_box(bb, rx + c - 1, 2 * ry + b - 1, 2 * rz + e - 1)) fl[c][b][e] = 0;
            uint32_t any = 0, allfull = 3, reach = 0;  // bit m
#pragma unroll
            for (int b = 1; b <= 2; ++b)
#pragma unroll
              for (int e = 1; e <= 2; ++e) {
                const int ty = 2 * ry + b - 1, tz = 2 * rz + e - 1;
                if (!tb_tile_in_box(bb, rx, ty, tz)) continue;
                const uint32_t f = fl[1][b][e];
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                  const uint32_t fm = f >> (16 * m);
                  any |= (fm & 1u) << m;
                  if (!((fm & 1u) && (fm & 2u))) allfull &= ~(1u << m);
                }
              }
            // voxels of the ring within eight steps of an empty mask's
            // region: any voxel of a y / z ring tile, and of an x
            // neighbour its eight columns toward the region
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                  if (c == 1 && b >= 1 && b <= 2 && e >= 1 && e <= 2) continue;
                  const uint32_t f = fl[c][b][e];
                  const uint32_t bit = c == 1 ? 0u : c == 0 ? 15u : 14u;
                  reach |= ((f >> bit) & 1u) | (((f >> (16 + bit)) & 1u) << 1);
                }
            for (int m = 0; m < 2; ++m) {
              if (m ? pdone[1] : pdone[0]) continue;
              uint32_t *rst = m ? v0.rstate[1] : v0.rstate[0];
              const bool an = (any >> m) & 1u;
              if (an && ((allfull >> m) & 1u)) {
                or_g(rst + rg, kFull);
              } else if (an || ((reach >> m) & 1u)) {
                uint32_t old = cas_g(rst + rg, 0u, 1u);
                bool rq = true;
                if (old == 2u) {  // claimed for period 1 by a period-0 step
                  old = cas_g(rst + rg, 2u, 1u) == 2u ? 0u : 1u;
                  rq = false;
                }
                if (old == 0u) {
                  const uint32_t pos = atomicAdd(&s_nseed, 1u);  // LDS: no contended counter
                  st_sc(seedq + (uint64_t)blockIdx.x * seg_cap + pos,
                        (uint64_t)tag(seed_epoch, kSeedMark | (rg << 2) | ((uint32_t)m << 1) | (rq ? 1u : 0u)));
                }
              }
            }
          }
          __syncthreads();  // every seed of this workgroup is counted in s_nseed
          // publish the count (tagged, in this workgroup's free-slot word,
          // which the first barrier overwrites later), then gather every
          // workgroup's count and take an equal contiguous share of all seeds
          if (threadIdx.x == 0) st_sc(&es->fslot[blockIdx.x], tag(seed_epoch, s_nseed));
          uint32_t part = 0;
          for (uint32_t w2 = 2 * threadIdx.x; w2 < 2 * threadIdx.x + 2; ++w2) {
            uint32_t cw = 0;
            if (w2 < gridDim.x) {
              const uint64_t t0 = wall_clock64();
              for (;;) {
                const unsigned long long x = ld_sc(&es->fslot[w2]);
                if (tagged(x, seed_epoch)) {
                  cw = (uint32_t)(x & kTagValue);
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > spin_limit) {
                  atomicOr(&es->error, 1u);
                  break;
                }
              }
            }
            s_seed[w2] = cw;
            part += cw;
          }
          // exclusive prefix over the workgroups' counts (two per thread)
          const uint32_t incl = wave_scan32(part);
          __shared__ uint32_t s_psum[kWavesPerWg];
          if (lane == 63) s_psum[wave] = incl;
          __syncthreads();
          uint32_t before = incl - part, total = 0;
#pragma unroll
          for (int k2 = 0; k2 < kWavesPerWg; ++k2) {
            if (k2 < wave) before += s_psum[k2];
            total += s_psum[k2];
          }
          const uint32_t c0w = s_seed[2 * threadIdx.x], c1w = s_seed[2 * threadIdx.x + 1];
          __syncthreads();
          s_seed[2 * threadIdx.x] = before;
          s_seed[2 * threadIdx.x + 1] = before + c0w;
          (void)c1w;
          __syncthreads();
          sj = (uint32_t)((uint64_t)total * blockIdx.x / gridDim.x);
          se = (uint32_t)((uint64_t)total * (blockIdx.x + 1) / gridDim.x);
          scanned = true;
          if (tr && blockIdx.x == 0 && threadIdx.x == 0) tr[5] = total;
        }
        // the next seed of this workgroup's share: global index sj, in the
        // segment of the workgroup whose prefix range holds it
        if (sj < se && threadIdx.x == 0) {
          uint32_t lo = 0, hi = gridDim.x - 1;  // last w with prefix[w] <= sj
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_seed[mid] <= sj)
              lo = mid;
            else
              hi = mid - 1;
          }
          const uint64_t *src = seedq + (uint64_t)lo * seg_cap + (sj - s_seed[lo]);
          const uint64_t t0 = wall_clock64();
          uint64_t w = 0;
          for (;;) {  // the count was published after the seed, but may be seen first
            w = ld_sc(src);
            if (is_seed(w)) break;
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > spin_limit) {
              atomicOr(&es->error, 1u);
              w = 0;
              break;
            }
          }
          s_ent = w;
        }
        if (sj < se) {
          ++sj;
        __syncthreads();
        const uint64_t w = s_ent;
        __syncthreads();
        if (is_seed(w)) {
          const uint32_t e = (uint32_t)(w & 0xFFFFFFFFull);
          ent = make_entry(e >> 2, 0u, (e >> 1) & 1u, 1, 1, v0.gen);
          requeue = (e & 1u) != 0;
          have = true;
        }
        }
      } else {
        if (phase == 0) {
          phase = 1;
          if (nloc) {
            ent = s_st[t & 1][0];
            have = true;
          } else if (s_fs[0] < L && s_fs[0] < s_fs[1]) {
            gi = s_fs[0];
            have = global = true;
          }
        }
        if (!have && phase == 1) {
          if (ei < L) {
            gi = ei;
            ei += gridDim.x;
            have = global = true;
          } else {
            phase = 2;
          }
        }
      }
      if (!have) break;
      if (global) {  // global entry gi of this period's queue
        if (threadIdx.x == 0) {
          int sq = 0;
          while (sq + 1 < kSubQ && s_pre[sq + 1] <= gi) ++sq;
          s_ent = ld_sc(sub_queue(q, qcap, tp & 1, sq) + (gi - s_pre[sq]));
        }
        __syncthreads();
        ent = s_ent;
        __syncthreads();
      }
      const int m = (int)((ent >> 48) & 1);
      if (m ? pdone[1] : pdone[0]) continue;  // a pass whose cut was reached stops (hello.cu:1748-1759)
      if (KS * tp + 1 >= (uint32_t)v0.hist_len) {  // beyond any distance in the volume
        if (threadIdx.x == 0) st_sc(&c0.status, -8);
        continue;
      }
      step_region(ent, tp, requeue);
    }
    if (tr) {
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint64_t now = wall_clock64();
        atomicMax((unsigned long long *)&tr[1], (unsigned long long)now);
        if (blockIdx.x == 0) tr[2] = now;
      }
    }
    const uint32_t nflushed =
        flush_local_queue(lq, sub_queue(q, qcap, (t + 1) & 1, kq), &es->qlen[(t + 1) % 4][kq]);
    if (threadIdx.x == 0) s_cnt[2] += nflushed, s_cnt[3] = s_stn[(t + 1) & 1];
    if (tp == 0 && blockIdx.x == 0 && threadIdx.x == 0) s_phase[2] = wall_clock64();
    const uint32_t tdone = t;
    ++t;
    if (!grid_barrier_single<KS>(es, eoff + t, spin_limit, &s_ok, s_cnt, &s_stn[tdone & 1], ps,
                                 v0, tdone, qstats, qstats_len, tr, s_pre, s_flags, s_red, 1,
                                 s_kept, s_fs, (t + 2) % 4, tdone == 0 ? gridDim.x : active,
                                 active))
      break;
    if (leaving) break;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // the master's state, for finalize
    for (int m = 0; m < 2; ++m) {
      st_sc(&c0.cov[m], ps.cov[m]);
      st_sc(&c0.sum_d[m], ps.sum[m]);
      st_sc(&c0.hd_last[m], ps.hd[m]);
      st_sc(&c0.done_from[m], ps.done_from[m]);
    }
  }
  if (blockIdx.x == 0) {
    const bool err = ld_sc(&es->error) != 0;
    if (threadIdx.x == 0) a.outs[0] = finalize_pair(d, c0, t, err, KS);
    if (a.ready) __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t qn = ps.next_total;
      st_sc(&es->t, t);
      st_sc(&es->qnext, qn);
      st_sc(&es->qtotal, ps.qtotal), st_sc(&es->qmax, ps.qmax);
      if (a.head) {
        EngineHead h;
        h.t = t;
        h.error = err ? 1u : 0u;
        h.qmax = ps.qmax;
        h.nwg = gridDim.x;
        h.qtotal = ps.qtotal;
        h.qcap = es->qcap;
        h.qnext = qn;
        h.phase[0] = s_phase[0];
        h.phase[1] = s_phase[1];
        h.phase[2] = s_phase[2];
        h.phase[3] = wall_clock64();
        *a.head = h;
      }
      if (a.ready) {  // the head too reaches the host before the marker
        __threadfence_system();
        __hip_atomic_store(a.ready, a.ready_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
}

// ---------------------------------------------------------------- near pass
// The single-pair launch for pairs whose voxels lie near the other mask
// (dilate_near_kernel; 64x8x8 tiles, 6 element).  It packs the pair's rows,
// counts and box as the two-step launch does, and marks, in a bitmap, every region of
// 1 x 2 x 2 tiles that holds voxels to be covered (FP for gold's dilation, FN
// for segm's); it writes no x-face words and only the tile flags the region
// launch reads (non-empty, full, voxels in the eight columns toward -x / +x).  After the one grid barrier that ends the pack, each workgroup
// reads the bitmap and takes an equal share of the marked (region, mask)
// entries -- every workgroup counts the marks itself, so no queue, counter or
// second barrier is needed -- and for each one dilates the *other* mask's
// original bits eight steps in the region's 32 x 32-row window (the region
// plus an eight-voxel halo, rows of 80 bits; the same window and step as the
// region launch).  After step k the region's own rows are exact, so the new
// hits res & ~src & ref of step k (hello.cu:1590-1626) are exactly the
// region's voxels at first-hit distance k: the per-step counts of all entries
// are the histogram bins 1 .. 8.  An entry stops stepping once every voxel it
// has to cover is covered.  Nothing is written back and no neighbour is
// claimed, because no later step reads the dilated bits.  Each workgroup then
// adds its counts per (pass, step) into spread slots and counts itself in;
// the one that arrives last folds the bins in order with the robust cut
// (hello.cu:1748-1759) and writes the results.  A pass whose cut is not
// reached within eight steps (HD > 8) hands the packed pair to the region
// launch (status 2), which computes it from the start; so does a pair whose
// HD bound from the tile boxes already exceeds eight steps, right after the
// pack.
// LDS of one near-pass entry: the window rows exchanged in z per step, and
// the entry's counts
struct NearLds {
  u64x2 win[2][8][2][kTbWin];
  uint32_t hs[kTbSteps];  // new hits per step
  uint32_t need;          // voxels to cover
};
// One near-pass entry by the whole workgroup: region rg of pair d (1 x 2 x 2
// tiles of 64x8x8), pass m.  It builds the region launch's window (32 x 32
// rows in (y, z), four rows of consecutive z per thread, each 64 bits plus
// eight halo columns on each side) from the original packed bits of mask m,
// and runs up to eight 6-neighbour steps there, stopping once every voxel of
// the other mask in the region (not in mask m) is covered.  L.hs and L.need
// must be zero on entry; on return, after a workgroup barrier, L.hs[k] holds
// the region's new hits of step k + 1, which the caller takes and clears
// (with L.need) before the next entry.  The packed masks must be visible to
// plain loads: written by an earlier launch, or written through in this one
// and read by no workgroup before.
__device__ __forceinline__ void near_window(const PairDesc &d, uint32_t rg, int m, NearLds &L) {
  constexpr int KS = kTbSteps;
  constexpr int R = 64;  // rows per tile
  const int lane = threadIdx.x & 63;
  // window mapping of the region launch: thread t holds rows (yy, zz0 .. zz0 + 3)
  const int yy = threadIdx.x & 31;
  const int seg = threadIdx.x >> 5;
  const int zz0 = 4 * seg;
  const int tyi = yy / 8 - 1, tzi = seg / 2 - 1;  // tile offsets in the region, -1 .. 2
  const bool core = yy >= 8 && yy < 24 && seg >= 2 && seg < 6;
  const int rgx = d.gx, rgy = (d.gy + 1) / 2;
  const int rx = (int)(rg % (uint32_t)rgx);
  const int ry = (int)((rg / (uint32_t)rgx) % (uint32_t)rgy);
  const int rz = (int)(rg / ((uint32_t)rgx * (uint32_t)rgy));
  // pass m dilates mask m and covers the other mask's voxels
  const uint64_t *orig_m = (const uint64_t *)(m ? d.orig[1] : d.orig[0]);
  const uint64_t *orig_o = (const uint64_t *)(m ? d.orig[0] : d.orig[1]);
  const int ty = 2 * ry + tyi, tz = 2 * rz + tzi;
  const int yl = yy & 7;
  const bool tyz = ty >= 0 && ty < d.gy && tz >= 0 && tz < d.gz;
  const bool yok = tyz && ty * 8 + yl < d.ny;
  bool tin[3];
  uint64_t vb[3];
  uint32_t tix[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int ux = rx + c - 1;
    tin[c] = tyz && ux >= 0 && ux < d.gx;
    vb[c] = tin[c] ? valid_bits<uint64_t>(ux, d.nx) : 0ull;
    tix[c] = tin[c] ? (uint32_t)ux + (uint32_t)d.gx * ((uint32_t)ty + (uint32_t)d.gy * (uint32_t)tz) : 0u;
  }
  // every load at once, at a valid address (tile 0 stands in for a tile
  // outside the grid), the stand-ins dropped afterwards; the packed masks
  // were written through by the pack and no workgroup read them before
  // the barrier, so plain loads see them
  bool rok[4];
  uint64_t w[4][3], oth[4];
  // the three tiles' flags with the rows: a mask empty in a tile has zero
  // rows there, which the near launch's pack does not write
  uint32_t tfl[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) tfl[c] = ldg(d.tflags + tix[c]);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int zl = (zz0 + i) & 7;
    rok[i] = yok && tz * 8 + zl < d.nz;
    const uint32_t r = (uint32_t)(yl + 8 * zl);
#pragma unroll
    for (int c = 0; c < 3; ++c) w[i][c] = ldg(orig_m + (uint64_t)tix[c] * R + r);
    oth[i] = ldg(orig_o + (uint64_t)tix[1] * R + r);
  }
  TbRow X[4], V[4];
  uint32_t need = 0;
  const bool oth_any = (tfl[1] >> (m ? 0 : 16)) & 1u;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!oth_any) oth[i] = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (!(rok[i] && tin[c] && ((tfl[c] >> (m ? 16 : 0)) & 1u))) w[i][c] = 0;
    V[i] = rok[i] ? tb_make(vb[0], vb[1], vb[2]) : TbRow{0, 0};
    X[i] = tb_make(w[i][0], w[i][1], w[i][2]);
    X[i].o &= V[i].o, X[i].h &= V[i].h;
    // the voxels to cover: the other mask's, not in the source (its own
    // rows only; the window's ring belongs to other entries)
    oth[i] = (core && rok[i] && tin[1]) ? oth[i] & ~X[i].o : 0ull;
    need += (uint32_t)__popcll(oth[i]);
  }
  need = (uint32_t)wave_reduce32((int)need, 0, [](int x, int y) {
    return (int)((uint32_t)x + (uint32_t)y);
  });
  if (lane == 0 && need) atomicAdd(&L.need, need);
  uint32_t covered = 0, want = 0;
#pragma unroll 1
  for (int k = 0; k < KS; ++k) {
    // rows z-1 / z+1 beyond the thread's four come through LDS; at the
    // window's edges the neighbour read is some other row, and a wrong
    // voxel there moves inward one row per step, so after eight steps it
    // is still one row short of the region's own rows
    L.win[k & 1][seg][0][yy] = u64x2{X[0].o, X[0].h};
    L.win[k & 1][seg][1][yy] = u64x2{X[3].o, X[3].h};
    __syncthreads();
    // every count read here was added before this barrier, so the whole
    // workgroup takes the same decision
    if (k == 0)
      want = L.need;
    else
      covered += L.hs[k - 1];
    if (covered >= want) break;  // every voxel of the entry is covered
    const u64x2 ezm = L.win[k & 1][seg > 0 ? seg - 1 : 0][1][yy];
    const u64x2 ezp = L.win[k & 1][seg < 7 ? seg + 1 : 7][0][yy];
    TbRow N[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const TbRow ym = tb_lane_nb<true>(X[i]), yp = tb_lane_nb<false>(X[i]);
      const TbRow zm = i > 0 ? X[i - 1] : TbRow{ezm.x, (uint32_t)ezm.y};
      const TbRow zp = i < 3 ? X[i + 1] : TbRow{ezp.x, (uint32_t)ezp.y};
      const TbRow xd = tb_xdil(X[i]);
      N[i] = TbRow{(xd.o | ym.o | yp.o | zm.o | zp.o) & V[i].o,
                   (xd.h | ym.h | yp.h | zm.h | zp.h) & V[i].h};
    }
    // new hits of step k + 1: res & ~src & ref (hello.cu:1590-1626)
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) c += (uint32_t)__popcll(N[i].o & ~X[i].o & oth[i]);
    c = (uint32_t)wave_reduce32((int)c, 0, [](int x, int y) {
      return (int)((uint32_t)x + (uint32_t)y);
    });
    if (lane == 0 && c) atomicAdd(&L.hs[k], c);
#pragma unroll
    for (int i = 0; i < 4; ++i) X[i] = N[i];
  }
  __syncthreads();  // every count of the entry is in L.hs
}

constexpr int kNearSteps = kTbSteps;
// bitmap words of the region marks a thread of the share count holds at most
// (256 threads: up to 2^16 regions of two passes)
constexpr uint32_t kNearWordsPerThread = 16;

__global__ __launch_bounds__(kWgThreads)
__attribute__((amdgpu_waves_per_eu(2))) void dilate_near_kernel(DilateArgs a) {
  constexpr int KS = kNearSteps;
  static_assert(KS == 8, "the window holds an eight-voxel halo");
  static_assert(2 * KS * kHitSlots == kWgThreads, "one hit slot per thread in the fold");
  PairCtr *ctrs = a.ctrs;
  Engine *es = a.es;
  const uint64_t spin_limit = a.spin;
  const PairDesc &d = a.desc;
  PairCtr &c0 = ctrs[0];
  const uint32_t qcap = a.qcap;

  __shared__ NearLds s_near;              // the running entry's window and counts
  __shared__ int s_ok;
  __shared__ unsigned long long s_tot[4];
  __shared__ int s_hand;                  // 1: the pair goes to the region launch
  __shared__ uint32_t s_hits[2 * KS];     // the workgroup's hits, index m + 2k
  __shared__ uint32_t s_list[kWgThreads]; // mark bit indices of the entries taken
  __shared__ uint32_t s_wsum[kWavesPerWg];
  __shared__ int s_last;
  __shared__ unsigned long long s_h[2 * KS];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint32_t nreg = pair_regions(d);

  if (threadIdx.x < 2 * KS) s_hits[threadIdx.x] = 0, s_h[threadIdx.x] = 0;
  if (threadIdx.x < KS) s_near.hs[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_near.need = 0;
  uint64_t ph0 = 0, ph1 = 0;  // workgroup 0: clock at its start and at the end of its pack
  if (blockIdx.x == 0) ph0 = wall_clock64();
  // trace (timing 2): one record of 8 words per workgroup -- [0] start, [1]
  // pack done, [2] past the pack barrier, [3] share counted, [4] entries
  // done, [5] entries taken, [6] partials read, [7] HW_ID | XCC_ID << 32 --
  // and one more for the fold of the workgroup that arrives last ([0] fold
  // start, [1] results written, [2] hit slots read, [3] fold computed, [4]
  // ready marker stored), where the trace holds them
  uint64_t *trw = (a.trace && blockIdx.x + 1 < (uint32_t)a.qstats_len)
                      ? a.trace + 8 * (uint64_t)blockIdx.x : nullptr;
  if (trw && threadIdx.x == 0) {
    trw[0] = wall_clock64();
    // where the workgroup runs: HW_ID (CU, shader array, engine) and XCC_ID
    trw[7] = (uint64_t)__builtin_amdgcn_s_getreg((31 << 11) | 4) |
             (uint64_t)__builtin_amdgcn_s_getreg((3 << 11) | 20) << 32;
  }
  PackSums *acc = &es->acc[a.acc_set & 1];
  // R: workgroup 0 zeroes the pair's counters and histograms, the engine's
  // state, and the pack partials (with the near pass's hit slots and arrival
  // counters) the next launch packs into (as the two-step launch does); every
  // workgroup zeroes its share of the region states a hand-off to the region
  // launch starts from and of the mark bitmap the next near launch fills.
  // Nothing here is read before the barrier that ends the pack, except by
  // the workgroup that wrote it.
  if (blockIdx.x == 0) {
    const int tid = threadIdx.x;
    if (tid == 0) {
      st_sc(&c0.n_gold, 0ull), st_sc(&c0.n_segm, 0ull);
      st_sc(&c0.n_fp, 0ull), st_sc(&c0.n_fn, 0ull);
      for (int m = 0; m < 2; ++m) {
        st_sc(&c0.cov[m], 0ull), st_sc(&c0.sum_d[m], 0ull);
        st_sc(&c0.done_from[m], 0x7FFFFFFF), st_sc(&c0.hd_last[m], 0u);
      }
      st_sc(&c0.status, 0), st_sc(&c0.near_done, 0);
      st_sc(&es->t, 0u), st_sc(&es->error, 0u), st_sc(&es->qmax, 0u);
      st_sc(&es->qtotal, 0ull), st_sc(&es->qcap, qcap), st_sc(&es->qnext, 0u);
    }
    PackSums *next = &es->acc[(a.acc_set & 1) ^ 1];
    for (int i = tid; i < 4 * kSumSlots; i += blockDim.x) st_sc(&next->sums[0][0] + i, 0ull);
    for (int i = tid; i < 18 * kSumSlots; i += blockDim.x) st_sc(&next->box[0][0] + i, 0);
    if (tid < 8) st_sc(&next->grab[32 * tid], 0u);
    for (int i = tid; i < 2 * kTbSteps * kHitSlots; i += blockDim.x)
      st_sc(&next->near_hits[0][0] + i, 0ull);
    if (tid == 0) st_sc(&next->near_done, 0u);
    for (int i = tid; i < d.hist_len; i += blockDim.x)
      st_sc(d.hist[0] + i, 0u), st_sc(d.hist[1] + i, 0u);
    for (int i = tid; i < 4 * kLineWords; i += blockDim.x) st_sc(&es->qlen[0][0] + i, 0u);
  }
  {
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, gn = gridDim.x * blockDim.x;
    for (uint32_t i = gt; i < nreg; i += gn) st_sc(d.rstate[0] + i, 0u), st_sc(d.rstate[1] + i, 0u);
    for (uint32_t i = gt; i < a.marks_zero_words; i += gn) st_sc(a.marks_zero + i, 0u);
  }
  // P: pack, every store written through, regions marked; lean: the rows,
  // counts, box, marks and the tile flags the region launch reads
  {
    __shared__ __align__(16) uint16_t s_pieces[kWavesPerWg][1024];
    pack_tiles<64, 8, 8, true, true, true>(&d, nullptr, ctrs, 1, a.total_tiles, s_pieces, acc,
                                     a.pack_pools ? &acc->grab[0] : nullptr, a.pack_pools);
  }
  if (blockIdx.x == 0) ph1 = wall_clock64();
  if (trw && threadIdx.x == 0) trw[1] = wall_clock64();
  bool ok = grid_sync_tagged(es, a.ebase + 2, spin_limit, &s_ok);
  if (trw && threadIdx.x == 0) trw[2] = wall_clock64();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st_sc(&es->near_phase[0], (unsigned long long)ph0);
    st_sc(&es->near_phase[1], (unsigned long long)ph1);
    st_sc(&es->near_phase[2], (unsigned long long)wall_clock64());
  }
  // The mark bitmap (agent-coherent: bits of several workgroups share a
  // word), thread t words [c t, c t + c), loaded together with the pack
  // partials: one round trip for both.
  const uint32_t W = 2 * nreg;                  // mark bits
  const uint32_t nw = (W + 31) / 32;            // mark words
  const uint32_t cw = (nw + kWgThreads - 1) / kWgThreads;  // words per thread (<= 16)
  uint32_t mw[kNearWordsPerThread];
#pragma unroll
  for (uint32_t q = 0; q < kNearWordsPerThread; ++q) {
    const uint32_t wi = threadIdx.x * cw + q;
    mw[q] = (ok && q < cw && wi < nw) ? ld_sc(a.marks + wi) : 0u;
  }
  // totals, box and the HD bound from the pack partials
  if (threadIdx.x < 64) {
    unsigned long long v[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) v[f] = wave_sum64(ld_sc(&acc->sums[f][lane]));
    int bx[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) {
      const int x = wave_max(ld_sc(&acc->box[k][lane]), lane);
      bx[k] = k % 6 < 3 ? box_dec_lo(x) : box_dec_hi(x);
    }
    const bool hand = a.tb_mode == 2 || hd_lower_bound(bx + 6, bx + 12) > KS;
    if (lane == 0) {
#pragma unroll
      for (int f = 0; f < 4; ++f) s_tot[f] = v[f];
      s_hand = hand ? 1 : 0;
      if (blockIdx.x == 0 && ok) {
        st_sc(&c0.n_gold, v[0]), st_sc(&c0.n_segm, v[1]);
        st_sc(&c0.n_fp, v[2]), st_sc(&c0.n_fn, v[3]);
#pragma unroll
        for (int k = 0; k < 3; ++k) st_sc(&c0.lo[k], bx[k]), st_sc(&c0.hi[k], bx[3 + k]);
      }
      if (trw) trw[6] = wall_clock64();
    }
  }
  __syncthreads();
  // The results (the workgroup that folds) or, for a hand-off, the tagged
  // hand-off word with the clock stamps of the start and the pack and, in
  // place of the results, status 2; then the ready marker, after both.  The
  // host queues the region launch on status 2.
  auto publish = [&](bool hand, bool err, uint32_t nent, uint64_t *trm, const PairOut *res) {
    if (hand) {
      st_sc(&es->handoff[1], ld_sc(&es->near_phase[0]));
      st_sc(&es->handoff[2], ld_sc(&es->near_phase[1]));
      st_sc(&es->handoff[0], tag(a.ebase, 2ull));
      __hip_atomic_store(&a.outs[0].status, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
      if (a.qstats && a.qstats_len > 0) a.qstats[0] = nent;
      a.outs[0] = res ? *res : finalize_pair(d, c0, 1, err, KS);
      st_sc(&es->t, 0u);
      st_sc(&es->qnext, 0u);
      st_sc(&es->qtotal, (unsigned long long)nent), st_sc(&es->qmax, nent);
      if (trm) trm[1] = wall_clock64();
      if (a.head) {
        EngineHead hh;
        // with a trace, t counts its records (one per workgroup and the
        // fold's), which the host copies out; the next launch's epochs start
        // past it
        hh.t = trm ? gridDim.x + 1 : 0;
        hh.error = err ? 1u : 0u;
        hh.qmax = nent;
        hh.nwg = gridDim.x;
        hh.qtotal = nent;
        hh.qcap = qcap;
        hh.qnext = 0;
        hh.phase[0] = ld_sc(&es->near_phase[0]);
        hh.phase[1] = ld_sc(&es->near_phase[1]);
        hh.phase[2] = ld_sc(&es->near_phase[2]);
        hh.phase[3] = wall_clock64();
        *a.head = hh;
      }
    }
    if (a.ready) {
      __threadfence_system();
      __hip_atomic_store(a.ready, a.ready_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (trm) trm[4] = wall_clock64();
  };
  // Entries only past a barrier that held and for a pair that stays: every
  // workgroup whose barrier held takes the same decision from the same
  // partials.  Every workgroup arrives below whatever it decided, so the one
  // that arrives last publishes (the hand-off, an error, or the results).
  const bool work = ok && !s_hand;

  // One entry: mark bit wi = region rg of pass m.
  auto near_entry = [&](uint32_t wi) {
    const int m = wi >= nreg ? 1 : 0;
    const uint32_t rg = wi - (m ? nreg : 0u);
    near_window(d, rg, m, s_near);
    if (threadIdx.x < KS) {
      s_hits[m + 2 * threadIdx.x] += s_near.hs[threadIdx.x];
      s_near.hs[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_near.need = 0;
    __syncthreads();  // cleared before the next entry adds to them
  };

  // The workgroup's share of the entries: every workgroup counts the marks
  // itself and takes entries [T b / G, T (b + 1) / G) of the T marked, in bit
  // order.
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t q = 0; q < kNearWordsPerThread; ++q) {
    const uint32_t b0 = (threadIdx.x * cw + q) * 32;  // bits past W never get set, but drop them
    if (b0 >= W || !work)
      mw[q] = 0;
    else if (W - b0 < 32)
      mw[q] &= (1u << (W - b0)) - 1u;
    mine += (uint32_t)__popc(mw[q]);
  }
  uint32_t before = 0, total = 0;
  {
    const uint32_t incl = wave_scan32(mine);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    before = incl - mine;
#pragma unroll
    for (int k = 0; k < kWavesPerWg; ++k) {
      if (k < wave) before += s_wsum[k];
      total += s_wsum[k];
    }
  }
  const uint32_t sb = (uint32_t)((uint64_t)total * blockIdx.x / gridDim.x);
  const uint32_t se = (uint32_t)((uint64_t)total * (blockIdx.x + 1) / gridDim.x);
  if (trw && threadIdx.x == 0) trw[3] = wall_clock64(), trw[5] = se - sb;
  for (uint32_t cs = sb; cs < se; cs += kWgThreads) {
    const uint32_t ce = min(se, cs + kWgThreads);
    // the entries [cs, ce) into s_list
    if (before < ce && before + mine > cs) {
      uint32_t g = before;
#pragma unroll
      for (uint32_t q = 0; q < kNearWordsPerThread; ++q) {
        uint32_t bits = mw[q];
        while (bits) {
          const int j = __builtin_ctz(bits);
          bits &= bits - 1;
          if (g >= cs && g < ce) s_list[g - cs] = (threadIdx.x * cw + q) * 32 + (uint32_t)j;
          ++g;
        }
      }
    }
    __syncthreads();
    for (uint32_t j = 0; j < ce - cs; ++j) near_entry(s_list[j]);
  }
  if (trw && threadIdx.x == 0) trw[4] = wall_clock64();
  // Arrival: the workgroup adds its hits into its slots and counts itself;
  // the last to arrive folds.  Every workgroup's adds are done (s_waitcnt 0)
  // before its count, so the one that folds reads them all.
  const uint32_t grp = blockIdx.x % kHitSlots;
  if (threadIdx.x < 2 * KS && s_hits[threadIdx.x])
    atomicAdd(&acc->near_hits[threadIdx.x][grp], (unsigned long long)s_hits[threadIdx.x]);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&acc->near_done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;

  // the fold, by the workgroup that arrived last: the hit slots and the
  // error flag in one round trip, the slots summed per bin in LDS, then wave
  // 0 folds the bins (one per lane) and thread 0 publishes
  uint64_t *trm = (a.trace && gridDim.x + 1 < (uint32_t)a.qstats_len)
                      ? a.trace + 8 * (uint64_t)gridDim.x : nullptr;
  if (trm && threadIdx.x == 0) trm[0] = wall_clock64();
  uint32_t eflag = 0;
  {
    const unsigned long long v = ld_sc(&acc->near_hits[0][0] + threadIdx.x);
    if (threadIdx.x < 64) eflag = ld_sc(&es->error);
    if (v) atomicAdd(&s_h[threadIdx.x / kHitSlots], v);
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  if (trm && threadIdx.x == 0) trm[2] = wall_clock64();
  const bool err = !ok || eflag != 0;
  if (err || s_hand) {  // a barrier timed out, or the pair goes to the region launch
    if (threadIdx.x == 0) publish(!err, err, 0, nullptr, nullptr);
    return;
  }
  // bins 1 .. 8 with the cut (hello.cu:1618 numbering); the bins taken go to
  // the histograms -- a pair handed over below has the region launch write
  // the same values into the same bins again
  unsigned long long cut[2], cov[2] = {0, 0}, sum[2] = {0, 0};
  uint32_t hd[2] = {0, 0};
  int dfrom[2] = {0x7FFFFFFF, 0x7FFFFFFF};
  cut[0] = robust_cut(d.robust, s_tot[2]);
  cut[1] = robust_cut(d.robust, s_tot[3]);
  fold_bins<KS>(cut, cov, sum, hd, dfrom, 1, [&](int i) { return s_h[i]; }, d.hist[0], d.hist[1],
                d.hist_len);
  if (threadIdx.x != 0) return;
  const bool done = (cut[0] == 0 || cov[0] >= cut[0]) && (cut[1] == 0 || cov[1] >= cut[1]);
  const bool empty = (s_tot[2] && !s_tot[0]) || (s_tot[3] && !s_tot[1]);
  if (!done && !empty) {  // HD beyond eight steps: the region launch computes it
    publish(true, false, 0, nullptr, nullptr);
    return;
  }
  const PairOut o = finalize_from(d, s_tot[0], s_tot[1], s_tot[2], s_tot[3], cov, sum, hd, dfrom,
                                  0, 1, err, KS);
  if (trm) trm[3] = wall_clock64();
  publish(false, err, total, trm, &o);
}

// ---------------------------------------------------------------- near pass, batches
// The same near pass over many pairs (hdm_batch_*, 64x8x8 tiles, 6 element),
// as three kernels between the pack (which sets each pair's bitmap) and the
// seed: near_list_kernel compacts the set bits of every pair's bitmap into one
// list of (pair, bit) entries; near_batch_kernel steps each entry in its
// window (near_window) and adds the per-step hits into the pair's spread
// slots; near_fold_kernel folds each pair's bins 1 .. 8 with the robust cut
// (hello.cu:1748-1759).  A pair whose two cuts are reached within eight steps
// is finished there (near_done, histogram bins, counters); the seed, the
// dilation and its fold leave it alone and compute the others from the start.

// One thread per bitmap word of all pairs (pair p's words start at
// idx[p].mark_base, its descriptor's marks pointing at them); each workgroup
// reserves room for its entries with one atomic add.  An entry is
// (pair << 32) | bit, bit = m * regions + rg.
__global__ __launch_bounds__(kWgThreads) void near_list_kernel(const PairDesc *__restrict__ descs,
                                                               const PackIndex *__restrict__ idx,
                                                               int npairs, uint64_t total_words,
                                                               uint64_t *list, uint32_t *count,
                                                               uint32_t cap) {
  __shared__ uint32_t s_wsum[kWavesPerWg];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t w0 = (uint64_t)blockIdx.x * kWgThreads; w0 < total_words;
       w0 += (uint64_t)gridDim.x * kWgThreads) {
    const uint64_t g = w0 + threadIdx.x;
    const bool in = g < total_words;
    int p = 0;
    uint32_t bits = 0, lw = 0;
    if (in) {
      p = find_pair(npairs, g, [&](int i) { return idx[i].mark_base; });
      lw = (uint32_t)(g - idx[p].mark_base);
      bits = descs[p].marks[lw];
      const uint32_t nbits = 2 * pair_regions(descs[p]);  // bits past the pair's end stay clear
      if (lw * 32 + 32 > nbits) bits &= (1u << (nbits - lw * 32)) - 1u;
    }
    const uint32_t n = (uint32_t)__popc(bits);
    const uint32_t incl = wave_scan32(n);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t before = incl - n, tot = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerWg; ++k) {
      if (k < wave) before += s_wsum[k];
      tot += s_wsum[k];
    }
    if (threadIdx.x == 0) s_base = tot ? atomicAdd(count, tot) : 0u;
    __syncthreads();
    uint32_t pos = s_base + before;
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      bits &= bits - 1;
      if (pos < cap) list[pos] = ((uint64_t)(uint32_t)p << 32) | (lw * 32 + b);
      ++pos;
    }
    __syncthreads();  // s_wsum and s_base are free for the next words
  }
}

// Entries taken round robin by the workgroups, one whole workgroup per entry.
// An entry past the list's capacity was never written (the host sizes the list
// for every bit of every bitmap, so none is).
__global__ __launch_bounds__(kWgThreads)
__attribute__((amdgpu_waves_per_eu(2))) void near_batch_kernel(const PairDesc *__restrict__ descs,
                                                              PairCtr *ctrs, const uint64_t *list,
                                                              const uint32_t *count, uint32_t cap) {
  __shared__ NearLds s_near;
  __shared__ uint64_t s_ent;
  if (threadIdx.x < kTbSteps) s_near.hs[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_near.need = 0;
  const uint32_t n = min(*count, cap);
  for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
    if (threadIdx.x == 0) s_ent = list[e];
    __syncthreads();
    const uint64_t ent = s_ent;
    const int p = (int)(ent >> 32);
    const uint32_t bit = (uint32_t)ent;
    const PairDesc &d = descs[p];
    const uint32_t nreg = pair_regions(d);
    const int m = bit >= nreg ? 1 : 0;
    near_window(d, bit - (m ? nreg : 0u), m, s_near);
    if (threadIdx.x < kTbSteps) {
      const uint32_t h = s_near.hs[threadIdx.x];
      if (h) atomicAdd(&ctrs[p].near_hits[m + 2 * threadIdx.x][blockIdx.x & 3], (unsigned long long)h);
      s_near.hs[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_near.need = 0;
    __syncthreads();  // counts cleared and s_ent read before the next entry
  }
}

// One wave per pair: the bins 1 .. 8 of both passes, folded in order up to
// the cut; a pair whose two cuts are reached is finished (its histogram bins,
// covered counts, sums of first hits, last hits and stop steps, and
// near_done); the others are left to the seed and the dilation.  A pair with
// an empty mask facing a non-empty one is left to them too (they report it).
__global__ __launch_bounds__(kWgThreads) void near_fold_kernel(const PairDesc *__restrict__ descs,
                                                              PairCtr *ctrs, int npairs) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * kWavesPerWg + (threadIdx.x >> 6);
  if (p >= npairs) return;
  PairCtr &c = ctrs[p];
  const PairDesc &d = descs[p];
  unsigned long long h = 0;
  if (lane < 2 * kTbSteps * 4) h = (&c.near_hits[0][0])[lane];
  // lanes 4i .. 4i + 3 hold the slots of index i (m + 2k): sum them
  h += dpp64<0x111, 0xF>(h);  // row_shr:1
  h += dpp64<0x112, 0xF>(h);  // row_shr:2
  unsigned long long bin[2 * kTbSteps];
#pragma unroll
  for (int i = 0; i < 2 * kTbSteps; ++i)
    bin[i] = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)h, 4 * i + 3) |
             ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(h >> 32), 4 * i + 3) << 32);
  if (lane != 0) return;
  const unsigned long long nf[2] = {c.n_fp, c.n_fn};
  if ((c.n_fp && !c.n_gold) || (c.n_fn && !c.n_segm)) return;
  unsigned long long cut[2], cov[2] = {0, 0}, sum[2] = {0, 0};
  uint32_t hd[2] = {0, 0}, taken[2] = {0, 0};
  int dfrom[2] = {0x7FFFFFFF, 0x7FFFFFFF};
  for (int m = 0; m < 2; ++m) {
    cut[m] = robust_cut(d.robust, nf[m]);
    for (int k = 0; k < kTbSteps; ++k) {
      const int it = k + 1;
      const unsigned long long hm = bin[m + 2 * k];
      if (cut[m] == 0 || dfrom[m] < it || hm == 0) continue;
      taken[m] |= 1u << k;
      cov[m] += hm;
      sum[m] += hm * (unsigned long long)it;
      hd[m] = (uint32_t)it;
      if (cov[m] >= cut[m]) dfrom[m] = it;
    }
    if (!(cut[m] == 0 || cov[m] >= cut[m])) return;  // beyond eight steps
  }
  for (int m = 0; m < 2; ++m) {
    for (int k = 0; k < kTbSteps; ++k)
      if (((taken[m] >> k) & 1u) && k + 1 < d.hist_len) d.hist[m][k + 1] = (uint32_t)bin[m + 2 * k];
    c.cov[m] = cov[m];
    c.cov2[m] = 0;
    c.sum_d[m] = sum[m];
    c.hd_last[m] = hd[m];
    c.done_from[m] = dfrom[m];
  }
  c.near_done = 1;
}

// ---------------------------------------------------------------- launchers
template <int TXB, int TY, int TZ, bool SKIP = false>
hipError_t launch_pack(dim3 grid, hipStream_t s, const PairDesc *d, const PackIndex *ix,
                       PairCtr *c, int np, uint64_t tw, uint32_t *grab) {
  hipLaunchKernelGGL((pack_kernel<TXB, TY, TZ, SKIP>), grid, dim3(kWgThreads), 0, s, d, ix, c, np,
                     tw, grab);
  return hipGetLastError();
}
// the rows a skipping batch pack left for the pairs its near pass leaves (64x8x8)
hipError_t launch_fill(dim3 grid, hipStream_t s, const PairDesc *d, const PairCtr *c, int np) {
  hipLaunchKernelGGL(fill_kernel, grid, dim3(kWgThreads), 0, s, d, c, np);
  return hipGetLastError();
}
template <int TXB, int TY, int TZ>
hipError_t launch_combine(dim3 grid, hipStream_t s, const PairDesc *d, PairCtr *c,
                          const uint32_t *gf, const uint32_t *sf, uint64_t tiles) {
  hipLaunchKernelGGL((combine_kernel<TXB, TY, TZ>), grid, dim3(kWgThreads), 0, s, d, c, gf, sf,
                     tiles);
  return hipGetLastError();
}
template <int TXB, int TY, int TZ, int CONN, bool DS>
hipError_t launch_seed(dim3 grid, hipStream_t s, const PairDesc *d, const PackIndex *ix,
                       PairCtr *c, int np, Engine *es, uint64_t *q, uint32_t qcap, uint64_t tt,
                       int p0) {
  hipLaunchKernelGGL((seed_kernel<TXB, TY, TZ, CONN, DS>), grid, dim3(kWgThreads), 0, s, d, ix,
                     c, np, es, q, qcap, tt, p0);
  return hipGetLastError();
}
template <int TXB, int TY, int TZ, int CONN, bool DS>
hipError_t launch_dilate(dim3 grid, hipStream_t s, const DilateArgs &a) {
  if (a.npairs == 1)
    hipLaunchKernelGGL((dilate_kernel<TXB, TY, TZ, true, false, CONN, DS>), grid,
                       dim3(kWgThreads), 0, s, a);
  else
    hipLaunchKernelGGL((dilate_kernel<TXB, TY, TZ, false, false, CONN, DS>), grid,
                       dim3(kWgThreads), 0, s, a);
  return hipGetLastError();
}

// The fused single-pair launch, plain or cooperative (a.coop): a cooperative
// launch of a grid larger than can be resident at once fails with
// hipErrorCooperativeLaunchTooLarge instead of starting.
template <typename K>
hipError_t launch_persistent(K kernel, dim3 grid, hipStream_t s, const DilateArgs &a) {
  if (a.coop) {
    DilateArgs copy = a;
    void *args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kernel), grid,
                                      dim3(kWgThreads), args, 0, s);
  }
  hipLaunchKernelGGL(kernel, grid, dim3(kWgThreads), 0, s, a);
  return hipGetLastError();
}
template <int TXB, int TY, int TZ, int CONN>
hipError_t launch_fused(dim3 grid, hipStream_t s, const DilateArgs &a) {
  return launch_persistent(dilate_kernel<TXB, TY, TZ, true, true, CONN>, grid, s, a);
}
// two dilation steps per barrier (64x8x8, 6 element)
hipError_t launch_fused_ds(dim3 grid, hipStream_t s, const DilateArgs &a) {
  return launch_persistent(dilate_ds_kernel, grid, s, a);
}
// regions of 1 x 2 x 2 tiles, eight dilation steps per barrier (64x8x8, 6 element)
hipError_t launch_fused_tb(dim3 grid, hipStream_t s, const DilateArgs &a) {
  return launch_persistent(dilate_tb_kernel, grid, s, a);
}
// pack and near pass (64x8x8, 6 element)
hipError_t launch_fused_near(dim3 grid, hipStream_t s, const DilateArgs &a) {
  return launch_persistent(dilate_near_kernel, grid, s, a);
}

// DS: every launch of the entry runs two dilation steps per engine iteration
// (its fused launch is dilate_ds_kernel)
template <int TXB, int TY, int TZ, int CONN = 6, bool DS = false>
TileLaunch make_launch() {
  constexpr bool kDS = TXB == 64 && TY == 8 && TZ == 8 && CONN == 6;
  static_assert(!DS || kDS, "two steps per barrier: 64x8x8 tiles, 6 element");
  return TileLaunch{&launch_pack<TXB, TY, TZ>,
                    &launch_seed<TXB, TY, TZ, CONN, DS>,
                    &launch_dilate<TXB, TY, TZ, CONN, DS>,
                    DS ? &launch_fused_ds : &launch_fused<TXB, TY, TZ, CONN>,
                    kDS ? &launch_fused_ds : nullptr,
                    kDS ? &launch_fused_tb : nullptr,
                    kDS ? &launch_fused_near : nullptr,
                    (const void *)&dilate_kernel<TXB, TY, TZ, false, false, CONN, DS>,
                    (const void *)&dilate_kernel<TXB, TY, TZ, true, false, CONN, DS>,
                    DS ? (const void *)&dilate_ds_kernel
                       : (const void *)&dilate_kernel<TXB, TY, TZ, true, true, CONN>,
                    kDS ? (const void *)&dilate_ds_kernel : nullptr,
                    kDS ? (const void *)&dilate_tb_kernel : nullptr,
                    kDS ? (const void *)&dilate_near_kernel : nullptr,
                    (const void *)&pack_kernel<TXB, TY, TZ>,
                    &launch_combine<TXB, TY, TZ>,
                    kDS ? &launch_pack<64, 8, 8, true> : nullptr,
                    kDS ? &launch_fill : nullptr,
                    TXB,
                    TY,
                    TZ,
                    CONN,
                    DS ? 2 : 1};
}

// ---------------------------------------------------------------- server admission groups
// The outcome of a server admission group's near pass, one thread per pair:
// a pair the near pass finished (or one with an empty mask) gets its
// results, the others status 2 (they join the engine); then the ready
// marker behind a system-scope fence, so the host takes the outcomes
// without waiting for the stream.
__global__ __launch_bounds__(kWgThreads) void group_out_kernel(const PairDesc *__restrict__ descs,
                                                               const PairCtr *__restrict__ ctrs,
                                                               int n, PairOut *outs,
                                                               uint32_t *ready, uint32_t seq) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const PairDesc &d = descs[k];
    const PairCtr &c = ctrs[k];
    const bool empty = (c.n_fp && !c.n_gold) || (c.n_fn && !c.n_segm);
    PairOut o{};
    if (ld_sc(&c.near_done) != 0 || empty)
      o = finalize_pair(d, c, d.t_start + kTbSteps, false, 1);
    else
      o.status = 2;
    outs[k] = o;
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(ready, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct MoveList {
  int n;
  uint16_t src[kSlotListMax], dst[kSlotListMax];
  uint32_t gen[kSlotListMax];
};
// Pairs of an admission group into engine slots, one workgroup per pair:
// the counters word by word, the descriptor with the slot's generation and
// no start iteration yet (the join sets it), a one-pair pack index.
__global__ __launch_bounds__(kWgThreads) void move_pairs_kernel(const PairDesc *__restrict__ gd,
                                                                const PairCtr *__restrict__ gc,
                                                                PairDesc *descs, PairCtr *ctrs,
                                                                PackIndex *idx, MoveList l) {
  const int i = blockIdx.x;
  if (i >= l.n) return;
  const int k = l.src[i], p = l.dst[i];
  const uint32_t *src = (const uint32_t *)&gc[k];
  uint32_t *dst = (uint32_t *)&ctrs[p];
  static_assert(sizeof(PairCtr) % 4 == 0, "counters copied as words");
  for (uint32_t w = threadIdx.x; w < sizeof(PairCtr) / 4; w += blockDim.x) dst[w] = src[w];
  if (threadIdx.x == 0) {
    PairDesc d = gd[k];
    d.gen = l.gen[i];
    d.t_start = 0xFFFFFFFFu;
    descs[p] = d;
    idx[p] = PackIndex{0, 0, 0};
  }
}

}  // namespace hdm

// table indexed by launch id: HDM_TILE_* (1..4) with the 6 element, then the
// 64x8x8 tile with the 18 (5) and the 26 (6) element, and the 64x8x8 tile with
// the 6 element at two dilation steps per engine iteration (7)
extern "C" const hdm::TileLaunch *hdm_tile_launch(int kind) {
  static const hdm::TileLaunch table[hdm::kLaunchIds] = {
      hdm::make_launch<64, 8, 8>(),     hdm::make_launch<64, 8, 8>(),
      hdm::make_launch<64, 4, 4>(),     hdm::make_launch<64, 8, 4>(),
      hdm::make_launch<32, 8, 8>(),     hdm::make_launch<64, 8, 8, 18>(),
      hdm::make_launch<64, 8, 8, 26>(), hdm::make_launch<64, 8, 8, 6, true>()};
  if (kind < 0 || kind >= hdm::kLaunchIds) return nullptr;
  return &table[kind];
}

extern "C" hipError_t hdm_launch_reset(hipStream_t s, hdm::PairDesc *descs,
                                       const hdm::PairDesc *single,
                                       const hdm::PackIndex *single_idx, hdm::PackIndex *idx,
                                       hdm::PairCtr *ctrs, int p0, int npairs, hdm::Engine *es,
                                       int reset_engine, uint32_t qcap) {
  hdm::PairDesc one{};
  hdm::PackIndex one_idx{};
  if (single) one = *single;
  if (single_idx) one_idx = *single_idx;
  // at least 16 blocks: the engine reset (about 140 KiB) is spread over them
  const int blocks = std::max(16, std::min(npairs, 1024));
  hipLaunchKernelGGL(hdm::reset_kernel, dim3(blocks), dim3(hdm::kWgThreads), 0, s, descs, one,
                     one_idx, idx, single ? 1 : 0, ctrs, p0, npairs, es, reset_engine, qcap);
  return hipGetLastError();
}

// The batched near pass, between the totals and the seed: the entry list
// (Engine::near_count entries, at most cap, into list), the entries on `grid`
// workgroups, and the fold of every pair.
extern "C" hipError_t hdm_launch_near_batch(hipStream_t s, const hdm::PairDesc *descs,
                                            const hdm::PackIndex *idx, hdm::PairCtr *ctrs,
                                            int npairs, uint64_t total_words, uint32_t *count,
                                            uint64_t *list, uint32_t cap, int grid) {
  const uint64_t lb = std::min<uint64_t>(4096, (total_words + hdm::kWgThreads - 1) / hdm::kWgThreads);
  hipLaunchKernelGGL(hdm::near_list_kernel, dim3((unsigned)std::max<uint64_t>(1, lb)),
                     dim3(hdm::kWgThreads), 0, s, descs, idx, npairs, total_words, list, count,
                     cap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hdm::near_batch_kernel, dim3(std::max(1, grid)), dim3(hdm::kWgThreads), 0, s,
                     descs, ctrs, list, count, cap);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(hdm::near_fold_kernel,
                     dim3((npairs + hdm::kWavesPerWg - 1) / hdm::kWavesPerWg),
                     dim3(hdm::kWgThreads), 0, s, descs, ctrs, npairs);
  return hipGetLastError();
}

extern "C" hipError_t hdm_launch_totals(hipStream_t s, hdm::PairCtr *ctrs, int p0, int npairs) {
  hipLaunchKernelGGL(hdm::totals_kernel, dim3(npairs), dim3(64), 0, s, ctrs, p0, npairs);
  return hipGetLastError();
}

extern "C" hipError_t hdm_launch_set_start(hipStream_t s, hdm::PairDesc *descs, const int *slots,
                                           int n, uint32_t t_start) {
  for (int i = 0; i < n; i += hdm::kSlotListMax) {
    hdm::SlotList l{};
    l.n = std::min(n - i, hdm::kSlotListMax);
    for (int k = 0; k < l.n; ++k) l.slot[k] = (uint16_t)slots[i + k];
    hipLaunchKernelGGL(hdm::set_start_kernel, dim3(1), dim3(hdm::kSlotListMax), 0, s, descs, l,
                       t_start);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

extern "C" hipError_t hdm_launch_group_out(hipStream_t s, const hdm::PairDesc *descs,
                                           const hdm::PairCtr *ctrs, int n, hdm::PairOut *outs,
                                           uint32_t *ready, uint32_t seq) {
  hipLaunchKernelGGL(hdm::group_out_kernel, dim3(1), dim3(hdm::kWgThreads), 0, s, descs, ctrs, n,
                     outs, ready, seq);
  return hipGetLastError();
}

extern "C" hipError_t hdm_launch_move_pairs(hipStream_t s, const hdm::PairDesc *gdescs,
                                            const hdm::PairCtr *gctrs, hdm::PairDesc *descs,
                                            hdm::PairCtr *ctrs, hdm::PackIndex *idx,
                                            const int *src, const int *dst, const uint32_t *gen,
                                            int n) {
  for (int i = 0; i < n; i += hdm::kSlotListMax) {
    hdm::MoveList l{};
    l.n = std::min(n - i, hdm::kSlotListMax);
    for (int k = 0; k < l.n; ++k)
      l.src[k] = (uint16_t)src[i + k], l.dst[k] = (uint16_t)dst[i + k], l.gen[k] = gen[i + k];
    hipLaunchKernelGGL(hdm::move_pairs_kernel, dim3(l.n), dim3(hdm::kWgThreads), 0, s, gdescs,
                       gctrs, descs, ctrs, idx, l);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

extern "C" hipError_t hdm_launch_combine_group(hipStream_t s, const hdm::PairDesc *descs,
                                               const hdm::PackIndex *idx, hdm::PairCtr *ctrs,
                                               int npairs, uint64_t total_tiles,
                                               const uint32_t *const *gfs,
                                               const uint32_t *const *sfs, int grid) {
  hipLaunchKernelGGL(hdm::combine_group_kernel, dim3(std::max(1, grid)), dim3(hdm::kWgThreads), 0,
                     s, descs, idx, ctrs, npairs, total_tiles, gfs, sfs);
  return hipGetLastError();
}
