/*
 * hdm.h - C ABI of the MI355X Hausdorff-distance engine (libhdm.so).
 *
 * Two 3-D masks plus their dims go in; the exact Hausdorff distance (HD) and
 * the exact average HD (64-bit integer sum and count, divided in FP64 on the
 * host) come out.  Morphological dilation, iterated until each mask covers
 * the other, as in the reference engine: the 6-neighbour cross by default,
 * or the 18- or 26-neighbour element (hdm_opts.connectivity).  The HDF5
 * loader is declared in hdm_h5.h.
 *
 * Conventions (all entry points):
 *   - volumes are dense, x fastest: index = x + nx*(y + ny*z)
 *     (reference: hello.cu:453, 615; forBench/Volume.h:51-54);
 *   - a voxel is "on" iff mask[v] == label (hello.cu:453-454, 615-616);
 *   - FP = segm & ~gold, FN = gold & ~segm (hello.cu:624-625);
 *   - first-hit d(v) of an FP voxel = least t >= 1 with v in D^t(gold),
 *     of an FN voxel = least t with v in D^t(segm) (hello.cu:1590-1626, 1618);
 *   - HD = max d(v); average HD = sum d(v) / |FP u FN|.
 *   - return value 0 = HDM_OK, negative = HDM_E_* (the library never exits
 *     the process and never resets the device).
 *   - device pointers are plain HIP device addresses; streams are hipStream_t
 *     passed as void*.  NULL stream = the legacy default stream (ordered with
 *     every blocking stream).  A single-pair call returns once its result is
 *     in host memory (the kernel stores it, then a ready marker): its launch
 *     may still be leaving, so work queued after it on the same stream runs
 *     after it, and the library itself waits for it before using another
 *     stream or replacing a buffer.  The inputs are no longer read by then.
 *     Calls that copy a map or histograms to the host, or time every phase,
 *     return after their stream is done.
 *
 * Differences from the reference that are deliberate (its quirks are not
 * reproduced; results follow the mathematical definitions above):
 *   identical masks give HD = 0 (reference: 1, hello.cu:1893);
 *   an empty mask with a non-empty partner returns HDM_E_EMPTY_MASK
 *   (reference: hangs, hello.cu:1887); dims need not be tile multiples
 *   (reference floors them, hello.cu:921-923); robust termination compares
 *   integers, cut = ceil(r * count) (reference: float product, hello.cu:1751).
 */
#ifndef HDM_H
#define HDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDM_ABI_VERSION 1

/* ---- error codes ------------------------------------------------------- */
#define HDM_OK 0
#define HDM_E_INVALID (-1)    /* bad argument (null pointer, dims, label type) */
#define HDM_E_EMPTY_MASK (-2) /* one mask empty while the other has voxels     */
#define HDM_E_HIP (-3)        /* a HIP runtime call failed                      */
#define HDM_E_NOMEM (-4)      /* device allocation failed                       */
#define HDM_E_TIMEOUT (-5)    /* device grid barrier exceeded its spin limit    */
#define HDM_E_BUDGET (-6)     /* request larger than the server's HBM budget    */
#define HDM_E_NODEVICE (-7)   /* no HIP device visible                          */
#define HDM_E_STUCK (-8)      /* work queue drained with targets uncovered      */
#define HDM_E_CAPACITY (-9)   /* a fixed-size table is full                     */

/* ---- option enums ------------------------------------------------------ */
#define HDM_PTR_DEVICE 0 /* gold/segm/map are device pointers           */
#define HDM_PTR_HOST 1   /* gold/segm/map are host pointers (copied)    */

#define HDM_DTYPE_U8 0
#define HDM_DTYPE_I32 1

/* Tile (block) shapes, x * y * z voxels.  One wavefront lane owns one x-row
 * of the tile (64 or 32 bits).  Packed layout is tile-major. */
#define HDM_TILE_AUTO 0
#define HDM_TILE_64x8x8 1
#define HDM_TILE_64x4x4 2
#define HDM_TILE_64x8x4 3
#define HDM_TILE_32x8x8 4
#define HDM_TILE_COUNT 5

typedef struct hdm_opts {
  double robustness;  /* (0,1]; 1 = exact HD.  cut = ceil(r * |FP|), ceil(r * |FN|) */
  int32_t ptr_kind;   /* HDM_PTR_DEVICE or HDM_PTR_HOST                               */
  int32_t tile;       /* HDM_TILE_*                                                    */
  int32_t grid_ctas;  /* persistent-kernel workgroups; 0 = occupancy-derived           */
  int32_t chunk;      /* queue entries a wave takes per claim; 0 = default             */
  int32_t crop;       /* 1 (default) = restrict dilation to the tile bounding box      */
  int32_t timing;     /* HIP-event timing (hdm_last_timings): 1 every phase, 2 also a
                         per-iteration trace (hdm_last_trace), 3 the dominant launch
                         only (hdm_last_timings) */
  void *stream;       /* hipStream_t or NULL                                           */
  uint16_t *map_out;  /* optional nx*ny*nz first-hit map: d(v) on FP u FN, 0 elsewhere */
  uint32_t *hist_out; /* optional 2*hist_len first-hit histograms (FP then FN), host   */
  int32_t hist_len;   /* length of each histogram in hist_out                          */
  int32_t connectivity; /* structuring element of the dilation: 0 or 6 the 6-neighbour
                           cross (hello.cu:373-376, 1310-1542); 18 the element of the
                           reference's second engine with CHEBYSHEV 1 (the six faces,
                           the four xy-plane diagonals and the eight corners,
                           forBench/HausdorffDistance.cu:37-60); 26 the 3x3x3 cube
                           (Chebyshev distance).  18 and 26 need tile AUTO or 64x8x8 */
} hdm_opts;

typedef struct hdm_result {
  int32_t hd;         /* Hausdorff distance (robust HD when robustness < 1)         */
  int32_t hd_fp;      /* max first-hit over FP (gold dilation) up to its cut        */
  int32_t hd_fn;      /* max first-hit over FN (segm dilation) up to its cut        */
  int32_t iterations; /* dilation iterations the device loop executed               */
  uint64_t sum;       /* sum_fp + sum_fn                                             */
  uint64_t count;     /* cnt_fp + cnt_fn                                             */
  uint64_t sum_fp, cnt_fp, sum_fn, cnt_fn;
  uint64_t n_gold, n_segm; /* voxels on in each mask                                */
  uint64_t n_fp, n_fn;     /* |FP|, |FN| before any robust cut                       */
  double avg_hd;      /* (double)sum / (double)count, 0 when count == 0             */
  uint64_t queue_total; /* sum over barrier periods of queued (tile, mask) entries;
                           an entry covers two dilation steps when the launch runs
                           two per barrier (hdm_last_timings ms[6])                 */
  uint32_t queue_max;   /* largest per-iteration queue                              */
  uint32_t tiles;       /* tiles of the (cropped) tile grid                         */
} hdm_result;

/* ---- library ----------------------------------------------------------- */
int hdm_abi_version(void);
const char *hdm_strerror(int code);
void hdm_opts_init(hdm_opts *opts);
/* Number of visible HIP devices (0 when none); never fails. */
int hdm_device_count(void);
/* Device bytes one pair of the given dims needs (workspace, excluding inputs). */
uint64_t hdm_workspace_bytes(int nx, int ny, int nz, int tile);
/* Device milliseconds of the phases of the last single-pair call on the
 * current device that ran with opts->timing = 1: ms[0] pack (+ reset), ms[1]
 * tile states and first queue, ms[2] persistent dilation (which also
 * finalizes), ms[3] 0, ms[4] all kernels; ms[5] is the number of launches of
 * the fused path (the default for a single pair: reset, pack and dilation in
 * one persistent kernel, phases then timed by the kernel's own wall clock
 * between its grid barriers): 1, or 2 when that launch handed the packed pair
 * to the region launch; 0 when it ran as separate kernels (phases timed by
 * HIP events on the stream).  With opts->timing = 3 only the dominant launch
 * is bracketed by a pair of HIP events, the others are -1: ms[2] the dilation
 * kernel of a separate-kernel call, ms[4] the whole fused path; such a call
 * does not wait for its events, this function does.  ms[6] is the number of
 * dilation steps the body that finished ran per grid barrier (8 for the near
 * pass and the region launch of 64x8x8 tiles with the 6 element, 2 for the
 * two-step launch, else 1).
 * Returns how many were written. */
int hdm_last_timings(float *ms, int n);
/* Device milliseconds of the dominant launch of each single-pair call made
 * with opts->timing = 3 on the current device since the log was last read
 * (at most the last 1024), oldest first; waits for their events, then clears
 * the log.  Returns how many were written (at most n). */
int hdm_timing_log(float *ms, int n);
/* Per-iteration device trace of the last single-pair call on the current
 * device that ran with opts->timing = 2 (observability; no counterpart in the
 * reference, which has only printf): 8 words per iteration, wall-clock ticks
 * [0] iteration start, [1] last workgroup done with its tiles, [2] workgroup 0
 * done, [3] every workgroup arrived at the barrier, [4] counters folded,
 * [5] queued (tile, mask) entries.  *tick_ns receives the tick length.
 * Returns the number of iterations written (at most max_iters). */
int hdm_last_trace(uint64_t *out, int max_iters, double *tick_ns);

/* ---- single pair -------------------------------------------------------
 * Replaces: template<T> int getHausdorffDistance_CUDA_Generic(T* goldStandard,
 *   T* algoOutput, int WIDTH, int HEIGHT, int DEPTH, float robustnessPercent,
 *   bool resIterNeeded, T numberToLookFor, bool res3DNeeded)
 *   (hello.cu:2025-2062), for T = uint8_t.  The int it returned is out->hd. */
int hdm_hausdorff_u8(const uint8_t *gold, const uint8_t *segm, int nx, int ny, int nz,
                     uint8_t label, const hdm_opts *opts, hdm_result *out);

/* Same, T = int (the reference instantiates its main pass for int,
 * hello.cu:1999). */
int hdm_hausdorff_i32(const int32_t *gold, const int32_t *segm, int nx, int ny, int nz,
                      int32_t label, const hdm_opts *opts, hdm_result *out);

/* Replaces: int HausdorffDistance::computeDistance(Volume* img1, Volume* img2)
 * (forBench/HausdorffDistance.cuh:17, .cu:101-187) for bool volumes
 * (forBench/Volume.h:10).  A voxel is on iff its byte is 1 (bool storage
 * holds 0 or 1; the reference allocates the volumes with calloc).  With
 * opts->connectivity = 18 the dilation uses that engine's structuring element
 * (CHEBYSHEV 1, forBench/HausdorffDistance.cuh:8, .cu:37-60) and the result's
 * hd is its computeDistance value; the default is the 6-neighbour cross of
 * the reference's main engine. */
int hdm_hausdorff_bool(const uint8_t *img1, const uint8_t *img2, int nx, int ny, int nz,
                       const hdm_opts *opts, hdm_result *out);

/* Replaces: getHausdorffDistance_CUDA_3Dres_local (hello.cu:2178-2241) and
 * getHausdorffDistance_CUDA_FullResList_local (hello.cu:2065-2103): writes
 * d(v) for every FP and FN voxel into map (nx*ny*nz uint16, x fastest),
 * 0 elsewhere; with robustness < 1 only voxels at d <= the cut are kept. */
int hdm_first_hit_map_u8(const uint8_t *gold, const uint8_t *segm, int nx, int ny, int nz,
                         uint8_t label, const hdm_opts *opts, uint16_t *map,
                         hdm_result *out);

/* ---- allocation-free single pair ----------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * or host memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same computation and results as
 * hdm_hausdorff_u8, with its whole workspace in a fixed device array of
 * hdm_noalloc_max_bytes() bytes (64 MiB: volumes up to 512x512x256 with
 * 64x8x8 tiles) that each GPU holds from the first call until the process
 * ends.  It calls no allocating routine.  Restrictions: device pointers only
 * (ptr_kind HDM_PTR_DEVICE), timing 0; map_out, when set, is a device
 * pointer; hist_out is host memory.  Returns HDM_E_CAPACITY when the pair's
 * workspace exceeds the array.  Calls on one GPU share the array: within a
 * process they are serialized (each returns after its stream is done); calls
 * that overlapped on one GPU would race. */
int hdm_hausdorff_u8_noalloc(const uint8_t *gold, const uint8_t *segm, int nx, int ny, int nz,
                             uint8_t label, const hdm_opts *opts, hdm_result *out);
/* Size of the fixed device array of hdm_hausdorff_u8_noalloc, in bytes. */
uint64_t hdm_noalloc_max_bytes(void);

/* ---- batched pairs -----------------------------------------------------
 * No single counterpart in the reference (it has one pair per call,
 * hello.cu:2025): many pairs of uneven size go through one persistent
 * launch; the create call captures pack + metadata + persistent dilation +
 * finalize into one hipGraph, which run() replays. */
typedef struct hdm_pair {
  const void *gold; /* device pointers */
  const void *segm;
  int32_t nx, ny, nz;
  int32_t dtype; /* HDM_DTYPE_* */
  int32_t label;
  int32_t reserved;
} hdm_pair;

typedef struct hdm_batch hdm_batch;

int hdm_batch_create(const hdm_pair *pairs, int n, const hdm_opts *opts, int use_graph,
                     hdm_batch **out);
int hdm_batch_run(hdm_batch *b);                    /* enqueue one run (async) */
int hdm_batch_sync(hdm_batch *b, hdm_result *out);  /* wait; out[n] results    */
/* Device time of the last run, from HIP events around it (ms). */
float hdm_batch_last_ms(hdm_batch *b);
/* (region, mask) entries the near pass of the last synced run stepped (the
 * 64x8x8 tile with the 6 element; pairs whose HD is at most 8 finish there),
 * 0 when it did not run. */
uint32_t hdm_batch_near_entries(hdm_batch *b);
void hdm_batch_destroy(hdm_batch *b);

/* One-shot convenience: create (no graph), run, sync, destroy. */
int hdm_hausdorff_batch(const hdm_pair *pairs, int n, const hdm_opts *opts, hdm_result *out);

/* ---- autotuner ---------------------------------------------------------
 * No counterpart in the reference (it takes launch shapes from the CUDA
 * occupancy API, hello.cu:1909-1954).  Times the batched launch for each
 * (tile, chunk, grid) candidate on the given pairs; best first. */
typedef struct hdm_tune_entry {
  int32_t tile, chunk, grid_ctas, status;
  float ms; /* median device ms of the batched run */
} hdm_tune_entry;

int hdm_autotune(const hdm_pair *pairs, int n, const int32_t *tiles, int n_tiles,
                 const int32_t *chunks, int n_chunks, const int32_t *grids, int n_grids,
                 int reps, hdm_tune_entry *out, int out_cap, int *n_out);

/* ---- serving engine ----------------------------------------------------
 * No counterpart in the reference.  Continuous batching: requests are
 * admitted against an HBM budget, join the running persistent engine at an
 * iteration boundary (each pair keeps its own iteration count), and retire
 * as soon as both of their masks are covered. */
typedef struct hdm_server hdm_server;

typedef struct hdm_server_opts {
  uint64_t hbm_budget;   /* bytes of device memory for inputs + workspace   */
  int32_t max_slots;     /* requests in flight at once (<= 4096)            */
  int32_t slice_iters;   /* dilation steps per persistent launch, rounded up
                            to whole engine iterations (two steps each on
                            64x8x8 tiles with the 6 element)               */
  int32_t tile;          /* HDM_TILE_*                                      */
  int32_t grid_ctas;     /* 0 = occupancy-derived                           */
  double robustness;
  int32_t max_nx, max_ny, max_nz; /* largest request dims (sizes the queue)  */
  int32_t connectivity;  /* structuring element, as hdm_opts.connectivity   */
} hdm_server_opts;

typedef struct hdm_completion {
  uint64_t id;       /* request id from hdm_server_submit        */
  int32_t status;    /* HDM_OK or an error code                  */
  int32_t reserved;
  hdm_result result;
  double t_submit, t_admit, t_done; /* seconds, library clock     */
} hdm_completion;

void hdm_server_opts_init(hdm_server_opts *o);
int hdm_server_create(const hdm_server_opts *o, hdm_server **out);
/* Queue a request.  With ptr_kind HDM_PTR_DEVICE, gold/segm are resident
 * device volumes (no copy), complete when submitted: the server's streams do
 * not wait for the caller's; with HDM_PTR_HOST the server copies them on its
 * copy stream once admitted (pinned host memory lets the copy overlap the
 * engine).  The volumes must stay valid until the request completes.
 * HDM_E_BUDGET if the request alone exceeds the HBM budget, HDM_E_CAPACITY
 * if its dims exceed the server's max dims. */
int hdm_server_submit(hdm_server *s, const void *gold, const void *segm, int nx, int ny,
                      int nz, int dtype, int label, int ptr_kind, uint64_t *id_out);
/* Drive the engine: admit, launch one slice, retire.  Fills up to cap
 * completions.  Returns HDM_OK.  A lone request (nothing in flight, nothing
 * else waiting) completes in the step as one one-pair launch on the server's
 * stream (the path of hdm_hausdorff_*, whose workspace is the library's and
 * not the arena). */
int hdm_server_step(hdm_server *s, hdm_completion *done, int cap, int *n_done);
/* Requests queued, in flight, and device bytes held. */
int hdm_server_load(hdm_server *s, int *queued, int *in_flight, uint64_t *bytes_used);

/* Resident masks.  No counterpart in the reference, which packs both volumes
 * again on every call (hello.cu:537-748, freed at 2001-2012).
 * hdm_server_register packs one volume (the voxels equal to label) with the
 * server's tile and keeps the packed bits, face words and tile flags in the
 * server's arena (charged to its HBM budget).  It returns once the mask is
 * packed; the volume may be freed then.  A request naming two registered
 * masks of equal dims (hdm_server_submit_masks) skips the pack at admission:
 * its admission reads the masks' 2 bits per voxel instead of the volumes.
 * HDM_E_BUDGET if the arena has no room, HDM_E_CAPACITY beyond max dims. */
int hdm_server_register(hdm_server *s, const void *vol, int nx, int ny, int nz, int dtype,
                        int label, int ptr_kind, uint64_t *handle_out);
/* HDM_E_INVALID for unknown handles or unequal dims. */
int hdm_server_submit_masks(hdm_server *s, uint64_t gold, uint64_t segm, uint64_t *id_out);
/* Releases a registered mask; its memory returns to the arena once no queued
 * or in-flight request names it.  HDM_E_INVALID for an unknown handle. */
int hdm_server_unregister(hdm_server *s, uint64_t handle);

typedef struct hdm_server_stats {
  uint64_t slices;            /* slices of the several-pair engine launched     */
  uint64_t pair_slices;       /* sum over those slices of the pairs in flight   */
  int32_t max_in_flight;      /* most pairs in flight in one slice              */
  int32_t reserved;
  uint64_t lone;              /* requests run as one one-pair call              */
  uint64_t admitted;          /* requests admitted through the slots            */
  uint64_t registered_bytes;  /* arena bytes held by registered masks           */
  uint64_t near_finished;     /* requests an admission group's near pass
                                 finished (64x8x8 tiles, 6 element): no slot    */
  uint64_t groups;            /* admission groups run                           */
  uint64_t group_pairs;       /* sum over those groups of their requests        */
} hdm_server_stats;
int hdm_server_stats_get(hdm_server *s, hdm_server_stats *out);
double hdm_server_now(void);
void hdm_server_destroy(hdm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* HDM_H */
