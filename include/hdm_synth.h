/*
 * hdm_synth.h - deterministic synthetic mask pairs for tests and benchmarks.
 *
 * No counterpart in the reference (its harness read a private HDF5 file,
 * forBench/hdf5Manag.cu:196-220).  Every shape test is integer arithmetic, so
 * the host and device generators give the same bytes:
 *
 *   ellipsoid e contains p  iff  |q_k| <= 1024*s_k for k = 0..2 and
 *       sum_k q_k^2 * prod_{j != k} s_j^2  <=  2^20 * prod_j s_j^2
 *   where q = R (p - c), R the rotation in Q10 fixed point (round(1024 r)).
 *
 *   flips: h = splitmix64 output number (v + 1) of the stream seeded with
 *   flip_seed, i.e. mix(flip_seed + (v + 1) * 0x9E3779B97F4A7C15);
 *   voxel v flips iff (h >> 11) < flip_thresh (= p * 2^53).
 *     flip_mode 1: on voxels of segm with an off 6-neighbour (or on the volume
 *                  border) turn off;
 *     flip_mode 2: voxels of segm that differ from an in-volume 6-neighbour
 *                  are toggled.
 */
#ifndef HDM_SYNTH_H
#define HDM_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDM_SYN_MAX_ELL 16

typedef struct hdm_ell {
  int32_t c[3]; /* centre, voxels                 */
  int32_t s[3]; /* semi-axes, voxels, 1..120       */
  int32_t r[9]; /* rotation, Q10, row-major         */
} hdm_ell;

typedef struct hdm_synth_spec {
  int32_t nx, ny, nz;
  int32_t n_gold, n_segm; /* ellipsoids in each mask's union          */
  int32_t flip_mode;      /* 0 none, 1 surface off, 2 boundary toggle */
  uint64_t flip_seed;
  uint64_t flip_thresh;
  int32_t cube_lo[3], cube_hi[3]; /* extra on-box [lo,hi) in segm; empty if lo>=hi */
  hdm_ell gold[HDM_SYN_MAX_ELL];
  hdm_ell segm[HDM_SYN_MAX_ELL];
} hdm_synth_spec;

/* Sphere of radius r at (cx,cy,cz) in gold; the same sphere shifted by
 * (dx,dy,dz) in segm, then a fraction p_flip of segm's surface voxels turned
 * off (flip_mode 1).  64^3, r 20, centre 32, shift (2,1,0), p 0.01 is the
 * small plain workload. */
int hdm_synth_spec_sphere(hdm_synth_spec *s, int nx, int ny, int nz, int cx, int cy, int cz,
                          int r, int dx, int dy, int dz, double p_flip, uint64_t seed);

/* Union of n seeded ellipsoids (semi-axes uniform in [amin,amax], centres
 * uniform in the middle 3/4 of each axis, uniform random rotations) in gold;
 * segm has each ellipsoid's semi-axes changed by an integer uniform in
 * [-dmax, dmax], then a fraction p_flip of its boundary voxels toggled. */
int hdm_synth_spec_ellipsoids(hdm_synth_spec *s, int nx, int ny, int nz, int n, int amin,
                              int amax, int dmax, double p_flip, uint64_t seed);

/* One axis-aligned ellipsoid (ax,ay,az) at (cx,cy,cz) in both masks, and a
 * cube of edge `edge` in segm only, whose near face is `gap` voxels beyond
 * the ellipsoid along +x (so the HD is about gap + edge). */
int hdm_synth_spec_long(hdm_synth_spec *s, int nx, int ny, int nz, int cx, int cy, int cz,
                        int ax, int ay, int az, int gap, int edge);

/* Evaluate on the host into two nx*ny*nz uint8 (0/1) volumes. */
int hdm_synth_host(const hdm_synth_spec *s, uint8_t *gold, uint8_t *segm);

/* Evaluate on the device (gold, segm device pointers), async on stream. */
int hdm_synth_device(const hdm_synth_spec *s, uint8_t *gold, uint8_t *segm, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* HDM_SYNTH_H */
