/*
 * hdm_oracle.c - CPU oracle for the Hausdorff dilation path.
 *
 * TEST INFRASTRUCTURE ONLY (the results tests/, the smoke check and the
 * benchmark's cpu_baseline leg compare against).  Never linked into or called
 * by the product.
 *
 * Follows the reference's algorithm (jakubMitura14/HousdorffHDF5cMake):
 *   - voxel on iff value == label ......................... hello.cu:453, 615
 *   - FP = segm & ~gold, FN = gold & ~segm ................. hello.cu:624-625
 *   - per-mask passes continue while count*r > covered .... hello.cu:1748-1759,
 *     the loop ends when both passes stop ................. hello.cu:1887
 *     (integer cut ceil(r*count) as old/testMainPassesKernels.cu:81-82)
 *   - one iteration dilates every live block with the 6-connectivity cross:
 *     in-row shift-or (bitDilatate) ....................... hello.cu:373-376
 *     plus the face neighbours ............................ hello.cu:1310-1542
 *   - newly covered voxels of the other mask: res & ~src & ref, recorded with
 *     iteration number iterationNumb+1 .................... hello.cu:1590-1626
 *   - blocks whose source is full stop ..................... hello.cu:1319,1634-1643
 *   - blocks whose face bits are set activate neighbours .. hello.cu:1649-1663
 *   - bounding-box crop to non-empty blocks ................ hello.cu:940-973
 *
 * Deliberate differences (the reference's quirks): identical masks give
 * HD 0, not 1 (hello.cu:1893); an empty mask with a non-empty partner is an
 * error (-2), not a hang (hello.cu:1887); tail blocks are masked, not
 * truncated (hello.cu:921-923).
 */
#include "hdm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ORACLE_MUTANT (tests only) plants a known mistake in oracle_tiled so the
 * test suite can show its golden checks catch it: 1 first-hit index off by
 * one, 2 the +z neighbour dropped, 3 the wrong mask as reference, 4 the x
 * shift one-sided.  0 (default) is the oracle. */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

#define E_INVALID (-1)
#define E_EMPTY (-2)
#define E_STUCK (-8)

static inline int is_on(const void *p, int dtype, int32_t label, uint64_t i) {
  if (dtype == ORACLE_DTYPE_U8) return ((const uint8_t *)p)[i] == (uint8_t)label;
  return ((const int32_t *)p)[i] == label;
}

static uint64_t cut_of(double r, uint64_t n) {
  if (n == 0) return 0;
  double c = ceil(r * (double)n);
  if (c < 0) c = 0;
  if (c > (double)n) c = (double)n;
  return (uint64_t)c;
}

void oracle_finish(const uint32_t *hf, const uint32_t *hn, int len, double r,
                   oracle_result *o) {
  const uint32_t *h[2] = {hf, hn};
  uint64_t tot[2] = {o->n_fp, o->n_fn};
  int32_t hd[2] = {0, 0};
  uint64_t sum[2] = {0, 0}, cnt[2] = {0, 0};
  for (int m = 0; m < 2; ++m) {
    uint64_t tgt = cut_of(r, tot[m]);
    if (tgt == 0) continue;
    for (int t = 1; t < len; ++t) {
      cnt[m] += h[m][t];
      sum[m] += (uint64_t)t * h[m][t];
      if (h[m][t]) hd[m] = t;
      if (cnt[m] >= tgt) break;
    }
  }
  o->hd_fp = hd[0], o->hd_fn = hd[1];
  o->hd = hd[0] > hd[1] ? hd[0] : hd[1];
  o->sum_fp = sum[0], o->cnt_fp = cnt[0], o->sum_fn = sum[1], o->cnt_fn = cnt[1];
  o->sum = sum[0] + sum[1];
  o->count = cnt[0] + cnt[1];
  o->avg_hd = o->count ? (double)o->sum / (double)o->count : 0.0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------
 * naive: one byte per voxel
 * ---------------------------------------------------------------------- */

/* Offsets of a structuring element: conn 6 the cross (hello.cu:373-376,
 * 1310-1542); 18 the element of forBench/HausdorffDistance.cu with CHEBYSHEV
 * 1 (HausdorffDistance.cuh:8): the six faces (.cu:37-44), the four xy-plane
 * diagonals (.cu:48-51) and the eight corners (.cu:53-60); 26 the cube.
 * Returns the number of offsets, 0 for an unknown conn. */
static int se_offsets(int conn, int off[26][3]) {
  int n = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int nz_ = (dx != 0) + (dy != 0) + (dz != 0);
        int in = 0;
        if (nz_ == 0) continue;
        if (conn == 6) in = nz_ == 1;
        else if (conn == 18) in = nz_ == 1 || (nz_ == 2 && dz == 0) || nz_ == 3;
        else if (conn == 26) in = 1;
        else return 0;
        if (in) off[n][0] = dx, off[n][1] = dy, off[n][2] = dz, ++n;
      }
  return n;
}

/* dst = src dilated by the element, clipped to the volume: dst[v] = OR over
 * offsets o of src[v + o] (the element is symmetric, so this equals the
 * reference's scatter of every set voxel to v + o). */
static void dilate_se(const uint8_t *src, uint8_t *dst, int nx, int ny, int nz, int n_off,
                      const int (*off)[3]) {
  const uint64_t sy = (uint64_t)nx, sz = (uint64_t)nx * ny;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const uint64_t i = x + sy * y + sz * z;
        uint8_t v = src[i];
        for (int k = 0; k < n_off && !v; ++k) {
          const int xx = x + off[k][0], yy = y + off[k][1], zz = z + off[k][2];
          if (xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
          v = src[xx + sy * yy + sz * zz];
        }
        dst[i] = v;
      }
}

int oracle_naive(const void *gold, const void *segm, int dtype, int32_t label, int nx, int ny,
                 int nz, double r, uint32_t *hist_fp, uint32_t *hist_fn, int hist_len,
                 uint16_t *map, oracle_result *out) {
  return oracle_naive_se(gold, segm, dtype, label, nx, ny, nz, r, 6, hist_fp, hist_fn, hist_len,
                         map, out);
}

int oracle_naive_se(const void *gold, const void *segm, int dtype, int32_t label, int nx,
                    int ny, int nz, double r, int conn, uint32_t *hist_fp, uint32_t *hist_fn,
                    int hist_len, uint16_t *map, oracle_result *out) {
  int off[26][3];
  const int n_off = se_offsets(conn, off);
  if (!gold || !segm || !out || !hist_fp || !hist_fn || nx < 1 || ny < 1 || nz < 1 ||
      !(r >= 0.0 && r <= 1.0) || hist_len < 2 || n_off == 0)
    return E_INVALID;
  memset(out, 0, sizeof(*out));
  const uint64_t n = (uint64_t)nx * ny * nz;
  uint8_t *g = malloc(n), *s = malloc(n), *cg = malloc(n), *cs = malloc(n), *nx_ = malloc(n);
  if (!g || !s || !cg || !cs || !nx_) {
    free(g), free(s), free(cg), free(cs), free(nx_);
    return E_INVALID;
  }
  memset(hist_fp, 0, sizeof(uint32_t) * hist_len);
  memset(hist_fn, 0, sizeof(uint32_t) * hist_len);
  if (map) memset(map, 0, sizeof(uint16_t) * n);
  for (uint64_t i = 0; i < n; ++i) {
    g[i] = (uint8_t)is_on(gold, dtype, label, i);
    s[i] = (uint8_t)is_on(segm, dtype, label, i);
    out->n_gold += g[i];
    out->n_segm += s[i];
    out->n_fp += s[i] && !g[i];
    out->n_fn += g[i] && !s[i];
  }
  int rc = 0;
  if ((out->n_fp && !out->n_gold) || (out->n_fn && !out->n_segm)) {
    rc = E_EMPTY;
    goto done;
  }
  memcpy(cg, g, n);
  memcpy(cs, s, n);
  const uint64_t tgt[2] = {cut_of(r, out->n_fp), cut_of(r, out->n_fn)};
  uint64_t cov[2] = {0, 0};
  int t = 0;
  while (cov[0] < tgt[0] || cov[1] < tgt[1]) {
    ++t;
    if (t >= hist_len) {
      rc = E_STUCK;
      goto done;
    }
    for (int m = 0; m < 2; ++m) {
      if (cov[m] >= tgt[m]) continue;
      uint8_t *cur = m == 0 ? cg : cs;
      const uint8_t *ref = m == 0 ? s : g; /* the other mask's original */
      uint32_t *hist = m == 0 ? hist_fp : hist_fn;
      dilate_se(cur, nx_, nx, ny, nz, n_off, (const int(*)[3])off);
      for (uint64_t i = 0; i < n; ++i) {
        if (nx_[i] && !cur[i] && ref[i]) { /* res & ~src & ref */
          hist[t]++;
          cov[m]++;
          if (map) map[i] = (uint16_t)t;
        }
      }
      memcpy(cur, nx_, n);
    }
  }
  out->iterations = t;
  oracle_finish(hist_fp, hist_fn, hist_len, r, out);
done:
  free(g), free(s), free(cg), free(cs), free(nx_);
  return rc;
}

/* ------------------------------------------------------------------------
 * tiled: the device engine's algorithm on the CPU
 * ---------------------------------------------------------------------- */

enum { ST_IDLE = 0, ST_ACTIVE = 1, ST_FULL = 2 };
#define TYC 8
#define TZC 8

typedef struct {
  int nx, ny, nz, wx;        /* wx: 64-bit words per x-row */
  int gx, gy, gz;            /* tile grid: wx * ceil(ny/8) * ceil(nz/8) */
  int lo[3], hi[3];          /* tile bounding box, inclusive */
  uint64_t last_valid;       /* valid bits of the last word of a row */
} geom_t;

static inline uint64_t valid_bits(const geom_t *G, int wx) {
  return wx == G->wx - 1 ? G->last_valid : ~0ull;
}
static inline uint64_t widx(const geom_t *G, int wx, int y, int z) {
  return (uint64_t)wx + (uint64_t)G->wx * ((uint64_t)y + (uint64_t)G->ny * z);
}
static inline int in_box(const geom_t *G, int tx, int ty, int tz) {
  return tx >= G->lo[0] && tx <= G->hi[0] && ty >= G->lo[1] && ty <= G->hi[1] &&
         tz >= G->lo[2] && tz <= G->hi[2];
}

typedef struct {
  uint32_t *v;
  uint64_t n, cap;
} vec_t;
static void vpush(vec_t *q, uint32_t e) {
  if (q->n == q->cap) {
    q->cap = q->cap ? 2 * q->cap : 256;
    q->v = realloc(q->v, q->cap * sizeof(uint32_t));
  }
  q->v[q->n++] = e;
}

/* word (wx, y, z) of a packed mask, 0 outside the volume */
static inline uint64_t word_at(const geom_t *G, const uint64_t *src, int wx, int y, int z) {
  if (y < 0 || z < 0 || y >= G->ny || z >= G->nz) return 0;
  return src[widx(G, wx, y, z)];
}

/* the x-neighbour bits of word (wx, y, z): w << 1 | w >> 1 plus the carries
 * from the adjacent words of the row (bitDilatate, hello.cu:373-376, and the
 * x faces, hello.cu:1310-1542); 0 outside the volume */
static inline uint64_t x_shift(const geom_t *G, const uint64_t *src, int wx, int y, int z) {
  if (y < 0 || z < 0 || y >= G->ny || z >= G->nz) return 0;
  const uint64_t i = widx(G, wx, y, z);
  const uint64_t w = src[i];
#if ORACLE_MUTANT == 4
  uint64_t d = w << 1;
#else
  uint64_t d = (w << 1) | (w >> 1);
#endif
  if (wx > 0) d |= src[i - 1] >> 63;
  if (wx < G->wx - 1) d |= src[i + 1] << 63;
  return d;
}

/* the element's dilation of word (tx, y, z):
 *   6  w | x(w) | y-1 | y+1 | z-1 | z+1
 *   18 (the forBench element) the faces, the xy diagonals (x of the y-1 and
 *      y+1 rows) and the corners (x of the four rows at y+-1, z+-1)
 *   26 every row of the 3x3 (y, z) neighbourhood with its x neighbours */
static inline uint64_t dilate_word(const geom_t *G, const uint64_t *src, int conn, int tx, int y,
                                   int z) {
  uint64_t d = word_at(G, src, tx, y, z) | x_shift(G, src, tx, y, z);
  if (conn == 26) {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        if (dy || dz) d |= word_at(G, src, tx, y + dy, z + dz) | x_shift(G, src, tx, y + dy, z + dz);
    return d;
  }
  d |= word_at(G, src, tx, y - 1, z) | word_at(G, src, tx, y + 1, z);
  d |= word_at(G, src, tx, y, z - 1);
#if ORACLE_MUTANT != 2
  d |= word_at(G, src, tx, y, z + 1);
#endif
  if (conn == 18) {
    d |= x_shift(G, src, tx, y - 1, z) | x_shift(G, src, tx, y + 1, z);
    for (int dz = -1; dz <= 1; dz += 2)
      for (int dy = -1; dy <= 1; dy += 2) d |= x_shift(G, src, tx, y + dy, z + dz);
  }
  return d;
}

/* Bit j = (ox+1) + 3(oy+1) + 9(oz+1) of the result is set when word d of row
 * (y, z) of tile (y0, z0) has a voxel on every tile face that offset (ox, oy,
 * oz) crosses: the neighbour tile at that offset can be reached from it in
 * one step of the 26 element (a superset of the 18 element's reach). */
static uint32_t reach_bits(uint64_t d, int y, int z, int y0, int z0) {
  if (!d) return 0;
  const int xs = 2 | (int)(d & 1) | (int)((d >> 63) << 2);
  const int ys = 2 | (y == y0) | ((y == y0 + TYC - 1) << 2);
  const int zs = 2 | (z == z0) | ((z == z0 + TZC - 1) << 2);
  uint32_t m = 0;
  for (int iz = 0; iz < 3; ++iz)
    for (int iy = 0; iy < 3; ++iy)
      for (int ix = 0; ix < 3; ++ix)
        if (((xs >> ix) & 1) && ((ys >> iy) & 1) && ((zs >> iz) & 1)) m |= 1u << (ix + 3 * iy + 9 * iz);
  return m & ~(1u << 13);
}

/* Dilate one tile of mask m from src into dst.  Returns the input-full flag;
 * faces[6] receives whether the output has bits on -x,+x,-y,+y,-z,+z and,
 * for conn 18 / 26, *reach the neighbours the output reaches (reach_bits). */
static int tile_step(const geom_t *G, const uint64_t *src, uint64_t *dst, const uint64_t *ref,
                     int conn, int tx, int ty, int tz, int t, uint16_t *map, uint64_t *cnt,
                     int faces[6], uint32_t *reach) {
  int full = 1;
  const int y0 = ty * TYC, y1 = (y0 + TYC < G->ny) ? y0 + TYC : G->ny;
  const int z0 = tz * TZC, z1 = (z0 + TZC < G->nz) ? z0 + TZC : G->nz;
  const uint64_t vb = valid_bits(G, tx);
  uint64_t fxl = 0, fxh = 0, fyl = 0, fyh = 0, fzl = 0, fzh = 0, c = 0;
  uint32_t rb = 0;
  for (int z = z0; z < z1; ++z)
    for (int y = y0; y < y1; ++y) {
      const uint64_t i = widx(G, tx, y, z);
      const uint64_t w = src[i];
      uint64_t d = dilate_word(G, src, conn, tx, y, z) & vb;
      if (w != vb) full = 0;
      const uint64_t nw = d & ~w & ref[i];
      if (nw) {
        c += (uint64_t)__builtin_popcountll(nw);
        if (map) {
          uint64_t b = nw;
          while (b) {
            const int k = __builtin_ctzll(b);
            b &= b - 1;
            const uint64_t x = (uint64_t)tx * 64 + k;
            map[x + (uint64_t)G->nx * ((uint64_t)y + (uint64_t)G->ny * z)] = (uint16_t)(t + 1);
          }
        }
      }
      dst[i] = d;
      fxl |= d & 1ull;
      fxh |= (d >> 63) & 1ull;
      if (y == y0) fyl |= d;
      if (y == y1 - 1) fyh |= d;
      if (z == z0) fzl |= d;
      if (z == z1 - 1) fzh |= d;
      if (conn != 6) rb |= reach_bits(d, y, z, y0, z0);
    }
  faces[0] = fxl != 0, faces[1] = fxh != 0, faces[2] = fyl != 0;
  faces[3] = fyh != 0, faces[4] = fzl != 0, faces[5] = fzh != 0;
  *cnt = c;
  *reach = rb;
  return full;
}

/* claims neighbour (ux, uy, uz) of mask m for the next queue when it is in
 * the grid, in the box and idle */
static inline void activate(const geom_t *G, uint8_t *state, int ux, int uy, int uz,
                            vec_t *out) {
  if (ux < 0 || uy < 0 || uz < 0 || ux >= G->gx || uy >= G->gy || uz >= G->gz) return;
  if (!in_box(G, ux, uy, uz)) return;
  const uint64_t ul = ux + (uint64_t)G->gx * (uy + (uint64_t)G->gy * uz);
  uint8_t expect = ST_IDLE;
  if (__atomic_compare_exchange_n(&state[ul], &expect, ST_ACTIVE, 0, __ATOMIC_RELAXED,
                                  __ATOMIC_RELAXED))
    vpush(out, (uint32_t)ul);
}

int oracle_tiled(const void *gold, const void *segm, int dtype, int32_t label, int nx, int ny,
                 int nz, double r, int threads, int crop, uint32_t *hist_fp, uint32_t *hist_fn,
                 int hist_len, uint16_t *map, oracle_result *out) {
  return oracle_tiled_se(gold, segm, dtype, label, nx, ny, nz, r, 6, threads, crop, hist_fp,
                         hist_fn, hist_len, map, out);
}

int oracle_tiled_se(const void *gold, const void *segm, int dtype, int32_t label, int nx,
                    int ny, int nz, double r, int conn, int threads, int crop,
                    uint32_t *hist_fp, uint32_t *hist_fn, int hist_len, uint16_t *map,
                    oracle_result *out) {
  if (!gold || !segm || !out || !hist_fp || !hist_fn || nx < 1 || ny < 1 || nz < 1 ||
      !(r >= 0.0 && r <= 1.0) || hist_len < 2 || (conn != 6 && conn != 18 && conn != 26))
    return E_INVALID;
  if (threads > 0) omp_set_num_threads(threads);
  memset(out, 0, sizeof(*out));
  geom_t G;
  G.nx = nx, G.ny = ny, G.nz = nz;
  G.wx = (nx + 63) / 64;
  G.last_valid = (nx % 64) ? ((1ull << (nx % 64)) - 1) : ~0ull;
  G.gx = G.wx, G.gy = (ny + TYC - 1) / TYC, G.gz = (nz + TZC - 1) / TZC;
  const uint64_t nw = (uint64_t)G.wx * ny * nz;
  const uint64_t ntile = (uint64_t)G.gx * G.gy * G.gz;
  uint64_t *orig[2], *buf[2][2];
  orig[0] = calloc(nw, 8), orig[1] = calloc(nw, 8);
  for (int m = 0; m < 2; ++m) buf[m][0] = malloc(nw * 8), buf[m][1] = malloc(nw * 8);
  uint8_t *state[2] = {calloc(ntile, 1), calloc(ntile, 1)};
  memset(hist_fp, 0, sizeof(uint32_t) * hist_len);
  memset(hist_fn, 0, sizeof(uint32_t) * hist_len);
  if (map) memset(map, 0, sizeof(uint16_t) * (uint64_t)nx * ny * nz);

  /* pack: bit x%64 of word (x/64, y, z); the label compare of
   * hello.cu:453 / 615 */
  uint64_t ng = 0, ns = 0, nfp = 0, nfn = 0;
#pragma omp parallel for collapse(2) reduction(+ : ng, ns, nfp, nfn) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int wx = 0; wx < G.wx; ++wx) {
        uint64_t a = 0, b = 0;
        const int xe = (wx * 64 + 64 < nx) ? 64 : nx - wx * 64;
        const uint64_t base = (uint64_t)wx * 64 + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * z);
        for (int k = 0; k < xe; ++k) {
          a |= (uint64_t)is_on(gold, dtype, label, base + k) << k;
          b |= (uint64_t)is_on(segm, dtype, label, base + k) << k;
        }
        const uint64_t i = widx(&G, wx, y, z);
        orig[0][i] = a, orig[1][i] = b;
        ng += __builtin_popcountll(a), ns += __builtin_popcountll(b);
        nfp += __builtin_popcountll(b & ~a), nfn += __builtin_popcountll(a & ~b);
      }
  out->n_gold = ng, out->n_segm = ns, out->n_fp = nfp, out->n_fn = nfn;
  out->tiles = (uint32_t)ntile;
  int rc = 0;
  vec_t q[2] = {{0}, {0}};
  if ((nfp && !ng) || (nfn && !ns)) {
    rc = E_EMPTY;
    goto done;
  }
  for (int m = 0; m < 2; ++m) {
    memcpy(buf[m][0], orig[m], nw * 8);
    memcpy(buf[m][1], orig[m], nw * 8);
  }
  /* tile classification and the bounding box of non-empty tiles */
  G.lo[0] = G.gx, G.lo[1] = G.gy, G.lo[2] = G.gz, G.hi[0] = G.hi[1] = G.hi[2] = -1;
  uint8_t *cls[2] = {calloc(ntile, 1), calloc(ntile, 1)}; /* 0 empty 1 mixed 2 full */
  for (uint64_t tl = 0; tl < ntile; ++tl) {
    const int tx = (int)(tl % G.gx), ty = (int)((tl / G.gx) % G.gy), tz = (int)(tl / ((uint64_t)G.gx * G.gy));
    const uint64_t vb = valid_bits(&G, tx);
    for (int m = 0; m < 2; ++m) {
      int any = 0, full = 1;
      for (int z = tz * TZC; z < nz && z < tz * TZC + TZC; ++z)
        for (int y = ty * TYC; y < ny && y < ty * TYC + TYC; ++y) {
          const uint64_t w = orig[m][widx(&G, tx, y, z)];
          any |= w != 0;
          full &= w == vb;
        }
      cls[m][tl] = full ? 2 : (any ? 1 : 0);
      if (any) {
        if (tx < G.lo[0]) G.lo[0] = tx;
        if (ty < G.lo[1]) G.lo[1] = ty;
        if (tz < G.lo[2]) G.lo[2] = tz;
        if (tx > G.hi[0]) G.hi[0] = tx;
        if (ty > G.hi[1]) G.hi[1] = ty;
        if (tz > G.hi[2]) G.hi[2] = tz;
      }
    }
  }
  if (!crop) {
    G.lo[0] = G.lo[1] = G.lo[2] = 0;
    G.hi[0] = G.gx - 1, G.hi[1] = G.gy - 1, G.hi[2] = G.gz - 1;
  } else if (conn == 18) {
    /* a shortest path of the 18 element may step one voxel outside the
     * voxels' box in x or y (offset (2,0,2) takes the corners (1,1,1) and
     * (1,-1,1)): one more tile on each side in x and y */
    for (int k = 0; k < 2; ++k) {
      const int g = k == 0 ? G.gx : G.gy;
      if (G.lo[k] > 0) --G.lo[k];
      if (G.hi[k] >= 0 && G.hi[k] < g - 1) ++G.hi[k];
    }
  }
  const uint64_t tgt[2] = {cut_of(r, nfp), cut_of(r, nfn)};
  uint64_t cov[2] = {0, 0};
  const int dtx[6] = {-1, 1, 0, 0, 0, 0}, dty[6] = {0, 0, -1, 1, 0, 0},
            dtz[6] = {0, 0, 0, 0, -1, 1};
  /* initial queue: non-empty non-full tiles, then empty neighbours that a
   * face of a non-empty tile reaches (the reference's padding pass,
   * hello.cu:1204, 1766-1863) */
  for (int m = 0; m < 2; ++m) {
    if (!tgt[m]) continue;
    for (uint64_t tl = 0; tl < ntile; ++tl) {
      if (cls[m][tl] == 1) {
        state[m][tl] = ST_ACTIVE;
        vpush(&q[m], (uint32_t)tl);
      } else if (cls[m][tl] == 2) {
        state[m][tl] = ST_FULL;
      }
    }
    for (uint64_t tl = 0; tl < ntile; ++tl) {
      if (!cls[m][tl]) continue;
      const int tx = (int)(tl % G.gx), ty = (int)((tl / G.gx) % G.gy), tz = (int)(tl / ((uint64_t)G.gx * G.gy));
      if (conn != 6) { /* the neighbours the tile's voxels reach */
        if (!in_box(&G, tx, ty, tz)) continue;
        uint32_t rb = 0;
        for (int z = tz * TZC; z < nz && z < tz * TZC + TZC; ++z)
          for (int y = ty * TYC; y < ny && y < ty * TYC + TYC; ++y)
            rb |= reach_bits(orig[m][widx(&G, tx, y, z)], y, z, ty * TYC, tz * TZC);
        for (int j = 0; j < 27; ++j)
          if ((rb >> j) & 1) activate(&G, state[m], tx + j % 3 - 1, ty + (j / 3) % 3 - 1, tz + j / 9 - 1, &q[m]);
        continue;
      }
      for (int f = 0; f < 6; ++f) {
        const int ux = tx + dtx[f], uy = ty + dty[f], uz = tz + dtz[f];
        if (ux < 0 || uy < 0 || uz < 0 || ux >= G.gx || uy >= G.gy || uz >= G.gz) continue;
        if (!in_box(&G, ux, uy, uz)) continue;
        const uint64_t ul = ux + (uint64_t)G.gx * (uy + (uint64_t)G.gy * uz);
        if (state[m][ul] != ST_IDLE) continue;
        /* face bits of this tile toward u */
        int hit = 0;
        for (int z = tz * TZC; z < nz && z < tz * TZC + TZC && !hit; ++z)
          for (int y = ty * TYC; y < ny && y < ty * TYC + TYC && !hit; ++y) {
            const uint64_t w = orig[m][widx(&G, tx, y, z)];
            switch (f) {
              case 0: hit = (w & 1ull) != 0; break;
              case 1: hit = (w >> 63) != 0; break;
              case 2: hit = y == ty * TYC && w; break;
              case 3: hit = y == ty * TYC + TYC - 1 && w; break;
              case 4: hit = z == tz * TZC && w; break;
              case 5: hit = z == tz * TZC + TZC - 1 && w; break;
            }
          }
        if (hit) {
          state[m][ul] = ST_ACTIVE;
          vpush(&q[m], (uint32_t)ul);
        }
      }
    }
  }
  free(cls[0]), free(cls[1]);

  int t = 0;
  while (cov[0] < tgt[0] || cov[1] < tgt[1]) {
    if (t + 1 >= hist_len) {
      rc = E_STUCK;
      goto done;
    }
    for (int m = 0; m < 2; ++m) {
      if (cov[m] >= tgt[m]) continue;
      if (q[m].n == 0) {
        rc = E_STUCK;
        goto done;
      }
      out->queue_total += q[m].n;
    }
    uint32_t qn = (uint32_t)((cov[0] < tgt[0] ? q[0].n : 0) + (cov[1] < tgt[1] ? q[1].n : 0));
    if (qn > out->queue_max) out->queue_max = qn;
    for (int m = 0; m < 2; ++m) {
      if (cov[m] >= tgt[m]) continue;
      const uint64_t *src = buf[m][t & 1];
      uint64_t *dst = buf[m][(t + 1) & 1];
#if ORACLE_MUTANT == 3
      const uint64_t *ref = orig[m];
#else
      const uint64_t *ref = orig[1 - m];
#endif
      uint64_t total = 0;
      vec_t nq = {0};
#pragma omp parallel reduction(+ : total)
      {
        vec_t local = {0};
#pragma omp for schedule(dynamic, 16)
        for (uint64_t e = 0; e < q[m].n; ++e) {
          const uint32_t tl = q[m].v[e];
          const int tx = (int)(tl % G.gx), ty = (int)((tl / G.gx) % G.gy),
                    tz = (int)(tl / ((uint64_t)G.gx * G.gy));
          int faces[6];
          uint64_t c;
          uint32_t rb;
          const int full = tile_step(&G, src, dst, ref, conn, tx, ty, tz, t, map, &c, faces, &rb);
          total += c;
          if (full) {
            state[m][tl] = ST_FULL;
          } else {
            vpush(&local, tl);
          }
          if (conn != 6) {
            for (int j = 0; j < 27; ++j)
              if ((rb >> j) & 1)
                activate(&G, state[m], tx + j % 3 - 1, ty + (j / 3) % 3 - 1, tz + j / 9 - 1, &local);
            continue;
          }
          for (int f = 0; f < 6; ++f)
            if (faces[f]) activate(&G, state[m], tx + dtx[f], ty + dty[f], tz + dtz[f], &local);
        }
#pragma omp critical
        {
          for (uint64_t k = 0; k < local.n; ++k) vpush(&nq, local.v[k]);
        }
        free(local.v);
      }
#if ORACLE_MUTANT == 1
      (m == 0 ? hist_fp : hist_fn)[t + 2 < hist_len ? t + 2 : t + 1] += (uint32_t)total;
#else
      (m == 0 ? hist_fp : hist_fn)[t + 1] += (uint32_t)total;
#endif
      cov[m] += total;
      free(q[m].v);
      q[m] = nq;
    }
    ++t;
  }
  out->iterations = t;
  oracle_finish(hist_fp, hist_fn, hist_len, r, out);
done:
  free(q[0].v), free(q[1].v);
  free(orig[0]), free(orig[1]);
  for (int m = 0; m < 2; ++m) free(buf[m][0]), free(buf[m][1]), free(state[m]);
  return rc;
}
