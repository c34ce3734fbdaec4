// Deterministic synthetic mask pairs: host spec builders, one integer voxel
// test shared by the host evaluator and the device kernel.  No counterpart in
// the reference; its benchmark read a private HDF5 volume
// (forBench/hdf5Manag.cu:196-220).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hdm.h"
#include "../../include/hdm_synth.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// splitmix64 stream: output k (k >= 1) of the stream seeded with s is
// mix64(s + k * golden).
struct Rng {
  uint64_t s;
  uint64_t next() {
    s += kGolden;
    return mix64(s);
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int range(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }
};

__host__ __device__ inline bool ell_contains(const hdm_ell &e, int x, int y, int z) {
  const int64_t d0 = x - e.c[0], d1 = y - e.c[1], d2 = z - e.c[2];
  const int64_t q0 = e.r[0] * d0 + e.r[1] * d1 + e.r[2] * d2;
  const int64_t q1 = e.r[3] * d0 + e.r[4] * d1 + e.r[5] * d2;
  const int64_t q2 = e.r[6] * d0 + e.r[7] * d1 + e.r[8] * d2;
  const uint64_t a0 = (uint64_t)(q0 < 0 ? -q0 : q0);
  const uint64_t a1 = (uint64_t)(q1 < 0 ? -q1 : q1);
  const uint64_t a2 = (uint64_t)(q2 < 0 ? -q2 : q2);
  if (a0 > 1024ull * (uint64_t)e.s[0] || a1 > 1024ull * (uint64_t)e.s[1] ||
      a2 > 1024ull * (uint64_t)e.s[2])
    return false;
  const uint64_t s0 = (uint64_t)e.s[0] * e.s[0], s1 = (uint64_t)e.s[1] * e.s[1],
                 s2 = (uint64_t)e.s[2] * e.s[2];
  const uint64_t lhs = a0 * a0 * (s1 * s2) + a1 * a1 * (s0 * s2) + a2 * a2 * (s0 * s1);
  const uint64_t rhs = (s0 * s1 * s2) << 20;
  return lhs <= rhs;
}

__host__ __device__ inline bool gold_at(const hdm_synth_spec &s, int x, int y, int z) {
  for (int i = 0; i < s.n_gold; ++i)
    if (ell_contains(s.gold[i], x, y, z)) return true;
  return false;
}

__host__ __device__ inline bool segm_base_at(const hdm_synth_spec &s, int x, int y, int z) {
  if (x >= s.cube_lo[0] && x < s.cube_hi[0] && y >= s.cube_lo[1] && y < s.cube_hi[1] &&
      z >= s.cube_lo[2] && z < s.cube_hi[2])
    return true;
  for (int i = 0; i < s.n_segm; ++i)
    if (ell_contains(s.segm[i], x, y, z)) return true;
  return false;
}

// Final segm value at (x,y,z): the base shape with the seeded flips applied.
__host__ __device__ inline bool segm_at(const hdm_synth_spec &s, int x, int y, int z) {
  const bool b = segm_base_at(s, x, y, z);
  if (s.flip_mode == 0) return b;
  const uint64_t v = (uint64_t)x + (uint64_t)s.nx * ((uint64_t)y + (uint64_t)s.ny * (uint64_t)z);
  const uint64_t h = mix64(s.flip_seed + (v + 1) * kGolden);
  if ((h >> 11) >= s.flip_thresh) return b;
  const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  if (s.flip_mode == 1) {
    if (!b) return false;
    for (int k = 0; k < 6; ++k) {
      const int u = x + nb[k][0], w = y + nb[k][1], t = z + nb[k][2];
      if (u < 0 || w < 0 || t < 0 || u >= s.nx || w >= s.ny || t >= s.nz) return false;
      if (!segm_base_at(s, u, w, t)) return false;
    }
    return true;
  }
  for (int k = 0; k < 6; ++k) {
    const int u = x + nb[k][0], w = y + nb[k][1], t = z + nb[k][2];
    if (u < 0 || w < 0 || t < 0 || u >= s.nx || w >= s.ny || t >= s.nz) continue;
    if (segm_base_at(s, u, w, t) != b) return !b;
  }
  return b;
}

__global__ void synth_kernel(hdm_synth_spec s, uint8_t *gold, uint8_t *segm, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int x = (int)(v % (uint64_t)s.nx);
    const uint64_t r = v / (uint64_t)s.nx;
    const int y = (int)(r % (uint64_t)s.ny);
    const int z = (int)(r / (uint64_t)s.ny);
    gold[v] = gold_at(s, x, y, z) ? 1 : 0;
    segm[v] = segm_at(s, x, y, z) ? 1 : 0;
  }
}

void identity_q10(int32_t r[9]) {
  for (int k = 0; k < 9; ++k) r[k] = (k % 4 == 0) ? 1024 : 0;
}

bool dims_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0; }

}  // namespace

extern "C" int hdm_synth_spec_sphere(hdm_synth_spec *s, int nx, int ny, int nz, int cx, int cy,
                                     int cz, int r, int dx, int dy, int dz, double p_flip,
                                     uint64_t seed) {
  if (!s || !dims_ok(nx, ny, nz) || r < 1 || r > 120 || p_flip < 0.0 || p_flip > 1.0)
    return HDM_E_INVALID;
  std::memset(s, 0, sizeof(*s));
  s->nx = nx, s->ny = ny, s->nz = nz;
  s->n_gold = s->n_segm = 1;
  hdm_ell &g = s->gold[0];
  g.c[0] = cx, g.c[1] = cy, g.c[2] = cz;
  g.s[0] = g.s[1] = g.s[2] = r;
  identity_q10(g.r);
  s->segm[0] = g;
  s->segm[0].c[0] += dx, s->segm[0].c[1] += dy, s->segm[0].c[2] += dz;
  s->flip_mode = p_flip > 0.0 ? 1 : 0;
  s->flip_seed = mix64(seed);
  s->flip_thresh = (uint64_t)(p_flip * 9007199254740992.0);
  return HDM_OK;
}

extern "C" int hdm_synth_spec_ellipsoids(hdm_synth_spec *s, int nx, int ny, int nz, int n,
                                         int amin, int amax, int dmax, double p_flip,
                                         uint64_t seed) {
  if (!s || !dims_ok(nx, ny, nz) || n < 1 || n > HDM_SYN_MAX_ELL || amin < 1 ||
      amax < amin || amax + dmax > 120 || dmax < 0 || p_flip < 0.0 || p_flip > 1.0)
    return HDM_E_INVALID;
  std::memset(s, 0, sizeof(*s));
  s->nx = nx, s->ny = ny, s->nz = nz;
  s->n_gold = s->n_segm = n;
  Rng g{seed};
  const int dims[3] = {nx, ny, nz};
  for (int i = 0; i < n; ++i) {
    hdm_ell &e = s->gold[i];
    for (int k = 0; k < 3; ++k) e.s[k] = g.range(amin, amax);
    for (int k = 0; k < 3; ++k) {
      const int m = dims[k] / 8;
      e.c[k] = g.range(m, std::max(m, dims[k] - 1 - m));
    }
    double w, x, y, z, n2;
    do {
      w = 2.0 * g.u01() - 1.0;
      x = 2.0 * g.u01() - 1.0;
      y = 2.0 * g.u01() - 1.0;
      z = 2.0 * g.u01() - 1.0;
      n2 = w * w + x * x + y * y + z * z;
    } while (n2 > 1.0 || n2 < 1e-4);
    const double nr = std::sqrt(n2);
    w /= nr, x /= nr, y /= nr, z /= nr;
    const double R[9] = {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - z * w),
                         2.0 * (x * z + y * w),       2.0 * (x * y + z * w),
                         1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - x * w),
                         2.0 * (x * z - y * w),       2.0 * (y * z + x * w),
                         1.0 - 2.0 * (x * x + y * y)};
    for (int k = 0; k < 9; ++k) e.r[k] = (int32_t)std::floor(R[k] * 1024.0 + 0.5);
    const int delta = g.range(-dmax, dmax);
    hdm_ell &f = s->segm[i];
    f = e;
    for (int k = 0; k < 3; ++k) f.s[k] = std::max(1, e.s[k] + delta);
  }
  s->flip_mode = p_flip > 0.0 ? 2 : 0;
  s->flip_seed = mix64(seed);
  s->flip_thresh = (uint64_t)(p_flip * 9007199254740992.0);
  return HDM_OK;
}

extern "C" int hdm_synth_spec_long(hdm_synth_spec *s, int nx, int ny, int nz, int cx, int cy,
                                   int cz, int ax, int ay, int az, int gap, int edge) {
  if (!s || !dims_ok(nx, ny, nz) || ax < 1 || ay < 1 || az < 1 || ax > 120 || ay > 120 ||
      az > 120 || gap < 1 || edge < 1)
    return HDM_E_INVALID;
  std::memset(s, 0, sizeof(*s));
  s->nx = nx, s->ny = ny, s->nz = nz;
  s->n_gold = s->n_segm = 1;
  hdm_ell &g = s->gold[0];
  g.c[0] = cx, g.c[1] = cy, g.c[2] = cz;
  g.s[0] = ax, g.s[1] = ay, g.s[2] = az;
  identity_q10(g.r);
  s->segm[0] = g;
  s->cube_lo[0] = cx + ax + gap;
  s->cube_lo[1] = cy - edge / 2;
  s->cube_lo[2] = cz - edge / 2;
  for (int k = 0; k < 3; ++k) s->cube_hi[k] = s->cube_lo[k] + edge;
  if (s->cube_hi[0] > nx || s->cube_lo[1] < 0 || s->cube_hi[1] > ny || s->cube_lo[2] < 0 ||
      s->cube_hi[2] > nz)
    return HDM_E_INVALID;
  return HDM_OK;
}

extern "C" int hdm_synth_host(const hdm_synth_spec *s, uint8_t *gold, uint8_t *segm) {
  if (!s || !gold || !segm || !dims_ok(s->nx, s->ny, s->nz)) return HDM_E_INVALID;
  const hdm_synth_spec sp = *s;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nt = std::min<unsigned>(nt, (unsigned)sp.nz);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    th.emplace_back([&, t]() {
      for (int z = (int)t; z < sp.nz; z += (int)nt)
        for (int y = 0; y < sp.ny; ++y) {
          const uint64_t row = (uint64_t)sp.nx * ((uint64_t)y + (uint64_t)sp.ny * z);
          for (int x = 0; x < sp.nx; ++x) {
            gold[row + x] = gold_at(sp, x, y, z) ? 1 : 0;
            segm[row + x] = segm_at(sp, x, y, z) ? 1 : 0;
          }
        }
    });
  }
  for (auto &t : th) t.join();
  return HDM_OK;
}

extern "C" int hdm_synth_device(const hdm_synth_spec *s, uint8_t *gold, uint8_t *segm,
                                void *stream) {
  if (!s || !gold || !segm || !dims_ok(s->nx, s->ny, s->nz)) return HDM_E_INVALID;
  const uint64_t n = (uint64_t)s->nx * s->ny * s->nz;
  const uint64_t want = (n + 255) / 256;
  const unsigned blocks = (unsigned)std::min<uint64_t>(want, 65536);
  hipLaunchKernelGGL(synth_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *s, gold,
                     segm, n);
  return hipGetLastError() == hipSuccess ? HDM_OK : HDM_E_HIP;
}
