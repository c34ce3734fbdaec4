/* hdm_compat.hpp - header-only C++ mirror of the reference's host entry point
 * over the C ABI of hdm.h.
 *
 * Replaces: template <typename T> int getHausdorffDistance_CUDA_Generic(
 *   T* goldStandard, T* algoOutput, int WIDTH, int HEIGHT, int DEPTH,
 *   float robustnessPercent, bool resIterNeeded, T numberToLookFor,
 *   bool res3DNeeded)   (hello.cu:2025-2062)
 *
 * Same names, argument order and meaning: device pointers owned by the
 * caller, dims in voxels (x fastest), robustness in [0, 1], the label to look
 * for; returns the integer HD.  Differences: failures throw hdm_compat::error
 * (the reference returns garbage or hangs on an empty mask), identical masks
 * give 0 (the reference returns 1), and the call is ordered on `stream`
 * (default: the legacy default stream) instead of a device-wide sync.
 * resIterNeeded and res3DNeeded asked the reference for its result list and
 * 3-D map; here use hdm_first_hit_map_u8 (or the average_hd overload below).
 * T may be any 1-byte type (bool, char, uint8_t) or a 4-byte integer.
 */
#ifndef HDM_COMPAT_HPP
#define HDM_COMPAT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "hdm.h"

namespace hdm_compat {

struct error : std::runtime_error {
  int code;
  explicit error(int c)
      : std::runtime_error(std::string("hdm: ") + hdm_strerror(c)), code(c) {}
};

template <typename T>
inline int run(const T *gold, const T *segm, int nx, int ny, int nz, double robustness,
               T label, void *stream, hdm_result *out) {
  static_assert(sizeof(T) == 1 || (sizeof(T) == 4 && std::is_integral<T>::value),
                "mask element type must be 1 byte or a 4-byte integer");
  hdm_opts o;
  hdm_opts_init(&o);
  o.robustness = robustness;
  o.ptr_kind = HDM_PTR_DEVICE;
  o.stream = stream;
  int rc;
  if (sizeof(T) == 1) {
    uint8_t lab;
    __builtin_memcpy(&lab, &label, 1);
    rc = hdm_hausdorff_u8(reinterpret_cast<const uint8_t *>(gold),
                          reinterpret_cast<const uint8_t *>(segm), nx, ny, nz, lab, &o, out);
  } else {
    int32_t lab;
    __builtin_memcpy(&lab, &label, 4);
    rc = hdm_hausdorff_i32(reinterpret_cast<const int32_t *>(gold),
                           reinterpret_cast<const int32_t *>(segm), nx, ny, nz, lab, &o, out);
  }
  if (rc != HDM_OK) throw error(rc);
  return out->hd;
}

template <typename T>
inline int getHausdorffDistance_CUDA_Generic(T *goldStandard, T *algoOutput, int WIDTH,
                                             int HEIGHT, int DEPTH, float robustnessPercent,
                                             bool resIterNeeded, T numberToLookFor,
                                             bool res3DNeeded, void *stream = nullptr) {
  (void)resIterNeeded;
  (void)res3DNeeded;
  hdm_result r;
  return run<T>(goldStandard, algoOutput, WIDTH, HEIGHT, DEPTH, robustnessPercent,
                numberToLookFor, stream, &r);
}

/* Average HD (sum of first hits / |FP u FN|, FP64 on the host) of the same call. */
template <typename T>
inline double averageHausdorffDistance(T *goldStandard, T *algoOutput, int WIDTH, int HEIGHT,
                                       int DEPTH, float robustnessPercent, T numberToLookFor,
                                       void *stream = nullptr) {
  hdm_result r;
  run<T>(goldStandard, algoOutput, WIDTH, HEIGHT, DEPTH, robustnessPercent, numberToLookFor,
         stream, &r);
  return r.avg_hd;
}

}  // namespace hdm_compat

#endif /* HDM_COMPAT_HPP */
