// Serving engine (continuous batching over the persistent dilation kernel).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "../../include/hdm.h"

extern "C" void hdm_server_opts_init(hdm_server_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->hbm_budget = 64ull << 30;
  o->max_slots = 64;
  o->slice_iters = 8;
  o->tile = HDM_TILE_64x8x8;
  o->robustness = 1.0;
}
extern "C" int hdm_server_create(const hdm_server_opts *, hdm_server **) { return HDM_E_CAPACITY; }
extern "C" int hdm_server_submit(hdm_server *, const void *, const void *, int, int, int, int,
                                 int, uint64_t *) {
  return HDM_E_INVALID;
}
extern "C" int hdm_server_step(hdm_server *, hdm_completion *, int, int *) { return HDM_E_INVALID; }
extern "C" int hdm_server_load(hdm_server *, int *, int *, uint64_t *) { return HDM_E_INVALID; }
extern "C" double hdm_server_now(void) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
extern "C" void hdm_server_destroy(hdm_server *) {}
