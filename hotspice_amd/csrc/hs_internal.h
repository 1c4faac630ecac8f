// Host-side helpers shared by the translation units of libhotspice_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/hotspice_hip.h"

int hs_set_error(int code, const char *fmt, ...);
void hs_clear_error();

#define HS_HIP_TRY(expr)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return hs_set_error(HS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define HS_LAUNCH_CHECK(name)                                                                 \
  do {                                                                                        \
    hipError_t e_ = hipGetLastError();                                                        \
    if (e_ != hipSuccess) return hs_set_error(HS_ERR_HIP, "launch of %s failed: %s", name, hipGetErrorString(e_)); \
  } while (0)

// Copies `bytes` of host memory to device memory `dst` in stream order, without waiting for the
// work queued before it: the bytes go through a pinned staging buffer that is refilled only once
// its previous copy has completed. `src` may be released as soon as the call returns.
int hs_upload_async(void *dst, const void *src, size_t bytes, hipStream_t s);

// Replays the kernel launches that `record(stream)` enqueues as one captured hipGraph on `s`
// (hs_core.hip). The first call with a given key on a device captures them on a private stream
// and instantiates the graph; later calls with that key replay it. The key must determine every
// launch `record` makes and every argument of those launches.
int hs_graph_launch(const std::vector<int64_t> &key, hipStream_t s, const std::function<int(hipStream_t)> &record);

// Steps per captured graph of the stepping loops (hs_set_graph_steps; 0 = plain launches).
int hs_graph_steps();

// Validates the fields every stepping kernel relies on (sizes, pointers, table bounds).
int hs_check_instance(const hs_instance_t *I);

// HS_ERR_ARG (naming `who`) for an instance with FP64 fields (E64), which only the recompute,
// readouts, Zeeman steps and small-grid Metropolis steps support.
int hs_check_f32_fields(const hs_instance_t *I, const char *who);

// Run-state helpers (hs_neel.hip): reset per-call fields of state[0]; copy state[from] to state[0].
__global__ void k_state_begin(hs_state_t *st, int begin);
__global__ void k_state_copy(hs_state_t *st, int from);
