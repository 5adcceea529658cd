// tfh_render_fwd_noalloc: the render_rays forward (render.py:105-279) for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc). Its residuals live in a
// file-scope device array of fixed size instead of a caller-owned workspace.
//
// Each device that runs the entry point holds its own copy of the array from
// the first call on that device until the process ends; the address of that
// copy is looked up once per device (hipGetSymbolAddress, with the device
// current) and kept. Every call on one device shares the array: calls that
// overlap on one device (two streams, two host threads) race on it.
#include <mutex>

#include "tfh_common.h"

namespace tfh {
int32_t render_fwd_scratch(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                           const tfh_render_config* cfg, const float* aabb, const float* origins,
                           const float* directions, const uint32_t* camera_indices,
                           int64_t num_rays, const float* jitter, const float* base_ts,
                           const float* base_steps, const float* gumbel, float* out,
                           void* (*scratch_of_device)(size_t cap), size_t cap, const char* who,
                           hipStream_t st);
}  // namespace tfh

namespace {

constexpr size_t kScratchBytes = (size_t)TFH_RENDER_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// The array's name is a host-side stand-in; kernels get the per-device
// address held in g_scratch.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

void* scratch_of_device(size_t cap) {
  if (cap > kScratchBytes) {
    tfh::fail(TFH_ERR_INVALID_ARGUMENT, "scratch of %zu bytes requested, the array holds %zu",
              cap, kScratchBytes);
    return nullptr;
  }
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    tfh::fail(TFH_ERR_LAUNCH, "tfh_render_fwd_noalloc: hipGetDevice failed (%s, device %d)",
              hipGetErrorString(e), dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      tfh::fail(TFH_ERR_LAUNCH, "tfh_render_fwd_noalloc: hipGetSymbolAddress failed: %s",
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" int32_t tfh_render_fwd_noalloc(const tfh_vm* density, const tfh_vm* appearance,
                                          const tfh_mlp* mlp, const tfh_render_config* cfg,
                                          const float* aabb, const float* origins,
                                          const float* directions,
                                          const uint32_t* camera_indices, int64_t num_rays,
                                          const float* jitter, const float* base_ts,
                                          const float* base_steps, const float* gumbel,
                                          float* out, tfh_stream_t stream) {
  return tfh::render_fwd_scratch(density, appearance, mlp, cfg, aabb, origins, directions,
                                 camera_indices, num_rays, jitter, base_ts, base_steps, gumbel,
                                 out, scratch_of_device, kScratchBytes, "tfh_render_fwd_noalloc",
                                 tfh::as_stream(stream));
}

