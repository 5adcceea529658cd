// Shared host and device helpers of libtensorf_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tensorf_hip.h"

namespace tfh {

// ---- host: errors ---------------------------------------------------------
// Records a printf-style message as the calling thread's last error and
// returns `code`.
int32_t fail(int32_t code, const char* fmt, ...);
void clear_error();
// TFH_ERR_NO_DEVICE when no HIP device is visible.
int32_t require_device();
// Converts the sticky launch error of the last kernel into a status.
int32_t check_launch(const char* what);

#define TFH_REQUIRE(cond, ...)                                       \
  do {                                                               \
    if (!(cond)) return ::tfh::fail(TFH_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
  } while (0)

#define TFH_HIP_TRY(call)                                                           \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::tfh::fail(TFH_ERR_LAUNCH, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define TFH_TRY(expr)              \
  do {                             \
    int32_t s_ = (expr);           \
    if (s_ != TFH_OK) return s_;   \
  } while (0)

inline hipStream_t as_stream(tfh_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Stage timing (tfh_profile_enable): while enabled, a ProfScope records a
// HIP event pair on `st` around the launches in its lifetime.
bool profiling();
struct ProfScope {
  int slot = -1;
  hipStream_t st;
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
};

// Validates a tfh_vm (non-null pointers, sane sizes).
int32_t check_vm(const tfh_vm* vm, const char* name);

// Bump allocator over a caller-owned workspace.
struct Carve {
  char* base;
  size_t cap;
  size_t off = 0;
  bool ok = true;
  Carve(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <class T>
  T* take(int64_t count) {
    size_t bytes = align_up(sizeof(T) * (size_t)(count > 0 ? count : 0), 256);
    if (off + bytes > cap || (base == nullptr && bytes > 0)) {
      ok = false;
      off += bytes;
      return nullptr;
    }
    T* p = reinterpret_cast<T*>(base + off);
    off += bytes;
    return p;
  }
};

// ---- XCD-ordered blocks --------------------------------------------------------
// The hardware deals a grid's blocks to the part's 8 XCDs in turn, each with
// its own L2. Over a grid of xcd_grid(units) blocks, block b takes unit
// (b % 8) * per + b / 8 (per = ceil(units / 8)), so that each XCD walks one
// contiguous range of units instead of every eighth one. Units past the end
// (at most 7) come out >= units; their blocks return at once. (The MLP
// forward that gathers its rows: uniform-512 2.30 -> 2.22-2.27 ms, fisheye
// 5.67 -> 4.99 ms, although its L2 hit rate did not rise; the ray-order
// line passes gained nothing from it, and the passes over sorted pairs
// lost: the proposal-128 line-cell walk 1.33 -> 1.60 ms.)
constexpr int kXcds = 8;
inline int64_t xcd_grid(int64_t units) { return kXcds * ceil_div(units, kXcds); }
__device__ __forceinline__ int64_t xcd_unit(int64_t units) {
  const int64_t per = (units + kXcds - 1) / kXcds;
  return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// ---- device helpers ---------------------------------------------------------
// Order-1 'nearest' taps of map_coordinates on a grid of `g` texels for a
// normalised coordinate x in [-1,1] (tensor_vm.py:150-153, 241-248):
// u = (x+1)/2*(g-1); i0 = floor(u); w1 = u - i0; indices clamped to [0, g-1].
// Any float (inf/NaN included) yields in-bounds indices.
struct Tap {
  int i0, i1;
  float w0, w1;
};

__device__ __forceinline__ Tap lin_tap(float x, int g) {
  float u = (x + 1.0f) / 2.0f;
  u = u * (float)(g - 1);
  float fl = floorf(u);
  Tap t;
  t.w1 = u - fl;
  t.w0 = 1.0f - t.w1;
  float flc = fminf(fmaxf(fl, -2.0f), (float)(g + 1));  // NaN -> -2
  int i = (int)flc;
  t.i0 = min(max(i, 0), g - 1);
  t.i1 = min(max(i + 1, 0), g - 1);
  return t;
}

// Mode m of TensorVM.interpolate reads (v, a, b) = ijk permuted by
// [0,1,2], [2,0,1], [1,2,0] (tensor_vm.py:50-52): the vector uses v, the
// matrix rows a and columns b.
struct ModeTaps {
  Tap v, a, b;
};

__device__ __forceinline__ void mode_taps(float x, float y, float z, int g, ModeTaps mt[3]) {
  Tap tx = lin_tap(x, g), ty = lin_tap(y, g), tz = lin_tap(z, g);
  mt[0].v = tx; mt[0].a = ty; mt[0].b = tz;
  mt[1].v = tz; mt[1].a = tx; mt[1].b = ty;
  mt[2].v = ty; mt[2].a = tz; mt[2].b = tx;
}

// jax.nn.softplus = logaddexp(x, 0) (render.py:207).
__device__ __forceinline__ float softplus(float x) {
  return fmaxf(x, 0.0f) + log1pf(expf(-fabsf(x)));
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ float wave_incl_scan(float v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    float t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over aligned groups of 16 lanes.
__device__ __forceinline__ float group16_sum(float v) {
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Order-preserving uint32 key of a float (larger float -> larger key).
__device__ __forceinline__ uint32_t float_key(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Counter-based uniform float in [0,1): splitmix64 finaliser of (seed, index).
__host__ __device__ __forceinline__ float uniform01(uint64_t seed, uint64_t idx) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + idx * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (float)(z >> 40) * (1.0f / 16777216.0f);
}

}  // namespace tfh
