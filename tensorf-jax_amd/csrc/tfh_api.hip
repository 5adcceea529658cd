// Library plumbing of libtensorf_hip: error state, device query, layout
// conversion, the counter-based RNG and the fused Adam step.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "tfh_common.h"

namespace tfh {

static thread_local char g_last_error[1024] = "";

int32_t fail(int32_t code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_last_error[0] = '\0'; }

int32_t require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(TFH_ERR_NO_DEVICE, "no HIP device visible");
  }
  return TFH_OK;
}

int32_t check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(TFH_ERR_LAUNCH, "%s: launch failed: %s", what, hipGetErrorString(e));
  return TFH_OK;
}

// ---- stage timing -------------------------------------------------------------
struct ProfRecord {
  std::string name;
  hipEvent_t start, stop;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRecord> g_prof;

bool profiling() { return g_prof_on; }

ProfScope::ProfScope(const char* name, hipStream_t s) : st(s) {
  if (!g_prof_on) return;
  ProfRecord r;
  r.name = name;
  if (hipEventCreate(&r.start) != hipSuccess || hipEventCreate(&r.stop) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  (void)hipEventRecord(r.start, st);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  slot = (int)g_prof.size();
  g_prof.push_back(r);
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  (void)hipEventRecord(g_prof[slot].stop, st);
}

int32_t check_vm(const tfh_vm* vm, const char* name) {
  TFH_REQUIRE(vm != nullptr, "%s: null tfh_vm", name);
  TFH_REQUIRE(vm->plane != nullptr && vm->line != nullptr, "%s: null plane/line pointer", name);
  TFH_REQUIRE(vm->grid_dim >= 2 && vm->grid_dim <= 4096, "%s: grid_dim %lld outside [2, 4096]",
              name, (long long)vm->grid_dim);
  TFH_REQUIRE(vm->channels >= 1 && vm->channels <= 64, "%s: channels %lld outside [1, 64]", name,
              (long long)vm->channels);
  return TFH_OK;
}

// ---- layout -----------------------------------------------------------------
// (3, C, G*G) <-> (3, G*G, C) via 32x32 LDS tiles; `to_cl` picks the direction.
template <bool TO_CL>
__global__ void __launch_bounds__(256) k_transpose_cg(const float* __restrict__ src,
                                                      float* __restrict__ dst, int64_t C,
                                                      int64_t S) {
  __shared__ float tile[32][33];
  const int64_t m = blockIdx.z;
  const int64_t s0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const float* sm = src + m * C * S;
  float* dm = dst + m * C * S;
  for (int j = ty; j < 32; j += 8) {
    int64_t c = c0 + j, s = s0 + tx;
    if (TO_CL) {  // src (C, S)
      if (c < C && s < S) tile[j][tx] = sm[c * S + s];
    } else {  // src (S, C): read row s = s0 + j, column c = c0 + tx
      int64_t s2 = s0 + j, c2 = c0 + tx;
      if (c2 < C && s2 < S) tile[tx][j] = sm[s2 * C + c2];
    }
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    if (TO_CL) {  // dst (S, C): row s = s0 + j, column c = c0 + tx
      int64_t s = s0 + j, c = c0 + tx;
      if (c < C && s < S) dm[s * C + c] = tile[tx][j];
    } else {  // dst (C, S)
      int64_t c = c0 + j, s = s0 + tx;
      if (c < C && s < S) dm[c * S + s] = tile[j][tx];
    }
  }
}

template <bool TO_CL>
static int32_t transpose_cg(const float* src, float* dst, int64_t C, int64_t S, hipStream_t st) {
  dim3 grid((unsigned)ceil_div(S, 32), (unsigned)ceil_div(C, 32), 3);
  hipLaunchKernelGGL(k_transpose_cg<TO_CL>, grid, dim3(256), 0, st, src, dst, C, S);
  return check_launch("transpose");
}

// ---- RNG ----------------------------------------------------------------------
__global__ void k_uniform(uint64_t seed, uint64_t offset, int64_t n, float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = uniform01(seed, offset + (uint64_t)i);
}

// ---- Adam -------------------------------------------------------------------
// optax.scale_by_adam (eps_root = 0) followed by scale(-lr) and the decay
// factor (training.py:184-205). bc1 = 1 - b1^count, bc2 = 1 - b2^count.
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ p, float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int64_t n, int64_t split, float lr_lo, float lr_hi,
                                              float b1, float b2, float bc1, float bc2, float eps,
                                              int zero_grad, const float* __restrict__ hyper) {
  if (hyper) {  // the step's scalars from device memory: -lr lo, -lr hi, 1-b1^t, 1-b2^t
    lr_lo = hyper[0];
    lr_hi = hyper[1];
    bc1 = hyper[2];
    bc2 = hyper[3];
  }
  int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  int64_t stride = (int64_t)gridDim.x * blockDim.x * 4;
  for (; i < n; i += stride) {
    if (i + 4 <= n && (((uintptr_t)(p + i) | (uintptr_t)(g + i) | (uintptr_t)(m + i) |
                        (uintptr_t)(v + i)) & 15) == 0) {
      float4 pp = *reinterpret_cast<float4*>(p + i);
      float4 gg = *reinterpret_cast<float4*>(g + i);
      float4 mm = *reinterpret_cast<float4*>(m + i);
      float4 vv = *reinterpret_cast<float4*>(v + i);
      float* pa = &pp.x; float* ga = &gg.x; float* ma = &mm.x; float* va = &vv.x;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float lr = (i + e < split) ? lr_lo : lr_hi;
        float gv = ga[e];
        float mv = (1.0f - b1) * gv + b1 * ma[e];
        float vv2 = (1.0f - b2) * (gv * gv) + b2 * va[e];
        float upd = (mv / bc1) / (sqrtf(vv2 / bc2) + eps);
        pa[e] = pa[e] + upd * lr;
        ma[e] = mv;
        va[e] = vv2;
      }
      *reinterpret_cast<float4*>(p + i) = pp;
      *reinterpret_cast<float4*>(m + i) = mm;
      *reinterpret_cast<float4*>(v + i) = vv;
      if (zero_grad) *reinterpret_cast<float4*>(g + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      for (int64_t j = i; j < i + 4 && j < n; ++j) {
        float lr = (j < split) ? lr_lo : lr_hi;
        float gv = g[j];
        float mv = (1.0f - b1) * gv + b1 * m[j];
        float vv2 = (1.0f - b2) * (gv * gv) + b2 * v[j];
        float upd = (mv / bc1) / (sqrtf(vv2 / bc2) + eps);
        p[j] = p[j] + upd * lr;
        m[j] = mv;
        v[j] = vv2;
        if (zero_grad) g[j] = 0.0f;
      }
    }
  }
}

int32_t adam_launch(float* p, float* g, float* m, float* v, int64_t n, int64_t split,
                    float lr_lo, float lr_hi, float decay, float b1, float b2, float eps,
                    int64_t count, int zero_grad, hipStream_t st, const float* hyper) {
  // the learning rate enters negated (gradient descent) and scaled by the decay
  float s_lo = -lr_lo * decay, s_hi = -lr_hi * decay;
  float bc1 = 1.0f - powf(b1, (float)count);
  float bc2 = 1.0f - powf(b2, (float)count);
  int64_t blocks = ceil_div(ceil_div(n, 4), 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  ProfScope ps("adam", st);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, st, p, g, m, v, n, split, s_lo,
                     s_hi, b1, b2, bc1, bc2, eps, zero_grad, hyper);
  return check_launch("adam");
}

}  // namespace tfh

using namespace tfh;

extern "C" {

int32_t tfh_abi_version(void) { return TFH_ABI_VERSION; }

const char* tfh_last_error(void) { return g_last_error; }

int32_t tfh_device_count(int32_t* count) {
  TFH_REQUIRE(count != nullptr, "tfh_device_count: null output");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return TFH_OK;
}

int32_t tfh_vm_to_channel_last(const float* matrix, const float* vector, int64_t channels,
                               int64_t grid_dim, float* plane, float* line, tfh_stream_t stream) {
  TFH_REQUIRE(matrix && vector && plane && line, "tfh_vm_to_channel_last: null pointer");
  TFH_REQUIRE(channels >= 1 && grid_dim >= 1, "tfh_vm_to_channel_last: bad shape");
  TFH_TRY(require_device());
  TFH_TRY(transpose_cg<true>(matrix, plane, channels, grid_dim * grid_dim, as_stream(stream)));
  return transpose_cg<true>(vector, line, channels, grid_dim, as_stream(stream));
}

int32_t tfh_vm_from_channel_last(const float* plane, const float* line, int64_t channels,
                                 int64_t grid_dim, float* matrix, float* vector,
                                 tfh_stream_t stream) {
  TFH_REQUIRE(matrix && vector && plane && line, "tfh_vm_from_channel_last: null pointer");
  TFH_REQUIRE(channels >= 1 && grid_dim >= 1, "tfh_vm_from_channel_last: bad shape");
  TFH_TRY(require_device());
  TFH_TRY(transpose_cg<false>(plane, matrix, channels, grid_dim * grid_dim, as_stream(stream)));
  return transpose_cg<false>(line, vector, channels, grid_dim, as_stream(stream));
}

int32_t tfh_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (on) {
    for (auto& r : g_prof) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.stop);
    }
    g_prof.clear();
  }
  g_prof_on = on != 0;
  return TFH_OK;
}

int32_t tfh_profile_read(int32_t max, char* names, double* total_ms, int64_t* counts,
                         int32_t* num_stages) {
  TFH_REQUIRE(names && total_ms && counts && num_stages && max > 0, "tfh_profile_read: bad args");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  std::vector<std::string> keys;
  std::vector<double> ms;
  std::vector<int64_t> cnt;
  for (auto& r : g_prof) {
    TFH_HIP_TRY(hipEventSynchronize(r.stop));
    float t = 0.f;
    TFH_HIP_TRY(hipEventElapsedTime(&t, r.start, r.stop));
    size_t j = 0;
    while (j < keys.size() && keys[j] != r.name) ++j;
    if (j == keys.size()) {
      keys.push_back(r.name);
      ms.push_back(0.0);
      cnt.push_back(0);
    }
    ms[j] += t;
    cnt[j] += 1;
  }
  int n = (int)std::min<size_t>(keys.size(), (size_t)max);
  for (int j = 0; j < n; ++j) {
    snprintf(names + 48 * j, 48, "%s", keys[j].c_str());
    total_ms[j] = ms[j];
    counts[j] = cnt[j];
  }
  *num_stages = n;
  return TFH_OK;
}

int32_t tfh_uniform(uint64_t seed, uint64_t offset, int64_t n, float* out, tfh_stream_t stream) {
  TFH_REQUIRE(n >= 0, "tfh_uniform: negative n");
  if (n == 0) return TFH_OK;
  TFH_REQUIRE(out != nullptr, "tfh_uniform: null output");
  TFH_TRY(require_device());
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_uniform, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), seed,
                     offset, n, out);
  return check_launch("tfh_uniform");
}

int32_t tfh_adam_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                      float lr, float decay, float b1, float b2, float eps, int64_t count,
                      int32_t zero_grad, tfh_stream_t stream) {
  TFH_REQUIRE(n >= 0, "tfh_adam_step: negative n");
  if (n == 0) return TFH_OK;
  TFH_REQUIRE(param && grad && exp_avg && exp_avg_sq, "tfh_adam_step: null pointer");
  TFH_REQUIRE(count >= 1, "tfh_adam_step: count must be >= 1 (got %lld)", (long long)count);
  TFH_TRY(require_device());
  return adam_launch(param, grad, exp_avg, exp_avg_sq, n, n, lr, lr, decay, b1, b2, eps, count,
                     zero_grad, as_stream(stream), nullptr);
}

}  // extern "C"
