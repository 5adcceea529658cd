// TensorVM.interpolate (tensor_vm.py:42-89, 140-167, 226-250), its VJP and
// TensorVM.resize (tensor_vm.py:91-100, 183-223) as gfx950 kernels.
#include <algorithm>

#include "tfh_vm.h"
#include "tfh_vm_kernels.h"

namespace tfh {

// Align-corners linear resampling of the channel-last factors; one thread per
// output float, channel fastest (coalesced). Output texel o samples input
// position o*(G-1)/(G'-1) (jax.image.scale_and_translate with
// scale=(G'-1)/(G-1), translation=0.5-scale/2, tensor_vm.py:210-223).
__device__ __forceinline__ void resample_pos(int64_t o, int64_t gin, int64_t gout, int& i0,
                                             int& i1, float& f) {
  double x = gout > 1 ? (double)o * (double)(gin - 1) / (double)(gout - 1) : 0.0;
  int i = (int)floor(x);
  if (i > gin - 1) i = (int)gin - 1;
  if (i < 0) i = 0;
  f = (float)(x - (double)i);
  i0 = i;
  i1 = i + 1 < gin ? i + 1 : (int)gin - 1;
}

__global__ void __launch_bounds__(256) k_resize_plane(const float* __restrict__ src,
                                                      float* __restrict__ dst, int64_t C,
                                                      int64_t gin, int64_t gout) {
  const int64_t total = 3 * gout * gout * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % C;
    int64_t t = e / C;
    const int64_t col = t % gout;
    t /= gout;
    const int64_t row = t % gout;
    const int64_t m = t / gout;
    int r0, r1, c0, c1;
    float fr, fc;
    resample_pos(row, gin, gout, r0, r1, fr);
    resample_pos(col, gin, gout, c0, c1, fc);
    const float* s = src + (size_t)m * gin * gin * C;
    const float v00 = s[((size_t)r0 * gin + c0) * C + c];
    const float v01 = s[((size_t)r0 * gin + c1) * C + c];
    const float v10 = s[((size_t)r1 * gin + c0) * C + c];
    const float v11 = s[((size_t)r1 * gin + c1) * C + c];
    // separable: along columns first, then rows
    const float top = (1.0f - fc) * v00 + fc * v01;
    const float bot = (1.0f - fc) * v10 + fc * v11;
    dst[e] = (1.0f - fr) * top + fr * bot;
  }
}

__global__ void __launch_bounds__(256) k_resize_line(const float* __restrict__ src,
                                                     float* __restrict__ dst, int64_t C,
                                                     int64_t gin, int64_t gout) {
  const int64_t total = 3 * gout * C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % C;
    int64_t t = e / C;
    const int64_t o = t % gout;
    const int64_t m = t / gout;
    int i0, i1;
    float f;
    resample_pos(o, gin, gout, i0, i1, f);
    const float* s = src + (size_t)m * gin * C;
    dst[e] = (1.0f - f) * s[(size_t)i0 * C + c] + f * s[(size_t)i1 * C + c];
  }
}

__global__ void __launch_bounds__(256) k_zero_copies(float* __restrict__ p, int64_t n) {
  for (int64_t e = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4; e < n;
       e += (int64_t)gridDim.x * 256 * 4) {
    if (e + 4 <= n)
      *reinterpret_cast<float4*>(p + e) = make_float4(0.f, 0.f, 0.f, 0.f);
    else
      for (int64_t j = e; j < n; ++j) p[j] = 0.0f;
  }
}

// A workgroup adds up 64 consecutive entries at a time; its kSumWaves waves
// split the copies (wave w adds copies w, w + kSumWaves, ...), so every load
// is one coalesced 256-byte segment, and the waves' sums meet in LDS in a
// fixed order. The chunk loop's trip count is the same for every thread.
constexpr int kSumWaves = 16;

__global__ void __launch_bounds__(64 * kSumWaves) k_sum_copies(const float* __restrict__ copies,
                                                              int ncopy, int64_t len,
                                                              float* __restrict__ dst) {
  __shared__ float red[kSumWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t e0 = (int64_t)blockIdx.x * 64; e0 < len; e0 += (int64_t)gridDim.x * 64) {
    const int64_t e = e0 + lane;
    float part = 0.0f;
    if (e < len) {
#pragma unroll 4
      for (int k = w; k < ncopy; k += kSumWaves) part += copies[(int64_t)k * len + e];
    }
    red[w][lane] = part;
    __syncthreads();
    if (w == 0 && e < len) {
      float sum = 0.0f;
#pragma unroll
      for (int i = 0; i < kSumWaves; ++i) sum += red[i][lane];
      dst[e] += sum;
    }
    __syncthreads();
  }
}

// p must be 16-byte aligned (a workspace region).
int32_t zero_copies(float* p, int64_t n, hipStream_t st) {
  if (n <= 0) return TFH_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div(n, 1024), 4096);
  hipLaunchKernelGGL(k_zero_copies, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
  return check_launch("clear gradient copies");
}

int32_t sum_copies(const float* copies, int ncopy, int64_t len, float* dst, hipStream_t st) {
  if (len <= 0) return TFH_OK;
  const int64_t blocks = std::min<int64_t>(ceil_div(len, 64), 8192);
  hipLaunchKernelGGL(k_sum_copies, dim3((unsigned)blocks), dim3(64 * kSumWaves), 0, st, copies,
                     ncopy, len, dst);
  return check_launch("sum gradient copies");
}

}  // namespace tfh

using namespace tfh;

extern "C" {

int32_t tfh_vm_interpolate(const tfh_vm* vm, const float* ijk, int64_t num_points, float* out,
                           tfh_stream_t stream) {
  TFH_TRY(check_vm(vm, "tfh_vm_interpolate"));
  TFH_REQUIRE(num_points >= 0, "tfh_vm_interpolate: negative num_points");
  if (num_points == 0) return TFH_OK;
  TFH_REQUIRE(ijk && out, "tfh_vm_interpolate: null pointer");
  TFH_TRY(require_device());
  PlanarPoints src{ijk, num_points};
  return launch_vm_gather<PlanarPoints, false>(*vm, src, num_points, out, as_stream(stream));
}

int32_t tfh_vm_interpolate_bwd(const tfh_vm* vm, const float* ijk, int64_t num_points,
                               const float* grad_out, const tfh_vm_grad* grad,
                               tfh_stream_t stream) {
  TFH_TRY(check_vm(vm, "tfh_vm_interpolate_bwd"));
  TFH_REQUIRE(num_points >= 0, "tfh_vm_interpolate_bwd: negative num_points");
  if (num_points == 0) return TFH_OK;
  TFH_REQUIRE(ijk && grad_out && grad && grad->plane && grad->line,
              "tfh_vm_interpolate_bwd: null pointer");
  TFH_TRY(require_device());
  PlanarPoints src{ijk, num_points};
  return launch_vm_scatter<PlanarPoints, false>(*vm, src, num_points, grad_out, *grad,
                                                as_stream(stream));
}

int32_t tfh_vm_resize(const tfh_vm* vm, int64_t grid_dim_out, float* out_plane, float* out_line,
                      tfh_stream_t stream) {
  TFH_TRY(check_vm(vm, "tfh_vm_resize"));
  TFH_REQUIRE(out_plane && out_line, "tfh_vm_resize: null output");
  TFH_REQUIRE(grid_dim_out >= vm->grid_dim && grid_dim_out <= 4096,
              "tfh_vm_resize: grid_dim_out %lld must be in [G=%lld, 4096] (upsampling only)",
              (long long)grid_dim_out, (long long)vm->grid_dim);
  TFH_TRY(require_device());
  const int64_t C = vm->channels, gin = vm->grid_dim, gout = grid_dim_out;
  int64_t blocks = ceil_div(3 * gout * gout * C, 256);
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_resize_plane, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                     vm->plane, out_plane, C, gin, gout);
  TFH_TRY(check_launch("tfh_vm_resize(plane)"));
  blocks = ceil_div(3 * gout * C, 256);
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_resize_line, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                     vm->line, out_line, C, gin, gout);
  return check_launch("tfh_vm_resize(line)");
}

}  // extern "C"
