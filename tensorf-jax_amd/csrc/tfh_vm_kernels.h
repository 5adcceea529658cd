// Templated VM gather / scatter kernels and their launchers. A point source
// functor maps a row to its normalised sample point:
//   PlanarPoints  : ijk (3, M), component stride M (TensorVM.interpolate API)
//   RaySamples    : pts (R*N, 3) interleaved, row = r*k + j selects sample
//                   idx[r*k + j] (or j when idx == NULL) of ray r.
// ROW_MAJOR selects a (M, 3C) feature layout (the MLP input) instead of the
// reference's (3C, M).
#pragma once

#include "tfh_vm.h"

namespace tfh {

// Each source also gives the number of rows it holds out of the launched M
// (a device-side count for CompactedPoints).
struct PlanarPoints {
  const float* p;
  int64_t m;
  __device__ __forceinline__ int64_t limit(int64_t M) const { return M; }
  __device__ __forceinline__ void get(int64_t s, float& x, float& y, float& z) const {
    x = p[s];
    y = p[m + s];
    z = p[2 * m + s];
  }
};

struct RaySamples {
  const float* pts;    // (R*N, 3)
  const int32_t* idx;  // (R, k) or null
  int64_t k, n;
  __device__ __forceinline__ int64_t limit(int64_t M) const { return M; }
  __device__ __forceinline__ void get(int64_t row, float& x, float& y, float& z) const {
    const int64_t r = row / k;
    const int64_t j = row - r * k;
    const int64_t i = idx ? (int64_t)idx[row] : j;
    const float* q = pts + (r * n + i) * 3;
    x = q[0];
    y = q[1];
    z = q[2];
  }
};

// Rows listed by a compaction: row r is point idx[r] of pts (N, 3); only the
// first *count rows exist.
struct CompactedPoints {
  const float* pts;
  const int32_t* idx;
  const int32_t* count;
  __device__ __forceinline__ int64_t limit(int64_t M) const {
    return min(M, (int64_t)*count);
  }
  __device__ __forceinline__ void get(int64_t row, float& x, float& y, float& z) const {
    const float* q = pts + (int64_t)idx[row] * 3;
    x = q[0];
    y = q[1];
    z = q[2];
  }
};

constexpr int kGatherThreads = 256;

// Gather: LPS lanes per sample, 256/LPS samples per block.
template <class Src, bool ROW_MAJOR, int LPS>
__global__ void __launch_bounds__(kGatherThreads) k_vm_gather(tfh_vm vm, Src src, int64_t M,
                                                              float* __restrict__ out) {
  constexpr int SPB = kGatherThreads / LPS;
  constexpr int NQ = 64 / LPS;  // channel slots per lane (channels <= 64)
  __shared__ float tile[ROW_MAJOR ? 1 : 3 * 64][SPB + 1];
  const int grp = threadIdx.x / LPS, l = threadIdx.x % LPS;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  const int64_t s = s0 + grp;
  const bool active = s < M;
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  float x = 0.f, y = 0.f, z = 0.f;
  if (active) src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, G, mt);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    TapAddr a;
    tap_addrs(mt[m], m, G, a);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = l + LPS * q;
      if (c < C && active) {
        float lv, pv;
        tap_values(vm.plane, vm.line, C, c, a, lv, pv);
        const float f = lv * pv;  // tensor_vm.py:167
        if (ROW_MAJOR)
          out[s * 3 * C + m * C + c] = f;
        else
          tile[m * C + c][grp] = f;
      }
    }
  }
  if (!ROW_MAJOR) {
    __syncthreads();
    const int nch = 3 * C;
    for (int e = threadIdx.x; e < nch * SPB; e += kGatherThreads) {
      const int ch = e / SPB, j = e % SPB;
      if (s0 + j < M) out[(int64_t)ch * M + s0 + j] = tile[ch][j];
    }
  }
}

// Row-major gather with 16-byte taps: LPS lanes per sample, lane l owns
// channels 4l .. 4l+3 of every mode (LPS = C/4 rounded up to a power of
// two, or 12 for 48 channels: five groups per wave and 4 lanes idle, where
// 16 lanes per sample leave a quarter of them idle, and every lane repeats
// its sample's tap and address work); a quarter of the load instructions of
// k_vm_gather, the same values.
template <class Src, int LPS>
__global__ void __launch_bounds__(kGatherThreads) k_vm_gather4(tfh_vm vm, Src src, int64_t M,
                                                               float* __restrict__ out) {
  constexpr int GPW = 64 / LPS, SPB = (kGatherThreads / 64) * GPW;  // groups per wave, block
  const int lane = threadIdx.x & 63;
  const int grp = (threadIdx.x >> 6) * GPW + lane / LPS, l = lane % LPS;
  const int64_t s = (int64_t)blockIdx.x * SPB + grp;
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  const int c4 = 4 * l;
  if (lane >= GPW * LPS || s >= M || c4 >= C) return;  // no barrier or lane exchange below
  float x = 0.f, y = 0.f, z = 0.f;
  src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, G, mt);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    TapAddr a;
    tap_addrs(mt[m], m, G, a);
    float4 lv, pv;
    tap_values4(vm.plane, vm.line, C, c4, a, lv, pv);
    // tensor_vm.py:167
    *reinterpret_cast<float4*>(out + s * 3 * C + m * C + c4) =
        make_float4(lv.x * pv.x, lv.y * pv.y, lv.z * pv.z, lv.w * pv.w);
  }
}

template <class Src, bool ROW_MAJOR>
// dense: the rows are samples along rays with more samples than the grid has
// cells, whose neighbours in a wave share most texels; 48 channels then take
// 12 lanes per sample (uniform-512 gather 0.55 -> 0.47 ms), while sparser
// rows keep 16 (proposal-64: 0.40 against 0.42 ms with 12, the five samples
// of a wave touching more distinct lines per load).
inline int32_t launch_vm_gather(const tfh_vm& vm, const Src& src, int64_t M, float* out,
                                hipStream_t st, bool dense = false) {
  const int C = (int)vm.channels;
  if (ROW_MAJOR && C % 4 == 0 && C <= 64 && (reinterpret_cast<uintptr_t>(out) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(vm.plane) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(vm.line) & 15) == 0) {
    const int lps = C <= 8 ? 2 : (C <= 16 ? 4 : (C <= 32 ? 8 : (C == 48 && dense ? 12 : 16)));
    const int64_t blocks = ceil_div(M, (kGatherThreads / 64) * (64 / lps));
    TFH_REQUIRE(blocks < (1ll << 31), "vm gather: too many points (%lld)", (long long)M);
#define TFH_G4(L)                                                                            \
  hipLaunchKernelGGL((k_vm_gather4<Src, L>), dim3((unsigned)blocks), dim3(kGatherThreads), 0, \
                     st, vm, src, M, out)
    if (lps == 2) TFH_G4(2);
    else if (lps == 4) TFH_G4(4);
    else if (lps == 8) TFH_G4(8);
    else if (lps == 12) TFH_G4(12);
    else TFH_G4(16);
#undef TFH_G4
    return check_launch("vm gather");
  }
  const bool narrow = vm.channels <= 8;
  const int spb = kGatherThreads / (narrow ? 8 : 16);
  const int64_t blocks = ceil_div(M, spb);
  TFH_REQUIRE(blocks < (1ll << 31), "vm gather: too many points (%lld)", (long long)M);
  if (narrow)
    hipLaunchKernelGGL((k_vm_gather<Src, ROW_MAJOR, 8>), dim3((unsigned)blocks),
                       dim3(kGatherThreads), 0, st, vm, src, M, out);
  else
    hipLaunchKernelGGL((k_vm_gather<Src, ROW_MAJOR, 16>), dim3((unsigned)blocks),
                       dim3(kGatherThreads), 0, st, vm, src, M, out);
  return check_launch("vm gather");
}

// VJP: grad (per sample, per channel) -> atomic scatter into the factor
// gradients. Each wave holds 64/LPS consecutive samples; equal texels among
// them, in any tap slot, are merged before the atomic (tap_plan). With
// npc / nlc > 1, grad holds that many consecutive copies of the plane / line
// gradient and block b adds into copy b % npc / b % nlc: float atomics are
// performed at the memory side, where adds to one address serialise, and a
// small field (the proposal's) draws thousands of adds per texel a step.
template <class Src, bool ROW_MAJOR, int LPS>
__global__ void __launch_bounds__(kGatherThreads) k_vm_scatter(tfh_vm vm, Src src, int64_t M,
                                                               const float* __restrict__ gout,
                                                               tfh_vm_grad grad, int npc,
                                                               int nlc) {
  constexpr int SPB = kGatherThreads / LPS;
  constexpr int NQ = 64 / LPS;
  __shared__ float tile[ROW_MAJOR ? 1 : 3 * 64][SPB + 1];
  const int grp = threadIdx.x / LPS, l = threadIdx.x % LPS;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  M = src.limit(M);
  if (s0 >= M) return;  // the whole block: no wave of it joins a merge
  const int64_t s = s0 + grp;
  const bool active = s < M;
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  grad.plane += (int64_t)(blockIdx.x % npc) * 3 * G * G * C;
  grad.line += (int64_t)(blockIdx.x % nlc) * 3 * G * C;
  if (!ROW_MAJOR) {
    const int nch = 3 * C;
    for (int e = threadIdx.x; e < nch * SPB; e += kGatherThreads) {
      const int ch = e / SPB, j = e % SPB;
      tile[ch][j] = (s0 + j < M) ? gout[(int64_t)ch * M + s0 + j] : 0.0f;
    }
    __syncthreads();
  }
  float x = 0.f, y = 0.f, z = 0.f;
  if (active) src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, G, mt);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    TapAddr a;
    tap_addrs(mt[m], m, G, a);
    // cross-slot merge for 16-lane groups (4 samples a wave); 8-lane groups
    // (narrow fields) merge per slot
    TapPlan<LPS> plan;
    if constexpr (LPS == 16) tap_plan<LPS>(a, active, plan);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = l + LPS * q;
      if (LPS * q < C) {  // wave-uniform: every lane joins the merged atomics
        const bool on = active && c < C;
        float lv = 0.f, pv = 0.f, g = 0.f;
        if (on) {
          tap_values(vm.plane, vm.line, C, c, a, lv, pv);
          g = ROW_MAJOR ? gout[s * 3 * C + m * C + c] : tile[m * C + c][grp];
        }
        if constexpr (LPS == 16)
          tap_scatter_plan<LPS>(grad.plane, grad.line, C, c, a, plan, g * pv, g * lv, on);
        else
          tap_scatter<LPS>(grad.plane, grad.line, C, c, a, g * pv, g * lv, on);
      }
    }
  }
}

// Gradient copies of a scatter (npc / nlc above): clear n floats; add the sum
// of ncopy consecutive copies of len floats into dst (tfh_vm.hip).
int32_t zero_copies(float* p, int64_t n, hipStream_t st);
// partial sums of a per-ray scalar (a loss), added up by sum_copies
constexpr int kLossSlots = 256;
int32_t sum_copies(const float* copies, int ncopy, int64_t len, float* dst, hipStream_t st);

template <class Src, bool ROW_MAJOR>
inline int32_t launch_vm_scatter(const tfh_vm& vm, const Src& src, int64_t M, const float* gout,
                                 const tfh_vm_grad& grad, hipStream_t st, int npc = 1,
                                 int nlc = 1) {
  const bool narrow = vm.channels <= 8;
  const int spb = kGatherThreads / (narrow ? 8 : 16);
  const int64_t blocks = ceil_div(M, spb);
  TFH_REQUIRE(blocks < (1ll << 31), "vm scatter: too many points (%lld)", (long long)M);
  if (narrow)
    hipLaunchKernelGGL((k_vm_scatter<Src, ROW_MAJOR, 8>), dim3((unsigned)blocks),
                       dim3(kGatherThreads), 0, st, vm, src, M, gout, grad, npc, nlc);
  else
    hipLaunchKernelGGL((k_vm_scatter<Src, ROW_MAJOR, 16>), dim3((unsigned)blocks),
                       dim3(kGatherThreads), 0, st, vm, src, M, gout, grad, npc, nlc);
  return check_launch("vm scatter");
}

}  // namespace tfh
