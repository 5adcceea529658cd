// MLP of the proposal density field (no counterpart in the reference):
// z = W3 relu(W2 relu(W1 f + b1) + b2) + b3, sigma = softplus(z), with
// f the 3C <= 32 VM features of a sample, 64 hidden units, FP32 MFMA.
// Forward: one 64-row tile per workgroup. Backward (input dL/dz): persistent
// workgroups recompute the forward of each tile, keep the 28 weight-gradient
// 16x16 tiles in registers (7 per wave) and store them once per workgroup;
// k_prop_reduce sums the workgroups.
#include <algorithm>

#include "tfh_mfma.h"

namespace tfh {

constexpr int kPH = 64;       // hidden width
constexpr int kPF = 32;       // padded feature width
constexpr int kPTM = 64;      // rows per tile
constexpr int kPNW = 4;       // waves per workgroup
constexpr int kPLdF = kPF + 2, kPLdH = kPH + 2, kPLdZ = 16 + 2;

struct PropW {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
  int F;
};

// packed gradient partial: w1 (F x 64) | b1 | w2 (64 x 64) | b2 | w3 (64) | b3
struct PropPack {
  int w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline PropPack prop_pack(int F) {
  PropPack p;
  int o = 0;
  p.w1 = o; o += F * kPH;
  p.b1 = o; o += kPH;
  p.w2 = o; o += kPH * kPH;
  p.b2 = o; o += kPH;
  p.w3 = o; o += kPH;
  p.b3 = o; o += 1;
  p.total = (o + 63) / 64 * 64;
  return p;
}

// Row `row` of the tile reads feature row idx[row] (or row when idx is null).
__device__ __forceinline__ void prop_load(const float* __restrict__ feats, int F, int64_t row0,
                                          int64_t M, float* Fs,
                                          const int32_t* __restrict__ idx = nullptr) {
  for (int e = threadIdx.x; e < kPTM * kPF; e += blockDim.x) {
    const int i = e / kPF, c = e % kPF;
    const int64_t row = row0 + i;
    float v = 0.0f;
    if (row < M && c < F) v = feats[(idx ? (int64_t)idx[row] : row) * F + c];
    Fs[i * kPLdF + c] = v;
  }
}

// Hidden layers of a tile: H1 = relu(Fs W1 + b1), H2 = relu(H1 W2 + b2).
__device__ __forceinline__ void prop_hidden(const PropW& w, const float* Fs, float* H1,
                                            float* H2) {
  constexpr int RT = kPTM / 16;
  const int wave = threadIdx.x >> 6, lane = lane_id();
  const int ct = wave;  // 4 column tiles, one per wave
  {
    f32x4 acc[RT];
    zero_acc(acc);
    mma_strip<RT, kPF>(Fs, kPLdF, kPF, GlobalB(w.w1, w.F, kPH), 16 * ct, acc);
    const int col = 16 * ct + (lane & 15);
    const float bias = w.b1[col];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        H1[(16 * rt + 4 * (lane >> 4) + e) * kPLdH + col] = fmaxf(acc[rt][e] + bias, 0.0f);
  }
  __syncthreads();
  {
    f32x4 acc[RT];
    zero_acc(acc);
    mma_strip<RT, kPH>(H1, kPLdH, kPH, GlobalB(w.w2, kPH, kPH), 16 * ct, acc);
    const int col = 16 * ct + (lane & 15);
    const float bias = w.b2[col];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        H2[(16 * rt + 4 * (lane >> 4) + e) * kPLdH + col] = fmaxf(acc[rt][e] + bias, 0.0f);
  }
  __syncthreads();
}

// z of row tile `wave` (16 rows) into Z[row] (column 0 of the D tile).
__device__ __forceinline__ void prop_out(const PropW& w, const float* H2, float* Z) {
  const int wave = threadIdx.x >> 6, lane = lane_id();
  f32x4 acc[1];
  zero_acc(acc);
  mma_strip<1, kPH>(H2 + 16 * wave * kPLdH, kPLdH, kPH, GlobalB(w.w3, kPH, 1), 0, acc);
  if ((lane & 15) == 0) {
    const float b3 = w.b3[0];
#pragma unroll
    for (int e = 0; e < 4; ++e) Z[16 * wave + 4 * (lane >> 4) + e] = acc[0][e] + b3;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kPNW * 64) k_prop_mlp_fwd(PropW w, const float* __restrict__ feats,
                                                            int64_t M, float* __restrict__ z,
                                                            float* __restrict__ sigma) {
  __shared__ float Fs[kPTM * kPLdF];
  __shared__ float H1[kPTM * kPLdH];
  __shared__ float H2[kPTM * kPLdH];
  __shared__ float Z[kPTM];
  const int64_t row0 = (int64_t)blockIdx.x * kPTM;
  prop_load(feats, w.F, row0, M, Fs);
  __syncthreads();
  prop_hidden(w, Fs, H1, H2);
  prop_out(w, H2, Z);
  for (int i = threadIdx.x; i < kPTM; i += blockDim.x) {
    const int64_t row = row0 + i;
    if (row < M) {
      const float zv = Z[i];
      if (z) z[row] = zv;
      sigma[row] = softplus(zv);
    }
  }
}

// Register slots per wave: dW3 4 tiles (64x16), dW2 16 (64x64), dW1 8 (32x64).
constexpr int kPC3 = 1, kPC2 = 4, kPC1 = 2, kPSlots = kPC3 + kPC2 + kPC1;

__device__ __forceinline__ void prop_store_tile(float* part, int off, int rows, int cols, int row0,
                                                int col0, const f32x4& acc) {
  const int lane = lane_id();
  const int col = col0 + (lane & 15);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int row = row0 + 4 * (lane >> 4) + e;
    if (row < rows && col < cols) part[off + row * cols + col] = acc[e];
  }
}

// Rows: all M, or the first min(M, *count) entries of idx (feature and dz row
// of each); dfeat is written per processed row.
__global__ void __launch_bounds__(kPNW * 64, 2) k_prop_mlp_bwd(PropW w, const float* __restrict__ feats,
                                                            const float* __restrict__ dz,
                                                            int64_t M_max,
                                                            const int32_t* __restrict__ idx,
                                                            const int32_t* __restrict__ count,
                                                            float* __restrict__ dfeat,
                                                            float* __restrict__ partials) {
  const int64_t M = count ? min(M_max, (int64_t)*count) : M_max;
  constexpr int RT = kPTM / 16;
  __shared__ float Fs[kPTM * kPLdF];
  __shared__ float H1[kPTM * kPLdH];
  __shared__ float H2[kPTM * kPLdH];
  __shared__ float DZ[kPTM * kPLdZ];
  const int wave = threadIdx.x >> 6, lane = lane_id();
  const PropPack pp = prop_pack(w.F);
  f32x4 wacc[kPSlots];
#pragma unroll
  for (int i = 0; i < kPSlots; ++i) wacc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db1 = 0.f, db2 = 0.f, db3 = 0.f;
  const int64_t ntiles = (M + kPTM - 1) / kPTM;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t row0 = tile * kPTM;
    __syncthreads();
    prop_load(feats, w.F, row0, M, Fs, idx);
    for (int e = threadIdx.x; e < kPTM * 16; e += blockDim.x) {
      const int i = e / 16, c = e % 16;
      const int64_t row = row0 + i;
      DZ[i * kPLdZ + c] = (c == 0 && row < M) ? dz[idx ? (int64_t)idx[row] : row] : 0.0f;
    }
    __syncthreads();
    prop_hidden(w, Fs, H1, H2);
    // dW3 (64 x 1) += H2^T dZ: row tile = wave
    mma_wgrad_acc<false, kPTM>(H2, kPLdH, 16 * wave, DZ, kPLdZ, 0, wacc[0]);
    __builtin_amdgcn_sched_barrier(0);
    if (threadIdx.x == 0)
      for (int i = 0; i < kPTM; ++i) db3 += DZ[i * kPLdZ];
    __syncthreads();
    // dH2 = dZ W3^T; dZ2 = dH2 * (h2 > 0) over H2 (dW3 is done)
    {
      f32x4 acc[RT];
      zero_acc(acc);
      mma_strip<RT, 4>(DZ, kPLdZ, 4, GlobalBT(w.w3, kPH, 1), 16 * wave, acc);
      const int col = 16 * wave + (lane & 15);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rt + 4 * (lane >> 4) + e;
          const float h = H2[row * kPLdH + col];
          H2[row * kPLdH + col] = h > 0.0f ? acc[rt][e] : 0.0f;
        }
    }
    __syncthreads();
    // dW2 (64 x 64) += H1^T dZ2: wave owns column tile `wave`, all 4 row tiles
#pragma unroll
    for (int s = 0; s < kPC2; ++s) {
      mma_wgrad_acc<false, kPTM>(H1, kPLdH, 16 * s, H2, kPLdH, 16 * wave, wacc[kPC3 + s]);
      // keep the operand loads of later tiles from being hoisted
      __builtin_amdgcn_sched_barrier(0);
    }
    if (threadIdx.x < kPH)
      for (int i = 0; i < kPTM; ++i) db2 += H2[i * kPLdH + threadIdx.x];
    __syncthreads();
    // dH1 = dZ2 W2^T; dZ1 = dH1 * (h1 > 0) over H1
    {
      f32x4 acc[RT];
      zero_acc(acc);
      mma_strip<RT, kPH>(H2, kPLdH, kPH, GlobalBT(w.w2, kPH, kPH), 16 * wave, acc);
      const int col = 16 * wave + (lane & 15);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rt + 4 * (lane >> 4) + e;
          const float h = H1[row * kPLdH + col];
          H1[row * kPLdH + col] = h > 0.0f ? acc[rt][e] : 0.0f;
        }
    }
    __syncthreads();
    // dW1 (32 x 64) += Fs^T dZ1: wave owns column tile `wave`, 2 row tiles
#pragma unroll
    for (int s = 0; s < kPC1; ++s) {
      mma_wgrad_acc<false, kPTM>(Fs, kPLdF, 16 * s, H1, kPLdH, 16 * wave,
                                 wacc[kPC3 + kPC2 + s]);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (threadIdx.x < kPH)
      for (int i = 0; i < kPTM; ++i) db1 += H1[i * kPLdH + threadIdx.x];
    // dF = dZ1 W1^T (64 x F): 2 column tiles on waves 0, 1
    if (dfeat && wave < 2) {
      f32x4 acc[RT];
      zero_acc(acc);
      mma_strip<RT, kPH>(H1, kPLdH, kPH, GlobalBT(w.w1, w.F, kPH), 16 * wave, acc);
      const int col = 16 * wave + (lane & 15);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t row = row0 + 16 * rt + 4 * (lane >> 4) + e;
          if (row < M && col < w.F) dfeat[row * w.F + col] = acc[rt][e];
        }
    }
  }
  float* part = partials + (int64_t)blockIdx.x * pp.total;
  prop_store_tile(part, pp.w3, kPH, 1, 16 * wave, 0, wacc[0]);
#pragma unroll
  for (int s = 0; s < kPC2; ++s)
    prop_store_tile(part, pp.w2, kPH, kPH, 16 * s, 16 * wave, wacc[kPC3 + s]);
#pragma unroll
  for (int s = 0; s < kPC1; ++s)
    prop_store_tile(part, pp.w1, w.F, kPH, 16 * s, 16 * wave, wacc[kPC3 + kPC2 + s]);
  if (threadIdx.x < kPH) {
    part[pp.b1 + threadIdx.x] = db1;
    part[pp.b2 + threadIdx.x] = db2;
  }
  if (threadIdx.x == 0) part[pp.b3] = db3;
}

// Sum of the workgroup partials: a workgroup owns 64 consecutive entries and
// its kPRedWaves waves split the partials (wave w adds rows w, w + kPRedWaves,
// ...), so every load is one coalesced 256-byte row segment; the waves' sums
// meet in LDS and are added in a fixed order.
constexpr int kPRedWaves = 16;

__global__ void __launch_bounds__(64 * kPRedWaves) k_prop_reduce(const float* __restrict__ partials,
                                                                int64_t nblk, PropPack pp,
                                                                tfh_prop_mlp_grad g) {
  __shared__ float red[kPRedWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = pp.b3 + 1;
  const int e = blockIdx.x * 64 + lane;
  float part = 0.0f;
  if (e < n) {
#pragma unroll 8
    for (int64_t b = w; b < nblk; b += kPRedWaves) part += partials[b * pp.total + e];
  }
  red[w][lane] = part;
  __syncthreads();
  if (w == 0 && e < n) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kPRedWaves; ++i) s += red[i][lane];
    float* dst;
    int o;
    if (e >= pp.b3) { dst = g.b3; o = e - pp.b3; }
    else if (e >= pp.w3) { dst = g.w3; o = e - pp.w3; }
    else if (e >= pp.b2) { dst = g.b2; o = e - pp.b2; }
    else if (e >= pp.w2) { dst = g.w2; o = e - pp.w2; }
    else if (e >= pp.b1) { dst = g.b1; o = e - pp.b1; }
    else { dst = g.w1; o = e - pp.w1; }
    dst[o] += s;
  }
}

static int prop_num_cus() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      n = v;
    else
      n = 256;
  }
  return n;
}

// Persistent backward: two workgroups per CU (the kernel's register budget
// is sized for that occupancy).
static int64_t prop_bwd_blocks(int64_t M) {
  return std::max<int64_t>(1, std::min<int64_t>(ceil_div(M, kPTM), 2 * (int64_t)prop_num_cus()));
}

int32_t check_prop_mlp(const tfh_prop_mlp* p, const char* who) {
  TFH_REQUIRE(p != nullptr, "%s: null tfh_prop_mlp", who);
  TFH_REQUIRE(p->w1 && p->b1 && p->w2 && p->b2 && p->w3 && p->b3, "%s: null proposal weight", who);
  TFH_REQUIRE(p->feat_dim >= 1 && p->feat_dim <= kPF, "%s: proposal feat_dim %lld outside [1, %d]",
              who, (long long)p->feat_dim, kPF);
  return TFH_OK;
}

static PropW to_pw(const tfh_prop_mlp* p) {
  return PropW{p->w1, p->b1, p->w2, p->b2, p->w3, p->b3, (int)p->feat_dim};
}

int32_t prop_mlp_fwd(const tfh_prop_mlp* p, const float* feats, int64_t M, float* z, float* sigma,
                     hipStream_t st) {
  if (M == 0) return TFH_OK;
  ProfScope ps("prop_mlp_fwd", st);
  hipLaunchKernelGGL(k_prop_mlp_fwd, dim3((unsigned)ceil_div(M, kPTM)), dim3(kPNW * 64), 0, st,
                     to_pw(p), feats, M, z, sigma);
  return check_launch("proposal mlp fwd");
}

size_t prop_mlp_bwd_workspace(const tfh_prop_mlp* p, int64_t M) {
  const int64_t nblk = prop_bwd_blocks(M);
  return align_up((size_t)nblk * prop_pack((int)p->feat_dim).total * sizeof(float), 256);
}

int32_t prop_mlp_bwd(const tfh_prop_mlp* p, const float* feats, const float* dz, int64_t M,
                     float* dfeat, const tfh_prop_mlp_grad* g, void* ws, size_t ws_bytes,
                     hipStream_t st, const int32_t* idx = nullptr,
                     const int32_t* count = nullptr) {
  if (M == 0) return TFH_OK;
  const int64_t nblk = prop_bwd_blocks(M);
  const PropPack pp = prop_pack((int)p->feat_dim);
  TFH_REQUIRE(ws_bytes >= prop_mlp_bwd_workspace(p, M), "proposal mlp bwd: workspace too small");
  float* partials = static_cast<float*>(ws);
  {
    ProfScope ps("prop_mlp_bwd", st);
    hipLaunchKernelGGL(k_prop_mlp_bwd, dim3((unsigned)nblk), dim3(kPNW * 64), 0, st, to_pw(p),
                       feats, dz, M, idx, count, dfeat, partials);
  }
  TFH_TRY(check_launch("proposal mlp bwd"));
  hipLaunchKernelGGL(k_prop_reduce, dim3((unsigned)ceil_div(pp.b3 + 1, 64)),
                     dim3(64 * kPRedWaves), 0, st, partials, nblk, pp, *g);
  return check_launch("proposal mlp reduce");
}

}  // namespace tfh

using namespace tfh;

extern "C" {

int32_t tfh_prop_mlp_fwd(const tfh_prop_mlp* mlp, const float* features, int64_t num_rows,
                         float* z, float* sigma, tfh_stream_t stream) {
  TFH_TRY(check_prop_mlp(mlp, "tfh_prop_mlp_fwd"));
  TFH_REQUIRE(num_rows >= 0, "tfh_prop_mlp_fwd: negative num_rows");
  if (num_rows == 0) return TFH_OK;
  TFH_REQUIRE(features && sigma, "tfh_prop_mlp_fwd: null pointer");
  TFH_TRY(require_device());
  return prop_mlp_fwd(mlp, features, num_rows, z, sigma, as_stream(stream));
}

size_t tfh_prop_mlp_bwd_workspace_size(const tfh_prop_mlp* mlp, int64_t num_rows) {
  if (check_prop_mlp(mlp, "tfh_prop_mlp_bwd_workspace_size") != TFH_OK || num_rows < 0) return 0;
  return prop_mlp_bwd_workspace(mlp, num_rows);
}

int32_t tfh_prop_mlp_bwd(const tfh_prop_mlp* mlp, const float* features, int64_t num_rows,
                         const float* grad_z, float* grad_features,
                         const tfh_prop_mlp_grad* grad, void* workspace, size_t workspace_bytes,
                         tfh_stream_t stream) {
  TFH_TRY(check_prop_mlp(mlp, "tfh_prop_mlp_bwd"));
  TFH_REQUIRE(num_rows >= 0, "tfh_prop_mlp_bwd: negative num_rows");
  if (num_rows == 0) return TFH_OK;
  TFH_REQUIRE(features && grad_z && grad && grad->w1 && grad->b1 && grad->w2 && grad->b2 &&
                  grad->w3 && grad->b3 && workspace,
              "tfh_prop_mlp_bwd: null pointer");
  TFH_TRY(require_device());
  return prop_mlp_bwd(mlp, features, grad_z, num_rows, grad_features, grad, workspace,
                      workspace_bytes, as_stream(stream));
}

}  // extern "C"
