// MFMA tile helpers shared by the MLP kernels (gfx950, wave64).
//
// D layout of a 16x16 tile: lane l, element r -> row 4*(l/16)+r, col l%16.
// FP32 (K=4):  A[l%16][k0 + l/16],        B[k0 + l/16][l%16]
// BF16 (K=32): A[l%16][k0 + 8(l/16) + e], B[k0 + 8(l/16) + e][l%16], e < 8
#pragma once

#include <type_traits>

#include "tfh_common.h"

namespace tfh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));


// The lane index behind an empty volatile asm: each call computes it afresh,
// so operand addresses derived from it are computed where a tile product
// runs. With the plain index, loop-invariant code motion hoists every
// product's addresses out of a persistent kernel's tile loop, where a kernel
// at its register limit spills them and reloads each one (a scratch round
// trip) in front of its MFMA.
__device__ __forceinline__ int lane_here() {
  int l = (int)(threadIdx.x & 63);
  asm volatile("" : "+v"(l));
  return l;
}

template <int NR>
__device__ __forceinline__ void zero_acc(f32x4 (&acc)[NR]) {
#pragma unroll
  for (int i = 0; i < NR; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
}

__device__ __forceinline__ bf16x8 to_bf16x8(const float4& lo, const float4& hi) {
  bf16x8 v;
  v[0] = (__bf16)lo.x; v[1] = (__bf16)lo.y; v[2] = (__bf16)lo.z; v[3] = (__bf16)lo.w;
  v[4] = (__bf16)hi.x; v[5] = (__bf16)hi.y; v[6] = (__bf16)hi.z; v[7] = (__bf16)hi.w;
  return v;
}

// Raw buffer resource over `bytes` bytes at `base`: loads take a 32-bit VGPR
// byte offset (no per-load 64-bit scalar address) and read 0 past the end.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
constexpr int kOOB = 0x7ffffff0;  // a byte offset past every buffer: loads return 0

// Row-major FP32 global matrix as a B operand with zero padding outside (rows, cols).
struct GlobalB {
  __amdgpu_buffer_rsrc_t r;
  int rows, cols;
  __device__ __forceinline__ GlobalB(const float* w, int rows_, int cols_)
      : r(buffer_rsrc(w, rows_ * cols_ * 4)), rows(rows_), cols(cols_) {}
  __device__ __forceinline__ float operator()(int k, int j) const {
    const int off = (k < rows && j < cols) ? (k * cols + j) * 4 : kOOB;
    // the builtin returns the raw 32 bits as an integer: reinterpret, do not convert
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  }
};

// Transposed view of a row-major FP32 (rows x cols) matrix W as a B operand:
// B(k, j) = W[j][k], zero outside.
struct GlobalBT {
  __amdgpu_buffer_rsrc_t r;
  int rows, cols;  // of W
  __device__ __forceinline__ GlobalBT(const float* w, int rows_, int cols_)
      : r(buffer_rsrc(w, rows_ * cols_ * 4)), rows(rows_), cols(cols_) {}
  __device__ __forceinline__ float operator()(int k, int j) const {
    const int off = (j < rows && k < cols) ? (j * cols + k) * 4 : kOOB;
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  }
};

// BF16 table T[j][k] (stride ld) as a B operand: B(k, j) = T[j][k].
struct PackedB {
  __amdgpu_buffer_rsrc_t r;
  int ld;
  __device__ __forceinline__ PackedB(const __bf16* t, int ld_, int rows)
      : r(buffer_rsrc(t, rows * ld_ * 2)), ld(ld_) {}
  __device__ __forceinline__ bf16x8 load8(int k, int j) const {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (j * ld + k) * 2, 0, 0);
    return __builtin_bit_cast(bf16x8, v);
  }
};

// An FP32 table copied into LDS (stride ld floats, 16-byte aligned rows) as a
// B operand: B(k, j) = T[j][k].
struct LdsB32 {
  const float* t;
  int ld;
  __device__ __forceinline__ float4 load4(int k, int j) const {
    return *reinterpret_cast<const float4*>(t + j * ld + k);
  }
};

// An FP32 table behind a generic pointer (LDS or global memory; flat loads)
// as a B operand: B(k, j) = T[j][k].
struct FlatB32 {
  const float* t;
  int ld;
  __device__ __forceinline__ float4 load4(int k, int j) const {
    return *reinterpret_cast<const float4*>(t + j * ld + k);
  }
};

// A BF16 table copied into LDS (stride ld elements, 16-byte aligned rows) as a
// B operand: B(k, j) = T[j][k].
struct LdsB {
  const __bf16* t;
  int ld;
  __device__ __forceinline__ bf16x8 load8(int k, int j) const {
    return *reinterpret_cast<const bf16x8*>(t + j * ld + k);
  }
};

// FP32 table T[j][k] (stride ld, zero padded) as a B operand: B(k, j) = T[j][k];
// four consecutive k of one column are one 16-byte load.
struct PackedB32 {
  __amdgpu_buffer_rsrc_t r;
  int ld;
  __device__ __forceinline__ PackedB32(const float* t, int ld_, int rows)
      : r(buffer_rsrc(t, rows * ld_ * 4)), ld(ld_) {}
  __device__ __forceinline__ float4 load4(int k, int j) const {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (j * ld + k) * 4, 0, 0);
    return __builtin_bit_cast(float4, v);
  }
};

// FP32 strip with a permuted K order: within each chunk of 16, MFMA c (c < 4)
// sums k = 16i + 4*kq + c over the four lane groups kq, so that a lane's four
// A operands (row r, k = 16i + 4kq .. +3) are one 16-byte LDS read and its
// four B operands one 16-byte buffer load. The sum over k is unchanged; only
// the order of the FP32 accumulation differs from a k-ascending chain.
// Requires K % 16 == 0, lda % 4 == 0 and a 16-byte aligned A.
// The B chunks an FP32 strip requests first (the ring's PD), loaded ahead of
// the strip: before the barrier that precedes its phase, so that their L2
// round trip overlaps the barrier wait instead of following it.
template <int KMAX>
struct PreF32 {
  static constexpr int NS = KMAX / 16, PD = NS < 4 ? NS : 4;
  float4 bb[PD];
};

template <int NR, int KMAX, class B4, bool PRE = false>
__device__ __forceinline__ void mma_strip_f32(const float* __restrict__ A, int lda, int K, B4 b,
                                              int col0, f32x4 (&acc)[NR],
                                              const PreF32<KMAX>* pre = nullptr) {
  const int lane = lane_here();
  const int r = lane & 15, kq = lane >> 4;
  constexpr int NS = KMAX / 16;
  for (int kc = 0; kc < K; kc += KMAX) {  // one pass for K <= KMAX
    // B chunks stream through a ring PD deep: chunk i + PD is requested
    // before chunk i is multiplied, so PD L2 round trips overlap the products.
    // The scheduling barriers pin that order; left to itself the scheduler
    // sinks each load next to its use under register pressure (one round trip
    // per chunk), and hoisting all NS loads costs NS * 4 registers.
    constexpr int PD = NS < 4 ? NS : 4;
    float4 bb[NS];
#pragma unroll
    for (int i = 0; i < PD; ++i)
      if (kc + 16 * i < K) {
        if (PRE && kc == 0)
          bb[i] = pre->bb[i];
        else
          bb[i] = b.load4(kc + 16 * i + 4 * kq, col0 + r);
      }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      if (i + PD < NS && kc + 16 * (i + PD) < K)
        bb[i + PD] = b.load4(kc + 16 * (i + PD) + 4 * kq, col0 + r);
      if (kc + 16 * i < K) {
#pragma unroll
        for (int rt = 0; rt < NR; ++rt) {
          const float4 a = *reinterpret_cast<const float4*>(A + (16 * rt + r) * lda + kc +
                                                            16 * i + 4 * kq);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, bb[i].x, acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, bb[i].y, acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, bb[i].z, acc[rt], 0, 0, 0);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, bb[i].w, acc[rt], 0, 0, 0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, PackedB32 b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_f32<NR, KMAX>(A, lda, K, b, col0, acc);
}

template <int KMAX>
__device__ __forceinline__ void strip_preload(PackedB32 b, int K, int col0, PreF32<KMAX>& p) {
  const int lane = lane_here();
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int i = 0; i < PreF32<KMAX>::PD; ++i)
    if (16 * i < K) p.bb[i] = b.load4(16 * i + 4 * kq, col0 + r);
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, PackedB32 b,
                                          int col0, f32x4 (&acc)[NR], const PreF32<KMAX>& p) {
  mma_strip_f32<NR, KMAX, PackedB32, true>(A, lda, K, b, col0, acc, &p);
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, LdsB32 b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_f32<NR, KMAX>(A, lda, K, b, col0, acc);
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, FlatB32 b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_f32<NR, KMAX>(A, lda, K, b, col0, acc);
}

// acc[rt] += A(rows 16*rt.., K) * B(K, cols col0..col0+15), rt < NR; A from
// LDS (row-major, lda), K <= KMAX. The whole B column strip is loaded into
// registers first (KMAX/4 floats or KMAX/32 bf16x8 per lane) so its L2
// latency overlaps; A is read from LDS per k-step.
template <int NR, int KMAX, class BF32>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, BF32 b,
                                          int col0, f32x4 (&acc)[NR]) {
  const int lane = lane_here();
  const int r = lane & 15, kq = lane >> 4;
  constexpr int NS = KMAX / 4;
  for (int kc = 0; kc < K; kc += KMAX) {  // one pass for K <= KMAX
    float bb[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) bb[i] = (kc + 4 * i < K) ? b(kc + 4 * i + kq, col0 + r) : 0.0f;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      if (kc + 4 * i < K) {
#pragma unroll
        for (int rt = 0; rt < NR; ++rt) {
          const float a = A[(16 * rt + r) * lda + kc + 4 * i + kq];
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb[i], acc[rt], 0, 0, 0);
        }
      }
    }
  }
}

// The B chunks of a BF16 strip (all of them: the strip loads its whole
// column before the first MFMA), loaded ahead of the strip.
template <int KMAX>
struct PreB16 {
  static constexpr int NS = (KMAX + 31) / 32;
  bf16x8 bb[NS];
};

// A lane's 8 A operands of a BF16 strip: FP32 rows rounded per use, or rows
// kept in BF16 (one 16-byte read).
__device__ __forceinline__ bf16x8 strip_a8(const float* ap) {
  return to_bf16x8(*reinterpret_cast<const float4*>(ap), *reinterpret_cast<const float4*>(ap + 4));
}
__device__ __forceinline__ bf16x8 strip_a8(const __bf16* ap) {
  return *reinterpret_cast<const bf16x8*>(ap);
}

template <int NR, int KMAX, class B16, bool PRE = false, class AT = float>
__device__ __forceinline__ void mma_strip_b16(const AT* __restrict__ A, int lda, int K, B16 b,
                                              int col0, f32x4 (&acc)[NR],
                                              const PreB16<KMAX>* pre = nullptr) {
  const int lane = lane_here();
  const int r = lane & 15, kq = lane >> 4;
  constexpr int NS = (KMAX + 31) / 32;
  for (int kc = 0; kc < K; kc += 32 * NS) {  // one pass for K <= KMAX
    bf16x8 bb[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (kc + 32 * i < K) {
        if (PRE && kc == 0)
          bb[i] = pre->bb[i];
        else
          bb[i] = b.load8(kc + 32 * i + 8 * kq, col0 + r);
      }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      if (kc + 32 * i < K) {
#pragma unroll
        for (int rt = 0; rt < NR; ++rt) {
          const bf16x8 av = strip_a8(A + (16 * rt + r) * lda + kc + 32 * i + 8 * kq);
          acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bb[i], acc[rt], 0, 0, 0);
        }
      }
    }
  }
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, PackedB b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_b16<NR, KMAX>(A, lda, K, b, col0, acc);
}

// BF16 A rows (lda in BF16 values, 16-byte aligned rows).
template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const __bf16* __restrict__ A, int lda, int K, PackedB b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_b16<NR, KMAX, PackedB, false, __bf16>(A, lda, K, b, col0, acc);
}

template <int KMAX>
__device__ __forceinline__ void strip_preload(PackedB b, int K, int col0, PreB16<KMAX>& p) {
  const int lane = lane_here();
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int i = 0; i < PreB16<KMAX>::NS; ++i)
    if (32 * i < K) p.bb[i] = b.load8(32 * i + 8 * kq, col0 + r);
}

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, PackedB b,
                                          int col0, f32x4 (&acc)[NR], const PreB16<KMAX>& p) {
  mma_strip_b16<NR, KMAX, PackedB, true>(A, lda, K, b, col0, acc, &p);
}

template <bool BF, int KMAX>
using StripPre = typename std::conditional<BF, PreB16<KMAX>, PreF32<KMAX>>::type;

template <int NR, int KMAX>
__device__ __forceinline__ void mma_strip(const float* __restrict__ A, int lda, int K, LdsB b,
                                          int col0, f32x4 (&acc)[NR]) {
  mma_strip_b16<NR, KMAX>(A, lda, K, b, col0, acc);
}

// Weight-gradient tile: acc(i, j) = sum_k X[k][row0 + i] * Y[k][col0 + j],
// k < K (the rows of the tile). Each lane reads single floats down a column,
// so the four lane groups of one read instruction hit the rows of four
// different k; the k order is permuted so that those rows are 4 apart
// (FP32: k = 16m + 4kq + c for MFMA c of chunk m; BF16: element e of lane
// group kq is k = 32m + 16(e/4) + 4kq + e%4). With row strides of 4 mod 8
// floats, 4 rows are 16 mod 32 floats apart: the four groups' 16-float
// segments fall in disjoint banks. The sum over k is unchanged; only the
// order of the FP32 accumulation differs. Requires K % 16 == 0 (32 for BF16).
// FRESH: the lane index from lane_here (operand addresses computed per call)
// or the plain one (addresses loop-invariant: hoisted out of a tile loop and
// kept in registers when the kernel has them to spare).
template <bool BF, int K, bool FRESH = true>
__device__ __forceinline__ void mma_wgrad_acc(const float* __restrict__ X, int ldx, int row0,
                                              const float* __restrict__ Y, int ldy, int col0,
                                              f32x4& acc) {
  const int lane = FRESH ? lane_here() : lane_id();
  const int r = lane & 15, kq = lane >> 4;
  if constexpr (BF) {
    static_assert(K % 32 == 0, "BF16 weight-gradient tiles take K in steps of 32");
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 32) {
      bf16x8 av, bv;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + 16 * (e >> 2) + 4 * kq + (e & 3);
        av[e] = (__bf16)X[k * ldx + row0 + r];
        bv[e] = (__bf16)Y[k * ldy + col0 + r];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
  } else {
    static_assert(K % 16 == 0, "FP32 weight-gradient tiles take K in steps of 16");
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = k0 + 4 * kq + c;
        const float a = X[k * ldx + row0 + r];
        const float b = Y[k * ldy + col0 + r];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
      }
    }
  }
}

// The same tile product split in two, for a software pipeline over tiles:
// wgrad_load reads a lane's K/4 operands of X and of Y (in the permuted k
// order of mma_wgrad_acc), wgrad_mma multiplies them into acc.
template <bool BF, int K>
__device__ __forceinline__ void wgrad_load(const float* __restrict__ X, int ldx, int row0,
                                           const float* __restrict__ Y, int ldy, int col0,
                                           int lane, float (&a)[K / 4], float (&b)[K / 4]) {
  static_assert(K % (BF ? 32 : 16) == 0, "weight-gradient K in whole MFMA steps");
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int m = 0; m < K / 4; ++m) {
    int k;
    if constexpr (BF) {
      const int e = m & 7;
      k = 32 * (m >> 3) + 16 * (e >> 2) + 4 * kq + (e & 3);
    } else {
      k = 16 * (m >> 2) + 4 * kq + (m & 3);
    }
    a[m] = X[k * ldx + row0 + r];
    b[m] = Y[k * ldy + col0 + r];
  }
}

template <bool BF, int K>
__device__ __forceinline__ void wgrad_mma(const float (&a)[K / 4], const float (&b)[K / 4],
                                          f32x4& acc) {
  if constexpr (BF) {
#pragma unroll
    for (int g = 0; g < K / 32; ++g) {
      bf16x8 av, bv;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        av[e] = (__bf16)a[8 * g + e];
        bv[e] = (__bf16)b[8 * g + e];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int m = 0; m < K / 4; ++m)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[m], b[m], acc, 0, 0, 0);
  }
}

template <bool BF, int K>
__device__ __forceinline__ f32x4 mma_wgrad(const float* __restrict__ X, int ldx, int row0,
                                           const float* __restrict__ Y, int ldy, int col0) {
  f32x4 acc{0.f, 0.f, 0.f, 0.f};
  mma_wgrad_acc<BF, K>(X, ldx, row0, Y, ldy, col0, acc);
  return acc;
}

}  // namespace tfh
