// FeatureMlp kernels on gfx950 MFMA, shared by five translation units:
// tfh_mlp.hip (forward, backward of the generic shapes), tfh_mlp_bwd_std.hip
// (backward of the standard shape), tfh_mlp_film.hip (the same with a FiLM
// camera embedding), tfh_mlp_bwd_act.hip and tfh_mlp_film_act.hip (the
// backward reading the activations the forward kept, tfh_mlp_bwd_kept.h,
// without and with the embedding). The standard backward sits at the 256-VGPR limit of two waves
// per SIMD, and the compiler allocates its registers worse when other
// instantiations share its translation unit (once 17/31 spilled VGPRs
// (BF16/FP32) on its own against 26/43 next to the generic kernels and 39/67
// next to the FiLM one).
//
// FeatureMlp (networks.py:38-121).
//
// Layers (kernels are (in, out) like Flax Dense):
//   h0 = f @ W0                       (F -> S, no bias)      networks.py:57-61
//   x  = [h0, d, FE(h0, nf), FE(d, nv)]  (E wide)            networks.py:68-76
//   h1 = relu(x @ W1 + b1)            (E -> U)               networks.py:86-90
//   h2 = relu(h1 @ W2 + b2)           (U -> U)               networks.py:94-98
//   h2[U/2:] = e[:U/2] * h2[U/2:] + e[U/2:],  e = embed[cam]  networks.py:103-111
//   rgb = sigmoid(h2 @ W3 + b3)       (U -> 3)               networks.py:114-120
//
// A workgroup owns a tile of TM rows. Activations live in LDS (FP32,
// row-major, padded strides); every layer is a set of 16x16 output tiles:
//   FP32: v_mfma_f32_16x16x4_f32 with a permuted K order so that a lane's
//         four consecutive A operands are one 16-byte LDS read and its four B
//         operands one 16-byte load from an FP32 weight table;
//   BF16: v_mfma_f32_16x16x32_bf16 with FP32 accumulation; activations are
//         rounded to BF16 when loaded as operands.
// Weights are packed once per call into zero-padded [out][in] (forward) and
// [in][out] (input-gradient) tables of the operand precision, L2-resident.
// The backward recomputes the forward of its tile (only layer 0 and the
// encoding when the forward kernel kept the hidden activations, the fused
// training steps' case: k_mlp_bwd_kept) and accumulates the weight
// gradients in registers of their owning wave (standard shape) or in a
// per-workgroup partial buffer (other shapes, plain read-modify-write, each
// partial tile owned by one wave), reduced across workgroups by k_mlp_reduce.
#pragma once

#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "tfh_mfma.h"

namespace tfh {

constexpr int kU = 128;   // hidden width
constexpr int kSP = 32;   // padded squash width

struct MlpDims {
  int F, FP;   // feature dim, padded
  int S;       // squash dim (<= 32)
  int E, EP;   // encoded dim, padded
  int nf, nv;
  int ncam;
  int Z3W;     // padded width of the layer-3 output / its gradient
  // LDS strides (floats)
  int ldF, ldE, ldU, ldS, ld3;
};

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Kept activations (the forward's relu(z1) and layer-3 input, 256 values a
// row, for the backward): FP32 rows, or BF16 rows with the BF16 MLP, whose
// backward only reads them as BF16 MFMA operands and as ReLU signs (the
// rounding keeps signs and zeros), so BF16 storage changes no result and
// halves the traffic.
template <bool BF>
constexpr int kKeptRowBytes = 256 * (BF ? 2 : 4);

__device__ __forceinline__ void bf16x8_to_f32(const int4& v, float4& lo, float4& hi) {
  // a BF16 value is the upper half of its FP32 bit pattern
  const unsigned w[4] = {(unsigned)v.x, (unsigned)v.y, (unsigned)v.z, (unsigned)v.w};
  lo = make_float4(__uint_as_float(w[0] << 16), __uint_as_float(w[0] & 0xffff0000u),
                   __uint_as_float(w[1] << 16), __uint_as_float(w[1] & 0xffff0000u));
  hi = make_float4(__uint_as_float(w[2] << 16), __uint_as_float(w[2] & 0xffff0000u),
                   __uint_as_float(w[3] << 16), __uint_as_float(w[3] & 0xffff0000u));
}

// Stores four kept values as FP32 or BF16 (round to nearest even).
template <bool BF>
__device__ __forceinline__ void store_kept4(void* row, int col, const float4& v) {
  if constexpr (BF) {
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    bf16x4 b;
    b[0] = (__bf16)v.x; b[1] = (__bf16)v.y; b[2] = (__bf16)v.z; b[3] = (__bf16)v.w;
    *reinterpret_cast<bf16x4*>(static_cast<__bf16*>(row) + col) = b;
  } else {
    *reinterpret_cast<float4*>(static_cast<float*>(row) + col) = v;
  }
}

// FP32 pads K to 16, BF16 to 32; row strides are 4 mod 8 floats: 16-byte
// aligned rows whose 128-bit reads by 8 consecutive rows hit 8 distinct bank
// groups.
inline MlpDims make_dims(const tfh_mlp* p, bool bf) {
  MlpDims d;
  const int kp = bf ? 32 : 16;
  const int pad = 4;
  d.F = (int)p->feat_dim;
  d.FP = round_up(d.F, kp);
  d.S = (int)p->squash_dim;
  d.nf = (int)p->feature_n_freqs;
  d.nv = (int)p->viewdir_n_freqs;
  d.E = d.S + 3 + 2 * d.nf * d.S + 2 * d.nv * 3;
  d.EP = round_up(d.E, kp);
  d.ncam = p->embed ? (int)p->num_cameras : 0;
  d.Z3W = bf ? 32 : 16;
  d.ldF = d.FP + pad;
  d.ldE = d.EP + pad;
  d.ldU = kU + pad;
  d.ldS = kSP + pad;
  d.ld3 = d.Z3W + pad;
  return d;
}

// Dims of the standard appearance MLP (48 channels x 3 modes, squash 27,
// 2+2 Fourier frequencies) as compile-time constants, so the kernels built for
// it fold every LDS offset and tile count; the camera count stays a run-time
// value (0 = no embedding).
template <bool BF>
__host__ __device__ constexpr MlpDims std_dims(int ncam = 0) {
  const int kp = BF ? 32 : 16, pad = 4;
  const int F = 144, S = 27, nf = 2, nv = 2;
  const int E = S + 3 + 2 * nf * S + 2 * nv * 3;
  const int FP = (F + kp - 1) / kp * kp, EP = (E + kp - 1) / kp * kp;
  const int Z3W = BF ? 32 : 16;
  return MlpDims{F, FP, S, E, EP, nf, nv, ncam, Z3W, FP + pad, EP + pad, 128 + pad, 32 + pad,
                 Z3W + pad};
}

inline bool is_std_dims(const MlpDims& d, bool bf) {
  const MlpDims s = bf ? std_dims<true>() : std_dims<false>();
  return d.F == s.F && d.S == s.S && d.nf == s.nf && d.nv == s.nv;
}

struct MlpW {  // FP32 parameters (device pointers)
  const float *w0, *w1, *b1, *w2, *b2, *w3, *b3, *embed;
};

// Operand tables built per call in the workspace: forward tables [out][in]
// and input-gradient tables [in][out], zero padded:
//   f0 [32][FP], f1 [U][EP], f2 [U][U], f3 [16][U]
//   g0 [FP][32], g1 [EP][U], g2 [U][U], g3 [U][32]
struct MlpPack {
  const __bf16 *f0, *f1, *f2, *f3, *g0, *g1, *g2, *g3;      // BF16 mode
  const float *ff0, *ff1, *ff2, *ff3, *fg0, *fg1, *fg2, *fg3;  // FP32 mode
};

// B operands of each GEMM, by precision.
template <bool BF>
struct Operands;

template <>
struct Operands<false> {
  const MlpDims& d;
  const MlpW& w;
  const MlpPack& p;
  __device__ PackedB32 l0() const { return PackedB32(p.ff0, d.FP, 32); }
  __device__ PackedB32 l1() const { return PackedB32(p.ff1, d.EP, kU); }
  __device__ PackedB32 l2() const { return PackedB32(p.ff2, kU, kU); }
  __device__ PackedB32 l3() const { return PackedB32(p.ff3, kU, 16); }
  __device__ PackedB32 g3() const { return PackedB32(p.fg3, 32, kU); }
  __device__ PackedB32 g2() const { return PackedB32(p.fg2, kU, kU); }
  __device__ PackedB32 g1() const { return PackedB32(p.fg1, kU, d.EP); }
  __device__ PackedB32 g0() const { return PackedB32(p.fg0, 32, d.FP); }
  __device__ int k0() const { return d.FP; }
  __device__ int k3() const { return 16; }  // K of dZ3 @ W3^T (3 padded to 16)
};

template <>
struct Operands<true> {
  const MlpDims& d;
  const MlpW& w;
  const MlpPack& p;
  __device__ PackedB l0() const { return PackedB(p.f0, d.FP, 32); }
  __device__ PackedB l1() const { return PackedB(p.f1, d.EP, kU); }
  __device__ PackedB l2() const { return PackedB(p.f2, kU, kU); }
  __device__ PackedB l3() const { return PackedB(p.f3, kU, 16); }
  __device__ PackedB g3() const { return PackedB(p.g3, 32, kU); }
  __device__ PackedB g2() const { return PackedB(p.g2, kU, kU); }
  __device__ PackedB g1() const { return PackedB(p.g1, kU, d.EP); }
  __device__ PackedB g0() const { return PackedB(p.g0, 32, d.FP); }
  __device__ int k0() const { return d.FP; }
  __device__ int k3() const { return 32; }
};

// ---- shared forward pieces ------------------------------------------------------
// Loads TM rows of features (zero padded to FP columns and past M), view
// directions and camera indices of the tile. Row `row` belongs to ray
// row / rpr (view direction and camera index are per ray).
template <int TM>
__device__ void load_tile(const MlpDims& d, const float* __restrict__ feats,
                          const float* __restrict__ dirs, const uint32_t* __restrict__ cams,
                          int64_t row0, int64_t M, int64_t rpr, float* Fs, float* Ds,
                          int* Cs) {
  constexpr int kMaxPer = 3;  // float4 loads per thread of the batched path
  if (d.F % 4 == 0 && (reinterpret_cast<uintptr_t>(feats) & 15) == 0 &&
      TM * d.FP / 4 <= kMaxPer * (int)blockDim.x) {
    // 16-byte loads, all issued before the first LDS store (one memory
    // latency per tile instead of one per loop trip)
    const int n4 = TM * d.FP / 4, fp4 = d.FP / 4, f4 = d.F / 4;
    float4 v[kMaxPer];
#pragma unroll
    for (int k = 0; k < kMaxPer; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      const int i = e / fp4, c4 = e - i * fp4;
      const int64_t row = row0 + i;
      v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (e < n4 && row < M && c4 < f4)
        v[k] = *reinterpret_cast<const float4*>(feats + row * d.F + 4 * c4);
    }
#pragma unroll
    for (int k = 0; k < kMaxPer; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      const int i = e / fp4, c4 = e - i * fp4;
      if (e < n4) *reinterpret_cast<float4*>(Fs + i * d.ldF + 4 * c4) = v[k];
    }
  } else {
    for (int e = threadIdx.x; e < TM * d.FP; e += blockDim.x) {
      const int i = e / d.FP, c = e % d.FP;
      const int64_t row = row0 + i;
      Fs[i * d.ldF + c] = (row < M && c < d.F) ? feats[row * d.F + c] : 0.0f;
    }
  }
  for (int e = threadIdx.x; e < TM * 3; e += blockDim.x) {
    const int i = e / 3, c = e % 3;
    const int64_t row = row0 + i;
    Ds[i * 3 + c] = row < M ? dirs[(row / rpr) * 3 + c] : 0.0f;
  }
  for (int i = threadIdx.x; i < TM; i += blockDim.x) {
    const int64_t row = row0 + i;
    int cam = (cams && row < M) ? (int)cams[row / rpr] : 0;
    if (cam < 0 || cam >= d.ncam) cam = 0;
    Cs[i] = cam;
  }
}

// Row stride (BF16 values) of the BF16 standard forward's feature rows in
// LDS: 84 dwords, 4 mod 8, so a 16-byte read per lane is conflict free.
constexpr int kFsLdB = 168;

// A tile in one pass: feature rows (Fs, zero padded to FP) and view
// directions (Ds). Each source is a buffer resource over the tile's valid
// rows, so an offset is 32 bits and rows past M read 0 with no branch; every
// load is issued before the first LDS store and the tile pays one memory
// latency. With CAMS also each row's camera index (Cs, clamped to [0, ncam),
// from the per-ray cams). FB16: the feature rows go to LDS in BF16 (Fs is
// then __bf16 with stride kFsLdB), rounded as the BF16 strips round their A
// operands. Requires F % 4 == 0 and 0 < rpr < 2^31.
template <int TM, int NT, bool CAMS = false, bool FB16 = false>
__device__ void load_tile_rows(const MlpDims& d, const float* __restrict__ feats,
                               const float* __restrict__ dirs, int64_t row0, int64_t M,
                               int64_t rpr, void* Fs, float* Ds,
                               const uint32_t* __restrict__ cams = nullptr, int* Cs = nullptr) {
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  static_assert(TM * 3 <= NT, "tile shape");
  const int nv = (int)min((int64_t)TM, M - row0);
  const int64_t ray0 = row0 / rpr;
  const unsigned rem0 = (unsigned)(row0 - ray0 * rpr);
  const int nray = (int)((row0 + nv - 1) / rpr - ray0) + 1;
  // the feature base goes through an empty asm: a load from a known
  // read-only argument counts as invariant, and the register allocator then
  // re-issues it at its use (after the other loads' waits) instead of keeping it
  const float* fbase = feats + row0 * d.F;
  asm volatile("" : "+s"(fbase));
  const auto rf = buffer_rsrc(fbase, nv * d.F * 4);
  const auto rd = buffer_rsrc(dirs + ray0 * 3, nray * 12);
  const auto rc = buffer_rsrc(CAMS ? static_cast<const void*>(cams + ray0) : feats,
                              CAMS ? nray * 4 : 0);
  const int fp4 = d.FP / 4, f4 = d.F / 4, n4 = TM * fp4;
  const int kF4 = (n4 + NT - 1) / NT;  // constant for the specialised shape
  const int t = threadIdx.x;
  // issue
  i32x4 vf[(TM * 160 / 4 + NT - 1) / NT];  // room for FP <= 160
#pragma unroll
  for (int k = 0; k < (TM * 160 / 4 + NT - 1) / NT; ++k) {
    if (k < kF4) {
      const int e = t + k * NT, i = e / fp4, c4 = e - i * fp4;
      const int off = (e < n4 && c4 < f4) ? (i * d.F + 4 * c4) * 4 : kOOB;
      vf[k] = __builtin_amdgcn_raw_buffer_load_b128(rf, off, 0, 0);
    }
  }
  const int i3 = t / 3, c3 = t - 3 * i3;
  const bool own3 = t < TM * 3;
  const unsigned rel = (rem0 + (unsigned)i3) / (unsigned)rpr;  // ray of row i3 after ray0
  const float dv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
      rd, (own3 && i3 < nv) ? (int)(rel * 3 + c3) * 4 : kOOB, 0, 0));
  int cam = 0;
  if constexpr (CAMS) {
    const unsigned relc = (rem0 + (unsigned)t) / (unsigned)rpr;  // ray of row t after ray0
    cam = __builtin_amdgcn_raw_buffer_load_b32(rc, (t < nv) ? (int)relc * 4 : kOOB, 0, 0);
  }
  // keeps the loads here: a load whose only use is a predicated store would
  // otherwise be sunk into that branch, after the waits of the others
  asm volatile("" ::: "memory");
  // store
#pragma unroll
  for (int k = 0; k < (TM * 160 / 4 + NT - 1) / NT; ++k) {
    const int e = t + k * NT, i = e / fp4, c4 = e - i * fp4;
    if (k < kF4 && e < n4) {
      const float4 v = __builtin_bit_cast(float4, vf[k]);
      if constexpr (FB16) {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        const bf16x4 b = {(__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
        *reinterpret_cast<bf16x4*>(static_cast<__bf16*>(Fs) + i * kFsLdB + 4 * c4) = b;
      } else {
        *reinterpret_cast<float4*>(static_cast<float*>(Fs) + i * d.ldF + 4 * c4) = v;
      }
    }
  }
  if (own3) Ds[i3 * 3 + c3] = dv;
  if constexpr (CAMS) {
    if (t < TM) Cs[t] = (cam < 0 || cam >= d.ncam) ? 0 : cam;
  }
}

// Fills X[:, S:EP] = [d, FE(h0, nf), FE(d, nv), 0...] from h0 = X[:, 0:S]
// (networks.py:13-35, 68-76): FE(v, n)[j*2n + q] = sin(2^q v_j) and
// [j*2n + n + q] = cos(2^q v_j).
// The forward (DOUBLE = false) takes one sincosf per (row, value, frequency).
// The backward's recompute (DOUBLE = true) takes one per (row, value): q = 0
// by sincosf, q = 1 from it by sin 2a = 2 sin a cos a and
// cos 2a = (cos a - sin a)(cos a + sin a), q >= 2 by sincosf again; its
// values feed only the weight gradient of layer 1 and the encoding's own
// backward, where the identities' ~2 ulp stay below the gradient tolerance
// (taken in the forward, their bias shifted the outputs a whole-batch
// gradient sums over a million rows). With XB (stride ldB), h0 is first
// formed as X[:, 0:S] + XB[:, 0:S] and written back to X (a layer 0
// computed in two K parts).
template <int TM, bool DOUBLE = false>
__device__ void encode_tile(const MlpDims& d, float* X, const float* Ds,
                            const float* XB = nullptr, int ldB = 0) {
  const int fe0 = d.S + 3;
  const int fe1 = fe0 + 2 * d.nf * d.S;
  const int fe2 = fe1 + 2 * d.nv * 3;
  // the view direction and the zero padding past the Fourier features
  const int nplain = 3 + (d.EP - fe2);
  for (int e = threadIdx.x; e < TM * nplain; e += blockDim.x) {
    const int i = e / nplain, t = e % nplain;
    const int c = t < 3 ? d.S + t : fe2 + (t - 3);
    X[i * d.ldE + c] = t < 3 ? Ds[i * 3 + t] : 0.0f;
  }
  if constexpr (!DOUBLE) {
    const int nfe = d.S * d.nf + 3 * d.nv;  // (value, frequency) pairs per row
    for (int e = threadIdx.x; e < TM * nfe; e += blockDim.x) {
      const int i = e / nfe, t = e % nfe;
      const bool feat = t < d.S * d.nf;
      const int u = feat ? t : t - d.S * d.nf;
      const int n = feat ? d.nf : d.nv;
      const int j = u / n, q = u % n;
      const float src = feat ? X[i * d.ldE + j] : Ds[i * 3 + j];
      float sv, cv;
      sincosf(src * (float)(1 << q), &sv, &cv);
      const int c = (feat ? fe0 : fe1) + j * 2 * n + q;
      X[i * d.ldE + c] = sv;
      X[i * d.ldE + c + n] = cv;
    }
    return;
  }
  const int nval = d.S + 3;  // values per row: h0, then the direction
  for (int e = threadIdx.x; e < TM * nval; e += blockDim.x) {
    const int i = e / nval, u = e % nval;
    const bool feat = u < d.S;
    const int j = feat ? u : u - d.S;
    const int n = feat ? d.nf : d.nv;
    float v;
    if (feat) {
      v = X[i * d.ldE + j];
      if (XB) {
        v += XB[i * ldB + j];
        X[i * d.ldE + j] = v;
      }
    } else {
      v = Ds[i * 3 + j];
    }
    if (n == 0) continue;
    float* fe = X + i * d.ldE + (feat ? fe0 : fe1) + j * 2 * n;
    float sv, cv;
    sincosf(v, &sv, &cv);
    fe[0] = sv;
    fe[n] = cv;
    if (n > 1) {
      fe[1] = 2.0f * sv * cv;
      fe[n + 1] = (cv - sv) * (cv + sv);
    }
    for (int q = 2; q < n; ++q) {
      float s2, c2;
      sincosf(v * (float)(1 << q), &s2, &c2);
      fe[q] = s2;
      fe[n + q] = c2;
    }
  }
}

// Runs layers 0..2 of the tile. On exit: X holds the encoded input, H0 (if
// non-null) a copy of h0, H1 = relu(z1), H2p = relu(z2) (pre-FiLM) and H2 the
// layer-3 input (== H2p without embedding). Fs is read by layer 0 only.
// Work items are (column tile, group of RTI row tiles): RTI = TM/16 keeps a
// column strip's B operands for all rows of the tile (few waves), RTI = 1
// spreads a layer over more waves.
// L12 == false stops after the encoding (H1 and H2 come from elsewhere).
// FT: the feature rows' type (float, stride ldF; or __bf16 with the BF16
// MLP, stride kFsLdB).
template <int TM, int NW, bool BF, int RTI, bool L12 = true, class FT = float>
// h2g (optional, rows < nvalid): relu(z2) also stored to global memory, rows
// of 256 kept values (FP32, or BF16 with the BF16 MLP: the forward's kept
// activations with an embedding).
__device__ void forward_hidden(const MlpDims& d, const MlpW& w, const Operands<BF>& op,
                               const FT* Fs, float* X, float* H0, float* H1, float* H2p,
                               float* H2, const float* Ds, const int* Cs,
                               void* __restrict__ h2g = nullptr, int nvalid = 0) {
  static_assert(std::is_same<FT, float>::value || BF, "BF16 feature rows need the BF16 MLP");
  const int ldFs = std::is_same<FT, float>::value ? d.ldF : kFsLdB;
  constexpr int RG = TM / 16 / RTI;  // row groups
  // B prefetch depth (K per pass): the whole strip with few waves, 64 when the
  // workgroup's waves share the register file (16 waves)
  constexpr int K0 = RTI == 1 ? 48 : 192, K1 = RTI == 1 ? 64 : 160, K2 = RTI == 1 ? 64 : kU;
  const int wave = threadIdx.x >> 6;
  const int lane = lane_id();
  // layer 0: (TM x F) @ (F x S) -> X[:, 0:S]: 2 column tiles of 16, one row
  // tile per item so that 2 * TM/16 waves share it
  for (int it = wave; it < 2 * (TM / 16); it += NW) {
    const int ct = it / (TM / 16), r0 = 16 * (it % (TM / 16));
    f32x4 acc[1];
    zero_acc(acc);
    mma_strip<1, K0>(Fs + r0 * ldFs, ldFs, op.k0(), op.l0(), 16 * ct, acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = r0 + 4 * (lane >> 4) + e, col = 16 * ct + (lane & 15);
      if (col < d.S) {
        X[row * d.ldE + col] = acc[0][e];
        if (H0) H0[row * d.ldS + col] = acc[0][e];
      }
    }
  }
  __syncthreads();
  encode_tile<TM>(d, X, Ds);
  __syncthreads();
  if constexpr (!L12) return;
  // layer 1: (TM x EP) @ (EP x U) + b1, relu -> H1
  for (int it = wave; it < (kU / 16) * RG; it += NW) {
    const int ct = it / RG, r0 = 16 * RTI * (it % RG);
    f32x4 acc[RTI];
    zero_acc(acc);
    mma_strip<RTI, K1>(X + r0 * d.ldE, d.ldE, d.EP, op.l1(), 16 * ct, acc);
    const int col = 16 * ct + (lane & 15);
    const float bias = w.b1[col];
#pragma unroll
    for (int rt = 0; rt < RTI; ++rt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 16 * rt + 4 * (lane >> 4) + e;
        H1[row * d.ldU + col] = fmaxf(acc[rt][e] + bias, 0.0f);
      }
  }
  __syncthreads();
  // layer 2: (TM x U) @ (U x U) + b2, relu (+ FiLM) -> H2p / H2
  for (int it = wave; it < (kU / 16) * RG; it += NW) {
    const int ct = it / RG, r0 = 16 * RTI * (it % RG);
    f32x4 acc[RTI];
    zero_acc(acc);
    mma_strip<RTI, K2>(H1 + r0 * d.ldU, d.ldU, kU, op.l2(), 16 * ct, acc);
    const int col = 16 * ct + (lane & 15);
    const float bias = w.b2[col];
    // when the tile's rows share a camera (a tile lies in one ray whenever
    // rays have at least TM rows) its scale and shift of this column are
    // read once
    static_assert(TM <= 64, "the camera check reads one tile row per lane");
    int tc = -1;
    float esc = 0.0f, esh = 0.0f;
    if (d.ncam > 0) {
      const int c0 = Cs[0];
      tc = __all(lane >= TM || Cs[lane] == c0) ? c0 : -1;
      if (tc >= 0 && col >= kU / 2) {
        esc = w.embed[(size_t)tc * kU + col - kU / 2];
        esh = w.embed[(size_t)tc * kU + col];
      }
    }
#pragma unroll
    for (int rt = 0; rt < RTI; ++rt)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 16 * rt + 4 * (lane >> 4) + e;
        const float h = fmaxf(acc[rt][e] + bias, 0.0f);
        float out = h;
        if (d.ncam > 0 && col >= kU / 2) {
          if (tc >= 0) {
            out = esc * h + esh;
          } else {
            const float* emb = w.embed + (size_t)Cs[row] * kU;
            out = emb[col - kU / 2] * h + emb[col];
          }
        }
        if (H2p != H2) H2p[row * d.ldU + col] = h;
        if (h2g && row < nvalid) {
          if constexpr (BF)
            static_cast<__bf16*>(h2g)[row * 256 + col] = (__bf16)h;
          else
            static_cast<float*>(h2g)[row * 256 + col] = h;
        }
        H2[row * d.ldU + col] = out;
      }
  }
  __syncthreads();
}

// layer 3 + sigmoid: R3[row][c] for c < 3 (other columns untouched).
// BF16: one 16x16x32 MFMA strip per row tile (4 products). FP32: the strip
// would be 32 dependent 16x16x4 products on two waves while the rest wait at
// the barrier, so every wave takes part on the vector ALUs instead: 16 lanes
// per row, lane q sums k = 8q .. 8q+7 for the three colours, then a 16-lane
// butterfly (mlp_fwd 1.98 -> 1.92 ms at 2.1 M rows; the same on the BF16
// MLP was 2% slower than its 4 MFMAs).
template <int TM, int NW, bool BF>
__device__ void forward_out(const MlpDims& d, const MlpW& w, const Operands<BF>& op,
                            const float* H2, float* R3) {
  if constexpr (BF) {
    constexpr int RT = TM / 16;
    const int wave = threadIdx.x >> 6;
    const int lane = lane_id();
    for (int rt = wave; rt < RT; rt += NW) {
      f32x4 acc[1];
      zero_acc(acc);
      mma_strip<1, kU>(H2 + 16 * rt * d.ldU, d.ldU, kU, op.l3(), 0, acc);
      const int col = lane & 15;
      if (col < 3) {
        const float bias = w.b3[col];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = 16 * rt + 4 * (lane >> 4) + e;
          const float z = acc[0][e] + bias;
          R3[row * d.ld3 + col] = 1.0f / (1.0f + expf(-z));
        }
      }
    }
  } else {
    static_assert(kU == 16 * 8, "16 lanes x 8 inputs per row");
    // TM * 16 is a multiple of 64: a wave's lanes all take a trip or none
    for (int t = threadIdx.x; t < TM * 16; t += NW * 64) {
      const int i = t >> 4, q = t & 15;
      const float4 h0 = *reinterpret_cast<const float4*>(H2 + i * d.ldU + 8 * q);
      const float4 h1 = *reinterpret_cast<const float4*>(H2 + i * d.ldU + 8 * q + 4);
      const float h[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
      float a[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float4 w0 = *reinterpret_cast<const float4*>(op.p.ff3 + c * kU + 8 * q);
        const float4 w1 = *reinterpret_cast<const float4*>(op.p.ff3 + c * kU + 8 * q + 4);
        const float wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) s += h[j] * wv[j];
        a[c] = s;
      }
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] += __shfl_xor(a[c], o, 64);
      if (q < 3) {
        const float z = (q == 0 ? a[0] : (q == 1 ? a[1] : a[2])) + w.b3[q];
        R3[i * d.ld3 + q] = 1.0f / (1.0f + expf(-z));
      }
    }
  }
  __syncthreads();
}

// ---- forward kernel -----------------------------------------------------------
template <int TM>
inline size_t fwd_lds_floats(const MlpDims& d, bool fs_b16 = false) {
  // the layer-3 output reuses region A (H1 is dead by then); BF16 feature
  // rows take half of it
  return (size_t)TM * (std::max(fs_b16 ? kFsLdB / 2 : d.ldF, d.ldU) + std::max(d.ldE, d.ldU) +
                       4 + 4);
}

// act (optional): each row's relu(z1) and layer-3 input (relu(z2) before
// the camera embedding when there is one), 256 values (FP32, or BF16 with
// the BF16 MLP), kept for the backward (k_mlp_bwd_kept).
// No tile source: the feature rows are read from feats.
struct NoGather {};

// GT (standard shape only): a tile source that forms the feature rows
// itself (gt.load<TM, NT, CAMS, FB16>(d, row0, M, rpr, Fs, Ds, Cs, dirs,
// cams): the appearance gather fused into the tile load, its rows also
// written to global memory for the backward), instead of reading feats.
template <int TM, int NW, bool BF, bool STD, class GT = NoGather>
__global__ void __launch_bounds__(NW * 64) k_mlp_fwd(MlpDims d_arg, MlpW w, MlpPack pk,
                                                     const float* __restrict__ feats,
                                                     const float* __restrict__ dirs,
                                                     const uint32_t* __restrict__ cams,
                                                     int64_t M, int64_t rpr,
                                                     float* __restrict__ rgb,
                                                     float* __restrict__ act, GT gt) {
  constexpr bool kGather = !std::is_same<GT, NoGather>::value;
  static_assert(!kGather || STD, "the fused gather serves the standard shape");
  const MlpDims d = STD ? std_dims<BF>(d_arg.ncam) : d_arg;
  extern __shared__ float smem[];
  // regions: A = Fs, later H1, last R3 (the layer-3 output); B = X, later
  // H2; small: Ds, Cs. 40 KB per workgroup for the FP32 standard shape: four
  // workgroups share a CU.
  // the BF16 standard shape keeps its feature rows in BF16 (layer 0 reads
  // them as BF16 operands either way): 39 KB of LDS, four workgroups per CU
  // instead of three
  constexpr bool FB16 = BF && STD;
  float* regA = smem;
  float* regB = regA + TM * std::max(FB16 ? kFsLdB / 2 : d.ldF, d.ldU);
  float* Ds = regB + TM * std::max(d.ldE, d.ldU);
  int* Cs = reinterpret_cast<int*>(Ds + TM * 4);
  float* R3 = regA;
  const Operands<BF> op{d, w, pk};
  int64_t tile = blockIdx.x;
  if constexpr (kGather) {
    // XCD-ordered tiles over xcd_grid(tiles) blocks: each XCD walks one
    // contiguous range of tiles
    tile = xcd_unit((M + TM - 1) / TM);
    if (tile * TM >= M) return;
  }
  const int64_t row0 = tile * TM;
  if constexpr (kGather) {
    if (d.ncam == 0)
      gt.template load<TM, NW * 64, false, FB16>(d, row0, M, rpr, regA, Ds, Cs, dirs, cams);
    else
      gt.template load<TM, NW * 64, true, FB16>(d, row0, M, rpr, regA, Ds, Cs, dirs, cams);
  } else if (STD && d.ncam == 0) {
    load_tile_rows<TM, NW * 64, false, FB16>(d, feats, dirs, row0, M, rpr, regA, Ds);
  } else if (STD) {
    load_tile_rows<TM, NW * 64, true, FB16>(d, feats, dirs, row0, M, rpr, regA, Ds, cams, Cs);
  } else {
    load_tile<TM>(d, feats, dirs, cams, row0, M, rpr, regA, Ds, Cs);
  }
  __syncthreads();
  // layer 1 reads X (regB) and writes H1 over Fs (regA, dead after layer 0);
  // layer 2 reads H1 and writes H2 over X (dead after layer 1)
  // kept activations per row: relu(z1), then relu(z2) before the embedding
  // (the layer-3 input without one); with an embedding layer 2 stores the
  // second half itself
  const bool film_keep = act && d.ncam > 0;
  char* kept = reinterpret_cast<char*>(act);
  constexpr int kRow = kKeptRowBytes<BF>;
  using FT = typename std::conditional<FB16, __bf16, float>::type;
  forward_hidden<TM, NW, BF, TM / 16, true, FT>(d, w, op, reinterpret_cast<const FT*>(regA), regB,
                                                nullptr, regA, regB, regB, Ds, Cs,
                                      film_keep ? kept + row0 * kRow + kRow / 2 : nullptr,
                                      (int)min((int64_t)TM, M - row0));
  if (act) {
    for (int e = threadIdx.x; e < TM * 64; e += blockDim.x) {
      const int i = e >> 6, c4 = e & 63;
      if (row0 + i < M && (c4 < 32 || !film_keep)) {
        const float* src = c4 < 32 ? regA + i * d.ldU + 4 * c4 : regB + i * d.ldU + 4 * (c4 - 32);
        store_kept4<BF>(kept + (row0 + i) * kRow, 4 * c4, *reinterpret_cast<const float4*>(src));
      }
    }
  }
  __syncthreads();  // H1 (region A) is read above before R3 overwrites it
  forward_out<TM, NW, BF>(d, w, op, regB, R3);
  for (int e = threadIdx.x; e < TM * 3; e += blockDim.x) {
    const int i = e / 3, c = e % 3;
    const int64_t row = row0 + i;
    if (row < M) rgb[row * 3 + c] = R3[i * d.ld3 + c];
  }
}

// ---- backward ---------------------------------------------------------------------
// Packed layout of one workgroup's gradient partial (and of the reduction).
struct GradPack {
  int64_t w0, w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ constexpr GradPack make_pack(const MlpDims& d) {
  GradPack p{};
  int64_t o = 0;
  p.w0 = o; o += (int64_t)d.F * d.S;
  p.w1 = o; o += (int64_t)d.E * kU;
  p.b1 = o; o += kU;
  p.w2 = o; o += (int64_t)kU * kU;
  p.b2 = o; o += kU;
  p.w3 = o; o += (int64_t)kU * 3;
  p.b3 = o; o += 3;
  p.total = (o + 63) / 64 * 64;
  return p;
}

// Operand tables written by k_pack_weights (element type T: __bf16 or
// float). One thread per table entry; entries outside the real matrix are zero.
template <typename T>
struct PackOut {
  T *f0, *f1, *f2, *f3, *g0, *g1, *g2, *g3;
};

struct PackSizes {
  int64_t f0, f1, f2, f3;      // forward tables
  int64_t g0, g1, g2, g3;      // input-gradient tables
};

inline PackSizes pack_sizes(const MlpDims& d) {
  PackSizes s;
  s.f0 = (int64_t)32 * d.FP; s.f1 = (int64_t)kU * d.EP; s.f2 = kU * kU; s.f3 = 16 * kU;
  s.g0 = (int64_t)d.FP * 32; s.g1 = (int64_t)d.EP * kU; s.g2 = kU * kU; s.g3 = kU * 32;
  return s;
}

__device__ __forceinline__ float wat(const float* w, int rows, int cols, int r, int c) {
  return (r < rows && c < cols) ? w[r * cols + c] : 0.0f;
}

template <typename T>
__global__ void k_pack_weights(MlpDims d, MlpW w, PackOut<T> o, PackSizes s) {
  const int64_t total = s.f0 + s.f1 + s.f2 + s.f3 + s.g0 + s.g1 + s.g2 + s.g3;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = e;
    // forward tables T[j][k] = W[k][j]
    if (f < s.f0) { const int j = f / d.FP, k = f % d.FP; o.f0[f] = (T)wat(w.w0, d.F, d.S, k, j); continue; }
    f -= s.f0;
    if (f < s.f1) { const int j = f / d.EP, k = f % d.EP; o.f1[f] = (T)wat(w.w1, d.E, kU, k, j); continue; }
    f -= s.f1;
    if (f < s.f2) { const int j = f / kU, k = f % kU; o.f2[f] = (T)wat(w.w2, kU, kU, k, j); continue; }
    f -= s.f2;
    if (f < s.f3) { const int j = f / kU, k = f % kU; o.f3[f] = (T)wat(w.w3, kU, 3, k, j); continue; }
    f -= s.f3;
    // input-gradient tables T[j][k] = W[j][k] (j = input unit, k = output unit)
    if (f < s.g0) { const int j = f / 32, k = f % 32; o.g0[f] = (T)wat(w.w0, d.F, d.S, j, k); continue; }
    f -= s.g0;
    if (f < s.g1) { const int j = f / kU, k = f % kU; o.g1[f] = (T)wat(w.w1, d.E, kU, j, k); continue; }
    f -= s.g1;
    if (f < s.g2) { const int j = f / kU, k = f % kU; o.g2[f] = (T)wat(w.w2, kU, kU, j, k); continue; }
    f -= s.g2;
    { const int j = f / 32, k = f % 32; o.g3[f] = (T)wat(w.w3, kU, 3, j, k); }
  }
}

// Adds a 16x16 weight-gradient tile (D layout) into the workgroup partial at
// rows [row0, row0+16) x cols [col0, col0+16) of a (rows x cols) block.
__device__ __forceinline__ void add_tile(float* __restrict__ part, int64_t off, int rows,
                                         int cols, int row0, int col0, const f32x4& acc) {
  const int lane = lane_id();
  const int col = col0 + (lane & 15);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int row = row0 + 4 * (lane >> 4) + e;
    if (row < rows && col < cols) part[off + (int64_t)row * cols + col] += acc[e];
  }
}

template <int TM>
inline size_t bwd_lds_floats(const MlpDims& d, int lds_tiles) {
  return (size_t)TM * (d.ldF + d.ldE + d.ldS + 2 * d.ldU + (d.ncam > 0 ? d.ldU : 0) + d.ld3 + 4 +
                       4) +
         (d.ncam > 0 ? kU + 4 : 0) +  // FiLM: running embedding gradient + 2 camera ids
         (size_t)lds_tiles * 256;     // weight-gradient tiles accumulated in LDS
}

// Weight-gradient tiles (16x16) of the four layers, numbered
// [dW3 | dW2 | dW1 | dW0]. With NT > 0 wave w accumulates tiles
// w, w+NW, ..., w+(NT-1)NW in registers over all its row tiles and stores
// them once at the end; with NT == 0 every row tile read-modify-writes the
// workgroup partial instead (used when the tiles do not fit NT*NW).
struct WTiles {
  int n3, n2, n1, n0, b2, b1, b0, total;
};

__host__ __device__ inline WTiles wtiles(const MlpDims& d) {
  WTiles t;
  t.n3 = kU / 16;
  t.n2 = (kU / 16) * (kU / 16);
  t.n1 = (d.EP / 16) * (kU / 16);
  t.n0 = (d.FP / 16) * 2;
  t.b2 = t.n3;
  t.b1 = t.b2 + t.n2;
  t.b0 = t.b1 + t.n1;
  t.total = t.b0 + t.n0;
  return t;
}

// Stores a 16x16 tile (D layout) into the (rows x cols) block at `off`.
__device__ __forceinline__ void store_tile(float* __restrict__ part, int64_t off, int rows,
                                           int cols, int row0, int col0, const f32x4& acc) {
  const int lane = lane_id();
  const int col = col0 + (lane & 15);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int row = row0 + 4 * (lane >> 4) + e;
    if (row < rows && col < cols) part[off + (int64_t)row * cols + col] = acc[e];
  }
}

// Weight-gradient phase of one layer: for every tile of the layer (n tiles,
// tcols column tiles per row of tiles) add X^T Y over the TM rows of the row
// tile. With NT > 0, wave w owns the layer's tiles w, w+NW, ..., held in the
// register slots [SLOT0, SLOT0 + CNT); otherwise tiles read-modify-write the
// workgroup partial.
template <int TM, int NW, bool BF, int NT, int SLOT0, int CNT, bool IN_LDS = false>
__device__ __forceinline__ void wgrad_phase(f32x4 (&wacc)[NT > 0 ? NT : 1], int wave, int n,
                                            int tcols, const float* Xm, int ldx,
                                            const float* Ym, int ldy, float* part, int64_t off,
                                            int rows, int cols, float* Wl = nullptr) {
  if constexpr (NT > 0 && IN_LDS) {
    // the wave's tiles in LDS slots [SLOT0, SLOT0 + CNT): each lane keeps its
    // own 4 accumulator values contiguous (one 16-byte read and write)
#pragma unroll
    for (int si = 0; si < CNT; ++si) {
      const int t = wave + NW * si;
      if (t < n) {
        const int ti = t / tcols, tj = t % tcols;
        f32x4* slot = reinterpret_cast<f32x4*>(Wl) + ((SLOT0 + si) * NW + wave) * 64 + lane_id();
        f32x4 acc = *slot;
        mma_wgrad_acc<BF, TM>(Xm, ldx, 16 * ti, Ym, ldy, 16 * tj, acc);
        *slot = acc;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else if constexpr (NT > 0) {
    // two-stage pipeline over the wave's tiles: the LDS reads of tile si + 1
    // are issued before the products of tile si, so their latency overlaps
    // the MFMAs (two operand sets live at a time)
    const int lane = lane_id();
    float a0[TM / 4], b0[TM / 4], a1[TM / 4], b1[TM / 4];
    if (wave < n) wgrad_load<BF, TM>(Xm, ldx, 16 * (wave / tcols), Ym, ldy, 16 * (wave % tcols),
                                     lane, a0, b0);
#pragma unroll
    for (int si = 0; si < CNT; ++si) {
      const int t = wave + NW * si, tn = t + NW;
      if (si + 1 < CNT && tn < n) {
        if (si & 1)
          wgrad_load<BF, TM>(Xm, ldx, 16 * (tn / tcols), Ym, ldy, 16 * (tn % tcols), lane, a0, b0);
        else
          wgrad_load<BF, TM>(Xm, ldx, 16 * (tn / tcols), Ym, ldy, 16 * (tn % tcols), lane, a1, b1);
      }
      if (t < n) {
        if (si & 1)
          wgrad_mma<BF, TM>(a1, b1, wacc[SLOT0 + si]);
        else
          wgrad_mma<BF, TM>(a0, b0, wacc[SLOT0 + si]);
      }
      // keeps the reads of tile si + 2 from being hoisted above these products
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    for (int t = wave; t < n; t += NW) {
      const int ti = t / tcols, tj = t % tcols;
      add_tile(part, off, rows, cols, 16 * ti, 16 * tj,
               mma_wgrad<BF, TM>(Xm, ldx, 16 * ti, Ym, ldy, 16 * tj));
    }
  }
}

// Weight-gradient tiles per wave (kBwdNW waves) of the standard shape: dW2
// (64 tiles) and dW1 (up to 80, encoded width <= 160) in register slots, the
// small dW3 (8 tiles) and dW0 (up to 20, features <= 160) in LDS slots, which
// keeps the kernel's registers within two waves per SIMD.
constexpr int kBwdNW = 8;
constexpr int kCnt2 = (64 + kBwdNW - 1) / kBwdNW, kSlot2 = 0;
constexpr int kCnt1 = (80 + kBwdNW - 1) / kBwdNW, kSlot1 = kSlot2 + kCnt2;
constexpr int kSlots = kSlot1 + kCnt1;
constexpr int kCnt3 = (8 + kBwdNW - 1) / kBwdNW, kLSlot3 = 0;
constexpr int kCnt0 = (20 + kBwdNW - 1) / kBwdNW, kLSlot0 = kLSlot3 + kCnt3;
constexpr int kLdsSlots = kLSlot0 + kCnt0;
// BF16 standard shape without an embedding: the rows of the dX strip's B
// table (g1 rows below fe1 = 138, 9 column tiles) are copied into LDS once per
// workgroup behind the weight-gradient slots, rows padded to 272 bytes (39 KB;
// 154 KB in all), so that strip reads LDS instead of L2.
constexpr int kG1Rows = 144, kG1Ld = 136;
constexpr size_t kG1LdsBytes = (size_t)kG1Rows * kG1Ld * 2;
// FP32: the first 80 rows (5 of the 9 column tiles, rows padded to 528 bytes,
// 41 KB; 153 KB in all); the other column tiles read L2.
constexpr int kG1RowsF = 80, kG1LdF = 132;
constexpr size_t kG1LdsBytesF = (size_t)kG1RowsF * kG1LdF * 4;
// With the FiLM embedding (the kept-activation backward, whose relu(z2)
// region takes 16.5 KB of the LDS): the first 80 BF16 rows (21 KB; 158 KB
// in all) or 48 FP32 rows (25 KB; 154 KB).
constexpr int kG1RowsFilm = 80, kG1RowsFilmF = 48;
constexpr size_t kG1FilmBytes = (size_t)kG1RowsFilm * kG1Ld * 2;
constexpr size_t kG1FilmBytesF = (size_t)kG1RowsFilmF * kG1LdF * 4;

// The backward that recomputes each tile's forward: callers without the
// activations kept by the forward kernel (the training steps' backward,
// which reads them, is k_mlp_bwd_kept in tfh_mlp_bwd_kept.h).
template <int TM, int NW, bool BF, int NT, bool STD, bool FILM>
__global__ void __launch_bounds__(NW * 64) k_mlp_bwd(MlpDims d_arg, MlpW w, MlpPack pk, GradPack gp_arg,
                                                     const float* __restrict__ feats,
                                                     const float* __restrict__ dirs,
                                                     const uint32_t* __restrict__ cams,
                                                     const float* __restrict__ grgb, int64_t M,
                                                     int64_t rpr, float* __restrict__ gfeat,
                                                     float* __restrict__ partials,
                                                     float* __restrict__ gembed) {
  constexpr int RT = TM / 16;
  static_assert(TM <= 64, "the FiLM camera check reads one tile row per lane");
  // row tiles per work item of the strip phases: all of them (the B strip is
  // loaded once per column tile) unless the workgroup has waves to spare
  constexpr int BRTI = NW >= 16 ? 1 : RT;
  constexpr int RG = RT / BRTI;
  // the specialised shape folds the camera count to 0 without an embedding
  const MlpDims d = STD ? (FILM ? std_dims<BF>(d_arg.ncam) : std_dims<BF>()) : d_arg;
  const GradPack gp = STD ? make_pack(std_dims<BF>()) : gp_arg;
  extern __shared__ float smem[];
  float* Fs = smem;                         // TM x ldF
  float* X = Fs + TM * d.ldF;               // TM x ldE   (later dX)
  float* H0 = X + TM * d.ldE;               // TM x ldS   (later dH0)
  float* H1 = H0 + TM * d.ldS;              // TM x ldU   (later dZ1)
  float* H2 = H1 + TM * d.ldU;              // TM x ldU   layer-3 input (later dZ2)
  float* after_h2 = H2 + TM * d.ldU;
  float* H2p = d.ncam > 0 ? after_h2 : H2;  // pre-FiLM relu(z2) (own region with FiLM)
  float* Z3 = d.ncam > 0 ? after_h2 + TM * d.ldU : after_h2;  // TM x ld3: rgb, then dZ3
  float* Ds = Z3 + TM * d.ld3;
  int* Cs = reinterpret_cast<int*>(Ds + TM * 4);
  // FiLM embedding gradient: tiles whose rows share one camera add their
  // column sums into Eacc (g * h2p for the scale half, g for the shift half);
  // Eacc goes to gembed with one atomic per column when the camera changes
  // and at the end. Tiles with several cameras add per row directly.
  float* Eacc = reinterpret_cast<float*>(Cs + 4 * TM);
  int* Run = reinterpret_cast<int*>(Eacc + kU);  // [0] tile camera or -1, [1] camera of Eacc or -1
  // LDS slots of the dW3 / dW0 tiles (NT > 0): after the FiLM state
  float* Wl = reinterpret_cast<float*>(Cs + 4 * TM) + (d.ncam > 0 ? kU + 4 : 0);
  const Operands<BF> op{d, w, pk};

  // masked to the block's waves (NW, a power of two): lets the compiler fold
  // each weight-gradient slot's tile offsets into constants
  static_assert((NW & (NW - 1)) == 0, "wave count a power of two");
  const int wave = (threadIdx.x >> 6) & (NW - 1);
  const int lane = lane_id();
  float* part = partials + (int64_t)blockIdx.x * gp.total;
  if (FILM && d.ncam > 0) {
    for (int c = threadIdx.x; c < kU; c += blockDim.x) Eacc[c] = 0.0f;
    if (threadIdx.x == 0) Run[1] = -1;
  }
  const WTiles wt = wtiles(d);
  f32x4 wacc[NT > 0 ? NT : 1];
#pragma unroll
  for (int i = 0; i < (NT > 0 ? NT : 1); ++i) wacc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Bias gradients, summed in registers: db1 / db2 per lane over the rows
  // 4(lane/16)+e of the row groups of column 16*(wave/RG) + lane%16 (the
  // strip phases that produce dZ1 and dZ2 give wave w the same item on every
  // tile), db3 per thread t < 3*TM over its (row t/3, colour t%3) slot;
  // reduced at the end.
  static_assert((kU / 16) * (RT / (NW >= 16 ? 1 : RT)) == NW,
                "one strip item per wave: the bias sums stay with their column");
  float db1 = 0.f, db2 = 0.f, db3 = 0.f;
  if (NT == 0)
    for (int64_t e = threadIdx.x; e < gp.total; e += blockDim.x) part[e] = 0.0f;
  else
    for (int e = threadIdx.x; e < kLdsSlots * NW * 256; e += blockDim.x) Wl[e] = 0.0f;
  constexpr bool kG1Lds = BF && STD && !FILM && NT > 0 && NW == 8;
  constexpr bool kG1LdsF = !BF && STD && !FILM && NT > 0 && NW == 8;
  __bf16* G1s = reinterpret_cast<__bf16*>(Wl + kLdsSlots * NW * 256);
  float* G1f = Wl + kLdsSlots * NW * 256;
  if constexpr (kG1Lds) {
    // the tile loop starts with a barrier
    for (int e = threadIdx.x; e < kG1Rows * (kU / 8); e += blockDim.x) {
      const int j = e / (kU / 8), c8 = e - j * (kU / 8);
      *reinterpret_cast<uint4*>(G1s + j * kG1Ld + 8 * c8) =
          *reinterpret_cast<const uint4*>(pk.g1 + j * kU + 8 * c8);
    }
  }
  if constexpr (kG1LdsF) {
    for (int e = threadIdx.x; e < kG1RowsF * (kU / 4); e += blockDim.x) {
      const int j = e / (kU / 4), c4 = e - j * (kU / 4);
      *reinterpret_cast<float4*>(G1f + j * kG1LdF + 4 * c4) =
          *reinterpret_cast<const float4*>(pk.fg1 + j * kU + 4 * c4);
    }
  }

  const int64_t ntiles = (M + TM - 1) / TM;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t row0 = tile * TM;
    __syncthreads();
    load_tile<TM>(d, feats, dirs, cams, row0, M, rpr, Fs, Ds, Cs);
    __syncthreads();
    if (FILM && d.ncam > 0 && wave == 0) {
      // read before the next use: forward_hidden synchronises the workgroup
      const int nvalid = (int)min((int64_t)TM, M - row0);
      const int c0 = Cs[0];
      const int tc = __all(lane >= nvalid || Cs[lane] == c0) ? c0 : -1;  // TM <= 64
      const int rc = Run[1];
      if (tc >= 0 && rc >= 0 && rc != tc) {
        for (int c = lane; c < kU; c += 64) {
          unsafeAtomicAdd(gembed + (size_t)rc * kU + c, Eacc[c]);
          Eacc[c] = 0.0f;
        }
      }
      if (lane == 0) {
        Run[0] = tc;
        if (tc >= 0) Run[1] = tc;
      }
    }
    forward_hidden<TM, NW, BF, BRTI>(d, w, op, Fs, X, H0, H1, H2p, H2, Ds, Cs);
    forward_out<TM, NW, BF>(d, w, op, H2, Z3);
    // dZ3 = grgb * s * (1 - s), zero outside the 3 colours and the valid rows
    for (int e = threadIdx.x; e < TM * d.Z3W; e += blockDim.x) {
      const int i = e / d.Z3W, c = e % d.Z3W;
      const int64_t row = row0 + i;
      float v = 0.0f;
      if (c < 3 && row < M) {
        const float s = Z3[i * d.ld3 + c];
        v = grgb[row * 3 + c] * s * (1.0f - s);
      }
      Z3[i * d.ld3 + c] = v;
    }
    __syncthreads();
    // dW3 (U x 3) += H2^T dZ3 ; db3
    wgrad_phase<TM, NW, BF, NT, kLSlot3, kCnt3, true>(wacc, wave, wt.n3, 1, H2, d.ldU, Z3, d.ld3,
                                                      part, gp.w3, kU, 3, Wl);
    if (threadIdx.x < 3 * TM) db3 += Z3[(threadIdx.x / 3) * d.ld3 + threadIdx.x % 3];
    __syncthreads();
    // Strip phases below run one work item per (column tile, row group).
    // dH2 = dZ3 @ W3^T; FiLM backward; dZ2 = dH2 * (h2 > 0) written over H2
    for (int it = wave; it < (kU / 16) * RG; it += NW) {
      const int ct = it / RG, r0 = 16 * BRTI * (it % RG);
      f32x4 acc[BRTI];
      zero_acc(acc);
      mma_strip<BRTI, 32>(Z3 + r0 * d.ld3, d.ld3, op.k3(), op.g3(), 16 * ct, acc);
      const int col = 16 * ct + (lane & 15);
      if constexpr (FILM) {
        const bool film = d.ncam > 0 && 16 * ct >= kU / 2;  // wave-uniform
        const int tcam = d.ncam > 0 ? Run[0] : -1;
        // the tile camera's scale of this column, read once
        const float tscale = (film && tcam >= 0) ? w.embed[(size_t)tcam * kU + col - kU / 2] : 0.0f;
        float sh = 0.0f, sg = 0.0f;
#pragma unroll
        for (int q = 0; q < 4 * BRTI; ++q) {
          const int rt = q / 4, e = q % 4;
          const int row = r0 + 16 * rt + 4 * (lane >> 4) + e;
          float g = acc[rt][e];
          const float hp = H2p[row * d.ldU + col];
          if (film) {
            if (row0 + row < M) {
              if (tcam >= 0) {
                sh += g * hp;
                sg += g;
              } else {
                float* gemb = gembed + (size_t)Cs[row] * kU;
                unsafeAtomicAdd(gemb + col - kU / 2, g * hp);
                unsafeAtomicAdd(gemb + col, g);
              }
            }
            g = g * (tcam >= 0 ? tscale : w.embed[(size_t)Cs[row] * kU + col - kU / 2]);
          }
          const float dz = hp > 0.0f ? g : 0.0f;
          H2[row * d.ldU + col] = dz;
          db2 += dz;
        }
        if (film && tcam >= 0) {
          sh += __shfl_xor(sh, 16, 64);
          sh += __shfl_xor(sh, 32, 64);
          sg += __shfl_xor(sg, 16, 64);
          sg += __shfl_xor(sg, 32, 64);
          if (lane < 16) {
            atomicAdd(Eacc + col - kU / 2, sh);  // LDS: row groups of one column may share it
            atomicAdd(Eacc + col, sg);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4 * BRTI; ++q) {
          const int rt = q / 4, e = q % 4;
          const int row = r0 + 16 * rt + 4 * (lane >> 4) + e;
          const float hp = H2p[row * d.ldU + col];
          const float dz = hp > 0.0f ? acc[rt][e] : 0.0f;
          H2[row * d.ldU + col] = dz;
          db2 += dz;
        }
      }
    }
    __syncthreads();
    // dW2 (U x U) += H1^T dZ2 ; db2
    wgrad_phase<TM, NW, BF, NT, kSlot2, kCnt2>(wacc, wave, wt.n2, kU / 16, H1, d.ldU, H2,
                                               d.ldU, part, gp.w2, kU, kU);
    __syncthreads();
    // dH1 = dZ2 @ W2^T ; dZ1 = dH1 * (h1 > 0) written over H1 (dW2 is done)
    for (int it = wave; it < (kU / 16) * RG; it += NW) {
      const int ct = it / RG, r0 = 16 * BRTI * (it % RG);
      f32x4 acc[BRTI];
      zero_acc(acc);
      mma_strip<BRTI, (BRTI == 1 ? 64 : kU)>(H2 + r0 * d.ldU, d.ldU, kU, op.g2(), 16 * ct, acc);
      const int col = 16 * ct + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4 * BRTI; ++q) {
        const int rt = q / 4, e = q % 4;
        const int row = r0 + 16 * rt + 4 * (lane >> 4) + e;
        const float h = H1[row * d.ldU + col];
        const float dz = h > 0.0f ? acc[rt][e] : 0.0f;
        H1[row * d.ldU + col] = dz;
        db1 += dz;
      }
    }
    __syncthreads();
    // dW1 (E x U) += X^T dZ1 ; db1
    wgrad_phase<TM, NW, BF, NT, kSlot1, kCnt1>(wacc, wave, wt.n1, kU / 16, X, d.ldE, H1,
                                               d.ldU, part, gp.w1, d.E, kU);
    __syncthreads();
    // dX = dZ1 @ W1^T for the columns that feed h0: [0, S) and the FE(h0)
    // block, i.e. the column tiles below fe1 (the 3 direction columns between
    // them are inside a tile); one row tile per item balances the waves
    {
      const int fe1 = d.S + 3 + 2 * d.nf * d.S;
      const int ndx = (fe1 + 15) / 16;
      for (int it = wave; it < ndx * RT; it += NW) {
        const int c0 = 16 * (it / RT), r0 = 16 * (it % RT);
        f32x4 acc[1];
        zero_acc(acc);
        if constexpr (kG1Lds) {
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, LdsB{G1s, kG1Ld}, c0, acc);
        } else if constexpr (kG1LdsF) {
          // one code path over a generic pointer: LDS rows or global rows
          // (two strip variants in a branch cost the kernel 7 more spills)
          const bool lds = c0 < kG1RowsF;  // wave-uniform
          const FlatB32 b{lds ? static_cast<const float*>(G1f) : pk.fg1, lds ? kG1LdF : kU};
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, b, c0, acc);
        } else {
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, op.g1(), c0, acc);
        }
        const int col = c0 + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = r0 + 4 * (lane >> 4) + e;
          X[row * d.ldE + col] = acc[0][e];  // dX (X is dead after dW1)
        }
      }
    }
    __syncthreads();
    // dH0[j] = dX[j] + sum_q dFE terms (networks.py:13-35); stored in H0 (zero padded)
    {
      const int fe0 = d.S + 3;
      for (int e = threadIdx.x; e < TM * kSP; e += blockDim.x) {
        const int i = e / kSP, j = e % kSP;
        float g = 0.0f;
        if (j < d.S) {
          const float h = H0[i * d.ldS + j];
          g = X[i * d.ldE + j];
          for (int q = 0; q < d.nf; ++q) {
            const float s = (float)(1 << q);
            float sv, cv;
            sincosf(h * s, &sv, &cv);
            // d/dh sin(s h) = s cos(s h), d/dh cos(s h) = -s sin(s h)
            g += X[i * d.ldE + fe0 + j * 2 * d.nf + q] * s * cv;
            g -= X[i * d.ldE + fe0 + j * 2 * d.nf + d.nf + q] * s * sv;
          }
        }
        H0[i * d.ldS + j] = g;
      }
    }
    __syncthreads();
    // dW0 (F x S) += F^T dH0
    wgrad_phase<TM, NW, BF, NT, kLSlot0, kCnt0, true>(wacc, wave, wt.n0, 2, Fs, d.ldF, H0, d.ldS,
                                                      part, gp.w0, d.F, d.S, Wl);
    // dF = dH0 @ W0^T -> global
    if (gfeat) {
      for (int it = wave; it < (d.FP / 16) * RT; it += NW) {
        const int ct = it / RT, r0 = 16 * (it % RT);
        f32x4 acc[1];
        zero_acc(acc);
        mma_strip<1, kSP>(H0 + r0 * d.ldS, d.ldS, kSP, op.g0(), 16 * ct, acc);
        const int col = 16 * ct + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = r0 + 4 * (lane >> 4) + e;
          const int64_t grow = row0 + row;
          if (grow < M && col < d.F) gfeat[grow * d.F + col] = acc[0][e];
        }
      }
    }
  }
  if (FILM && d.ncam > 0) {
    __syncthreads();
    const int rc = Run[1];
    if (rc >= 0)
      for (int c = threadIdx.x; c < kU; c += blockDim.x)
        unsafeAtomicAdd(gembed + (size_t)rc * kU + c, Eacc[c]);
  }
  // bias sums: db1 / db2 over the lane groups, then over the waves of a
  // column tile and db3 over the threads of a colour, in X (every wave is
  // past its last tile after the barrier): [db1 | db2 | db3]
  db1 += __shfl_xor(db1, 16, 64);
  db1 += __shfl_xor(db1, 32, 64);
  db2 += __shfl_xor(db2, 16, 64);
  db2 += __shfl_xor(db2, 32, 64);
  float* bsum = X;
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * kU + 3; e += blockDim.x) bsum[e] = 0.0f;
  __syncthreads();
  if (lane < 16) {
    const int bcol = 16 * (wave / RG) + lane;
    atomicAdd(bsum + bcol, db1);
    atomicAdd(bsum + kU + bcol, db2);
  }
  if (threadIdx.x < 3 * TM) atomicAdd(bsum + 2 * kU + threadIdx.x % 3, db3);
  __syncthreads();
  // the workgroup's gradient sums -> its partial
  if constexpr (NT > 0) {
#pragma unroll
    for (int si = 0; si < kSlots; ++si) {
      if (si < kSlot1) {
        const int t = wave + NW * (si - kSlot2);
        if (t < wt.n2)
          store_tile(part, gp.w2, kU, kU, 16 * (t / (kU / 16)), 16 * (t % (kU / 16)), wacc[si]);
      } else {
        const int t = wave + NW * (si - kSlot1);
        if (t < wt.n1)
          store_tile(part, gp.w1, d.E, kU, 16 * (t / (kU / 16)), 16 * (t % (kU / 16)), wacc[si]);
      }
    }
    // the LDS slots are the wave's own: no barrier needed
    const f32x4* wl = reinterpret_cast<const f32x4*>(Wl);
    for (int si = 0; si < kLdsSlots; ++si) {
      const f32x4 v = wl[(si * NW + wave) * 64 + lane];
      if (si < kLSlot0) {
        const int t = wave + NW * (si - kLSlot3);
        if (t < wt.n3) store_tile(part, gp.w3, kU, 3, 16 * t, 0, v);
      } else {
        const int t = wave + NW * (si - kLSlot0);
        if (t < wt.n0) store_tile(part, gp.w0, d.F, d.S, 16 * (t / 2), 16 * (t % 2), v);
      }
    }
    if (threadIdx.x < kU) {
      part[gp.b1 + threadIdx.x] = bsum[threadIdx.x];
      part[gp.b2 + threadIdx.x] = bsum[kU + threadIdx.x];
    }
    if (threadIdx.x < 3) part[gp.b3 + threadIdx.x] = bsum[2 * kU + threadIdx.x];
  } else {
    if (threadIdx.x < kU) {
      part[gp.b1 + threadIdx.x] += bsum[threadIdx.x];
      part[gp.b2 + threadIdx.x] += bsum[kU + threadIdx.x];
    }
    if (threadIdx.x < 3) part[gp.b3 + threadIdx.x] += bsum[2 * kU + threadIdx.x];
  }
}

constexpr int kFwdTM = 32, kFwdNW = 8;
constexpr int kBwdTM = 32;
constexpr int kBwdNT = kSlots;  // register-resident weight-gradient tiles per wave

// Dynamic LDS above 64 KB needs the per-kernel opt-in.
template <class K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    (void)hipGetLastError();
  }
}

// Backward launches of the standard shape without (tfh_mlp_bwd_std.hip) and
// with (tfh_mlp_film.hip) a camera embedding.
void launch_mlp_bwd_std(bool bf, unsigned nblk, size_t lds, hipStream_t st, const MlpDims& d,
                        const MlpW& w, const MlpPack& pk, const GradPack& gp,
                        const float* feats, const float* dirs, const uint32_t* cams,
                        const float* grgb, int64_t M, int64_t rpr, float* gfeat,
                        float* partials, float* gembed);
void launch_mlp_bwd_std_act(bool bf, unsigned nblk, size_t lds, hipStream_t st,
                            const MlpDims& d, const MlpW& w, const MlpPack& pk,
                            const GradPack& gp, const float* feats, const float* dirs,
                            const float* grgb, int64_t M, int64_t rpr, float* gfeat,
                            float* partials, const float* act, const float* rgb);
void launch_mlp_bwd_std_film_act(bool bf, unsigned nblk, size_t lds, hipStream_t st,
                                 const MlpDims& d, const MlpW& w, const MlpPack& pk,
                                 const GradPack& gp, const float* feats, const float* dirs,
                                 const uint32_t* cams, const float* grgb, int64_t M, int64_t rpr,
                                 float* gfeat, float* partials, float* gembed, const float* act,
                                 const float* rgb);
void launch_mlp_bwd_std_film(bool bf, unsigned nblk, size_t lds, hipStream_t st,
                             const MlpDims& d, const MlpW& w, const MlpPack& pk,
                             const GradPack& gp, const float* feats, const float* dirs,
                             const uint32_t* cams, const float* grgb, int64_t M, int64_t rpr,
                             float* gfeat, float* partials, float* gembed);

}  // namespace tfh
