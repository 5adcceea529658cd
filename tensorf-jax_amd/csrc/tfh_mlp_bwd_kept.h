// Backward of the standard FeatureMlp shape (networks.py:38-121; 144
// features, squash 27, 2+2 frequencies) that reads the activations the
// forward kernel kept (relu(z1), the layer-3 input -- relu(z2) before the
// FiLM camera embedding when there is one -- and the output): the fused
// training steps' backward.
//
// One workgroup (8 waves) per CU walks tiles of 32 rows; the weight gradients
// of dW1 / dW2 stay in registers of their owning wave and those of dW3 / dW0
// in LDS slots for the whole kernel (as in k_mlp_bwd). Per tile:
//   commit  the tile's inputs, loaded into registers during the previous
//           tile's dW2 phase, go to LDS; dZ3 = grgb * s * (1 - s)
//   T1      layer 0 as 8 items (2 column tiles x 2 row tiles x 2 K parts,
//           one per wave; the parts land in X[:, 0:S] and H0) ; dW3 ; db3
//   T2      encoding (h0 = the two parts' sum; sin / cos of the first
//           frequency by sincosf, the doubled one by the double-angle
//           identities) ; dZ2 = (dZ3 @ W3^T) * (h2 > 0) over H2 ; db2
//   T3      the next tile's loads are issued (dW2 reads LDS only, so no
//           wait of this phase queues behind them; in-order vmcnt) ; dW2
//   T4      dZ1 = (dZ2 @ W2^T) * (h1 > 0) over H1 ; db1
//   T5      dW1 ; dX = dZ1 @ W1^T (the columns that feed h0) over H2 / Z3
//   T6      dH0 from dX and the sin / cos the encoding left in X
//   T7      dW0 ; dF = dH0 @ W0^T
// with one barrier after each step: 8 per tile. With the camera embedding
// (FILM, networks.py:103-111) the kept relu(z2) goes to its own region H2p,
// T1 forms the layer-3 input H2 = e[:U/2] * h2 + e[U/2:] on the second half
// of the columns (e = embed[cam]; read once per tile when its rows share a
// camera), T2 runs dW3 beside the dZ2 strip, whose epilogue also takes the
// embedding's gradient and leaves dZ2 in H2p, and the LDS copy of the dX
// strip's B table holds fewer rows (the rest read L2).
#pragma once

#include "tfh_mlp_kernels.h"

namespace tfh {

constexpr int kKeptTM = 32, kKeptNW = 8;
constexpr int kKeptLdDX = 148;  // row stride of dX (144 columns)

// The tile's inputs in registers between issue and commit (kept activations
// in BF16 with the BF16 MLP).
template <int TM, int NT, bool BF>
struct KeptRegs {
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  static constexpr int kF = (TM * 160 / 4 + NT - 1) / NT;  // float4 of features per thread
  static constexpr int kA = TM * (BF ? 32 : 64) / NT;      // 16-byte activation loads per thread
  static_assert(TM * (BF ? 32 : 64) % NT == 0, "activation loads fill the block");
  i32x4 vf[kF];
  i32x4 va[kA];
  float s, g, dv;
  int cam;  // camera of row threadIdx.x (< TM), with an embedding
};

// Issues the loads of the tile at row0 (features, view directions, kept
// activations, output and its gradient) into r; every source is a buffer
// resource over the tile's valid rows, so rows past M read 0 (all of them
// for a tile past the end: the loop issues one unconditionally, which keeps
// the registers of r dead between commit and the next issue).
template <int TM, int NT, bool BF, bool FILM = false>
__device__ __forceinline__ void kept_issue(const MlpDims& d, const float* __restrict__ feats,
                                           const float* __restrict__ dirs,
                                           const float* __restrict__ act,
                                           const float* __restrict__ rgb,
                                           const float* __restrict__ grgb, int64_t row0,
                                           int64_t M, int64_t rpr, KeptRegs<TM, NT, BF>& r,
                                           const uint32_t* __restrict__ cams = nullptr) {
  const int nv = (int)max((int64_t)0, min((int64_t)TM, M - row0));
  const int64_t ray0 = row0 / rpr;
  const unsigned rem0 = (unsigned)(row0 - ray0 * rpr);
  const int nray = nv > 0 ? (int)((row0 + nv - 1) / rpr - ray0) + 1 : 0;
  const float* fbase = feats + row0 * d.F;
  asm volatile("" : "+s"(fbase));
  const auto rf = buffer_rsrc(fbase, nv * d.F * 4);
  const auto ra = buffer_rsrc(reinterpret_cast<const char*>(act) + row0 * kKeptRowBytes<BF>,
                              nv * kKeptRowBytes<BF>);
  const auto rs = buffer_rsrc(rgb + row0 * 3, nv * 12);
  const auto rg = buffer_rsrc(grgb + row0 * 3, nv * 12);
  const auto rd = buffer_rsrc(dirs + ray0 * 3, nray * 12);
  const int fp4 = d.FP / 4, f4 = d.F / 4, n4 = TM * fp4;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < KeptRegs<TM, NT, BF>::kF; ++k) {
    const int e = t + k * NT, i = e / fp4, c4 = e - i * fp4;
    const int off = (e < n4 && c4 < f4) ? (i * d.F + 4 * c4) * 4 : kOOB;
    r.vf[k] = __builtin_amdgcn_raw_buffer_load_b128(rf, off, 0, 0);
  }
#pragma unroll
  for (int k = 0; k < KeptRegs<TM, NT, BF>::kA; ++k)
    r.va[k] = __builtin_amdgcn_raw_buffer_load_b128(ra, (t + k * NT) * 16, 0, 0);
  const int i3 = t / 3, c3 = t - 3 * i3;
  const bool own3 = t < TM * 3;
  r.s = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, own3 ? t * 4 : kOOB, 0, 0));
  r.g = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rg, own3 ? t * 4 : kOOB, 0, 0));
  const unsigned rel = (rem0 + (unsigned)i3) / (unsigned)rpr;  // ray of row i3 after ray0
  r.dv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
      rd, (own3 && i3 < nv) ? (int)(rel * 3 + c3) * 4 : kOOB, 0, 0));
  if constexpr (FILM) {
    const auto rc = buffer_rsrc(cams + ray0, nray * 4);
    const unsigned relc = (rem0 + (unsigned)t) / (unsigned)rpr;  // ray of row t after ray0
    r.cam = (int)__builtin_amdgcn_raw_buffer_load_b32(rc, t < nv ? (int)relc * 4 : kOOB, 0, 0);
  }
  // the loads stay here, a phase ahead of their use: nothing moves a memory
  // access across this statement
  asm volatile("" ::: "memory");
}

// Stores the issued tile: features (Fs, zero padded to FP), directions (Ds),
// relu(z1) (H1), the layer-3 input (H2) and dZ3 (Z3, zero past the colours).
// With FILM the second half (relu(z2) before the embedding) goes to H2
// (the caller passes H2p) and each row's camera index, clamped to
// [0, ncam), to Cs.
template <int TM, int NT, bool BF, bool FILM = false>
__device__ __forceinline__ void kept_commit(const MlpDims& d, const KeptRegs<TM, NT, BF>& r,
                                            float* Fs, float* Ds, float* H1, float* H2,
                                            float* Z3, int* Cs = nullptr, int ncam = 0) {
  const int fp4 = d.FP / 4, n4 = TM * fp4;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < KeptRegs<TM, NT, BF>::kF; ++k) {
    const int e = t + k * NT, i = e / fp4, c4 = e - i * fp4;
    if (e < n4)
      *reinterpret_cast<float4*>(Fs + i * d.ldF + 4 * c4) = __builtin_bit_cast(float4, r.vf[k]);
  }
  const int i3 = t / 3, c3 = t - 3 * i3;
  if (t < TM * 3) {
    Ds[i3 * 3 + c3] = r.dv;
    Z3[i3 * d.ld3 + c3] = r.g * r.s * (1.0f - r.s);
  }
#pragma unroll
  for (int k = 0; k < KeptRegs<TM, NT, BF>::kA; ++k) {
    if constexpr (BF) {
      const int e = t + k * NT, i = e >> 5, c8 = e & 31;
      float* dst = c8 < 16 ? H1 + i * d.ldU + 8 * c8 : H2 + i * d.ldU + 8 * (c8 - 16);
      float4 lo, hi;
      bf16x8_to_f32(__builtin_bit_cast(int4, r.va[k]), lo, hi);
      *reinterpret_cast<float4*>(dst) = lo;
      *reinterpret_cast<float4*>(dst + 4) = hi;
    } else {
      const int e = t + k * NT, i = e >> 6, c4 = e & 63;
      float* dst = c4 < 32 ? H1 + i * d.ldU + 4 * c4 : H2 + i * d.ldU + 4 * (c4 - 32);
      *reinterpret_cast<float4*>(dst) = __builtin_bit_cast(float4, r.va[k]);
    }
  }
  for (int e = t; e < TM * (d.Z3W - 3); e += NT) {
    const int i = e / (d.Z3W - 3), c = 3 + e % (d.Z3W - 3);
    Z3[i * d.ld3 + c] = 0.0f;
  }
  if constexpr (FILM) {
    if (t < TM) Cs[t] = (r.cam < 0 || r.cam >= ncam) ? 0 : r.cam;
  }
}

// FILM: cams (per ray), gembed (ncam x U, added to) and ncam >= 1.
template <bool BF, bool FILM = false>
__global__ void __launch_bounds__(kKeptNW * 64) k_mlp_bwd_kept(
    MlpW w, MlpPack pk, const float* __restrict__ feats, const float* __restrict__ dirs,
    const float* __restrict__ grgb, int64_t M, int64_t rpr, float* __restrict__ gfeat,
    float* __restrict__ partials, const float* __restrict__ act,
    const float* __restrict__ rgb_in, const uint32_t* __restrict__ cams,
    float* __restrict__ gembed, int ncam) {
  constexpr int TM = kKeptTM, NW = kKeptNW, NT = NW * 64, RT = TM / 16;
  constexpr MlpDims d = std_dims<BF>();
  constexpr GradPack gp = make_pack(d);
  static_assert(TM * kKeptLdDX <= TM * (d.ldU + d.ld3), "dX fits over H2 and Z3");
  static_assert(kSP <= d.ldS && d.S <= kSP, "layer-0 part B fits H0");
  extern __shared__ float smem[];
  float* Fs = smem;                 // TM x ldF
  float* X = Fs + TM * d.ldF;       // TM x ldE
  float* H0 = X + TM * d.ldE;       // TM x ldS: layer-0 part B, later dH0
  float* H1 = H0 + TM * d.ldS;      // TM x ldU: relu(z1), later dZ1
  float* H2 = H1 + TM * d.ldU;      // TM x ldU: layer-3 input, later dZ2 (no FILM), then dX
  // FILM: relu(z2) before the embedding, later dZ2 (without FILM: H2 itself)
  float* H2p = FILM ? H2 + TM * d.ldU : H2;
  float* Z3 = H2p + TM * d.ldU;     // TM x ld3: dZ3
  float* Ds = Z3 + TM * d.ld3;      // TM x 4
  int* Cs = reinterpret_cast<int*>(Ds + TM * 4);  // TM x 4: camera of each row (FILM)
  // FILM: running embedding gradient of camera Run[1] (g * h2 for the scale
  // half, g for the shift half), added to gembed when the camera changes and
  // at the end; Run[0]: the tile's camera when its rows share one, else -1
  float* Eacc = reinterpret_cast<float*>(Cs + TM * 4);
  int* Run = reinterpret_cast<int*>(Eacc + kU);
  float* Wl = FILM ? Eacc + kU + 4 : Eacc;  // weight-gradient LDS slots
  float* DX = H2;                   // TM x kKeptLdDX (over H2 and the dead dZ2 / dZ3 rows)
  float* DZ2 = H2p;
  const Operands<BF> op{d, w, pk};
  const int wave = (threadIdx.x >> 6) & (NW - 1);  // < NW: slot tile offsets fold to constants
  const int lane = lane_id();
  float* part = partials + (int64_t)blockIdx.x * gp.total;
  const WTiles wt = wtiles(d);
  f32x4 wacc[kSlots];
#pragma unroll
  for (int i = 0; i < kSlots; ++i) wacc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  // bias gradients: db1 / db2 per lane over the rows of column 16*wave +
  // lane%16 (the strip items of wave w are column tile w), db3 per thread
  // t < 3*TM over its (row t/3, colour t%3) slot; reduced at the end
  float db1 = 0.f, db2 = 0.f, db3 = 0.f;
  for (int e = threadIdx.x; e < kLdsSlots * NW * 256; e += blockDim.x) Wl[e] = 0.0f;
  if constexpr (FILM) {
    for (int c = threadIdx.x; c < kU; c += blockDim.x) Eacc[c] = 0.0f;
    if (threadIdx.x == 0) Run[1] = -1;
  }
  // the first rows of the dX strip's B table in LDS (as k_mlp_bwd; fewer
  // with FILM)
  __bf16* G1s = reinterpret_cast<__bf16*>(Wl + kLdsSlots * NW * 256);
  float* G1f = Wl + kLdsSlots * NW * 256;
  constexpr int kG1B = FILM ? kG1RowsFilm : kG1Rows, kG1F = FILM ? kG1RowsFilmF : kG1RowsF;
  if constexpr (BF) {
    for (int e = threadIdx.x; e < kG1B * (kU / 8); e += blockDim.x) {
      const int j = e / (kU / 8), c8 = e - j * (kU / 8);
      *reinterpret_cast<uint4*>(G1s + j * kG1Ld + 8 * c8) =
          *reinterpret_cast<const uint4*>(pk.g1 + j * kU + 8 * c8);
    }
  } else {
    for (int e = threadIdx.x; e < kG1F * (kU / 4); e += blockDim.x) {
      const int j = e / (kU / 4), c4 = e - j * (kU / 4);
      *reinterpret_cast<float4*>(G1f + j * kG1LdF + 4 * c4) =
          *reinterpret_cast<const float4*>(pk.fg1 + j * kU + 4 * c4);
    }
  }
  // layer-0 K split: part A over K [0, KA), part B over [KA, FP)
  constexpr int KA = BF ? 96 : 80;
  static_assert(KA % (BF ? 32 : 16) == 0 && (d.FP - KA) % (BF ? 32 : 16) == 0 && KA < d.FP,
                "layer-0 K parts in whole MFMA steps");

  const int64_t ntiles = (M + TM - 1) / TM;
  KeptRegs<TM, NT, BF> nx;
  kept_issue<TM, NT, BF, FILM>(d, feats, dirs, act, rgb_in, grgb, (int64_t)blockIdx.x * TM, M, rpr,
                               nx, cams);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t row0 = tile * TM;
    __syncthreads();
    kept_commit<TM, NT, BF, FILM>(d, nx, Fs, Ds, H1, H2p, Z3, Cs, ncam);
    // the layer-0 item's B chunks (item wave: K part wave & 1, column tile
    // (wave >> 1) & 1), requested before the barrier
    constexpr int K0A = BF ? 96 : 80;
    StripPre<BF, K0A> pre0l;
    {
      const int kp = wave & 1, ct = (wave >> 1) & 1;
      if constexpr (BF)
        strip_preload<K0A>(PackedB(pk.f0 + (kp ? K0A : 0), d.FP, 32), kp ? d.FP - K0A : K0A,
                           16 * ct, pre0l);
      else
        strip_preload<K0A>(PackedB32(pk.ff0 + (kp ? K0A : 0), d.FP, 32),
                           kp ? d.FP - K0A : K0A, 16 * ct, pre0l);
    }
    __syncthreads();
    // T1: layer 0, one (column tile, row tile, K part) per wave -> X / H0 ;
    // dW3 += H2^T dZ3 ; db3
    for (int it = wave; it < 2 * RT * 2; it += NW) {
      const int kp = it & 1, ct = (it >> 1) & 1, r0 = 16 * (it >> 2);
      f32x4 acc[1];
      zero_acc(acc);
      static_assert(K0A == KA, "preloaded layer-0 chunks match the K split");
      if constexpr (BF) {
        const PackedB b(pk.f0 + (kp ? KA : 0), d.FP, 32);
        if (it == wave)
          mma_strip<1, KA>(Fs + r0 * d.ldF + (kp ? KA : 0), d.ldF, kp ? d.FP - KA : KA, b,
                           16 * ct, acc, pre0l);
        else
          mma_strip<1, KA>(Fs + r0 * d.ldF + (kp ? KA : 0), d.ldF, kp ? d.FP - KA : KA, b,
                           16 * ct, acc);
      } else {
        const PackedB32 b(pk.ff0 + (kp ? KA : 0), d.FP, 32);
        if (it == wave)
          mma_strip<1, KA>(Fs + r0 * d.ldF + (kp ? KA : 0), d.ldF, kp ? d.FP - KA : KA, b,
                           16 * ct, acc, pre0l);
        else
          mma_strip<1, KA>(Fs + r0 * d.ldF + (kp ? KA : 0), d.ldF, kp ? d.FP - KA : KA, b,
                           16 * ct, acc);
      }
      float* dst = kp ? H0 : X;
      const int ld = kp ? d.ldS : d.ldE;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + 4 * (lane >> 4) + e, col = 16 * ct + (lane & 15);
        if (col < d.S) dst[row * ld + col] = acc[0][e];
      }
    }
    if constexpr (FILM) {
      // the tile's camera (every wave alike) ; wave 0 hands the running
      // embedding gradient over when the camera changes ; H2 = the layer-3
      // input from relu(z2) (forward_hidden's layer 2), a thread per column
      const int nvalid = (int)min((int64_t)TM, M - row0);
      const int c0 = Cs[0];
      const int tc = __all(lane >= nvalid || Cs[lane] == c0) ? c0 : -1;  // TM <= 64
      if (wave == 0) {
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
      static_assert(NT % kU == 0, "a thread keeps one column");
      const int c = threadIdx.x % kU;
      const bool film = c >= kU / 2;
      float esc = 1.0f, esh = 0.0f;
      if (film && tc >= 0) {
        esc = w.embed[(size_t)tc * kU + c - kU / 2];
        esh = w.embed[(size_t)tc * kU + c];
      }
      for (int i = threadIdx.x / kU; i < TM; i += NT / kU) {
        const float h = H2p[i * d.ldU + c];
        float out = h;
        if (film) {
          if (tc >= 0) {
            out = esc * h + esh;
          } else {
            const float* emb = w.embed + (size_t)Cs[i] * kU;
            out = emb[c - kU / 2] * h + emb[c];
          }
        }
        H2[i * d.ldU + c] = out;
      }
    } else {
      wgrad_phase<TM, NW, BF, kSlots, kLSlot3, kCnt3, true>(wacc, wave, wt.n3, 1, H2, d.ldU, Z3,
                                                            d.ld3, part, gp.w3, kU, 3, Wl);
      if (threadIdx.x < 3 * TM) db3 += Z3[(threadIdx.x / 3) * d.ld3 + threadIdx.x % 3];
    }
    // the dZ2 strip's B chunks (column tile wave), requested before the barrier
    StripPre<BF, 32> pre3;
    strip_preload<32>(op.g3(), op.k3(), 16 * wave, pre3);
    __syncthreads();
    // T2: encoding ; dZ2 = (dZ3 @ W3^T) * (h2 > 0) over H2 ; db2
    encode_tile<TM, true>(d, X, Ds, H0, d.ldS);
    {
      const int ct = wave;  // 8 column tiles, all row tiles per item
      f32x4 acc[RT];
      zero_acc(acc);
      mma_strip<RT, 32>(Z3, d.ld3, op.k3(), op.g3(), 16 * ct, acc, pre3);
      const int col = 16 * ct + (lane & 15);
      if constexpr (FILM) {
        // dH2 -> the embedding's gradient (tile camera: column sums into
        // Eacc; several cameras: per row into gembed) and d relu(z2) =
        // dH2 * scale on the second half; dZ2 over H2p
        const bool film = 16 * ct >= kU / 2;  // wave-uniform
        const int tcam = Run[0];
        const float tscale = (film && tcam >= 0) ? w.embed[(size_t)tcam * kU + col - kU / 2] : 0.0f;
        float sh = 0.0f, sg = 0.0f;
#pragma unroll
        for (int q = 0; q < 4 * RT; ++q) {
          const int rt = q / 4, e = q % 4;
          const int row = 16 * rt + 4 * (lane >> 4) + e;
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
          H2p[row * d.ldU + col] = dz;
          db2 += dz;
        }
        if (film && tcam >= 0) {
          sh += __shfl_xor(sh, 16, 64);
          sh += __shfl_xor(sh, 32, 64);
          sg += __shfl_xor(sg, 16, 64);
          sg += __shfl_xor(sg, 32, 64);
          if (lane < 16) {  // the item covers every row of the column: plain adds
            Eacc[col - kU / 2] += sh;
            Eacc[col] += sg;
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4 * RT; ++q) {
          const int rt = q / 4, e = q % 4;
          const int row = 16 * rt + 4 * (lane >> 4) + e;
          const float h = H2[row * d.ldU + col];
          const float dz = h > 0.0f ? acc[rt][e] : 0.0f;
          H2[row * d.ldU + col] = dz;
          db2 += dz;
        }
      }
    }
    if constexpr (FILM) {
      // dW3 += H2^T dZ3 ; db3 (H2 stays the layer-3 input through this phase)
      wgrad_phase<TM, NW, BF, kSlots, kLSlot3, kCnt3, true>(wacc, wave, wt.n3, 1, H2, d.ldU, Z3,
                                                            d.ld3, part, gp.w3, kU, 3, Wl);
      if (threadIdx.x < 3 * TM) db3 += Z3[(threadIdx.x / 3) * d.ld3 + threadIdx.x % 3];
    }
    __syncthreads();
    // T3: the next tile's loads (nothing in this phase waits on vmcnt) ;
    // dW2 += H1^T dZ2
    kept_issue<TM, NT, BF, FILM>(d, feats, dirs, act, rgb_in, grgb,
                                 row0 + (int64_t)gridDim.x * TM, M, rpr, nx, cams);
    wgrad_phase<TM, NW, BF, kSlots, kSlot2, kCnt2>(wacc, wave, wt.n2, kU / 16, H1, d.ldU, DZ2,
                                                   d.ldU, part, gp.w2, kU, kU);
    // the dZ1 strip's first B chunks, requested before the barrier
    StripPre<BF, kU> pre2;
    strip_preload<kU>(op.g2(), kU, 16 * wave, pre2);
    __syncthreads();
    // T4: dZ1 = (dZ2 @ W2^T) * (h1 > 0) over H1 ; db1
    {
      const int ct = wave;
      f32x4 acc[RT];
      zero_acc(acc);
      mma_strip<RT, kU>(DZ2, d.ldU, kU, op.g2(), 16 * ct, acc, pre2);
      const int col = 16 * ct + (lane & 15);
#pragma unroll
      for (int q = 0; q < 4 * RT; ++q) {
        const int rt = q / 4, e = q % 4;
        const int row = 16 * rt + 4 * (lane >> 4) + e;
        const float h = H1[row * d.ldU + col];
        const float dz = h > 0.0f ? acc[rt][e] : 0.0f;
        H1[row * d.ldU + col] = dz;
        db1 += dz;
      }
    }
    __syncthreads();
    // T5: dW1 += X^T dZ1 ; dX = dZ1 @ W1^T for the column tiles below fe1
    // (the columns that feed h0) -> DX
    wgrad_phase<TM, NW, BF, kSlots, kSlot1, kCnt1>(wacc, wave, wt.n1, kU / 16, X, d.ldE, H1,
                                                   d.ldU, part, gp.w1, d.E, kU);
    {
      constexpr int fe1 = d.S + 3 + 2 * d.nf * d.S;
      constexpr int ndx = (fe1 + 15) / 16;
      static_assert(16 * ndx <= kKeptLdDX, "dX columns fit the stride");
      for (int it = wave; it < ndx * RT; it += NW) {
        const int c0 = 16 * (it / RT), r0 = 16 * (it % RT);
        f32x4 acc[1];
        zero_acc(acc);
        if constexpr (FILM && BF) {
          // LDS rows or L2 rows through one generic pointer
          const bool lds = c0 < kG1B;  // wave-uniform
          const LdsB b{lds ? static_cast<const __bf16*>(G1s) : pk.g1, lds ? kG1Ld : kU};
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, b, c0, acc);
        } else if constexpr (BF) {
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, LdsB{G1s, kG1Ld}, c0, acc);
        } else {
          const bool lds = c0 < kG1F;  // wave-uniform
          const FlatB32 b{lds ? static_cast<const float*>(G1f) : pk.fg1, lds ? kG1LdF : kU};
          mma_strip<1, kU>(H1 + r0 * d.ldU, d.ldU, kU, b, c0, acc);
        }
        const int col = c0 + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = r0 + 4 * (lane >> 4) + e;
          DX[row * kKeptLdDX + col] = acc[0][e];
        }
      }
    }
    __syncthreads();
    // T6: dH0[j] = dX[j] + sum_q 2^q (dX[sin_q] cos(2^q h) - dX[cos_q] sin(2^q h))
    // (networks.py:13-35), the sines and cosines read back from X; zero
    // padded to kSP columns
    {
      constexpr int fe0 = d.S + 3;
      for (int e = threadIdx.x; e < TM * kSP; e += blockDim.x) {
        const int i = e / kSP, j = e % kSP;
        float g = 0.0f;
        if (j < d.S) {
          const float* dx = DX + i * kKeptLdDX;
          const float* xe = X + i * d.ldE + fe0 + j * 2 * d.nf;
          g = dx[j];
#pragma unroll
          for (int q = 0; q < d.nf; ++q) {
            const float s = (float)(1 << q);
            const float sv = xe[q], cv = xe[d.nf + q];
            g += dx[fe0 + j * 2 * d.nf + q] * s * cv;
            g -= dx[fe0 + j * 2 * d.nf + d.nf + q] * s * sv;
          }
        }
        H0[i * d.ldS + j] = g;
      }
    }
    // the first dF item's B chunks. The dF items count down from the last
    // wave while the dW0 tiles count up from wave 0, so a wave with a third
    // dW0 tile has two dF items (FP32: 9 tile products per SIMD instead of
    // 10 on two of them)
    const int it0 = NW - 1 - wave;
    StripPre<BF, kSP> pre0;
    strip_preload<kSP>(op.g0(), kSP, 16 * (it0 / RT), pre0);
    __syncthreads();
    // T7: dW0 += Fs^T dH0 ; dF = dH0 @ W0^T -> gfeat
    wgrad_phase<TM, NW, BF, kSlots, kLSlot0, kCnt0, true>(wacc, wave, wt.n0, 2, Fs, d.ldF, H0,
                                                          d.ldS, part, gp.w0, d.F, d.S, Wl);
    for (int it = it0; it < (d.FP / 16) * RT; it += NW) {
      const int ct = it / RT, r0 = 16 * (it % RT);
      f32x4 acc[1];
      zero_acc(acc);
      if (it == it0)
        mma_strip<1, kSP>(H0 + r0 * d.ldS, d.ldS, kSP, op.g0(), 16 * ct, acc, pre0);
      else
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
  if constexpr (FILM) {
    __syncthreads();
    const int rc = Run[1];
    if (rc >= 0)
      for (int c = threadIdx.x; c < kU; c += blockDim.x)
        unsafeAtomicAdd(gembed + (size_t)rc * kU + c, Eacc[c]);
  }
  // bias sums: db1 / db2 over the lane groups, db3 over the threads of a
  // colour, in X (every wave is past its last tile after the barrier)
  db1 += __shfl_xor(db1, 16, 64);
  db1 += __shfl_xor(db1, 32, 64);
  db2 += __shfl_xor(db2, 16, 64);
  db2 += __shfl_xor(db2, 32, 64);
  float* bsum = X;
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * kU + 3; e += blockDim.x) bsum[e] = 0.0f;
  __syncthreads();
  if (lane < 16) {
    bsum[16 * wave + lane] = db1;
    bsum[kU + 16 * wave + lane] = db2;
  }
  if (threadIdx.x < 3 * TM) atomicAdd(bsum + 2 * kU + threadIdx.x % 3, db3);
  __syncthreads();
  // the workgroup's gradient sums -> its partial
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
}

}  // namespace tfh
