// render_rays (render.py:105-279), its backward and the fused training step
// (training.py:101-205) on gfx950.
//
// Forward (one ray per wave64):
//   k_density_pass   sampling (bbox or contraction) -> normalise -> density VM
//                    gather (16 lanes per sample) -> softplus -> log-space
//                    transmittance scan (lane per sample) -> weights ->
//                    distance outputs or per-ray top-k + unbiasing terms
//   k_vm_gather      appearance features of the selected samples, (M, 3Ca)
//   mlp_fwd          FeatureMlp on MFMA
//   k_composite_rgb  U * sum_k w_k c_k + E_{N-1} (+ loss gradient when training)
// Backward:
//   k_composite_bwd  reverse suffix scans -> dL/dsigma -> density scatter
//                    (wave pre-reduced atomics) and dL/dc for the MLP
//   mlp_bwd          recompute + backward of the MLP -> dL/dfeatures
//   k_vm_scatter     appearance factor gradients
#include <algorithm>

#include "tfh_prop.h"
#include "tfh_scatter_sorted.h"

namespace tfh {

int32_t check_mlp(const tfh_mlp* p, const char* who);
// act (optional): (M, 256) activations the forward keeps for the backward,
// which then also reads the forward's rgb; only for mlp_keeps_activations().
int32_t mlp_fwd_rows(const tfh_mlp* p, const float* feats, const float* dirs, const uint32_t* cams,
                     int64_t M, int64_t rows_per_ray, float* rgb, int32_t precision, void* ws,
                     size_t ws_bytes, hipStream_t st, float* act = nullptr);
int32_t mlp_fwd_rows_gather(const tfh_mlp* p, const tfh_vm& vm, const RaySamples& src,
                            float* feats, const float* dirs, const uint32_t* cams, int64_t M,
                            int64_t rpr, float* rgb, int32_t precision, void* ws,
                            size_t ws_bytes, hipStream_t st, float* act, bool* done);
size_t mlp_bwd_workspace(const tfh_mlp* p, int64_t M, int32_t precision);
int32_t mlp_bwd_rows(const tfh_mlp* p, const float* feats, const float* dirs, const uint32_t* cams,
                     int64_t M, int64_t rows_per_ray, const float* grgb, float* gfeat,
                     const tfh_mlp_grad* g, int32_t precision, void* ws, size_t ws_bytes,
                     hipStream_t st, const float* act = nullptr, const float* rgb = nullptr);
bool mlp_keeps_activations(const tfh_mlp* p, int32_t precision);
// hyper (optional, device): -lr_lo * decay, -lr_hi * decay, 1 - b1^count,
// 1 - b2^count, read by the kernel in place of the host values.
int32_t adam_launch(float* p, float* g, float* m, float* v, int64_t n, int64_t split, float lr_lo,
                    float lr_hi, float decay, float b1, float b2, float eps, int64_t count,
                    int zero_grad, hipStream_t st, const float* hyper = nullptr);

// ---- per-ray selection (one wave) -----------------------------------------------
// Top-k of the keys (LDS, order-preserving uint32) -> ascending indices
// into idx_out (global, k entries). Ties: lower index first. Returns the sum
// of s_w over the selection (wave-uniform).
__device__ float wave_topk(const uint32_t* key, const float* s_w, int n, int k,
                           int32_t* __restrict__ idx_out) {
  const int lane = lane_id();
  uint32_t T = 0;
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t cand = T | (1u << bit);
    int cnt = 0;
    for (int c0 = 0; c0 < n; c0 += 64) {
      const int i = c0 + lane;
      cnt += __popcll(__ballot(i < n && key[i] >= cand));
    }
    if (cnt >= k) T = cand;
  }
  int n_gt = 0;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int i = c0 + lane;
    n_gt += __popcll(__ballot(i < n && key[i] > T));
  }
  const int need_eq = k - n_gt;
  int base = 0, eq_base = 0;
  float wsum = 0.0f;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int i = c0 + lane;
    const bool act = i < n;
    const uint32_t kv = act ? key[i] : 0u;
    const uint64_t eq_mask = __ballot(act && kv == T);
    const bool eq_sel = act && kv == T && (eq_base + lanes_below(eq_mask)) < need_eq;
    const bool sel = (act && kv > T) || eq_sel;
    const uint64_t sel_mask = __ballot(sel);
    if (sel) {
      idx_out[base + lanes_below(sel_mask)] = i;
      wsum += s_w[i];
    }
    base += __popcll(sel_mask);
    eq_base += __popcll(eq_mask);
  }
  return wave_sum(wsum);
}

struct DensityOut {
  float* pts;    // (R*N, 3) normalised points
  float* ts;     // (R*N) or null
  float* sigma;  // (R*N)
  float* delta;  // (R*N)
  float* w;      // (R*N)
  float* e;      // (R*N)
  int32_t* idx;  // (R*k) when k < N
  float* ray;    // (R*4): E_{N-1}, sum_K w, unbias
  float* out;    // (R) distance outputs
};

__global__ void __launch_bounds__(64) k_density_pass(tfh_vm dvm, RayIn in, DensityOut o) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  const int lane = threadIdx.x;
  const int N = in.N;
  float* s_a = sm;       // a_i = sigma_i * delta_i  (later: top-k keys)
  float* s_e = sm + N;   // E_i
  float* s_w = s_e + N;  // w_i
  const float ox = in.o[3 * r], oy = in.o[3 * r + 1], oz = in.o[3 * r + 2];
  const float dx = in.d[3 * r], dy = in.d[3 * r + 1], dz = in.d[3 * r + 2];
  float t0 = 0.f, step = 0.f;
  if (!in.contraction && in.sampling == TFH_SAMPLING_SCHEDULE)
    bbox_segment(ox, oy, oz, dx, dy, dz, in.aabb, N, t0, step);
  const float a0x = in.aabb[0], a0y = in.aabb[1], a0z = in.aabb[2];
  const float a1x = in.aabb[3], a1y = in.aabb[4], a1z = in.aabb[5];
  const int G = (int)dvm.grid_dim, C = (int)dvm.channels;
  // 16-byte taps when the channels allow: lps lanes per sample (C/4 rounded
  // up to a power of two), 64/lps samples per step; else 16 lanes and
  // 4-byte taps
  const bool wide = in.wide_taps;
  const int lps = !wide ? 16 : (C <= 8 ? 2 : (C <= 16 ? 4 : (C <= 32 ? 8 : 16)));
  const int gs = lane / lps, ls = lane % lps;

  for (int i0 = 0; i0 < N; i0 += 64 / lps) {
    const int i = i0 + gs;
    const bool act = i < N;
    float t = 0.f, dlt = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    if (act) sample_point(in, r, i, ox, oy, oz, dx, dy, dz, t0, step, t, dlt, px, py, pz);
    const float xh = normalise(px, a0x, a1x), yh = normalise(py, a0y, a1y),
                zh = normalise(pz, a0z, a1z);
    ModeTaps mt[3];
    mode_taps(xh, yh, zh, G, mt);
    float acc = 0.0f;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      TapAddr a;
      tap_addrs(mt[m], m, G, a);
      if (wide) {
        if (act && 4 * ls < C) {
          float4 lv, pv;
          tap_values4(dvm.plane, dvm.line, C, 4 * ls, a, lv, pv);
          acc += lv.x * pv.x;
          acc += lv.y * pv.y;
          acc += lv.z * pv.z;
          acc += lv.w * pv.w;
        }
      } else {
#pragma unroll
        for (int q = 0; q < kMaxQ; ++q) {
          const int c = ls + 16 * q;
          if (c < C && act) {
            float lv, pv;
            tap_values(dvm.plane, dvm.line, C, c, a, lv, pv);
            acc += lv * pv;
          }
        }
      }
    }
    // sum over the 3C channels: the sample's lps lanes
    for (int o = 1; o < lps; o <<= 1) acc += __shfl_xor(acc, o, 64);
    if (act && ls == 0) {
      const float sig = softplus(acc + in.density_shift);  // render.py:207 (+10 there)
      s_a[i] = sig * dlt;                          // render.py:322
      const int64_t s = r * N + i;
      o.sigma[s] = sig;
      o.delta[s] = dlt;
      if (o.ts) o.ts[s] = t;
      o.pts[3 * s] = xh;
      o.pts[3 * s + 1] = yh;
      o.pts[3 * s + 2] = zh;
    }
  }
  __syncthreads();
  wave_segment_probs(s_a, s_e, s_w, N);
  __syncthreads();
  for (int i = lane; i < N; i += 64) {
    o.w[r * N + i] = s_w[i];
    o.e[r * N + i] = s_e[i];
  }
  const float e_last = s_e[N - 1];

  if (in.mode == TFH_MODE_DIST_MEAN) {  // render.py:269-276
    float acc = 0.0f;
    for (int i = lane; i < N; i += 64) {
      float t, dlt, px, py, pz;
      sample_point(in, r, i, ox, oy, oz, dx, dy, dz, t0, step, t, dlt, px, py, pz);
      acc += s_w[i] * t;
      if (i == N - 1) acc += e_last * t;
    }
    acc = wave_sum(acc);
    if (lane == 0) o.out[r] = acc;
    return;
  }
  if (in.mode == TFH_MODE_DIST_MEDIAN_AS_WRITTEN) {
    // render.py:250-266 as written: sum(mask * [ts, +inf]) with mask[N] =
    // (1 - E_{N-1} <= 0.5); the last column is 0 * inf = NaN or 1 * inf, and
    // the finite columns cannot undo either.
    if (lane == 0) o.out[r] = (1.0f - e_last) > 0.5f ? NAN : INFINITY;
    return;
  }
  if (in.mode == TFH_MODE_DIST_MEDIAN) {  // render.py:248-267
    int first = N;
    for (int c0 = 0; c0 < N && first == N; c0 += 64) {
      const int i = c0 + lane;
      const uint64_t m = __ballot(i < N && (1.0f - s_e[i]) > 0.5f);
      if (m) first = c0 + __builtin_ctzll(m);
    }
    if (lane == 0) {
      float depth;
      if (first == N) {
        depth = INFINITY;
      } else if (first == 0) {
        depth = 0.0f;
      } else {
        float t, dlt, px, py, pz;
        sample_point(in, r, first, ox, oy, oz, dx, dy, dz, t0, step, t, dlt, px, py, pz);
        depth = t;
      }
      o.out[r] = depth;
    }
    return;
  }
  // RGB: appearance sample selection (render.py:459-470) and unbiasing terms
  float ssum;
  if (in.k < N) {
    uint32_t* key = reinterpret_cast<uint32_t*>(s_a);
    for (int i = lane; i < N; i += 64) {
      const float wv = s_w[i];
      key[i] = in.gumbel ? float_key(logf(wv) + in.gumbel[i]) : float_key(wv);
    }
    __syncthreads();
    ssum = wave_topk(key, s_w, N, in.k, o.idx + r * in.k);
  } else {
    float acc = 0.0f;
    for (int i = lane; i < N; i += 64) acc += s_w[i];
    ssum = wave_sum(acc);
  }
  if (lane == 0) {
    o.ray[4 * r] = e_last;
    o.ray[4 * r + 1] = ssum;
    o.ray[4 * r + 2] = (1.0f - e_last + 1e-8f) / (ssum + 1e-8f);  // render.py:537-546
  }
}

// C = U * sum_j w_{i_j} c_j + E_{N-1} (render.py:233-244). With colors != null
// also g = 2 (C - target) / (3R) and mse[r % nslot] += |C - target|^2 / (3R):
// adds to one address serialise at the memory side, so the rays spread over
// nslot partial sums that the caller adds up.
__global__ void __launch_bounds__(64) k_composite_rgb(const float* __restrict__ w,
                                                      const int32_t* __restrict__ idx,
                                                      const float* __restrict__ rgb,
                                                      const float* __restrict__ ray, int N, int k,
                                                      float* __restrict__ out,
                                                      const float* __restrict__ colors,
                                                      float* __restrict__ gout,
                                                      float* __restrict__ mse, int nslot,
                                                      float inv_count) {
  const int64_t r = blockIdx.x;
  const int lane = threadIdx.x;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int j = lane; j < k; j += 64) {
    const int i = idx ? idx[r * k + j] : j;
    const float wv = w[r * N + i];
    const float* c = rgb + (r * k + j) * 3;
    acc[0] += c[0] * wv;
    acc[1] += c[1] * wv;
    acc[2] += c[2] * wv;
  }
  const float u = ray[4 * r + 2], e_last = ray[4 * r];
  float sq = 0.0f;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float v = wave_sum(acc[ch]) * u + e_last * 1.0f;
    if (lane == 0) {
      out[r * 3 + ch] = v;
      if (colors) {
        const float diff = v - colors[r * 3 + ch];
        gout[r * 3 + ch] = 2.0f * diff * inv_count;
        sq += diff * diff;
      }
    }
  }
  if (colors && lane == 0) unsafeAtomicAdd(mse + r % nslot, sq * inv_count);
}

// Backward of the compositing (RGB mode) for one ray per wave, followed by
// the density scatter. With A = sum_K w_k c_k, S = sum_K w_k, E = E_{N-1},
// U = (1-E+eps)/(S+eps), a_i = sigma_i delta_i, g = dL/dC:
//   dA/da_i = [i in K] E_i c_i - sum_{k in K,k>i} w_k c_k
//   dS/da_i = [i in K] E_i     - sum_{k in K,k>i} w_k
//   dU/da_i = E/(S+eps) - U dS/da_i/(S+eps),  dE/da_i = -E
//   dL/da_i = g.(dU/da_i A + U dA/da_i - E 1);  dL/ds_i = dL/da_i delta_i sigmoid(s_i)
//   dL/dc_k = U w_k g
// sigmoid(s) is recovered from sigma = softplus(s) as 1 - exp(-sigma).
struct BwdIn {
  const float *pts, *sigma, *delta, *w, *e, *ray, *rgb, *gout;
  const int32_t* idx;
  int N, k;
};

// One workgroup of kCompBwdWaves waves per ray: wave 0 runs the scans, every
// wave then scatters 4 samples at a time.
constexpr int kCompBwdWaves = 4;

// With ds_out != null (one wave per ray) the kernel stops after the scans
// and writes dL/ds of every sample there instead of scattering it.
__global__ void __launch_bounds__(64 * kCompBwdWaves) k_composite_bwd(tfh_vm dvm, BwdIn in,
                                                                     float* __restrict__ dc,
                                                                     tfh_vm_grad dgrad, int nlc,
                                                                     float* __restrict__ ds_out) {
  // nlc consecutive copies of the line gradient: ray r adds into copy r % nlc
  dgrad.line += (int64_t)(blockIdx.x % nlc) * 3 * dvm.grid_dim * dvm.channels;
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int N = in.N, k = in.k;
  float* s_c = sm;              // (N, 3) colour of selected samples, 0 otherwise
  float* s_sel = s_c + 3 * N;   // 1 for selected samples
  float* s_ds = s_sel + N;      // dL/ds_i
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    s_c[3 * i] = 0.f;
    s_c[3 * i + 1] = 0.f;
    s_c[3 * i + 2] = 0.f;
    s_sel[i] = 0.f;
  }
  __syncthreads();
  const float e_last = in.ray[4 * r], ssum = in.ray[4 * r + 1], u = in.ray[4 * r + 2];
  const float g0 = in.gout[3 * r], g1 = in.gout[3 * r + 1], g2 = in.gout[3 * r + 2];
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const int i = in.idx ? in.idx[r * k + j] : j;
    const float* c = in.rgb + (r * k + j) * 3;
    s_c[3 * i] = c[0];
    s_c[3 * i + 1] = c[1];
    s_c[3 * i + 2] = c[2];
    s_sel[i] = 1.0f;
    const float wv = in.w[r * N + i];
    const float f = u * wv;
    dc[(r * k + j) * 3] = f * g0;
    dc[(r * k + j) * 3 + 1] = f * g1;
    dc[(r * k + j) * 3 + 2] = f * g2;
  }
  __syncthreads();
  if (wave == 0) {
    // A = sum_K w c
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int i = lane; i < N; i += 64) {
      const float wv = in.w[r * N + i];
      a0 += wv * s_c[3 * i];
      a1 += wv * s_c[3 * i + 1];
      a2 += wv * s_c[3 * i + 2];
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    const float inv_s = 1.0f / (ssum + 1e-8f);
    const float gsum = g0 + g1 + g2;
    const float gA = g0 * a0 + g1 * a1 + g2 * a2;
    // reverse chunks: exclusive suffix sums of (w c) and (w sel)
    float c0s = 0.f, c1s = 0.f, c2s = 0.f, css = 0.f;  // carries from later chunks
    const int nchunks = (N + 63) / 64;
    for (int cb = nchunks - 1; cb >= 0; --cb) {
      const int i = cb * 64 + lane;
      const bool act = i < N;
      const float wv = act ? in.w[r * N + i] : 0.f;
      const float sel = act ? s_sel[i] : 0.f;
      const float cx = act ? s_c[3 * i] : 0.f, cy = act ? s_c[3 * i + 1] : 0.f,
                  cz = act ? s_c[3 * i + 2] : 0.f;
      float v0 = wv * cx, v1 = wv * cy, v2 = wv * cz, vs = wv * sel;
      // inclusive suffix within the chunk
  #pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const float t0 = __shfl_down(v0, o, 64), t1 = __shfl_down(v1, o, 64),
                    t2 = __shfl_down(v2, o, 64), ts = __shfl_down(vs, o, 64);
        if (lane + o < 64) {
          v0 += t0;
          v1 += t1;
          v2 += t2;
          vs += ts;
        }
      }
      v0 += c0s;
      v1 += c1s;
      v2 += c2s;
      vs += css;
      float x0 = __shfl_down(v0, 1, 64), x1 = __shfl_down(v1, 1, 64), x2 = __shfl_down(v2, 1, 64),
            xs = __shfl_down(vs, 1, 64);
      if (lane == 63) {
        x0 = c0s;
        x1 = c1s;
        x2 = c2s;
        xs = css;
      }
      if (act) {
        const float ei = in.e[r * N + i];
        const float dA0 = sel * ei * cx - x0, dA1 = sel * ei * cy - x1, dA2 = sel * ei * cz - x2;
        const float dS = sel * ei - xs;
        const float dU = e_last * inv_s - u * dS * inv_s;
        const float dl_da = dU * gA + u * (g0 * dA0 + g1 * dA1 + g2 * dA2) - e_last * gsum;
        const float sig = in.sigma[r * N + i];
        s_ds[i] = dl_da * in.delta[r * N + i] * (-expm1f(-sig));
      }
      c0s = __shfl(v0, 0, 64);
      c1s = __shfl(v1, 0, 64);
      c2s = __shfl(v2, 0, 64);
      css = __shfl(vs, 0, 64);
    }
  }
  __syncthreads();
  if (ds_out) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) ds_out[r * N + i] = s_ds[i];
    return;
  }
  // density scatter: every channel of every mode receives dL/ds_i
  const int grp = lane >> 4, l16 = lane & 15;
  const int G = (int)dvm.grid_dim, C = (int)dvm.channels;
  for (int i0 = 4 * wave; i0 < N; i0 += 4 * kCompBwdWaves) {
    const int i = i0 + grp;
    const bool act = i < N;
    float xh = 0.f, yh = 0.f, zh = 0.f, g = 0.f;
    if (act) {
      const float* p = in.pts + (r * N + i) * 3;
      xh = p[0];
      yh = p[1];
      zh = p[2];
      g = s_ds[i];
    }
    ModeTaps mt[3];
    mode_taps(xh, yh, zh, G, mt);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      TapAddr a;
      tap_addrs(mt[m], m, G, a);
      // merged per tap slot only: building the cross-slot plan (tap_plan)
      // costs more here than the atomics it saves (composite_bwd 1.41 ->
      // 1.29 ms on the 4096-ray, 512-sample step)
#pragma unroll
      for (int q = 0; q < kMaxQ; ++q) {
        if (16 * q < C) {
          const int c = l16 + 16 * q;
          const bool on = act && c < C;
          float lv = 0.f, pv = 0.f;
          if (on) tap_values(dvm.plane, dvm.line, C, c, a, lv, pv);
          tap_scatter<16>(dgrad.plane, dgrad.line, C, c, a, g * pv, g * lv, on);
        }
      }
    }
  }
}

// ---- standalone API kernels ---------------------------------------------------------
__global__ void k_sample_api(RayIn in, float* __restrict__ points, float* __restrict__ ts,
                             float* __restrict__ steps) {
  const int64_t total = in.R * in.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / in.N;
    const int i = (int)(e - r * in.N);
    const float ox = in.o[3 * r], oy = in.o[3 * r + 1], oz = in.o[3 * r + 2];
    const float dx = in.d[3 * r], dy = in.d[3 * r + 1], dz = in.d[3 * r + 2];
    float t0 = 0.f, step = 0.f;
    if (!in.contraction && in.sampling == TFH_SAMPLING_SCHEDULE)
      bbox_segment(ox, oy, oz, dx, dy, dz, in.aabb, in.N, t0, step);
    float t, dlt, px, py, pz;
    sample_point(in, r, i, ox, oy, oz, dx, dy, dz, t0, step, t, dlt, px, py, pz);
    if (points) {
      points[e] = px;
      points[total + e] = py;
      points[2 * total + e] = pz;
    }
    if (ts) ts[e] = t;
    if (steps) steps[e] = dlt;
  }
}

__global__ void __launch_bounds__(64) k_segment_api(const float* __restrict__ sigmas,
                                                    const float* __restrict__ steps, int N,
                                                    float* __restrict__ p_exits,
                                                    float* __restrict__ p_term) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  float* s_a = sm;
  float* s_e = sm + N;
  float* s_w = s_e + N;
  for (int i = threadIdx.x; i < N; i += 64) s_a[i] = sigmas[r * N + i] * steps[r * N + i];
  __syncthreads();
  wave_segment_probs(s_a, s_e, s_w, N);
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += 64) {
    p_exits[r * N + i] = s_e[i];
    p_term[r * N + i] = s_w[i];
  }
}

__global__ void __launch_bounds__(64) k_topk_api(const float* __restrict__ w, int N, int k,
                                                 const float* __restrict__ gumbel,
                                                 int32_t* __restrict__ idx) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  uint32_t* key = reinterpret_cast<uint32_t*>(sm);
  float* s_w = sm + N;
  for (int i = threadIdx.x; i < N; i += 64) {
    const float wv = w[r * N + i];
    s_w[i] = wv;
    key[i] = gumbel ? float_key(logf(wv) + gumbel[i]) : float_key(wv);
  }
  __syncthreads();
  (void)wave_topk(key, s_w, N, k, idx + r * k);
}

__global__ void k_fill(float* __restrict__ p, int64_t n, float v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// ---- host orchestration -------------------------------------------------------------
struct RenderWs {
  DensityOut d;
  float* feats;  // (M, F)
  float* rgb;    // (M, 3)
  float* gout;   // (R, 3)
  float* dc;     // (M, 3)
  float* dfeat;  // (M, F)
  void* mlp_ws;
  size_t mlp_ws_bytes;
  float* act = nullptr;  // (M, 256) MLP activations kept for the backward, or null
  float* mse_part = nullptr;  // kLossSlots partial sums of the MSE
  // copies of the density / appearance line gradients for the scatters, or null
  float* dlcopy = nullptr;
  float* alcopy = nullptr;
  int dnlc = 1, anlc = 1;
  SortWs sort;           // the sorted scatters (RGB mode)
  float* ds = nullptr;   // (R*N) dL/ds of every sample, for the sorted density scatter
};

static int32_t check_cfg(const tfh_render_config* cfg, const char* who) {
  TFH_REQUIRE(cfg != nullptr, "%s: null config", who);
  TFH_REQUIRE(cfg->density_samples_per_ray >= 1 && cfg->density_samples_per_ray <= kMaxSamples,
              "%s: density_samples_per_ray %d outside [1, %d]", who, cfg->density_samples_per_ray,
              kMaxSamples);
  TFH_REQUIRE(cfg->appearance_samples_per_ray >= 1 &&
                  cfg->appearance_samples_per_ray <= cfg->density_samples_per_ray,
              "%s: appearance_samples_per_ray %d outside [1, N]", who,
              cfg->appearance_samples_per_ray);
  TFH_REQUIRE(cfg->sampling == TFH_SAMPLING_SCHEDULE || cfg->sampling == TFH_SAMPLING_EDGES,
              "%s: unknown sampling %d", who, cfg->sampling);
  TFH_REQUIRE(cfg->mode == TFH_MODE_RGB || cfg->mode == TFH_MODE_DIST_MEDIAN ||
                  cfg->mode == TFH_MODE_DIST_MEAN || cfg->mode == TFH_MODE_DIST_MEDIAN_AS_WRITTEN,
              "%s: unknown mode %d", who, cfg->mode);
  TFH_REQUIRE(cfg->mlp_precision == TFH_PREC_FP32 || cfg->mlp_precision == TFH_PREC_BF16,
              "%s: unknown precision %d", who, cfg->mlp_precision);
  return TFH_OK;
}

static size_t carve_render(Carve& cv, const tfh_render_config* cfg, const tfh_mlp* mlp, int64_t F,
                           int64_t R, RenderWs& ws) {
  const int64_t N = cfg->density_samples_per_ray, k = cfg->appearance_samples_per_ray;
  const int64_t M = R * k;
  ws.d.pts = cv.take<float>(R * N * 3);
  ws.d.ts = cv.take<float>(R * N);
  ws.d.sigma = cv.take<float>(R * N);
  ws.d.delta = cv.take<float>(R * N);
  ws.d.w = cv.take<float>(R * N);
  ws.d.e = cv.take<float>(R * N);
  ws.d.idx = cv.take<int32_t>(k < N ? R * k : 0);
  if (k >= N) ws.d.idx = nullptr;
  ws.d.ray = cv.take<float>(R * 4);
  ws.d.out = nullptr;
  ws.feats = cv.take<float>(M * F);
  ws.rgb = cv.take<float>(M * 3);
  ws.gout = cv.take<float>(R * 3);
  ws.dc = cv.take<float>(M * 3);
  ws.dfeat = cv.take<float>(M * F);
  ws.mse_part = cv.take<float>(kLossSlots);
  // one region serves the forward weight tables and, later, the backward's
  // tables + partials (the backward's need is the larger)
  ws.mlp_ws_bytes = mlp ? mlp_bwd_workspace(mlp, M, cfg->mlp_precision) : 0;
  ws.mlp_ws = cv.take<char>((int64_t)ws.mlp_ws_bytes);
  if (mlp) {
    ws.ds = cv.take<float>(R * N);
    take_sort_ws(cv, R * N, ws.sort);  // pairs of every sample (k <= N)
  }
  return cv.off;
}

// The fused training steps keep the MLP's hidden activations between its
// forward and backward (the backward then recomputes only layer 0 and the
// encoding) when the shape allows it and the workspace has the room: 1 KB per
// shaded sample past the render workspace.
static size_t act_bytes(const tfh_render_config* cfg, const tfh_mlp* mlp, int64_t R) {
  if (!mlp_keeps_activations(mlp, cfg->mlp_precision)) return 0;
  return align_up((size_t)R * cfg->appearance_samples_per_ray * 256 * sizeof(float), 256);
}

static void take_act(Carve& cv, const tfh_render_config* cfg, const tfh_mlp* mlp, int64_t R,
                     RenderWs& ws) {
  ws.act = nullptr;
  const size_t bytes = act_bytes(cfg, mlp, R);
  if (bytes == 0 || !cv.ok || cv.off + bytes > cv.cap) return;
  ws.act = cv.take<float>((int64_t)(bytes / sizeof(float)));
}

// The training steps' scatters add the factor-line gradients into copies
// (block b into copy b % n) that one kernel then sums: float atomics are
// performed at the memory side, where adds to one address serialise, and each
// line texel draws thousands of adds per step (a plane texel tens). The size
// queries reserve 8 MB (appearance) + 4 MB (density) for them; up to 256
// copies of each fit.
constexpr size_t kALineCopyBytes = 8u << 20, kDLineCopyBytes = 4u << 20;

static size_t line_copy_bytes() {
  return align_up(kALineCopyBytes, 256) + align_up(kDLineCopyBytes, 256);
}

static int line_copies(const tfh_vm& vm, size_t budget) {
  const int64_t bytes = 3 * vm.grid_dim * vm.channels * (int64_t)sizeof(float);
  return (int)std::min<int64_t>(256, (int64_t)budget / bytes);
}

static void take_line_copies(Carve& cv, const tfh_vm* density, const tfh_vm* appearance,
                             RenderWs& ws) {
  ws.dlcopy = ws.alcopy = nullptr;
  ws.dnlc = ws.anlc = 1;
  const int an = line_copies(*appearance, kALineCopyBytes);
  const int dn = line_copies(*density, kDLineCopyBytes);
  const int64_t af = an * 3 * appearance->grid_dim * appearance->channels;
  const int64_t df = dn * 3 * density->grid_dim * density->channels;
  if (an >= 2 && cv.ok && cv.off + align_up((size_t)af * 4, 256) <= cv.cap) {
    ws.alcopy = cv.take<float>(af);
    ws.anlc = an;
  }
  if (dn >= 2 && cv.ok && cv.off + align_up((size_t)df * 4, 256) <= cv.cap) {
    ws.dlcopy = cv.take<float>(df);
    ws.dnlc = dn;
  }
}

static RayIn make_rayin(const tfh_render_config* cfg, const float* aabb, const float* o,
                        const float* d, int64_t R, const float* jitter, const float* base_ts,
                        const float* base_steps, const float* gumbel) {
  RayIn in;
  in.o = o;
  in.d = d;
  in.aabb = aabb;
  in.R = R;
  in.N = cfg->density_samples_per_ray;
  in.k = cfg->appearance_samples_per_ray;
  in.mode = cfg->mode;
  in.contraction = cfg->scene_contraction ? 1 : 0;
  in.sampling = cfg->sampling;
  in.density_shift = cfg->density_shift;
  in.jitter = jitter;
  in.base_ts = base_ts;
  in.base_steps = base_steps;
  in.gumbel = gumbel;
  return in;
}

static int32_t render_forward(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                              const tfh_render_config* cfg, const RayIn& in,
                              const uint32_t* cams, float* out, const float* colors, float* mse,
                              RenderWs& ws, hipStream_t st) {
  const int64_t R = in.R;
  const int N = in.N, k = in.k;
  DensityOut dout = ws.d;
  dout.out = out;
  {
  ProfScope ps("density_pass", st);
  RayIn din = in;
  din.wide_taps = density->channels % 4 == 0 && density->channels <= 64 &&
                  (reinterpret_cast<uintptr_t>(density->plane) & 15) == 0 &&
                  (reinterpret_cast<uintptr_t>(density->line) & 15) == 0;
  hipLaunchKernelGGL(k_density_pass, dim3((unsigned)R), dim3(64), 3 * N * sizeof(float), st,
                     *density, din, dout);
  }
  TFH_TRY(check_launch("render: density pass"));
  if (cfg->mode != TFH_MODE_RGB) return TFH_OK;
  const int64_t M = R * k;
  RaySamples src{ws.d.pts, ws.d.idx, k, N};
  // the standard MLP gathers its feature rows in its own tile load (the
  // rows also land in ws.feats for the backward)
  bool fused = false;
  {
    ProfScope ps("mlp_fwd", st);
    TFH_TRY(mlp_fwd_rows_gather(mlp, *appearance, src, ws.feats, in.d, cams, M, k, ws.rgb,
                                cfg->mlp_precision, ws.mlp_ws, ws.mlp_ws_bytes, st, ws.act,
                                &fused));
  }
  if (!fused) {
    {
      ProfScope ps("appearance_gather", st);
      TFH_TRY((launch_vm_gather<RaySamples, true>(*appearance, src, M, ws.feats, st,
                                                   k >= (int)appearance->grid_dim)));
    }
    ProfScope ps("mlp_fwd", st);
    TFH_TRY(mlp_fwd_rows(mlp, ws.feats, in.d, cams, M, k, ws.rgb, cfg->mlp_precision, ws.mlp_ws,
                         ws.mlp_ws_bytes, st, ws.act));
  }
  ProfScope ps("composite_rgb", st);
  const bool part = colors && mse && ws.mse_part;
  if (part) TFH_TRY(zero_copies(ws.mse_part, kLossSlots, st));
  hipLaunchKernelGGL(k_composite_rgb, dim3((unsigned)R), dim3(64), 0, st, ws.d.w, ws.d.idx, ws.rgb,
                     ws.d.ray, N, k, out, colors, ws.gout, part ? ws.mse_part : mse,
                     part ? kLossSlots : 1, 1.0f / (3.0f * (float)R));
  TFH_TRY(check_launch("render: composite"));
  return part ? sum_copies(ws.mse_part, kLossSlots, 1, mse, st) : TFH_OK;
}

static int32_t render_backward(const tfh_vm* density, const tfh_vm* appearance,
                               const tfh_mlp* mlp, const tfh_render_config* cfg, int64_t R,
                               const float* dirs, const uint32_t* cams, const float* gout,
                               const tfh_vm_grad* dgrad, const tfh_vm_grad* agrad,
                               const tfh_mlp_grad* mgrad, RenderWs& ws, hipStream_t st) {
  const int N = cfg->density_samples_per_ray, k = cfg->appearance_samples_per_ray;
  const int64_t M = R * k;
  BwdIn bi{ws.d.pts, ws.d.sigma, ws.d.delta, ws.d.w, ws.d.e, ws.d.ray, ws.rgb, gout, ws.d.idx, N, k};
  const int64_t dln = 3 * density->grid_dim * density->channels;
  const int64_t aln = 3 * appearance->grid_dim * appearance->channels;
  // every density sample, row r*N + i
  const RaySamples dsrc{ws.d.pts, nullptr, N, N};
  const bool dsorted = ws.ds && sorted_scatter_ok(*density, R * N, ws.sort);
  // every sample shaded, one grid: one sort and one pass per scatter serve
  // both fields, after the MLP backward (dL/dsigma waits in the workspace)
  const RowGrad gsrc{ws.dfeat, (int)appearance->channels};
  const bool fuse = dsorted && ws.d.idx == nullptr && sorted_scatter_ok(*appearance, M, ws.sort) &&
                    two_field_scatter_ok(*density, *appearance, gsrc);
  {
    ProfScope ps(fuse ? "composite_bwd" : "composite_bwd_density_scatter", st);
    tfh_vm_grad dg = *dgrad;
    if (ws.dlcopy) {
      TFH_TRY(zero_copies(ws.dlcopy, ws.dnlc * dln, st));
      dg.line = ws.dlcopy;
    }
    const int nlc = ws.dlcopy ? ws.dnlc : 1;
    if (dsorted) {
      // the scans write dL/ds; the scatter then runs on the pairs sorted by cell
      hipLaunchKernelGGL(k_composite_bwd, dim3((unsigned)R), dim3(64), 5 * N * sizeof(float), st,
                         *density, bi, ws.dc, dg, nlc, ws.ds);
      TFH_TRY(check_launch("render bwd: compositing"));
      TFH_TRY(sort_pairs((int)density->grid_dim, dsrc, R * N, ws.sort, "density_sort", st));
      if (!fuse)
        TFH_TRY(scatter_sorted_pairs(*density, dsrc, R * N, ScalarGrad{ws.ds}, dg, nlc, ws.sort,
                                     "density_scatter_planes", "density_scatter_lines", st));
    } else {
      hipLaunchKernelGGL(k_composite_bwd, dim3((unsigned)R), dim3(64 * kCompBwdWaves),
                         5 * N * sizeof(float), st, *density, bi, ws.dc, dg, nlc,
                         (float*)nullptr);
      TFH_TRY(check_launch("render bwd: compositing"));
    }
    if (ws.dlcopy && !fuse) TFH_TRY(sum_copies(ws.dlcopy, ws.dnlc, dln, dgrad->line, st));
  }
  TFH_TRY(mlp_bwd_rows(mlp, ws.feats, dirs, cams, M, k, ws.dc, ws.dfeat, mgrad, cfg->mlp_precision,
                       ws.mlp_ws, ws.mlp_ws_bytes, st, ws.act, ws.act ? ws.rgb : nullptr));
  if (fuse) {
    ProfScope ps("factor_scatter", st);
    if (ws.alcopy) TFH_TRY(zero_copies(ws.alcopy, ws.anlc * aln, st));
    const FieldOut fa{*appearance, agrad->plane, ws.alcopy ? ws.alcopy : agrad->line,
                      ws.alcopy ? ws.anlc : 1};
    const FieldOut fd{*density, dgrad->plane, ws.dlcopy ? ws.dlcopy : dgrad->line,
                      ws.dlcopy ? ws.dnlc : 1};
    // the line pass in line-cell order (a second sort; adds straight into
    // the line gradients) for rays with fewer samples than the grid has
    // cells, whose consecutive samples rarely share a line texel (proposal
    // rays: 64 samples, 20.7 M memory-side transactions in ray order, line
    // pass 1.07 -> 0.69 ms + 0.08 ms sort); dense rays merge well in ray order
    // and keep it (512 samples: 1.07 ms against 1.15 + 0.11 sorted)
    const bool sorted_lines = N < (int)appearance->grid_dim;
    // with the scene contraction the far samples of a ray crowd into few
    // cells, and summing a ray's runs per line cell in registers beats the
    // four-sample merge (contracted fisheye step: lines 4.74 -> 3.78 ms); on
    // bounded scenes the runs are short and the merge is faster (uniform-512:
    // 1.08 against 1.19 ms)
    const bool ray_walk = cfg->scene_contraction != 0;
    TFH_TRY(scatter_two_fields(dsrc, M, ws.sort, gsrc, fa, ScalarGrad{ws.ds}, fd, st,
                               sorted_lines, agrad->line, dgrad->line, ray_walk));
    if (sorted_lines) return TFH_OK;
    if (ws.dlcopy) TFH_TRY(sum_copies(ws.dlcopy, ws.dnlc, dln, dgrad->line, st));
    if (ws.alcopy) TFH_TRY(sum_copies(ws.alcopy, ws.anlc, aln, agrad->line, st));
    return TFH_OK;
  }
  RaySamples src{ws.d.pts, ws.d.idx, k, N};
  ProfScope ps("appearance_scatter", st);
  const bool sorted = sorted_scatter_ok(*appearance, M, ws.sort);
  if (sorted) {
    // with every sample shaded the rows are the density scatter's, and on the
    // same grid its sorted pairs serve again
    const bool shared = dsorted && ws.d.idx == nullptr && appearance->grid_dim == density->grid_dim;
    if (!shared)
      TFH_TRY(sort_pairs((int)appearance->grid_dim, src, M, ws.sort, "appearance_sort", st));
  }
  if (!ws.alcopy) {
    if (sorted)
      return scatter_sorted_pairs(*appearance, src, M, gsrc, *agrad, 1, ws.sort,
                                  "appearance_scatter_planes", "appearance_scatter_lines", st);
    return launch_vm_scatter<RaySamples, true>(*appearance, src, M, ws.dfeat, *agrad, st);
  }
  TFH_TRY(zero_copies(ws.alcopy, ws.anlc * aln, st));
  const tfh_vm_grad ag{agrad->plane, ws.alcopy};
  if (sorted)
    TFH_TRY(scatter_sorted_pairs(*appearance, src, M, gsrc, ag, ws.anlc, ws.sort,
                                 "appearance_scatter_planes", "appearance_scatter_lines", st));
  else
    TFH_TRY((launch_vm_scatter<RaySamples, true>(*appearance, src, M, ws.dfeat, ag, st, 1,
                                                 ws.anlc)));
  return sum_copies(ws.alcopy, ws.anlc, aln, agrad->line, st);
}

static int32_t check_render_args(const tfh_vm* density, const tfh_vm* appearance,
                                 const tfh_mlp* mlp, const tfh_render_config* cfg,
                                 const float* aabb, const float* o, const float* d, int64_t R,
                                 const float* jitter, const float* base_ts,
                                 const float* base_steps, const char* who,
                                 bool samples = true) {
  TFH_TRY(check_cfg(cfg, who));
  TFH_TRY(check_vm(density, who));
  TFH_REQUIRE(R >= 0 && R < (1ll << 31), "%s: num_rays %lld out of range", who, (long long)R);
  // (an empty batch may pass the null pointers of empty arrays)
  TFH_REQUIRE(aabb && (R == 0 || (o && d)), "%s: null aabb/origins/directions", who);
  if (cfg->sampling == TFH_SAMPLING_EDGES && samples)
    TFH_REQUIRE(base_ts, "%s: interval-edge sampling needs the edges (base_ts)", who);
  else if (cfg->scene_contraction && samples)
    TFH_REQUIRE(base_ts && base_steps, "%s: contraction needs base_ts and base_steps", who);
  (void)jitter;
  if (cfg->mode == TFH_MODE_RGB) {
    TFH_TRY(check_vm(appearance, who));
    TFH_TRY(check_mlp(mlp, who));
    TFH_REQUIRE(mlp->feat_dim == 3 * appearance->channels,
                "%s: mlp feat_dim %lld != 3 * appearance channels %lld", who,
                (long long)mlp->feat_dim, (long long)appearance->channels);
  }
  return TFH_OK;
}

size_t mlp_fwd_workspace(const tfh_mlp* p, int32_t precision);

// The residuals the forward alone writes (no backward buffers): sample
// points, distances, densities, step sizes, weights and transmittances,
// top-k indices, per-ray terms, appearance features and colours, and the
// MLP's weight tables of the forward's precision.
static size_t carve_render_fwd(Carve& cv, const tfh_render_config* cfg, const tfh_mlp* mlp,
                               int64_t F, int64_t R, RenderWs& ws) {
  const int64_t N = cfg->density_samples_per_ray, k = cfg->appearance_samples_per_ray;
  const int64_t M = R * k;
  ws.d.pts = cv.take<float>(R * N * 3);
  ws.d.ts = cv.take<float>(R * N);
  ws.d.sigma = cv.take<float>(R * N);
  ws.d.delta = cv.take<float>(R * N);
  ws.d.w = cv.take<float>(R * N);
  ws.d.e = cv.take<float>(R * N);
  ws.d.idx = cv.take<int32_t>(k < N ? R * k : 0);
  if (k >= N) ws.d.idx = nullptr;
  ws.d.ray = cv.take<float>(R * 4);
  ws.d.out = nullptr;
  ws.feats = cv.take<float>(M * F);
  ws.rgb = cv.take<float>(M * 3);
  ws.gout = ws.dc = ws.dfeat = ws.mse_part = nullptr;
  ws.mlp_ws_bytes = mlp ? mlp_fwd_workspace(mlp, cfg->mlp_precision) : 0;
  ws.mlp_ws = cv.take<char>((int64_t)ws.mlp_ws_bytes);
  return cv.off;
}

// tfh_render_fwd with its residuals in `scratch` (cap bytes, owned by the
// caller of this function) instead of a workspace argument. Every kernel the
// forward launches overwrites what it reads later, so the scratch needs no
// clearing between calls.
int32_t render_fwd_scratch(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                           const tfh_render_config* cfg, const float* aabb, const float* origins,
                           const float* directions, const uint32_t* camera_indices,
                           int64_t num_rays, const float* jitter, const float* base_ts,
                           const float* base_steps, const float* gumbel, float* out,
                           void* (*scratch_of_device)(size_t cap), size_t cap, const char* who,
                           hipStream_t st) {
  TFH_TRY(check_render_args(density, appearance, mlp, cfg, aabb, origins, directions, num_rays,
                            jitter, base_ts, base_steps, who));
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(out != nullptr, "%s: null output", who);
  const bool rgb = cfg->mode == TFH_MODE_RGB;
  const int64_t F = rgb ? 3 * appearance->channels : 0;
  {
    Carve size(nullptr, 0);
    RenderWs probe;
    const size_t need = carve_render_fwd(size, cfg, rgb ? mlp : nullptr, F, num_rays, probe);
    if (need > cap)
      return fail(TFH_ERR_WORKSPACE,
                  "%s: %lld rays need %zu bytes of scratch, more than its fixed %zu bytes; "
                  "render fewer rays per call or use tfh_render_fwd with a workspace",
                  who, (long long)num_rays, need, cap);
  }
  TFH_TRY(require_device());
  void* scratch = scratch_of_device(cap);
  if (scratch == nullptr) return TFH_ERR_LAUNCH;  // the message is already recorded
  Carve cv(scratch, cap);
  RenderWs ws;
  carve_render_fwd(cv, cfg, rgb ? mlp : nullptr, F, num_rays, ws);
  if (!cv.ok) return fail(TFH_ERR_WORKSPACE, "%s: scratch carve failed", who);
  RayIn in = make_rayin(cfg, aabb, origins, directions, num_rays, jitter, base_ts, base_steps,
                        gumbel);
  return render_forward(density, appearance, mlp, cfg, in, camera_indices, out, nullptr, nullptr,
                        ws, st);
}
// tfh_train_step; with hyper != null the Adam step reads its scalars from
// device memory (see adam_launch) instead of lr_tensor, lr_mlp, decay and
// adam_count.
int32_t train_step_run(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices,
                       const float* colors, int64_t num_rays, const float* jitter,
                       const float* base_ts, const float* base_steps, const float* gumbel,
                       const tfh_vm_grad* density_grad, const tfh_vm_grad* appearance_grad,
                       const tfh_mlp_grad* mlp_grad, const tfh_train_buffers* buffers,
                       float lr_tensor, float lr_mlp, float decay, int64_t adam_count,
                       const float* hyper, float* mse_out, void* workspace,
                       size_t workspace_bytes, hipStream_t st) {
  TFH_TRY(check_render_args(density, appearance, mlp, cfg, aabb, origins, directions, num_rays,
                            jitter, base_ts, base_steps, "tfh_train_step"));
  TFH_REQUIRE(cfg->mode == TFH_MODE_RGB, "tfh_train_step: the training loss is on RGB");
  TFH_REQUIRE(num_rays >= 1, "tfh_train_step: empty batch");
  TFH_REQUIRE(colors && mse_out && buffers && density_grad && appearance_grad && mlp_grad,
              "tfh_train_step: null argument");
  TFH_REQUIRE(buffers->params && buffers->grads && buffers->exp_avg && buffers->exp_avg_sq &&
                  buffers->total > 0 && buffers->tensor_count >= 0 &&
                  buffers->tensor_count <= buffers->total,
              "tfh_train_step: bad parameter buffers");
  TFH_REQUIRE(hyper != nullptr || adam_count >= 1, "tfh_train_step: adam_count must be >= 1");
  TFH_TRY(require_device());
  Carve cv(workspace, workspace_bytes);
  RenderWs ws;
  carve_render(cv, cfg, mlp, 3 * appearance->channels, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_train_step: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  take_act(cv, cfg, mlp, num_rays, ws);
  take_line_copies(cv, density, appearance, ws);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, mse_out, (int64_t)1, 0.0f);
  TFH_TRY(check_launch("train step: clear mse"));
  RayIn in = make_rayin(cfg, aabb, origins, directions, num_rays, jitter, base_ts, base_steps,
                        gumbel);
  // the rendered colours are not needed by the caller: keep them in ws.gout's
  // neighbour (dc is free until the backward)
  TFH_TRY(render_forward(density, appearance, mlp, cfg, in, camera_indices, ws.dc, colors, mse_out,
                         ws, st));
  TFH_TRY(render_backward(density, appearance, mlp, cfg, num_rays, directions, camera_indices,
                          ws.gout, density_grad, appearance_grad, mlp_grad, ws, st));
  return adam_launch(buffers->params, buffers->grads, buffers->exp_avg, buffers->exp_avg_sq,
                     buffers->total, buffers->tensor_count, lr_tensor, lr_mlp, decay, 0.9f, 0.99f,
                     1e-8f, hyper ? 1 : adam_count, 1, st, hyper);
}

}  // namespace tfh

using namespace tfh;

extern "C" {

int32_t tfh_sample_bbox(const float* origins, const float* directions, int64_t num_rays,
                        const float* aabb, int64_t samples_per_ray, const float* jitter,
                        float* points, float* ts, float* step_sizes, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 1 && samples_per_ray <= (1 << 20),
              "tfh_sample_bbox: bad shape");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(origins && directions && aabb, "tfh_sample_bbox: null input");
  TFH_TRY(require_device());
  tfh_render_config cfg{};
  cfg.density_samples_per_ray = (int32_t)samples_per_ray;
  cfg.appearance_samples_per_ray = 1;
  RayIn in = make_rayin(&cfg, aabb, origins, directions, num_rays, jitter, nullptr, nullptr,
                        nullptr);
  int64_t blocks = std::min<int64_t>(ceil_div(num_rays * samples_per_ray, 256), 65535);
  hipLaunchKernelGGL(k_sample_api, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), in,
                     points, ts, step_sizes);
  return check_launch("tfh_sample_bbox");
}

int32_t tfh_sample_contracted(const float* origins, const float* directions, int64_t num_rays,
                              int64_t samples_per_ray, const float* base_ts,
                              const float* base_steps, const float* jitter, float* points,
                              float* ts, float* step_sizes, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 2 && samples_per_ray <= (1 << 20),
              "tfh_sample_contracted: bad shape");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(origins && directions && base_ts && base_steps,
              "tfh_sample_contracted: null input");
  TFH_TRY(require_device());
  tfh_render_config cfg{};
  cfg.density_samples_per_ray = (int32_t)samples_per_ray;
  cfg.appearance_samples_per_ray = 1;
  cfg.scene_contraction = 1;
  RayIn in = make_rayin(&cfg, nullptr, origins, directions, num_rays, jitter, base_ts, base_steps,
                        nullptr);
  int64_t blocks = std::min<int64_t>(ceil_div(num_rays * samples_per_ray, 256), 65535);
  hipLaunchKernelGGL(k_sample_api, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), in,
                     points, ts, step_sizes);
  return check_launch("tfh_sample_contracted");
}

int32_t tfh_segment_probabilities(const float* sigmas, const float* step_sizes, int64_t num_rays,
                                  int64_t samples_per_ray, float* p_exits, float* p_terminates,
                                  tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 1 && samples_per_ray <= 4096,
              "tfh_segment_probabilities: samples_per_ray %lld outside [1, 4096]",
              (long long)samples_per_ray);
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(sigmas && step_sizes && p_exits && p_terminates,
              "tfh_segment_probabilities: null pointer");
  TFH_TRY(require_device());
  hipLaunchKernelGGL(k_segment_api, dim3((unsigned)num_rays), dim3(64),
                     3 * samples_per_ray * sizeof(float), as_stream(stream), sigmas, step_sizes,
                     (int)samples_per_ray, p_exits, p_terminates);
  return check_launch("tfh_segment_probabilities");
}

int32_t tfh_topk_select(const float* p_terminates, int64_t num_rays, int64_t samples_per_ray,
                        int64_t k, const float* gumbel, int32_t* indices, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 1 && samples_per_ray <= 4096,
              "tfh_topk_select: samples_per_ray %lld outside [1, 4096]",
              (long long)samples_per_ray);
  TFH_REQUIRE(k >= 1 && k <= samples_per_ray, "tfh_topk_select: k %lld outside [1, N]",
              (long long)k);
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(p_terminates && indices, "tfh_topk_select: null pointer");
  TFH_TRY(require_device());
  hipLaunchKernelGGL(k_topk_api, dim3((unsigned)num_rays), dim3(64),
                     2 * samples_per_ray * sizeof(float), as_stream(stream), p_terminates,
                     (int)samples_per_ray, (int)k, gumbel, indices);
  return check_launch("tfh_topk_select");
}

size_t tfh_render_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                 int64_t appearance_channels, int64_t num_rays) {
  if (check_cfg(cfg, "tfh_render_workspace_size") != TFH_OK || num_rays < 0) return 0;
  if (mlp && check_mlp(mlp, "tfh_render_workspace_size") != TFH_OK) return 0;
  Carve cv(nullptr, 0);
  RenderWs ws;
  return carve_render(cv, cfg, mlp, 3 * appearance_channels, num_rays, ws);
}

size_t tfh_train_step_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                     int64_t appearance_channels, int64_t num_rays) {
  if (check_cfg(cfg, "tfh_train_step_workspace_size") != TFH_OK || num_rays < 0 || !mlp ||
      check_mlp(mlp, "tfh_train_step_workspace_size") != TFH_OK)
    return 0;
  Carve cv(nullptr, 0);
  RenderWs ws;
  return carve_render(cv, cfg, mlp, 3 * appearance_channels, num_rays, ws) +
         act_bytes(cfg, mlp, num_rays) + line_copy_bytes();
}

int32_t tfh_render_fwd(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices, int64_t num_rays,
                       const float* jitter, const float* base_ts, const float* base_steps,
                       const float* gumbel, float* out, void* workspace, size_t workspace_bytes,
                       tfh_stream_t stream) {
  TFH_TRY(check_render_args(density, appearance, mlp, cfg, aabb, origins, directions, num_rays,
                            jitter, base_ts, base_steps, "tfh_render_fwd"));
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(out != nullptr, "tfh_render_fwd: null output");
  TFH_TRY(require_device());
  const int64_t F = cfg->mode == TFH_MODE_RGB ? 3 * appearance->channels : 0;
  Carve cv(workspace, workspace_bytes);
  RenderWs ws;
  carve_render(cv, cfg, cfg->mode == TFH_MODE_RGB ? mlp : nullptr, F, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_render_fwd: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  RayIn in = make_rayin(cfg, aabb, origins, directions, num_rays, jitter, base_ts, base_steps,
                        gumbel);
  return render_forward(density, appearance, mlp, cfg, in, camera_indices, out, nullptr, nullptr,
                        ws, as_stream(stream));
}

int32_t tfh_render_bwd(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices, int64_t num_rays,
                       const float* grad_out, const tfh_vm_grad* density_grad,
                       const tfh_vm_grad* appearance_grad, const tfh_mlp_grad* mlp_grad,
                       void* workspace, size_t workspace_bytes, tfh_stream_t stream) {
  TFH_TRY(check_render_args(density, appearance, mlp, cfg, aabb, origins, directions, num_rays,
                            nullptr, nullptr, nullptr, "tfh_render_bwd", false));
  TFH_REQUIRE(cfg->mode == TFH_MODE_RGB, "tfh_render_bwd: only the RGB mode is differentiable");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(grad_out && density_grad && appearance_grad && mlp_grad,
              "tfh_render_bwd: null gradient argument");
  TFH_REQUIRE(density_grad->plane && density_grad->line && appearance_grad->plane &&
                  appearance_grad->line,
              "tfh_render_bwd: null factor gradient");
  TFH_TRY(require_device());
  Carve cv(workspace, workspace_bytes);
  RenderWs ws;
  carve_render(cv, cfg, mlp, 3 * appearance->channels, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_render_bwd: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  return render_backward(density, appearance, mlp, cfg, num_rays, directions, camera_indices,
                         grad_out, density_grad, appearance_grad, mlp_grad, ws,
                         as_stream(stream));
}

int32_t tfh_train_step(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices,
                       const float* colors, int64_t num_rays, const float* jitter,
                       const float* base_ts, const float* base_steps, const float* gumbel,
                       const tfh_vm_grad* density_grad, const tfh_vm_grad* appearance_grad,
                       const tfh_mlp_grad* mlp_grad, const tfh_train_buffers* buffers,
                       float lr_tensor, float lr_mlp, float decay, int64_t adam_count,
                       float* mse_out, void* workspace, size_t workspace_bytes,
                       tfh_stream_t stream) {
  return train_step_run(density, appearance, mlp, cfg, aabb, origins, directions, camera_indices,
                        colors, num_rays, jitter, base_ts, base_steps, gumbel, density_grad,
                        appearance_grad, mlp_grad, buffers, lr_tensor, lr_mlp, decay, adam_count,
                        nullptr, mse_out, workspace, workspace_bytes, as_stream(stream));
}

size_t tfh_train_step_proposal_workspace_size(const tfh_render_config* cfg, const tfh_mlp* mlp,
                                              int64_t appearance_channels,
                                              const tfh_proposal_config* pc,
                                              const tfh_vm* field, int64_t num_rays) {
  if (check_cfg(cfg, "tfh_train_step_proposal_workspace_size") != TFH_OK || num_rays < 0 ||
      !mlp || check_mlp(mlp, "tfh_train_step_proposal_workspace_size") != TFH_OK ||
      !pc || !field || check_vm(field, "tfh_train_step_proposal_workspace_size") != TFH_OK)
    return 0;
  Carve cv(nullptr, 0);
  RenderWs ws;
  carve_render(cv, cfg, mlp, 3 * appearance_channels, num_rays, ws);
  cv.take<float>(num_rays * (cfg->density_samples_per_ray + 1));
  return cv.off + prop_workspace_bytes(pc, field, num_rays) + act_bytes(cfg, mlp, num_rays) +
         line_copy_bytes();
}

int32_t tfh_train_step_proposal(const tfh_vm* density, const tfh_vm* appearance,
                                const tfh_mlp* mlp, const tfh_render_config* cfg,
                                const tfh_vm* field, const tfh_prop_mlp* prop_mlp,
                                const tfh_proposal_config* pc, const float* aabb,
                                const float* origins, const float* directions,
                                const uint32_t* camera_indices, const float* colors,
                                int64_t num_rays, const float* jitter, const float* base_ts,
                                const float* base_steps, const float* gumbel,
                                const tfh_vm_grad* density_grad,
                                const tfh_vm_grad* appearance_grad, const tfh_mlp_grad* mlp_grad,
                                const tfh_vm_grad* field_grad,
                                const tfh_prop_mlp_grad* prop_mlp_grad,
                                const tfh_train_buffers* buffers,
                                const tfh_train_buffers* prop_buffers, float lr_tensor,
                                float lr_mlp, float decay, int64_t adam_count, float* mse_out,
                                float* prop_loss_out, void* workspace, size_t workspace_bytes,
                                tfh_stream_t stream) {
  const char* who = "tfh_train_step_proposal";
  TFH_TRY(check_render_args(density, appearance, mlp, cfg, aabb, origins, directions, num_rays,
                            nullptr, nullptr, nullptr, who, false));
  TFH_TRY(check_prop_args(field, prop_mlp, pc, num_rays, who));
  TFH_REQUIRE(cfg->mode == TFH_MODE_RGB, "%s: the training loss is on RGB", who);
  TFH_REQUIRE(cfg->sampling == TFH_SAMPLING_EDGES, "%s: cfg->sampling must be EDGES", who);
  TFH_REQUIRE(cfg->density_samples_per_ray == pc->samples_out,
              "%s: density_samples_per_ray %d != samples_out %d", who,
              cfg->density_samples_per_ray, pc->samples_out);
  TFH_REQUIRE(cfg->scene_contraction == pc->scene_contraction,
              "%s: scene_contraction differs between the configs", who);
  TFH_REQUIRE(num_rays >= 1, "%s: empty batch", who);
  TFH_REQUIRE(colors && mse_out && prop_loss_out && buffers && prop_buffers && density_grad &&
                  appearance_grad && mlp_grad && field_grad && field_grad->plane &&
                  field_grad->line && prop_mlp_grad,
              "%s: null argument", who);
  if (pc->scene_contraction)
    TFH_REQUIRE(base_ts && base_steps, "%s: contraction needs the round-0 schedule", who);
  for (const tfh_train_buffers* b : {buffers, prop_buffers})
    TFH_REQUIRE(b->params && b->grads && b->exp_avg && b->exp_avg_sq && b->total > 0 &&
                    b->tensor_count >= 0 && b->tensor_count <= b->total,
                "%s: bad parameter buffers", who);
  TFH_REQUIRE(adam_count >= 1, "%s: adam_count must be >= 1", who);
  TFH_TRY(require_device());
  hipStream_t st = as_stream(stream);
  const int64_t R = num_rays;
  const int N = cfg->density_samples_per_ray;
  Carve cv(workspace, workspace_bytes);
  RenderWs ws;
  carve_render(cv, cfg, mlp, 3 * appearance->channels, R, ws);
  float* edges = cv.take<float>(R * (N + 1));
  PropWs pws;
  carve_prop(cv, pc, field, prop_mlp, R, pws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "%s: workspace of %zu bytes, need %zu", who, workspace_bytes,
                cv.off);
  take_act(cv, cfg, mlp, R, ws);
  take_line_copies(cv, density, appearance, ws);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, mse_out, (int64_t)1, 0.0f);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, prop_loss_out, (int64_t)1, 0.0f);
  TFH_TRY(check_launch("train step: clear losses"));
  const int prop_samples = pc->samples_round0 + pc->samples_round1;
  TFH_TRY(prop_sample(field, prop_mlp, pc, aabb, origins, directions, R, jitter, base_ts,
                      base_steps, edges, pws, st));
  const float* fjit = jitter ? jitter + R * prop_samples : nullptr;
  RayIn in = make_rayin(cfg, aabb, origins, directions, R, fjit, edges, nullptr, gumbel);
  TFH_TRY(render_forward(density, appearance, mlp, cfg, in, camera_indices, ws.dc, colors, mse_out,
                         ws, st));
  TFH_TRY(render_backward(density, appearance, mlp, cfg, R, directions, camera_indices, ws.gout,
                          density_grad, appearance_grad, mlp_grad, ws, st));
  TFH_TRY(prop_loss_bwd(field, prop_mlp, pc, R, edges, ws.d.w, field_grad, prop_mlp_grad,
                        prop_loss_out, pws, st));
  TFH_TRY(adam_launch(buffers->params, buffers->grads, buffers->exp_avg, buffers->exp_avg_sq,
                      buffers->total, buffers->tensor_count, lr_tensor, lr_mlp, decay, 0.9f,
                      0.99f, 1e-8f, adam_count, 1, st));
  return adam_launch(prop_buffers->params, prop_buffers->grads, prop_buffers->exp_avg,
                     prop_buffers->exp_avg_sq, prop_buffers->total, prop_buffers->tensor_count,
                     lr_tensor, lr_mlp, decay, 0.9f, 0.99f, 1e-8f, adam_count, 1, st);
}

}  // extern "C"
