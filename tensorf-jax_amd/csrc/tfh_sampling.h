// Ray sampling and per-ray compositing helpers shared by the field renderer
// (tfh_render.hip) and the proposal sampler (tfh_prop.hip).
#pragma once

#include "tfh_vm_kernels.h"

namespace tfh {

constexpr int kMaxSamples = 1024;

// ray_segment_from_bounding_box (render.py:399-434) + the stratified step
// (render.py:369): returns t_min and the step size (t_max - t_min) / N.
__device__ __forceinline__ void bbox_segment(float ox, float oy, float oz, float dx, float dy,
                                             float dz, const float* __restrict__ aabb, int n,
                                             float& t0, float& step) {
  const float o[3] = {ox, oy, oz};
  const float d[3] = {dx, dy, dz};
  float tmin_axis_max = -INFINITY, tmax_axis_min = INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float den = d[a] + 1e-8f;                 // render.py:413-415
    const float t_lo = (aabb[a] - o[a]) / den;
    const float t_hi = (aabb[3 + a] - o[a]) / den;
    tmin_axis_max = fmaxf(tmin_axis_max, fminf(t_lo, t_hi));  // render.py:418
    tmax_axis_min = fminf(tmax_axis_min, fmaxf(t_lo, t_hi));  // render.py:419
  }
  const float t_min = fmaxf(0.0f, tmin_axis_max);   // render.py:423
  const float t_max = tmax_axis_min;                // render.py:424
  const float t_max_clipped = fmaxf(t_max, t_min + 1e-3f);  // render.py:425
  const bool valid = t_min < t_max;                 // render.py:429
  t0 = valid ? t_min : 0.0f;                        // render.py:432-433
  const float t1 = valid ? t_max_clipped : 1e-3f;
  step = (t1 - t0) / (float)n;                      // render.py:369
}

// L-inf scene contraction p -> (2 - 1/|p|) p/|p| outside the unit cube
// (render.py:168-173).
__device__ __forceinline__ void contract(float& px, float& py, float& pz) {
  const float nrm = fmaxf(fmaxf(fabsf(px), fabsf(py)), fabsf(pz));  // render.py:170
  if (!(nrm <= 1.0f)) {                                              // render.py:171
    const float f = 2.0f - 1.0f / nrm;
    px = f * px / nrm;
    py = f * py / nrm;
    pz = f * pz / nrm;
  }
}

// ((p - a0) / (a1 - a0) - 0.5) * 2 (render.py:193-195)
__device__ __forceinline__ float normalise(float p, float a0, float a1) {
  return ((p - a0) / (a1 - a0) - 0.5f) * 2.0f;
}

struct RayIn {
  const float* o;
  const float* d;
  const float* aabb;  // (2,3) = 6 floats
  int64_t R;
  int N, k, mode, contraction, sampling;
  float density_shift;      // softplus(sum + shift), +10 in the reference
  const float* jitter;      // (N) bbox, (R,N) contraction or edges, or null
  const float* base_ts;     // (N) contraction schedule, or (R, N+1) interval edges
  const float* base_steps;  // (N)
  const float* gumbel;      // (N) or null
  int wide_taps = 0;        // density gather with 16-byte taps (set at launch)
};

// Sample i of ray r: distance t, step size delta and the world (or
// contracted) point (render.py:122-186; interval edges: t = e_i + u (e_{i+1} - e_i)).
__device__ __forceinline__ void sample_point(const RayIn& in, int64_t r, int i, float ox, float oy,
                                             float oz, float dx, float dy, float dz, float t0,
                                             float step, float& t, float& dlt, float& px,
                                             float& py, float& pz) {
  if (in.sampling == TFH_SAMPLING_EDGES) {
    const float* e = in.base_ts + r * (in.N + 1);
    const float e0 = e[i], e1 = e[i + 1];
    const float u = in.jitter ? in.jitter[r * in.N + i] : 0.5f;
    dlt = e1 - e0;
    t = e0 + u * dlt;
    px = ox + t * dx;
    py = oy + t * dy;
    pz = oz + t * dz;
    if (in.contraction) contract(px, py, pz);
  } else if (!in.contraction) {
    const float u = in.jitter ? in.jitter[i] : 0.0f;
    t = ((float)i + u) * step;                     // render.py:372-380
    t = t0 + t;                                    // render.py:381
    dlt = step;                                    // render.py:186
    px = ox + dx * t;                              // render.py:384-386
    py = oy + dy * t;
    pz = oz + dz * t;
  } else {
    const float bs = in.base_steps[i];
    const float u = in.jitter ? in.jitter[r * in.N + i] : 0.0f;
    t = in.base_ts[i] + bs * u;                    // render.py:161
    dlt = bs;                                      // render.py:154-155
    px = ox + t * dx;                              // render.py:164-167
    py = oy + t * dy;
    pz = oz + t * dz;
    contract(px, py, pz);
  }
}

// ---- per-ray scans (one wave) ---------------------------------------------------
// p_exits E_i = exp(sum_{j<=i} -a_j), p_terminates w_i = (1-exp(-a_i)) E_{i-1}
// (render.py:321-341), from a (LDS) into e / w (LDS).
__device__ inline void wave_segment_probs(const float* s_a, float* s_e, float* s_w, int n) {
  const int lane = lane_id();
  float carry = 0.0f;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int i = c0 + lane;
    const bool act = i < n;
    const float neg = act ? -s_a[i] : 0.0f;
    const float incl = wave_incl_scan(neg) + carry;
    float prev = __shfl_up(incl, 1, 64);
    if (lane == 0) prev = carry;
    if (act) {
      s_e[i] = expf(incl);
      s_w[i] = (1.0f - expf(neg)) * expf(prev);
    }
    carry = __shfl(incl, 63, 64);
  }
}

// Exclusive prefix sum of v[0..n) into out[0..n] (out[n] = total), one wave.
__device__ inline void wave_excl_prefix(const float* v, float* out, int n) {
  const int lane = lane_id();
  float carry = 0.0f;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int i = c0 + lane;
    const float x = i < n ? v[i] : 0.0f;
    const float incl = wave_incl_scan(x) + carry;
    if (i < n) out[i + 1] = incl;
    carry = __shfl(incl, 63, 64);
  }
  if (lane == 0) out[0] = 0.0f;
}

// VJP of the segment probabilities w.r.t. a_i = sigma_i delta_i for
// upstream gradients gw (of w, may be null) and ge (of E, may be null):
//   dL/da_i = gw_i E_i - sum_{j>i} gw_j w_j - sum_{j>=i} ge_j E_j
// written to out[i]. w, e: the forward's values (LDS or global).
__device__ inline void wave_segment_vjp(const float* gw, const float* ge, const float* w,
                                        const float* e, float* out, int n) {
  const int lane = lane_id();
  float carry_w = 0.0f, carry_e = 0.0f;  // sums over later chunks
  const int nchunks = (n + 63) / 64;
  for (int cb = nchunks - 1; cb >= 0; --cb) {
    const int i = cb * 64 + lane;
    const bool act = i < n;
    const float ei = act ? e[i] : 0.0f;
    const float gwi = (act && gw) ? gw[i] : 0.0f;
    const float gei = (act && ge) ? ge[i] : 0.0f;
    float vw = gwi * (act ? w[i] : 0.0f), ve = gei * ei;
    // inclusive suffix sums within the chunk
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const float tw = __shfl_down(vw, o, 64), te = __shfl_down(ve, o, 64);
      if (lane + o < 64) {
        vw += tw;
        ve += te;
      }
    }
    vw += carry_w;
    ve += carry_e;
    float xw = __shfl_down(vw, 1, 64);  // exclusive suffix of gw w
    if (lane == 63) xw = carry_w;
    if (act) out[i] = gwi * ei - xw - ve;
    carry_w = __shfl(vw, 0, 64);
    carry_e = __shfl(ve, 0, 64);
  }
}

// Number of entries of the sorted array a[0..n) that are < x (lower bound)
// or <= x (upper bound).
__device__ __forceinline__ int lower_bound(const float* a, int n, float x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int upper_bound(const float* a, int n, float x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

}  // namespace tfh
