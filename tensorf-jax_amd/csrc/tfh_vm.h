// Device-side VM-decomposition gather / scatter shared by the kernels of
// tfh_vm.hip and tfh_render.hip.
//
// Mapping: one sample per aligned group of LPS lanes (16, or 8 for fields of
// at most 8 channels); lane `l` of the group owns channels c = l + LPS*q, so
// each tap of a sample is one contiguous C*4-byte read (or atomic) per group
// in the channel-last layout plane[m][row][col][c], line[m][v][c].
#pragma once

#include <climits>

#include "tfh_common.h"

namespace tfh {

constexpr int kMaxQ = 4;  // channels <= 64

struct TapAddr {
  // texel indices (without the channel) of the 2 line and 4 plane taps
  int line0, line1;
  int p00, p01, p10, p11;
  float wl0, wl1;
  float w00, w01, w10, w11;
};

__device__ __forceinline__ void tap_addrs(const ModeTaps& t, int m, int g, TapAddr& a) {
  a.line0 = m * g + t.v.i0;
  a.line1 = m * g + t.v.i1;
  const int base = m * g;
  a.p00 = (base + t.a.i0) * g + t.b.i0;
  a.p01 = (base + t.a.i0) * g + t.b.i1;
  a.p10 = (base + t.a.i1) * g + t.b.i0;
  a.p11 = (base + t.a.i1) * g + t.b.i1;
  a.wl0 = t.v.w0;
  a.wl1 = t.v.w1;
  // (wa*wb) products in the order JAX's map_coordinates forms them
  a.w00 = t.a.w0 * t.b.w0;
  a.w01 = t.a.w0 * t.b.w1;
  a.w10 = t.a.w1 * t.b.w0;
  a.w11 = t.a.w1 * t.b.w1;
}

// line and plane values of channel c for one mode.
__device__ __forceinline__ void tap_values(const float* __restrict__ plane,
                                           const float* __restrict__ line, int C, int c,
                                           const TapAddr& a, float& lv, float& pv) {
  const float l0 = line[(size_t)a.line0 * C + c];
  const float l1 = line[(size_t)a.line1 * C + c];
  lv = a.wl0 * l0 + a.wl1 * l1;
  const float t00 = plane[(size_t)a.p00 * C + c];
  const float t01 = plane[(size_t)a.p01 * C + c];
  const float t10 = plane[(size_t)a.p10 * C + c];
  const float t11 = plane[(size_t)a.p11 * C + c];
  pv = ((a.w00 * t00 + a.w01 * t01) + a.w10 * t10) + a.w11 * t11;
}

// line and plane values of channels c4 .. c4+3 for one mode: one 16-byte
// load per tap (C and c4 multiples of 4); each channel's arithmetic is that
// of tap_values.
__device__ __forceinline__ void tap_values4(const float* __restrict__ plane,
                                            const float* __restrict__ line, int C, int c4,
                                            const TapAddr& a, float4& lv, float4& pv) {
  const float4 l0 = *reinterpret_cast<const float4*>(line + (size_t)a.line0 * C + c4);
  const float4 l1 = *reinterpret_cast<const float4*>(line + (size_t)a.line1 * C + c4);
  const float4 t00 = *reinterpret_cast<const float4*>(plane + (size_t)a.p00 * C + c4);
  const float4 t01 = *reinterpret_cast<const float4*>(plane + (size_t)a.p01 * C + c4);
  const float4 t10 = *reinterpret_cast<const float4*>(plane + (size_t)a.p10 * C + c4);
  const float4 t11 = *reinterpret_cast<const float4*>(plane + (size_t)a.p11 * C + c4);
  lv.x = a.wl0 * l0.x + a.wl1 * l1.x;
  lv.y = a.wl0 * l0.y + a.wl1 * l1.y;
  lv.z = a.wl0 * l0.z + a.wl1 * l1.z;
  lv.w = a.wl0 * l0.w + a.wl1 * l1.w;
  pv.x = ((a.w00 * t00.x + a.w01 * t01.x) + a.w10 * t10.x) + a.w11 * t11.x;
  pv.y = ((a.w00 * t00.y + a.w01 * t01.y) + a.w10 * t10.y) + a.w11 * t11.y;
  pv.z = ((a.w00 * t00.z + a.w01 * t01.z) + a.w10 * t10.z) + a.w11 * t11.z;
  pv.w = ((a.w00 * t00.w + a.w01 * t01.w) + a.w10 * t10.w) + a.w11 * t11.w;
}

// Adds `val` at dst[key*C + c] once per distinct `key` among the 64/LPS
// groups of LPS lanes of the wave (all lanes of the wave must call it
// together): each group sums the values of the groups with its key in the
// same tap slot and only the lowest such group issues the atomic. Lanes with
// active == false contribute nothing. Used for 8-lane groups, where matching
// every slot of 7 partner groups (tap_plan) costs more than it saves.
template <int LPS = 16>
__device__ __forceinline__ void merged_atomic_add(float* __restrict__ dst, int C, int c, int key,
                                                  float val, bool active) {
  const int grp = lane_id() / LPS;
  const int k = active ? key : -1 - grp;
  const float v = active ? val : 0.0f;
  float acc = v;
  bool leader = active;
#pragma unroll
  for (int o = LPS; o < 64; o += LPS) {
    const int pk = __shfl_xor(k, o, 64);
    const float pv = __shfl_xor(v, o, 64);
    if (pk == k) {
      acc += pv;
      if ((grp ^ (o / LPS)) < grp) leader = false;
    }
  }
  if (leader && acc != 0.0f) unsafeAtomicAdd(dst + (size_t)k * C + c, acc);
}

// Scatter of one mode, same-slot merge: dL/dline = gl * (per-tap weight),
// dL/dplane = gp * weight.
template <int LPS = 16>
__device__ __forceinline__ void tap_scatter(float* __restrict__ gplane, float* __restrict__ gline,
                                            int C, int c, const TapAddr& a, float gl, float gp,
                                            bool active) {
  merged_atomic_add<LPS>(gline, C, c, a.line0, a.wl0 * gl, active);
  merged_atomic_add<LPS>(gline, C, c, a.line1, a.wl1 * gl, active);
  merged_atomic_add<LPS>(gplane, C, c, a.p00, a.w00 * gp, active);
  merged_atomic_add<LPS>(gplane, C, c, a.p01, a.w01 * gp, active);
  merged_atomic_add<LPS>(gplane, C, c, a.p10, a.w10 * gp, active);
  merged_atomic_add<LPS>(gplane, C, c, a.p11, a.w11 * gp, active);
}

// Cross-slot merge plan of one mode's six taps (slots 0-1 line, 2-5 plane)
// over the 64/LPS groups of a wave, built from the texel keys alone and
// reused for every channel slot. Consecutive samples of a ray share texels
// in different slots (the upper line tap of one sample is often the lower
// tap of the next), so a key is matched against every slot of every partner
// group; per distinct key the slot of the lowest group issues one atomic
// carrying the sum of all its copies. Duplicate taps of one sample (clamped
// at the grid edge) fold into their first slot.
template <int LPS>
struct TapPlan {
  static constexpr int NO = 64 / LPS - 1;  // partner groups
  int key[6];
  int src[NO > 0 ? NO : 1][6];  // partner slot holding this slot's key, or -1
  int fold[6];                  // slot this one folds into, or -1
  bool lead[6];
};

template <int LPS>
__device__ __forceinline__ void tap_plan(const TapAddr& a, bool active, TapPlan<LPS>& P) {
  constexpr int NO = TapPlan<LPS>::NO;
  const int grp = lane_id() / LPS;
  int k[6] = {a.line0, a.line1, a.p00, a.p01, a.p10, a.p11};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    P.fold[s] = -1;
    if (!active) k[s] = -1 - (grp * 6 + s);
  }
  if (k[1] == k[0]) P.fold[1] = 0;
#pragma unroll
  for (int s = 3; s < 6; ++s)
#pragma unroll
    for (int t = 2; t < s; ++t)
      if (P.fold[s] < 0 && P.fold[t] < 0 && k[s] == k[t]) P.fold[s] = t;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    if (P.fold[s] >= 0) k[s] = INT_MIN + grp * 6 + s;  // matches no other slot
    P.key[s] = k[s];
    P.lead[s] = active && P.fold[s] < 0;
  }
#pragma unroll
  for (int oi = 0; oi < NO; ++oi) {
    const int o = LPS * (oi + 1);
    const bool lower = (grp ^ (oi + 1)) < grp;
    int pk[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) pk[t] = __shfl_xor(k[t], o, 64);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      int src = -1;
      const int t0 = s < 2 ? 0 : 2, t1 = s < 2 ? 2 : 6;
#pragma unroll
      for (int t = t0; t < t1; ++t)
        if (pk[t] == k[s]) src = t;
      P.src[oi][s] = src;
      if (src >= 0 && lower) P.lead[s] = false;
    }
  }
}

// Scatter of one mode and channel c with a plan: dL/dline = gl * (line tap
// weight), dL/dplane = gp * (plane tap weight). All lanes of the wave call it
// together; on == false contributes nothing.
template <int LPS>
__device__ __forceinline__ void tap_scatter_plan(float* __restrict__ gplane,
                                                 float* __restrict__ gline, int C, int c,
                                                 const TapAddr& a, const TapPlan<LPS>& P,
                                                 float gl, float gp, bool on) {
  constexpr int NO = TapPlan<LPS>::NO;
  float v[6];
  v[0] = on ? a.wl0 * gl : 0.0f;
  v[1] = on ? a.wl1 * gl : 0.0f;
  v[2] = on ? a.w00 * gp : 0.0f;
  v[3] = on ? a.w01 * gp : 0.0f;
  v[4] = on ? a.w10 * gp : 0.0f;
  v[5] = on ? a.w11 * gp : 0.0f;
#pragma unroll
  for (int s = 1; s < 6; ++s) {
    if (P.fold[s] >= 0) {
#pragma unroll
      for (int t = 0; t < s; ++t)
        if (P.fold[s] == t) v[t] += v[s];
      v[s] = 0.0f;
    }
  }
  float acc[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) acc[s] = v[s];
#pragma unroll
  for (int oi = 0; oi < NO; ++oi) {
    const int o = LPS * (oi + 1);
    float pv[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) pv[t] = __shfl_xor(v[t], o, 64);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int src = P.src[oi][s];
      float add = 0.0f;
#pragma unroll
      for (int t = (s < 2 ? 0 : 2); t < (s < 2 ? 2 : 6); ++t) add = src == t ? pv[t] : add;
      acc[s] += add;
    }
  }
#pragma unroll
  for (int s = 0; s < 6; ++s)
    if (P.lead[s] && acc[s] != 0.0f)
      unsafeAtomicAdd((s < 2 ? gline : gplane) + (size_t)P.key[s] * C + c, acc[s]);
}

}  // namespace tfh
