// Device-side VM-decomposition gather / scatter shared by the kernels of
// tfh_vm.hip and tfh_render.hip.
//
// Mapping: one sample per aligned group of LPS lanes (16, or 8 for fields of
// at most 8 channels); lane `l` of the group owns channels c = l + LPS*q, so
// each tap of a sample is one contiguous C*4-byte read (or atomic) per group
// in the channel-last layout plane[m][row][col][c], line[m][v][c].
#pragma once

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

// Adds `val` at dst[key*C + c] once per distinct `key` among the 64/LPS
// groups of LPS lanes of the wave (all lanes of the wave must call it
// together): each group sums the values of the groups with its key and only
// the lowest such group issues the atomic. Lanes with active == false
// contribute nothing.
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

// Scatter of one mode: dL/dline = gl * (per-tap weight), dL/dplane = gp * weight.
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

}  // namespace tfh
