// Scatter of VM factor gradients with far fewer memory-side atomics: the
// transpose of TensorVM.interpolate (tensor_vm.py:42-89, 226-250) for a
// batch of rows, split into
//   * planes: the (row, mode) pairs are counting-sorted by plane cell (the 2x2
//     texel block a row's bilinear taps read); a 16-lane group walks a run of
//     pairs of one cell and sums their four tap gradients in registers, so a
//     cell costs one add per texel and channel instead of one per pair;
//   * lines: rows in their own (ray) order, one 16-lane group per row, equal
//     line texels among the four rows of a wave merged before the add.
// Float atomics to device memory are performed at the memory side on this
// part and adds to one address serialise there; the transaction count is
// what bounds a scatter.
#pragma once

#include "tfh_vm_kernels.h"

namespace tfh {

// Grids up to this size use the sorted plane scatter (bins: 3 (G+1)^2).
constexpr int kSortMaxGrid = 512;
constexpr int64_t kSortMaxBins = 3ll * (kSortMaxGrid + 1) * (kSortMaxGrid + 1);
constexpr int kScanChunk = 4096;  // counts scanned by one workgroup
constexpr int64_t kSortMaxChunks = (kSortMaxBins + kScanChunk - 1) / kScanChunk;

// Code of one axis' tap pair: the lower texel when the pair is (i, i+1),
// G-1 for the pair clamped at the low edge (0, 0), G for the high edge
// (G-1, G-1) -- lin_tap never yields other equal pairs.
__device__ __forceinline__ int tap_code(const Tap& t, int g) {
  if (t.i1 == t.i0 + 1) return t.i0;
  return t.i0 == 0 ? g - 1 : g;
}

// Taps of mode m only: (v, a, b) = (x, y, z), (z, x, y), (y, z, x)
// (tensor_vm.py:50-52).
__device__ __forceinline__ void one_mode_taps(float x, float y, float z, int g, int m, Tap& v,
                                              Tap& a, Tap& b) {
  const float pv = m == 0 ? x : (m == 1 ? z : y);
  const float pa = m == 0 ? y : (m == 1 ? x : z);
  const float pb = m == 0 ? z : (m == 1 ? y : x);
  v = lin_tap(pv, g);
  a = lin_tap(pa, g);
  b = lin_tap(pb, g);
}

__device__ __forceinline__ int plane_bin(int m, const Tap& a, const Tap& b, int g) {
  return (m * (g + 1) + tap_code(a, g)) * (g + 1) + tap_code(b, g);
}

// Upstream gradient of a row's feature (mode m, channel c): a row-major
// (M, 3C) array (appearance features), or one scalar per row that every
// channel of every mode receives (the density head: sigma = softplus(sum of
// the 3C features + shift), so dL/dfeature = dL/ds for all of them).
struct RowGrad {
  const float* g;
  int C;
  __device__ __forceinline__ float at(int64_t row, int m, int c) const {
    return g[row * 3 * C + m * C + c];
  }
};

struct ScalarGrad {
  const float* g;
  __device__ __forceinline__ float at(int64_t row, int, int) const { return g[row]; }
};

struct SortWs {
  uint32_t* order = nullptr;  // (3 M) pairs row*4 + mode, sorted by bin
  int32_t* counts = nullptr;  // (bins) pairs per bin, then the running cursor
  int32_t* starts = nullptr;  // (bins) first pair of each bin
  int32_t* chunk = nullptr;   // (chunks) pairs per scan chunk, then its offset
};

// Pairs per bin. One thread per row: its three (row, mode) pairs.
template <class Src>
__global__ void __launch_bounds__(256) k_bin_count(Src src, int64_t M, int g,
                                                   int32_t* __restrict__ counts) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= M) return;
  float x, y, z;
  src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, g, mt);
#pragma unroll
  for (int m = 0; m < 3; ++m) atomicAdd(counts + plane_bin(m, mt[m].a, mt[m].b, g), 1);
}

// Exclusive scan of the bin counts in chunks of kScanChunk: each workgroup
// (256 threads, 16 counts each) scans its chunk in place and records the
// chunk total.
__global__ void __launch_bounds__(256) k_scan_chunks(int32_t* __restrict__ counts,
                                                     int32_t* __restrict__ starts, int64_t bins,
                                                     int32_t* __restrict__ chunk_total) {
  __shared__ int32_t wsum[4];
  const int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * 16;
  int32_t v[16];
  int32_t run = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    v[i] = (base + i < bins) ? counts[base + i] : 0;
    run += v[i];
  }
  // exclusive scan of the per-thread totals over the workgroup
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t incl = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int32_t before = 0;
  for (int i = 0; i < w; ++i) before += wsum[i];
  int32_t off = before + incl - run;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (base + i < bins) starts[base + i] = off;
    off += v[i];
  }
  if (threadIdx.x == 255) chunk_total[blockIdx.x] = before + incl;
}

// Chunk offsets (one workgroup), then every bin's start += its chunk's offset
// and the cursor (counts) = start.
__global__ void __launch_bounds__(1024) k_scan_offsets(int32_t* __restrict__ chunk_total,
                                                       int64_t chunks) {
  __shared__ int32_t wsum[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i = threadIdx.x;
  const int32_t v = i < chunks ? chunk_total[i] : 0;
  int32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int32_t before = 0;
  for (int j = 0; j < w; ++j) before += wsum[j];
  if (i < chunks) chunk_total[i] = before + incl - v;
}

__global__ void __launch_bounds__(256) k_scan_finish(int32_t* __restrict__ counts,
                                                     int32_t* __restrict__ starts, int64_t bins,
                                                     const int32_t* __restrict__ chunk_off) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= bins) return;
  const int32_t s = starts[i] + chunk_off[i / kScanChunk];
  starts[i] = s;
  counts[i] = s;
}

// Placement: pair (row, mode) goes to the next slot of its bin.
template <class Src>
__global__ void __launch_bounds__(256) k_bin_place(Src src, int64_t M, int g,
                                                   int32_t* __restrict__ cursor,
                                                   uint32_t* __restrict__ order) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= M) return;
  float x, y, z;
  src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, g, mt);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int32_t pos = atomicAdd(cursor + plane_bin(m, mt[m].a, mt[m].b, g), 1);
    order[pos] = (uint32_t)s * 4u + (uint32_t)m;
  }
}

// Plane gradients from the sorted pairs: dL/dplane[m][texel][c] +=
// w_texel * line_m(c) * gout[row][m*C + c]. Each 16-lane group owns kPairsPerGroup
// consecutive sorted pairs; lane l owns channels l, l+16, ... . The group
// walks its pairs in batches of 16 (one pair's index and point loaded per
// lane, then broadcast), sums the four taps' gradients of the current bin in
// registers and adds them when the bin changes and at the end.
constexpr int kPairsPerGroup = 64;

template <class Src, class GS>
__global__ void __launch_bounds__(256) k_plane_scatter_sorted(tfh_vm vm, Src src, int64_t pairs,
                                                              const uint32_t* __restrict__ order,
                                                              GS gsrc,
                                                              float* __restrict__ gplane) {
  constexpr int LPS = 16, NQ = 64 / LPS;
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  const int l = threadIdx.x & 15;
  const int gbase = threadIdx.x & ~15;  // first lane of the group in the block
  const int64_t first = ((int64_t)blockIdx.x * (256 / LPS) + threadIdx.x / LPS) * kPairsPerGroup;
  if (first >= pairs) return;  // whole group: no cross-group exchange below
  const int64_t last = min(first + (int64_t)kPairsPerGroup, pairs);
  float acc[4][NQ];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[t][q] = 0.0f;
  int cur = -1;  // bin being summed
  int tex[4] = {0, 0, 0, 0};
  auto flush = [&]() {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = l + LPS * q;
        if (c < C && acc[t][q] != 0.0f) unsafeAtomicAdd(gplane + (size_t)tex[t] * C + c, acc[t][q]);
        acc[t][q] = 0.0f;
      }
  };
  for (int64_t b0 = first; b0 < last; b0 += LPS) {
    // lane l fetches pair b0 + l: its packed index and the row's point
    const int64_t e = b0 + l;
    uint32_t pk = 0xffffffffu;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (e < last) {
      pk = order[e];
      src.get((int64_t)(pk >> 2), px, py, pz);
    }
    const int nb = (int)min((int64_t)LPS, last - b0);
    for (int j = 0; j < nb; ++j) {
      const uint32_t p = __shfl(pk, gbase + j, 64);
      const float x = __shfl(px, gbase + j, 64), y = __shfl(py, gbase + j, 64),
                  z = __shfl(pz, gbase + j, 64);
      const int64_t row = (int64_t)(p >> 2);
      const int m = (int)(p & 3u);
      Tap tv, ta, tb;
      one_mode_taps(x, y, z, G, m, tv, ta, tb);
      const int bin = plane_bin(m, ta, tb, G);
      if (bin != cur) {
        if (cur >= 0) flush();
        cur = bin;
        const int base = m * G;
        tex[0] = (base + ta.i0) * G + tb.i0;
        tex[1] = (base + ta.i0) * G + tb.i1;
        tex[2] = (base + ta.i1) * G + tb.i0;
        tex[3] = (base + ta.i1) * G + tb.i1;
      }
      const float w[4] = {ta.w0 * tb.w0, ta.w0 * tb.w1, ta.w1 * tb.w0, ta.w1 * tb.w1};
      const float* lrow0 = vm.line + (size_t)(m * G + tv.i0) * C;
      const float* lrow1 = vm.line + (size_t)(m * G + tv.i1) * C;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int c = l + LPS * q;
        if (LPS * q < C && c < C) {
          const float lv = tv.w0 * lrow0[c] + tv.w1 * lrow1[c];
          const float gl = gsrc.at(row, m, c) * lv;
#pragma unroll
          for (int t = 0; t < 4; ++t) acc[t][q] += w[t] * gl;
        }
      }
    }
  }
  if (cur >= 0) flush();
}

// Line gradients in row order: dL/dline[m][v][c] += w_v * plane_m(c) *
// gout[row][m*C + c]; 16-lane groups, equal line texels of the wave's four
// rows (in either tap slot) added once. grad_line holds nlc copies; block b
// adds into copy b % nlc.
template <class Src, class GS>
__global__ void __launch_bounds__(kGatherThreads) k_line_scatter(tfh_vm vm, Src src, int64_t M,
                                                                 GS gsrc,
                                                                 float* __restrict__ gline,
                                                                 int nlc) {
  constexpr int LPS = 16, SPB = kGatherThreads / LPS, NQ = 64 / LPS;
  const int grp = threadIdx.x / LPS, l = threadIdx.x % LPS;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  if (s0 >= M) return;
  const int64_t s = s0 + grp;
  const bool active = s < M;
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  gline += (int64_t)(blockIdx.x % nlc) * 3 * G * C;
  float x = 0.f, y = 0.f, z = 0.f;
  if (active) src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, G, mt);
  const int wg = grp & 3;  // group within the wave
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    TapAddr a;
    tap_addrs(mt[m], m, G, a);
    // merge plan of the two line slots over the wave's four groups
    int k0 = active ? a.line0 : -1 - 2 * wg, k1 = active ? a.line1 : -2 - 2 * wg;
    const bool fold = active && k1 == k0;  // clamped at the grid edge
    if (fold) k1 = INT_MIN + wg;
    int src0[3], src1[3];
    bool lead0 = active, lead1 = active && !fold;
#pragma unroll
    for (int oi = 0; oi < 3; ++oi) {
      const int o = LPS * (oi + 1);
      const bool lower = (wg ^ (oi + 1)) < wg;
      const int p0 = __shfl_xor(k0, o, 64), p1 = __shfl_xor(k1, o, 64);
      src0[oi] = p0 == k0 ? 0 : (p1 == k0 ? 1 : -1);
      src1[oi] = p0 == k1 ? 0 : (p1 == k1 ? 1 : -1);
      if (src0[oi] >= 0 && lower) lead0 = false;
      if (src1[oi] >= 0 && lower) lead1 = false;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = l + LPS * q;
      if (LPS * q < C) {  // wave-uniform: every lane joins the exchange
        const bool on = active && c < C;
        float v0 = 0.f, v1 = 0.f;
        if (on) {
          float lv, pv;
          tap_values(vm.plane, vm.line, C, c, a, lv, pv);
          const float gp = gsrc.at(s, m, c) * pv;
          v0 = a.wl0 * gp;
          v1 = a.wl1 * gp;
          if (fold) {
            v0 += v1;
            v1 = 0.f;
          }
        }
        float acc0 = v0, acc1 = v1;
#pragma unroll
        for (int oi = 0; oi < 3; ++oi) {
          const int o = LPS * (oi + 1);
          const float q0 = __shfl_xor(v0, o, 64), q1 = __shfl_xor(v1, o, 64);
          acc0 += src0[oi] == 0 ? q0 : (src0[oi] == 1 ? q1 : 0.0f);
          acc1 += src1[oi] == 0 ? q0 : (src1[oi] == 1 ? q1 : 0.0f);
        }
        if (lead0 && acc0 != 0.0f) unsafeAtomicAdd(gline + (size_t)k0 * C + c, acc0);
        if (lead1 && acc1 != 0.0f) unsafeAtomicAdd(gline + (size_t)k1 * C + c, acc1);
      }
    }
  }
}

// Bytes of the sort's workspace for M rows.
inline size_t sort_ws_bytes(int64_t M) {
  return align_up((size_t)(3 * M) * 4, 256) + 2 * align_up((size_t)kSortMaxBins * 4, 256) +
         align_up((size_t)kSortMaxChunks * 4, 256);
}

inline void take_sort_ws(Carve& cv, int64_t M, SortWs& s) {
  s.order = cv.take<uint32_t>(3 * M);
  s.counts = cv.take<int32_t>(kSortMaxBins);
  s.starts = cv.take<int32_t>(kSortMaxBins);
  s.chunk = cv.take<int32_t>(kSortMaxChunks);
}

// Whether a field takes the sorted scatter: 9..64 channels (16-lane
// groups), a grid within the bin budget, and a row count whose packed
// (row, mode) index fits 32 bits.
inline bool sorted_scatter_ok(const tfh_vm& vm, int64_t M, const SortWs& s) {
  return s.order && vm.channels > 8 && vm.channels <= 64 && vm.grid_dim <= kSortMaxGrid &&
         M < (1ll << 30);
}

// Sorts the (row, mode) pairs of M rows by plane cell of a grid of size g.
template <class Src>
inline int32_t sort_pairs(int g, const Src& src, int64_t M, const SortWs& s, const char* stage,
                          hipStream_t st) {
  const int64_t bins = 3ll * (g + 1) * (g + 1);
  const int64_t chunks = ceil_div(bins, kScanChunk);
  const int64_t rb = ceil_div(M, 256);
  ProfScope ps(stage, st);
  TFH_TRY(zero_copies(reinterpret_cast<float*>(s.counts), bins, st));  // all-zero bits
  hipLaunchKernelGGL(k_bin_count<Src>, dim3((unsigned)rb), dim3(256), 0, st, src, M, g, s.counts);
  hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)chunks), dim3(256), 0, st, s.counts, s.starts,
                     bins, s.chunk);
  hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, st, s.chunk, chunks);
  hipLaunchKernelGGL(k_scan_finish, dim3((unsigned)ceil_div(bins, 256)), dim3(256), 0, st,
                     s.counts, s.starts, bins, s.chunk);
  hipLaunchKernelGGL(k_bin_place<Src>, dim3((unsigned)rb), dim3(256), 0, st, src, M, g, s.counts,
                     s.order);
  return check_launch("vm scatter: pair sort");
}

// The scatter of a field's feature gradients (gsrc) for M rows whose pairs
// sort_pairs has ordered: planes into grad.plane, lines into grad.line (nlc
// copies). stages: profile names of the plane and line passes.
template <class Src, class GS>
inline int32_t scatter_sorted_pairs(const tfh_vm& vm, const Src& src, int64_t M, const GS& gsrc,
                                    const tfh_vm_grad& grad, int nlc, const SortWs& s,
                                    const char* plane_stage, const char* line_stage,
                                    hipStream_t st) {
  if (M <= 0) return TFH_OK;
  {
    ProfScope ps(plane_stage, st);
    const int64_t pairs = 3 * M;
    const int64_t groups = ceil_div(pairs, kPairsPerGroup);
    hipLaunchKernelGGL((k_plane_scatter_sorted<Src, GS>), dim3((unsigned)ceil_div(groups, 16)),
                       dim3(256), 0, st, vm, src, pairs, s.order, gsrc, grad.plane);
    TFH_TRY(check_launch("vm scatter: planes"));
  }
  ProfScope ps(line_stage, st);
  hipLaunchKernelGGL((k_line_scatter<Src, GS>), dim3((unsigned)ceil_div(M, kGatherThreads / 16)),
                     dim3(kGatherThreads), 0, st, vm, src, M, gsrc, grad.line, nlc);
  return check_launch("vm scatter: lines");
}

}  // namespace tfh
