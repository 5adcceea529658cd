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

// Grids up to this size use the sorted plane scatter.
constexpr int kSortMaxGrid = 512;

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
  // channels c4 .. c4+3 (16-byte aligned rows: C % 4 == 0)
  __device__ __forceinline__ float4 at4(int64_t row, int m, int c4) const {
    return *reinterpret_cast<const float4*>(g + row * 3 * C + m * C + c4);
  }
  bool aligned16() const { return (reinterpret_cast<uintptr_t>(g) & 15) == 0 && C % 4 == 0; }
};

struct ScalarGrad {
  const float* g;
  __device__ __forceinline__ float at(int64_t row, int, int) const { return g[row]; }
  __device__ __forceinline__ float4 at4(int64_t row, int, int) const {
    const float v = g[row];
    return make_float4(v, v, v, v);
  }
  bool aligned16() const { return true; }
};

// Two-level sort of the (row, mode) pairs by plane cell. Cells are grouped
// into tiles of kTile x kTile cells per mode. (A) every workgroup ranks its
// pairs within their tiles in LDS and reserves each tile's share with one
// global atomic; (B) one workgroup scans the tile counts; (C) every pair
// moves to its tile's segment; (D) every chunk of up to kSortChunk pairs of
// one tile is counting-sorted by cell in LDS in place. Equal cells are then
// adjacent within a chunk (a tile larger than a chunk repeats a cell in each
// of its chunks). Global atomics: one per (workgroup, tile) instead of two
// per pair.
constexpr int kTile = 16;
constexpr int kSortMaxTiles = 3 * ((kSortMaxGrid + 1 + kTile - 1) / kTile) *
                              ((kSortMaxGrid + 1 + kTile - 1) / kTile);
constexpr int kRowsPerRankGroup = 1024;  // rows of one workgroup of pass A
constexpr int kSortChunk = 4096;         // pairs sorted by cell in one workgroup

struct SortWs {
  uint32_t* order = nullptr;    // (3 M) pairs row*4 + mode, sorted by cell
  uint16_t* tile_of = nullptr;  // (3 M) tile of pair row*3 + mode
  uint32_t* rank_of = nullptr;  // (3 M) position of the pair within its tile
  int32_t* tile_count = nullptr;  // (tiles) pairs per tile
  int32_t* tile_start = nullptr;  // (tiles + 1) first pair of each tile
  int32_t* chunk_start = nullptr; // (tiles + 1) first chunk of each tile
};

__device__ __forceinline__ int tiles_per_axis(int g) { return (g + 1 + kTile - 1) / kTile; }

// Tile and cell (0..kTile^2-1) of mode m's plane taps.
__device__ __forceinline__ void pair_tile(int m, const Tap& a, const Tap& b, int g, int& tile,
                                          int& cell) {
  const int ca = tap_code(a, g), cb = tap_code(b, g), nt = tiles_per_axis(g);
  tile = (m * nt + ca / kTile) * nt + cb / kTile;
  cell = (ca % kTile) * kTile + cb % kTile;
}

// (A) Per workgroup of kRowsPerRankGroup rows: LDS ranks within tiles, one
// global reservation per tile present. LINE: the bins are instead the line
// cells of each mode (m (g+1) + the line tap code, 3 (g+1) bins), with no
// cell pass after.
template <class Src, bool LINE = false>
__global__ void __launch_bounds__(256) k_tile_rank(Src src, int64_t M, int g,
                                                   int32_t* __restrict__ tile_count,
                                                   uint16_t* __restrict__ tile_of,
                                                   uint32_t* __restrict__ rank_of) {
  constexpr int RPT = kRowsPerRankGroup / 256;
  __shared__ int32_t cnt[kSortMaxTiles];
  const int nt = tiles_per_axis(g), tiles = LINE ? 3 * (g + 1) : 3 * nt * nt;
  for (int t = threadIdx.x; t < tiles; t += 256) cnt[t] = 0;
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerRankGroup;
  int tl[RPT][3], rk[RPT][3];
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int64_t s = r0 + j * 256 + threadIdx.x;
#pragma unroll
    for (int m = 0; m < 3; ++m) tl[j][m] = -1;
    if (s < M) {
      float x, y, z;
      src.get(s, x, y, z);
      ModeTaps mt[3];
      mode_taps(x, y, z, g, mt);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        if constexpr (LINE) {
          tl[j][m] = m * (g + 1) + tap_code(mt[m].v, g);
        } else {
          int cell;
          pair_tile(m, mt[m].a, mt[m].b, g, tl[j][m], cell);
        }
        rk[j][m] = atomicAdd(&cnt[tl[j][m]], 1);
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < tiles; t += 256) {
    const int c = cnt[t];
    if (c > 0) cnt[t] = atomicAdd(tile_count + t, c);  // this workgroup's base in the tile
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int64_t s = r0 + j * 256 + threadIdx.x;
    if (s < M) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        tile_of[3 * s + m] = (uint16_t)tl[j][m];
        rank_of[3 * s + m] = (uint32_t)(cnt[tl[j][m]] + rk[j][m]);
      }
    }
  }
}

// (B) Exclusive scans of the tile counts and of the chunks per tile (one
// workgroup; tiles <= kSortMaxTiles).
__global__ void __launch_bounds__(1024) k_tile_scan(const int32_t* __restrict__ tile_count,
                                                    int tiles, int32_t* __restrict__ tile_start,
                                                    int32_t* __restrict__ chunk_start) {
  __shared__ int32_t wsum[2][16];
  __shared__ int32_t carry[2];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (int t0 = 0; t0 < tiles; t0 += 1024) {
    const int t = t0 + threadIdx.x;
    const int32_t c = t < tiles ? tile_count[t] : 0;
    const int32_t ch = (c + kSortChunk - 1) / kSortChunk;
    int32_t ic = c, ih = ch;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t a = __shfl_up(ic, o, 64), b = __shfl_up(ih, o, 64);
      if (lane >= o) {
        ic += a;
        ih += b;
      }
    }
    if (lane == 63) {
      wsum[0][w] = ic;
      wsum[1][w] = ih;
    }
    __syncthreads();
    int32_t bc = carry[0], bh = carry[1];
    for (int j = 0; j < w; ++j) {
      bc += wsum[0][j];
      bh += wsum[1][j];
    }
    if (t < tiles) {
      tile_start[t] = bc + ic - c;
      chunk_start[t] = bh + ih - ch;
    }
    __syncthreads();
    if (threadIdx.x == 1023) {
      carry[0] = bc + ic;
      carry[1] = bh + ih;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    tile_start[tiles] = carry[0];
    chunk_start[tiles] = carry[1];
  }
}

// (C) Every pair to its tile's segment.
__global__ void __launch_bounds__(256) k_tile_place(int64_t pairs,
                                                    const uint16_t* __restrict__ tile_of,
                                                    const uint32_t* __restrict__ rank_of,
                                                    const int32_t* __restrict__ tile_start,
                                                    uint32_t* __restrict__ order) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= pairs) return;
  const int64_t row = p / 3;
  const int m = (int)(p - 3 * row);
  order[tile_start[tile_of[p]] + rank_of[p]] = (uint32_t)row * 4u + (uint32_t)m;
}

// (D) One chunk of a tile: counting sort by cell in LDS, written back in
// place. Workgroup w finds its tile by binary search of chunk_start;
// workgroups past the last chunk return.
template <class Src>
__global__ void __launch_bounds__(256) k_chunk_sort(Src src, int g, int tiles,
                                                    const int32_t* __restrict__ tile_start,
                                                    const int32_t* __restrict__ chunk_start,
                                                    uint32_t* __restrict__ order) {
  constexpr int PT = kSortChunk / 256, NC = kTile * kTile;
  __shared__ int32_t hist[NC];
  __shared__ int32_t wsum[4];
  const int w = blockIdx.x;
  if (w >= chunk_start[tiles]) return;
  int lo = 0, hi = tiles - 1;  // last tile with chunk_start <= w
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunk_start[mid] <= w) lo = mid; else hi = mid - 1;
  }
  const int t = lo;
  const int64_t base = (int64_t)tile_start[t] + (int64_t)(w - chunk_start[t]) * kSortChunk;
  const int len = (int)min((int64_t)kSortChunk, (int64_t)tile_start[t + 1] - base);
  for (int c = threadIdx.x; c < NC; c += 256) hist[c] = 0;
  __syncthreads();
  uint32_t pk[PT];
  int cl[PT], rk[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const int i = j * 256 + threadIdx.x;
    cl[j] = -1;
    if (i < len) {
      pk[j] = order[base + i];
      float x, y, z;
      src.get((int64_t)(pk[j] >> 2), x, y, z);
      Tap tv, ta, tb;
      one_mode_taps(x, y, z, g, (int)(pk[j] & 3u), tv, ta, tb);
      int tile;
      pair_tile((int)(pk[j] & 3u), ta, tb, g, tile, cl[j]);
      rk[j] = atomicAdd(&hist[cl[j]], 1);
    }
  }
  __syncthreads();
  // exclusive scan of the 256 cell counts (one per thread)
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int32_t c0 = hist[threadIdx.x];
  int32_t inc = c0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t a = __shfl_up(inc, o, 64);
    if (lane >= o) inc += a;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  int32_t before = 0;
  for (int j = 0; j < wv; ++j) before += wsum[j];
  __syncthreads();
  hist[threadIdx.x] = before + inc - c0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PT; ++j)
    if (cl[j] >= 0) order[base + hist[cl[j]] + rk[j]] = pk[j];
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

// Channel 16q + l of a 16-lane group whose lane j holds channels
// 4j .. 4j+3 (v): the sums go back to one channel per lane before they are
// added, so that each atomic instruction of a group covers 64 contiguous
// bytes (one memory-side transaction) as in the scalar kernel; with the
// quads the lanes' adds were 16 bytes apart, three transactions per
// instruction (the plane pass took 1.54 ms instead of 0.69). All 16 lanes of
// the group must call it (gl0: the group's first lane in the wave).
__device__ __forceinline__ float quad_to_lane(const float (&v)[4], int q, int l, int gl0) {
  const int from = gl0 + 4 * q + (l >> 2);
  const float t0 = __shfl(v[0], from, 64), t1 = __shfl(v[1], from, 64);
  const float t2 = __shfl(v[2], from, 64), t3 = __shfl(v[3], from, 64);
  const int u = l & 3;
  return u == 0 ? t0 : (u == 1 ? t1 : (u == 2 ? t2 : t3));
}

// k_plane_scatter_sorted with 16-byte loads: lane l of a 16-lane group of 64
// pairs owns channels 4l .. 4l+3 (C % 4 == 0, 16 < C <= 64). The same sums
// per channel, a third of the loads for 48 channels.
template <class Src, class GS, int LPS>
__global__ void __launch_bounds__(256) k_plane_scatter_sorted4(tfh_vm vm, Src src, int64_t pairs,
                                                               const uint32_t* __restrict__ order,
                                                               GS gsrc,
                                                               float* __restrict__ gplane) {
  static_assert(LPS == 16, "one channel per lane for the adds");
  const int G = (int)vm.grid_dim, C = (int)vm.channels;
  const int l = threadIdx.x % LPS;
  const int gbase = threadIdx.x - l;  // first lane of the group in the block
  const int64_t first = ((int64_t)blockIdx.x * (256 / LPS) + threadIdx.x / LPS) * kPairsPerGroup;
  if (first >= pairs) return;  // whole group: no cross-group exchange below
  const int64_t last = min(first + (int64_t)kPairsPerGroup, pairs);
  const int c4 = 4 * l;
  const bool lane_on = c4 < C;
  float4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  int cur = -1;  // bin being summed
  int tex[4] = {0, 0, 0, 0};
  const int gl0 = gbase & 63;
  auto flush = [&]() {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float v[4] = {acc[t].x, acc[t].y, acc[t].z, acc[t].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (16 * q < C) {  // group-uniform
          const float val = quad_to_lane(v, q, l, gl0);
          const int c = 16 * q + l;
          if (c < C && val != 0.0f) unsafeAtomicAdd(gplane + (size_t)tex[t] * C + c, val);
        }
      }
      acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
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
      if (lane_on) {
        const float w[4] = {ta.w0 * tb.w0, ta.w0 * tb.w1, ta.w1 * tb.w0, ta.w1 * tb.w1};
        const float4 l0 = *reinterpret_cast<const float4*>(vm.line + (size_t)(m * G + tv.i0) * C + c4);
        const float4 l1 = *reinterpret_cast<const float4*>(vm.line + (size_t)(m * G + tv.i1) * C + c4);
        const float4 g4 = gsrc.at4(row, m, c4);
        const float4 gl = make_float4(g4.x * (tv.w0 * l0.x + tv.w1 * l1.x),
                                      g4.y * (tv.w0 * l0.y + tv.w1 * l1.y),
                                      g4.z * (tv.w0 * l0.z + tv.w1 * l1.z),
                                      g4.w * (tv.w0 * l0.w + tv.w1 * l1.w));
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[t].x += w[t] * gl.x;
          acc[t].y += w[t] * gl.y;
          acc[t].z += w[t] * gl.z;
          acc[t].w += w[t] * gl.w;
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

// ---- both fields in one pass -------------------------------------------------
// When every sample is shaded and the density (CD <= 16 channels, upstream
// gradient dL/dsigma per row) and appearance (16 < CA <= 64, C % 4 == 0,
// upstream gradient the feature row) factors share the grid, the pairs, the
// taps and the line merge plan are the same for both: one pass computes them
// once and scatters both fields.
struct FieldOut {
  tfh_vm vm;
  float* gplane;
  float* gline;  // nlc copies of the line gradient
  int nlc;
};

// The plane pass of both fields: k_plane_scatter_sorted4 for the appearance
// channels (lane l: 4l .. 4l+3) beside the scalar density channels (lane l:
// channel l), one bin change flushing both.
template <class Src>
__global__ void __launch_bounds__(256) k_plane_scatter_two(Src src, int64_t pairs,
                                                           const uint32_t* __restrict__ order,
                                                           RowGrad ga, FieldOut fa, ScalarGrad gd,
                                                           FieldOut fd) {
  constexpr int LPS = 16;
  const int G = (int)fa.vm.grid_dim, CA = (int)fa.vm.channels, CD = (int)fd.vm.channels;
  const int l = threadIdx.x & 15;
  const int gbase = threadIdx.x & ~15;
  const int gl0 = gbase & 63;
  const int64_t first = ((int64_t)blockIdx.x * (256 / LPS) + threadIdx.x / LPS) * kPairsPerGroup;
  if (first >= pairs) return;  // whole group: no cross-group exchange below
  const int64_t last = min(first + (int64_t)kPairsPerGroup, pairs);
  const int c4 = 4 * l;
  const bool a_on = c4 < CA, d_on = l < CD;
  float4 acc[4];
  float accd[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    accd[t] = 0.0f;
  }
  int cur = -1;
  int tex[4] = {0, 0, 0, 0};
  auto flush = [&]() {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float v[4] = {acc[t].x, acc[t].y, acc[t].z, acc[t].w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (16 * q < CA) {  // group-uniform
          const float val = quad_to_lane(v, q, l, gl0);
          const int c = 16 * q + l;
          if (c < CA && val != 0.0f) unsafeAtomicAdd(fa.gplane + (size_t)tex[t] * CA + c, val);
        }
      }
      if (d_on && accd[t] != 0.0f) unsafeAtomicAdd(fd.gplane + (size_t)tex[t] * CD + l, accd[t]);
      acc[t] = make_float4(0.f, 0.f, 0.f, 0.f);
      accd[t] = 0.0f;
    }
  };
  for (int64_t b0 = first; b0 < last; b0 += LPS) {
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
      // both fields' line taps and upstream gradients requested before the
      // flush: one memory round trip per pair instead of one per field
      float4 l0 = make_float4(0.f, 0.f, 0.f, 0.f), l1 = l0, g4 = l0;
      float dl0 = 0.f, dl1 = 0.f, dg = 0.f;
      if (a_on) {
        l0 = *reinterpret_cast<const float4*>(fa.vm.line + (size_t)(m * G + tv.i0) * CA + c4);
        l1 = *reinterpret_cast<const float4*>(fa.vm.line + (size_t)(m * G + tv.i1) * CA + c4);
        g4 = ga.at4(row, m, c4);
      }
      if (d_on) {
        dl0 = fd.vm.line[(size_t)(m * G + tv.i0) * CD + l];
        dl1 = fd.vm.line[(size_t)(m * G + tv.i1) * CD + l];
        dg = gd.at(row, m, l);
      }
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
      if (a_on) {
        const float4 gl = make_float4(g4.x * (tv.w0 * l0.x + tv.w1 * l1.x),
                                      g4.y * (tv.w0 * l0.y + tv.w1 * l1.y),
                                      g4.z * (tv.w0 * l0.z + tv.w1 * l1.z),
                                      g4.w * (tv.w0 * l0.w + tv.w1 * l1.w));
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          acc[t].x += w[t] * gl.x;
          acc[t].y += w[t] * gl.y;
          acc[t].z += w[t] * gl.z;
          acc[t].w += w[t] * gl.w;
        }
      }
      if (d_on) {
        const float gl = dg * (tv.w0 * dl0 + tv.w1 * dl1);
#pragma unroll
        for (int t = 0; t < 4; ++t) accd[t] += w[t] * gl;
      }
    }
  }
  if (cur >= 0) flush();
}

// The line pass of both fields: k_line_scatter's groups, taps and merge plan
// once; channel slots 0 .. NQA-1 (16 channels each) of the appearance, then
// the density's (CD <= 16).
template <class Src>
__global__ void __launch_bounds__(kGatherThreads) k_line_scatter_two(Src src, int64_t M,
                                                                     RowGrad ga, FieldOut fa,
                                                                     ScalarGrad gd, FieldOut fd) {
  constexpr int LPS = 16, SPB = kGatherThreads / LPS;
  const int grp = threadIdx.x / LPS, l = threadIdx.x % LPS;
  const int64_t s0 = (int64_t)blockIdx.x * SPB;
  if (s0 >= M) return;
  const int64_t s = s0 + grp;
  const bool active = s < M;
  const int G = (int)fa.vm.grid_dim, CA = (int)fa.vm.channels, CD = (int)fd.vm.channels;
  float* glA = fa.gline + (int64_t)(blockIdx.x % fa.nlc) * 3 * G * CA;
  float* glD = fd.gline + (int64_t)(blockIdx.x % fd.nlc) * 3 * G * CD;
  float x = 0.f, y = 0.f, z = 0.f;
  if (active) src.get(s, x, y, z);
  ModeTaps mt[3];
  mode_taps(x, y, z, G, mt);
  const int wg = grp & 3;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    TapAddr a;
    tap_addrs(mt[m], m, G, a);
    int k0 = active ? a.line0 : -1 - 2 * wg, k1 = active ? a.line1 : -2 - 2 * wg;
    const bool fold = active && k1 == k0;
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
    // slots 0..3 appearance (while 16 q < CA), slot 4 density
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const bool dens = q == 4;
      const int C = dens ? CD : CA;
      if (!dens && LPS * q >= CA) continue;  // wave-uniform
      const int c = dens ? l : l + LPS * q;
      const bool on = active && c < C;
      float v0 = 0.f, v1 = 0.f;
      if (on) {
        float lv, pv;
        tap_values(dens ? fd.vm.plane : fa.vm.plane, dens ? fd.vm.line : fa.vm.line, C, c, a, lv,
                   pv);
        const float gp = (dens ? gd.at(s, m, c) : ga.at(s, m, c)) * pv;
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
      float* gl = dens ? glD : glA;
      if (lead0 && acc0 != 0.0f) unsafeAtomicAdd(gl + (size_t)k0 * C + c, acc0);
      if (lead1 && acc1 != 0.0f) unsafeAtomicAdd(gl + (size_t)k1 * C + c, acc1);
    }
  }
}

// The line pass of both fields over pairs ordered by line cell (sort_lines):
// a 16-lane group walks 64 pairs, sums the two line taps' gradients of a
// cell's run in registers (appearance: 16-byte plane taps, lane l channels
// 4l .. 4l+3; density: lane l channel l) and adds them when the cell
// changes. A line texel collects the adds of every sample on its plane
// slice, so in ray order it draws one memory-side transaction per sample
// and 16 channels wherever consecutive samples of a ray move to a new
// texel; in cell order it draws a few per run of thousands of pairs. The
// adds go straight to the line gradients (no copies).
static_assert(3 * (kSortMaxGrid + 1) <= kSortMaxTiles, "line bins fit the sort's tile table");

// RAY: the pairs in ray order instead, mode by mode (pair e = (row e % M,
// mode e / M), M rows): a group walks 64 consecutive samples of one ray and
// mode, and a run is a stretch of the ray inside one line cell; the adds go
// to fa's / fd's line copies (block b into copy b % nlc), since every ray
// crosses the same few line texels.
template <class Src, bool RAY = false>
__global__ void __launch_bounds__(256) k_line_walk_two(Src src, int64_t pairs,
                                                       const uint32_t* __restrict__ order,
                                                       RowGrad ga, FieldOut fa, ScalarGrad gd,
                                                       FieldOut fd, int64_t M = 0) {
  constexpr int LPS = 16;
  const int G = (int)fa.vm.grid_dim, CA = (int)fa.vm.channels, CD = (int)fd.vm.channels;
  float* glA = fa.gline + (int64_t)(blockIdx.x % fa.nlc) * 3 * G * CA;
  float* glD = fd.gline + (int64_t)(blockIdx.x % fd.nlc) * 3 * G * CD;
  const int l = threadIdx.x & 15;
  const int gbase = threadIdx.x & ~15;
  const int gl0 = gbase & 63;
  const int64_t first = ((int64_t)blockIdx.x * (256 / LPS) + threadIdx.x / LPS) * kPairsPerGroup;
  if (first >= pairs) return;  // whole group: no cross-group exchange below
  const int64_t last = min(first + (int64_t)kPairsPerGroup, pairs);
  const int c4 = 4 * l;
  const bool a_on = c4 < CA, d_on = l < CD;
  float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f), a1 = a0;
  float d0 = 0.0f, d1 = 0.0f;
  int cur = -1, cell0 = 0, cell1 = 0;
  auto flush = [&]() {
    const float v0[4] = {a0.x, a0.y, a0.z, a0.w}, v1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (16 * q < CA) {  // group-uniform
        const float x0 = quad_to_lane(v0, q, l, gl0), x1 = quad_to_lane(v1, q, l, gl0);
        const int c = 16 * q + l;
        if (c < CA) {
          if (x0 != 0.0f) unsafeAtomicAdd(glA + (size_t)cell0 * CA + c, x0);
          if (x1 != 0.0f) unsafeAtomicAdd(glA + (size_t)cell1 * CA + c, x1);
        }
      }
    }
    if (d_on) {
      if (d0 != 0.0f) unsafeAtomicAdd(glD + (size_t)cell0 * CD + l, d0);
      if (d1 != 0.0f) unsafeAtomicAdd(glD + (size_t)cell1 * CD + l, d1);
    }
    a0 = a1 = make_float4(0.f, 0.f, 0.f, 0.f);
    d0 = d1 = 0.0f;
  };
  for (int64_t b0 = first; b0 < last; b0 += LPS) {
    const int64_t e = b0 + l;
    uint32_t pk = 0xffffffffu;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (e < last) {
      if constexpr (RAY) {
        const int64_t m = e / M;
        pk = (uint32_t)(((e - m * M) << 2) | m);
      } else {
        pk = order[e];
      }
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
      // the four plane taps of mode m (tap_addrs' order and weights) and the
      // upstream gradients of both fields, requested before the flush
      const int base = m * G;
      const int64_t p00 = (int64_t)(base + ta.i0) * G + tb.i0, p01 = (int64_t)(base + ta.i0) * G + tb.i1;
      const int64_t p10 = (int64_t)(base + ta.i1) * G + tb.i0, p11 = (int64_t)(base + ta.i1) * G + tb.i1;
      float4 t00 = make_float4(0.f, 0.f, 0.f, 0.f), t01 = t00, t10 = t00, t11 = t00, g4 = t00;
      float d00 = 0.f, d01 = 0.f, d10 = 0.f, d11 = 0.f, dg = 0.f;
      if (a_on) {
        const float* pl = fa.vm.plane + c4;
        t00 = *reinterpret_cast<const float4*>(pl + p00 * CA);
        t01 = *reinterpret_cast<const float4*>(pl + p01 * CA);
        t10 = *reinterpret_cast<const float4*>(pl + p10 * CA);
        t11 = *reinterpret_cast<const float4*>(pl + p11 * CA);
        g4 = ga.at4(row, m, c4);
      }
      if (d_on) {
        const float* pl = fd.vm.plane + l;
        d00 = pl[p00 * CD];
        d01 = pl[p01 * CD];
        d10 = pl[p10 * CD];
        d11 = pl[p11 * CD];
        dg = gd.at(row, m, l);
      }
      const int bin = m * (G + 1) + tap_code(tv, G);
      if (bin != cur) {
        if (cur >= 0) flush();
        cur = bin;
        cell0 = m * G + tv.i0;
        cell1 = m * G + tv.i1;
      }
      const float w00 = ta.w0 * tb.w0, w01 = ta.w0 * tb.w1, w10 = ta.w1 * tb.w0, w11 = ta.w1 * tb.w1;
      if (a_on) {
        const float gx = g4.x * (((w00 * t00.x + w01 * t01.x) + w10 * t10.x) + w11 * t11.x);
        const float gy = g4.y * (((w00 * t00.y + w01 * t01.y) + w10 * t10.y) + w11 * t11.y);
        const float gz = g4.z * (((w00 * t00.z + w01 * t01.z) + w10 * t10.z) + w11 * t11.z);
        const float gw = g4.w * (((w00 * t00.w + w01 * t01.w) + w10 * t10.w) + w11 * t11.w);
        a0.x += tv.w0 * gx; a0.y += tv.w0 * gy; a0.z += tv.w0 * gz; a0.w += tv.w0 * gw;
        a1.x += tv.w1 * gx; a1.y += tv.w1 * gy; a1.z += tv.w1 * gz; a1.w += tv.w1 * gw;
      }
      if (d_on) {
        const float pv = ((w00 * d00 + w01 * d01) + w10 * d10) + w11 * d11;
        const float gp = dg * pv;
        d0 += tv.w0 * gp;
        d1 += tv.w1 * gp;
      }
    }
  }
  if (cur >= 0) flush();
}

// Orders the (row, mode) pairs by line cell (3 (g+1) bins: rank, scan,
// place; the bin is the whole key).
template <class Src>
inline int32_t sort_lines(int g, const Src& src, int64_t M, const SortWs& s, hipStream_t st) {
  const int bins = 3 * (g + 1);
  const int64_t pairs = 3 * M;
  ProfScope ps("line_sort", st);
  TFH_TRY(zero_copies(reinterpret_cast<float*>(s.tile_count), bins, st));  // all-zero bits
  hipLaunchKernelGGL((k_tile_rank<Src, true>), dim3((unsigned)ceil_div(M, kRowsPerRankGroup)),
                     dim3(256), 0, st, src, M, g, s.tile_count, s.tile_of, s.rank_of);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, s.tile_count, bins, s.tile_start,
                     s.chunk_start);
  hipLaunchKernelGGL(k_tile_place, dim3((unsigned)ceil_div(pairs, 256)), dim3(256), 0, st, pairs,
                     s.tile_of, s.rank_of, s.tile_start, s.order);
  return check_launch("vm scatter: line sort");
}

// Whether the two fields can take the one-pass scatter.
inline bool two_field_scatter_ok(const tfh_vm& den, const tfh_vm& app, const RowGrad& ga) {
  return den.grid_dim == app.grid_dim && den.channels >= 1 && den.channels <= 16 &&
         app.channels > 16 && app.channels <= 64 && app.channels % 4 == 0 &&
         (reinterpret_cast<uintptr_t>(app.plane) & 15) == 0 &&
         (reinterpret_cast<uintptr_t>(app.line) & 15) == 0 && ga.aligned16();
}

// Both fields' plane and line passes over the pairs sort_pairs ordered.
// lines_sorted: the line pass on pairs re-sorted by line cell (the sort's
// arrays are reused after the plane pass), adding straight into lines_a /
// lines_d (the fields' line gradients); else a ray-order pass into fa's
// and fd's line copies: ray_walk, runs of a ray's samples in one line cell
// summed in registers (k_line_walk_two in ray order; pays where many
// consecutive samples share a cell), or k_line_scatter_two's merge of each
// wave's four samples.
template <class Src>
inline int32_t scatter_two_fields(const Src& src, int64_t M, const SortWs& s, const RowGrad& ga,
                                  const FieldOut& fa, const ScalarGrad& gd, const FieldOut& fd,
                                  hipStream_t st, bool lines_sorted = false,
                                  float* lines_a = nullptr, float* lines_d = nullptr,
                                  bool ray_walk = false) {
  if (M <= 0) return TFH_OK;
  const int64_t pairs = 3 * M;
  const int64_t groups = ceil_div(pairs, kPairsPerGroup);
  {
    ProfScope ps("factor_scatter_planes", st);
    hipLaunchKernelGGL((k_plane_scatter_two<Src>), dim3((unsigned)ceil_div(groups, 16)), dim3(256),
                       0, st, src, pairs, s.order, ga, fa, gd, fd);
    TFH_TRY(check_launch("vm scatter: planes of both fields"));
  }
  if (lines_sorted) {
    TFH_TRY(sort_lines((int)fa.vm.grid_dim, src, M, s, st));
    ProfScope ps("factor_scatter_lines", st);
    FieldOut la = fa, ld = fd;
    la.gline = lines_a;
    la.nlc = 1;
    ld.gline = lines_d;
    ld.nlc = 1;
    hipLaunchKernelGGL((k_line_walk_two<Src>), dim3((unsigned)ceil_div(groups, 16)), dim3(256), 0,
                       st, src, pairs, s.order, ga, la, gd, ld);
    return check_launch("vm scatter: sorted lines of both fields");
  }
  ProfScope ps("factor_scatter_lines", st);
  if (ray_walk)
    hipLaunchKernelGGL((k_line_walk_two<Src, true>), dim3((unsigned)ceil_div(groups, 16)),
                       dim3(256), 0, st, src, pairs, nullptr, ga, fa, gd, fd, M);
  else
    hipLaunchKernelGGL((k_line_scatter_two<Src>), dim3((unsigned)ceil_div(M, kGatherThreads / 16)),
                       dim3(kGatherThreads), 0, st, src, M, ga, fa, gd, fd);
  return check_launch("vm scatter: lines of both fields");
}

// Bytes of the sort's workspace for up to M rows.
inline size_t sort_ws_bytes(int64_t M) {
  return align_up((size_t)(3 * M) * 4, 256) * 2 + align_up((size_t)(3 * M) * 2, 256) +
         3 * align_up((size_t)(kSortMaxTiles + 1) * 4, 256);
}

inline void take_sort_ws(Carve& cv, int64_t M, SortWs& s) {
  s.order = cv.take<uint32_t>(3 * M);
  s.rank_of = cv.take<uint32_t>(3 * M);
  s.tile_of = cv.take<uint16_t>(3 * M);
  s.tile_count = cv.take<int32_t>(kSortMaxTiles + 1);
  s.tile_start = cv.take<int32_t>(kSortMaxTiles + 1);
  s.chunk_start = cv.take<int32_t>(kSortMaxTiles + 1);
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
  const int nt = (g + 1 + kTile - 1) / kTile, tiles = 3 * nt * nt;
  const int64_t pairs = 3 * M;
  ProfScope ps(stage, st);
  TFH_TRY(zero_copies(reinterpret_cast<float*>(s.tile_count), tiles, st));  // all-zero bits
  hipLaunchKernelGGL(k_tile_rank<Src>, dim3((unsigned)ceil_div(M, kRowsPerRankGroup)), dim3(256),
                     0, st, src, M, g, s.tile_count, s.tile_of, s.rank_of);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, s.tile_count, tiles, s.tile_start,
                     s.chunk_start);
  hipLaunchKernelGGL(k_tile_place, dim3((unsigned)ceil_div(pairs, 256)), dim3(256), 0, st, pairs,
                     s.tile_of, s.rank_of, s.tile_start, s.order);
  // at most one partial chunk per tile beyond pairs / kSortChunk full ones
  const int64_t chunk_bound = pairs / kSortChunk + tiles;
  hipLaunchKernelGGL(k_chunk_sort<Src>, dim3((unsigned)chunk_bound), dim3(256), 0, st, src, g,
                     tiles, s.tile_start, s.chunk_start, s.order);
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
  // 16-byte loads in the plane pass when the channels and every tap source
  // allow; only for fields of more than 16 channels: a 16-channel field's
  // 16-lane groups read a tap in one instruction either way, and narrower
  // groups would split its adds over several transactions
  const bool wide = vm.channels % 4 == 0 && vm.channels <= 64 &&
                    (reinterpret_cast<uintptr_t>(vm.plane) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(vm.line) & 15) == 0 && gsrc.aligned16();
  {
    ProfScope ps(plane_stage, st);
    const int64_t pairs = 3 * M;
    const int64_t groups = ceil_div(pairs, kPairsPerGroup);
    const int C = (int)vm.channels;
    if (wide && C > 16) {
      hipLaunchKernelGGL((k_plane_scatter_sorted4<Src, GS, 16>),
                         dim3((unsigned)ceil_div(groups, 16)), dim3(256), 0, st, vm, src, pairs,
                         s.order, gsrc, grad.plane);
    } else {
      hipLaunchKernelGGL((k_plane_scatter_sorted<Src, GS>), dim3((unsigned)ceil_div(groups, 16)),
                         dim3(256), 0, st, vm, src, pairs, s.order, gsrc, grad.plane);
    }
    TFH_TRY(check_launch("vm scatter: planes"));
  }
  // the line pass keeps 4-byte taps: with 16-byte ones its merged adds need
  // the sums moved back to one channel per lane, which cost more than the
  // loads saved (0.83 -> 0.89 ms for 48 channels at 2.1 M rows)
  ProfScope ps(line_stage, st);
  hipLaunchKernelGGL((k_line_scatter<Src, GS>), dim3((unsigned)ceil_div(M, kGatherThreads / 16)),
                     dim3(kGatherThreads), 0, st, vm, src, M, gsrc, grad.line, nlc);
  return check_launch("vm scatter: lines");
}

}  // namespace tfh
