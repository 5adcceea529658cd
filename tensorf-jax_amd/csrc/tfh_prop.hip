// Learned proposal sampler (no counterpart in the reference, which stops at a
// comment, render.py:126) and the VJP of the segment probabilities.
//
// Forward, per round (n intervals per ray):
//   k_prop_points    interval edges (round 0: box segment split evenly or the
//                    contracted schedule) -> sample t = e_i + u (e_{i+1} - e_i)
//                    -> contracted / normalised point
//   k_vm_gather      proposal VM features (M, 3C)
//   prop_mlp_fwd     sigma = softplus(MLP(features)) on FP32 MFMA
//   k_prop_resample  one wave per ray: weights (log-space transmittance scan),
//                    padded histogram CDF, inverse-CDF edges of the next round
// Backward (interval loss of mip-NeRF 360 against the field's weights):
//   k_prop_loss_bwd  one wave per ray: bounds by binary search over the edges,
//                    dL/dw of the round by sums over the overlapping index
//                    ranges, segment VJP -> dL/dz
//   prop_mlp_bwd     persistent MLP backward -> dL/dfeatures
//   k_vm_scatter     proposal factor gradients
#include <algorithm>

#include "tfh_prop.h"

namespace tfh {

int32_t check_prop_mlp(const tfh_prop_mlp* p, const char* who);
int32_t prop_mlp_fwd(const tfh_prop_mlp* p, const float* feats, int64_t M, float* z, float* sigma,
                     hipStream_t st);
size_t prop_mlp_bwd_workspace(const tfh_prop_mlp* p, int64_t M);
// idx / count (device): process only the first *count rows listed in idx
int32_t prop_mlp_bwd(const tfh_prop_mlp* p, const float* feats, const float* dz, int64_t M,
                     float* dfeat, const tfh_prop_mlp_grad* g, void* ws, size_t ws_bytes,
                     hipStream_t st, const int32_t* idx = nullptr,
                     const int32_t* count = nullptr);

constexpr float kLossEps = 1e-7f;

// ---- device building blocks (one wave per ray) ---------------------------------------
// Inverse-CDF edges: h_j = w_j^power + pad / n, F = prefix sums of h, out[m] =
// F^{-1}(m / n_out * F_n) on the piecewise-linear CDF over the edges, with
// out[0] = edge[0] and out[n_out] = edge[n].
__device__ inline void wave_resample(const float* s_edge, const float* s_w, int n, int n_out,
                                     float pad, float power, float* s_h, float* s_cdf,
                                     float* __restrict__ out) {
  const int lane = lane_id();
  const float floor_mass = pad / (float)n;
  for (int j = lane; j < n; j += 64) {
    const float w = s_w[j];
    s_h[j] = (power == 1.0f ? w : (power == 0.0f ? 1.0f : powf(w, power))) + floor_mass;
  }
  __syncthreads();
  wave_excl_prefix(s_h, s_cdf, n);
  __syncthreads();
  const float total = s_cdf[n];
  for (int m = lane; m <= n_out; m += 64) {
    float s;
    if (m == 0) {
      s = s_edge[0];
    } else if (m == n_out) {
      s = s_edge[n];
    } else {
      const float q = total * ((float)m / (float)n_out);
      // largest j in [0, n-1] with cdf[j] <= q
      const int j = min(max(upper_bound(s_cdf, n, q) - 1, 0), n - 1);
      const float lo = s_cdf[j], den = s_cdf[j + 1] - lo;
      const float frac = den > 0.0f ? fminf(fmaxf((q - lo) / den, 0.0f), 1.0f) : 0.0f;
      const float e0 = s_edge[j], e1 = s_edge[j + 1];
      s = fminf(fmaxf(e0 + frac * (e1 - e0), e0), e1);
    }
    out[m] = s;
  }
}

// Interval loss of one ray. Field intervals [fe_i, fe_{i+1}) with weights fw_i
// (i < nf), proposal intervals [pe_j, pe_{j+1}) with weights pw_j (j < n):
//   bound_i = sum of pw_j over the j with pe_{j+1} > fe_i and pe_j < fe_{i+1}
//   loss = scale * sum_i max(0, fw_i - bound_i)^2 / (fw_i + eps)
// and s_G[j] = dloss/dpw_j. Returns the loss (wave-uniform). Scratch: s_g
// (nf). Both edge arrays are sorted, so the proposal intervals overlapping
// field interval i are an index range [jlo, jend) and the field intervals
// overlapping proposal interval j are a range [c_hi, c_lo); both sums run
// over their ranges directly (about nf + n terms in all). Differences of
// prefix sums would cancel in FP32, and dloss/dbound = -2 (w - bound) / w
// amplifies an absolute error of the bound by 1/w.
__device__ inline float wave_interval_loss(const float* s_fe, const float* __restrict__ fw, int nf,
                                           const float* s_pe, const float* s_pw, int n,
                                           float scale, float* s_g, float* s_G) {
  const int lane = lane_id();
  float loss = 0.0f;
  for (int i = lane; i < nf; i += 64) {
    const float a = s_fe[i], b = s_fe[i + 1], wi = fw[i];
    const int jlo = upper_bound(s_pe + 1, n, a);   // first j with pe_{j+1} > a
    const int jend = lower_bound(s_pe, n, b);      // one past the last j with pe_j < b
    float bound = 0.0f;
    for (int j = jlo; j < jend; ++j) bound += s_pw[j];
    const float dlt = fmaxf(wi - bound, 0.0f);
    const float inv = 1.0f / (wi + kLossEps);
    loss += dlt * dlt * inv;
    s_g[i] = -2.0f * dlt * inv * scale;
  }
  __syncthreads();
  for (int j = lane; j < n; j += 64) {
    const int c_lo = lower_bound(s_fe, nf, s_pe[j + 1]);  // i with fe_i < pe_{j+1}
    const int c_hi = upper_bound(s_fe + 1, nf, s_pe[j]);  // i with fe_{i+1} <= pe_j
    float acc = 0.0f;
    for (int i = c_hi; i < c_lo; ++i) acc += s_g[i];
    s_G[j] = acc;
  }
  __syncthreads();
  return wave_sum(loss) * scale;
}

// ---- proposal kernels -------------------------------------------------------------------
struct PropPts {
  const float *o, *d, *aabb;
  int64_t R;
  int n, contraction, round0;
  const float* base_ts;     // round 0 with contraction: (n) schedule
  const float* base_steps;  // (n)
  const float* jitter;      // jitter[r * jstride + joff + i] or null
  int jstride, joff;
  float* edges;             // (R, n+1): written in round 0, read otherwise
  float* pts;               // (R*n, 3) normalised
};

__global__ void __launch_bounds__(256) k_prop_points(PropPts p) {
  const int64_t total = p.R * p.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / p.n;
    const int i = (int)(e - r * p.n);
    const float ox = p.o[3 * r], oy = p.o[3 * r + 1], oz = p.o[3 * r + 2];
    const float dx = p.d[3 * r], dy = p.d[3 * r + 1], dz = p.d[3 * r + 2];
    float* edge = p.edges + r * (p.n + 1);
    float e0, e1;
    if (p.round0) {
      if (p.contraction) {
        e0 = p.base_ts[i];
        e1 = i + 1 < p.n ? p.base_ts[i + 1] : p.base_ts[i] + p.base_steps[i];
      } else {
        float t0, step;
        bbox_segment(ox, oy, oz, dx, dy, dz, p.aabb, p.n, t0, step);
        e0 = t0 + (float)i * step;
        e1 = t0 + (float)(i + 1) * step;
      }
      edge[i] = e0;
      if (i == p.n - 1) edge[p.n] = e1;
    } else {
      e0 = edge[i];
      e1 = edge[i + 1];
    }
    const float u = p.jitter ? p.jitter[r * p.jstride + p.joff + i] : 0.5f;
    const float t = e0 + u * (e1 - e0);
    float px = ox + t * dx, py = oy + t * dy, pz = oz + t * dz;
    if (p.contraction) contract(px, py, pz);
    float* q = p.pts + e * 3;
    q[0] = normalise(px, p.aabb[0], p.aabb[3]);
    q[1] = normalise(py, p.aabb[1], p.aabb[4]);
    q[2] = normalise(pz, p.aabb[2], p.aabb[5]);
  }
}

// LDS layout (in floats) of k_prop_resample, shared by the kernel and its launch.
struct ResampleLds {
  int edge, a, e, w, cdf, total;
  __host__ __device__ explicit ResampleLds(int n) {
    edge = 0;          // n+1 interval edges
    a = edge + n + 1;  // n optical depths, then the histogram
    e = a + n;         // n transmittances
    w = e + n;         // n weights
    cdf = w + n;       // n+1 histogram prefix sums
    total = cdf + n + 1;
  }
};

// Weights of the round from sigma and its edges, then the next round's edges.
__global__ void __launch_bounds__(64) k_prop_resample(const float* __restrict__ edges,
                                                      const float* __restrict__ sigma, int n,
                                                      int n_out, float pad, float power,
                                                      float* __restrict__ w_out,
                                                      float* __restrict__ e_out,
                                                      float* __restrict__ next_edges) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  const ResampleLds L(n);
  float* s_edge = sm + L.edge;
  float* s_a = sm + L.a;
  float* s_e = sm + L.e;
  float* s_w = sm + L.w;
  float* s_cdf = sm + L.cdf;
  for (int j = threadIdx.x; j <= n; j += 64) s_edge[j] = edges[r * (n + 1) + j];
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 64)
    s_a[j] = sigma[r * n + j] * (s_edge[j + 1] - s_edge[j]);
  __syncthreads();
  wave_segment_probs(s_a, s_e, s_w, n);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 64) {
    w_out[r * n + j] = s_w[j];
    e_out[r * n + j] = s_e[j];
  }
  wave_resample(s_edge, s_w, n, n_out, pad, power, s_a, s_cdf, next_edges + r * (n_out + 1));
}

// LDS layout (in floats) of k_prop_loss_bwd, shared by the kernel and its launch.
struct LossLds {
  int fe, pe, pw, g, G, e, da, total;
  __host__ __device__ LossLds(int nf, int n) {
    fe = 0;             // nf+1 field edges
    pe = fe + nf + 1;   // n+1 round edges
    pw = pe + n + 1;    // n round weights
    g = pw + n;         // nf loss gradients of the field intervals
    G = g + nf;         // n dloss/dw of the round
    e = G + n;          // n round transmittances
    da = e + n;         // n dloss/da of the round
    total = da + n;
  }
};

// Interval loss of one round and its backward to dL/dz of the round's samples.
__global__ void __launch_bounds__(64) k_prop_loss_bwd(
    const float* __restrict__ fedges, const float* __restrict__ fw, int nf,
    const float* __restrict__ pedges, const float* __restrict__ pw, const float* __restrict__ pe,
    const float* __restrict__ sigma, int n, float scale, float* __restrict__ dz,
    float* __restrict__ loss_out, int nslot) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  const LossLds L(nf, n);
  float* s_fe = sm + L.fe;
  float* s_pe = sm + L.pe;
  float* s_pw = sm + L.pw;
  float* s_g = sm + L.g;
  float* s_G = sm + L.G;
  float* s_e = sm + L.e;
  float* s_da = sm + L.da;
  for (int i = threadIdx.x; i <= nf; i += 64) s_fe[i] = fedges[r * (nf + 1) + i];
  for (int j = threadIdx.x; j <= n; j += 64) s_pe[j] = pedges[r * (n + 1) + j];
  for (int j = threadIdx.x; j < n; j += 64) {
    s_pw[j] = pw[r * n + j];
    s_e[j] = pe[r * n + j];
  }
  __syncthreads();
  const float loss =
      wave_interval_loss(s_fe, fw + r * nf, nf, s_pe, s_pw, n, scale, s_g, s_G);
  // ray r adds into partial sum r % nslot (adds to one address serialise)
  if (threadIdx.x == 0 && loss != 0.0f) unsafeAtomicAdd(loss_out + r % nslot, loss);
  wave_segment_vjp(s_G, nullptr, s_pw, s_e, s_da, n);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 64) {
    const float sig = sigma[r * n + j];
    // sigma = softplus(z): dsigma/dz = sigmoid(z) = 1 - exp(-sigma)
    dz[r * n + j] = s_da[j] * (s_pe[j + 1] - s_pe[j]) * (-expm1f(-sig));
  }
}

// ---- standalone kernels -------------------------------------------------------------
__global__ void __launch_bounds__(64) k_resample_api(const float* __restrict__ edges,
                                                     const float* __restrict__ w, int n,
                                                     int n_out, float pad, float power,
                                                     float* __restrict__ out) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  float* s_edge = sm;
  float* s_w = s_edge + n + 1;
  float* s_h = s_w + n;
  float* s_cdf = s_h + n;
  for (int j = threadIdx.x; j <= n; j += 64) s_edge[j] = edges[r * (n + 1) + j];
  for (int j = threadIdx.x; j < n; j += 64) s_w[j] = w[r * n + j];
  __syncthreads();
  wave_resample(s_edge, s_w, n, n_out, pad, power, s_h, s_cdf, out + r * (n_out + 1));
}

__global__ void __launch_bounds__(64) k_interval_loss_api(
    const float* __restrict__ fedges, const float* __restrict__ fw, int nf,
    const float* __restrict__ pedges, const float* __restrict__ pw, int n, float scale,
    float* __restrict__ loss, float* __restrict__ gpw) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  float* s_fe = sm;
  float* s_pe = s_fe + nf + 1;
  float* s_pw = s_pe + n + 1;
  float* s_g = s_pw + n;
  float* s_G = s_g + nf;
  for (int i = threadIdx.x; i <= nf; i += 64) s_fe[i] = fedges[r * (nf + 1) + i];
  for (int j = threadIdx.x; j <= n; j += 64) s_pe[j] = pedges[r * (n + 1) + j];
  for (int j = threadIdx.x; j < n; j += 64) s_pw[j] = pw[r * n + j];
  __syncthreads();
  const float l =
      wave_interval_loss(s_fe, fw + r * nf, nf, s_pe, s_pw, n, scale, s_g, s_G);
  if (threadIdx.x == 0) loss[r] = l;
  if (gpw)
    for (int j = threadIdx.x; j < n; j += 64) gpw[r * n + j] = s_G[j];
}

__global__ void __launch_bounds__(64) k_segment_vjp_api(
    const float* __restrict__ sigmas, const float* __restrict__ steps,
    const float* __restrict__ e, const float* __restrict__ w, const float* __restrict__ ge,
    const float* __restrict__ gw, int n, float* __restrict__ gsig) {
  extern __shared__ float sm[];
  const int64_t r = blockIdx.x;
  float* s_da = sm;
  (void)sigmas;
  wave_segment_vjp(gw ? gw + r * n : nullptr, ge ? ge + r * n : nullptr, w + r * n, e + r * n,
                   s_da, n);
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += 64) gsig[r * n + j] = s_da[j] * steps[r * n + j];
}

// ---- host orchestration -------------------------------------------------------------
static int32_t check_prop_cfg(const tfh_proposal_config* pc, const char* who) {
  TFH_REQUIRE(pc != nullptr, "%s: null proposal config", who);
  TFH_REQUIRE(pc->samples_round0 >= 2 && pc->samples_round0 <= kMaxSamples,
              "%s: samples_round0 %d outside [2, %d]", who, pc->samples_round0, kMaxSamples);
  TFH_REQUIRE(pc->samples_round1 >= 0 && pc->samples_round1 <= kMaxSamples,
              "%s: samples_round1 %d outside [0, %d]", who, pc->samples_round1, kMaxSamples);
  TFH_REQUIRE(pc->samples_out >= 1 && pc->samples_out <= kMaxSamples,
              "%s: samples_out %d outside [1, %d]", who, pc->samples_out, kMaxSamples);
  TFH_REQUIRE(pc->histogram_pad >= 0.0f, "%s: negative histogram_pad", who);
  TFH_REQUIRE(pc->weight_power >= 0.0f && pc->weight_power <= 1.0f,
              "%s: weight_power %g outside [0, 1]", who, pc->weight_power);
  return TFH_OK;
}

size_t carve_prop(Carve& cv, const tfh_proposal_config* pc, const tfh_vm* field,
                  const tfh_prop_mlp* mlp, int64_t R, PropWs& ws) {
  const int64_t F = 3 * field->channels;
  ws.rounds = pc->samples_round1 > 0 ? 2 : 1;
  ws.rd[0].n = pc->samples_round0;
  ws.rd[1].n = pc->samples_round1;
  ws.rows = R * (ws.rd[0].n + ws.rd[1].n);
  ws.pts = cv.take<float>(ws.rows * 3);
  ws.feats = cv.take<float>(ws.rows * F);
  ws.dz = cv.take<float>(ws.rows);
  ws.dfeat = cv.take<float>(ws.rows * F);
  float* sigma = cv.take<float>(ws.rows);
  float* w = cv.take<float>(ws.rows);
  float* e = cv.take<float>(ws.rows);
  int64_t row0 = 0;
  for (int k = 0; k < 2; ++k) {
    PropRound& rd = ws.rd[k];
    rd.edges = cv.take<float>(k < ws.rounds ? R * (rd.n + 1) : 0);
    if (cv.ok && k < ws.rounds) {
      rd.pts = ws.pts + row0 * 3;
      rd.feats = ws.feats + row0 * F;
      rd.sigma = sigma + row0;
      rd.w = w + row0;
      rd.e = e + row0;
      rd.dz = ws.dz + row0;
    }
    row0 += R * rd.n;
  }
  ws.mlp_ws_bytes = prop_mlp_bwd_workspace(mlp, std::max<int64_t>(ws.rows, 1));
  ws.mlp_ws = cv.take<char>((int64_t)ws.mlp_ws_bytes);
  ws.nz_nblk = (ws.rows + kNzBlock - 1) / kNzBlock;
  ws.nz_idx = cv.take<int32_t>(ws.rows);
  ws.nz_blk = cv.take<int32_t>(ws.nz_nblk + 1);
  // gradient copies: up to 32 of the planes within 16 MB, up to 256 of the
  // lines within 2 MB (14 MB at 64^3 x 8 channels)
  const int64_t pn = 3 * field->grid_dim * field->grid_dim * field->channels;
  const int64_t ln = 3 * field->grid_dim * field->channels;
  ws.npc = (int)std::max<int64_t>(1, std::min<int64_t>(32, (16 << 20) / (pn * 4)));
  ws.nlc = (int)std::max<int64_t>(1, std::min<int64_t>(256, (2 << 20) / (ln * 4)));
  ws.gcopy = cv.take<float>(ws.npc * pn + ws.nlc * ln);
  ws.loss_part = cv.take<float>(kLossSlots);
  return cv.off;
}

size_t prop_workspace_bytes(const tfh_proposal_config* pc, const tfh_vm* field, int64_t R) {
  // the MLP workspace depends only on the feature width
  tfh_prop_mlp probe{};
  probe.feat_dim = 3 * field->channels;
  Carve cv(nullptr, 0);
  PropWs ws;
  return carve_prop(cv, pc, field, &probe, R, ws);
}

int32_t prop_sample(const tfh_vm* field, const tfh_prop_mlp* mlp, const tfh_proposal_config* pc,
                    const float* aabb, const float* o, const float* d, int64_t R,
                    const float* jitter, const float* base_ts, const float* base_steps,
                    float* edges_out, PropWs& ws, hipStream_t st) {
  const int jstride = ws.rd[0].n + ws.rd[1].n;
  int joff = 0;
  for (int k = 0; k < ws.rounds; ++k) {
    PropRound& rd = ws.rd[k];
    const int64_t M = R * rd.n;
    PropPts pp{o, d, aabb, R, rd.n, pc->scene_contraction ? 1 : 0, k == 0, base_ts, base_steps,
               jitter, jstride, joff, rd.edges, rd.pts};
    {
      ProfScope ps("prop_points", st);
      const int64_t blocks = std::min<int64_t>(ceil_div(M, 256), 1 << 16);
      hipLaunchKernelGGL(k_prop_points, dim3((unsigned)blocks), dim3(256), 0, st, pp);
    }
    TFH_TRY(check_launch("proposal points"));
    {
      ProfScope ps("prop_gather", st);
      RaySamples src{rd.pts, nullptr, 1, 1};
      TFH_TRY((launch_vm_gather<RaySamples, true>(*field, src, M, rd.feats, st)));
    }
    TFH_TRY(prop_mlp_fwd(mlp, rd.feats, M, nullptr, rd.sigma, st));
    const bool last = k + 1 == ws.rounds;
    const int n_out = last ? pc->samples_out : ws.rd[k + 1].n;
    float* next = last ? edges_out : ws.rd[k + 1].edges;
    {
      ProfScope ps("prop_resample", st);
      const size_t lds = (size_t)ResampleLds(rd.n).total * sizeof(float);
      hipLaunchKernelGGL(k_prop_resample, dim3((unsigned)R), dim3(64), lds, st, rd.edges,
                         rd.sigma, rd.n, n_out, pc->histogram_pad, pc->weight_power, rd.w, rd.e,
                         next);
    }
    TFH_TRY(check_launch("proposal resample"));
    joff += rd.n;
  }
  return TFH_OK;
}

// ---- compaction of the rows with a non-zero loss gradient ----------------------------
// Most proposal samples satisfy the bound once training settles (their dL/dz
// is exactly 0: 1-3% of round-0 and 3-8% of round-1 samples remain after 1000
// steps on the blob scene), and a zero row contributes nothing to the
// proposal MLP's gradients or to the factor scatter. The list keeps row order
// (deterministic): per-block counts, one scan, then ordered writes.
__device__ __forceinline__ int block_excl_scan256(int flag, int* s_w, int& total) {
  // 256 threads = 4 waves
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const uint64_t b = __ballot(flag);
  const int within = __popcll(b & ((1ull << lane) - 1));
  if (lane == 0) s_w[wave] = __popcll(b);
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += s_w[w];
  total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();
  return before + within;
}

__global__ void __launch_bounds__(256) k_nz_count(const float* __restrict__ dz, int64_t M,
                                                  int32_t* __restrict__ blk) {
  __shared__ int s_w[4];
  const int64_t base = (int64_t)blockIdx.x * kNzBlock;
  int sum = 0;
  for (int p = 0; p < kNzBlock / 256; ++p) {
    const int64_t row = base + p * 256 + threadIdx.x;
    int total;
    (void)block_excl_scan256(row < M && dz[row] != 0.0f, s_w, total);
    sum += total;
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = sum;
}

// Exclusive scan of nb block counts in place, total into blk[nb]; one block.
__global__ void __launch_bounds__(1024) k_nz_scan(int32_t* __restrict__ blk, int64_t nb) {
  __shared__ int s_part[1024];
  const int64_t per = (nb + 1023) / 1024;
  const int64_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  int local = 0;
  for (int64_t b = b0; b < b1; ++b) local += blk[b];
  s_part[threadIdx.x] = local;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
    const int v = threadIdx.x >= o ? s_part[threadIdx.x - o] : 0;
    __syncthreads();
    s_part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s_part[threadIdx.x] - local;
  for (int64_t b = b0; b < b1; ++b) {
    const int c = blk[b];
    blk[b] = run;
    run += c;
  }
  if (threadIdx.x == 1023) blk[nb] = s_part[1023];
}

__global__ void __launch_bounds__(256) k_nz_write(const float* __restrict__ dz, int64_t M,
                                                  const int32_t* __restrict__ blk,
                                                  int32_t* __restrict__ idx) {
  __shared__ int s_w[4];
  const int64_t base = (int64_t)blockIdx.x * kNzBlock;
  int out = blk[blockIdx.x];
  for (int p = 0; p < kNzBlock / 256; ++p) {
    const int64_t row = base + p * 256 + threadIdx.x;
    const bool nz = row < M && dz[row] != 0.0f;
    int total;
    const int pos = block_excl_scan256(nz, s_w, total);
    if (nz) idx[out + pos] = (int32_t)row;
    out += total;
  }
}

int32_t prop_loss_bwd(const tfh_vm* field, const tfh_prop_mlp* mlp,
                      const tfh_proposal_config* pc, int64_t R, const float* fedges,
                      const float* fw, const tfh_vm_grad* fgrad, const tfh_prop_mlp_grad* mgrad,
                      float* loss_out, PropWs& ws, hipStream_t st) {
  const int nf = pc->samples_out;
  const float scale = pc->loss_weight / (float)R;
  {
    ProfScope ps("prop_loss_bwd", st);
    TFH_TRY(zero_copies(ws.loss_part, kLossSlots, st));
    for (int k = 0; k < ws.rounds; ++k) {
      const PropRound& rd = ws.rd[k];
      const size_t lds = (size_t)LossLds(nf, rd.n).total * sizeof(float);
      hipLaunchKernelGGL(k_prop_loss_bwd, dim3((unsigned)R), dim3(64), lds, st, fedges, fw, nf,
                         rd.edges, rd.w, rd.e, rd.sigma, rd.n, scale, rd.dz, ws.loss_part,
                         kLossSlots);
      TFH_TRY(check_launch("proposal loss bwd"));
    }
    TFH_TRY(sum_copies(ws.loss_part, kLossSlots, 1, loss_out, st));
  }
  {
    ProfScope ps("prop_compact", st);
    TFH_REQUIRE(ws.rows < (1ll << 31), "proposal loss bwd: %lld rows", (long long)ws.rows);
    hipLaunchKernelGGL(k_nz_count, dim3((unsigned)ws.nz_nblk), dim3(256), 0, st, ws.dz, ws.rows,
                       ws.nz_blk);
    hipLaunchKernelGGL(k_nz_scan, dim3(1), dim3(1024), 0, st, ws.nz_blk, ws.nz_nblk);
    hipLaunchKernelGGL(k_nz_write, dim3((unsigned)ws.nz_nblk), dim3(256), 0, st, ws.dz, ws.rows,
                       ws.nz_blk, ws.nz_idx);
    TFH_TRY(check_launch("proposal loss bwd: compaction"));
  }
  const int32_t* count = ws.nz_blk + ws.nz_nblk;
  TFH_TRY(prop_mlp_bwd(mlp, ws.feats, ws.dz, ws.rows, ws.dfeat, mgrad, ws.mlp_ws, ws.mlp_ws_bytes,
                       st, ws.nz_idx, count));
  ProfScope ps("prop_scatter", st);
  CompactedPoints src{ws.pts, ws.nz_idx, count};
  const int64_t pn = 3 * field->grid_dim * field->grid_dim * field->channels;
  const int64_t ln = 3 * field->grid_dim * field->channels;
  const int64_t ncopy = ws.npc * pn + ws.nlc * ln;
  TFH_TRY(zero_copies(ws.gcopy, ncopy, st));
  const tfh_vm_grad priv{ws.gcopy, ws.gcopy + ws.npc * pn};
  TFH_TRY((launch_vm_scatter<CompactedPoints, true>(*field, src, ws.rows, ws.dfeat, priv, st,
                                                    ws.npc, ws.nlc)));
  TFH_TRY(sum_copies(ws.gcopy, ws.npc, pn, fgrad->plane, st));
  return sum_copies(ws.gcopy + ws.npc * pn, ws.nlc, ln, fgrad->line, st);
}

int32_t check_prop_args(const tfh_vm* field, const tfh_prop_mlp* mlp,
                        const tfh_proposal_config* pc, int64_t R, const char* who) {
  TFH_TRY(check_prop_cfg(pc, who));
  TFH_TRY(check_vm(field, who));
  TFH_TRY(check_prop_mlp(mlp, who));
  TFH_REQUIRE(mlp->feat_dim == 3 * field->channels,
              "%s: proposal mlp feat_dim %lld != 3 * proposal channels %lld", who,
              (long long)mlp->feat_dim, (long long)field->channels);
  TFH_REQUIRE(R >= 0 && R < (1ll << 31), "%s: num_rays %lld out of range", who, (long long)R);
  return TFH_OK;
}

}  // namespace tfh

using namespace tfh;

extern "C" {

size_t tfh_proposal_workspace_size(const tfh_proposal_config* pc, const tfh_vm* field,
                                   int64_t num_rays) {
  if (check_prop_cfg(pc, "tfh_proposal_workspace_size") != TFH_OK ||
      check_vm(field, "tfh_proposal_workspace_size") != TFH_OK || num_rays < 0)
    return 0;
  return prop_workspace_bytes(pc, field, num_rays);
}

int32_t tfh_proposal_sample(const tfh_vm* field, const tfh_prop_mlp* mlp,
                            const tfh_proposal_config* pc, const float* aabb,
                            const float* origins, const float* directions, int64_t num_rays,
                            const float* jitter, const float* base_ts, const float* base_steps,
                            float* edges_out, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream) {
  TFH_TRY(check_prop_args(field, mlp, pc, num_rays, "tfh_proposal_sample"));
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(aabb && origins && directions && edges_out, "tfh_proposal_sample: null pointer");
  if (pc->scene_contraction)
    TFH_REQUIRE(base_ts && base_steps, "tfh_proposal_sample: contraction needs the schedule");
  TFH_TRY(require_device());
  Carve cv(workspace, workspace_bytes);
  PropWs ws;
  carve_prop(cv, pc, field, mlp, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_proposal_sample: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  return prop_sample(field, mlp, pc, aabb, origins, directions, num_rays, jitter, base_ts,
                     base_steps, edges_out, ws, as_stream(stream));
}

int32_t tfh_proposal_round(const tfh_proposal_config* pc, const tfh_vm* field, int64_t num_rays,
                           int32_t round, const void* workspace, size_t workspace_bytes,
                           const float** edges, const float** weights) {
  TFH_TRY(check_prop_cfg(pc, "tfh_proposal_round"));
  TFH_TRY(check_vm(field, "tfh_proposal_round"));
  TFH_REQUIRE(num_rays >= 0, "tfh_proposal_round: negative num_rays");
  TFH_REQUIRE(edges && weights && workspace, "tfh_proposal_round: null pointer");
  tfh_prop_mlp probe{};
  probe.feat_dim = 3 * field->channels;
  Carve cv(const_cast<void*>(workspace), workspace_bytes);
  PropWs ws;
  carve_prop(cv, pc, field, &probe, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_proposal_round: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  TFH_REQUIRE(round >= 0 && round < ws.rounds, "tfh_proposal_round: round %d of %d", round,
              ws.rounds);
  *edges = ws.rd[round].edges;
  *weights = ws.rd[round].w;
  return TFH_OK;
}

int32_t tfh_proposal_loss_bwd(const tfh_vm* field, const tfh_prop_mlp* mlp,
                              const tfh_proposal_config* pc, int64_t num_rays, const float* edges,
                              const float* field_w, const tfh_vm_grad* field_grad,
                              const tfh_prop_mlp_grad* mlp_grad, float* loss_out, void* workspace,
                              size_t workspace_bytes, tfh_stream_t stream) {
  TFH_TRY(check_prop_args(field, mlp, pc, num_rays, "tfh_proposal_loss_bwd"));
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(edges && field_w && field_grad && field_grad->plane && field_grad->line &&
                  mlp_grad && loss_out,
              "tfh_proposal_loss_bwd: null pointer");
  TFH_TRY(require_device());
  Carve cv(workspace, workspace_bytes);
  PropWs ws;
  carve_prop(cv, pc, field, mlp, num_rays, ws);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "tfh_proposal_loss_bwd: workspace of %zu bytes, need %zu",
                workspace_bytes, cv.off);
  return prop_loss_bwd(field, mlp, pc, num_rays, edges, field_w, field_grad, mlp_grad, loss_out,
                       ws, as_stream(stream));
}

int32_t tfh_proposal_resample(const float* edges, const float* weights, int64_t num_rays,
                              int64_t num_intervals, int64_t num_out, float pad, float power,
                              float* edges_out, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && num_intervals >= 1 && num_intervals <= 4096 && num_out >= 1 &&
                  num_out <= 4096,
              "tfh_proposal_resample: bad sizes");
  TFH_REQUIRE(pad >= 0.0f, "tfh_proposal_resample: negative pad");
  TFH_REQUIRE(power >= 0.0f && power <= 1.0f, "tfh_proposal_resample: power outside [0, 1]");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(edges && weights && edges_out, "tfh_proposal_resample: null pointer");
  TFH_TRY(require_device());
  const size_t lds = (size_t)(4 * num_intervals + 2) * sizeof(float);
  hipLaunchKernelGGL(k_resample_api, dim3((unsigned)num_rays), dim3(64), lds, as_stream(stream),
                     edges, weights, (int)num_intervals, (int)num_out, pad, power, edges_out);
  return check_launch("tfh_proposal_resample");
}

int32_t tfh_interval_loss(const float* edges, const float* weights, int64_t num_intervals,
                          const float* prop_edges, const float* prop_weights,
                          int64_t num_prop_intervals, int64_t num_rays, float scale,
                          float* loss, float* grad_prop_weights, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && num_intervals >= 1 && num_intervals <= 4096 &&
                  num_prop_intervals >= 1 && num_prop_intervals <= 4096,
              "tfh_interval_loss: bad sizes");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(edges && weights && prop_edges && prop_weights && loss,
              "tfh_interval_loss: null pointer");
  TFH_TRY(require_device());
  const int64_t nf = num_intervals, n = num_prop_intervals;
  const size_t lds = (size_t)(2 * (nf + 1) + 4 * n + 2) * sizeof(float);
  hipLaunchKernelGGL(k_interval_loss_api, dim3((unsigned)num_rays), dim3(64), lds,
                     as_stream(stream), edges, weights, (int)nf, prop_edges, prop_weights, (int)n,
                     scale, loss, grad_prop_weights);
  return check_launch("tfh_interval_loss");
}

int32_t tfh_segment_probabilities_bwd(const float* sigmas, const float* step_sizes,
                                      int64_t num_rays, int64_t samples_per_ray,
                                      const float* p_exits, const float* p_terminates,
                                      const float* grad_p_exits, const float* grad_p_terminates,
                                      float* grad_sigmas, tfh_stream_t stream) {
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 1 && samples_per_ray <= 4096,
              "tfh_segment_probabilities_bwd: samples_per_ray %lld outside [1, 4096]",
              (long long)samples_per_ray);
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(step_sizes && p_exits && p_terminates && grad_sigmas,
              "tfh_segment_probabilities_bwd: null pointer");
  TFH_REQUIRE(grad_p_exits || grad_p_terminates,
              "tfh_segment_probabilities_bwd: no upstream gradient");
  TFH_TRY(require_device());
  hipLaunchKernelGGL(k_segment_vjp_api, dim3((unsigned)num_rays), dim3(64),
                     samples_per_ray * sizeof(float), as_stream(stream), sigmas, step_sizes,
                     p_exits, p_terminates, grad_p_exits, grad_p_terminates, (int)samples_per_ray,
                     grad_sigmas);
  return check_launch("tfh_segment_probabilities_bwd");
}

}  // extern "C"
