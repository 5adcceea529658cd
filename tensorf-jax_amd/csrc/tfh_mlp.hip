// FeatureMlp (networks.py:38-121) forward and backward: weight packing, the
// launches and the C entry points. The kernels are in tfh_mlp_kernels.h.
#include "tfh_mlp_kernels.h"
#include "tfh_vm_kernels.h"

namespace tfh {

// grads[e] += sum_b partials[b][e], mapped onto the separate gradient arrays.
// Sum of the workgroup partials: a workgroup owns 64 consecutive entries and
// its kRedWaves waves split the partials (wave w adds rows w, w + kRedWaves,
// ...), so every load is one coalesced 256-byte row segment; the waves' sums
// meet in LDS and are added in a fixed order.
constexpr int kRedWaves = 16;

__global__ void __launch_bounds__(64 * kRedWaves) k_mlp_reduce(const float* __restrict__ partials,
                                                              int64_t nblk, GradPack gp,
                                                              tfh_mlp_grad g) {
  __shared__ float red[kRedWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t n = gp.b3 + 3;
  const int64_t e = (int64_t)blockIdx.x * 64 + lane;
  float part = 0.0f;
  if (e < n) {
#pragma unroll 8
    for (int64_t b = w; b < nblk; b += kRedWaves) part += partials[b * gp.total + e];
  }
  red[w][lane] = part;
  __syncthreads();
  if (w == 0 && e < n) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < kRedWaves; ++i) s += red[i][lane];
    float* dst;
    int64_t o;
    if (e >= gp.b3) { dst = g.b3; o = e - gp.b3; }
    else if (e >= gp.w3) { dst = g.w3; o = e - gp.w3; }
    else if (e >= gp.b2) { dst = g.b2; o = e - gp.b2; }
    else if (e >= gp.w2) { dst = g.w2; o = e - gp.w2; }
    else if (e >= gp.b1) { dst = g.b1; o = e - gp.b1; }
    else if (e >= gp.w1) { dst = g.w1; o = e - gp.w1; }
    else { dst = g.w0; o = e - gp.w0; }
    dst[o] += s;
  }
}

static int g_num_cus = 0;

static int num_cus() {
  if (g_num_cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      g_num_cus = n;
    else
      g_num_cus = 256;
  }
  return g_num_cus;
}


int32_t check_mlp(const tfh_mlp* p, const char* who) {
  TFH_REQUIRE(p != nullptr, "%s: null tfh_mlp", who);
  TFH_REQUIRE(p->w0 && p->w1 && p->b1 && p->w2 && p->b2 && p->w3 && p->b3,
              "%s: null MLP parameter pointer", who);
  TFH_REQUIRE(p->units == kU, "%s: units=%lld unsupported (kernels are built for %d)", who,
              (long long)p->units, kU);
  TFH_REQUIRE(p->squash_dim >= 1 && p->squash_dim <= kSP, "%s: squash_dim %lld outside [1, %d]",
              who, (long long)p->squash_dim, kSP);
  TFH_REQUIRE(p->feat_dim >= 1 && p->feat_dim <= 192, "%s: feat_dim %lld outside [1, 192]", who,
              (long long)p->feat_dim);
  TFH_REQUIRE(p->feature_n_freqs >= 0 && p->feature_n_freqs <= 8 && p->viewdir_n_freqs >= 0 &&
                  p->viewdir_n_freqs <= 8,
              "%s: Fourier frequency counts must be in [0, 8]", who);
  TFH_REQUIRE(p->embed == nullptr || p->num_cameras >= 1, "%s: embed given with num_cameras=%lld",
              who, (long long)p->num_cameras);
  MlpDims d = make_dims(p, true);
  TFH_REQUIRE(d.EP <= 448, "%s: encoded width %d too large", who, d.E);
  return TFH_OK;
}

static MlpW to_w(const tfh_mlp* p) {
  return MlpW{p->w0, p->w1, p->b1, p->w2, p->b2, p->w3, p->b3, p->embed};
}

// Workspace of the weight tables of one precision.
static size_t pack_bytes(const MlpDims& d, bool bf) {
  PackSizes s = pack_sizes(d);
  const size_t el = bf ? sizeof(__bf16) : sizeof(float);
  size_t n = 0;
  for (int64_t v : {s.f0, s.f1, s.f2, s.f3, s.g0, s.g1, s.g2, s.g3})
    n += align_up((size_t)v * el, 256);
  return n;
}

template <typename T>
static PackOut<T> carve_tables(Carve& cv, const PackSizes& s) {
  PackOut<T> o;
  o.f0 = cv.take<T>(s.f0); o.f1 = cv.take<T>(s.f1); o.f2 = cv.take<T>(s.f2);
  o.f3 = cv.take<T>(s.f3); o.g0 = cv.take<T>(s.g0); o.g1 = cv.take<T>(s.g1);
  o.g2 = cv.take<T>(s.g2); o.g3 = cv.take<T>(s.g3);
  return o;
}

// Carves and fills the weight tables of the operand precision.
static int32_t build_pack(const tfh_mlp* p, const MlpDims& d, bool bf, Carve& cv, MlpPack& pk,
                          hipStream_t st) {
  pk = MlpPack{};
  PackSizes s = pack_sizes(d);
  if (bf) {
    PackOut<__bf16> o = carve_tables<__bf16>(cv, s);
    if (!cv.ok) return fail(TFH_ERR_WORKSPACE, "mlp: workspace too small for the weight tables");
    ProfScope ps("mlp_pack", st);
    hipLaunchKernelGGL(k_pack_weights<__bf16>, dim3(128), dim3(256), 0, st, d, to_w(p), o, s);
    pk.f0 = o.f0; pk.f1 = o.f1; pk.f2 = o.f2; pk.f3 = o.f3;
    pk.g0 = o.g0; pk.g1 = o.g1; pk.g2 = o.g2; pk.g3 = o.g3;
  } else {
    PackOut<float> o = carve_tables<float>(cv, s);
    if (!cv.ok) return fail(TFH_ERR_WORKSPACE, "mlp: workspace too small for the weight tables");
    ProfScope ps("mlp_pack", st);
    hipLaunchKernelGGL(k_pack_weights<float>, dim3(128), dim3(256), 0, st, d, to_w(p), o, s);
    pk.ff0 = o.f0; pk.ff1 = o.f1; pk.ff2 = o.f2; pk.ff3 = o.f3;
    pk.fg0 = o.g0; pk.fg1 = o.g1; pk.fg2 = o.g2; pk.fg3 = o.g3;
  }
  return check_launch("mlp: weight tables");
}

size_t mlp_fwd_workspace(const tfh_mlp* p, int32_t precision) {
  const bool bf = precision == TFH_PREC_BF16;
  return pack_bytes(make_dims(p, bf), bf);
}

// The standard shape keeps relu(z1) and relu(z2) (before any camera
// embedding) for the backward when the caller provides the room (the fused
// training steps).
bool mlp_keeps_activations(const tfh_mlp* p, int32_t precision) {
  const bool bf = precision == TFH_PREC_BF16;
  return p && is_std_dims(make_dims(p, bf), bf);
}

int32_t mlp_fwd_rows(const tfh_mlp* p, const float* feats, const float* dirs, const uint32_t* cams,
                     int64_t M, int64_t rpr, float* rgb, int32_t precision, void* ws,
                     size_t ws_bytes, hipStream_t st, float* act = nullptr) {
  const bool bf = precision == TFH_PREC_BF16;
  MlpDims d = make_dims(p, bf);
  Carve cv(ws, ws_bytes);
  MlpPack pk;
  TFH_TRY(build_pack(p, d, bf, cv, pk, st));
  const bool sd = is_std_dims(d, bf);
  // (the BF16 standard shape keeps its feature rows in BF16: k_mlp_fwd)
  const size_t lds = fwd_lds_floats<kFwdTM>(d, bf && sd) * sizeof(float);
  TFH_REQUIRE(lds <= 160 * 1024, "mlp fwd: tile needs %zu bytes of LDS", lds);
  const int64_t blocks = ceil_div(M, kFwdTM);
  ProfScope ps("mlp_fwd_kernel", st);
#define TFH_LAUNCH_FWD(BFV, STDV)                                                             \
  do {                                                                                       \
    allow_lds(k_mlp_fwd<kFwdTM, kFwdNW, BFV, STDV>, lds);                                    \
    hipLaunchKernelGGL((k_mlp_fwd<kFwdTM, kFwdNW, BFV, STDV>), dim3((unsigned)blocks),       \
                       dim3(kFwdNW * 64), lds, st, d, to_w(p), pk, feats, dirs, cams, M, rpr,  \
                       rgb, act, NoGather{});                                                     \
  } while (0)
  if (bf) {
    if (sd) TFH_LAUNCH_FWD(true, true); else TFH_LAUNCH_FWD(true, false);
  } else {
    if (sd) TFH_LAUNCH_FWD(false, true); else TFH_LAUNCH_FWD(false, false);
  }
#undef TFH_LAUNCH_FWD
  return check_launch("mlp fwd");
}

// The appearance gather (TensorVM.interpolate, tensor_vm.py:42-89; the
// arithmetic of k_vm_gather4, so the rows are the same values) as the tile
// source of the standard forward: a 16-lane group per row, lane l channels
// 4l .. 4l+3 of each mode (48 channels: lanes 12..15 write the BF16 rows'
// zero padding). The rows also go to feats (M, 144) for the backward. The
// four workgroups of a CU overlap one tile's tap latency with the others'
// products.
struct GatherTile {
  tfh_vm vm;         // appearance field, 48 channels, 16-byte aligned
  RaySamples src;    // the sample points of the rows
  float* feats;      // (M, 144) out
  template <int TM, int NT, bool CAMS, bool FB16>
  __device__ void load(const MlpDims& d, int64_t row0, int64_t M, int64_t rpr, float* Fs,
                       float* Ds, int* Cs, const float* __restrict__ dirs,
                       const uint32_t* __restrict__ cams) const {
    constexpr int C = 48;
    static_assert(TM * 16 == NT, "a 16-lane group per row");
    const int t = threadIdx.x, i = t >> 4, l = t & 15, c4 = 4 * l;
    const int G = (int)vm.grid_dim;
    const int64_t row = row0 + i;
    const bool valid = row < M;
    // the view directions and cameras first, so that their loads overlap
    // the taps'
    float dv = 0.f;
    if (t < TM * 3) {
      const int i3 = t / 3, c3 = t - 3 * i3;
      const int64_t r = row0 + i3;
      if (r < M) dv = dirs[(r / rpr) * 3 + c3];
    }
    int cam = 0;
    if constexpr (CAMS) {
      if (t < TM) {
        const int64_t r = row0 + t;
        if (r < M) cam = (int)cams[r / rpr];
      }
    }
    float x = 0.f, y = 0.f, z = 0.f;
    if (valid) src.get(row, x, y, z);
    ModeTaps mt[3];
    mode_taps(x, y, z, G, mt);
    // the rows go to feats after every mode's taps are loaded: a store to
    // feats between them would order the next mode's loads behind it (feats
    // may alias the factors as far as the compiler knows). The 64-register
    // budget of four workgroups per CU keeps one mode's taps in flight at a
    // time; holding two took 98 registers (or spilled at 64).
    float4 f[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      f[m] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (valid && c4 < C) {
        TapAddr a;
        tap_addrs(mt[m], m, G, a);
        float4 lv, pv;
        tap_values4(vm.plane, vm.line, C, c4, a, lv, pv);
        // tensor_vm.py:167
        f[m] = make_float4(lv.x * pv.x, lv.y * pv.y, lv.z * pv.z, lv.w * pv.w);
      }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      if (c4 < C) {
        if constexpr (FB16) {
          typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
          const bf16x4 b = {(__bf16)f[m].x, (__bf16)f[m].y, (__bf16)f[m].z, (__bf16)f[m].w};
          *reinterpret_cast<bf16x4*>(reinterpret_cast<__bf16*>(Fs) + i * kFsLdB + m * C + c4) = b;
        } else {
          *reinterpret_cast<float4*>(Fs + i * d.ldF + m * C + c4) = f[m];
        }
      }
    }
    if (valid && c4 < C) {
#pragma unroll
      for (int m = 0; m < 3; ++m)
        *reinterpret_cast<float4*>(feats + row * 3 * C + m * C + c4) = f[m];
    }
    if constexpr (FB16) {
      // columns 3C .. FP - 1 (the K padding of layer 0) are zero
      static_assert(3 * C + 16 == std_dims<true>().FP, "BF16 rows pad 144 to 160");
      if (l >= 12) {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        const bf16x4 zero = {(__bf16)0.f, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f};
        *reinterpret_cast<bf16x4*>(reinterpret_cast<__bf16*>(Fs) + i * kFsLdB + 3 * C +
                                   4 * (l - 12)) = zero;
      }
    }
    if (t < TM * 3) Ds[t] = dv;
    if constexpr (CAMS) {
      if (t < TM) Cs[t] = (cam < 0 || cam >= d.ncam) ? 0 : cam;
    }
  }
};

// The standard forward with the appearance gather as its tile source
// (GatherTile); *done = false when the shape, channels or alignment do not
// allow it (the caller then gathers and calls mlp_fwd_rows).
int32_t mlp_fwd_rows_gather(const tfh_mlp* p, const tfh_vm& vm, const RaySamples& src,
                            float* feats, const float* dirs, const uint32_t* cams, int64_t M,
                            int64_t rpr, float* rgb, int32_t precision, void* ws,
                            size_t ws_bytes, hipStream_t st, float* act, bool* done) {
  *done = false;
  const bool bf = precision == TFH_PREC_BF16;
  MlpDims d = make_dims(p, bf);
  if (!is_std_dims(d, bf) || vm.channels != 48 || 3 * vm.channels != d.F ||
      (reinterpret_cast<uintptr_t>(vm.plane) & 15) || (reinterpret_cast<uintptr_t>(vm.line) & 15) ||
      (reinterpret_cast<uintptr_t>(feats) & 15) || M <= 0)
    return TFH_OK;
  Carve cv(ws, ws_bytes);
  MlpPack pk;
  TFH_TRY(build_pack(p, d, bf, cv, pk, st));
  const size_t lds = fwd_lds_floats<kFwdTM>(d, bf) * sizeof(float);
  TFH_REQUIRE(lds <= 160 * 1024, "mlp fwd: tile needs %zu bytes of LDS", lds);
  const int64_t blocks = xcd_grid(ceil_div(M, kFwdTM));  // XCD-ordered tiles
  const GatherTile gt{vm, src, feats};
  ProfScope ps("mlp_fwd_kernel", st);
  if (bf) {
    allow_lds(k_mlp_fwd<kFwdTM, kFwdNW, true, true, GatherTile>, lds);
    hipLaunchKernelGGL((k_mlp_fwd<kFwdTM, kFwdNW, true, true, GatherTile>), dim3((unsigned)blocks),
                       dim3(kFwdNW * 64), lds, st, d, to_w(p), pk, nullptr, dirs, cams, M, rpr,
                       rgb, act, gt);
  } else {
    allow_lds(k_mlp_fwd<kFwdTM, kFwdNW, false, true, GatherTile>, lds);
    hipLaunchKernelGGL((k_mlp_fwd<kFwdTM, kFwdNW, false, true, GatherTile>),
                       dim3((unsigned)blocks), dim3(kFwdNW * 64), lds, st, d, to_w(p), pk,
                       nullptr, dirs, cams, M, rpr, rgb, act, gt);
  }
  *done = true;
  return check_launch("mlp fwd with the gather");
}

static int64_t bwd_blocks(int64_t M) {
  int64_t nblk = std::min<int64_t>(ceil_div(M, kBwdTM), (int64_t)num_cus());
  return nblk < 1 ? 1 : nblk;
}

size_t mlp_bwd_workspace(const tfh_mlp* p, int64_t M, int32_t precision) {
  const bool bf = precision == TFH_PREC_BF16;
  MlpDims d = make_dims(p, bf);
  GradPack gp = make_pack(d);
  return pack_bytes(d, bf) + align_up((size_t)bwd_blocks(M) * gp.total * sizeof(float), 256);
}

int32_t mlp_bwd_rows(const tfh_mlp* p, const float* feats, const float* dirs, const uint32_t* cams,
                     int64_t M, int64_t rpr, const float* grgb, float* gfeat,
                     const tfh_mlp_grad* g, int32_t precision, void* ws, size_t ws_bytes,
                     hipStream_t st, const float* act = nullptr,
                     const float* rgb = nullptr) {
  const bool bf = precision == TFH_PREC_BF16;
  MlpDims d = make_dims(p, bf);
  GradPack gp = make_pack(d);
  const int64_t nblk = bwd_blocks(M);
  Carve cv(ws, ws_bytes);
  MlpPack pk;
  TFH_TRY(build_pack(p, d, bf, cv, pk, st));
  float* partials = cv.take<float>(nblk * gp.total);
  if (!cv.ok)
    return fail(TFH_ERR_WORKSPACE, "mlp bwd: workspace of %zu bytes, need %zu", ws_bytes,
                mlp_bwd_workspace(p, M, precision));
  // TFH_MLP_WGRAD=partial forces the read-modify-write accumulation of the
  // weight gradients (the register-resident one is the default when the
  // shape has a specialised kernel)
  static const bool force_partial = [] {
    const char* v = getenv("TFH_MLP_WGRAD");
    return v && strcmp(v, "partial") == 0;
  }();
  const bool sd = is_std_dims(d, bf) && !force_partial;
  const bool kept_film = sd && act && d.ncam > 0;  // k_mlp_bwd_kept<BF, true>
  const size_t lds = bwd_lds_floats<kBwdTM>(d, sd ? kLdsSlots * kBwdNW : 0) * sizeof(float) +
                     (sd && d.ncam == 0 && kBwdNW == 8 ? (bf ? kG1LdsBytes : kG1LdsBytesF) : 0) +
                     (kept_film ? (bf ? kG1FilmBytes : kG1FilmBytesF) : 0);
  TFH_REQUIRE(lds <= 160 * 1024, "mlp bwd: tile needs %zu bytes of LDS", lds);
  {
    ProfScope ps("mlp_bwd", st);
#define TFH_LAUNCH_BWD(BFV, NTV, STDV, FILMV)                                                 \
  do {                                                                                        \
    allow_lds(k_mlp_bwd<kBwdTM, kBwdNW, BFV, NTV, STDV, FILMV>, lds);                         \
    hipLaunchKernelGGL((k_mlp_bwd<kBwdTM, kBwdNW, BFV, NTV, STDV, FILMV>),                    \
                       dim3((unsigned)nblk), dim3(kBwdNW * 64), lds, st, d, to_w(p), pk, gp,  \
                       feats, dirs, cams, grgb, M, rpr, gfeat, partials, g->embed);          \
  } while (0)
    const bool film = d.ncam > 0;
    if (sd && act && film)
      launch_mlp_bwd_std_film_act(bf, (unsigned)nblk, lds, st, d, to_w(p), pk, gp, feats, dirs,
                                  cams, grgb, M, rpr, gfeat, partials, g->embed, act, rgb);
    else if (sd && act && !film)
      launch_mlp_bwd_std_act(bf, (unsigned)nblk, lds, st, d, to_w(p), pk, gp, feats, dirs, grgb, M,
                             rpr, gfeat, partials, act, rgb);
    else if (sd && film)
      launch_mlp_bwd_std_film(bf, (unsigned)nblk, lds, st, d, to_w(p), pk, gp, feats, dirs, cams,
                              grgb, M, rpr, gfeat, partials, g->embed);
    else if (sd)
      launch_mlp_bwd_std(bf, (unsigned)nblk, lds, st, d, to_w(p), pk, gp, feats, dirs, cams, grgb,
                         M, rpr, gfeat, partials, g->embed);
    else if (bf)
      TFH_LAUNCH_BWD(true, 0, false, true);
    else
      TFH_LAUNCH_BWD(false, 0, false, true);
#undef TFH_LAUNCH_BWD
  }
  TFH_TRY(check_launch("mlp bwd"));
  ProfScope ps("mlp_reduce", st);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((unsigned)ceil_div(gp.b3 + 3, 64)),
                     dim3(64 * kRedWaves), 0, st, partials, nblk, gp, *g);
  return check_launch("mlp bwd: reduce");
}

}  // namespace tfh

using namespace tfh;

extern "C" {

size_t tfh_feature_mlp_fwd_workspace_size(const tfh_mlp* mlp, int32_t precision) {
  if (check_mlp(mlp, "tfh_feature_mlp_fwd_workspace_size") != TFH_OK) return 0;
  return mlp_fwd_workspace(mlp, precision);
}

int32_t tfh_feature_mlp_fwd(const tfh_mlp* mlp, const float* features, const float* viewdirs,
                            const uint32_t* camera_indices, int64_t num_rows, float* rgb,
                            int32_t precision, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream) {
  TFH_TRY(check_mlp(mlp, "tfh_feature_mlp_fwd"));
  TFH_REQUIRE(num_rows >= 0, "tfh_feature_mlp_fwd: negative num_rows");
  TFH_REQUIRE(precision == TFH_PREC_FP32 || precision == TFH_PREC_BF16,
              "tfh_feature_mlp_fwd: unknown precision %d", precision);
  if (num_rows == 0) return TFH_OK;
  TFH_REQUIRE(features && viewdirs && rgb, "tfh_feature_mlp_fwd: null pointer");
  TFH_REQUIRE(mlp->embed == nullptr || camera_indices,
              "tfh_feature_mlp_fwd: camera_indices required with an embedding");
  TFH_TRY(require_device());
  return mlp_fwd_rows(mlp, features, viewdirs, camera_indices, num_rows, 1, rgb, precision,
                      workspace, workspace_bytes, as_stream(stream));
}

size_t tfh_feature_mlp_bwd_workspace_size(const tfh_mlp* mlp, int64_t num_rows,
                                          int32_t precision) {
  if (check_mlp(mlp, "tfh_feature_mlp_bwd_workspace_size") != TFH_OK || num_rows < 0) return 0;
  return mlp_bwd_workspace(mlp, num_rows, precision);
}

int32_t tfh_feature_mlp_bwd(const tfh_mlp* mlp, const float* features, const float* viewdirs,
                            const uint32_t* camera_indices, int64_t num_rows,
                            const float* grad_rgb, float* grad_features, const tfh_mlp_grad* grad,
                            int32_t precision, void* workspace, size_t workspace_bytes,
                            tfh_stream_t stream) {
  TFH_TRY(check_mlp(mlp, "tfh_feature_mlp_bwd"));
  TFH_REQUIRE(num_rows >= 0, "tfh_feature_mlp_bwd: negative num_rows");
  TFH_REQUIRE(precision == TFH_PREC_FP32 || precision == TFH_PREC_BF16,
              "tfh_feature_mlp_bwd: unknown precision %d", precision);
  if (num_rows == 0) return TFH_OK;
  TFH_REQUIRE(features && viewdirs && grad_rgb && grad, "tfh_feature_mlp_bwd: null pointer");
  TFH_REQUIRE(grad->w0 && grad->w1 && grad->b1 && grad->w2 && grad->b2 && grad->w3 && grad->b3,
              "tfh_feature_mlp_bwd: null gradient pointer");
  TFH_REQUIRE(mlp->embed == nullptr || (camera_indices && grad->embed),
              "tfh_feature_mlp_bwd: camera_indices and grad->embed required with an embedding");
  TFH_TRY(require_device());
  return mlp_bwd_rows(mlp, features, viewdirs, camera_indices, num_rows, 1, grad_rgb,
                      grad_features, grad, precision, workspace, workspace_bytes,
                      as_stream(stream));
}

}  // extern "C"
