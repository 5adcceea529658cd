// XLA custom-call entry points (the original GPU custom-call form,
// api_version 0): render_rays's forward and VJP (render.py:105-279) and the
// fused training step (training.py:101-205) with their operands in XLA's
// buffer list and an ASCII descriptor as the opaque backend config. The
// buffer order and the descriptor's fields are documented in tensorf_hip.h.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>

#include "tfh_common.h"

namespace tfh {
int32_t train_step_run(const tfh_vm* density, const tfh_vm* appearance, const tfh_mlp* mlp,
                       const tfh_render_config* cfg, const float* aabb, const float* origins,
                       const float* directions, const uint32_t* camera_indices,
                       const float* colors, int64_t num_rays, const float* jitter,
                       const float* base_ts, const float* base_steps, const float* gumbel,
                       const tfh_vm_grad* density_grad, const tfh_vm_grad* appearance_grad,
                       const tfh_mlp_grad* mlp_grad, const tfh_train_buffers* buffers,
                       float lr_tensor, float lr_mlp, float decay, int64_t adam_count,
                       const float* hyper, float* mse_out, void* workspace,
                       size_t workspace_bytes, hipStream_t st);
}  // namespace tfh

namespace {

using tfh::fail;

__global__ void k_xla_fill(float* __restrict__ p, int64_t n, float v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

void fill(float* p, int64_t n, float v, hipStream_t st) {
  if (!p || n <= 0) return;
  const int64_t blocks = std::min<int64_t>(tfh::ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_xla_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

// The parsed descriptor.
struct Desc {
  tfh_render_config cfg{};
  int64_t R = 0, dG = 0, dC = 0, aG = 0, aC = 0;
  int64_t squash = 27, units = 128, nf = 2, nv = 2, ncam = 0;
  bool jitter = false, schedule = false, gumbel = false;
  size_t ws = 0;
  int64_t offsets[12] = {};
  int64_t total = 0, tcount = 0;
  bool has_layout = false;
};

int32_t parse(const char* opaque, size_t len, Desc& d, const char* who) {
  if (opaque == nullptr || len == 0) return fail(TFH_ERR_INVALID_ARGUMENT, "%s: empty descriptor", who);
  std::string text(opaque, len);
  std::map<std::string, std::string> kv;
  size_t pos = 0;
  while (pos < text.size()) {
    while (pos < text.size() && (text[pos] == ' ' || text[pos] == '\0')) ++pos;
    size_t end = pos;
    while (end < text.size() && text[end] != ' ' && text[end] != '\0') ++end;
    if (end > pos) {
      const std::string tok = text.substr(pos, end - pos);
      const size_t eq = tok.find('=');
      if (eq == std::string::npos || eq == 0)
        return fail(TFH_ERR_INVALID_ARGUMENT, "%s: descriptor token '%s' is not key=value", who,
                    tok.c_str());
      kv[tok.substr(0, eq)] = tok.substr(eq + 1);
    }
    pos = end;
  }
  if (kv["tfhx"] != "1")
    return fail(TFH_ERR_INVALID_ARGUMENT, "%s: descriptor does not start a tfhx=1 record", who);
  const char* need[] = {"near", "far", "mode", "N", "k", "contraction", "precision", "sampling",
                        "shift", "R", "dG", "dC", "aG", "aC", "squash", "units", "nf", "nv",
                        "ncam", "jitter", "schedule", "gumbel", "ws"};
  for (const char* key : need)
    if (kv.find(key) == kv.end())
      return fail(TFH_ERR_INVALID_ARGUMENT, "%s: descriptor lacks '%s'", who, key);
  auto I = [&](const char* k) { return (int64_t)strtoll(kv[k].c_str(), nullptr, 10); };
  auto F = [&](const char* k) { return (float)strtod(kv[k].c_str(), nullptr); };
  d.cfg.near_plane = F("near");
  d.cfg.far_plane = F("far");
  d.cfg.mode = (int32_t)I("mode");
  d.cfg.density_samples_per_ray = (int32_t)I("N");
  d.cfg.appearance_samples_per_ray = (int32_t)I("k");
  d.cfg.scene_contraction = (int32_t)I("contraction");
  d.cfg.mlp_precision = (int32_t)I("precision");
  d.cfg.sampling = (int32_t)I("sampling");
  d.cfg.density_shift = F("shift");
  d.R = I("R");
  d.dG = I("dG");
  d.dC = I("dC");
  d.aG = I("aG");
  d.aC = I("aC");
  d.squash = I("squash");
  d.units = I("units");
  d.nf = I("nf");
  d.nv = I("nv");
  d.ncam = I("ncam");
  d.jitter = I("jitter") != 0;
  d.schedule = I("schedule") != 0;
  d.gumbel = I("gumbel") != 0;
  d.ws = (size_t)strtoull(kv["ws"].c_str(), nullptr, 10);
  if (kv.count("offsets")) {
    const char* p = kv["offsets"].c_str();
    for (int i = 0; i < 12; ++i) {
      char* e = nullptr;
      d.offsets[i] = (int64_t)strtoll(p, &e, 10);
      if (e == p) return fail(TFH_ERR_INVALID_ARGUMENT, "%s: offsets needs 12 entries", who);
      p = (*e == ',') ? e + 1 : e;
    }
    if (!kv.count("total") || !kv.count("tcount"))
      return fail(TFH_ERR_INVALID_ARGUMENT, "%s: offsets without total/tcount", who);
    d.total = I("total");
    d.tcount = I("tcount");
    d.has_layout = true;
  }
  if (d.R < 0 || d.dG < 2 || d.aG < 2 || d.dC < 1 || d.aC < 1 || d.ncam < 0)
    return fail(TFH_ERR_INVALID_ARGUMENT, "%s: bad sizes in the descriptor", who);
  return TFH_OK;
}

int64_t encoded_dim(const Desc& d) { return d.squash + 3 + 2 * d.nf * d.squash + 2 * d.nv * 3; }

// Element counts of the twelve parameter views (factor plane/line of the
// density and appearance fields, then the MLP leaves).
void view_sizes(const Desc& d, int64_t n[12]) {
  const int64_t F = 3 * d.aC, E = encoded_dim(d), U = d.units;
  n[0] = 3 * d.dG * d.dG * d.dC;
  n[1] = 3 * d.dG * d.dC;
  n[2] = 3 * d.aG * d.aG * d.aC;
  n[3] = 3 * d.aG * d.aC;
  n[4] = F * d.squash;
  n[5] = E * U;
  n[6] = U;
  n[7] = U * U;
  n[8] = U;
  n[9] = U * 3;
  n[10] = 3;
  n[11] = d.ncam * U;
}

tfh_mlp make_mlp(const Desc& d, float* const w[8]) {
  tfh_mlp m{};
  m.w0 = w[0]; m.w1 = w[1]; m.b1 = w[2]; m.w2 = w[3]; m.b2 = w[4]; m.w3 = w[5]; m.b3 = w[6];
  m.embed = d.ncam > 0 ? w[7] : nullptr;
  m.feat_dim = 3 * d.aC;
  m.squash_dim = d.squash;
  m.units = d.units;
  m.feature_n_freqs = d.nf;
  m.viewdir_n_freqs = d.nv;
  m.num_cameras = d.ncam;
  return m;
}

void report(const char* who) {
  fprintf(stderr, "%s failed: %s\n", who, tfh_last_error());
}

float* F32(void* p) { return static_cast<float*>(p); }

}  // namespace

extern "C" {

void tfh_xla_render_fwd(void* stream, void** b, const char* opaque, size_t opaque_len) {
  const char* who = "tfh_xla_render_fwd";
  hipStream_t st = tfh::as_stream(stream);
  Desc d;
  int32_t s = parse(opaque, opaque_len, d, who);
  const bool parsed = s == TFH_OK;  // only then are the result sizes known
  if (s == TFH_OK) {
    const tfh_vm dvm{F32(b[0]), F32(b[1]), d.dG, d.dC};
    const tfh_vm avm{F32(b[2]), F32(b[3]), d.aG, d.aC};
    float* w[8] = {F32(b[4]), F32(b[5]), F32(b[6]), F32(b[7]), F32(b[8]), F32(b[9]), F32(b[10]),
                   F32(b[11])};
    const tfh_mlp mlp = make_mlp(d, w);
    s = tfh_render_fwd(&dvm, &avm, &mlp, &d.cfg, F32(b[12]), F32(b[13]), F32(b[14]),
                       static_cast<const uint32_t*>(b[15]), d.R, d.jitter ? F32(b[16]) : nullptr,
                       d.schedule ? F32(b[17]) : nullptr, d.schedule ? F32(b[18]) : nullptr,
                       d.gumbel ? F32(b[19]) : nullptr, F32(b[20]), b[21], d.ws, stream);
  }
  if (s != TFH_OK) {
    report(who);
    if (parsed) fill(F32(b[20]), d.R * (d.cfg.mode == TFH_MODE_RGB ? 3 : 1), NAN, st);
  }
}

void tfh_xla_render_bwd(void* stream, void** b, const char* opaque, size_t opaque_len) {
  const char* who = "tfh_xla_render_bwd";
  hipStream_t st = tfh::as_stream(stream);
  Desc d;
  int32_t s = parse(opaque, opaque_len, d, who);
  const bool parsed = s == TFH_OK;  // only then are the result sizes known
  int64_t n[12];
  view_sizes(d, n);
  if (parsed)
    for (int i = 0; i < 12; ++i) fill(F32(b[18 + i]), n[i], 0.0f, st);  // VJP outputs start at 0
  if (s == TFH_OK && b[30] != b[17] && d.ws > 0) {
    if (hipMemcpyAsync(b[30], b[17], d.ws, hipMemcpyDeviceToDevice, st) != hipSuccess)
      s = fail(TFH_ERR_LAUNCH, "%s: copy of the workspace failed", who);
  }
  if (s == TFH_OK) {
    const tfh_vm dvm{F32(b[0]), F32(b[1]), d.dG, d.dC};
    const tfh_vm avm{F32(b[2]), F32(b[3]), d.aG, d.aC};
    float* w[8] = {F32(b[4]), F32(b[5]), F32(b[6]), F32(b[7]), F32(b[8]), F32(b[9]), F32(b[10]),
                   F32(b[11])};
    const tfh_mlp mlp = make_mlp(d, w);
    const tfh_vm_grad dg{F32(b[18]), F32(b[19])}, ag{F32(b[20]), F32(b[21])};
    tfh_mlp_grad mg{F32(b[22]), F32(b[23]), F32(b[24]), F32(b[25]), F32(b[26]), F32(b[27]),
                    F32(b[28]), d.ncam > 0 ? F32(b[29]) : nullptr};
    s = tfh_render_bwd(&dvm, &avm, &mlp, &d.cfg, F32(b[12]), F32(b[13]), F32(b[14]),
                       static_cast<const uint32_t*>(b[15]), d.R, F32(b[16]), &dg, &ag, &mg, b[30],
                       d.ws, stream);
  }
  if (s != TFH_OK) {
    report(who);
    if (parsed)
      for (int i = 0; i < 12; ++i) fill(F32(b[18 + i]), n[i], NAN, st);
  }
}

void tfh_xla_train_step(void* stream, void** b, const char* opaque, size_t opaque_len) {
  const char* who = "tfh_xla_train_step";
  hipStream_t st = tfh::as_stream(stream);
  Desc d;
  int32_t s = parse(opaque, opaque_len, d, who);
  const bool parsed = s == TFH_OK;
  if (s == TFH_OK && !d.has_layout)
    s = fail(TFH_ERR_INVALID_ARGUMENT, "%s: the descriptor needs offsets, total and tcount", who);
  const size_t bytes = (size_t)(s == TFH_OK ? d.total : 0) * sizeof(float);
  for (int i = 0; i < 3 && s == TFH_OK; ++i)  // parameters and moments are updated in place
    if (b[13 + i] != b[i] &&
        hipMemcpyAsync(b[13 + i], b[i], bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
      s = fail(TFH_ERR_LAUNCH, "%s: copy of operand %d failed", who, i);
  if (s == TFH_OK) {
    int64_t n[12];
    view_sizes(d, n);
    for (int i = 0; i < 12; ++i)
      if (d.offsets[i] < 0 || d.offsets[i] + n[i] > d.total)
        s = fail(TFH_ERR_INVALID_ARGUMENT, "%s: view %d outside the flat buffer", who, i);
  }
  if (s == TFH_OK) {
    float* P = F32(b[13]);
    float* G = F32(b[17]);
    fill(G, d.total, 0.0f, st);  // the step needs zero gradients on entry
    const int64_t* o = d.offsets;
    const tfh_vm dvm{P + o[0], P + o[1], d.dG, d.dC};
    const tfh_vm avm{P + o[2], P + o[3], d.aG, d.aC};
    float* w[8] = {P + o[4], P + o[5], P + o[6], P + o[7], P + o[8], P + o[9], P + o[10],
                   P + o[11]};
    const tfh_mlp mlp = make_mlp(d, w);
    const tfh_vm_grad dg{G + o[0], G + o[1]}, ag{G + o[2], G + o[3]};
    tfh_mlp_grad mg{G + o[4], G + o[5], G + o[6], G + o[7], G + o[8], G + o[9], G + o[10],
                    d.ncam > 0 ? G + o[11] : nullptr};
    const tfh_train_buffers tb{P, G, F32(b[14]), F32(b[15]), d.total, d.tcount};
    s = tfh::train_step_run(&dvm, &avm, &mlp, &d.cfg, F32(b[3]), F32(b[4]), F32(b[5]),
                            static_cast<const uint32_t*>(b[6]), F32(b[7]), d.R,
                            d.jitter ? F32(b[8]) : nullptr, d.schedule ? F32(b[9]) : nullptr,
                            d.schedule ? F32(b[10]) : nullptr, d.gumbel ? F32(b[11]) : nullptr,
                            &dg, &ag, &mg, &tb, 0.0f, 0.0f, 0.0f, 0, F32(b[12]), F32(b[16]),
                            b[18], d.ws, st);
  }
  if (s != TFH_OK) {
    report(who);
    if (parsed) fill(F32(b[16]), 1, NAN, st);
  }
}

}  // extern "C"
