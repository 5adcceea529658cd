// Proposal sampler workspace and host entry points shared by tfh_prop.hip
// and the fused training step of tfh_render.hip.
#pragma once

#include "tfh_sampling.h"

namespace tfh {

struct PropRound {
  int n = 0;
  float* edges = nullptr;  // (R, n+1)
  float* pts = nullptr;    // (R*n, 3)
  float* feats = nullptr;  // (R*n, F)
  float* sigma = nullptr;  // (R*n)
  float* w = nullptr;      // (R*n)
  float* e = nullptr;      // (R*n)
  float* dz = nullptr;     // (R*n)
};

struct PropWs {
  PropRound rd[2];
  int rounds = 1;
  int64_t rows = 0;        // R * (n0 + n1)
  float* pts = nullptr;    // all rounds' rows, round 0 first
  float* feats = nullptr;
  float* dz = nullptr;
  float* dfeat = nullptr;
  void* mlp_ws = nullptr;
  size_t mlp_ws_bytes = 0;
  // rows whose interval-loss gradient is non-zero: their indices and, per
  // block of kNzBlock rows, counts then offsets (entry nz_nblk: the total)
  int32_t* nz_idx = nullptr;
  int32_t* nz_blk = nullptr;
  int64_t nz_nblk = 0;
  // npc copies of the field's plane gradient, then nlc of its line gradient:
  // the scatter's blocks add into them round-robin, one kernel sums them
  float* gcopy = nullptr;
  int npc = 1, nlc = 1;
  float* loss_part = nullptr;  // kLossSlots partial sums of the interval loss
};

constexpr int kNzBlock = 1024;

size_t carve_prop(Carve& cv, const tfh_proposal_config* pc, const tfh_vm* field,
                  const tfh_prop_mlp* mlp, int64_t R, PropWs& ws);
size_t prop_workspace_bytes(const tfh_proposal_config* pc, const tfh_vm* field, int64_t R);
int32_t check_prop_args(const tfh_vm* field, const tfh_prop_mlp* mlp,
                        const tfh_proposal_config* pc, int64_t R, const char* who);
int32_t prop_sample(const tfh_vm* field, const tfh_prop_mlp* mlp, const tfh_proposal_config* pc,
                    const float* aabb, const float* o, const float* d, int64_t R,
                    const float* jitter, const float* base_ts, const float* base_steps,
                    float* edges_out, PropWs& ws, hipStream_t st);
int32_t prop_loss_bwd(const tfh_vm* field, const tfh_prop_mlp* mlp,
                      const tfh_proposal_config* pc, int64_t R, const float* fedges,
                      const float* fw, const tfh_vm_grad* fgrad, const tfh_prop_mlp_grad* mgrad,
                      float* loss_out, PropWs& ws, hipStream_t st);

}  // namespace tfh
