// Backward of the standard FeatureMlp shape with a FiLM camera embedding
// (networks.py:103-111) that reads H1, relu(z2) before the embedding and the
// output kept by the forward kernel instead of recomputing layers 1-3: the
// kept-activation kernel's FILM variant, in a translation unit of its own
// (register allocation, see tfh_mlp_kernels.h).
#include "tfh_mlp_bwd_kept.h"

namespace tfh {

void launch_mlp_bwd_std_film_act(bool bf, unsigned nblk, size_t lds, hipStream_t st,
                                 const MlpDims& d, const MlpW& w, const MlpPack& pk,
                                 const GradPack& gp, const float* feats, const float* dirs,
                                 const uint32_t* cams, const float* grgb, int64_t M, int64_t rpr,
                                 float* gfeat, float* partials, float* gembed, const float* act,
                                 const float* rgb) {
  (void)gp;
  if (bf) {
    allow_lds(k_mlp_bwd_kept<true, true>, lds);
    hipLaunchKernelGGL((k_mlp_bwd_kept<true, true>), dim3(nblk), dim3(kKeptNW * 64), lds, st, w,
                       pk, feats, dirs, grgb, M, rpr, gfeat, partials, act, rgb, cams, gembed,
                       d.ncam);
  } else {
    allow_lds(k_mlp_bwd_kept<false, true>, lds);
    hipLaunchKernelGGL((k_mlp_bwd_kept<false, true>), dim3(nblk), dim3(kKeptNW * 64), lds, st, w,
                       pk, feats, dirs, grgb, M, rpr, gfeat, partials, act, rgb, cams, gembed,
                       d.ncam);
  }
}

}  // namespace tfh
