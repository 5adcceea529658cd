// Backward of the standard FeatureMlp shape without a camera embedding that
// reads H1, H2 and the output kept by the forward kernel instead of
// recomputing layers 1-3 (k_mlp_bwd_kept), in a translation unit of its own
// (register allocation, see tfh_mlp_kernels.h).
#include "tfh_mlp_bwd_kept.h"

namespace tfh {

void launch_mlp_bwd_std_act(bool bf, unsigned nblk, size_t lds, hipStream_t st,
                            const MlpDims& d, const MlpW& w, const MlpPack& pk,
                            const GradPack& gp, const float* feats, const float* dirs,
                            const float* grgb, int64_t M, int64_t rpr, float* gfeat,
                            float* partials, const float* act, const float* rgb) {
  (void)d;
  (void)gp;
  if (bf) {
    allow_lds(k_mlp_bwd_kept<true>, lds);
    hipLaunchKernelGGL(k_mlp_bwd_kept<true>, dim3(nblk), dim3(kKeptNW * 64), lds, st, w, pk,
                       feats, dirs, grgb, M, rpr, gfeat, partials, act, rgb, nullptr, nullptr, 0);
  } else {
    allow_lds(k_mlp_bwd_kept<false>, lds);
    hipLaunchKernelGGL(k_mlp_bwd_kept<false>, dim3(nblk), dim3(kKeptNW * 64), lds, st, w, pk,
                       feats, dirs, grgb, M, rpr, gfeat, partials, act, rgb, nullptr, nullptr, 0);
  }
}

}  // namespace tfh
