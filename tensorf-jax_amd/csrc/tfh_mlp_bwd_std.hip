// Backward of the standard FeatureMlp shape (48 channels x 3 modes, squash 27,
// 2+2 Fourier frequencies) without a camera embedding, in a translation unit
// of its own (register allocation, see tfh_mlp_kernels.h).
#include "tfh_mlp_kernels.h"

namespace tfh {

void launch_mlp_bwd_std(bool bf, unsigned nblk, size_t lds, hipStream_t st, const MlpDims& d,
                        const MlpW& w, const MlpPack& pk, const GradPack& gp,
                        const float* feats, const float* dirs, const uint32_t* cams,
                        const float* grgb, int64_t M, int64_t rpr, float* gfeat,
                        float* partials, float* gembed) {
  if (bf) {
    allow_lds(k_mlp_bwd<kBwdTM, kBwdNW, true, kBwdNT, true, false>, lds);
    hipLaunchKernelGGL((k_mlp_bwd<kBwdTM, kBwdNW, true, kBwdNT, true, false>), dim3(nblk),
                       dim3(kBwdNW * 64), lds, st, d, w, pk, gp, feats, dirs, cams, grgb, M, rpr,
                       gfeat, partials, gembed);
  } else {
    allow_lds(k_mlp_bwd<kBwdTM, kBwdNW, false, kBwdNT, true, false>, lds);
    hipLaunchKernelGGL((k_mlp_bwd<kBwdTM, kBwdNW, false, kBwdNT, true, false>), dim3(nblk),
                       dim3(kBwdNW * 64), lds, st, d, w, pk, gp, feats, dirs, cams, grgb, M, rpr,
                       gfeat, partials, gembed);
  }
}

}  // namespace tfh
