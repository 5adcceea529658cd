// Procedural benchmark scenes and their ground truth: a density field made of
// isotropic Gaussian blobs with per-blob albedo, rendered by emission-absorption
// quadrature in FP64 (white background). Replaces the image datasets of the
// reference (data.py) for synthetic training and evaluation.
#include "tfh_common.h"

namespace tfh {

// blob b: [cx, cy, cz, sigma, peak, r, g, b]
// sigma(x) = sum_b peak_b exp(-|x - c_b|^2 / (2 sigma_b^2))
// colour(x) = sum_b sigma_b(x) albedo_b / sigma(x)
__global__ void __launch_bounds__(256) k_blob_render(const double* __restrict__ blobs, int nb,
                                                     const float* __restrict__ origins,
                                                     const float* __restrict__ dirs, int64_t R,
                                                     double t_near, double t_far, int n_samples,
                                                     int contracted_far, float* __restrict__ rgb,
                                                     float* __restrict__ opacity,
                                                     float* __restrict__ median_depth) {
  extern __shared__ double sb[];
  for (int i = threadIdx.x; i < nb * 8; i += blockDim.x) sb[i] = blobs[i];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const double ox = origins[3 * r], oy = origins[3 * r + 1], oz = origins[3 * r + 2];
  const double dx = dirs[3 * r], dy = dirs[3 * r + 1], dz = dirs[3 * r + 2];
  double logT = 0.0, cr = 0.0, cg = 0.0, cb = 0.0;
  double depth = INFINITY;  // where the transmittance crosses 1/2
  const double log_half = -0.69314718055994530942;
  // midpoint rule; with contracted_far the samples are uniform in 1/t past t=1
  // (unbounded scenes), otherwise uniform in t over [t_near, t_far]
  for (int i = 0; i < n_samples; ++i) {
    double t, dt, ta;
    const double u0 = (double)i / n_samples, u1 = (double)(i + 1) / n_samples;
    if (contracted_far) {
      // s in [0, 1]: t = t_near + s * (1 - t_near) for s <= 0.5 (linear),
      // and 1 / t linear from 1 to 1 / t_far for s > 0.5
      auto tmap = [&](double s) {
        if (s <= 0.5) return t_near + 2.0 * s * (1.0 - t_near);
        const double q = 2.0 * (s - 0.5);
        return 1.0 / (1.0 + q * (1.0 / t_far - 1.0));
      };
      ta = tmap(u0);
      const double tb = tmap(u1);
      t = 0.5 * (ta + tb);
      dt = tb - ta;
    } else {
      dt = (t_far - t_near) / n_samples;
      t = t_near + (i + 0.5) * dt;
      ta = t - 0.5 * dt;
    }
    const double px = ox + t * dx, py = oy + t * dy, pz = oz + t * dz;
    double sig = 0.0, ar = 0.0, ag = 0.0, ab = 0.0;
    for (int b = 0; b < nb; ++b) {
      const double* q = sb + 8 * b;
      const double ex = px - q[0], ey = py - q[1], ez = pz - q[2];
      const double s = q[4] * exp(-(ex * ex + ey * ey + ez * ez) / (2.0 * q[3] * q[3]));
      sig += s;
      ar += s * q[5];
      ag += s * q[6];
      ab += s * q[7];
    }
    if (sig > 0.0) {
      const double a = sig * dt;
      const double w = exp(logT) * (-expm1(-a));
      cr += w * ar / sig;
      cg += w * ag / sig;
      cb += w * ab / sig;
      // the density is taken constant over the step: log T falls linearly
      if (logT > log_half && logT - a <= log_half) depth = ta + (logT - log_half) / a * dt;
      logT -= a;
    }
  }
  const double T = exp(logT);
  rgb[3 * r] = (float)(cr + T);
  rgb[3 * r + 1] = (float)(cg + T);
  rgb[3 * r + 2] = (float)(cb + T);
  if (opacity) opacity[r] = (float)(1.0 - T);
  if (median_depth) median_depth[r] = (float)depth;
}

}  // namespace tfh

using namespace tfh;

extern "C" int32_t tfh_blob_scene_render(const double* blobs, int32_t num_blobs,
                                         const float* origins, const float* directions,
                                         int64_t num_rays, double t_near, double t_far,
                                         int32_t samples_per_ray, int32_t contracted_far,
                                         float* rgb, float* opacity, float* median_depth,
                                         tfh_stream_t stream) {
  TFH_REQUIRE(num_blobs >= 1 && num_blobs <= 1024, "tfh_blob_scene_render: num_blobs %d", num_blobs);
  TFH_REQUIRE(num_rays >= 0 && samples_per_ray >= 1, "tfh_blob_scene_render: bad sizes");
  TFH_REQUIRE(t_far > t_near && t_near >= 0.0, "tfh_blob_scene_render: need 0 <= t_near < t_far");
  if (num_rays == 0) return TFH_OK;
  TFH_REQUIRE(blobs && origins && directions && rgb, "tfh_blob_scene_render: null pointer");
  TFH_TRY(require_device());
  const int64_t blocks = ceil_div(num_rays, 256);
  TFH_REQUIRE(blocks < (1ll << 31), "tfh_blob_scene_render: too many rays");
  hipLaunchKernelGGL(k_blob_render, dim3((unsigned)blocks), dim3(256),
                     (size_t)num_blobs * 8 * sizeof(double), as_stream(stream), blobs, num_blobs,
                     origins, directions, num_rays, t_near, t_far, samples_per_ray,
                     contracted_far, rgb, opacity, median_depth);
  return check_launch("tfh_blob_scene_render");
}
