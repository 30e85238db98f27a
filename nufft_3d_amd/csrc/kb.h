// Kaiser-Bessel weights evaluated on the fly (replaces the stored values of H).
//
// Restates @nufft_3d/kernel.m:2-34 on the kernel support, where the argument
// alpha*sqrt(1 - d2/(J/2)^2) is real and >= 0 (the imaginary continuation is
// only reached by deap, which runs on the host in double).
//   bessi0: Numerical-Recipes polynomial I0 (kernel.m:19-30)
//   value : bessi0(arg)/bessi0(alpha) (kernel.m:9)
// The reference stores every weight as single (nufft_3d.m:160,195,205):
// the fp64 path computes in double and rounds to float before widening, the
// fp32 path computes directly in float. Which grid points a sample reaches
// (the support test) is decided once, at plan time, in double.
#pragma once
#include <hip/hip_runtime.h>

namespace nufft3d {

struct KBParams {
  double J2;         // J^2 (support test 4*d2 <= J^2, nufft_3d.m:184-185,189)
  double half2;      // (J/2)^2 (kernel.m:4)
  double alpha;
  double b0;         // bessi0(alpha)
  float J2f, inv_half2f, alphaf, inv_b0f;
  float inv_P0f;     // 1/P(3.75/alpha), P the large-argument polynomial of bessi0
  int large_b0;      // alpha >= 3.75 (bessi0(alpha) in the exponential branch)
  int radial;
};

__device__ __forceinline__ double bessi0_d(double ax) {
#pragma clang fp contract(off)
  if (ax < 3.75) {
    double y = ax / 3.75;
    y = y * y;
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
                 y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  double y = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) *
         (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
          y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
          y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

// kernel(dist2) in double, as kernel.m evaluates it. The double-precision
// geometry and kernel code below is compiled without FMA contraction: the
// reference (MATLAB) rounds every product and sum, and the support decision
// taken at plan time must see the same dist2 as the weight evaluation.
__device__ __forceinline__ double kb_d(double dist2, const KBParams& p) {
#pragma clang fp contract(off)
  double arg = p.alpha * sqrt(fmax(1.0 - dist2 / p.half2, 0.0));
  return bessi0_d(arg) / p.b0;
}

// fp32 weight. For arguments in the exponential branch the ratio is taken as
// exp(ax - alpha) * sqrt(alpha/ax) * P(3.75/ax) / P(3.75/alpha) with
// ax - alpha = -alpha*q/(1 + sqrt(1-q)), q = d2/(J/2)^2: the same value as
// bessi0(ax)/bessi0(alpha) (kernel.m:9,28-30) without exponentiating the
// rounding error of ax ~ alpha, which would cost ~alpha ulps on the largest
// weights. Both branches are evaluated and one is selected (no divergence),
// on the hardware square root, reciprocal, 2^x and reciprocal square root
// (each within ~1 ulp): the weight carries a few ulps (~5e-7 relative),
// against the 2^-24 of rounding the reference's double weight to single.
__device__ __forceinline__ float kb_f(float dist2, const KBParams& p) {
  const float q = dist2 * p.inv_half2f;
  const float t = fmaxf(1.0f - q, 0.0f);
  const float s = __builtin_amdgcn_sqrtf(t);
  const float ax = p.alphaf * s;
  // ax < 3.75 (kernel.m:19-23)
  float y = ax * (1.0f / 3.75f);
  y = y * y;
  const float small = (1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f +
                       y * (0.360768e-1f + y * 0.45813e-2f)))))) * p.inv_b0f;
  // ax >= 3.75 (kernel.m:26-30), as the ratio to bessi0(alpha)
  const float e = -p.alphaf * q * __builtin_amdgcn_rcpf(1.0f + s);
  const float z = 3.75f * __builtin_amdgcn_rcpf(ax);
  const float P = 0.39894228f + z * (0.1328592e-1f + z * (0.225319e-2f + z * (-0.157565e-2f +
                  z * (0.916281e-2f + z * (-0.2057706e-1f + z * (0.2635537e-1f +
                  z * (-0.1647633e-1f + z * 0.392377e-2f)))))));
  const float large = __builtin_amdgcn_exp2f(e * 1.44269504088896341f) *
                      __builtin_amdgcn_rsqf(s) * P * p.inv_P0f;
  return (ax < 3.75f || !p.large_b0) ? small : large;
}

}  // namespace nufft3d
