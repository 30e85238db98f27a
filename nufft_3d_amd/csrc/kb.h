// Kaiser-Bessel weights evaluated on the fly (replaces the stored values of H).
//
// Restates @nufft_3d/kernel.m:2-34 on the kernel support, where the argument
// alpha*sqrt(1 - d2/(J/2)^2) is real and >= 0 (the imaginary continuation is
// only reached by deap, which runs on the host in double).
//   bessi0: Numerical-Recipes polynomial I0 (kernel.m:19-30)
//   value : bessi0(arg)/bessi0(alpha) (kernel.m:9)
// The reference stores every weight as single (nufft_3d.m:160,195,205):
// the fp64 path computes in double and rounds to float before widening, the
// fp32 path computes directly in float.
#pragma once
#include <hip/hip_runtime.h>

namespace nufft3d {

struct KBParams {
  double J2;         // J^2 (support test 4*d2 <= J^2, nufft_3d.m:184-185,189)
  double half2;      // (J/2)^2 (kernel.m:4)
  double alpha;
  double b0;         // bessi0(alpha)
  float J2f, inv_half2f, alphaf, inv_b0f;
  int radial;
};

__device__ __forceinline__ double bessi0_d(double ax) {
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

__device__ __forceinline__ float bessi0_f(float ax) {
  if (ax < 3.75f) {
    float y = ax * (1.0f / 3.75f);
    y = y * y;
    return 1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f +
                  y * (0.360768e-1f + y * 0.45813e-2f)))));
  }
  float y = 3.75f / ax;
  return (expf(ax) * rsqrtf(ax)) *
         (0.39894228f + y * (0.1328592e-1f + y * (0.225319e-2f + y * (-0.157565e-2f +
          y * (0.916281e-2f + y * (-0.2057706e-1f + y * (0.2635537e-1f +
          y * (-0.1647633e-1f + y * 0.392377e-2f))))))));
}

// kernel(dist2) in double, as kernel.m evaluates it
__device__ __forceinline__ double kb_d(double dist2, const KBParams& p) {
  double arg = p.alpha * sqrt(1.0 - dist2 / p.half2);
  return bessi0_d(arg) / p.b0;
}

__device__ __forceinline__ float kb_f(float dist2, const KBParams& p) {
  float t = fmaxf(1.0f - dist2 * p.inv_half2f, 0.0f);
  return bessi0_f(p.alphaf * sqrtf(t)) * p.inv_b0f;
}

// Weight of one candidate given its per-axis squared distances, or 0 when
// outside the support. Precision-generic entry point.
template <typename R>
__device__ __forceinline__ R kb_weight(R dx2, R dy2, R dz2, const KBParams& p);

template <>
__device__ __forceinline__ double kb_weight<double>(double dx2, double dy2, double dz2,
                                                    const KBParams& p) {
  if (p.radial) {
    double dist2 = dx2 + dy2 + dz2;                  // nufft_3d.m:184
    if (!(4.0 * dist2 <= p.J2)) return 0.0;          // nufft_3d.m:185
    return (double)(float)kb_d(dist2, p);            // nufft_3d.m:187,160,205
  }
  if (!(4.0 * dx2 <= p.J2) || !(4.0 * dy2 <= p.J2) || !(4.0 * dz2 <= p.J2)) return 0.0;
  return (double)(float)(kb_d(dx2, p) * kb_d(dy2, p) * kb_d(dz2, p));  // nufft_3d.m:191
}

template <>
__device__ __forceinline__ float kb_weight<float>(float dx2, float dy2, float dz2,
                                                  const KBParams& p) {
  if (p.radial) {
    float dist2 = dx2 + dy2 + dz2;
    if (!(4.0f * dist2 <= p.J2f)) return 0.0f;
    return kb_f(dist2, p);
  }
  if (!(4.0f * dx2 <= p.J2f) || !(4.0f * dy2 <= p.J2f) || !(4.0f * dz2 <= p.J2f)) return 0.0f;
  return kb_f(dx2, p) * kb_f(dy2, p) * kb_f(dz2, p);
}

}  // namespace nufft3d
