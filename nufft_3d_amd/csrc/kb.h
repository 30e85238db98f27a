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
  float tief;        // |4*d2 - J^2| below which the fp32 test is redone in double
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
// weights.
__device__ __forceinline__ float kb_f(float dist2, const KBParams& p) {
  const float q = dist2 * p.inv_half2f;
  const float t = fmaxf(1.0f - q, 0.0f);
  const float s = sqrtf(t);
  const float ax = p.alphaf * s;
  if (ax < 3.75f || !p.large_b0) return bessi0_f(ax) * p.inv_b0f;
  const float e = -p.alphaf * q / (1.0f + s);
  const float y = 3.75f / ax;
  const float P = 0.39894228f + y * (0.1328592e-1f + y * (0.225319e-2f + y * (-0.157565e-2f +
                  y * (0.916281e-2f + y * (-0.2057706e-1f + y * (0.2635537e-1f +
                  y * (-0.1647633e-1f + y * 0.392377e-2f)))))));
  return expf(e) * rsqrtf(s) * P * p.inv_P0f;
}

// Weight of one candidate given its per-axis squared distances, or 0 when
// outside the support. Precision-generic entry point.
template <typename R>
__device__ __forceinline__ R kb_weight(R dx2, R dy2, R dz2, const KBParams& p);

template <>
__device__ __forceinline__ double kb_weight<double>(double dx2, double dy2, double dz2,
                                                    const KBParams& p) {
#pragma clang fp contract(off)
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

// fp32 data path: the support test and the weight in float, except that a
// candidate within tief of the support edge is decided again in double from
// the exact offsets (dx - f), so the support set is the reference's one
// (nufft_3d.m:184-189) while the common case stays in float.
__device__ __forceinline__ bool edge_in_d(int d, double f, double J2) {
  const double e = (double)d - f;
  return 4.0 * (e * e) <= J2;
}

__device__ __forceinline__ float kb_weight_f32(float dx2, float dy2, float dz2, int dx, int dy,
                                               int dz, const double* __restrict__ f,
                                               const KBParams& p) {
  if (p.radial) {
    const float dist2 = dx2 + dy2 + dz2;
    const float q = 4.0f * dist2 - p.J2f;
    if (q > p.tief) return 0.0f;
    if (q >= -p.tief) {
      const double ex = (double)dx - f[0], ey = (double)dy - f[1], ez = (double)dz - f[2];
      const double d2 = ex * ex + ey * ey + ez * ez;
      if (!(4.0 * d2 <= p.J2)) return 0.0f;
    }
    return kb_f(dist2, p);
  }
  const float qx = 4.0f * dx2 - p.J2f, qy = 4.0f * dy2 - p.J2f, qz = 4.0f * dz2 - p.J2f;
  if (qx > p.tief || qy > p.tief || qz > p.tief) return 0.0f;
  if (qx >= -p.tief && !edge_in_d(dx, f[0], p.J2)) return 0.0f;
  if (qy >= -p.tief && !edge_in_d(dy, f[1], p.J2)) return 0.0f;
  if (qz >= -p.tief && !edge_in_d(dz, f[2], p.J2)) return 0.0f;
  return kb_f(dx2, p) * kb_f(dy2, p) * kb_f(dz2, p);
}

}  // namespace nufft3d
