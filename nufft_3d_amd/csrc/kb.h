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
  double inv_half2;  // 1/(J/2)^2
  double inv_b0;     // 1/bessi0(alpha)
  double inv_P0;     // 1/P(3.75/alpha)
  double c375a;      // 3.75/alpha
  double qb;         // 1 - (3.75/alpha)^2: bessi0's branch point in q = dist2/(J/2)^2
                     // (-1 when alpha < 3.75: the polynomial branch throughout)
  double tscale;     // kKbIntervals / qb
  double cy;         // (alpha/3.75)^2
};

// Piecewise-polynomial table of the exponential branch of kernel(q):
// kKbIntervals equal intervals of q over [0, qb], a degree-kKbDeg polynomial
// in the local variable t in [-1, 1] on each (interpolated at Chebyshev
// nodes from an extended-precision evaluation of kernel.m's formula; within
// ~2e-16 relative for alpha <= 11, ~2e-15 at alpha = 14).
constexpr int kKbIntervals = 128;
constexpr int kKbDeg = 7;
constexpr int kKbCoef = kKbDeg + 1;

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

// kernel(dist2) in double from the table (tab: kKbIntervals x kKbCoef
// coefficients, lowest degree first) for fp64 arithmetic behind fp32 arrays:
// the exponential branch from the table, the polynomial branch (kernel.m:
// 19-23) directly, as a polynomial in y = (alpha/3.75)^2 (1 - q). Within a
// few ulps of kernel.m; rounded to float by the caller, as the reference's
// single weights are, it differs from kb_d's only where the double values
// straddle a float rounding boundary (a few parts in 1e6 of hits, by one
// float ulp).
__device__ __forceinline__ double kb_tab_d(double dist2, const double* tab, const KBParams& p) {
  const double q = dist2 * p.inv_half2;
  const double y = __builtin_fma(-p.cy, q, p.cy);
  const double small = (1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
                        y * (0.360768e-1 + y * 0.45813e-2)))))) * p.inv_b0;
  const double x = q * p.tscale;
  int i = (int)x;
  i = i < kKbIntervals - 1 ? i : kKbIntervals - 1;
  const double t = 2.0 * (x - (double)i) - 1.0;
  // coefficient-major table: coefficient k of interval i at k*kKbIntervals + i,
  // so lanes reading different intervals hit different LDS banks
  const double* c = tab + i;
  double g = c[kKbDeg * kKbIntervals];
#pragma unroll
  for (int k = kKbDeg - 1; k >= 0; --k) g = __builtin_fma(g, t, c[k * kKbIntervals]);
  return q > p.qb ? small : g;
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
