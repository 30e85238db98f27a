// 3-D FFT on the oversampled grid with fused deapodization, zero-pad and
// crop: replaces fft3_pad / ifft3_crop (nufft_3d.m:285-302) and the U
// multiplies of fNUFT.m:5 / aNUFT.m:8.
//
// Geometry (per axis j, image index i, x = i - N_j/2): fft3_pad places image
// point x at grid index x mod K_j (symmetric pad followed by fftshift) and
// takes a forward DFT (e^{-2 pi i k x / K}); ifft3_crop is its exact adjoint
// (unnormalised inverse DFT, then the same index map read back).
//
// Passes (pruned: the padded zeros are never transformed):
//   forward : x-pass image[c][z][y][x] -> T1[z][y][kx][c]   (U fused)
//             y-pass T1 -> T2[z][ky][kx][c]
//             z-pass T2 -> G[kz][ky][kx][c]
//   adjoint : the reverse, cropping as early as possible, U and 1/sqrt(nnz)
//             (and the damp^2 x term of iprojection) fused into the x-pass.
// Each work-group transforms a bundle of B lines with a mixed-radix
// (16, 8, 4, 2, 3, 5, 7) Stockham autosort whose stages exchange data in LDS
// (element-major layout, line fastest); the first stage reads global memory
// and the last writes it wherever the access pattern stays coalesced.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <cmath>
#include <vector>

#include "nufft3d_internal.h"

namespace nufft3d {

struct Stages {
  int L;
  int n;
  int r[24];
};

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  cplx<R> o;
  o.x = a.x * b.x - a.y * b.y;
  o.y = a.x * b.y + a.y * b.x;
  return o;
}

template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { cplx<R> o; o.x = a.x + b.x; o.y = a.y + b.y; return o; }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { cplx<R> o; o.x = a.x - b.x; o.y = a.y - b.y; return o; }

// DFT of size RAD in registers: X[q] = sum_r v[r] e^{sign 2 pi i r q / RAD}
template <typename R, int RAD>
struct Bfly {
  static __device__ __forceinline__ void run(cplx<R>* v, int sign, const cplx<R>* tw, int L) {
    cplx<R> o[RAD];
    const int step = L / RAD;
#pragma unroll
    for (int q = 0; q < RAD; ++q) {
      cplx<R> acc = v[0];
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        cplx<R> w = tw[((r * q) % RAD) * step];
        if (sign > 0) w.y = -w.y;
        acc = cadd<R>(acc, cmul<R>(v[r], w));
      }
      o[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < RAD; ++q) v[q] = o[q];
  }
};

template <typename R>
struct Bfly<R, 2> {
  static __device__ __forceinline__ void run(cplx<R>* v, int, const cplx<R>*, int) {
    cplx<R> a = v[0], b = v[1];
    v[0] = cadd<R>(a, b);
    v[1] = csub<R>(a, b);
  }
};

template <typename R>
struct Bfly<R, 4> {
  static __device__ __forceinline__ void run(cplx<R>* v, int sign, const cplx<R>*, int) {
    cplx<R> t0 = cadd<R>(v[0], v[2]), t1 = csub<R>(v[0], v[2]);
    cplx<R> t2 = cadd<R>(v[1], v[3]), d = csub<R>(v[1], v[3]);
    cplx<R> t3;  // d * (sign * i)
    if (sign < 0) { t3.x = d.y; t3.y = -d.x; } else { t3.x = -d.y; t3.y = d.x; }
    v[0] = cadd<R>(t0, t2);
    v[2] = csub<R>(t0, t2);
    v[1] = cadd<R>(t1, t3);
    v[3] = csub<R>(t1, t3);
  }
};

template <typename R>
struct Bfly<R, 3> {
  static __device__ __forceinline__ void run(cplx<R>* v, int sign, const cplx<R>*, int) {
    const R c = (R)-0.5;
    const R s = (R)(sign < 0 ? -0.86602540378443864676 : 0.86602540378443864676);
    cplx<R> a = v[0], b = v[1], e = v[2];
    cplx<R> sum = cadd<R>(b, e), dif = csub<R>(b, e);
    cplx<R> m; m.x = a.x + c * sum.x; m.y = a.y + c * sum.y;
    cplx<R> j; j.x = -s * dif.y; j.y = s * dif.x;   // i*s*dif
    v[0] = cadd<R>(a, sum);
    v[1] = cadd<R>(m, j);
    v[2] = csub<R>(m, j);
  }
};

// Constant roots of unity e^{2 pi i k/16}: cos and sin of 2 pi k / 16
__device__ __constant__ const double kC16[10] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
    -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613, -1.0,
    -0.92387953251128675613};
__device__ __constant__ const double kS16[10] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613, 1.0,
    0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0,
    -0.38268343236508977173};

// v * e^{sign 2 pi i k / 16}
template <typename R>
__device__ __forceinline__ cplx<R> rot16(cplx<R> v, int k, int sign) {
  if (k == 0) return v;
  cplx<R> w;
  w.x = (R)kC16[k];
  w.y = (R)(sign < 0 ? -kS16[k] : kS16[k]);
  return cmul<R>(v, w);
}

// DFT of size P*Q as Q-point DFTs over r2 of v[r1 + P r2], twiddles
// w^{r1 q2} (w = e^{sign 2 pi i/(PQ)}), then P-point DFTs over r1:
// X[q2 + Q q1]. Used for 16 = 4x4 and 8 = 2x4.
template <typename R, int P, int Q>
__device__ __forceinline__ void bfly_composite(cplx<R>* v, int sign) {
  cplx<R> y[P][Q];
#pragma unroll
  for (int r1 = 0; r1 < P; ++r1) {
    cplx<R> t[Q];
#pragma unroll
    for (int r2 = 0; r2 < Q; ++r2) t[r2] = v[r1 + P * r2];
    Bfly<R, Q>::run(t, sign, nullptr, 0);
#pragma unroll
    for (int q2 = 0; q2 < Q; ++q2) y[r1][q2] = rot16<R>(t[q2], (16 / (P * Q)) * r1 * q2, sign);
  }
#pragma unroll
  for (int q2 = 0; q2 < Q; ++q2) {
    cplx<R> t[P];
#pragma unroll
    for (int r1 = 0; r1 < P; ++r1) t[r1] = y[r1][q2];
    Bfly<R, P>::run(t, sign, nullptr, 0);
#pragma unroll
    for (int q1 = 0; q1 < P; ++q1) v[q2 + Q * q1] = t[q1];
  }
}

template <typename R>
struct Bfly<R, 8> {
  static __device__ __forceinline__ void run(cplx<R>* v, int sign, const cplx<R>*, int) {
    bfly_composite<R, 2, 4>(v, sign);
  }
};

template <typename R>
struct Bfly<R, 16> {
  static __device__ __forceinline__ void run(cplx<R>* v, int sign, const cplx<R>*, int) {
    bfly_composite<R, 4, 4>(v, sign);
  }
};

// Elements held in registers per thread and stage (L*B <= 256*MaxE<R>)
template <typename R> struct MaxE { static constexpr int v = 16; };

// One Stockham stage over B lines of length L (Ns = product of the earlier
// radices). Every thread loads the inputs of its butterflies into registers
// - from LDS (element-major, a[pos*B + line]) or, for the first stage when
// FG, straight from global memory through ld(line, pos) - then twiddles,
// transforms and writes the outputs to their autosorted places - in LDS, or
// for the last stage when LG straight to global memory through
// sto(line, pos, value). Barriers: before reading LDS written by the previous
// stage, and between the reads and the in-place writes of one stage.
template <typename R, int RAD, bool FG, bool LG, typename LD, typename ST>
__device__ __forceinline__ void fft_stage(cplx<R>* buf, int L, int B, int Ns,
                                          const cplx<R>* __restrict__ tw, int sign, bool first,
                                          bool last, const LD& ld, const ST& sto) {
  constexpr int NB = (MaxE<R>::v + RAD - 1) / RAD;   // butterflies per thread (upper bound)
  const int Lr = L / RAD;
  const int tstep = L / (Ns * RAD);
  const int total = Lr * B;
  const bool gin = FG && first, gout = LG && last;
  cplx<R> v[NB][RAD];
  if (!first) __syncthreads();
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int idx = threadIdx.x + k * blockDim.x;
    if (idx < total) {
      const int line = idx % B;
      const int j = idx / B;
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        if (gin) v[k][r] = ld(line, j + r * Lr);
        else v[k][r] = buf[(j + r * Lr) * B + line];
      }
    }
  }
  if (!gin && !gout) __syncthreads();
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int idx = threadIdx.x + k * blockDim.x;
    if (idx < total) {
      const int line = idx % B;
      const int j = idx / B;
      const int kk = j % Ns;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          cplx<R> w = tw[kk * r * tstep];
          if (sign > 0) w.y = -w.y;
          v[k][r] = cmul<R>(v[k][r], w);
        }
      }
      Bfly<R, RAD>::run(v[k], sign, tw, L);
      const int idxD = (j / Ns) * Ns * RAD + kk;
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        if (gout) sto(line, idxD + r * Ns, v[k][r]);
        else buf[(idxD + r * Ns) * B + line] = v[k][r];
      }
    }
  }
}

// Mixed-radix Stockham FFT of B lines of length L (L*B <= blockDim.x*MaxE)
// for any stage sequence factorize() emits: 16s, one 8/4/2 (or 8 then 4),
// 3s, 5s, 7s. FG: the lines are first copied from ld(line, pos) into buf;
// otherwise the input is in buf, written and synchronised by the caller.
// LG: the result is finally copied out through sto(line, pos, v); otherwise
// it is left in buf (the caller synchronises before reading it). The LDS
// twiddles tw must be visible. Every stage works LDS to LDS in one loop per
// radix, so no code region holds more than one radix's registers (a single
// loop over a switch of radices, or global loads fused into the first and
// last stages, make the register allocation cover all radices at once).
template <typename R, bool FG, bool LG, typename LD, typename ST>
__device__ void fft_lines(cplx<R>* buf, int B, const Stages& st, const cplx<R>* __restrict__ tw,
                          int sign, const LD& ld, const ST& sto) {
  const int L = st.L;
  const int n = st.n;
  const int total = L * B;
  if constexpr (FG) {
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) buf[idx] = ld(idx % B, idx / B);
  }
  int Ns = 1, k = 0;
#define N3D_RUN(RAD, LOOP)                                                           \
  LOOP (k < n && st.r[k] == RAD) {                                                 \
    fft_stage<R, RAD, false, false>(buf, L, B, Ns, tw, sign, false, false, ld, sto); \
    Ns *= RAD;                                                                      \
    ++k;                                                                            \
  }
  N3D_RUN(16, while)
  N3D_RUN(8, if)
  N3D_RUN(4, if)
  N3D_RUN(2, if)
  N3D_RUN(3, while)
  N3D_RUN(5, while)
  N3D_RUN(7, while)
  N3D_RUN(11, while)
  N3D_RUN(13, while)
#undef N3D_RUN
  if constexpr (LG) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) sto(idx % B, idx / B, buf[idx]);
  }
}

// Stage plans compiled as straight-line code for the common grid sizes
// (a runtime loop over stages makes the register allocation cover every
// radix at once). Plan 0 is the runtime loop for any other size.
template <int PL> struct PlanRadix { static constexpr int r0 = 1, r1 = 1, r2 = 1; };
template <> struct PlanRadix<1> { static constexpr int r0 = 16, r1 = 16, r2 = 1; };  // 256
template <> struct PlanRadix<2> { static constexpr int r0 = 16, r1 = 8, r2 = 1; };   // 128
template <> struct PlanRadix<3> { static constexpr int r0 = 16, r1 = 8, r2 = 3; };   // 384
template <> struct PlanRadix<4> { static constexpr int r0 = 16, r1 = 4, r2 = 5; };   // 320
template <> struct PlanRadix<5> { static constexpr int r0 = 16, r1 = 4, r2 = 1; };   // 64
template <> struct PlanRadix<6> { static constexpr int r0 = 16, r1 = 8, r2 = 4; };   // 512
template <> struct PlanRadix<7> { static constexpr int r0 = 16, r1 = 4, r2 = 3; };   // 192
constexpr int kNumPlans = 8;

template <typename R, bool FG, bool LG, int PL, typename LD, typename ST>
__device__ __forceinline__ void fft_lines_pl(cplx<R>* buf, int B, const Stages& st,
                                             const cplx<R>* __restrict__ tw, int sign,
                                             const LD& ld, const ST& sto) {
  if constexpr (PL == 0) {
    fft_lines<R, FG, LG>(buf, B, st, tw, sign, ld, sto);
  } else {
    constexpr int r0 = PlanRadix<PL>::r0, r1 = PlanRadix<PL>::r1, r2 = PlanRadix<PL>::r2;
    constexpr int L = r0 * r1 * r2;
    fft_stage<R, r0, FG, LG>(buf, L, B, 1, tw, sign, true, false, ld, sto);
    fft_stage<R, r1, FG, LG>(buf, L, B, r0, tw, sign, false, r2 == 1, ld, sto);
    if constexpr (r2 > 1) fft_stage<R, r2, FG, LG>(buf, L, B, r0 * r1, tw, sign, false, true, ld, sto);
  }
}

// grid position of image index i on an axis of image size n, grid size L
__device__ __forceinline__ int padpos(int i, int n, int L) {
  int x = i - n / 2;
  return x < 0 ? x + L : x;
}

// image index at grid position p (inverse of padpos), or -1 in the padding
__device__ __forceinline__ int unpad(int p, int n, int L) {
  const int h = n / 2;
  if (p < n - h) return p + h;
  if (p >= L - h) return p - L + h;
  return -1;
}

template <typename R>
__device__ __forceinline__ void stage_twiddles(cplx<R>* stw, const cplx<R>* __restrict__ tw, int L) {
  for (int e = threadIdx.x; e < L; e += blockDim.x) stw[e] = tw[e];
}

// ---------------------------------------------------------------- x-pass
// Lines are l = yz*nc + c. A work-group owns lines [l0, l0+B); either B is a
// multiple of nc (whole coil sets, contiguous output block) or B divides nc.
struct XGeom {
  int N1, K1, nc, B;
  long long nyz;      // N2*N3 (image rows)
  long long nlines;   // nyz*nc
};

// Forward: image rows (x fastest) are staged through LDS with coalesced
// reads; the last stage writes [yz][kx][c] directly (lines = coils fastest).
template <typename R, int PL, typename IO = R, int BT = 0>
__global__ void __launch_bounds__(512) k_fft_x_fwd(const cplx<IO>* __restrict__ img,
                                                   const R* __restrict__ U,
                                                   cplx<R>* __restrict__ out, XGeom g, Stages st,
                                                   const cplx<R>* __restrict__ tw) {
  const int B = BT > 0 ? BT : g.B;   // lines per bundle (a constant in the common case)
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)g.K1 * B;
  stage_twiddles<R>(stw, tw, g.K1);
  const long long l0 = (long long)blockIdx.x * B;
  // a bundle of one image row's coils (B == nc): row blockIdx.x, coil = line
  const bool one = B == g.nc;
  // zeros at the padding positions only ([N1 - N1/2, K1 - N1/2) of each line)
  {
    const int p0 = g.N1 - g.N1 / 2;
    for (int e = threadIdx.x; e < (g.K1 - g.N1) * B; e += blockDim.x) {
      cplx<R> z; z.x = 0; z.y = 0;
      a[(p0 + e / B) * B + e % B] = z;
    }
  }
  for (int e = threadIdx.x; e < g.N1 * B; e += blockDim.x) {
    const int li = e / g.N1, x = e % g.N1;
    const long long l = l0 + li;
    if (l >= g.nlines) continue;
    const long long yz = one ? (long long)blockIdx.x : l / g.nc;
    const int c = one ? li : (int)(l % g.nc);
    const cplx<IO> vi = img[x + (long long)g.N1 * (yz + g.nyz * c)];
    const R u = U[x + (long long)g.N1 * yz];
    cplx<R> v;
    v.x = (R)vi.x * u; v.y = (R)vi.y * u;
    a[padpos(x, g.N1, g.K1) * B + li] = v;
  }
  __syncthreads();
  auto ld = [](int, int) { cplx<R> z; z.x = 0; z.y = 0; return z; };
  auto sto = [=](int li, int kx, cplx<R> v) {
    const long long l = l0 + li;
    if (l >= g.nlines) return;
    if (one) out[((long long)blockIdx.x * g.K1 + kx) * B + li] = v;
    else out[((l / g.nc) * g.K1 + kx) * g.nc + (l % g.nc)] = v;
  };
  fft_lines_pl<R, false, true, PL>(a, B, st, stw, -1, ld, sto);
}

// Adjoint: the first stage reads [yz][kx][c] directly; the result is cropped
// out of LDS with coalesced image-row writes (x fastest), with U, the scale
// and the optional add term. The crop loop steps (line, x) without integer
// divisions (fp64, K=320, 8 coils: 1.28 -> 0.90 ms).
template <typename R, int PL, typename IO = R, int BT = 0>
__global__ void __launch_bounds__(512) k_fft_x_adj(const cplx<R>* __restrict__ in,
                                                   const R* __restrict__ U,
                                                   cplx<IO>* __restrict__ img, XGeom g, Stages st,
                                                   const cplx<R>* __restrict__ tw, R scale,
                                                   const cplx<R>* __restrict__ add, R add_coef) {
  const int B = BT > 0 ? BT : g.B;   // lines per bundle (a constant in the common case)
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)g.K1 * B;
  stage_twiddles<R>(stw, tw, g.K1);
  __syncthreads();
  const long long l0 = (long long)blockIdx.x * B;
  // a bundle of one image row's coils (B == nc): row blockIdx.x, coil = line
  const bool one = B == g.nc;
  auto ld = [=](int li, int kx) {
    const long long l = l0 + li;
    cplx<R> v; v.x = 0; v.y = 0;
    if (one) v = in[((long long)blockIdx.x * g.K1 + kx) * B + li];
    else if (l < g.nlines) v = in[((l / g.nc) * g.K1 + kx) * g.nc + (l % g.nc)];
    return v;
  };
  auto sto = [](int, int, cplx<R>) {};
  fft_lines_pl<R, true, false, PL>(a, B, st, stw, +1, ld, sto);
  __syncthreads();
  const cplx<R>* r = a;
  // element (li, x) of the B image rows, stepped without divisions
  const int dli = blockDim.x / g.N1, dx = blockDim.x % g.N1;
  for (int li = threadIdx.x / g.N1, x = threadIdx.x % g.N1; li < B;) {
    const long long l = l0 + li;
    if (l >= g.nlines) break;
    const long long yz = one ? (long long)blockIdx.x : l / g.nc;
    const int c = one ? li : (int)(l % g.nc);
    cplx<R> v = r[padpos(x, g.N1, g.K1) * B + li];
    const R u = U[x + (long long)g.N1 * yz] * scale;
    v.x *= u; v.y *= u;
    const long long o = x + (long long)g.N1 * (yz + g.nyz * c);
    if (add) { const cplx<R> q = add[o]; v.x += add_coef * q.x; v.y += add_coef * q.y; }
    cplx<IO> vo;
    vo.x = (IO)v.x; vo.y = (IO)v.y;
    img[o] = vo;
    li += dli;
    x += dx;
    if (x >= g.N1) { x -= g.N1; ++li; }
  }
}

// ---------------------------------------------------------------- y/z pass
// in [outer][Lin][inner] -> out [outer][Lout][inner]; one work-group per
// (outer, bundle of B consecutive inner columns). The first stage reads the
// columns from global memory (zero in the padding) and the last stage writes
// them back (cropped rows skipped), so the only LDS traffic is between
// stages.
struct SGeom {
  int Lin, Lout, n_img;   // n_img: image size on this axis (pad/crop map)
  int mode;               // 0: pad (Lin = n_img, Lout = L), 1: crop (Lin = L, Lout = n_img),
                          // 2: plain (Lin = Lout = L, no index map)
  int B;
  long long inner;
  long long nbundle;      // ceil(inner / B)
};

template <typename R, int PL, int BT = 0>
__global__ void __launch_bounds__(1024) k_fft_strided(const cplx<R>* __restrict__ in,
                                                     cplx<R>* __restrict__ out, SGeom g,
                                                     Stages st, const cplx<R>* __restrict__ tw,
                                                     int sign) {
  const int B = BT > 0 ? BT : g.B;   // lines per bundle (a constant in the common case)
  extern __shared__ unsigned char smem_raw[];
  const int L = st.L;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)L * B;
  stage_twiddles<R>(stw, tw, L);
  __syncthreads();
  const long long outer = blockIdx.x / g.nbundle;
  const long long col0 = (blockIdx.x % g.nbundle) * B;
  const cplx<R>* src = in + outer * g.Lin * g.inner;
  cplx<R>* dst = out + outer * g.Lout * g.inner;
  auto ld = [=](int li, int p) {
    const long long col = col0 + li;
    const int i = g.mode == 0 ? unpad(p, g.n_img, L) : p;
    cplx<R> v; v.x = 0; v.y = 0;
    if (i >= 0 && col < g.inner) v = src[(long long)i * g.inner + col];
    return v;
  };
  auto sto = [=](int li, int p, cplx<R> v) {
    const long long col = col0 + li;
    const int i = g.mode == 1 ? unpad(p, g.n_img, L) : p;
    if (i >= 0 && col < g.inner) dst[(long long)i * g.inner + col] = v;
  };
  fft_lines_pl<R, true, true, PL>(a, B, st, stw, sign, ld, sto);
}

// Bluestein (chirp-z) transform of B lines of length L with a prime factor
// above 13, in the geometry of k_fft_strided (pad / crop / plain): with
// c[n] = exp(sign i pi n^2 / L), X[k] = c[k] sum_n (x[n] c[n]) conj(c[k-n]),
// a linear convolution evaluated as a circular one of the smooth length L2
// >= 2L - 1 by two L2-point transforms in LDS, with the conjugate chirp's
// spectrum (over L2) precomputed per sign.
template <typename R>
__global__ void __launch_bounds__(256) k_fft_bluestein(const cplx<R>* __restrict__ in,
                                                       cplx<R>* __restrict__ out, SGeom g,
                                                       Stages st2, const cplx<R>* __restrict__ tw2,
                                                       const cplx<R>* __restrict__ chirp,
                                                       const cplx<R>* __restrict__ bhat, int L,
                                                       int sign) {
  extern __shared__ unsigned char smem_raw[];
  const int L2 = st2.L;
  const int B = g.B;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)L2 * B;
  stage_twiddles<R>(stw, tw2, L2);
  const long long outer = blockIdx.x / g.nbundle;
  const long long col0 = (blockIdx.x % g.nbundle) * B;
  const cplx<R>* src = in + outer * g.Lin * g.inner;
  cplx<R>* dst = out + outer * g.Lout * g.inner;
  auto chirp_at = [&](int n) {
    cplx<R> c = chirp[n];
    if (sign > 0) c.y = -c.y;
    return c;
  };
  for (int idx = threadIdx.x; idx < L2 * B; idx += blockDim.x) {
    const int li = idx % B, n = idx / B;
    const long long col = col0 + li;
    cplx<R> v;
    v.x = 0;
    v.y = 0;
    if (n < L && col < g.inner) {
      const int i = g.mode == 0 ? unpad(n, g.n_img, L) : n;
      if (i >= 0) v = cmul<R>(src[(long long)i * g.inner + col], chirp_at(n));
    }
    a[idx] = v;
  }
  __syncthreads();
  auto none_ld = [](int, int) { cplx<R> z; z.x = 0; z.y = 0; return z; };
  auto none_st = [](int, int, cplx<R>) {};
  fft_lines<R, false, false>(a, B, st2, stw, -1, none_ld, none_st);
  __syncthreads();
  for (int idx = threadIdx.x; idx < L2 * B; idx += blockDim.x) a[idx] = cmul<R>(a[idx], bhat[idx / B]);
  __syncthreads();
  fft_lines<R, false, false>(a, B, st2, stw, +1, none_ld, none_st);
  __syncthreads();
  for (int idx = threadIdx.x; idx < L * B; idx += blockDim.x) {
    const int li = idx % B, k = idx / B;
    const long long col = col0 + li;
    const int i = g.mode == 1 ? unpad(k, g.n_img, L) : k;
    if (i >= 0 && col < g.inner) dst[(long long)i * g.inner + col] = cmul<R>(a[idx], chirp_at(k));
  }
}

// Layout changes of the transform path for plans with a Bluestein axis:
// image [nc][N3][N2][N1] times U <-> coil-fastest [N3][N2][N1][nc] (the
// layout the strided passes then take along x, y and z).
template <typename R, typename IO = R>
__global__ void k_img_to_cl(const cplx<IO>* __restrict__ img, const R* __restrict__ U,
                            cplx<R>* __restrict__ w, long long nvox, int nc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nvox * nc) return;
  const long long v = t / nc;
  const int c = (int)(t - v * nc);
  const cplx<IO> a = img[(long long)c * nvox + v];
  const R u = U[v];
  cplx<R> o;
  o.x = (R)a.x * u;
  o.y = (R)a.y * u;
  w[t] = o;
}

template <typename R, typename IO = R>
__global__ void k_cl_to_img(const cplx<R>* __restrict__ w, const R* __restrict__ U,
                            cplx<IO>* __restrict__ img, long long nvox, int nc, R scale,
                            const cplx<R>* __restrict__ add, R add_coef) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nvox * nc) return;
  const int c = (int)(t / nvox);
  const long long v = t - (long long)c * nvox;
  const cplx<R> a = w[v * nc + c];
  const R u = U[v] * scale;
  cplx<R> o;
  o.x = a.x * u;
  o.y = a.y * u;
  if (add) {
    o.x += add_coef * add[t].x;
    o.y += add_coef * add[t].y;
  }
  cplx<IO> oo;
  oo.x = (IO)o.x; oo.y = (IO)o.y;
  img[t] = oo;
}

// ---------------------------------------------------------------- host side
#define N3D_PLAN_SWITCH(ID, LAUNCH)     \
  switch (ID) {                         \
    case 1: LAUNCH(1); break;           \
    case 2: LAUNCH(2); break;           \
    case 3: LAUNCH(3); break;           \
    case 4: LAUNCH(4); break;           \
    case 5: LAUNCH(5); break;           \
    case 6: LAUNCH(6); break;           \
    case 7: LAUNCH(7); break;           \
    default: LAUNCH(0); break;          \
  }
// the same with the bundle width as a template constant when it is 8 (the
// usual fp64 bundle: index arithmetic by shifts and masks). Used by the x
// passes (fp64, K=320, 8 coils: 1.43 -> 1.35 ms forward, 1.28 -> 1.15 ms
// adjoint); the y and z passes measured 3-5% slower with it and keep the
// runtime width.
#define N3D_PLAN_SWITCH_B(ID, LAUNCH, BT)  \
  switch (ID) {                            \
    case 1: LAUNCH(1, BT); break;          \
    case 2: LAUNCH(2, BT); break;          \
    case 3: LAUNCH(3, BT); break;          \
    case 4: LAUNCH(4, BT); break;          \
    case 5: LAUNCH(5, BT); break;          \
    case 6: LAUNCH(6, BT); break;          \
    case 7: LAUNCH(7, BT); break;          \
    default: LAUNCH(0, BT); break;         \
  }
#define N3D_BUNDLE_SWITCH(B, ID, LAUNCH)                 \
  do {                                                   \
    if ((B) == 8) { N3D_PLAN_SWITCH_B(ID, LAUNCH, 8) }   \
    else { N3D_PLAN_SWITCH_B(ID, LAUNCH, 0) }            \
  } while (0)
// Radices 16 first (the fewest stages, hence LDS exchanges), the leftover
// power of two as one 2, 4 or 8 (2^(4k+1) as 8*4 instead of 16*2), then the
// odd primes 3, 5, 7, 11, 13 (11 and 13 as direct DFT butterflies in the
// runtime stage loop). Other lengths: n = -1 (Bluestein).
static void factorize(int L, Stages& st) {
  st.L = L;
  st.n = 0;
  int m = L, p2 = 0;
  while (m % 2 == 0) { m /= 2; ++p2; }
  int n16 = p2 / 4, rest = p2 % 4;
  const bool split = rest == 1 && n16 > 0;   // 2^(4k+1) = 16^(k-1) * 8 * 4
  if (split) --n16;
  for (int k = 0; k < n16; ++k) st.r[st.n++] = 16;
  if (split) { st.r[st.n++] = 8; st.r[st.n++] = 4; }
  else if (rest > 0) st.r[st.n++] = 1 << rest;
  const int primes[] = {3, 5, 7, 11, 13};
  for (int q : primes)
    while (m % q == 0) { st.r[st.n++] = q; m /= q; }
  if (m != 1) st.n = -1;
}

// compile-time plan matching a factorization (0: the runtime loop)
static int plan_id_of(const Stages& st) {
  const int r0 = st.n > 0 ? st.r[0] : 1, r1 = st.n > 1 ? st.r[1] : 1, r2 = st.n > 2 ? st.r[2] : 1;
  if (st.n < 2 || st.n > 3) return 0;
  const int ids[][3] = {{0, 0, 0}, {16, 16, 1}, {16, 8, 1}, {16, 8, 3}, {16, 4, 5}, {16, 4, 1},
                        {16, 8, 4}, {16, 4, 3}};
  for (int i = 1; i < kNumPlans; ++i)
    if (ids[i][0] == r0 && ids[i][1] == r1 && ids[i][2] == r2) return i;
  return 0;
}

// exp(sign i pi n^2 / L) in extended precision (n^2 reduced mod 2L exactly)
static void chirp_ld(int L, int sign, std::vector<long double>& re, std::vector<long double>& im,
                     int count) {
  const long double pi = 3.141592653589793238462643383279502884L;
  re.resize(count);
  im.resize(count);
  for (int n = 0; n < count; ++n) {
    const long long m = ((long long)n * n) % (2LL * L);
    const long double ang = sign * pi * (long double)m / (long double)L;
    re[n] = cosl(ang);
    im[n] = sinl(ang);
  }
}

// Bluestein set-up of an axis of length L: the convolution length L2, its
// roots of unity (d_tw), the chirp and the two conjugate-chirp spectra
static int setup_bluestein(nufft3d_plan* p, FftAxis& a, int L, Stages& st) {
  int L2 = 2 * L - 1;
  for (;; ++L2) {
    factorize(L2, st);
    if (st.n > 0) break;
  }
  a.bs = 1;
  a.L2 = L2;
  const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
  auto put = [&](std::vector<unsigned char>& h, int i, long double x, long double y) {
    if (p->prec == 2) reinterpret_cast<double2*>(h.data())[i] = make_double2((double)x, (double)y);
    else reinterpret_cast<float2*>(h.data())[i] = make_float2((float)x, (float)y);
  };
  std::vector<long double> cr, ci;
  chirp_ld(L, -1, cr, ci, L);
  std::vector<unsigned char> hc(esz * L);
  for (int n = 0; n < L; ++n) put(hc, n, cr[n], ci[n]);
  N3D_HIP(dev_malloc(&a.d_chirp, esz * L));
  N3D_HIP(hipMemcpy(a.d_chirp, hc.data(), esz * L, hipMemcpyHostToDevice));
  // b[m] = exp(-s i pi m^2 / L) on m in (-L, L), wrapped to length L2; its
  // L2-point DFT (e^{-2 pi i jk / L2}) over L2, by direct sums
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int si = 0; si < 2; ++si) {
    const int s = si == 0 ? -1 : +1;
    std::vector<long double> br(L2, 0), bi(L2, 0), er, ei;
    chirp_ld(L, -s, er, ei, L);
    for (int m = 0; m < L; ++m) {
      br[m] = er[m];
      bi[m] = ei[m];
      if (m > 0) { br[L2 - m] = er[m]; bi[L2 - m] = ei[m]; }
    }
    std::vector<long double> tc(L2), ts(L2);
    for (int t = 0; t < L2; ++t) {
      tc[t] = cosl(2 * pi * t / L2);
      ts[t] = -sinl(2 * pi * t / L2);
    }
    std::vector<unsigned char> hb(esz * L2);
    for (int k = 0; k < L2; ++k) {
      long double xr = 0, xi = 0;
      for (int j = 0; j < L2; ++j) {
        if (br[j] == 0 && bi[j] == 0) continue;
        const int t = (int)(((long long)j * k) % L2);
        xr += br[j] * tc[t] - bi[j] * ts[t];
        xi += br[j] * ts[t] + bi[j] * tc[t];
      }
      put(hb, k, xr / L2, xi / L2);
    }
    N3D_HIP(dev_malloc(&a.d_bhat[si], esz * L2));
    N3D_HIP(hipMemcpy(a.d_bhat[si], hb.data(), esz * L2, hipMemcpyHostToDevice));
  }
  return NUFFT3D_OK;
}

int fft_setup_axis(nufft3d_plan* p, int axis) {
  const int L = p->K[axis];
  FftAxis& a = p->ax[axis];
  a.L = L;
  Stages st;
  factorize(L, st);
  if (st.n < 0) N3D_TRY(setup_bluestein(p, a, L, st));
  const int Ltw = a.bs ? a.L2 : L;   // length of the staged transforms
  a.nstage = st.n;
  for (int i = 0; i < st.n; ++i) a.radix[i] = st.r[i];
  a.plan_id = plan_id_of(st);
  const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
  std::vector<unsigned char> host(esz * Ltw);
  // roots of unity in extended precision (x86 long double), rounded once:
  // the angle 2 pi t / L formed in double is off by up to ~6 ulps, and the
  // fp64 plans' transforms feed the deapodization, whose range (up to 2e7
  // at u = 1.25, J = 6) amplifies any rounding at the field-of-view edge
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < Ltw; ++t) {
    double c, s;
    if ((8LL * t) % Ltw == 0) {   // multiples of pi/4: exact values
      static const double c8[8] = {1, 0.70710678118654752440, 0, -0.70710678118654752440, -1,
                                   -0.70710678118654752440, 0, 0.70710678118654752440};
      const int o = (int)(8LL * t / Ltw);
      c = c8[o];
      s = -c8[(o + 6) % 8];   // sin(-x) = -cos(x - pi/2)
    } else {
      const long double ang = -two_pi * (long double)t / (long double)Ltw;
      c = (double)cosl(ang);
      s = (double)sinl(ang);
    }
    if (p->prec == 2) {
      double2 v; v.x = c; v.y = s;
      reinterpret_cast<double2*>(host.data())[t] = v;
    } else {
      float2 v; v.x = (float)c; v.y = (float)s;
      reinterpret_cast<float2*>(host.data())[t] = v;
    }
  }
  N3D_HIP(dev_malloc(&a.d_tw, esz * Ltw));
  N3D_HIP(hipMemcpy(a.d_tw, host.data(), esz * Ltw, hipMemcpyHostToDevice));
  return NUFFT3D_OK;
}

static Stages stages_of(const FftAxis& a) {
  Stages st;
  st.L = a.bs ? a.L2 : a.L;
  st.n = a.nstage;
  for (int i = 0; i < a.nstage; ++i) st.r[i] = a.radix[i];
  return st;
}

// lines per LDS bundle: L*B elements staged in registers by 256 threads
// (at most MaxE each) and held in one LDS buffer of at most 64 KiB
// `want` = requested lines per work-group (0 = automatic), from the plan
static int bundle_for(int L, size_t esz, int want, int threads, size_t lds) {
  int B = threads * (esz == sizeof(double2) ? MaxE<double>::v : MaxE<float>::v) / L;
  const int bl = (int)(lds / ((size_t)L * esz)) - 1;   // the L twiddles share the buffer
  if (bl < B) B = bl;
  if (want > 0 && want <= B) {
    int b = 1;
    while (b * 2 <= want) b *= 2;
    return b;
  }
  if (B >= 64 && threads > 512) return 64;
  if (B >= 32 && threads > 256) return 32;
  if (B >= 16) return 16;
  if (B >= 8) return 8;
  if (B >= 4) return 4;
  if (B >= 2) return 2;
  return 1;
}

// dynamic LDS allowed to the z-pass work-groups (above the 64 KiB default)
constexpr size_t kStridedLds = 160 * 1024;

// threads per z-pass work-group (planned axes, measured at K=256, 16
// coils): the forward pass reads half the rows it writes and runs best with
// 1024 threads and 64-line bundles (512-byte row pieces: 4.1 -> 4.8 TB/s),
// the adjoint with 512 threads and 32-line bundles (4.3 TB/s; 1024: 3.9).
// Fewer where the lines are too long for those bundles; 256 for the generic
// stage loop. fft_zthreads overrides both (256, 512 or 1024).
static int z_threads(const nufft3d_plan* p, size_t esz, bool fwd) {
  if (p->fft_zthreads == 256 || p->fft_zthreads == 512 || p->fft_zthreads == 1024)
    return p->fft_zthreads;
  if (p->ax[2].plan_id == 0) return 256;
  if (fwd && bundle_for(p->K[2], esz, p->fft_bundle, 1024, kStridedLds) >= 64) return 1024;
  return bundle_for(p->K[2], esz, p->fft_bundle, 512, kStridedLds) >= 32 ? 512 : 256;
}

// raise a kernel's dynamic LDS limit once when a launch needs more than 64 KiB
// (one flag per kernel instantiation: instantiations share a function type)
template <typename K>
static void allow_lds(K* kernel, size_t bytes) {
  static std::mutex mu;
  static std::vector<const void*> granted;
  if (bytes <= 65536) return;
  const void* f = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(mu);
  if (std::find(granted.begin(), granted.end(), f) != granted.end()) return;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStridedLds) ==
      hipSuccess)
    granted.push_back(f);
  (void)hipGetLastError();
}

// threads per x-pass work-group: 512 on axes with a compiled stage plan
// (K=256, 16 coils: 3.3 -> 3.6 TB/s forward, 3.5 -> 3.9 adjoint, the same
// bundles served by twice the threads); 256 for the generic stage loop
// (slower with 512). fft_xthreads overrides (256 or 512).
static int x_threads(const nufft3d_plan* p) {
  if (p->fft_xthreads == 256 || p->fft_xthreads == 512) return p->fft_xthreads;
  return p->ax[0].plan_id != 0 ? 512 : 256;
}

static int bundle_x(int L, size_t esz, int nc, int want, int threads = 256) {
  const int Bmax = bundle_for(L, esz, want, threads, 65536);
  if (nc <= Bmax) return nc * (Bmax / nc);
  for (int B = Bmax; B >= 1; --B)
    if (nc % B == 0) return B;
  return 1;
}

// One pass of B-line bundles along an axis for the transform path of plans
// with a Bluestein axis (and the plain image-size transforms): the compiled
// or runtime stage loop for a smooth length, k_fft_bluestein otherwise.
template <typename R>
static int strided_pass(nufft3d_plan* p, const FftAxis& a, const void* in, void* out, SGeom g,
                        long long outer, int sign, const char* name) {
  const size_t esz = sizeof(cplx<R>);
  ProfScope ps(p, name);
  if (a.bs) {
    g.B = bundle_for(a.L2, esz, p->fft_bundle, 256, kStridedLds);
    if (g.B < 1) {
      set_error("transform length " + std::to_string(a.L) + " needs a Bluestein length " +
                std::to_string(a.L2) + " beyond the work-group's LDS");
      return NUFFT3D_ERR_UNSUPPORTED;
    }
    if (g.B > g.inner) g.B = (int)g.inner;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)a.L2 * (g.B + 1) * esz;
    allow_lds(k_fft_bluestein<R>, sh);
    hipLaunchKernelGGL(k_fft_bluestein<R>, dim3((unsigned)(g.nbundle * outer)), dim3(256), sh,
                       p->stream, (const cplx<R>*)in, (cplx<R>*)out, g, stages_of(a),
                       (const cplx<R>*)a.d_tw, (const cplx<R>*)a.d_chirp,
                       (const cplx<R>*)a.d_bhat[sign < 0 ? 0 : 1], a.L, sign);
  } else {
    g.B = bundle_for(a.L, esz, p->fft_bundle, 256, 65536);
    if (g.B > g.inner) g.B = (int)g.inner;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)a.L * (g.B + 1) * esz;
#define N3D_L(PL) hipLaunchKernelGGL((k_fft_strided<R, PL>), dim3((unsigned)(g.nbundle * outer)), \
                                    dim3(256), sh, p->stream, (const cplx<R>*)in, (cplx<R>*)out,  \
                                    g, stages_of(a), (const cplx<R>*)a.d_tw, sign)
    N3D_PLAN_SWITCH(a.plan_id, N3D_L);
#undef N3D_L
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

static bool has_bluestein(const nufft3d_plan* p) {
  return p->ax[0].bs || p->ax[1].bs || p->ax[2].bs;
}

// Transform path of plans with a Bluestein axis: U and the coil-fastest
// layout first, then pad + transform along x, y, z (forward), or the reverse
// with crops and U, scale and the add term last (adjoint). T0 (the
// coil-fastest image) shares t2's buffer, which it precedes (forward) or
// follows (adjoint).
template <typename R, typename IO>
static int fft_forward_generic(nufft3d_plan* p, const void* img, int nc, void* grid, void* t1buf,
                               void* t2buf) {
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const size_t esz = sizeof(cplx<R>);
  const long long nvox = (long long)N1 * N2 * N3;
  const size_t t1 = (size_t)K1 * N2 * N3 * nc;
  const size_t t2 = std::max((size_t)K1 * K2 * N3, (size_t)nvox) * nc;
  void* T1 = t1buf;
  void* T2 = t2buf;
  if (!T1) { N3D_TRY(ensure(p->ws_t1, t1 * esz)); T1 = p->ws_t1.p; }
  if (!T2) { N3D_TRY(ensure(p->ws_t2, t2 * esz)); T2 = p->ws_t2.p; }
  {
    ProfScope ps(p, "k_img_to_cl");
    const long long n = nvox * nc;
    hipLaunchKernelGGL((k_img_to_cl<R, IO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       p->stream, (const cplx<IO>*)img, (const R*)p->d_U, (cplx<R>*)T2, nvox, nc);
  }
  SGeom g;
  g.mode = 0;
  g.Lin = N1; g.Lout = K1; g.n_img = N1; g.inner = nc;
  N3D_TRY(strided_pass<R>(p, p->ax[0], T2, T1, g, (long long)N2 * N3, -1, "k_fft_x_fwd"));
  g.Lin = N2; g.Lout = K2; g.n_img = N2; g.inner = (long long)K1 * nc;
  N3D_TRY(strided_pass<R>(p, p->ax[1], T1, K3 > 1 ? T2 : grid, g, N3, -1, "k_fft_y_fwd"));
  if (K3 > 1) {
    g.Lin = N3; g.Lout = K3; g.n_img = N3; g.inner = (long long)K2 * K1 * nc;
    N3D_TRY(strided_pass<R>(p, p->ax[2], T2, grid, g, 1, -1, "k_fft_z_fwd"));
  }
  return NUFFT3D_OK;
}

template <typename R, typename IO>
static int fft_adjoint_generic(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                               const void* add, double add_coef) {
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const size_t esz = sizeof(cplx<R>);
  const long long nvox = (long long)N1 * N2 * N3;
  const size_t t1 = (size_t)K1 * N2 * N3 * nc;
  const size_t t2 = std::max((size_t)K1 * K2 * N3, (size_t)nvox) * nc;
  N3D_TRY(ensure(p->ws_t1, t1 * esz));
  N3D_TRY(ensure(p->ws_t2, t2 * esz));
  void* T1 = p->ws_t1.p;
  void* T2 = p->ws_t2.p;
  SGeom g;
  g.mode = 1;
  const void* yin = grid;
  if (K3 > 1) {
    g.Lin = K3; g.Lout = N3; g.n_img = N3; g.inner = (long long)K2 * K1 * nc;
    N3D_TRY(strided_pass<R>(p, p->ax[2], grid, T2, g, 1, +1, "k_fft_z_adj"));
    yin = T2;
  }
  g.Lin = K2; g.Lout = N2; g.n_img = N2; g.inner = (long long)K1 * nc;
  N3D_TRY(strided_pass<R>(p, p->ax[1], yin, T1, g, N3, +1, "k_fft_y_adj"));
  g.Lin = K1; g.Lout = N1; g.n_img = N1; g.inner = nc;
  N3D_TRY(strided_pass<R>(p, p->ax[0], T1, T2, g, (long long)N2 * N3, +1, "k_fft_x_adj"));
  {
    ProfScope ps(p, "k_cl_to_img");
    const long long n = nvox * nc;
    hipLaunchKernelGGL((k_cl_to_img<R, IO>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       p->stream, (const cplx<R>*)T2, (const R*)p->d_U, (cplx<IO>*)img, nvox, nc,
                       (R)scale, (const cplx<R>*)add, (R)add_coef);
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R, typename IO = R>
static int fft_forward_t(nufft3d_plan* p, const void* img, int nc, void* grid, void* t1buf,
                         void* t2buf) {
  if (has_bluestein(p)) return fft_forward_generic<R, IO>(p, img, nc, grid, t1buf, t2buf);
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const size_t esz = sizeof(cplx<R>);
  const size_t t1 = (size_t)K1 * N2 * N3 * nc, t2 = (size_t)K1 * K2 * N3 * nc;
  void* T1 = t1buf;
  if (!T1) {
    N3D_TRY(ensure(p->ws_t1, t1 * esz));
    T1 = p->ws_t1.p;
  }
  void* T2 = grid;
  if (K3 > 1) {
    T2 = t2buf;
    if (!T2) {
      N3D_TRY(ensure(p->ws_t2, t2 * esz));
      T2 = p->ws_t2.p;
    }
  }
  {  // x-pass
    XGeom g;
    g.N1 = N1; g.K1 = K1; g.nc = nc;
    const int xt = x_threads(p);
    g.B = bundle_x(K1, esz, nc, p->fft_bundle, xt);
    g.nyz = (long long)N2 * N3;
    g.nlines = g.nyz * nc;
    const long long nb = (g.nlines + g.B - 1) / g.B;
    const size_t sh = (size_t)K1 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_x_fwd");
#define N3D_L(PL, BT) hipLaunchKernelGGL((k_fft_x_fwd<R, PL, IO, BT>), dim3((unsigned)nb), dim3(xt), \
                                        sh, p->stream, (const cplx<IO>*)img, (const R*)p->d_U,  \
                                        (cplx<R>*)T1, g, stages_of(p->ax[0]),                   \
                                        (const cplx<R>*)p->ax[0].d_tw)
    N3D_BUNDLE_SWITCH(g.B, p->ax[0].plan_id, N3D_L);
#undef N3D_L
  }
  {  // y-pass: [N3][N2][K1*nc] -> [N3][K2][K1*nc]
    SGeom g;
    g.Lin = N2; g.Lout = K2; g.n_img = N2; g.mode = 0;
    const int yt = p->fft_ythreads == 512 ? 512 : 256;
    g.B = bundle_for(K2, esz, p->fft_bundle, yt, 65536);
    g.inner = (long long)K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K2 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_y_fwd");
#define N3D_L(PL, BT) hipLaunchKernelGGL((k_fft_strided<R, PL, BT>), dim3((unsigned)(g.nbundle * N3)), \
                                        dim3(yt), sh, p->stream, (const cplx<R>*)T1, (cplx<R>*)T2, \
                                        g, stages_of(p->ax[1]), (const cplx<R>*)p->ax[1].d_tw, -1)
    N3D_PLAN_SWITCH_B(p->ax[1].plan_id, N3D_L, 0);
#undef N3D_L
  }
  if (K3 > 1) {  // z-pass: [N3][K2*K1*nc] -> [K3][K2*K1*nc]
    SGeom g;
    g.Lin = N3; g.Lout = K3; g.n_img = N3; g.mode = 0;
    const int zt = z_threads(p, esz, true);
    g.B = bundle_for(K3, esz, p->fft_bundle, zt, kStridedLds);
    g.inner = (long long)K2 * K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K3 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_z_fwd");
#define N3D_L(PL, BT) allow_lds(k_fft_strided<R, PL, BT>, sh);                                   \
                  hipLaunchKernelGGL((k_fft_strided<R, PL, BT>), dim3((unsigned)g.nbundle),       \
                                    dim3(zt), sh, p->stream, (const cplx<R>*)T2, (cplx<R>*)grid, \
                                    g, stages_of(p->ax[2]), (const cplx<R>*)p->ax[2].d_tw, -1)
    N3D_PLAN_SWITCH_B(p->ax[2].plan_id, N3D_L, 0);
#undef N3D_L
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R, typename IO = R>
static int fft_adjoint_t(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                         const void* add, double add_coef) {
  if (has_bluestein(p)) return fft_adjoint_generic<R, IO>(p, grid, nc, img, scale, add, add_coef);
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const size_t esz = sizeof(cplx<R>);
  const size_t t1 = (size_t)K1 * N2 * N3 * nc, t2 = (size_t)K1 * K2 * N3 * nc;
  N3D_TRY(ensure(p->ws_t1, t1 * esz));
  void* T1 = p->ws_t1.p;
  const void* T2 = grid;
  if (K3 > 1) {  // z-pass: [K3][K2*K1*nc] -> [N3][K2*K1*nc]
    N3D_TRY(ensure(p->ws_t2, t2 * esz));
    SGeom g;
    g.Lin = K3; g.Lout = N3; g.n_img = N3; g.mode = 1;
    const int zt = z_threads(p, esz, false);
    g.B = bundle_for(K3, esz, p->fft_bundle, zt, kStridedLds);
    g.inner = (long long)K2 * K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K3 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_z_adj");
#define N3D_L(PL, BT) allow_lds(k_fft_strided<R, PL, BT>, sh);                                   \
                  hipLaunchKernelGGL((k_fft_strided<R, PL, BT>), dim3((unsigned)g.nbundle),       \
                                    dim3(zt), sh, p->stream, (const cplx<R>*)grid,               \
                                    (cplx<R>*)p->ws_t2.p, g, stages_of(p->ax[2]),                \
                                    (const cplx<R>*)p->ax[2].d_tw, +1)
    N3D_PLAN_SWITCH_B(p->ax[2].plan_id, N3D_L, 0);
#undef N3D_L
    T2 = p->ws_t2.p;
  }
  {  // y-pass: [N3][K2][K1*nc] -> [N3][N2][K1*nc]
    SGeom g;
    g.Lin = K2; g.Lout = N2; g.n_img = N2; g.mode = 1;
    const int yt = p->fft_ythreads == 512 ? 512 : 256;
    g.B = bundle_for(K2, esz, p->fft_bundle, yt, 65536);
    g.inner = (long long)K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K2 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_y_adj");
#define N3D_L(PL, BT) hipLaunchKernelGGL((k_fft_strided<R, PL, BT>), dim3((unsigned)(g.nbundle * N3)), \
                                        dim3(yt), sh, p->stream, (const cplx<R>*)T2, (cplx<R>*)T1, \
                                        g, stages_of(p->ax[1]), (const cplx<R>*)p->ax[1].d_tw, +1)
    N3D_PLAN_SWITCH_B(p->ax[1].plan_id, N3D_L, 0);
#undef N3D_L
  }
  {  // x-pass with crop, U, scale and the optional add term
    XGeom g;
    g.N1 = N1; g.K1 = K1; g.nc = nc;
    const int xt = x_threads(p);
    g.B = bundle_x(K1, esz, nc, p->fft_bundle, xt);
    g.nyz = (long long)N2 * N3;
    g.nlines = g.nyz * nc;
    const long long nb = (g.nlines + g.B - 1) / g.B;
    const size_t sh = (size_t)K1 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_x_adj");
#define N3D_L(PL, BT) hipLaunchKernelGGL((k_fft_x_adj<R, PL, IO, BT>), dim3((unsigned)nb), dim3(xt), \
                                        sh, p->stream, (const cplx<R>*)T1, (const R*)p->d_U,   \
                                        (cplx<IO>*)img, g, stages_of(p->ax[0]),                \
                                        (const cplx<R>*)p->ax[0].d_tw, (R)scale,               \
                                        (const cplx<R>*)add, (R)add_coef)
    N3D_BUNDLE_SWITCH(g.B, p->ax[0].plan_id, N3D_L);
#undef N3D_L
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

// Unshifted 3-D DFT of an image-size array [nc][N3][N2][N1] (out of place,
// sign -1), as fft(fft(fft(x,[],1),[],2),[],3) of iNUFT.m:204.
template <typename R>
static int fft_plain_t(nufft3d_plan* p, const void* in, int nc, void* out) {
  const size_t esz = sizeof(cplx<R>);
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  for (int a = 0; a < 3; ++a)
    if (p->N[a] > 1 && !p->axN[a].d_tw) {
      const int keep = p->K[a];
      p->K[a] = p->N[a];
      FftAxis save = p->ax[a];
      p->ax[a] = FftAxis();
      const int st = fft_setup_axis(p, a);
      p->axN[a] = p->ax[a];
      p->ax[a] = save;
      p->K[a] = keep;
      if (st) return st;
    }
  const size_t total = (size_t)N1 * N2 * N3 * nc;
  N3D_TRY(ensure(p->ws_t1, total * esz));
  const void* src = in;
  void* bufs[2] = {p->ws_t1.p, out};
  int which = 0;
  // after three passes the result must land in out: order the ping-pong so
  const int npass = (N1 > 1) + (N2 > 1) + (N3 > 1);
  which = npass % 2 == 1 ? 1 : 0;
  struct AxisPass { int n; long long outer, inner; int a; } passes[3] = {
      {N1, (long long)nc * N3 * N2, 1LL, 0},
      {N2, (long long)nc * N3, (long long)N1, 1},
      {N3, (long long)nc, (long long)N1 * N2, 2}};
  for (auto& ps : passes) {
    if (ps.n <= 1) continue;
    void* dst = bufs[which];
    if (p->axN[ps.a].bs) {   // Bluestein length
      SGeom gb;
      gb.Lin = ps.n; gb.Lout = ps.n; gb.n_img = ps.n; gb.mode = 2; gb.inner = ps.inner;
      N3D_TRY(strided_pass<R>(p, p->axN[ps.a], src, dst, gb, ps.outer, -1, "k_fft_plain"));
      src = dst;
      which ^= 1;
      continue;
    }
    SGeom g;
    g.Lin = ps.n; g.Lout = ps.n; g.n_img = ps.n; g.mode = 2;
    g.B = ps.inner >= 16 ? bundle_for(ps.n, esz, p->fft_bundle, 256, 65536) : 1;
    g.inner = ps.inner;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)ps.n * (g.B + 1) * esz;
    ProfScope prs(p, "k_fft_plain");
#define N3D_L(PL) hipLaunchKernelGGL((k_fft_strided<R, PL>), dim3((unsigned)(g.nbundle * ps.outer)), \
                                    dim3(256), sh, p->stream, (const cplx<R>*)src, (cplx<R>*)dst,  \
                                    g, stages_of(p->axN[ps.a]), (const cplx<R>*)p->axN[ps.a].d_tw, -1)
    N3D_PLAN_SWITCH(p->axN[ps.a].plan_id, N3D_L);
#undef N3D_L
    src = dst;
    which ^= 1;
  }
  if (npass == 0)
    N3D_HIP(hipMemcpyAsync(out, in, total * esz, hipMemcpyDeviceToDevice, p->stream));
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int fft_plain(nufft3d_plan* p, const void* in, int nc, void* out) {
  return p->prec == 2 ? fft_plain_t<double>(p, in, nc, out) : fft_plain_t<float>(p, in, nc, out);
}

int fft_forward(nufft3d_plan* p, const void* img, int nc, void* grid, void* t1, void* t2, int ep) {
  if (ep == 1 && p->prec == 2) return fft_forward_t<double, float>(p, img, nc, grid, t1, t2);
  return p->prec == 2 ? fft_forward_t<double>(p, img, nc, grid, t1, t2)
                      : fft_forward_t<float>(p, img, nc, grid, t1, t2);
}

int fft_adjoint(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                const void* add, double add_coef, int ep) {
  if (ep == 1 && p->prec == 2) {
    if (add) {
      set_error("an fp32 image out of an fp64 transform takes no add term");
      return NUFFT3D_ERR_INVALID;
    }
    return fft_adjoint_t<double, float>(p, grid, nc, img, scale, nullptr, 0.0);
  }
  return p->prec == 2 ? fft_adjoint_t<double>(p, grid, nc, img, scale, add, add_coef)
                      : fft_adjoint_t<float>(p, grid, nc, img, scale, add, add_coef);
}

}  // namespace nufft3d
