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
// Each work-group transforms a bundle of B lines staged in LDS with
// element-major layout (line fastest) and runs a mixed-radix (4,2,3,5,7,..)
// Stockham autosort between two LDS buffers.
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

// Elements held in registers per thread and stage (L*B <= 256*MaxE<R>)
template <typename R> struct MaxE { static constexpr int v = sizeof(R) == 8 ? 16 : 32; };

// One in-place Stockham stage (Ns = product of the earlier radices): every
// thread first loads the inputs of its butterflies into registers, the block
// synchronises, then the outputs are written to their autosorted places.
template <typename R, int RAD>
__device__ __forceinline__ void stockham_stage(cplx<R>* buf, int L, int B, int Ns,
                                               const cplx<R>* __restrict__ tw, int sign) {
  constexpr int NB = (MaxE<R>::v + RAD - 1) / RAD;   // butterflies per thread (upper bound)
  const int Lr = L / RAD;
  const int tstep = L / (Ns * RAD);
  const int total = Lr * B;
  cplx<R> v[NB][RAD];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int idx = threadIdx.x + k * blockDim.x;
    if (idx < total) {
      const int line = idx % B;
      const int j = idx / B;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[k][r] = buf[(j + r * Lr) * B + line];
    }
  }
  __syncthreads();
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
      for (int r = 0; r < RAD; ++r) buf[(idxD + r * Ns) * B + line] = v[k][r];
    }
  }
  __syncthreads();
}

// In-LDS in-place Stockham FFT of B lines of length L stored a[i*B + line]
// (L*B <= blockDim.x * MaxE).
template <typename R>
__device__ void lds_fft(cplx<R>* a, int B, const Stages& st, const cplx<R>* __restrict__ tw,
                        int sign) {
  const int L = st.L;
  int Ns = 1;
  for (int k = 0; k < st.n; ++k) {
    const int rad = st.r[k];
    switch (rad) {
      case 4: stockham_stage<R, 4>(a, L, B, Ns, tw, sign); break;
      case 2: stockham_stage<R, 2>(a, L, B, Ns, tw, sign); break;
      case 3: stockham_stage<R, 3>(a, L, B, Ns, tw, sign); break;
      case 5: stockham_stage<R, 5>(a, L, B, Ns, tw, sign); break;
      default: stockham_stage<R, 7>(a, L, B, Ns, tw, sign); break;
    }
    Ns *= rad;
  }
}

// grid position of image index i on an axis of image size n, grid size L
__device__ __forceinline__ int padpos(int i, int n, int L) {
  int x = i - n / 2;
  return x < 0 ? x + L : x;
}

// ---------------------------------------------------------------- x-pass
// Lines are l = yz*nc + c. A work-group owns lines [l0, l0+B); either B is a
// multiple of nc (whole coil sets, contiguous output block) or B divides nc.
struct XGeom {
  int N1, K1, nc, B;
  long long nyz;      // N2*N3 (image rows)
  long long nlines;   // nyz*nc
};

template <typename R>
__device__ __forceinline__ void xline_of(int e, const XGeom& g, int& li, int& kx) {
  if (g.B % g.nc == 0) {
    const int per = g.K1 * g.nc;
    const int j = e / per, rem = e % per;
    kx = rem / g.nc;
    li = j * g.nc + rem % g.nc;
  } else {
    kx = e / g.B;
    li = e % g.B;
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_x_fwd(const cplx<R>* __restrict__ img,
                                                   const R* __restrict__ U,
                                                   cplx<R>* __restrict__ out, XGeom g, Stages st,
                                                   const cplx<R>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)g.K1 * g.B;
  for (int e = threadIdx.x; e < g.K1; e += blockDim.x) stw[e] = tw[e];
  const long long l0 = (long long)blockIdx.x * g.B;
  for (int e = threadIdx.x; e < g.K1 * g.B; e += blockDim.x) { a[e].x = 0; a[e].y = 0; }
  __syncthreads();
  for (int e = threadIdx.x; e < g.N1 * g.B; e += blockDim.x) {
    const int li = e / g.N1, x = e % g.N1;
    const long long l = l0 + li;
    if (l >= g.nlines) continue;
    const long long yz = l / g.nc;
    const int c = (int)(l % g.nc);
    cplx<R> v = img[x + (long long)g.N1 * (yz + g.nyz * c)];
    const R u = U[x + (long long)g.N1 * yz];
    v.x *= u; v.y *= u;
    a[padpos(x, g.N1, g.K1) * g.B + li] = v;
  }
  __syncthreads();
  lds_fft<R>(a, g.B, st, stw, -1);
  const cplx<R>* r = a;
  for (int e = threadIdx.x; e < g.K1 * g.B; e += blockDim.x) {
    int li, kx;
    xline_of<R>(e, g, li, kx);
    const long long l = l0 + li;
    if (l >= g.nlines) continue;
    const long long yz = l / g.nc;
    const int c = (int)(l % g.nc);
    out[(yz * g.K1 + kx) * g.nc + c] = r[kx * g.B + li];
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_x_adj(const cplx<R>* __restrict__ in,
                                                   const R* __restrict__ U,
                                                   cplx<R>* __restrict__ img, XGeom g, Stages st,
                                                   const cplx<R>* __restrict__ tw, R scale,
                                                   const cplx<R>* __restrict__ add, R add_coef) {
  extern __shared__ unsigned char smem_raw[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)g.K1 * g.B;
  for (int e = threadIdx.x; e < g.K1; e += blockDim.x) stw[e] = tw[e];
  const long long l0 = (long long)blockIdx.x * g.B;
  for (int e = threadIdx.x; e < g.K1 * g.B; e += blockDim.x) {
    int li, kx;
    xline_of<R>(e, g, li, kx);
    const long long l = l0 + li;
    cplx<R> v; v.x = 0; v.y = 0;
    if (l < g.nlines) {
      const long long yz = l / g.nc;
      const int c = (int)(l % g.nc);
      v = in[(yz * g.K1 + kx) * g.nc + c];
    }
    a[kx * g.B + li] = v;
  }
  __syncthreads();
  lds_fft<R>(a, g.B, st, stw, +1);
  const cplx<R>* r = a;
  for (int e = threadIdx.x; e < g.N1 * g.B; e += blockDim.x) {
    const int li = e / g.N1, x = e % g.N1;
    const long long l = l0 + li;
    if (l >= g.nlines) continue;
    const long long yz = l / g.nc;
    const int c = (int)(l % g.nc);
    cplx<R> v = r[padpos(x, g.N1, g.K1) * g.B + li];
    const R u = U[x + (long long)g.N1 * yz] * scale;
    v.x *= u; v.y *= u;
    const long long o = x + (long long)g.N1 * (yz + g.nyz * c);
    if (add) { const cplx<R> q = add[o]; v.x += add_coef * q.x; v.y += add_coef * q.y; }
    img[o] = v;
  }
}

// ---------------------------------------------------------------- y/z pass
// in [outer][Lin][inner] -> out [outer][Lout][inner]; one work-group per
// (outer, bundle of B consecutive inner columns).
struct SGeom {
  int Lin, Lout, n_img;   // n_img: image size on this axis (pad/crop map)
  int mode;               // 0: pad (Lin = n_img, Lout = L), 1: crop (Lin = L, Lout = n_img),
                          // 2: plain (Lin = Lout = L, no index map)
  int B;
  long long inner;
  long long nbundle;      // ceil(inner / B)
};

template <typename R>
__global__ void __launch_bounds__(256) k_fft_strided(const cplx<R>* __restrict__ in,
                                                     cplx<R>* __restrict__ out, SGeom g,
                                                     Stages st, const cplx<R>* __restrict__ tw,
                                                     int sign) {
  extern __shared__ unsigned char smem_raw[];
  const int L = st.L;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* stw = a + (size_t)L * g.B;
  for (int e = threadIdx.x; e < L; e += blockDim.x) stw[e] = tw[e];
  const long long outer = blockIdx.x / g.nbundle;
  const long long col0 = (blockIdx.x % g.nbundle) * g.B;
  const cplx<R>* src = in + outer * g.Lin * g.inner;
  cplx<R>* dst = out + outer * g.Lout * g.inner;
  if (g.mode == 0) {
    for (int e = threadIdx.x; e < L * g.B; e += blockDim.x) { a[e].x = 0; a[e].y = 0; }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < g.Lin * g.B; e += blockDim.x) {
    const int i = e / g.B, li = e % g.B;
    const long long col = col0 + li;
    if (col >= g.inner) continue;
    const cplx<R> v = src[(long long)i * g.inner + col];
    const int p = g.mode == 0 ? padpos(i, g.n_img, L) : i;
    a[p * g.B + li] = v;
  }
  __syncthreads();
  lds_fft<R>(a, g.B, st, stw, sign);
  const cplx<R>* r = a;
  for (int e = threadIdx.x; e < g.Lout * g.B; e += blockDim.x) {
    const int i = e / g.B, li = e % g.B;
    const long long col = col0 + li;
    if (col >= g.inner) continue;
    const int p = g.mode == 1 ? padpos(i, g.n_img, L) : i;
    dst[(long long)i * g.inner + col] = r[p * g.B + li];
  }
}

// ---------------------------------------------------------------- host side
static void factorize(int L, Stages& st) {
  st.L = L;
  st.n = 0;
  int m = L;
  while (m % 4 == 0) { st.r[st.n++] = 4; m /= 4; }
  const int primes[] = {2, 3, 5, 7};
  for (int q : primes)
    while (m % q == 0) { st.r[st.n++] = q; m /= q; }
  if (m != 1) st.n = -1;
}

int fft_setup_axis(nufft3d_plan* p, int axis) {
  const int L = p->K[axis];
  FftAxis& a = p->ax[axis];
  a.L = L;
  Stages st;
  factorize(L, st);
  if (st.n < 0) {
    set_error("oversampled size K=" + std::to_string(L) +
              " has a prime factor above 7 (radices 2, 3, 4, 5, 7 supported)");
    return NUFFT3D_ERR_UNSUPPORTED;
  }
  a.nstage = st.n;
  for (int i = 0; i < st.n; ++i) a.radix[i] = st.r[i];
  const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
  std::vector<unsigned char> host(esz * L);
  for (int t = 0; t < L; ++t) {
    const double ang = -2.0 * M_PI * (double)t / (double)L;
    const double c = cos(ang), s = sin(ang);
    if (p->prec == 2) {
      double2 v; v.x = c; v.y = s;
      reinterpret_cast<double2*>(host.data())[t] = v;
    } else {
      float2 v; v.x = (float)c; v.y = (float)s;
      reinterpret_cast<float2*>(host.data())[t] = v;
    }
  }
  N3D_HIP(hipMalloc(&a.d_tw, esz * L));
  N3D_HIP(hipMemcpy(a.d_tw, host.data(), esz * L, hipMemcpyHostToDevice));
  return NUFFT3D_OK;
}

static Stages stages_of(const FftAxis& a) {
  Stages st;
  st.L = a.L;
  st.n = a.nstage;
  for (int i = 0; i < a.nstage; ++i) st.r[i] = a.radix[i];
  return st;
}

// lines per LDS bundle: L*B elements staged in registers by 256 threads
// (at most MaxE each) and held in one LDS buffer of at most 64 KiB
// `want` = requested lines per work-group (0 = automatic), from the plan
static int bundle_for(int L, size_t esz, int want) {
  int B = 256 * (esz == sizeof(double2) ? MaxE<double>::v : MaxE<float>::v) / L;
  const int bl = (int)(65536 / ((size_t)L * esz));
  if (bl < B) B = bl;
  if (want > 0 && want <= B) {
    int b = 1;
    while (b * 2 <= want) b *= 2;
    return b;
  }
  if (B >= 16) return 16;
  if (B >= 8) return 8;
  if (B >= 4) return 4;
  if (B >= 2) return 2;
  return 1;
}

static int bundle_x(int L, size_t esz, int nc, int want) {
  const int Bmax = bundle_for(L, esz, want);
  if (nc <= Bmax) return nc * (Bmax / nc);
  for (int B = Bmax; B >= 1; --B)
    if (nc % B == 0) return B;
  return 1;
}

template <typename R>
static int fft_forward_t(nufft3d_plan* p, const void* img, int nc, void* grid) {
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const size_t esz = sizeof(cplx<R>);
  const size_t t1 = (size_t)K1 * N2 * N3 * nc, t2 = (size_t)K1 * K2 * N3 * nc;
  N3D_TRY(ensure(p->ws_t1, t1 * esz));
  void* T1 = p->ws_t1.p;
  void* T2 = grid;
  if (K3 > 1) {
    N3D_TRY(ensure(p->ws_t2, t2 * esz));
    T2 = p->ws_t2.p;
  }
  {  // x-pass
    XGeom g;
    g.N1 = N1; g.K1 = K1; g.nc = nc;
    g.B = bundle_x(K1, esz, nc, p->fft_bundle);
    g.nyz = (long long)N2 * N3;
    g.nlines = g.nyz * nc;
    const long long nb = (g.nlines + g.B - 1) / g.B;
    const size_t sh = (size_t)K1 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_x_fwd");
    hipLaunchKernelGGL((k_fft_x_fwd<R>), dim3((unsigned)nb), dim3(256), sh, p->stream,
                       (const cplx<R>*)img, (const R*)p->d_U, (cplx<R>*)T1, g,
                       stages_of(p->ax[0]), (const cplx<R>*)p->ax[0].d_tw);
  }
  {  // y-pass: [N3][N2][K1*nc] -> [N3][K2][K1*nc]
    SGeom g;
    g.Lin = N2; g.Lout = K2; g.n_img = N2; g.mode = 0;
    g.B = bundle_for(K2, esz, p->fft_bundle);
    g.inner = (long long)K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K2 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_y_fwd");
    hipLaunchKernelGGL((k_fft_strided<R>), dim3((unsigned)(g.nbundle * N3)), dim3(256), sh,
                       p->stream, (const cplx<R>*)T1, (cplx<R>*)T2, g, stages_of(p->ax[1]),
                       (const cplx<R>*)p->ax[1].d_tw, -1);
  }
  if (K3 > 1) {  // z-pass: [N3][K2*K1*nc] -> [K3][K2*K1*nc]
    SGeom g;
    g.Lin = N3; g.Lout = K3; g.n_img = N3; g.mode = 0;
    g.B = bundle_for(K3, esz, p->fft_bundle);
    g.inner = (long long)K2 * K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K3 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_z_fwd");
    hipLaunchKernelGGL((k_fft_strided<R>), dim3((unsigned)g.nbundle), dim3(256), sh, p->stream,
                       (const cplx<R>*)T2, (cplx<R>*)grid, g, stages_of(p->ax[2]),
                       (const cplx<R>*)p->ax[2].d_tw, -1);
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R>
static int fft_adjoint_t(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                         const void* add, double add_coef) {
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
    g.B = bundle_for(K3, esz, p->fft_bundle);
    g.inner = (long long)K2 * K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K3 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_z_adj");
    hipLaunchKernelGGL((k_fft_strided<R>), dim3((unsigned)g.nbundle), dim3(256), sh, p->stream,
                       (const cplx<R>*)grid, (cplx<R>*)p->ws_t2.p, g, stages_of(p->ax[2]),
                       (const cplx<R>*)p->ax[2].d_tw, +1);
    T2 = p->ws_t2.p;
  }
  {  // y-pass: [N3][K2][K1*nc] -> [N3][N2][K1*nc]
    SGeom g;
    g.Lin = K2; g.Lout = N2; g.n_img = N2; g.mode = 1;
    g.B = bundle_for(K2, esz, p->fft_bundle);
    g.inner = (long long)K1 * nc;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)K2 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_y_adj");
    hipLaunchKernelGGL((k_fft_strided<R>), dim3((unsigned)(g.nbundle * N3)), dim3(256), sh,
                       p->stream, (const cplx<R>*)T2, (cplx<R>*)T1, g, stages_of(p->ax[1]),
                       (const cplx<R>*)p->ax[1].d_tw, +1);
  }
  {  // x-pass with crop, U, scale and the optional add term
    XGeom g;
    g.N1 = N1; g.K1 = K1; g.nc = nc;
    g.B = bundle_x(K1, esz, nc, p->fft_bundle);
    g.nyz = (long long)N2 * N3;
    g.nlines = g.nyz * nc;
    const long long nb = (g.nlines + g.B - 1) / g.B;
    const size_t sh = (size_t)K1 * (g.B + 1) * esz;
    ProfScope ps(p, "k_fft_x_adj");
    hipLaunchKernelGGL((k_fft_x_adj<R>), dim3((unsigned)nb), dim3(256), sh, p->stream,
                       (const cplx<R>*)T1, (const R*)p->d_U, (cplx<R>*)img, g,
                       stages_of(p->ax[0]), (const cplx<R>*)p->ax[0].d_tw, (R)scale,
                       (const cplx<R>*)add, (R)add_coef);
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
    SGeom g;
    g.Lin = ps.n; g.Lout = ps.n; g.n_img = ps.n; g.mode = 2;
    g.B = ps.inner >= 16 ? bundle_for(ps.n, esz, p->fft_bundle) : 1;
    g.inner = ps.inner;
    g.nbundle = (g.inner + g.B - 1) / g.B;
    const size_t sh = (size_t)ps.n * (g.B + 1) * esz;
    void* dst = bufs[which];
    ProfScope prs(p, "k_fft_plain");
    hipLaunchKernelGGL((k_fft_strided<R>), dim3((unsigned)(g.nbundle * ps.outer)), dim3(256), sh,
                       p->stream, (const cplx<R>*)src, (cplx<R>*)dst, g, stages_of(p->axN[ps.a]),
                       (const cplx<R>*)p->axN[ps.a].d_tw, -1);
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

int fft_forward(nufft3d_plan* p, const void* img, int nc, void* grid) {
  return p->prec == 2 ? fft_forward_t<double>(p, img, nc, grid)
                      : fft_forward_t<float>(p, img, nc, grid);
}

int fft_adjoint(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                const void* add, double add_coef) {
  return p->prec == 2 ? fft_adjoint_t<double>(p, grid, nc, img, scale, add, add_coef)
                      : fft_adjoint_t<float>(p, grid, nc, img, scale, add, add_coef);
}

}  // namespace nufft3d
