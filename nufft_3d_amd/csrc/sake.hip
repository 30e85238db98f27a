// SAKE Cadzow matrix completion (sake.m:100-171, undo_data_matrix at
// sake.m:227-284) on the device.
//
// The loop state is the image x = ifft3(ksp) [nc][N3][N2][N1] in the plan
// precision (fft3 = unshifted unnormalised 3-D DFT). Per iteration:
//   data consistency (sake.m:108-111): with b0 = A'W data formed once,
//     b = A'W(A x - data) = A'WA x - b0, t = MINRES(A'WA + damp^2 I, b,
//     1e-4, dc_iters), x -= t   (ksp - fft3(t) = fft3(x - t));
//   Hankel Gram of ksp = fft3(x) (sake.m:121), its eigen-decomposition (the
//     svd of the PSD Gram, sake.m:124), noise floor and minimum-variance
//     filter on the host (sake.m:127-145), F = V diag(f) V' (sake.m:146);
//   anti-diagonal averaging (sake.m:149). In k-space undo_data_matrix is
//     ksp'(v,d) = 1/(m nk) sum_e sum_c [ksp(v+e,c) G_e[c,d] + g(v+e,c) H_e[c,d]]
//     with g(u) = conj(ksp(flip(u))), m = 2 with loraks (else 1 and H = 0),
//       G_e = sum_{s_k - s_j = e} F(c,j,d,k) + conj(F(c,j+nk,d,k+nk)),
//       H_e = sum_{s_j + s_k = e} F(c,j+nk,d,k) + conj(F(c,j,d,k+nk)).
//     A k-space correlation is a per-voxel product in the image domain:
//       x'(r) = x(r) B_G(r) + xt(r) B_H(r),  B(r) = sum_e K_e e^{-2 pi i e.r/N},
//       xt(r) = conj(x(r)) e^{+2 pi i (2 c0).r/N},  c0 = center - 1,
//     so the averaging costs one lag-to-voxel transform of F's nc x nc blocks
//     (separable DFT over the (4h+1)^3 lag cube) and one coil mixing per voxel
//     instead of nk^2 circshifted [V x nc][nc x nc] products.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "nufft3d_internal.h"

namespace nufft3d {

static inline unsigned nblk_s(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <typename R>
__global__ void k_sake_sub(cplx<R>* __restrict__ x, const cplx<R>* __restrict__ t, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<R> a = x[i];
  const cplx<R> b = t[i];
  a.x -= b.x;
  a.y -= b.y;
  x[i] = a;
}

// out[:, j] = V[:, j] * f[j]  (n x n, column-major)
__global__ void k_scale_cols(const double2* __restrict__ V, const double* __restrict__ f, int n,
                             double2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * n) return;
  const double s = f[i / n];
  double2 v = V[i];
  v.x *= s;
  v.y *= s;
  out[i] = v;
}

// Lag kernels of the averaging operator from F ((m nk nc) square, column-major,
// F(c,J,d,K) = F[(c + nc J) + n2 (d + nc K)]). Output GH[a][h][d + nc c]:
// a indexes the lag cube stored at offset -e (index -e + (w-1) per axis, x
// fastest), so that lag_field's e^{+2 pi i (index-(w-1)).r/N} gives
// e^{-2 pi i e.r/N}; h = 0 is G (tap differences), h = 1 is H (tap sums).
// tapmap[(x+hw) + w((y+hw) + w(z+hw))] = tap index or -1, hw = (w-1)/2.
__global__ void k_sake_lags(const double2* __restrict__ F, int n2, const int4* __restrict__ taps,
                            const int* __restrict__ tapmap, int nk, int nc, int w, int m,
                            double2* __restrict__ GH) {
  const int D = 2 * w - 1;
  const long long cc = (long long)m * nc * nc;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)D * D * D * cc) return;
  const int q = (int)(idx % cc);
  const int a = (int)(idx / cc);
  const int h = q / (nc * nc);
  const int r = q % (nc * nc);
  const int d = r % nc, c = r / nc;
  const int hw = (w - 1) / 2;
  // e = -(index - (w-1))
  const int ex = (w - 1) - a % D, ey = (w - 1) - (a / D) % D, ez = (w - 1) - a / (D * D);
  double re = 0, im = 0;
  for (int j = 0; j < nk; ++j) {
    const int4 sj = taps[j];
    // G: s_k = e + s_j; H: s_k = e - s_j
    const int kx = h == 0 ? ex + sj.x : ex - sj.x;
    const int ky = h == 0 ? ey + sj.y : ey - sj.y;
    const int kz = h == 0 ? ez + sj.z : ez - sj.z;
    if (kx < -hw || kx > hw || ky < -hw || ky > hw || kz < -hw || kz > hw) continue;
    const int k = tapmap[(kx + hw) + w * ((ky + hw) + w * (kz + hw))];
    if (k < 0) continue;
    if (h == 0) {
      const double2 v = F[(c + (long long)nc * j) + (long long)n2 * (d + (long long)nc * k)];
      re += v.x;
      im += v.y;
      if (m == 2) {
        const double2 u = F[(c + (long long)nc * (j + nk)) + (long long)n2 * (d + (long long)nc * (k + nk))];
        re += u.x;
        im -= u.y;
      }
    } else {
      const double2 v = F[(c + (long long)nc * (j + nk)) + (long long)n2 * (d + (long long)nc * k)];
      const double2 u = F[(c + (long long)nc * j) + (long long)n2 * (d + (long long)nc * (k + nk))];
      re += v.x + u.x;
      im += v.y - u.y;
    }
  }
  double2 o;
  o.x = re;
  o.y = im;
  GH[idx] = o;
}

// y[d][v] = sum_c B[v][d + nc c] x[c][v]  (+ sum_c B[v][nc^2 + d + nc c] xt[c][v])
// with xt = conj(x) e^{+2 pi i (2 c0).r/N}; phase per axis as (2 c0 r mod N)/N.
template <typename R>
__global__ void k_sake_apply(const cplx<R>* __restrict__ B, const cplx<R>* __restrict__ x,
                             cplx<R>* __restrict__ y, int N1, int N2, int N3, int nc, int m,
                             int c2x, int c2y, int c2z) {
  const long long V = (long long)N1 * N2 * N3;
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const long long cc = (long long)m * nc * nc;
  const cplx<R>* b = B + v * cc;
  R pc = 1, ps = 0;
  if (m == 2) {
    const int rx = (int)(v % N1), ry = (int)((v / N1) % N2), rz = (int)(v / ((long long)N1 * N2));
    const double fr = (double)(((long long)c2x * rx) % N1) / N1 + (double)(((long long)c2y * ry) % N2) / N2 +
                      (double)(((long long)c2z * rz) % N3) / N3;
    double sn, cs;
    sincospi(2.0 * fr, &sn, &cs);
    pc = (R)cs;
    ps = (R)sn;
  }
  for (int d = 0; d < nc; ++d) {
    R re = 0, im = 0;
    for (int c = 0; c < nc; ++c) {
      const cplx<R> xv = x[(long long)c * V + v];
      const cplx<R> g = b[d + nc * c];
      re += g.x * xv.x - g.y * xv.y;
      im += g.x * xv.y + g.y * xv.x;
      if (m == 2) {
        // xt = conj(xv) * (pc + i ps)
        const R tx = xv.x * pc + xv.y * ps;
        const R ty = xv.x * ps - xv.y * pc;
        const cplx<R> hgt = b[nc * nc + d + nc * c];
        re += hgt.x * tx - hgt.y * ty;
        im += hgt.x * ty + hgt.y * tx;
      }
    }
    cplx<R> o;
    o.x = re;
    o.y = im;
    y[(long long)d * V + v] = o;
  }
}

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)dev_free(p);
  }
};

int dev_alloc(DevBuf& b, size_t bytes) {
  if (dev_malloc(&b.p, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
    return NUFFT3D_ERR_NOMEM;
  }
  return NUFFT3D_OK;
}

}  // namespace

int sake_run(nufft3d_plan* p, const void* data, int nc, const nufft3d_sake_opts& o,
             const void* wsamp, void* out, double* history, nufft3d_sake_info* info) {
  const bool dbl = p->prec == 2;
  const size_t esz = dbl ? sizeof(double2) : sizeof(float2);
  const int N[3] = {p->N[0], p->N[1], p->N[2]};
  const long long V = (long long)N[0] * N[1] * N[2];
  const size_t nimg = (size_t)V * nc;
  std::vector<int4> taps;
  int w = 0;
  kernel_taps(o.width, o.radial, false, taps, w);                     // sake.m:62-71
  const int nk = (int)taps.size();
  const int m = o.loraks ? 2 : 1;
  const int n2 = m * nk * nc;
  const int D = 2 * w - 1;
  const long long cc = (long long)m * nc * nc;
  const int hw = (w - 1) / 2;
  std::vector<int> tapmap((size_t)w * w * w, -1);
  for (int t = 0; t < nk; ++t)
    tapmap[(taps[t].x + hw) + w * ((taps[t].y + hw) + w * (taps[t].z + hw))] = t;
  const int cen[3] = {o.center[0], o.center[1], o.center[2]};
  const int dims[4] = {N[0], N[1], N[2], nc};

  DevBuf b0, bres, tt, xa, xb, B, AA, VF, F, GH, Wd, fd, dtaps, dmap;
  N3D_TRY(dev_alloc(b0, nimg * esz));
  N3D_TRY(dev_alloc(bres, nimg * esz));
  N3D_TRY(dev_alloc(tt, nimg * esz));
  N3D_TRY(dev_alloc(xa, nimg * esz));
  N3D_TRY(dev_alloc(xb, nimg * esz));
  N3D_TRY(dev_alloc(B, (size_t)V * cc * esz));
  N3D_TRY(dev_alloc(AA, sizeof(double2) * (size_t)n2 * n2));
  N3D_TRY(dev_alloc(VF, sizeof(double2) * (size_t)n2 * n2));
  N3D_TRY(dev_alloc(F, sizeof(double2) * (size_t)n2 * n2));
  N3D_TRY(dev_alloc(GH, sizeof(double2) * (size_t)D * D * D * cc));
  N3D_TRY(dev_alloc(Wd, sizeof(double) * n2));
  N3D_TRY(dev_alloc(fd, sizeof(double) * n2));
  N3D_TRY(dev_alloc(dtaps, sizeof(int4) * nk));
  N3D_TRY(dev_alloc(dmap, sizeof(int) * tapmap.size()));
  N3D_HIP(hipMemcpyAsync(dtaps.p, taps.data(), sizeof(int4) * nk, hipMemcpyHostToDevice, p->stream));
  N3D_HIP(hipMemcpyAsync(dmap.p, tapmap.data(), sizeof(int) * tapmap.size(), hipMemcpyHostToDevice,
                         p->stream));

  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  N3D_TRY(ensure(p->ws_grid, ngrid * esz));
  N3D_TRY(ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz));
  const double sc = 1.0 / sqrt((double)p->nnz);
  // b0 = aNUFT(W .* data)
  N3D_TRY(permute_in(p, data, nc, p->ws_samp.p, wsamp));
  N3D_TRY(spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr));
  N3D_TRY(fft_adjoint(p, p->ws_grid.p, nc, b0.p, sc, nullptr, 0.0));
  N3D_HIP(hipMemsetAsync(xa.p, 0, nimg * esz, p->stream));            // ksp = 0 (sake.m:103)
  void* x = xa.p;
  void* y = xb.p;

  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    set_error("rocblas_create_handle failed");
    return NUFFT3D_ERR_DEVICE;
  }
  rocblas_set_stream(h, p->stream);
  struct HandleGuard {
    rocblas_handle h;
    ~HandleGuard() { rocblas_destroy_handle(h); }
  } hg{h};

  const double density = std::min(1.0, (double)p->nnz / (double)V);   // sake.m:96 (obj.sd)
  double noise = o.noise;
  double sigma = 0, prev = 0;
  int iters = 0, rank = 0;
  std::vector<double> lam(n2), wabs(n2), f(n2);
  for (int it = 1; it <= o.maxit; ++it) {
    // data consistency (sake.m:108-111)
    N3D_TRY(fft_forward(p, x, nc, p->ws_grid.p));
    N3D_TRY(interp(p, p->ws_grid.p, nc, p->ws_samp.p, 0, wsamp, sc));
    N3D_TRY(spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr));
    N3D_TRY(fft_adjoint(p, p->ws_grid.p, nc, bres.p, sc, b0.p, -1.0));
    const void* tcorr = bres.p;                                       // maxit 0: regridding
    if (o.dc_iters > 0) {
      N3D_TRY(minres_solve(p, bres.p, nc, o.dc_iters, o.damp, wsamp, 1e-4, tt.p, nullptr, false));
      tcorr = tt.p;
    }
    {
      ProfScope ps(p, "k_sake_sub");
      if (dbl)
        hipLaunchKernelGGL((k_sake_sub<double>), dim3(nblk_s((long long)nimg)), dim3(256), 0, p->stream,
                           (double2*)x, (const double2*)tcorr, (long long)nimg);
      else
        hipLaunchKernelGGL((k_sake_sub<float>), dim3(nblk_s((long long)nimg)), dim3(256), 0, p->stream,
                           (float2*)x, (const float2*)tcorr, (long long)nimg);
    }
    // Gram of ksp = fft3(x) and its eigen-decomposition (sake.m:121-125)
    N3D_TRY(fft_plain(p, x, nc, y));
    N3D_TRY(gram_device(p, y, dims, taps, (double2*)AA.p, o.loraks != 0, cen));
    N3D_TRY(heev_device(p, (double2*)AA.p, n2, (double*)Wd.p));
    N3D_HIP(hipMemcpyAsync(lam.data(), Wd.p, sizeof(double) * n2, hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
    // singular values |lambda| in descending order (svd of the PSD Gram)
    for (int i = 0; i < n2; ++i) wabs[i] = std::fabs(lam[i]);
    std::vector<double> wdesc(wabs);
    std::sort(wdesc.begin(), wdesc.end(), [](double a, double b) { return a > b; });
    if (!(noise >= 0)) {                                              // sake.m:128-138
      std::vector<double> fdum;
      int nvdum = 0;
      N3D_TRY(pruno_select(wdesc, -1.0, density, N[0], N[1], 0.0, noise, fdum, nvdum));
    }
    sigma = noise * std::sqrt(density * N[0] * N[1]);                 // sake.m:139
    rank = 0;
    for (int i = 0; i < n2; ++i) {                                    // sake.m:142-145
      const double Ws = std::sqrt(wabs[i]);
      f[i] = Ws > 0 ? std::max(0.0, 1.0 - sigma * sigma / (Ws * Ws)) : 0.0;
      rank += f[i] > 0;
    }
    N3D_HIP(hipMemcpyAsync(fd.p, f.data(), sizeof(double) * n2, hipMemcpyHostToDevice, p->stream));
    // F = V diag(f) V' (sake.m:146)
    hipLaunchKernelGGL(k_scale_cols, dim3(nblk_s((long long)n2 * n2)), dim3(256), 0, p->stream,
                       (const double2*)AA.p, (const double*)fd.p, n2, (double2*)VF.p);
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    if (rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, n2, n2, n2,
                      &one, (const rocblas_double_complex*)VF.p, n2,
                      (const rocblas_double_complex*)AA.p, n2, &zero,
                      (rocblas_double_complex*)F.p, n2) != rocblas_status_success) {
      set_error("rocblas_zgemm failed");
      return NUFFT3D_ERR_DEVICE;
    }
    // anti-diagonal averaging (sake.m:149, 227-284) as a per-voxel coil mixing
    {
      ProfScope ps(p, "k_sake_lags");
      hipLaunchKernelGGL(k_sake_lags, dim3(nblk_s((long long)D * D * D * cc)), dim3(256), 0, p->stream,
                         (const double2*)F.p, n2, (const int4*)dtaps.p, (const int*)dmap.p, nk, nc, w, m,
                         (double2*)GH.p);
    }
    N3D_TRY(lag_field(p, (const double2*)GH.p, w, cc, N, 1.0 / (m * nk), B.p));
    {
      ProfScope ps(p, "k_sake_apply");
      const int c2[3] = {2 * (cen[0] - 1), 2 * (cen[1] - 1), 2 * (cen[2] - 1)};
      if (dbl)
        hipLaunchKernelGGL((k_sake_apply<double>), dim3(nblk_s(V)), dim3(256), 0, p->stream,
                           (const double2*)B.p, (const double2*)x, (double2*)y, N[0], N[1], N[2], nc, m,
                           c2[0], c2[1], c2[2]);
      else
        hipLaunchKernelGGL((k_sake_apply<float>), dim3(nblk_s(V)), dim3(256), 0, p->stream,
                           (const float2*)B.p, (const float2*)x, (float2*)y, N[0], N[1], N[2], nc, m,
                           c2[0], c2[1], c2[2]);
    }
    N3D_HIP(hipGetLastError());
    std::swap(x, y);
    // convergence (sake.m:151-169)
    double n_inf = 0, n_1 = 0, n_2 = 0;
    for (int i = 0; i < n2; ++i) {
      const double Ws = std::sqrt(wabs[i]);
      n_inf = std::max(n_inf, Ws);
      n_1 += Ws;
      n_2 += wabs[i];
    }
    n_2 = std::sqrt(n_2);
    const double tl = it == 1 ? o.tol : std::fabs(n_inf - prev) / n_inf;
    prev = n_inf;
    iters = it;
    if (history) {
      double* hrow = history + 5 * (size_t)(it - 1);
      hrow[0] = n_inf;
      hrow[1] = n_1;
      hrow[2] = n_2;
      hrow[3] = tl;
      hrow[4] = rank;
    }
    if (tl < o.tol) break;
  }
  if (o.output == 1)
    N3D_HIP(hipMemcpyAsync(out, x, nimg * esz, hipMemcpyDeviceToDevice, p->stream));
  else
    N3D_TRY(fft_plain(p, x, nc, out));
  N3D_HIP(hipStreamSynchronize(p->stream));
  if (info) {
    info->iters = iters;
    info->rank = rank;
    info->noise = noise >= 0 ? noise : 0.0;
    info->sigma = sigma;
  }
  return NUFFT3D_OK;
}

}  // namespace nufft3d
