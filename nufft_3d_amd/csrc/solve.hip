// Normal operator and MINRES (replaces iprojection, nufft_3d.m:305-313, and
// the MATLAB minres call of iNUFT.m:136).
//
// MINRES is Paige-Saunders with every scalar of the recurrence kept on the
// device in double, one independent solve per coil (right-hand side), so an
// iteration is a fixed launch sequence with no host round trip:
//   y = A v                       (fwd FFT, interp*W, spread, adj FFT + damp^2 v)
//   phase A : y -= (beta/oldb) r1, partial <v,y>        (fp64 block partials)
//   scalar A: alfa = sum partials (fixed order)
//   phase B : y -= (alfa/beta) r2, partial ||y||^2
//   scalar B: beta, plane rotation, phi, relres, convergence
//   phase C : w = (v - oldeps w1 - delta w2)/gamma; x += phi w; v = y/beta
// Inner products and norms accumulate in double in a fixed order, so the
// solve is bitwise reproducible.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <functional>
#include <vector>

#include "nufft3d_internal.h"

namespace nufft3d {

struct MrState {
  double beta1, beta, oldb, alfa, dbar, epsln, oldeps, phibar, cs, sn, gamma, delta, phi, relres;
  double bnorm;   // ||b||: relres and the stopping test are relative to it
  int active, upd, iters, pad;
};

constexpr int kRedThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// partial[c*nblk + b] = sum over block b of Re(conj(a) b) or |a|^2
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_dot_partial(const cplx<R>* __restrict__ a,
                                                             const cplx<R>* __restrict__ b,
                                                             long long n, double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  const int c = blockIdx.y;
  const int nblk = gridDim.x;
  const cplx<R>* pa = a + (long long)c * n;
  const cplx<R>* pb = b + (long long)c * n;
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)nblk * blockDim.x) {
    const cplx<R> x = pa[i], y = pb[i];
    acc += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(long long)c * nblk + blockIdx.x] = s;
}

// out[c] = sum_b partial[c*nblk + b] in a fixed tree order
__global__ void __launch_bounds__(kRedThreads) k_sum_partials(const double* __restrict__ partial,
                                                              int nblk, double* __restrict__ out) {
  __shared__ double sh[kRedThreads];
  const int c = blockIdx.x;
  double acc = 0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc += partial[(long long)c * nblk + i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[c] = s;
}

template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_sum_real_partial(const R* __restrict__ v, long long n,
                                                                  double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += (double)v[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---------------------------------------------------------------- MINRES
// r0 = b - A x0 (warm start) or b; beta1 = ||r0||. A column starts active
// unless ||r0|| <= tol ||b|| already; b = 0 gives x = 0 and no iteration
// (MATLAB minres returns the zero vector for a zero right-hand side).
template <typename R>
__global__ void k_mr_init(const cplx<R>* __restrict__ r0, long long n, const double* __restrict__ rn2,
                          const double* __restrict__ bn2, MrState* __restrict__ st,
                          cplx<R>* __restrict__ x, cplx<R>* __restrict__ r2, cplx<R>* __restrict__ v,
                          double tol, int warm) {
  const int c = blockIdx.y;
  const double beta1 = sqrt(rn2[c]), bnorm = sqrt(bn2[c]);
  const bool zero_x = !warm || !(bnorm > 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MrState s = {};
    s.beta1 = beta1; s.beta = beta1; s.oldb = 0; s.dbar = 0; s.epsln = 0;
    s.phibar = beta1; s.cs = -1; s.sn = 0; s.bnorm = bnorm;
    s.relres = bnorm > 0 ? beta1 / bnorm : 0.0;
    s.active = bnorm > 0 && beta1 > 0 && beta1 / bnorm > tol;
    s.upd = 0; s.iters = 0;
    st[c] = s;
  }
  const R inv = beta1 > 0 ? (R)(1.0 / beta1) : (R)0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long o = (long long)c * n + i;
    const cplx<R> bv = r0[o];
    if (zero_x) {
      cplx<R> z; z.x = 0; z.y = 0;
      x[o] = z;
    }
    r2[o] = bv;
    cplx<R> vv; vv.x = bv.x * inv; vv.y = bv.y * inv;
    v[o] = vv;
  }
}

// y = b - y over all columns
template <typename R>
__global__ void k_sub_from(const cplx<R>* __restrict__ b, cplx<R>* __restrict__ y, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cplx<R> bv = b[i];
    cplx<R> yv = y[i];
    yv.x = bv.x - yv.x;
    yv.y = bv.y - yv.y;
    y[i] = yv;
  }
}

// partial[b] = sum over block b of |a - c|^2
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_diff_norm_partial(const cplx<R>* __restrict__ a,
                                                                   const cplx<R>* __restrict__ c,
                                                                   long long n,
                                                                   double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cplx<R> x = a[i], y = c[i];
    const double dx = (double)x.x - (double)y.x, dy = (double)x.y - (double)y.y;
    acc += dx * dx + dy * dy;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y -= (beta/oldb) r1 (itn>=2); partial <v, y>
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_mr_phaseA(cplx<R>* __restrict__ y,
                                                           const cplx<R>* __restrict__ r1,
                                                           const cplx<R>* __restrict__ v, long long n,
                                                           const MrState* __restrict__ st, int itn,
                                                           double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  const int c = blockIdx.y;
  const int nblk = gridDim.x;
  const MrState s = st[c];
  const R coef = (itn >= 2 && s.oldb > 0) ? (R)(s.beta / s.oldb) : (R)0;
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)nblk * blockDim.x) {
    const long long o = (long long)c * n + i;
    cplx<R> yv = y[o];
    if (coef != (R)0) {
      const cplx<R> r = r1[o];
      yv.x -= coef * r.x;
      yv.y -= coef * r.y;
      y[o] = yv;
    }
    const cplx<R> vv = v[o];
    acc += (double)vv.x * (double)yv.x + (double)vv.y * (double)yv.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(long long)c * nblk + blockIdx.x] = t;
}

__global__ void k_mr_scalarA(MrState* __restrict__ st, const double* __restrict__ sums) {
  const int c = blockIdx.x;
  if (threadIdx.x == 0) st[c].alfa = sums[c];
}

// y -= (alfa/beta) r2; partial ||y||^2
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_mr_phaseB(cplx<R>* __restrict__ y,
                                                           const cplx<R>* __restrict__ r2, long long n,
                                                           const MrState* __restrict__ st,
                                                           double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  const int c = blockIdx.y;
  const int nblk = gridDim.x;
  const MrState s = st[c];
  const R coef = s.beta > 0 ? (R)(s.alfa / s.beta) : (R)0;
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)nblk * blockDim.x) {
    const long long o = (long long)c * n + i;
    cplx<R> yv = y[o];
    const cplx<R> r = r2[o];
    yv.x -= coef * r.x;
    yv.y -= coef * r.y;
    y[o] = yv;
    acc += (double)yv.x * (double)yv.x + (double)yv.y * (double)yv.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(long long)c * nblk + blockIdx.x] = t;
}

__global__ void k_mr_scalarB(MrState* __restrict__ st, const double* __restrict__ sums, double tol) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0) return;
  MrState s = st[c];
  s.upd = s.active;
  if (!s.active) { st[c] = s; return; }
  s.oldb = s.beta;
  s.beta = sqrt(sums[c]);
  s.oldeps = s.epsln;
  s.delta = s.cs * s.dbar + s.sn * s.alfa;
  const double gbar = s.sn * s.dbar - s.cs * s.alfa;
  s.epsln = s.sn * s.beta;
  s.dbar = -s.cs * s.beta;
  double gamma = hypot(gbar, s.beta);
  if (gamma < 2.220446049250313e-16) gamma = 2.220446049250313e-16;
  s.gamma = gamma;
  s.cs = gbar / gamma;
  s.sn = s.beta / gamma;
  s.phi = s.cs * s.phibar;
  s.phibar = s.sn * s.phibar;
  s.iters += 1;
  s.relres = s.bnorm > 0 ? s.phibar / s.bnorm : 0.0;
  s.active = s.relres > tol;
  st[c] = s;
}

// w_new = (v - oldeps w1 - delta w2)/gamma, stored over the w1 buffer;
// x += phi w_new; v <- y / beta
template <typename R>
__global__ void k_mr_phaseC(cplx<R>* __restrict__ v, cplx<R>* __restrict__ w1,
                            const cplx<R>* __restrict__ w2, cplx<R>* __restrict__ x,
                            const cplx<R>* __restrict__ y, long long n,
                            const MrState* __restrict__ st) {
  const int c = blockIdx.y;
  const MrState s = st[c];
  if (!s.upd) return;
  const R oe = (R)s.oldeps, de = (R)s.delta, ig = (R)(1.0 / s.gamma), ph = (R)s.phi;
  const R ib = s.beta > 0 ? (R)(1.0 / s.beta) : (R)0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long o = (long long)c * n + i;
    const cplx<R> vv = v[o], a = w1[o], b = w2[o];
    cplx<R> w;
    w.x = (vv.x - oe * a.x - de * b.x) * ig;
    w.y = (vv.y - oe * a.y - de * b.y) * ig;
    w1[o] = w;
    cplx<R> xv = x[o];
    xv.x += ph * w.x;
    xv.y += ph * w.y;
    x[o] = xv;
    const cplx<R> yv = y[o];
    cplx<R> nv; nv.x = yv.x * ib; nv.y = yv.y * ib;
    v[o] = nv;
  }
}

// ---------------------------------------------------------------- host side
static int red_blocks(long long n) {
  long long nb = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
  if (nb > 512) nb = 512;
  if (nb < 1) nb = 1;
  return (int)nb;
}

int sum_real(nufft3d_plan* p, const void* v, int64_t n, double* out) {
  const int nb = red_blocks(n);
  N3D_TRY(ensure(p->ws_red, (size_t)(nb + 8) * sizeof(double)));
  double* part = (double*)p->ws_red.p;
  if (p->prec == 2)
    hipLaunchKernelGGL((k_sum_real_partial<double>), dim3(nb), dim3(kRedThreads), 0, p->stream,
                       (const double*)v, (long long)n, part);
  else
    hipLaunchKernelGGL((k_sum_real_partial<float>), dim3(nb), dim3(kRedThreads), 0, p->stream,
                       (const float*)v, (long long)n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, p->stream, part, nb, part + nb);
  N3D_HIP(hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  return NUFFT3D_OK;
}

// ||a - c||^2 over n complex entries, accumulated in double in a fixed order
int diff_norm2(nufft3d_plan* p, const void* a, const void* c, int64_t n, double* out) {
  const int nb = red_blocks(n);
  N3D_TRY(ensure(p->ws_red, (size_t)(nb + 8) * sizeof(double)));
  double* part = (double*)p->ws_red.p;
  if (p->prec == 2)
    hipLaunchKernelGGL((k_diff_norm_partial<double>), dim3(nb), dim3(kRedThreads), 0, p->stream,
                       (const cplx<double>*)a, (const cplx<double>*)c, (long long)n, part);
  else
    hipLaunchKernelGGL((k_diff_norm_partial<float>), dim3(nb), dim3(kRedThreads), 0, p->stream,
                       (const cplx<float>*)a, (const cplx<float>*)c, (long long)n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, p->stream, part, nb, part + nb);
  N3D_HIP(hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  return NUFFT3D_OK;
}

// y = A'WA x + damp^2 x (iprojection with coils independent)
int apply_normal(nufft3d_plan* p, const void* x, int nc, double damp, const void* wsamp, void* y) {
  const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  N3D_TRY(ensure(p->ws_grid, ngrid * esz));
  N3D_TRY(ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz));
  const double sc = 1.0 / sqrt((double)p->nnz);
  N3D_TRY(fft_forward(p, x, nc, p->ws_grid.p));
  N3D_TRY(interp(p, p->ws_grid.p, nc, p->ws_samp.p, 0, wsamp, sc));
  N3D_TRY(spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr));
  // the damping term damp^2 x (none when damp = 0: the x pass then skips
  // reading x)
  N3D_TRY(fft_adjoint(p, p->ws_grid.p, nc, y, sc, damp != 0.0 ? x : nullptr, damp * damp));
  return NUFFT3D_OK;
}

// MINRES on nc right-hand sides of length n stored [nc][n]; op(v, y) applies
// the Hermitian operator to all right-hand sides at once. `opkey` identifies
// the operator (its scalars and extra arrays) for the cached iteration graph.
// With `warm` the solve starts from the x0 held in xout (MATLAB's
// minres(A,b,tol,maxit,[],[],x0)): the recurrence runs on r0 = b - A x0.
//
// Iteration 1 runs eagerly (it also grows every workspace the operator
// uses). After it, blocks of kGraphBlock iterations - the period of the
// buffer rotation (w/w2 swap, r1/r2/y cycle), so a block leaves the buffers
// where it found them - are captured once as a HIP graph and replayed; the
// host checks for convergence between blocks. The graph is cached on the
// plan while every pointer and scalar it captured is unchanged. Without a
// capturable stream (the null stream, profiling on, a caller already
// capturing, NUFFT3D_GRAPH=0) the blocks run eagerly: same kernels, same
// order, bitwise the same result.
constexpr int kGraphBlock = 6;

template <typename R>
static int minres_core(nufft3d_plan* p, const void* b, long long n, int nc, int maxit, double tol,
                       void* xout, nufft3d_info* info,
                       const std::function<int(const void*, void*)>& op,
                       const std::vector<uint64_t>& opkey, bool warm = false) {
  using C = cplx<R>;
  const size_t vec = (size_t)n * nc * sizeof(C);
  // vectors: v, r1, r2, y, w, w2 ; x is the caller's output
  N3D_TRY(ensure(p->ws_solve, 6 * vec));
  unsigned char* base = (unsigned char*)p->ws_solve.p;
  C* v = (C*)(base + 0 * vec);
  C* r1 = (C*)(base + 1 * vec);
  C* r2 = (C*)(base + 2 * vec);
  C* y = (C*)(base + 3 * vec);
  C* w = (C*)(base + 4 * vec);
  C* w2 = (C*)(base + 5 * vec);
  C* x = (C*)xout;
  const int nb = red_blocks(n);
  N3D_TRY(ensure(p->ws_red, (size_t)nc * (nb + 2) * sizeof(double) + nc * sizeof(MrState) + 64));
  double* part = (double*)p->ws_red.p;
  double* sums = part + (size_t)nc * nb;
  double* bn2 = sums + nc;
  MrState* st = (MrState*)(bn2 + nc);
  const dim3 grid(nb, nc);

  hipLaunchKernelGGL((k_dot_partial<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)b,
                     (const C*)b, n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, bn2);
  const C* r0 = (const C*)b;
  const double* rn2 = bn2;
  if (warm) {   // r0 = b - A x0, held in y until iteration 1 overwrites it
    N3D_TRY(op(x, y));
    hipLaunchKernelGGL((k_sub_from<R>), dim3(red_blocks(n * nc)), dim3(kRedThreads), 0, p->stream,
                       (const C*)b, y, n * nc);
    hipLaunchKernelGGL((k_dot_partial<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)y,
                       (const C*)y, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    r0 = y;
    rn2 = sums;
  }
  hipLaunchKernelGGL((k_mr_init<R>), grid, dim3(kRedThreads), 0, p->stream, r0, n, rn2,
                     (const double*)bn2, st, x, r2, v, tol, warm ? 1 : 0);
  N3D_HIP(hipMemsetAsync(w, 0, vec, p->stream));
  N3D_HIP(hipMemsetAsync(w2, 0, vec, p->stream));

  auto iterate = [&](int itn) -> int {
    N3D_TRY(op(v, y));
    hipLaunchKernelGGL((k_mr_phaseA<R>), grid, dim3(kRedThreads), 0, p->stream, y, (const C*)r1,
                       (const C*)v, n, (const MrState*)st, itn, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    hipLaunchKernelGGL(k_mr_scalarA, dim3(nc), dim3(1), 0, p->stream, st, (const double*)sums);
    hipLaunchKernelGGL((k_mr_phaseB<R>), grid, dim3(kRedThreads), 0, p->stream, y, (const C*)r2, n,
                       (const MrState*)st, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    hipLaunchKernelGGL(k_mr_scalarB, dim3(nc), dim3(1), 0, p->stream, st, (const double*)sums, tol);
    // w_new = (v - oldeps*w2 - delta*w)/gamma is written over w2, whose old
    // value (the next iteration's w1) is dead after this update
    hipLaunchKernelGGL((k_mr_phaseC<R>), grid, dim3(kRedThreads), 0, p->stream, v, w2,
                       (const C*)w, x, (const C*)y, n, (const MrState*)st);
    {
      C* t = w2;
      w2 = w;
      w = t;
    }
    C* t = r1;
    r1 = r2;
    r2 = y;
    y = t;
    N3D_HIP(hipGetLastError());
    return NUFFT3D_OK;
  };
  auto any_active = [&](bool& any) -> int {   // early exit once every column stopped
    std::vector<MrState> hs(nc);
    N3D_HIP(hipMemcpyAsync(hs.data(), st, nc * sizeof(MrState), hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
    any = false;
    for (int c = 0; c < nc; ++c) any = any || hs[c].active;
    return NUFFT3D_OK;
  };

  int itn = 1;
  if (maxit >= 1) {
    bool any = true;
    if (warm) N3D_TRY(any_active(any));   // x0 may already meet the test
    if (any) N3D_TRY(iterate(itn++));
    else maxit = 0;
  }
  // the graph's identity: every pointer and scalar its kernels captured
  std::vector<uint64_t> key(opkey);
  const void* ptrs[] = {b, xout, p->ws_solve.p, p->ws_red.p, p->ws_grid.p, p->ws_samp.p,
                        p->ws_t1.p, p->ws_t2.p, p->ws_part.p, p->stream};
  for (const void* q : ptrs) key.push_back((uint64_t)(uintptr_t)q);
  key.push_back((uint64_t)n);
  key.push_back((uint64_t)nc);
  key.push_back(__builtin_bit_cast(uint64_t, tol));
  key.push_back((uint64_t)sizeof(R));
  bool graphs = p->use_graphs && !p->prof.on && p->stream != nullptr;
  const int test_capture_fail_at = p->capture_fail_at;
  bool stop = false;
  while (!stop && itn + kGraphBlock - 1 <= maxit) {
    bool any = true;
    N3D_TRY(any_active(any));
    if (!any) { stop = true; break; }
    bool done = false;
    if (graphs) {
      if (!(p->mr_exec && p->mr_key == key)) {
        if (p->mr_exec) { (void)hipGraphExecDestroy(p->mr_exec); p->mr_exec = nullptr; }
        hipGraph_t gph = nullptr;
        // the rotation state at the block's start: a capture that stops
        // part-way leaves the pointers rotated by fewer than a full period
        C* const rot[6] = {v, r1, r2, y, w, w2};
        bool ok;
        int st_cap = NUFFT3D_OK;
        {
          capture_scope cs;   // no device allocation or release anywhere meanwhile
          ok = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
          if (ok) {
            for (int k = 0; k < kGraphBlock && st_cap == NUFFT3D_OK; ++k)
              st_cap = (test_capture_fail_at == k + 1) ? NUFFT3D_ERR_DEVICE : iterate(itn + k);
            ok = hipStreamEndCapture(p->stream, &gph) == hipSuccess && st_cap == NUFFT3D_OK && gph;
          }
        }
        if (!ok) {   // nothing captured ran: back to the block's start
          v = rot[0]; r1 = rot[1]; r2 = rot[2]; y = rot[3]; w = rot[4]; w2 = rot[5];
        }
        hipGraphExec_t ex = nullptr;
        if (ok) ok = hipGraphInstantiate(&ex, gph, nullptr, nullptr, 0) == hipSuccess;
        if (gph) (void)hipGraphDestroy(gph);
        (void)hipGetLastError();
        if (ok) {
          p->mr_exec = ex;
          p->mr_key = key;
          ++p->graphs_captured;
        } else {
          ++p->graphs_failed;
          // nothing captured ran and the buffers are back at the block's
          // start: run this block eagerly from here on
          graphs = false;
        }
      }
      if (graphs) {
        N3D_HIP(hipGraphLaunch(p->mr_exec, p->stream));
        done = true;
      }
    }
    if (!done)
      for (int k = 0; k < kGraphBlock; ++k) N3D_TRY(iterate(itn + k));
    itn += kGraphBlock;
  }
  if (!stop && itn <= maxit) {
    bool any = true;
    N3D_TRY(any_active(any));
    if (any)
      for (; itn <= maxit; ++itn) N3D_TRY(iterate(itn));
  }
  std::vector<MrState> hs(nc);
  N3D_HIP(hipMemcpyAsync(hs.data(), st, nc * sizeof(MrState), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  if (info) {
    info->nrhs = nc;
    for (int c = 0; c < nc && c < 64; ++c) {
      info->iters[c] = hs[c].iters;
      info->relres[c] = hs[c].relres;
    }
  }
  return NUFFT3D_OK;
}

template <typename R>
static int minres_t(nufft3d_plan* p, const void* b, int nc, int maxit, double damp,
                    const void* wsamp, double tol, void* xout, nufft3d_info* info, bool warm) {
  const long long n = (long long)p->N[0] * p->N[1] * p->N[2];
  const std::vector<uint64_t> key = {1, __builtin_bit_cast(uint64_t, damp), (uint64_t)(uintptr_t)wsamp};
  return minres_core<R>(p, b, n, nc, maxit, tol, xout, info, [&](const void* v, void* y) {
    return apply_normal(p, v, nc, damp, wsamp, y);
  }, key, warm);
}

// Coils coupled through the PRUNO null term: one right-hand side of length
// N^3*nc (iNUFT.m:214-218, nufft_3d.m:317-320).
template <typename R>
static int minres_coupled_t(nufft3d_plan* p, const void* b, int nc, int maxit, double damp,
                            const void* wsamp, const void* B, double tol, void* xout,
                            nufft3d_info* info) {
  const long long n = (long long)p->N[0] * p->N[1] * p->N[2] * nc;
  const std::vector<uint64_t> key = {2, __builtin_bit_cast(uint64_t, damp), (uint64_t)(uintptr_t)wsamp,
                                      (uint64_t)(uintptr_t)B, (uint64_t)nc};
  return minres_core<R>(p, b, n, 1, maxit, tol, xout, info, [&](const void* v, void* y) {
    N3D_TRY(apply_normal(p, v, nc, damp, wsamp, y));
    return null_apply(p, B, v, y, nc);
  }, key);
}

int minres_coupled(nufft3d_plan* p, const void* b, int nc, int maxit, double damp,
                   const void* wsamp, const void* B, double tol, void* x, nufft3d_info* info) {
  return p->prec == 2 ? minres_coupled_t<double>(p, b, nc, maxit, damp, wsamp, B, tol, x, info)
                      : minres_coupled_t<float>(p, b, nc, maxit, damp, wsamp, B, tol, x, info);
}

int minres_solve(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, const void* wsamp,
                 double tol, void* x, nufft3d_info* info, bool warm) {
  return p->prec == 2 ? minres_t<double>(p, b, nc, maxit, damp, wsamp, tol, x, info, warm)
                      : minres_t<float>(p, b, nc, maxit, damp, wsamp, tol, x, info, warm);
}

// ---------------------------------------------------------------- phase constraint
// iNUFT.m:144-166 with pprojection (nufft_3d.m:325-334): conjugate gradients
// on the real-linear operator  y = conj(P) (A'WA + damp^2)(P x) + i lam^2 Im(x)
// with the real inner product Re<a,b> (the method of the external pcgpc,
// Bydder & Robson MRM 2005;53:1393), zero start, one solve per coil, every
// scalar on the device in double (same reductions as MINRES).

struct CgState {
  double rho, rho_old, beta, alpha, bnorm, relres;
  int active, upd, iters, pad;
};

template <typename R>
__global__ void k_cg_init(const cplx<R>* __restrict__ b, long long n, const double* __restrict__ bn2,
                          CgState* __restrict__ st, cplx<R>* __restrict__ x, cplx<R>* __restrict__ r,
                          cplx<R>* __restrict__ pv, double tol) {
  const int c = blockIdx.y;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    CgState s = {};
    s.bnorm = sqrt(bn2[c]);
    s.relres = s.bnorm > 0 ? 1.0 : 0.0;
    s.active = s.bnorm > 0 && 1.0 > tol;
    st[c] = s;
  }
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long o = (long long)c * n + i;
    cplx<R> z; z.x = 0; z.y = 0;
    x[o] = z;
    pv[o] = z;
    r[o] = b[o];
  }
}

// rho = Re<r,r>; beta = rho/rho_old (0 on the first iteration)
__global__ void k_cg_scalar_rho(CgState* __restrict__ st, const double* __restrict__ sums, int itn) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0) return;
  CgState s = st[c];
  s.upd = s.active;
  if (s.active) {
    s.rho = sums[c];
    s.beta = (itn > 1 && s.rho_old > 0) ? s.rho / s.rho_old : 0.0;
  }
  st[c] = s;
}

// p = r + beta p
template <typename R>
__global__ void k_cg_dir(cplx<R>* __restrict__ pv, const cplx<R>* __restrict__ r, long long n,
                         const CgState* __restrict__ st) {
  const int c = blockIdx.y;
  const CgState s = st[c];
  if (!s.upd) return;
  const R be = (R)s.beta;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long o = (long long)c * n + i;
    const cplx<R> rv = r[o];
    cplx<R> pp = pv[o];
    pp.x = rv.x + be * pp.x;
    pp.y = rv.y + be * pp.y;
    pv[o] = pp;
  }
}

// alpha = rho / Re<p,q>; a non-positive curvature stops the column
__global__ void k_cg_scalar_alpha(CgState* __restrict__ st, const double* __restrict__ sums) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0) return;
  CgState s = st[c];
  if (s.upd) {
    const double pq = sums[c];
    if (pq > 0 && isfinite(pq)) {
      s.alpha = s.rho / pq;
    } else {
      s.alpha = 0;
      s.upd = 0;
      s.active = 0;
    }
  }
  st[c] = s;
}

// x += alpha p; r -= alpha q; partial ||r||^2
template <typename R>
__global__ void __launch_bounds__(kRedThreads) k_cg_step(cplx<R>* __restrict__ x, cplx<R>* __restrict__ r,
                                                         const cplx<R>* __restrict__ pv,
                                                         const cplx<R>* __restrict__ q, long long n,
                                                         const CgState* __restrict__ st,
                                                         double* __restrict__ partial) {
  __shared__ double sh[kRedThreads];
  const int c = blockIdx.y;
  const int nblk = gridDim.x;
  const CgState s = st[c];
  const R al = (R)s.alpha;
  double acc = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)nblk * blockDim.x) {
    const long long o = (long long)c * n + i;
    cplx<R> rv = r[o];
    if (s.upd) {
      const cplx<R> pp = pv[o], qq = q[o];
      cplx<R> xv = x[o];
      xv.x += al * pp.x;
      xv.y += al * pp.y;
      x[o] = xv;
      rv.x -= al * qq.x;
      rv.y -= al * qq.y;
      r[o] = rv;
    }
    acc += (double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[(long long)c * nblk + blockIdx.x] = t;
}

__global__ void k_cg_scalar_res(CgState* __restrict__ st, const double* __restrict__ sums, double tol) {
  const int c = blockIdx.x;
  if (threadIdx.x != 0) return;
  CgState s = st[c];
  if (s.upd) {
    s.relres = sqrt(sums[c]) / s.bnorm;
    s.rho_old = s.rho;
    s.iters += 1;
    s.active = s.relres > tol;
  }
  st[c] = s;
}

// P = exp(i angle(sum of the in-bounds 3x3x3 neighbours)) per coil
// (convn(x, ones(3), 'same'), iNUFT.m:147-149; angle(0) = 0)
template <typename R>
__global__ void k_phase_est(const cplx<R>* __restrict__ x, cplx<R>* __restrict__ P, int N1, int N2,
                            int N3, int nc) {
  const long long V = (long long)N1 * N2 * N3;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= V * nc) return;
  const int c = (int)(idx / V);
  const long long v = idx % V;
  const int ix = (int)(v % N1), iy = (int)((v / N1) % N2), iz = (int)(v / ((long long)N1 * N2));
  const cplx<R>* xc = x + (long long)c * V;
  double re = 0, im = 0;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = iz + dz;
    if (z < 0 || z >= N3) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = iy + dy;
      if (y < 0 || y >= N2) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = ix + dx;
        if (xx < 0 || xx >= N1) continue;
        const cplx<R> a = xc[xx + (long long)N1 * (y + (long long)N2 * z)];
        re += (double)a.x;
        im += (double)a.y;
      }
    }
  }
  const double th = atan2(im, re);
  double sn, cs;
  sincos(th, &sn, &cs);
  cplx<R> o; o.x = (R)cs; o.y = (R)sn;
  P[idx] = o;
}

// y = P x (conj = 0) or conj(P) x (conj = 1)
template <typename R>
__global__ void k_phase_mul(const cplx<R>* __restrict__ P, const cplx<R>* __restrict__ x,
                            cplx<R>* __restrict__ y, long long n, int conj) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> a = P[i], b = x[i];
  const R ay = conj ? -a.y : a.y;
  cplx<R> o;
  o.x = a.x * b.x - ay * b.y;
  o.y = a.x * b.y + ay * b.x;
  y[i] = o;
}

// y = conj(P) y + i lam2 Im(x)
template <typename R>
__global__ void k_phase_finish(const cplx<R>* __restrict__ P, const cplx<R>* __restrict__ x,
                               cplx<R>* __restrict__ y, long long n, R lam2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> a = P[i], b = y[i];
  cplx<R> o;
  o.x = a.x * b.x + a.y * b.y;
  o.y = a.x * b.y - a.y * b.x + lam2 * x[i].y;
  y[i] = o;
}

template <typename R>
static int phase_cg_t(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, double lambda,
                      const void* wsamp, const void* P, double tol, void* xout, nufft3d_info* info) {
  using C = cplx<R>;
  const long long n = (long long)p->N[0] * p->N[1] * p->N[2];
  const long long tot = n * nc;
  const size_t vec = (size_t)tot * sizeof(C);
  // vectors: r, p, q, u (P p) ; x is the caller's output
  N3D_TRY(ensure(p->ws_solve, 4 * vec));
  unsigned char* base = (unsigned char*)p->ws_solve.p;
  C* r = (C*)(base + 0 * vec);
  C* pv = (C*)(base + 1 * vec);
  C* q = (C*)(base + 2 * vec);
  C* u = (C*)(base + 3 * vec);
  C* x = (C*)xout;
  const int nb = red_blocks(n);
  N3D_TRY(ensure(p->ws_red, (size_t)nc * (nb + 2) * sizeof(double) + nc * sizeof(CgState) + 64));
  double* part = (double*)p->ws_red.p;
  double* sums = part + (size_t)nc * nb;
  double* bn2 = sums + nc;
  CgState* st = (CgState*)(bn2 + nc);
  const dim3 grid(nb, nc);
  const R lam2 = (R)(lambda * lambda);
  hipLaunchKernelGGL((k_dot_partial<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)b,
                     (const C*)b, n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, bn2);
  hipLaunchKernelGGL((k_cg_init<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)b, n,
                     (const double*)bn2, st, x, r, pv, tol);
  std::vector<CgState> hs(nc);
  auto fetch = [&]() -> int {
    N3D_HIP(hipMemcpyAsync(hs.data(), st, nc * sizeof(CgState), hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
    return NUFFT3D_OK;
  };
  N3D_TRY(fetch());
  for (int itn = 1; itn <= maxit; ++itn) {
    bool any = false;
    for (int c = 0; c < nc; ++c) any = any || hs[c].active;
    if (!any) break;
    hipLaunchKernelGGL((k_dot_partial<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)r,
                       (const C*)r, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    hipLaunchKernelGGL(k_cg_scalar_rho, dim3(nc), dim3(1), 0, p->stream, st, (const double*)sums, itn);
    hipLaunchKernelGGL((k_cg_dir<R>), grid, dim3(kRedThreads), 0, p->stream, pv, (const C*)r, n,
                       (const CgState*)st);
    {
      ProfScope ps(p, "k_phase_mul");
      hipLaunchKernelGGL((k_phase_mul<R>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, p->stream,
                         (const C*)P, (const C*)pv, u, tot, 0);
    }
    N3D_TRY(apply_normal(p, u, nc, damp, wsamp, q));
    {
      ProfScope ps(p, "k_phase_finish");
      hipLaunchKernelGGL((k_phase_finish<R>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         p->stream, (const C*)P, (const C*)pv, q, tot, lam2);
    }
    hipLaunchKernelGGL((k_dot_partial<R>), grid, dim3(kRedThreads), 0, p->stream, (const C*)pv,
                       (const C*)q, n, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    hipLaunchKernelGGL(k_cg_scalar_alpha, dim3(nc), dim3(1), 0, p->stream, st, (const double*)sums);
    hipLaunchKernelGGL((k_cg_step<R>), grid, dim3(kRedThreads), 0, p->stream, x, r, (const C*)pv,
                       (const C*)q, n, (const CgState*)st, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(nc), dim3(kRedThreads), 0, p->stream, part, nb, sums);
    hipLaunchKernelGGL(k_cg_scalar_res, dim3(nc), dim3(1), 0, p->stream, st, (const double*)sums, tol);
    N3D_HIP(hipGetLastError());
    N3D_TRY(fetch());
  }
  N3D_TRY(fetch());
  if (info) {
    info->nrhs = nc;
    for (int c = 0; c < nc && c < 64; ++c) {
      info->iters[c] = hs[c].iters;
      info->relres[c] = hs[c].relres;
    }
  }
  return NUFFT3D_OK;
}

int phase_estimate(nufft3d_plan* p, const void* x, int nc, void* P) {
  const long long tot = (long long)p->N[0] * p->N[1] * p->N[2] * nc;
  ProfScope ps(p, "k_phase_est");
  if (p->prec == 2)
    hipLaunchKernelGGL((k_phase_est<double>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       p->stream, (const double2*)x, (double2*)P, p->N[0], p->N[1], p->N[2], nc);
  else
    hipLaunchKernelGGL((k_phase_est<float>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       p->stream, (const float2*)x, (float2*)P, p->N[0], p->N[1], p->N[2], nc);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int phase_mul(nufft3d_plan* p, const void* P, const void* x, void* y, int nc, bool conj) {
  const long long tot = (long long)p->N[0] * p->N[1] * p->N[2] * nc;
  if (p->prec == 2)
    hipLaunchKernelGGL((k_phase_mul<double>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       p->stream, (const double2*)P, (const double2*)x, (double2*)y, tot, conj ? 1 : 0);
  else
    hipLaunchKernelGGL((k_phase_mul<float>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       p->stream, (const float2*)P, (const float2*)x, (float2*)y, tot, conj ? 1 : 0);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int phase_cg(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, double lambda,
             const void* wsamp, const void* P, double tol, void* x, nufft3d_info* info) {
  return p->prec == 2 ? phase_cg_t<double>(p, b, nc, maxit, damp, lambda, wsamp, P, tol, x, info)
                      : phase_cg_t<float>(p, b, nc, maxit, damp, lambda, wsamp, P, tol, x, info);
}

// ---------------------------------------------------------------- compressed sensing
// iNUFT.m:170-189: FISTA on the normal equations (A'WA + damp^2) x = A'W raw
// with a Haar-domain sparsity prox (the method of the external fistaN/HWT):
//   z = y - (A y - b)/L; x' = Q' soft(Q z, t), t zeroing a fraction lambda
//   of all coefficients; y = x' + ((t_k - 1)/t_{k+1})(x' - x).
// All coils form one vector (the reference reshapes b to one column).
// Q: orthonormal 3-D Haar, per level butterflies (a +- b)/sqrt(2) along x,
// y, z of the low corner (low half first), levels while every non-singleton
// active length is even, at most 3.

constexpr int kHaarMaxLevels = 3;

static int haar_levels(const int N[3]) {
  int a[3] = {N[0], N[1], N[2]};
  int lv = 0;
  while (lv < kHaarMaxLevels) {
    bool ok = false, all_even = true;
    for (int d = 0; d < 3; ++d)
      if (N[d] > 1) {
        ok = true;
        all_even = all_even && a[d] % 2 == 0;
      }
    if (!ok || !all_even) break;
    for (int d = 0; d < 3; ++d)
      if (N[d] > 1) a[d] /= 2;
    ++lv;
  }
  return lv;
}

// one butterfly pass along `axis` over the active block [a1 a2 a3]; every
// other element is copied through
template <typename R>
__global__ void k_haar_axis(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int N1, int N2,
                            int N3, long long total, int a1, int a2, int a3, int axis, int inverse) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long V = (long long)N1 * N2 * N3;
  const long long v = i % V;
  const long long cb = i - v;
  const int x = (int)(v % N1), y = (int)((v / N1) % N2), z = (int)(v / ((long long)N1 * N2));
  if (x >= a1 || y >= a2 || z >= a3) {
    out[i] = in[i];
    return;
  }
  const int len = axis == 0 ? a1 : (axis == 1 ? a2 : a3);
  const int pos = axis == 0 ? x : (axis == 1 ? y : z);
  const long long stride = axis == 0 ? 1 : (axis == 1 ? (long long)N1 : (long long)N1 * N2);
  const long long base = cb + v - (long long)pos * stride;
  const int h = len / 2;
  const R r2 = (R)0.70710678118654752440;
  cplx<R> o;
  if (!inverse) {
    const int j = pos < h ? pos : pos - h;
    const cplx<R> a = in[base + (long long)(2 * j) * stride], b = in[base + (long long)(2 * j + 1) * stride];
    if (pos < h) { o.x = (a.x + b.x) * r2; o.y = (a.y + b.y) * r2; }
    else { o.x = (a.x - b.x) * r2; o.y = (a.y - b.y) * r2; }
  } else {
    const int j = pos / 2;
    const cplx<R> lo = in[base + (long long)j * stride], hi = in[base + (long long)(h + j) * stride];
    if ((pos & 1) == 0) { o.x = (lo.x + hi.x) * r2; o.y = (lo.y + hi.y) * r2; }
    else { o.x = (lo.x - hi.x) * r2; o.y = (lo.y - hi.y) * r2; }
  }
  out[i] = o;
}

template <typename R>
static int haar_t(nufft3d_plan* p, cplx<R>* buf, cplx<R>* tmp, int nc, bool inverse) {
  const int N[3] = {p->N[0], p->N[1], p->N[2]};
  const int L = haar_levels(N);
  const long long total = (long long)N[0] * N[1] * N[2] * nc;
  const unsigned nb = (unsigned)((total + 255) / 256);
  cplx<R>* src = buf;
  cplx<R>* dst = tmp;
  ProfScope ps(p, "k_haar_axis");
  for (int k = 0; k < L; ++k) {
    const int lv = inverse ? L - 1 - k : k;
    const int a[3] = {N[0] > 1 ? N[0] >> lv : 1, N[1] > 1 ? N[1] >> lv : 1, N[2] > 1 ? N[2] >> lv : 1};
    for (int q = 0; q < 3; ++q) {
      const int axis = inverse ? 2 - q : q;
      if (N[axis] <= 1) continue;
      hipLaunchKernelGGL((k_haar_axis<R>), dim3(nb), dim3(256), 0, p->stream, (const cplx<R>*)src, dst,
                         N[0], N[1], N[2], total, a[0], a[1], a[2], axis, inverse ? 1 : 0);
      std::swap(src, dst);
    }
  }
  if (src != buf)
    N3D_HIP(hipMemcpyAsync(buf, src, (size_t)total * sizeof(cplx<R>), hipMemcpyDeviceToDevice, p->stream));
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R>
__global__ void k_magnitudes(const cplx<R>* __restrict__ w, long long n, double* __restrict__ m) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> a = w[i];
  m[i] = sqrt((double)a.x * (double)a.x + (double)a.y * (double)a.y);
}

// w *= max(0, 1 - t/|w|) with t = sorted[k] (zero where |w| <= t)
template <typename R>
__global__ void k_soft_thresh(cplx<R>* __restrict__ w, long long n, const double* __restrict__ sorted,
                              long long k) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double t = sorted[k];
  const cplx<R> a = w[i];
  const double m = sqrt((double)a.x * (double)a.x + (double)a.y * (double)a.y);
  const double s = m > t ? 1.0 - t / m : 0.0;
  cplx<R> o;
  o.x = (R)((double)a.x * s);
  o.y = (R)((double)a.y * s);
  w[i] = o;
}

// z = y - (g - b) / L
template <typename R>
__global__ void k_fista_grad(const cplx<R>* __restrict__ y, const cplx<R>* __restrict__ g,
                             const cplx<R>* __restrict__ b, cplx<R>* __restrict__ z, long long n, double invL) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> yv = y[i], gv = g[i], bv = b[i];
  const R s = (R)invL;
  cplx<R> o;
  o.x = yv.x - (gv.x - bv.x) * s;
  o.y = yv.y - (gv.y - bv.y) * s;
  z[i] = o;
}

// y = xn + c (xn - x); x = xn
template <typename R>
__global__ void k_fista_momentum(const cplx<R>* __restrict__ xn, cplx<R>* __restrict__ x,
                                 cplx<R>* __restrict__ y, long long n, double c) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cplx<R> a = xn[i], b = x[i];
  const R cc = (R)c;
  cplx<R> o;
  o.x = a.x + cc * (a.x - b.x);
  o.y = a.y + cc * (a.y - b.y);
  y[i] = o;
  x[i] = a;
}

template <typename R>
__global__ void k_scale_vec(cplx<R>* __restrict__ v, long long n, double s) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<R> a = v[i];
  a.x = (R)((double)a.x * s);
  a.y = (R)((double)a.y * s);
  v[i] = a;
}

template <typename R>
__global__ void k_fill_vec(cplx<R>* __restrict__ v, long long n, double re) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<R> a;
  a.x = (R)re;
  a.y = 0;
  v[i] = a;
}

template <typename R>
static int norm2_all(nufft3d_plan* p, const cplx<R>* v, long long n, double* out) {
  const int nb = red_blocks(n);
  N3D_TRY(ensure(p->ws_red, (size_t)(nb + 8) * sizeof(double)));
  double* part = (double*)p->ws_red.p;
  hipLaunchKernelGGL((k_dot_partial<R>), dim3(nb, 1), dim3(kRedThreads), 0, p->stream, v, v, n, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, p->stream, part, nb, part + nb);
  N3D_HIP(hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  return NUFFT3D_OK;
}

template <typename R>
static int fista_t(nufft3d_plan* p, const void* bv, int nc, int maxit, double damp, double frac,
                   const void* wsamp, double tol, void* xout, nufft3d_info* info) {
  using C = cplx<R>;
  const long long n = (long long)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t vec = (size_t)n * sizeof(C);
  const unsigned nb = (unsigned)((n + 255) / 256);
  N3D_TRY(ensure(p->ws_solve, 4 * vec));
  unsigned char* base = (unsigned char*)p->ws_solve.p;
  C* y = (C*)(base + 0 * vec);
  C* z = (C*)(base + 1 * vec);
  C* g = (C*)(base + 2 * vec);
  C* t = (C*)(base + 3 * vec);
  C* x = (C*)xout;
  const C* b = (const C*)bv;
  const long long nz = (long long)std::llround(frac * (double)n);
  double *keys = nullptr, *sorted = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_bytes = 0;
  struct Free {
    void** q[3];
    ~Free() { for (void** a : q) if (*a) (void)dev_free(*a); }
  } fr{{(void**)&keys, (void**)&sorted, &sort_tmp}};
  if (nz > 0) {
    N3D_HIP(dev_malloc(&keys, sizeof(double) * n));
    N3D_HIP(dev_malloc(&sorted, sizeof(double) * n));
    N3D_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys, sorted, (int)n, 0, 64,
                                              p->stream));
    N3D_HIP(dev_malloc(&sort_tmp, sort_bytes ? sort_bytes : 16));
  }
  if (info) {
    info->nrhs = 1;
    info->iters[0] = 0;
    info->relres[0] = 0;
  }
  hipLaunchKernelGGL((k_fill_vec<R>), dim3(nb), dim3(256), 0, p->stream, x, n, 0.0);
  double bn2 = 0;
  N3D_TRY(norm2_all<R>(p, b, n, &bn2));
  if (!(bn2 > 0) || maxit <= 0) return NUFFT3D_OK;
  // Lipschitz constant: 20 power iterations from the constant vector, x 1.05
  double L = 0;
  hipLaunchKernelGGL((k_fill_vec<R>), dim3(nb), dim3(256), 0, p->stream, y, n, 1.0 / std::sqrt((double)n));
  for (int k = 0; k < 20; ++k) {
    N3D_TRY(apply_normal(p, y, nc, damp, wsamp, g));
    double g2 = 0;
    N3D_TRY(norm2_all<R>(p, g, n, &g2));
    L = std::sqrt(g2);
    if (!(L > 0)) break;
    N3D_HIP(hipMemcpyAsync(y, g, vec, hipMemcpyDeviceToDevice, p->stream));
    hipLaunchKernelGGL((k_scale_vec<R>), dim3(nb), dim3(256), 0, p->stream, y, n, 1.0 / L);
  }
  if (!(L > 0)) { set_error("compressed sensing: the normal operator is zero"); return NUFFT3D_ERR_INVALID; }
  L *= 1.05;
  hipLaunchKernelGGL((k_fill_vec<R>), dim3(nb), dim3(256), 0, p->stream, y, n, 0.0);
  double tk = 1.0, rel = 0;
  int it = 0;
  for (it = 1; it <= maxit; ++it) {
    N3D_TRY(apply_normal(p, y, nc, damp, wsamp, g));
    hipLaunchKernelGGL((k_fista_grad<R>), dim3(nb), dim3(256), 0, p->stream, (const C*)y, (const C*)g, b,
                       z, n, 1.0 / L);
    N3D_TRY(haar_t<R>(p, z, t, nc, false));
    if (nz > 0) {
      ProfScope ps(p, "k_soft_thresh");
      hipLaunchKernelGGL((k_magnitudes<R>), dim3(nb), dim3(256), 0, p->stream, (const C*)z, n, keys);
      N3D_HIP(hipcub::DeviceRadixSort::SortKeys(sort_tmp, sort_bytes, keys, sorted, (int)n, 0, 64,
                                                p->stream));
      hipLaunchKernelGGL((k_soft_thresh<R>), dim3(nb), dim3(256), 0, p->stream, z, n,
                         (const double*)sorted, std::min(nz, n) - 1);
    }
    N3D_TRY(haar_t<R>(p, z, t, nc, true));
    double d2 = 0, x2 = 0;
    N3D_TRY(diff_norm2(p, z, x, n, &d2));
    N3D_TRY(norm2_all<R>(p, z, n, &x2));
    const double tn = (1.0 + std::sqrt(1.0 + 4.0 * tk * tk)) / 2.0;
    hipLaunchKernelGGL((k_fista_momentum<R>), dim3(nb), dim3(256), 0, p->stream, (const C*)z, x, y, n,
                       (tk - 1.0) / tn);
    N3D_HIP(hipGetLastError());
    tk = tn;
    rel = x2 > 0 ? std::sqrt(d2 / x2) : 0.0;
    if (rel <= tol) break;
  }
  N3D_HIP(hipStreamSynchronize(p->stream));
  if (info) {
    info->iters[0] = std::min(it, maxit);
    info->relres[0] = rel;
  }
  return NUFFT3D_OK;
}

int fista_solve(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, double frac,
                const void* wsamp, double tol, void* x, nufft3d_info* info) {
  if (p->N[0] * (long long)p->N[1] * p->N[2] * nc > 2147483647LL) {
    set_error("compressed sensing: more than 2^31 coefficients");
    return NUFFT3D_ERR_UNSUPPORTED;
  }
  return p->prec == 2 ? fista_t<double>(p, b, nc, maxit, damp, frac, wsamp, tol, x, info)
                      : fista_t<float>(p, b, nc, maxit, damp, frac, wsamp, tol, x, info);
}

}  // namespace nufft3d
