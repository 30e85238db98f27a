// Gridding kernels: forward interpolation (H*g) and deterministic adjoint
// spread (H'*k), both with Kaiser-Bessel weights computed on the fly from
// the bin-sorted samples instead of a stored sparse H.
//
// Reference: spmv / spmv_t (nufft_3d.m:259-282) over the H assembled at
// nufft_3d.m:153-212. Candidate neighbours of a sample are
// round(u*om)+d, d in -ceil(J/2)..ceil(J/2) per axis (nufft_3d.m:154,168-170),
// wrapped mod K (nufft_3d.m:178-180), kept when inside the radial or
// separable support with <= (nufft_3d.m:182-191).
#include "kb.h"
#include "nufft3d_internal.h"

namespace nufft3d {

struct GridGeom {
  int K1, K2, K3;
  int W1, W2, W3;
  int nc;     // coils in the grid / sample arrays
  int c0;     // first coil of this launch's coil group
  int cg;     // coils in this launch's group
  int nnz;
  long long M;
};

__device__ __forceinline__ int wrapk(int g, int K) {
  return g < 0 ? g + K : (g >= K ? g - K : g);
}

// Per-sample offsets u*om - round(u*om), double [3*nnz] in both precisions:
// the support test and the weights are evaluated in double and rounded to
// float exactly as the reference builds H (nufft_3d.m:173-191,160) so the
// fp32 path applies the same interpolation matrix as the fp64 path; only
// the data arithmetic differs.
__device__ __forceinline__ void load_off(const double* p, int s, double& x, double& y, double& z) {
  const double* q = p + 3LL * s;
  x = q[0]; y = q[1]; z = q[2];
}

// ---------------------------------------------------------------------------
// Forward: k[m,c] = scale * w_m * sum_g H(m,g) G[g,c]. One thread per sorted
// sample, candidates visited z-y-x so grid reads walk contiguous rows.
// user_layout: out[orig[s] + c*M] (reference [M nc]); else out[s*nc + c].
template <typename R, int CG>
__global__ void __launch_bounds__(256) k_interp(const int4* __restrict__ base,
                                                const double* __restrict__ off,
                                                const int* __restrict__ orig,
                                                const cplx<R>* __restrict__ grid,
                                                cplx<R>* __restrict__ out, int user_layout,
                                                const R* __restrict__ wsamp, R scale,
                                                GridGeom g, KBParams kp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.nnz) return;
  const int4 b = base[s];
  double fx, fy, fz;
  load_off(off, s, fx, fy, fz);
  R ar[CG], ai[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) { ar[c] = 0; ai[c] = 0; }
  for (int dz = -g.W3; dz <= g.W3; ++dz) {
    const double tz = (double)dz - fz;
    const double dz2 = tz * tz;
    const int gz = wrapk(b.z + dz, g.K3);
    for (int dy = -g.W2; dy <= g.W2; ++dy) {
      const double ty = (double)dy - fy;
      const double dy2 = ty * ty;
      const int gy = wrapk(b.y + dy, g.K2);
      const long long row = ((long long)gz * g.K2 + gy) * g.K1;
      for (int dx = -g.W1; dx <= g.W1; ++dx) {
        const double tx = (double)dx - fx;
        const double wd = kb_weight<double>(tx * tx, dy2, dz2, kp);
        if (wd == 0.0) continue;
        const R w = (R)wd;
        const int gx = wrapk(b.x + dx, g.K1);
        const cplx<R>* gp = grid + (row + gx) * g.nc + g.c0;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
          if (c < g.cg) {
            const cplx<R> v = gp[c];
            ar[c] += w * v.x;
            ai[c] += w * v.y;
          }
        }
      }
    }
  }
  R sc = scale;
  if (wsamp) sc *= wsamp[s];
  if (user_layout) {
    const long long m = orig[s];
#pragma unroll
    for (int c = 0; c < CG; ++c)
      if (c < g.cg) {
        cplx<R> v; v.x = ar[c] * sc; v.y = ai[c] * sc;
        out[m + (long long)(g.c0 + c) * g.M] = v;
      }
  } else {
#pragma unroll
    for (int c = 0; c < CG; ++c)
      if (c < g.cg) {
        cplx<R> v; v.x = ar[c] * sc; v.y = ai[c] * sc;
        out[(long long)s * g.nc + g.c0 + c] = v;
      }
  }
}

// ---------------------------------------------------------------------------
// Adjoint spread, output driven: one work-group per chunk of a tile's sample
// list, one thread per grid point of the tile. Each thread accumulates the
// samples of the list in list order, so the summation order is fixed by the
// plan alone (no atomics, bitwise reproducible, independent of batching).
// A tile whose list spans several chunks writes partial tiles that
// k_spread_reduce sums in chunk order.
struct TileGeom {
  int T1, T2, T3;       // tile edge per axis
  int nt1, nt2;         // tiles per axis (x, y)
  int npts;             // T1*T2*T3
};

template <typename R>
struct SampleStage {
  int bx, by, bz, pad;
  double fx, fy, fz;
  R w;
};

template <typename R, int CG, int SB>
__global__ void __launch_bounds__(512) k_spread(const SpreadChunk* __restrict__ chunks,
                                                const int* __restrict__ list,
                                                const int4* __restrict__ base,
                                                const double* __restrict__ off,
                                                const int* __restrict__ orig,
                                                const cplx<R>* __restrict__ samp, int user_layout,
                                                const R* __restrict__ wsamp,
                                                cplx<R>* __restrict__ grid,
                                                cplx<R>* __restrict__ partial, GridGeom g,
                                                TileGeom t, KBParams kp) {
  __shared__ SampleStage<R> sh[SB];
  __shared__ cplx<R> sv[SB * CG];
  const SpreadChunk ch = chunks[blockIdx.x];
  const int tid = threadIdx.x;
  const int tx = ch.tile % t.nt1;
  const int ty = (ch.tile / t.nt1) % t.nt2;
  const int tz = ch.tile / (t.nt1 * t.nt2);
  const int lx = tid % t.T1;
  const int ly = (tid / t.T1) % t.T2;
  const int lz = tid / (t.T1 * t.T2);
  const int gx = tx * t.T1 + lx;
  const int gy = ty * t.T2 + ly;
  const int gz = tz * t.T3 + lz;
  const bool own = tid < t.npts && gx < g.K1 && gy < g.K2 && gz < g.K3;
  R ar[CG], ai[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) { ar[c] = 0; ai[c] = 0; }

  for (int e0 = 0; e0 < ch.count; e0 += SB) {
    const int n = min(SB, ch.count - e0);
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      const int s = list[ch.start + e0 + i];
      const int4 b = base[s];
      SampleStage<R> st;
      st.bx = b.x; st.by = b.y; st.bz = b.z; st.pad = 0;
      load_off(off, s, st.fx, st.fy, st.fz);
      st.w = wsamp ? wsamp[s] : (R)1;
      sh[i] = st;
    }
    for (int i = tid; i < n * g.cg; i += blockDim.x) {
      const int e = i / g.cg, c = i % g.cg;
      const int s = list[ch.start + e0 + e];
      const long long idx = user_layout ? (long long)orig[s] + (long long)(g.c0 + c) * g.M
                                        : (long long)s * g.nc + g.c0 + c;
      sv[e * CG + c] = samp[idx];
    }
    __syncthreads();
    if (own) {
      for (int i = 0; i < n; ++i) {
        const SampleStage<R> st = sh[i];
        int dx = gx - st.bx, dy = gy - st.by, dz = gz - st.bz;
        if (dx > g.W1) dx -= g.K1; else if (dx < -g.W1) dx += g.K1;
        if (dy > g.W2) dy -= g.K2; else if (dy < -g.W2) dy += g.K2;
        if (dz > g.W3) dz -= g.K3; else if (dz < -g.W3) dz += g.K3;
        if (abs(dx) > g.W1 || abs(dy) > g.W2 || abs(dz) > g.W3) continue;
        const double ex = (double)dx - st.fx, ey = (double)dy - st.fy, ez = (double)dz - st.fz;
        const double wd = kb_weight<double>(ex * ex, ey * ey, ez * ez, kp);
        if (wd == 0.0) continue;
        const R w = (R)wd * st.w;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
          if (c < g.cg) {
            const cplx<R> v = sv[i * CG + c];
            ar[c] += w * v.x;
            ai[c] += w * v.y;
          }
        }
      }
    }
  }
  if (!own) return;
  if (ch.slot < 0) {
    cplx<R>* gp = grid + (((long long)gz * g.K2 + gy) * g.K1 + gx) * g.nc + g.c0;
#pragma unroll
    for (int c = 0; c < CG; ++c)
      if (c < g.cg) { cplx<R> v; v.x = ar[c]; v.y = ai[c]; gp[c] = v; }
  } else {
    cplx<R>* pp = partial + ((long long)ch.slot * t.npts + tid) * g.cg;
#pragma unroll
    for (int c = 0; c < CG; ++c)
      if (c < g.cg) { cplx<R> v; v.x = ar[c]; v.y = ai[c]; pp[c] = v; }
  }
}

// Sum the partial tiles of each split tile in chunk order (deterministic).
template <typename R>
__global__ void k_spread_reduce(const SplitTile* __restrict__ split,
                                const cplx<R>* __restrict__ partial,
                                cplx<R>* __restrict__ grid, GridGeom g, TileGeom t) {
  const SplitTile st = split[blockIdx.x];
  const int tx = st.tile % t.nt1;
  const int ty = (st.tile / t.nt1) % t.nt2;
  const int tz = st.tile / (t.nt1 * t.nt2);
  for (int i = threadIdx.x; i < t.npts * g.cg; i += blockDim.x) {
    const int p = i / g.cg, c = i % g.cg;
    const int lx = p % t.T1, ly = (p / t.T1) % t.T2, lz = p / (t.T1 * t.T2);
    const int gx = tx * t.T1 + lx, gy = ty * t.T2 + ly, gz = tz * t.T3 + lz;
    if (gx >= g.K1 || gy >= g.K2 || gz >= g.K3) continue;
    R sr = 0, si = 0;
    for (int j = 0; j < st.nslots; ++j) {
      const cplx<R> v = partial[((long long)(st.first_slot + j) * t.npts + p) * g.cg + c];
      sr += v.x;
      si += v.y;
    }
    cplx<R> o; o.x = sr; o.y = si;
    grid[(((long long)gz * g.K2 + gy) * g.K1 + gx) * g.nc + g.c0 + c] = o;
  }
}

template <typename R>
__global__ void k_zero_oob(const int* __restrict__ oob, long long noob, long long M, int nc,
                           cplx<R>* __restrict__ k) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= noob * nc) return;
  const long long j = i % noob, c = i / noob;
  cplx<R> z; z.x = 0; z.y = 0;
  k[oob[j] + c * M] = z;
}

// ks[s*nc + c] = w_s * k_user[orig[s] + c*M]
template <typename R>
__global__ void k_permute_in(const int* __restrict__ orig, int nnz, long long M, int nc,
                             const cplx<R>* __restrict__ ku, const R* __restrict__ wsamp,
                             cplx<R>* __restrict__ ks) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nnz * nc) return;
  const long long s = i / nc;
  const int c = (int)(i % nc);
  cplx<R> v = ku[orig[s] + (long long)c * M];
  if (wsamp) { const R w = wsamp[s]; v.x *= w; v.y *= w; }
  ks[i] = v;
}

// ---------------------------------------------------------------------------
static KBParams make_kb(const nufft3d_plan* p) {
  KBParams k;
  k.J2 = p->J * p->J;
  k.half2 = (p->J / 2) * (p->J / 2);
  k.alpha = p->alpha;
  // bessi0(alpha) on the host with the same polynomial (kernel.m:9)
  double a = p->alpha, b0;
  if (a < 3.75) {
    double y = a / 3.75; y = y * y;
    b0 = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
         y * (0.360768e-1 + y * 0.45813e-2)))));
  } else {
    double y = 3.75 / a;
    b0 = (exp(a) / sqrt(a)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
         y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
         y * (-0.1647633e-1 + y * 0.392377e-2))))))));
  }
  k.b0 = b0;
  k.J2f = (float)k.J2;
  k.inv_half2f = (float)(1.0 / k.half2);
  k.alphaf = (float)a;
  k.inv_b0f = (float)(1.0 / b0);
  k.radial = p->radial;
  return k;
}

static GridGeom make_geom(const nufft3d_plan* p, int nc, int c0, int cg) {
  GridGeom g;
  g.K1 = p->K[0]; g.K2 = p->K[1]; g.K3 = p->K[2];
  g.W1 = p->W[0]; g.W2 = p->W[1]; g.W3 = p->W[2];
  g.nc = nc; g.c0 = c0; g.cg = cg;
  g.nnz = (int)p->nnz;
  g.M = p->M;
  return g;
}

static TileGeom make_tiles(const nufft3d_plan* p) {
  TileGeom t;
  t.T1 = p->T[0]; t.T2 = p->T[1]; t.T3 = p->T[2];
  t.nt1 = p->ntile[0]; t.nt2 = p->ntile[1];
  t.npts = t.T1 * t.T2 * t.T3;
  return t;
}

template <typename R, int CG>
static void launch_interp(nufft3d_plan* p, const void* grid, int nc, int c0, int cg, void* out,
                          int user_layout, const void* wsamp, double scale) {
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  const int nb = (int)((p->nnz + 255) / 256);
  ProfScope ps(p, "k_interp");
  hipLaunchKernelGGL((k_interp<R, CG>), dim3(nb), dim3(256), 0, p->stream, p->d_base, (const double*)p->d_off,
                     p->d_orig, (const cplx<R>*)grid, (cplx<R>*)out, user_layout,
                     (const R*)wsamp, (R)scale, g, kp);
}

template <typename R>
static int interp_t(nufft3d_plan* p, const void* grid, int nc, void* out, int user_layout,
                    const void* wsamp, double scale) {
  if (p->nnz == 0) return NUFFT3D_OK;
  for (int c0 = 0; c0 < nc; c0 += 16) {
    const int cg = nc - c0 < 16 ? nc - c0 : 16;
    if (cg <= 1) launch_interp<R, 1>(p, grid, nc, c0, cg, out, user_layout, wsamp, scale);
    else if (cg <= 2) launch_interp<R, 2>(p, grid, nc, c0, cg, out, user_layout, wsamp, scale);
    else if (cg <= 4) launch_interp<R, 4>(p, grid, nc, c0, cg, out, user_layout, wsamp, scale);
    else if (cg <= 8) launch_interp<R, 8>(p, grid, nc, c0, cg, out, user_layout, wsamp, scale);
    else launch_interp<R, 16>(p, grid, nc, c0, cg, out, user_layout, wsamp, scale);
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int interp(nufft3d_plan* p, const void* grid, int nc, void* out, int user_layout,
           const void* wsamp, double scale) {
  return p->prec == 2 ? interp_t<double>(p, grid, nc, out, user_layout, wsamp, scale)
                      : interp_t<float>(p, grid, nc, out, user_layout, wsamp, scale);
}

template <typename R, int CG>
static void launch_spread(nufft3d_plan* p, const void* samp, int nc, int c0, int cg, void* grid,
                          void* partial, int user_layout, const void* wsamp) {
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  TileGeom t = make_tiles(p);
  const int threads = ((t.npts + 63) / 64) * 64;
  constexpr int SB = sizeof(R) == 8 ? 64 : 128;
  {
  ProfScope ps(p, "k_spread");
  hipLaunchKernelGGL((k_spread<R, CG, SB>), dim3(p->nchunks), dim3(threads), 0, p->stream,
                     p->d_chunks, p->d_list, p->d_base, (const double*)p->d_off, p->d_orig,
                     (const cplx<R>*)samp, user_layout, (const R*)wsamp, (cplx<R>*)grid,
                     (cplx<R>*)partial, g, t, kp);
  }
  if (p->nsplit > 0) {
    ProfScope ps(p, "k_spread_reduce");
    hipLaunchKernelGGL((k_spread_reduce<R>), dim3(p->nsplit), dim3(256), 0, p->stream,
                       p->d_split, (const cplx<R>*)partial, (cplx<R>*)grid, g, t);
  }
}

template <typename R>
static int spread_t(nufft3d_plan* p, const void* samp, int nc, void* grid, int user_layout,
                    const void* wsamp) {
  const int npts = p->T[0] * p->T[1] * p->T[2];
  const int cgmax = nc < 16 ? nc : 16;
  void* partial = nullptr;
  if (p->nslots > 0) {
    N3D_TRY(ensure(p->ws_part, (size_t)p->nslots * npts * cgmax * sizeof(cplx<R>)));
    partial = p->ws_part.p;
  }
  for (int c0 = 0; c0 < nc; c0 += 16) {
    const int cg = nc - c0 < 16 ? nc - c0 : 16;
    if (cg <= 1) launch_spread<R, 1>(p, samp, nc, c0, cg, grid, partial, user_layout, wsamp);
    else if (cg <= 2) launch_spread<R, 2>(p, samp, nc, c0, cg, grid, partial, user_layout, wsamp);
    else if (cg <= 4) launch_spread<R, 4>(p, samp, nc, c0, cg, grid, partial, user_layout, wsamp);
    else if (cg <= 8) launch_spread<R, 8>(p, samp, nc, c0, cg, grid, partial, user_layout, wsamp);
    else launch_spread<R, 16>(p, samp, nc, c0, cg, grid, partial, user_layout, wsamp);
  }
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int spread(nufft3d_plan* p, const void* samp, int nc, void* grid, int user_layout,
           const void* wsamp) {
  return p->prec == 2 ? spread_t<double>(p, samp, nc, grid, user_layout, wsamp)
                      : spread_t<float>(p, samp, nc, grid, user_layout, wsamp);
}

int zero_oob(nufft3d_plan* p, void* k_user, int nc) {
  const long long noob = p->M - p->nnz;
  if (noob <= 0) return NUFFT3D_OK;
  const long long n = noob * nc;
  const int nb = (int)((n + 255) / 256);
  ProfScope ps(p, "k_zero_oob");
  if (p->prec == 2)
    hipLaunchKernelGGL((k_zero_oob<double>), dim3(nb), dim3(256), 0, p->stream, p->d_oob, noob,
                       (long long)p->M, nc, (double2*)k_user);
  else
    hipLaunchKernelGGL((k_zero_oob<float>), dim3(nb), dim3(256), 0, p->stream, p->d_oob, noob,
                       (long long)p->M, nc, (float2*)k_user);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int permute_in(nufft3d_plan* p, const void* k_user, int nc, void* ks, const void* wsamp) {
  const long long n = p->nnz * (long long)nc;
  if (n == 0) return NUFFT3D_OK;
  const int nb = (int)((n + 255) / 256);
  ProfScope ps(p, "k_permute_in");
  if (p->prec == 2)
    hipLaunchKernelGGL((k_permute_in<double>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                       (int)p->nnz, (long long)p->M, nc, (const double2*)k_user,
                       (const double*)wsamp, (double2*)ks);
  else
    hipLaunchKernelGGL((k_permute_in<float>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                       (int)p->nnz, (long long)p->M, nc, (const float2*)k_user,
                       (const float*)wsamp, (float2*)ks);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

}  // namespace nufft3d
