// Plan construction and the C ABI (include/nufft3d.h).
//
// Reference constructor: @nufft_3d/nufft_3d.m:54-251.
//   - argument checks, alpha (Beatty), K = 2*ceil(N*u/2)   nufft_3d.m:60-111
//   - om_s = u*double(om), inclusive bounds, nnz           nufft_3d.m:120-127
//   - H (nufft_3d.m:129-212) is NOT stored: samples are sorted into spread
//     tiles and the Kaiser-Bessel weights are recomputed on the fly
//   - U = deap() (deap.m), d = density() (density.m)      nufft_3d.m:224-238
// Entry points: fNUFT.m, aNUFT.m, iNUFT.m.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nufft3d_internal.h"

namespace nufft3d {

std::shared_mutex& alloc_capture_mutex() {
  static std::shared_mutex m;
  return m;
}
thread_local bool t_capturing = false;

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

// Grow a workspace to at least `bytes`. hipFree waits for the whole device,
// stalling every other stream (the serving engine runs several), so growth
// over-allocates by a quarter: a stream of slightly larger batches does not
// free and reallocate each time.
int ensure(Workspace& w, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (w.bytes >= bytes) return NUFFT3D_OK;
  if (w.p) {
    if (dev_free(w.p) != hipSuccess) { (void)hipGetLastError(); }
    w.p = nullptr;
    w.bytes = 0;
    bytes += bytes / 4;
  }
  if (dev_malloc(&w.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    w.p = nullptr;
    set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
    return NUFFT3D_ERR_NOMEM;
  }
  w.bytes = bytes;
  return NUFFT3D_OK;
}

int ensure_host(Workspace& w, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (w.bytes >= bytes) return NUFFT3D_OK;
  if (w.p) { (void)hipHostFree(w.p); w.p = nullptr; w.bytes = 0; }
  if (hipHostMalloc(&w.p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    w.p = nullptr;
    set_error("pinned host allocation of " + std::to_string(bytes) + " bytes failed");
    return NUFFT3D_ERR_NOMEM;
  }
  w.bytes = bytes;
  return NUFFT3D_OK;
}

void prof_begin(nufft3d_plan* p, const char* name) {
  if (!p->prof.on) return;
  hipEvent_t e[2];
  for (int i = 0; i < 2; ++i) {
    if (!p->prof.pool.empty()) { e[i] = p->prof.pool.back(); p->prof.pool.pop_back(); }
    else if (hipEventCreate(&e[i]) != hipSuccess) { (void)hipGetLastError(); return; }
  }
  (void)hipEventRecord(e[0], p->stream);
  p->prof.recs.push_back({name, e[0], e[1]});
}

void prof_end(nufft3d_plan* p) {
  if (!p->prof.on || p->prof.recs.empty()) return;
  (void)hipEventRecord(p->prof.recs.back().b, p->stream);
}

static void release(Workspace& w, bool host = false) {
  if (!w.p) return;
  if (host) (void)hipHostFree(w.p); else (void)dev_free(w.p);
  w.p = nullptr;
  w.bytes = 0;
}

static bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) { (void)hipGetLastError(); return false; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// ---------------------------------------------------------------- deap (host, fp64)
// kernel.m:2-34 with complex arguments (deap evaluates beyond the support)
static std::complex<double> bessi0_c(std::complex<double> ax) {
  if (std::abs(ax) < 3.75) {
    std::complex<double> y = ax / 3.75;
    y = y * y;
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
                 y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  std::complex<double> y = 3.75 / ax;
  std::complex<double> r = (std::exp(ax) / std::sqrt(ax)) *
      (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
       y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
       y * (-0.1647633e-1 + y * 0.392377e-2))))))));
  if (ax.real() == 0) r *= 2.0;                          // kernel.m:33-34
  return r;
}

static double bessi0_r(double ax) {
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

static double kernel_host(double dist2, double J, double alpha) {
  const double t = 1.0 - dist2 / ((J / 2) * (J / 2));
  const double b0 = bessi0_r(alpha);
  if (t >= 0) return bessi0_r(alpha * sqrt(t)) / b0;
  const std::complex<double> arg = alpha * std::sqrt(std::complex<double>(t, 0.0));
  return bessi0_c(arg).real() / b0;
}

// deap.m:5-36: table at half-integer offsets, symmetric inverse FFT of the
// table padded to 2K per axis (closed form 0.5*(T0 + 2 sum T_m cos(pi m x/K))),
// crop to x in [-N/2, N/2), fftshift; U = 1/that (deap.m:98).
static void deap_host(nufft3d_plan* p) {
  const int nJ = (int)floor(p->J);
  const bool is2d = p->N[2] == 1;
  const int t1 = nJ + 1, t2 = nJ + 1, t3 = is2d ? 1 : nJ + 1;
  std::vector<double> T((size_t)t1 * t2 * t3);
  for (int dx = 0; dx < t1; ++dx)
    for (int dy = 0; dy < t2; ++dy)
      for (int dz = 0; dz < t3; ++dz) {
        const double dx2 = (dx / 2.0) * (dx / 2.0);
        const double dy2 = (dy / 2.0) * (dy / 2.0);
        const double dz2 = (dz / 2.0) * (dz / 2.0);
        double v;
        if (p->radial) v = kernel_host(dx2 + dy2 + dz2, p->J, p->alpha);
        else v = kernel_host(dx2, p->J, p->alpha) * kernel_host(dy2, p->J, p->alpha) *
                 kernel_host(dz2, p->J, p->alpha);
        T[dx + (size_t)t1 * (dy + (size_t)t2 * dz)] = v;
      }
  auto axis = [&](int n, int K, int nt) {
    std::vector<double> A((size_t)n * nt);
    for (int i = 0; i < n; ++i) {
      const double x = i - n / 2;
      for (int m = 0; m < nt; ++m)
        A[(size_t)i * nt + m] = 0.5 * (m == 0 ? 1.0 : 2.0) * cos(M_PI * m * x / K);
    }
    return A;
  };
  const int N1 = p->N[0], N2 = p->N[1], N3 = p->N[2];
  std::vector<double> A = axis(N1, p->K[0], t1), B = axis(N2, p->K[1], t2);
  std::vector<double> Cz = is2d ? std::vector<double>(1, 1.0) : axis(N3, p->K[2], t3);
  // stage 1: x
  std::vector<double> U1((size_t)N1 * t2 * t3, 0.0);
  for (int c = 0; c < t3; ++c)
    for (int b = 0; b < t2; ++b)
      for (int i = 0; i < N1; ++i) {
        double s = 0;
        for (int a = 0; a < t1; ++a) s += A[(size_t)i * t1 + a] * T[a + (size_t)t1 * (b + (size_t)t2 * c)];
        U1[i + (size_t)N1 * (b + (size_t)t2 * c)] = s;
      }
  std::vector<double> U2((size_t)N1 * N2 * t3, 0.0);
  for (int c = 0; c < t3; ++c)
    for (int j = 0; j < N2; ++j)
      for (int i = 0; i < N1; ++i) {
        double s = 0;
        for (int b = 0; b < t2; ++b) s += B[(size_t)j * t2 + b] * U1[i + (size_t)N1 * (b + (size_t)t2 * c)];
        U2[i + (size_t)N1 * (j + (size_t)N2 * c)] = s;
      }
  p->h_U.assign((size_t)N1 * N2 * N3, 0.0);
  for (int k = 0; k < N3; ++k)
    for (int j = 0; j < N2; ++j)
      for (int i = 0; i < N1; ++i) {
        double s = 0;
        if (is2d) s = U2[i + (size_t)N1 * j];
        else
          for (int c = 0; c < t3; ++c) s += Cz[(size_t)k * t3 + c] * U2[i + (size_t)N1 * (j + (size_t)N2 * c)];
        p->h_U[i + (size_t)N1 * (j + (size_t)N2 * k)] = 1.0 / s;
      }
}

// ---------------------------------------------------------------- binning kernels
struct BinGeom {
  double u;
  int K1, K2, K3;
  int T1, T2, T3;
  int nt1, nt2, nt3;
  int ntiles;
  int W1, W2, W3;
  int is2d;
  int mbits;    // bits per axis of the Morton tile key (0 = linear tile order)
};

// Interleave the low `bits` bits of (x, y, z) as z y x ... z y x (x lowest):
// the Z-order curve, under which consecutive keys stay in compact 3-D blocks
__host__ __device__ __forceinline__ unsigned morton3(unsigned x, unsigned y, unsigned z, int bits) {
  unsigned k = 0;
  for (int b = 0; b < bits; ++b)
    k |= (((x >> b) & 1u) << (3 * b)) | (((y >> b) & 1u) << (3 * b + 1)) |
         (((z >> b) & 1u) << (3 * b + 2));
  return k;
}

__device__ __forceinline__ bool sample_ok(const double* om, long long m, const BinGeom& g,
                                          double& x, double& y, double& z) {
  x = g.u * om[3 * m + 0];
  y = g.u * om[3 * m + 1];
  z = g.u * om[3 * m + 2];
  // nufft_3d.m:123-125 (the 2-D bug `a & b | c` makes every 2-D sample in bounds)
  bool ok = x >= -g.K1 / 2.0 && x <= g.K1 / 2.0 - 1;
  ok = y >= -g.K2 / 2.0 && y <= g.K2 / 2.0 - 1 && ok;
  ok = (z >= -g.K3 / 2.0 && z <= g.K3 / 2.0 - 1 && ok) || g.is2d;
  return ok;
}

__device__ __forceinline__ int wrapb(double r, int K) {
  int b = (int)r;
  b %= K;
  return b < 0 ? b + K : b;
}

__global__ void k_keys(const double* __restrict__ om, long long M, BinGeom g,
                       unsigned* __restrict__ keys, int* __restrict__ vals, int* __restrict__ counts) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double x, y, z;
  // out-of-bounds samples sort last
  unsigned key = g.mbits > 0 ? (1u << (3 * g.mbits)) : (unsigned)g.ntiles;
  if (sample_ok(om, m, g, x, y, z)) {
    const int bx = wrapb(round(x), g.K1), by = wrapb(round(y), g.K2), bz = wrapb(round(z), g.K3);
    const int tx = bx / g.T1, ty = by / g.T2, tz = bz / g.T3;
    key = g.mbits > 0 ? morton3(tx, ty, tz, g.mbits)
                      : (unsigned)(tz * g.nt2 * g.nt1 + ty * g.nt1 + tx);
  } else {
    atomicAdd(counts, 1);
  }
  keys[m] = key;
  vals[m] = (int)m;
}

// inv[orig[s]] = s
__global__ void k_invert(const int* __restrict__ orig, long long nnz, int* __restrict__ inv) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnz) inv[orig[s]] = (int)s;
}

__global__ void k_gather(const double* __restrict__ om, const unsigned* __restrict__ skeys,
                         const int* __restrict__ svals, long long M, long long nnz, BinGeom g,
                         int* __restrict__ orig, int* __restrict__ oob, int4* __restrict__ base,
                         double* __restrict__ off) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int m = svals[s];
  if (s >= nnz) { oob[s - nnz] = m; return; }
  double x, y, z;
  sample_ok(om, m, g, x, y, z);
  // MATLAB round is half away from zero, as C round (nufft_3d.m:168-170)
  const double rx = round(x), ry = round(y), rz = round(z);
  int4 b;
  b.x = wrapb(rx, g.K1);
  b.y = wrapb(ry, g.K2);
  b.z = wrapb(rz, g.K3);
  b.w = (int)skeys[s];
  base[s] = b;
  orig[s] = m;
  // om_s - round(om_s) is exact in double; dx = d - f reproduces (x - om) of
  // nufft_3d.m:173-175 bit for bit
  double* o = off + 3 * s;
  o[0] = x - rx; o[1] = y - ry; o[2] = z - rz;
}

// The interpolation pattern of H: candidates round(u*om)+d, |d| <= ceil(J/2)
// per axis (nufft_3d.m:154,168-170), kept when 4*dist2 <= J^2 (radial) or
// 4*d^2 <= J^2 per axis (separable), evaluated in double exactly as
// nufft_3d.m:173-191 (dx = d - f with f = u*om - round(u*om) exact).
struct HitGeom {
  int K1, K2, K3, W1, W2, W3;
  double J2;
  int radial;
};

__device__ __forceinline__ bool hit_test(double tx, double ty, double tz, const HitGeom& g) {
#pragma clang fp contract(off)
  const double dx2 = tx * tx, dy2 = ty * ty, dz2 = tz * tz;
  if (g.radial) return 4.0 * (dx2 + dy2 + dz2) <= g.J2;
  return 4.0 * dx2 <= g.J2 && 4.0 * dy2 <= g.J2 && 4.0 * dz2 <= g.J2;
}

__device__ __forceinline__ int wrapg(int v, int K) {   // mod(v, K) (nufft_3d.m:178-180)
  const int r = v % K;
  return r < 0 ? r + K : r;
}

__global__ void k_hits_count(const int4* __restrict__ base, const double* __restrict__ off,
                             long long nnz, HitGeom g, long long* __restrict__ cnt_s,
                             unsigned long long* __restrict__ cnt_g) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int4 b = base[s];
  const double fx = off[3 * s], fy = off[3 * s + 1], fz = off[3 * s + 2];
  int n = 0;
  for (int dz = -g.W3; dz <= g.W3; ++dz)
    for (int dy = -g.W2; dy <= g.W2; ++dy)
      for (int dx = -g.W1; dx <= g.W1; ++dx) {
        if (!hit_test((double)dx - fx, (double)dy - fy, (double)dz - fz, g)) continue;
        const long long gi = ((long long)wrapg(b.z + dz, g.K3) * g.K2 + wrapg(b.y + dy, g.K2)) * g.K1 +
                             wrapg(b.x + dx, g.K1);
        atomicAdd(&cnt_g[gi], 1ULL);
        ++n;
      }
  cnt_s[s] = n;
}

// hits in sample-major, z-y-x candidate order: key = grid index, value =
// (sample << 16) | code, code = (dx+W1) | (dy+W2) << 5 | (dz+W3) << 10
__global__ void k_hits_fill(const int4* __restrict__ base, const double* __restrict__ off,
                            long long nnz, HitGeom g, const long long* __restrict__ spos,
                            unsigned* __restrict__ keys, unsigned long long* __restrict__ vals,
                            unsigned short* __restrict__ fcode) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int4 b = base[s];
  const double fx = off[3 * s], fy = off[3 * s + 1], fz = off[3 * s + 2];
  long long o = spos[s];
  for (int dz = -g.W3; dz <= g.W3; ++dz)
    for (int dy = -g.W2; dy <= g.W2; ++dy)
      for (int dx = -g.W1; dx <= g.W1; ++dx) {
        if (!hit_test((double)dx - fx, (double)dy - fy, (double)dz - fz, g)) continue;
        const long long gi = ((long long)wrapg(b.z + dz, g.K3) * g.K2 + wrapg(b.y + dy, g.K2)) * g.K1 +
                             wrapg(b.x + dx, g.K1);
        const unsigned code = (unsigned)(dx + g.W1) | ((unsigned)(dy + g.W2) << 5) |
                              ((unsigned)(dz + g.W3) << 10);
        keys[o] = (unsigned)gi;
        vals[o] = ((unsigned long long)s << 16) | code;
        fcode[o] = (unsigned short)code;
        ++o;
      }
}

__global__ void k_split_vals(const unsigned long long* __restrict__ vals, long long n,
                             uint2* __restrict__ hits) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long v = vals[i];
  hits[i] = make_uint2((unsigned)(v >> 16), (unsigned)(v & 0xffffu));
}

// brick lists (BrickLists): key (brick << 32 | sample) per hit of a point
// with at most rmax hits, all ones for the hits of denser points
__global__ void k_bl_keys(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                          long long ngrid, int K1, int K2, const int* __restrict__ brick_of,
                          int bx, int by, int rmax, unsigned long long* __restrict__ keys) {
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ngrid) return;
  const long long e0 = rowptr[gi], e1 = rowptr[gi + 1];
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const unsigned long long b =
      (unsigned long long)brick_of[((long long)(z >> 3) * by + (y >> 3)) * bx + (x >> 3)];
  const bool dense = e1 - e0 > rmax;
  for (long long e = e0; e < e1; ++e)
    keys[e] = dense ? ~0ull : (b << 32) | hits[e].x;
}

__global__ void k_bl_samp(const unsigned long long* __restrict__ ukeys, long long n,
                          int* __restrict__ samp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) samp[i] = (int)(ukeys[i] & 0xffffffffull);
}

__device__ __forceinline__ long long lower_bound_u64(const unsigned long long* a, long long n,
                                                     unsigned long long v) {
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ptr[b] = first unique key of brick b (b = 0..nbrick)
__global__ void k_bl_ptr(const unsigned long long* __restrict__ ukeys, long long n, long long nbrick,
                         long long* __restrict__ ptr) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbrick) return;
  ptr[b] = lower_bound_u64(ukeys, n, (unsigned long long)b << 32);
}

// per hit of a light point: (slot of its sample in the brick list) << 15 | code
__global__ void k_bl_ent(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                         long long ngrid, int K1, int K2,
                         const int* __restrict__ brick_of, int bx, int by, int rmax,
                         const unsigned long long* __restrict__ ukeys,
                         const long long* __restrict__ ptr, unsigned* __restrict__ ent) {
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ngrid) return;
  const long long e0 = rowptr[gi], e1 = rowptr[gi + 1];
  if (e1 - e0 > rmax) return;
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const int b = brick_of[((long long)(z >> 3) * by + (y >> 3)) * bx + (x >> 3)];
  const long long base = ptr[b], len = ptr[b + 1] - base;
  long long lo = 0;   // hits are in ascending sample order: search from the last slot
  for (long long e = e0; e < e1; ++e) {
    const uint2 h = hits[e];
    const unsigned long long key = ((unsigned long long)b << 32) | h.x;
    lo += lower_bound_u64(ukeys + base + lo, len - lo, key);
    ent[e] = ((unsigned)lo << 15) | h.y;
  }
}

// Split-adjoint lists, one z-slab of bricks (grid points [g0, g1), hits
// [e0, ...)) at a time: keys (brick << 32) | sample of the light points'
// hits, all ones for the dense points'
__global__ void k_wl_keys(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                          long long g0, long long g1, long long e0, int K1, int K2,
                          const int* __restrict__ brick_of, int bx, int by, int rmax,
                          unsigned long long* __restrict__ keys) {
  const long long gi = g0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const unsigned long long br =
      (unsigned long long)brick_of[((long long)(z >> 3) * by + (y >> 3)) * bx + (x >> 3)];
  const bool dense = b - a > rmax;
  for (long long e = a; e < b; ++e) keys[e - e0] = dense ? ~0ull : (br << 32) | hits[e].x;
}

// the slab's unique keys u[0, n) (ascending, light only) -> samp[off + i],
// and each brick's range [beg, end) in samp
__global__ void k_wl_lists(const unsigned long long* __restrict__ u, long long n, long long off,
                           int* __restrict__ samp, long long* __restrict__ beg,
                           long long* __restrict__ end) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long b = u[i] >> 32;
  samp[off + i] = (int)(u[i] & 0xffffffffull);
  if (i == 0 || (u[i - 1] >> 32) != b) beg[b] = off + i;
  if (i == n - 1 || (u[i + 1] >> 32) != b) end[b] = off + i + 1;
}

// per light hit of the slab: (slot of its sample in the brick's list, weight bits)
__global__ void k_wl_ent(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                         long long g0, long long g1, int K1, int K2,
                         const int* __restrict__ brick_of, int bx, int by, int rmax,
                         const int* __restrict__ samp, const long long* __restrict__ beg,
                         const long long* __restrict__ end, uint2* __restrict__ went) {
  const long long gi = g0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  if (b - a > rmax) return;
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const int br = brick_of[((long long)(z >> 3) * by + (y >> 3)) * bx + (x >> 3)];
  const long long base = beg[br], len = end[br] - base;
  long long lo = 0;   // hits are in ascending sample order: search from the last slot
  for (long long e = a; e < b; ++e) {
    const uint2 h = hits[e];
    long long l = lo, r = len;
    while (l < r) {
      const long long m = (l + r) >> 1;
      if ((unsigned)samp[base + m] < h.x) l = m + 1; else r = m;
    }
    lo = l;
    went[e] = make_uint2((unsigned)lo, h.y);
  }
}

// first sample of each dense-point segment (its launch order key)
__global__ void k_seg_first(const uint2* __restrict__ hits, const long long* __restrict__ starts,
                            int n, unsigned* __restrict__ first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) first[i] = hits[starts[i]].x;
}

__global__ void k_off32(const double* __restrict__ off, long long nnz, float4* __restrict__ o32) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  float4 v;
  v.x = (float)off[3 * s]; v.y = (float)off[3 * s + 1]; v.z = (float)off[3 * s + 2]; v.w = 0.f;
  o32[s] = v;
}

// ---------------------------------------------------------------- density kernels
template <typename R>
__global__ void k_fill_dens(cplx<R>* __restrict__ ks, R* __restrict__ d, long long n) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  d[s] = (R)1;
  cplx<R> v; v.x = 1; v.y = 0;
  ks[s] = v;
}

// d = d ./ max(H*H'*d, eps) (density.m:151)
template <typename R>
__global__ void k_dens_update(cplx<R>* __restrict__ ks, const cplx<R>* __restrict__ hhd,
                              R* __restrict__ d, long long n) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const R e = (R)2.220446049250313e-16;
  const R h = hhd[s].x;
  const R v = d[s] / (h > e ? h : e);
  d[s] = v;
  cplx<R> o; o.x = v; o.y = 0;
  ks[s] = o;
}

template <typename R>
__global__ void k_div_scalar(R* __restrict__ d, long long n, R q) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) d[s] = d[s] / q;
}

template <typename R>
__global__ void k_gather_real(const double* __restrict__ src, const int* __restrict__ orig,
                              const R* __restrict__ mul, long long n, R* __restrict__ dst) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  R v = (R)src[orig[s]];
  if (mul) v *= mul[s];
  dst[s] = v;
}

template <typename R>
__global__ void k_scatter_dens(const R* __restrict__ d, const int* __restrict__ orig,
                               const int* __restrict__ oob, long long nnz, long long M,
                               R* __restrict__ out) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nnz) out[orig[s]] = d[s];
  else if (s < M) out[oob[s - nnz]] = (R)0;
}

static inline unsigned nblocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <typename R>
static int density_t(nufft3d_plan* p) {
  const long long n = p->nnz;
  using C = cplx<R>;
  N3D_HIP(dev_malloc(&p->d_dens, sizeof(R) * (n > 0 ? n : 1)));
  if (n == 0) return NUFFT3D_OK;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2];
  N3D_TRY(ensure(p->ws_grid, ngrid * sizeof(C)));
  N3D_TRY(ensure(p->ws_samp, 2 * (size_t)n * sizeof(C)));
  C* ks = (C*)p->ws_samp.p;
  C* hhd = ks + n;
  R* d = (R*)p->d_dens;
  hipLaunchKernelGGL((k_fill_dens<R>), dim3(nblocks(n)), dim3(256), 0, p->stream, ks, d, n);
  for (int it = 0; it < 20; ++it) {                       // density.m:5,11-14
    N3D_TRY(spread(p, ks, 1, p->ws_grid.p, 0, nullptr));  // H'd
    N3D_TRY(interp(p, p->ws_grid.p, 1, hhd, 0, nullptr, 1.0));  // H(H'd)
    hipLaunchKernelGGL((k_dens_update<R>), dim3(nblocks(n)), dim3(256), 0, p->stream, ks,
                       (const C*)hhd, d, n);
  }
  double sum = 0;
  N3D_TRY(sum_real(p, d, n, &sum));
  const R q = (R)(sum / (double)p->nnz);                   // density.m:42
  hipLaunchKernelGGL((k_div_scalar<R>), dim3(nblocks(n)), dim3(256), 0, p->stream, d, n, q);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R>
static int user_density_t(nufft3d_plan* p, const double* dh, int64_t nd) {
  const long long n = p->nnz;
  N3D_HIP(dev_malloc(&p->d_dens, sizeof(R) * (n > 0 ? n : 1)));
  double* dd = nullptr;
  N3D_HIP(dev_malloc(&dd, sizeof(double) * nd));
  N3D_HIP(hipMemcpy(dd, dh, sizeof(double) * nd, hipMemcpyHostToDevice));
  if (n > 0)
    hipLaunchKernelGGL((k_gather_real<R>), dim3(nblocks(n)), dim3(256), 0, p->stream, dd,
                       p->d_orig, (const R*)nullptr, n, (R*)p->d_dens);
  N3D_HIP(hipStreamSynchronize(p->stream));
  N3D_HIP(dev_free(dd));
  // warning of nufft_3d.m:235-237: mean over all rows after zeroing out of bounds
  std::vector<double> tmp(dh, dh + nd);
  double s = 0;
  std::vector<char> okm(nd, 0);
  std::vector<int> orig(n);
  if (n > 0) N3D_HIP(hipMemcpy(orig.data(), p->d_orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  for (long long i = 0; i < n; ++i) okm[orig[i]] = 1;
  for (int64_t i = 0; i < nd; ++i) s += okm[i] ? tmp[i] : 0.0;
  if (s / (double)nd != 1.0) {
    char b[96];
    snprintf(b, sizeof b, "mean(density) should be 1 (%f)", s / (double)nd);
    p->warnings.push_back(b);
    if (p->verbose) printf("Warning: %s\n", b);
  }
  return NUFFT3D_OK;
}

// ---------------------------------------------------------------- plan build
// Segment list of one dense-point set: points with more than ds.rmax hits,
// in segments of ds.chunk_len hits. Processing order: by segment index
// within the point, then by the point's place in the Z-ordered 8^3 bricks
// (neighbouring dense points hold mostly the same spokes in ascending sample
// order, so their k-th segments read similar samples side by side); each
// segment keeps the partial slot of its place within its point, so the sums
// are independent of that order.
static int build_dense(nufft3d_plan* p, const std::vector<long long>& hcnt, long long ngrid,
                       DenseSet& ds) {
  std::vector<HeavySeg> hseg;
  std::vector<HeavyPt> hpt;
  std::vector<long long> seg_k;
  long long pos = 0;
  for (long long gi = 0; gi < ngrid; ++gi) {
    const int c = (int)hcnt[gi];
    if (c > ds.rmax) {
      const int nseg = (c + ds.chunk_len - 1) / ds.chunk_len;
      hpt.push_back({(int)gi, (int)hseg.size(), nseg, 0});
      for (int j = 0; j < nseg; ++j) {
        const long long a = pos + (long long)j * ds.chunk_len;
        const long long e = std::min(pos + c, a + ds.chunk_len);
        hseg.push_back({a, e, (int)gi, (int)hseg.size()});
        seg_k.push_back(j);
      }
    }
    pos += c;
  }
  if (!hseg.empty() && !(p->tune_flags & 1024)) {
    // launch order by the segment's first sample: segments running together
    // read nearby sample rows (the k-space centre's samples stay in L2
    // instead of being fetched again by every point's segment of the same
    // index); the partial slots, hence the sums, are unchanged
    const int n = (int)hseg.size();
    std::vector<long long> st(n);
    for (int i = 0; i < n; ++i) st[i] = hseg[i].start;
    long long* d_st = nullptr;
    unsigned* d_first = nullptr;
    if (dev_malloc(&d_st, sizeof(long long) * n) != hipSuccess ||
        dev_malloc(&d_first, sizeof(unsigned) * n) != hipSuccess) {
      if (d_st) (void)dev_free(d_st);
      (void)hipGetLastError();
      set_error("dense segment ordering: allocation failed");
      return NUFFT3D_ERR_NOMEM;
    }
    std::vector<unsigned> first(n);
    int st2 = hipMemcpy(d_st, st.data(), sizeof(long long) * n, hipMemcpyHostToDevice) != hipSuccess;
    if (!st2) {
      hipLaunchKernelGGL(k_seg_first, dim3((n + 255) / 256), dim3(256), 0, p->stream, p->d_hit,
                         (const long long*)d_st, n, d_first);
      st2 = hipStreamSynchronize(p->stream) != hipSuccess ||
            hipMemcpy(first.data(), d_first, sizeof(unsigned) * n, hipMemcpyDeviceToHost) != hipSuccess;
    }
    (void)dev_free(d_st);
    (void)dev_free(d_first);
    if (st2) {
      (void)hipGetLastError();
      set_error("dense segment ordering failed");
      return NUFFT3D_ERR_DEVICE;
    }
    std::vector<std::pair<unsigned long long, int>> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = {((unsigned long long)first[i] << 32) | (unsigned)i, i};
    std::sort(ord.begin(), ord.end());
    std::vector<HeavySeg> h2(n);
    for (int i = 0; i < n; ++i) h2[i] = hseg[ord[i].second];
    hseg.swap(h2);
  } else if (!hseg.empty()) {
    int bits = 1;
    while ((1 << bits) < std::max((p->K[0] + 7) / 8, std::max((p->K[1] + 7) / 8, (p->K[2] + 7) / 8)))
      ++bits;
    std::vector<std::pair<unsigned long long, int>> ord(hseg.size());
    for (size_t i = 0; i < hseg.size(); ++i) {
      const long long gi = hseg[i].g;
      const int x = (int)(gi % p->K[0]), y = (int)((gi / p->K[0]) % p->K[1]),
                z = (int)(gi / ((long long)p->K[0] * p->K[1]));
      const unsigned long long brick = morton3(x >> 3, y >> 3, z >> 3, bits);
      const unsigned local = (unsigned)((x & 7) | ((y & 7) << 3) | ((z & 7) << 6));
      ord[i] = {((unsigned long long)seg_k[i] << 40) | (brick << 9) | local, (int)i};
    }
    std::stable_sort(ord.begin(), ord.end());
    std::vector<HeavySeg> h2(hseg.size());
    for (size_t i = 0; i < ord.size(); ++i) h2[i] = hseg[ord[i].second];
    hseg.swap(h2);
  }
  ds.nhseg = (int)hseg.size();
  ds.nhpt = (int)hpt.size();
  N3D_HIP(dev_malloc(&ds.d_hseg, sizeof(HeavySeg) * std::max<size_t>(1, hseg.size())));
  N3D_HIP(dev_malloc(&ds.d_hpt, sizeof(HeavyPt) * std::max<size_t>(1, hpt.size())));
  if (!hseg.empty()) {
    N3D_HIP(hipMemcpy(ds.d_hseg, hseg.data(), sizeof(HeavySeg) * hseg.size(), hipMemcpyHostToDevice));
    N3D_HIP(hipMemcpy(ds.d_hpt, hpt.data(), sizeof(HeavyPt) * hpt.size(), hipMemcpyHostToDevice));
  }
  return NUFFT3D_OK;
}

// BrickLists of the plan for points with at most rmax hits: sort the
// (brick, sample) pairs of all their hits, keep the unique ones, and code
// every hit with its sample's slot in the brick's list.
static int build_brick_lists(nufft3d_plan* p, const std::vector<int>& brick_of, int bx, int by,
                             int rmax) {
  BrickLists& bl = p->bl;
  const long long ngrid = (long long)p->K[0] * p->K[1] * p->K[2];
  const long long nh = p->nhits;
  if (nh <= 0 || nh >= 0x7fffffffLL || p->prec != 1 || p->K[2] == 1) return NUFFT3D_OK;
  int bbits = 1;
  while ((1LL << bbits) <= p->nbrick) ++bbits;
  int* d_bof = nullptr;
  unsigned long long *keys = nullptr, *skeys = nullptr;
  int* d_nsel = nullptr;
  void* tmp = nullptr;
  auto cleanup = [&]() {
    if (d_bof) (void)dev_free(d_bof);
    if (keys) (void)dev_free(keys);
    if (skeys) (void)dev_free(skeys);
    if (d_nsel) (void)dev_free(d_nsel);
    if (tmp) (void)dev_free(tmp);
  };
  auto fail = [&](const char* what) {
    cleanup();
    set_error(std::string("brick list construction: ") + what);
    return NUFFT3D_ERR_NOMEM;
  };
  if (dev_malloc(&d_bof, sizeof(int) * brick_of.size()) != hipSuccess ||
      dev_malloc(&keys, sizeof(unsigned long long) * nh) != hipSuccess ||
      dev_malloc(&skeys, sizeof(unsigned long long) * nh) != hipSuccess ||
      dev_malloc(&d_nsel, sizeof(int)) != hipSuccess)
    return fail("allocation failed");
  N3D_HIP(hipMemcpy(d_bof, brick_of.data(), sizeof(int) * brick_of.size(), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_bl_keys, dim3(nblocks(ngrid)), dim3(256), 0, p->stream, p->d_rowptr,
                     p->d_hit, ngrid, p->K[0], p->K[1], d_bof, bx, by, rmax, keys);
  size_t tb1 = 0, tb2 = 0;
  N3D_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, keys, skeys, (int)nh, 0, 32 + bbits,
                                            p->stream));
  N3D_HIP(hipcub::DeviceSelect::Unique(nullptr, tb2, skeys, keys, d_nsel, (int)nh, p->stream));
  if (dev_malloc(&tmp, std::max(tb1, tb2) + 16) != hipSuccess) return fail("allocation failed");
  // the dense points' all-ones keys sort last within the sorted bits too
  // (every brick index is below 2^bbits - 1)
  N3D_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb1, keys, skeys, (int)nh, 0, 32 + bbits,
                                            p->stream));
  N3D_HIP(hipcub::DeviceSelect::Unique(tmp, tb2, skeys, keys, d_nsel, (int)nh, p->stream));
  int nsel = 0;
  N3D_HIP(hipMemcpyAsync(&nsel, d_nsel, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  unsigned long long last = 0;
  if (nsel > 0)
    N3D_HIP(hipMemcpy(&last, keys + nsel - 1, sizeof(last), hipMemcpyDeviceToHost));
  const long long total = (nsel > 0 && (last >> 32) >= (unsigned long long)p->nbrick) ? nsel - 1 : nsel;
  std::vector<long long> hptr(p->nbrick + 1);
  if (dev_malloc(&bl.d_ptr, sizeof(long long) * (p->nbrick + 1)) != hipSuccess ||
      dev_malloc(&bl.d_samp, sizeof(int) * std::max(1LL, total)) != hipSuccess)
    return fail("allocation failed");
  hipLaunchKernelGGL(k_bl_ptr, dim3(nblocks(p->nbrick + 1)), dim3(256), 0, p->stream,
                     (const unsigned long long*)keys, total, p->nbrick, bl.d_ptr);
  if (total > 0)
    hipLaunchKernelGGL(k_bl_samp, dim3(nblocks(total)), dim3(256), 0, p->stream,
                       (const unsigned long long*)keys, total, bl.d_samp);
  N3D_HIP(hipMemcpyAsync(hptr.data(), bl.d_ptr, sizeof(long long) * (p->nbrick + 1),
                         hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  long long mx = 0;
  for (long long b = 0; b < p->nbrick; ++b) mx = std::max(mx, hptr[b + 1] - hptr[b]);
  if (mx >= (1LL << 17)) {   // slots would not fit the 17 bits of an entry
    cleanup();
    (void)dev_free(bl.d_ptr); (void)dev_free(bl.d_samp);
    bl.d_ptr = nullptr; bl.d_samp = nullptr;
    return NUFFT3D_OK;
  }
  if (dev_malloc(&bl.d_ent, sizeof(unsigned) * nh) != hipSuccess) return fail("allocation failed");
  N3D_HIP(hipMemsetAsync(bl.d_ent, 0, sizeof(unsigned) * nh, p->stream));
  hipLaunchKernelGGL(k_bl_ent, dim3(nblocks(ngrid)), dim3(256), 0, p->stream, p->d_rowptr,
                     p->d_hit, ngrid, p->K[0], p->K[1], d_bof, bx, by, rmax,
                     (const unsigned long long*)keys, (const long long*)bl.d_ptr, bl.d_ent);
  N3D_HIP(hipGetLastError());
  N3D_HIP(hipStreamSynchronize(p->stream));
  cleanup();
  bl.rmax = rmax;
  bl.total = total;
  bl.max_len = (int)mx;
  return NUFFT3D_OK;
}

// BrickLists on first need (their sort is ~15 ms and 4 B per hit at
// 3M samples)
static int ensure_brick_lists(nufft3d_plan* p) {
  if (p->bl.d_ent || p->bl_tried || p->h_brick_of.empty()) return NUFFT3D_OK;
  p->bl_tried = true;
  const int bx = (p->K[0] + 7) / 8, by = (p->K[1] + 7) / 8;
  return build_brick_lists(p, p->h_brick_of, bx, by, p->dense[1].rmax);
}

// WideLists of a split-precision plan (points with at most rmax hits of the
// plan's weight table): per z-slab of bricks, sort the (brick, sample) keys
// of the slab's hits and keep the unique ones (pass 1 counts, pass 2 writes
// the lists and each hit's slot). A slab of 2^31 hits or more leaves the
// lists unbuilt (the gather spread is used).
static int build_wide_lists(nufft3d_plan* p, const std::vector<int>& brick_of, int bx, int by,
                            int rmax) {
  WideLists& wl = p->wl;
  const long long nh = p->nhits;
  if (nh <= 0 || !p->d_hitw || p->K[2] == 1 || p->nbrick <= 0) return NUFFT3D_OK;
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const long long plane = (long long)K1 * K2;
  const int nbz = (K3 + 7) / 8;
  std::vector<long long> eb(nbz + 1);
  for (int bz = 0; bz <= nbz; ++bz) {
    const long long gz = std::min<long long>(8LL * bz, K3) * plane;
    N3D_HIP(hipMemcpy(&eb[bz], p->d_rowptr + gz, sizeof(long long), hipMemcpyDeviceToHost));
  }
  long long smax = 0;
  for (int bz = 0; bz < nbz; ++bz) smax = std::max(smax, eb[bz + 1] - eb[bz]);
  if (smax >= 0x7fffffffLL) return NUFFT3D_OK;
  int bbits = 1;
  while ((1LL << bbits) <= p->nbrick) ++bbits;
  int* d_bof = nullptr;
  unsigned long long *keys = nullptr, *skeys = nullptr;
  int* d_nsel = nullptr;
  void* tmp = nullptr;
  auto cleanup = [&]() {
    if (d_bof) (void)dev_free(d_bof);
    if (keys) (void)dev_free(keys);
    if (skeys) (void)dev_free(skeys);
    if (d_nsel) (void)dev_free(d_nsel);
    if (tmp) (void)dev_free(tmp);
  };
  auto drop = [&]() {
    for (void* q : {(void*)wl.d_beg, (void*)wl.d_end, (void*)wl.d_samp, (void*)wl.d_went})
      if (q) (void)dev_free(q);
    wl.d_beg = nullptr; wl.d_end = nullptr; wl.d_samp = nullptr; wl.d_went = nullptr;
  };
  auto fail = [&](const char* what) {
    cleanup();
    drop();
    (void)hipGetLastError();
    set_error(std::string("split-adjoint list construction: ") + what);
    return NUFFT3D_ERR_NOMEM;
  };
  const size_t nk = (size_t)std::max<long long>(smax, 1);
  if (dev_malloc(&d_bof, sizeof(int) * brick_of.size()) != hipSuccess ||
      dev_malloc(&keys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&skeys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&d_nsel, sizeof(int)) != hipSuccess)
    return fail("allocation failed");
  N3D_HIP(hipMemcpy(d_bof, brick_of.data(), sizeof(int) * brick_of.size(), hipMemcpyHostToDevice));
  size_t tb1 = 0, tb2 = 0;
  N3D_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, keys, skeys, (int)nk, 0, 32 + bbits,
                                            p->stream));
  N3D_HIP(hipcub::DeviceSelect::Unique(nullptr, tb2, skeys, keys, d_nsel, (int)nk, p->stream));
  if (dev_malloc(&tmp, std::max(tb1, tb2) + 16) != hipSuccess) return fail("allocation failed");
  // one slab: its unique light keys left in keys[0, return value)
  auto slab = [&](int bz, long long& nlight) -> int {
    const long long g0 = 8LL * bz * plane, g1 = std::min<long long>(8LL * bz + 8, K3) * plane;
    const int n = (int)(eb[bz + 1] - eb[bz]);
    nlight = 0;
    if (n == 0) return NUFFT3D_OK;
    hipLaunchKernelGGL(k_wl_keys, dim3(nblocks(g1 - g0)), dim3(256), 0, p->stream, p->d_rowptr,
                       p->d_hitw, g0, g1, eb[bz], K1, K2, d_bof, bx, by, rmax, keys);
    N3D_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb1, keys, skeys, n, 0, 32 + bbits, p->stream));
    N3D_HIP(hipcub::DeviceSelect::Unique(tmp, tb2, skeys, keys, d_nsel, n, p->stream));
    int nsel = 0;
    N3D_HIP(hipMemcpyAsync(&nsel, d_nsel, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
    unsigned long long last = 0;
    if (nsel > 0) N3D_HIP(hipMemcpy(&last, keys + nsel - 1, sizeof(last), hipMemcpyDeviceToHost));
    nlight = (nsel > 0 && last == ~0ull) ? nsel - 1 : nsel;   // the dense points' all-ones key
    return NUFFT3D_OK;
  };
  long long total = 0;
  std::vector<long long> nl(nbz);
  for (int bz = 0; bz < nbz; ++bz) {
    if (slab(bz, nl[bz])) return fail("sort failed");
    total += nl[bz];
  }
  if (dev_malloc(&wl.d_beg, sizeof(long long) * p->nbrick) != hipSuccess ||
      dev_malloc(&wl.d_end, sizeof(long long) * p->nbrick) != hipSuccess ||
      dev_malloc(&wl.d_samp, sizeof(int) * std::max(1LL, total)) != hipSuccess ||
      dev_malloc(&wl.d_went, sizeof(uint2) * nh) != hipSuccess)
    return fail("allocation failed");
  N3D_HIP(hipMemsetAsync(wl.d_beg, 0, sizeof(long long) * p->nbrick, p->stream));
  N3D_HIP(hipMemsetAsync(wl.d_end, 0, sizeof(long long) * p->nbrick, p->stream));
  N3D_HIP(hipMemsetAsync(wl.d_went, 0, sizeof(uint2) * nh, p->stream));
  long long off = 0;
  for (int bz = 0; bz < nbz; ++bz) {
    long long n2 = 0;
    if (slab(bz, n2) || n2 != nl[bz]) return fail("sort failed");
    if (n2 > 0)
      hipLaunchKernelGGL(k_wl_lists, dim3(nblocks(n2)), dim3(256), 0, p->stream,
                         (const unsigned long long*)keys, n2, off, wl.d_samp, wl.d_beg, wl.d_end);
    const long long g0 = 8LL * bz * plane, g1 = std::min<long long>(8LL * bz + 8, K3) * plane;
    hipLaunchKernelGGL(k_wl_ent, dim3(nblocks(g1 - g0)), dim3(256), 0, p->stream, p->d_rowptr,
                       p->d_hitw, g0, g1, K1, K2, d_bof, bx, by, rmax, (const int*)wl.d_samp,
                       (const long long*)wl.d_beg, (const long long*)wl.d_end, wl.d_went);
    N3D_HIP(hipGetLastError());
    off += n2;
  }
  std::vector<long long> hb(p->nbrick), he(p->nbrick);
  N3D_HIP(hipMemcpyAsync(hb.data(), wl.d_beg, sizeof(long long) * p->nbrick, hipMemcpyDeviceToHost,
                         p->stream));
  N3D_HIP(hipMemcpyAsync(he.data(), wl.d_end, sizeof(long long) * p->nbrick, hipMemcpyDeviceToHost,
                         p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  long long mx = 0;
  for (long long b = 0; b < p->nbrick; ++b) mx = std::max(mx, he[b] - hb[b]);
  cleanup();
  wl.rmax = rmax;
  wl.total = total;
  wl.max_len = mx;
  return NUFFT3D_OK;
}

// WideLists when asked for (tune flag 8 at plan creation): 8 B per hit and a
// few seconds of sorting at 2.3G hits, for a kernel measured no faster than
// the gather spread on radial trajectories (about 10% faster on uniform
// density), so the autotuner times it only when the lists exist
static int ensure_wide_lists(nufft3d_plan* p) {
  if (p->wl.d_went || p->wl.tried || p->h_brick_of.empty() || !split_precision(p)) return NUFFT3D_OK;
  p->wl.tried = true;
  const int bx = (p->K[0] + 7) / 8, by = (p->K[1] + 7) / 8;
  return build_wide_lists(p, p->h_brick_of, bx, by, p->dense[1].rmax);
}

// ---------------------------------------------------------------- block-staged adjoint lists
// keys (block << 32 | sample) of every hit of grid points [g0, g1) (hits
// [e0, ...)), blocks numbered in launch order by blk_of
__global__ void k_bk_keys(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                          long long g0, long long g1, long long e0, int K1, int K2,
                          const int* __restrict__ blk_of, int nbx, int nby, int sxy,
                          unsigned long long* __restrict__ keys) {
  const long long gi = g0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const unsigned long long bk =
      (unsigned long long)blk_of[((long long)(z >> 2) * nby + (y >> sxy)) * nbx + (x >> sxy)];
  for (long long e = a; e < b; ++e) keys[e - e0] = (bk << 32) | hits[e].x;
}

// per hit of grid points [g0, g1): the slot of its sample in its block's
// ascending list asc[beg[b], end[b])
__global__ void k_bk_slot(const long long* __restrict__ rowptr, const uint2* __restrict__ hits,
                          long long g0, long long g1, int K1, int K2,
                          const int* __restrict__ blk_of, int nbx, int nby, int sxy,
                          const int* __restrict__ asc, const long long* __restrict__ beg,
                          const long long* __restrict__ end, unsigned* __restrict__ hslot) {
  const long long gi = g0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= g1) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  if (a == b) return;
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  const int bk = blk_of[((long long)(z >> 2) * nby + (y >> sxy)) * nbx + (x >> sxy)];
  const long long base = beg[bk], len = end[bk] - base;
  long long lo = 0;   // hits are in ascending sample order: search from the last slot
  for (long long e = a; e < b; ++e) {
    const unsigned s = hits[e].x;
    long long l = lo, r = len;
    while (l < r) {
      const long long m = (l + r) >> 1;
      if ((unsigned)asc[base + m] < s) l = m + 1; else r = m;
    }
    lo = l;
    hslot[e] = (unsigned)l;
  }
}

// a block's list dealt into its chunks: list entry i -> chunk i % nch, slot
// i / nch; the unused slots of a chunk hold its first sample (never read)
__global__ void k_bk_samp(const int* __restrict__ asc, const long long* __restrict__ beg,
                          const long long* __restrict__ end, const long long* __restrict__ cbase,
                          int C, int* __restrict__ samp) {
  const long long b = blockIdx.x;
  const long long base = beg[b], u = end[b] - base;
  const long long c0 = cbase[b], nch = cbase[b + 1] - c0;
  for (long long i = threadIdx.x; i < nch * C; i += blockDim.x) {
    const long long c = i / C, o = i % C;
    const long long li = o * nch + c;
    samp[(c0 + c) * C + o] = asc[base + (li < u ? li : c)];
  }
}

// a point's block and its thread in the block's work: lane (x&3, y&3, z&3)
// of wavefront ((y>>2)&1, (x>>2)&1) (wide 8x8x4 blocks; one wavefront of
// 4x4x4 points otherwise)
__device__ __forceinline__ void bk_point(long long gi, int K1, int K2, const int* blk_of, int nbx,
                                         int nby, int sxy, int& bk, int& t) {
  const int x = (int)(gi % K1), y = (int)((gi / K1) % K2), z = (int)(gi / ((long long)K1 * K2));
  bk = blk_of[((long long)(z >> 2) * nby + (y >> sxy)) * nbx + (x >> sxy)];
  t = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
  if (sxy == 3) t |= (((x >> 2) & 1) | (((y >> 2) & 1) << 1)) << 6;
}

// per point: the rank of each hit among the point's hits of the same chunk
// (its step in that chunk); the point's thread owns its counters, so the
// ranks follow the point's list order
__global__ void k_bk_rank(const long long* __restrict__ rowptr, long long ngrid, int K1, int K2,
                          const int* __restrict__ blk_of, int nbx, int nby, int sxy, int T,
                          const long long* __restrict__ cbase, const unsigned* __restrict__ hslot,
                          unsigned short* __restrict__ cnt, unsigned short* __restrict__ hk) {
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ngrid) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  if (a == b) return;
  int bk, t;
  bk_point(gi, K1, K2, blk_of, nbx, nby, sxy, bk, t);
  const long long c0 = cbase[bk], nch = cbase[bk + 1] - c0;
  for (long long e = a; e < b; ++e) {
    const long long c = (long long)hslot[e] % nch;
    unsigned short* q = cnt + (c0 + c) * T + t;
    const unsigned short k = *q;
    *q = (unsigned short)(k + 1);
    hk[e] = k;
  }
}

// entry groups of each (chunk, wavefront) of blocks of T points: the
// wavefront's largest per-lane step count over 4, rounded up
__global__ void k_bk_groups(const unsigned short* __restrict__ cnt, long long nchunk, int T,
                            long long* __restrict__ groups) {
  const int nw = T / 64;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nchunk * nw) return;
  const long long c = i / nw;
  const int w = (int)(i % nw);
  int m = 0;
  for (int l = 0; l < 64; ++l) m = max(m, (int)cnt[c * T + w * 64 + l]);
  groups[i] = (m + 3) / 4;
}

__global__ void k_bk_fill(const long long* __restrict__ rowptr, long long ngrid, int K1, int K2,
                          const int* __restrict__ blk_of, int nbx, int nby, int sxy,
                          const long long* __restrict__ cbase, const long long* __restrict__ eoff,
                          const unsigned* __restrict__ hslot, const unsigned short* __restrict__ hk,
                          const uint2* __restrict__ hitw, unsigned char* __restrict__ slot,
                          float* __restrict__ w) {
  const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ngrid) return;
  const long long a = rowptr[gi], b = rowptr[gi + 1];
  if (a == b) return;
  int bk, t;
  bk_point(gi, K1, K2, blk_of, nbx, nby, sxy, bk, t);
  const int nw = sxy == 3 ? 4 : 1, wv = t >> 6, lane = t & 63;
  const long long c0 = cbase[bk], nch = cbase[bk + 1] - c0;
  for (long long e = a; e < b; ++e) {
    const long long sl = hslot[e];
    const long long c = sl % nch, o = sl / nch;
    const int k = hk[e];
    const long long q = ((eoff[(c0 + c) * nw + wv] + k / 4) * 64 + lane) * 4 + (k & 3);
    slot[q] = (unsigned char)o;
    w[q] = __uint_as_float(hitw[e].y);
  }
}

static void drop_block_lists(nufft3d_plan* p) {
  BlockLists& bk = p->bk;
  for (void* q : {(void*)bk.d_blk, (void*)bk.d_cbase, (void*)bk.d_samp, (void*)bk.d_eoff,
                  (void*)bk.d_slot, (void*)bk.d_w, (void*)bk.d_piece, (void*)bk.d_red})
    if (q) (void)dev_free(q);
  const bool tried = bk.tried;
  bk = BlockLists();
  bk.tried = tried;
}

// BlockLists of a split-precision 3-D plan, from the plan's point-major hit
// lists and weight table. The (block, sample) pairs are sorted one z-slab of
// blocks at a time (any total hit count; a slab needs fewer than 2^31 hits).
static int build_block_lists(nufft3d_plan* p) {
  BlockLists& bk = p->bk;
  // wide: 8x8x4 blocks, four wavefronts, chunks of kBlkWideRows samples
  const bool wide = p->blk_wide != 0;
  const int C = wide ? kBlkWideRows : p->blk_chunk == 128 ? 128 : 64;   // samples staged per chunk
  const int T = wide ? 256 : 64, NW = T / 64, sxy = wide ? 3 : 2;
  const long long nh = p->nhits;
  if (nh <= 0 || !p->d_hitw || p->K[2] == 1) return NUFFT3D_OK;
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  const long long plane = (long long)K1 * K2, ngrid = plane * K3;
  const int bxy = 1 << sxy;
  const int nbx = (K1 + bxy - 1) / bxy, nby = (K2 + bxy - 1) / bxy, nbz = (K3 + 3) / 4;
  const long long nblk = (long long)nbx * nby * nbz;
  if (nblk >= (1LL << 30)) return NUFFT3D_OK;
  std::vector<long long> eb(nbz + 1);
  for (int z = 0; z <= nbz; ++z) {
    const long long gz = std::min<long long>(4LL * z, K3) * plane;
    N3D_HIP(hipMemcpy(&eb[z], p->d_rowptr + gz, sizeof(long long), hipMemcpyDeviceToHost));
  }
  long long smax = 0;
  for (int z = 0; z < nbz; ++z) smax = std::max(smax, eb[z + 1] - eb[z]);
  if (smax >= 0x7fffffffLL) return NUFFT3D_OK;
  // launch order: Z-order of the block coordinates
  std::vector<int> blk(nblk), blk_of(nblk);
  {
    int bits = 1;
    while ((1 << bits) < std::max(nbx, std::max(nby, nbz))) ++bits;
    std::vector<std::pair<unsigned long long, int>> ord;
    ord.reserve(nblk);
    for (int z = 0; z < nbz; ++z)
      for (int y = 0; y < nby; ++y)
        for (int x = 0; x < nbx; ++x)
          ord.push_back({(unsigned long long)morton3(x, y, z, bits), x | (y << 10) | (z << 20)});
    std::sort(ord.begin(), ord.end());
    for (long long i = 0; i < nblk; ++i) {
      blk[i] = ord[i].second;
      const int x = blk[i] & 1023, y = (blk[i] >> 10) & 1023, z = blk[i] >> 20;
      blk_of[((long long)z * nby + y) * nbx + x] = (int)i;
    }
  }
  int bbits = 1;
  while ((1LL << bbits) <= nblk) ++bbits;
  int* d_bof = nullptr;
  unsigned long long *keys = nullptr, *skeys = nullptr;
  int* d_nsel = nullptr;
  void* tmp = nullptr;
  int* asc = nullptr;
  long long *beg = nullptr, *end = nullptr, *grp = nullptr;
  unsigned* hslot = nullptr;
  unsigned short *cnt = nullptr, *hk = nullptr;
  auto cleanup = [&]() {
    for (void* q : {(void*)d_bof, (void*)keys, (void*)skeys, (void*)d_nsel, tmp, (void*)asc,
                    (void*)beg, (void*)end, (void*)grp, (void*)hslot, (void*)cnt, (void*)hk})
      if (q) (void)dev_free(q);
    d_bof = nullptr; keys = skeys = nullptr; d_nsel = nullptr; tmp = nullptr; asc = nullptr;
    beg = end = grp = nullptr; hslot = nullptr; cnt = hk = nullptr;
  };
  // a plan that cannot hold the lists keeps the gather spread
  auto fail = [&](const char* what) {
    cleanup();
    drop_block_lists(p);
    (void)hipGetLastError();
    if (p->verbose) printf("  block-staged adjoint lists not built: %s\n", what);
    return NUFFT3D_OK;
  };
  const size_t nk = (size_t)std::max<long long>(smax, 1);
  if (dev_malloc(&d_bof, sizeof(int) * nblk) != hipSuccess ||
      dev_malloc(&keys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&skeys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&d_nsel, sizeof(int)) != hipSuccess ||
      dev_malloc(&hslot, sizeof(unsigned) * nh) != hipSuccess ||
      dev_malloc(&beg, sizeof(long long) * nblk) != hipSuccess ||
      dev_malloc(&end, sizeof(long long) * nblk) != hipSuccess)
    return fail("allocation failed");
  if (hipMemcpy(d_bof, blk_of.data(), sizeof(int) * nblk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(beg, 0, sizeof(long long) * nblk, p->stream) != hipSuccess ||
      hipMemsetAsync(end, 0, sizeof(long long) * nblk, p->stream) != hipSuccess)
    return fail("copy failed");
  size_t tb1 = 0, tb2 = 0;
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, keys, skeys, (int)nk, 0, 32 + bbits,
                                        p->stream) != hipSuccess ||
      hipcub::DeviceSelect::Unique(nullptr, tb2, skeys, keys, d_nsel, (int)nk, p->stream) != hipSuccess ||
      dev_malloc(&tmp, std::max(tb1, tb2) + 16) != hipSuccess)
    return fail("sort set-up failed");
  // pass 1: per slab, the unique (block, sample) keys -> each block's list
  // length; the lists themselves are kept (total = sum of the lengths)
  std::vector<long long> nsel_slab(nbz, 0);
  auto slab_keys = [&](int z, int& nsel) -> bool {
    const long long g0 = 4LL * z * plane, g1 = std::min<long long>(4LL * z + 4, K3) * plane;
    const int n = (int)(eb[z + 1] - eb[z]);
    nsel = 0;
    if (n == 0) return true;
    hipLaunchKernelGGL(k_bk_keys, dim3(nblocks(g1 - g0)), dim3(256), 0, p->stream, p->d_rowptr,
                       p->d_hitw, g0, g1, eb[z], K1, K2, d_bof, nbx, nby, sxy, keys);
    if (hipcub::DeviceRadixSort::SortKeys(tmp, tb1, keys, skeys, n, 0, 32 + bbits, p->stream) != hipSuccess ||
        hipcub::DeviceSelect::Unique(tmp, tb2, skeys, keys, d_nsel, n, p->stream) != hipSuccess ||
        hipMemcpyAsync(&nsel, d_nsel, sizeof(int), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      return false;
    return true;
  };
  long long total = 0;
  for (int z = 0; z < nbz; ++z) {
    int ns = 0;
    if (!slab_keys(z, ns)) return fail("sort failed");
    nsel_slab[z] = ns;
    total += ns;
  }
  if (dev_malloc(&asc, sizeof(int) * std::max(1LL, total)) != hipSuccess)
    return fail("allocation failed");
  long long off = 0;
  for (int z = 0; z < nbz; ++z) {
    int ns = 0;
    if (!slab_keys(z, ns) || ns != nsel_slab[z]) return fail("sort failed");
    if (ns > 0)
      hipLaunchKernelGGL(k_wl_lists, dim3(nblocks(ns)), dim3(256), 0, p->stream,
                         (const unsigned long long*)keys, (long long)ns, off, asc, beg, end);
    const long long g0 = 4LL * z * plane, g1 = std::min<long long>(4LL * z + 4, K3) * plane;
    hipLaunchKernelGGL(k_bk_slot, dim3(nblocks(g1 - g0)), dim3(256), 0, p->stream, p->d_rowptr,
                       p->d_hitw, g0, g1, K1, K2, d_bof, nbx, nby, sxy, (const int*)asc,
                       (const long long*)beg, (const long long*)end, hslot);
    if (hipGetLastError() != hipSuccess) return fail("slot kernels failed");
    off += ns;
  }
  (void)dev_free(keys); keys = nullptr;
  (void)dev_free(skeys); skeys = nullptr;
  (void)dev_free(tmp); tmp = nullptr;
  // chunks per block (launch order), then the chunked lists
  std::vector<long long> hb(nblk), he(nblk), cbase(nblk + 1, 0);
  if (hipMemcpyAsync(hb.data(), beg, sizeof(long long) * nblk, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipMemcpyAsync(he.data(), end, sizeof(long long) * nblk, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess)
    return fail("copy failed");
  for (long long b = 0; b < nblk; ++b)
    cbase[b + 1] = cbase[b] + (he[b] - hb[b] + C - 1) / C;
  const long long nchunk = cbase[nblk];
  const long long nent = nchunk * NW;   // (chunk, wavefront) entry ranges
  if (dev_malloc(&bk.d_blk, sizeof(int) * nblk) != hipSuccess ||
      dev_malloc(&bk.d_cbase, sizeof(long long) * (nblk + 1)) != hipSuccess ||
      dev_malloc(&bk.d_samp, sizeof(int) * std::max(1LL, nchunk * C)) != hipSuccess ||
      dev_malloc(&bk.d_eoff, sizeof(long long) * (nent + 1)) != hipSuccess ||
      dev_malloc(&cnt, sizeof(unsigned short) * std::max(1LL, nchunk * T)) != hipSuccess ||
      dev_malloc(&hk, sizeof(unsigned short) * nh) != hipSuccess ||
      dev_malloc(&grp, sizeof(long long) * (nent + 1)) != hipSuccess)
    return fail("allocation failed");
  if (hipMemcpy(bk.d_blk, blk.data(), sizeof(int) * nblk, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(bk.d_cbase, cbase.data(), sizeof(long long) * (nblk + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(cnt, 0, sizeof(unsigned short) * std::max(1LL, nchunk * T), p->stream) != hipSuccess)
    return fail("copy failed");
  hipLaunchKernelGGL(k_bk_samp, dim3((unsigned)nblk), dim3(256), 0, p->stream, (const int*)asc,
                     (const long long*)beg, (const long long*)end, (const long long*)bk.d_cbase, C,
                     bk.d_samp);
  hipLaunchKernelGGL(k_bk_rank, dim3(nblocks(ngrid)), dim3(256), 0, p->stream, p->d_rowptr, ngrid,
                     K1, K2, d_bof, nbx, nby, sxy, T, (const long long*)bk.d_cbase,
                     (const unsigned*)hslot, cnt, hk);
  if (nent > 0)
    hipLaunchKernelGGL(k_bk_groups, dim3(nblocks(nent)), dim3(256), 0, p->stream,
                       (const unsigned short*)cnt, nchunk, T, grp);
  if (hipGetLastError() != hipSuccess) return fail("list kernels failed");
  std::vector<long long> hgrp(nent + 1, 0), eoff(nent + 1, 0);
  if (nent > 0 &&
      (hipMemcpyAsync(hgrp.data(), grp, sizeof(long long) * nent, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
       hipStreamSynchronize(p->stream) != hipSuccess))
    return fail("copy failed");
  for (long long c = 0; c < nent; ++c) eoff[c + 1] = eoff[c] + hgrp[c];
  const long long ngroups = eoff[nent];
  if (dev_malloc(&bk.d_slot, sizeof(unsigned) * std::max(1LL, ngroups * 64)) != hipSuccess ||
      dev_malloc(&bk.d_w, sizeof(float4) * std::max(1LL, ngroups * 64)) != hipSuccess)
    return fail("allocation failed");
  if (hipMemcpy(bk.d_eoff, eoff.data(), sizeof(long long) * (nent + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(bk.d_slot, 0xff, sizeof(unsigned) * std::max(1LL, ngroups * 64), p->stream) != hipSuccess ||
      hipMemsetAsync(bk.d_w, 0, sizeof(float4) * std::max(1LL, ngroups * 64), p->stream) != hipSuccess)
    return fail("copy failed");
  hipLaunchKernelGGL(k_bk_fill, dim3(nblocks(ngrid)), dim3(256), 0, p->stream, p->d_rowptr, ngrid,
                     K1, K2, d_bof, nbx, nby, sxy, (const long long*)bk.d_cbase,
                     (const long long*)bk.d_eoff, (const unsigned*)hslot, (const unsigned short*)hk,
                     (const uint2*)p->d_hitw, (unsigned char*)bk.d_slot, (float*)bk.d_w);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess)
    return fail("fill failed");
  // pieces: a block's chunks in runs of at most kBlkPieceRows entry rows
  // (at least one chunk each; a chunk's rows are its largest wavefront's);
  // blocks in several pieces sum partials
  std::vector<long long> crow(nchunk, 0);
  for (long long c = 0; c < nchunk; ++c)
    for (int w = 0; w < NW; ++w) crow[c] = std::max(crow[c], 4 * hgrp[c * NW + w]);
  // (wide blocks: four times the rows, so fewer partials to write and add)
  const long long prow = p->blk_piece > 0 ? p->blk_piece : wide ? 4 * kBlkPieceRows : kBlkPieceRows;
  std::vector<int4> pieces, red;
  long long npslot = 0;
  pieces.reserve(nblk + nblk / 4);
  for (long long b = 0; b < nblk; ++b) {
    const long long c0 = cbase[b], c1 = cbase[b + 1];
    if (c0 == c1) {
      pieces.push_back(make_int4((int)b, 0, 0, -1));
      continue;
    }
    std::vector<std::pair<int, int>> runs;
    long long c = c0;
    while (c < c1) {
      long long rows = 0, e = c;
      while (e < c1 && (e == c || rows + crow[e] <= prow)) rows += crow[e++];
      runs.push_back({(int)(c - c0), (int)(e - c0)});
      c = e;
    }
    if (runs.size() == 1) {
      pieces.push_back(make_int4((int)b, runs[0].first, runs[0].second, -1));
    } else {
      red.push_back(make_int4((int)b, (int)npslot, (int)runs.size(), 0));
      for (auto& r : runs) pieces.push_back(make_int4((int)b, r.first, r.second, (int)npslot++));
    }
  }
  if (npslot >= 0x7fffffffLL || (long long)pieces.size() >= 0x7fffffffLL)
    return fail("too many pieces");
  if (dev_malloc(&bk.d_piece, sizeof(int4) * pieces.size()) != hipSuccess ||
      dev_malloc(&bk.d_red, sizeof(int4) * std::max<size_t>(1, red.size())) != hipSuccess)
    return fail("allocation failed");
  if (hipMemcpy(bk.d_piece, pieces.data(), sizeof(int4) * pieces.size(), hipMemcpyHostToDevice) != hipSuccess ||
      (!red.empty() &&
       hipMemcpy(bk.d_red, red.data(), sizeof(int4) * red.size(), hipMemcpyHostToDevice) != hipSuccess))
    return fail("copy failed");
  cleanup();
  bk.C = C;
  bk.T = T;
  bk.nblk = nblk;
  bk.nchunk = nchunk;
  bk.ngroups = ngroups;
  bk.npiece = (long long)pieces.size();
  bk.nred = (long long)red.size();
  bk.npslot = npslot;
  if (p->verbose)
    printf("  block-staged adjoint: %lld blocks of %d points, %lld chunks of %d samples, %lld entry "
           "groups (%.2f B/hit), %lld pieces, %lld partial slots\n", nblk, T, nchunk, C, ngroups,
           5.0 * 4 * 64 * ngroups / (double)nh, bk.npiece, npslot);
  return NUFFT3D_OK;
}

// BlockLists when asked for (tune flag 4096 at plan creation) or timed by
// the autotuner: fp64-arithmetic 3-D plans with the weight table
int ensure_block_lists(nufft3d_plan* p) {
  if (p->bk.d_piece || p->bk.tried || p->prec != 2 || p->K[2] == 1 || !p->d_hitw)
    return NUFFT3D_OK;
  p->bk.tried = true;
  return build_block_lists(p);
}

// ---------------------------------------------------------------- block-staged forward lists
// grid index of a forward hit (code: candidate offset biased by W, 5 bits per
// axis) of a sample with wrapped base b, K >= 2W + 1 per axis
__device__ __forceinline__ long long fk_point(int4 b, unsigned code, int K1, int K2, int K3, int W) {
  const int dx = (int)(code & 31u) - W, dy = (int)((code >> 5) & 31u) - W, dz = (int)(code >> 10) - W;
  return ((long long)wrapg(b.z + dz, K3) * K2 + wrapg(b.y + dy, K2)) * K1 + wrapg(b.x + dx, K1);
}

// keys (group << 32 | grid index) of the hits of samples [s0, s1) (hits from e0)
__global__ void k_fk_keys(const long long* __restrict__ spos, const unsigned short* __restrict__ fcode,
                          const int4* __restrict__ base, long long s0, long long s1, long long e0,
                          int K1, int K2, int K3, int W, int G, unsigned long long* __restrict__ keys) {
  const long long s = s0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= s1) return;
  const int4 b = base[s];
  const unsigned long long grp = (unsigned long long)(s / G) << 32;
  for (long long e = spos[s]; e < spos[s + 1]; ++e)
    keys[e - e0] = grp | (unsigned long long)fk_point(b, fcode[e], K1, K2, K3, W);
}

// per hit of samples [s0, s1): the slot of its grid point in its group's
// ascending list lst[beg[j], end[j])
__global__ void k_fk_slot(const long long* __restrict__ spos, const unsigned short* __restrict__ fcode,
                          const int4* __restrict__ base, long long s0, long long s1, int K1, int K2,
                          int K3, int W, int G, const int* __restrict__ lst,
                          const long long* __restrict__ beg, const long long* __restrict__ end,
                          unsigned short* __restrict__ hslot) {
  const long long s = s0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= s1) return;
  const int4 b = base[s];
  const long long j = s / G;
  const long long bs = beg[j], len = end[j] - bs;
  for (long long e = spos[s]; e < spos[s + 1]; ++e) {
    const int gi = (int)fk_point(b, fcode[e], K1, K2, K3, W);
    long long l = 0, r = len;
    while (l < r) {
      const long long m = (l + r) >> 1;
      if (lst[bs + m] < gi) l = m + 1; else r = m;
    }
    hslot[e] = (unsigned short)l;
  }
}

// per sample: the rank of each hit among the sample's hits of the same chunk
__global__ void k_fk_rank(const long long* __restrict__ spos, long long nnz, int G,
                          const long long* __restrict__ cbase, const unsigned short* __restrict__ hslot,
                          unsigned short* __restrict__ cnt, unsigned short* __restrict__ hk) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const long long j = s / G;
  const int t = (int)(s % G);
  const long long c0 = cbase[j], nch = cbase[j + 1] - c0;
  for (long long e = spos[s]; e < spos[s + 1]; ++e) {
    const long long c = hslot[e] % nch;
    unsigned short* q = cnt + (c0 + c) * G + t;
    const unsigned short k = *q;
    *q = (unsigned short)(k + 1);
    hk[e] = k;
  }
}

// entry groups of each (chunk, wavefront) of a group of G samples: the
// wavefront's largest per-lane step count over 4, rounded up
__global__ void k_fk_groups(const unsigned short* __restrict__ cnt, long long nchunk, int G,
                            long long* __restrict__ groups) {
  const int nw = G / 64;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nchunk * nw) return;
  const long long c = i / nw;
  const int w = (int)(i % nw);
  int m = 0;
  for (int l = 0; l < 64; ++l) m = max(m, (int)cnt[c * G + w * 64 + l]);
  groups[i] = (m + 3) / 4;
}

static void drop_fwd_lists(nufft3d_plan* p) {
  FwdLists& fk = p->fk;
  for (void* q : {(void*)fk.d_cbase, (void*)fk.d_gpt, (void*)fk.d_eoff, (void*)fk.d_slot,
                  (void*)fk.d_w})
    if (q) (void)dev_free(q);
  const bool tried = fk.tried;
  fk = FwdLists();
  fk.tried = tried;
}

// FwdLists of an fp64-arithmetic 3-D plan with K >= 2*ceil(J/2) + 2 per axis
// (a grid point is reached by at most one candidate of a sample), from the
// sample-major hit codes. The (group, grid point) pairs are sorted a batch
// of groups at a time.
static int build_fwd_lists(nufft3d_plan* p, bool weights) {
  FwdLists& fk = p->fk;
  const long long nnz = p->nnz, nh = p->nhits;
  if (nh <= 0 || p->prec != 2 || p->K[2] == 1 || !p->d_fcode) return NUFFT3D_OK;
  const int W = p->W[0];
  for (int a = 0; a < 3; ++a)
    if (p->W[a] != W || p->K[a] < 2 * W + 2) return NUFFT3D_OK;
  const int K1 = p->K[0], K2 = p->K[1], K3 = p->K[2];
  // samples per group (one wavefront per 64) and grid rows per chunk; a
  // group's union of points is indexed by 16-bit slots
  const int G = p->fwd_group == 256 && 256LL * W * W * W < 65535 ? 256 : 64;
  const int C = G == 256 ? 255 : 64, NW = G / 64;
  const long long ngrp = (nnz + G - 1) / G;
  std::vector<long long> hspos(nnz + 1);
  N3D_HIP(hipMemcpy(hspos.data(), p->d_spos, sizeof(long long) * (nnz + 1), hipMemcpyDeviceToHost));
  // batches of whole groups with fewer than 2^28 hits each (one group alone if larger)
  std::vector<long long> bs;   // first sample of each batch
  for (long long s = 0; s < nnz;) {
    bs.push_back(s);
    long long e = s;
    while (e < nnz) {
      const long long n = std::min(e + G, nnz);
      if (e > s && hspos[n] - hspos[s] > (1LL << 28)) break;
      e = n;
    }
    s = e;
  }
  bs.push_back(nnz);
  const int nb = (int)bs.size() - 1;
  long long bmax = 0;
  for (int i = 0; i < nb; ++i) bmax = std::max(bmax, hspos[bs[i + 1]] - hspos[bs[i]]);
  if (bmax >= 0x7fffffffLL) return NUFFT3D_OK;
  int gbits = 1;
  while ((1LL << gbits) <= ngrp) ++gbits;
  unsigned long long *keys = nullptr, *skeys = nullptr;
  int* d_nsel = nullptr;
  void* tmp = nullptr;
  int* lst = nullptr;
  long long *beg = nullptr, *end = nullptr, *grp = nullptr;
  unsigned short *hslot = nullptr, *cnt = nullptr, *hk = nullptr;
  auto cleanup = [&]() {
    for (void* q : {(void*)keys, (void*)skeys, (void*)d_nsel, tmp, (void*)lst, (void*)beg,
                    (void*)end, (void*)grp, (void*)hslot, (void*)cnt, (void*)hk})
      if (q) (void)dev_free(q);
    keys = skeys = nullptr; d_nsel = nullptr; tmp = nullptr; lst = nullptr;
    beg = end = grp = nullptr; hslot = cnt = hk = nullptr;
  };
  auto fail = [&](const char* what) {
    cleanup();
    drop_fwd_lists(p);
    (void)hipGetLastError();
    if (p->verbose) printf("  block-staged forward lists not built: %s\n", what);
    return NUFFT3D_OK;
  };
  const size_t nk = (size_t)std::max<long long>(bmax, 1);
  if (dev_malloc(&keys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&skeys, sizeof(unsigned long long) * nk) != hipSuccess ||
      dev_malloc(&d_nsel, sizeof(int)) != hipSuccess ||
      dev_malloc(&hslot, sizeof(unsigned short) * nh) != hipSuccess ||
      dev_malloc(&beg, sizeof(long long) * ngrp) != hipSuccess ||
      dev_malloc(&end, sizeof(long long) * ngrp) != hipSuccess)
    return fail("allocation failed");
  if (hipMemsetAsync(beg, 0, sizeof(long long) * ngrp, p->stream) != hipSuccess ||
      hipMemsetAsync(end, 0, sizeof(long long) * ngrp, p->stream) != hipSuccess)
    return fail("memset failed");
  size_t tb1 = 0, tb2 = 0;
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, tb1, keys, skeys, (int)nk, 0, 32 + gbits,
                                        p->stream) != hipSuccess ||
      hipcub::DeviceSelect::Unique(nullptr, tb2, skeys, keys, d_nsel, (int)nk, p->stream) != hipSuccess ||
      dev_malloc(&tmp, std::max(tb1, tb2) + 16) != hipSuccess)
    return fail("sort set-up failed");
  auto batch_keys = [&](int i, int& nsel) -> bool {
    const long long s0 = bs[i], s1 = bs[i + 1];
    const int n = (int)(hspos[s1] - hspos[s0]);
    nsel = 0;
    if (n == 0) return true;
    hipLaunchKernelGGL(k_fk_keys, dim3(nblocks(s1 - s0)), dim3(256), 0, p->stream, p->d_spos,
                       p->d_fcode, p->d_base, s0, s1, hspos[s0], K1, K2, K3, W, G, keys);
    return hipcub::DeviceRadixSort::SortKeys(tmp, tb1, keys, skeys, n, 0, 32 + gbits, p->stream) == hipSuccess &&
           hipcub::DeviceSelect::Unique(tmp, tb2, skeys, keys, d_nsel, n, p->stream) == hipSuccess &&
           hipMemcpyAsync(&nsel, d_nsel, sizeof(int), hipMemcpyDeviceToHost, p->stream) == hipSuccess &&
           hipStreamSynchronize(p->stream) == hipSuccess;
  };
  std::vector<long long> nsel_b(nb, 0);
  long long total = 0;
  for (int i = 0; i < nb; ++i) {
    int ns = 0;
    if (!batch_keys(i, ns)) return fail("sort failed");
    nsel_b[i] = ns;
    total += ns;
  }
  if (dev_malloc(&lst, sizeof(int) * std::max(1LL, total)) != hipSuccess) return fail("allocation failed");
  long long off = 0;
  for (int i = 0; i < nb; ++i) {
    int ns = 0;
    if (!batch_keys(i, ns) || ns != nsel_b[i]) return fail("sort failed");
    if (ns > 0)
      hipLaunchKernelGGL(k_wl_lists, dim3(nblocks(ns)), dim3(256), 0, p->stream,
                         (const unsigned long long*)keys, (long long)ns, off, lst, beg, end);
    hipLaunchKernelGGL(k_fk_slot, dim3(nblocks(bs[i + 1] - bs[i])), dim3(256), 0, p->stream,
                       p->d_spos, p->d_fcode, p->d_base, bs[i], bs[i + 1], K1, K2, K3, W, G,
                       (const int*)lst, (const long long*)beg, (const long long*)end, hslot);
    if (hipGetLastError() != hipSuccess) return fail("slot kernels failed");
    off += ns;
  }
  (void)dev_free(keys); keys = nullptr;
  (void)dev_free(skeys); skeys = nullptr;
  (void)dev_free(tmp); tmp = nullptr;
  std::vector<long long> hb(ngrp), he(ngrp), cbase(ngrp + 1, 0);
  if (hipMemcpyAsync(hb.data(), beg, sizeof(long long) * ngrp, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipMemcpyAsync(he.data(), end, sizeof(long long) * ngrp, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess)
    return fail("copy failed");
  for (long long j = 0; j < ngrp; ++j) {
    if (he[j] - hb[j] > 65535) return fail("a group's point list exceeds 16-bit slots");
    cbase[j + 1] = cbase[j] + (he[j] - hb[j] + C - 1) / C;
  }
  const long long nchunk = cbase[ngrp];
  const long long nent = nchunk * NW;   // (chunk, wavefront) entry ranges
  if (dev_malloc(&fk.d_cbase, sizeof(long long) * (ngrp + 1)) != hipSuccess ||
      dev_malloc(&fk.d_gpt, sizeof(int) * std::max(1LL, nchunk * C)) != hipSuccess ||
      dev_malloc(&fk.d_eoff, sizeof(long long) * (nent + 1)) != hipSuccess ||
      dev_malloc(&cnt, sizeof(unsigned short) * std::max(1LL, nchunk * G)) != hipSuccess ||
      dev_malloc(&hk, sizeof(unsigned short) * nh) != hipSuccess ||
      dev_malloc(&grp, sizeof(long long) * (nent + 1)) != hipSuccess)
    return fail("allocation failed");
  if (hipMemcpy(fk.d_cbase, cbase.data(), sizeof(long long) * (ngrp + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(cnt, 0, sizeof(unsigned short) * std::max(1LL, nchunk * G), p->stream) != hipSuccess)
    return fail("copy failed");
  // chunked point lists (the same dealing as the adjoint's sample lists)
  hipLaunchKernelGGL(k_bk_samp, dim3((unsigned)ngrp), dim3(256), 0, p->stream, (const int*)lst,
                     (const long long*)beg, (const long long*)end, (const long long*)fk.d_cbase, C,
                     fk.d_gpt);
  hipLaunchKernelGGL(k_fk_rank, dim3(nblocks(nnz)), dim3(256), 0, p->stream,
                     (const long long*)p->d_spos, nnz, G, (const long long*)fk.d_cbase,
                     (const unsigned short*)hslot, cnt, hk);
  if (nent > 0)
    hipLaunchKernelGGL(k_fk_groups, dim3(nblocks(nent)), dim3(256), 0, p->stream,
                       (const unsigned short*)cnt, nchunk, G, grp);
  if (hipGetLastError() != hipSuccess) return fail("list kernels failed");
  std::vector<long long> hgrp(nent + 1, 0), eoff(nent + 1, 0);
  if (nent > 0 &&
      (hipMemcpyAsync(hgrp.data(), grp, sizeof(long long) * nent, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
       hipStreamSynchronize(p->stream) != hipSuccess))
    return fail("copy failed");
  for (long long c = 0; c < nent; ++c) eoff[c + 1] = eoff[c] + hgrp[c];
  const long long ngroups = eoff[nent];
  if (dev_malloc(&fk.d_slot, sizeof(unsigned) * std::max(1LL, ngroups * 64)) != hipSuccess ||
      (weights && dev_malloc(&fk.d_w, sizeof(float4) * std::max(1LL, ngroups * 64)) != hipSuccess))
    return fail("allocation failed");
  fk.ngroups = ngroups;
  fk.G = G;
  fk.C = C;
  if (hipMemcpy(fk.d_eoff, eoff.data(), sizeof(long long) * (nent + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(fk.d_slot, 0xff, sizeof(unsigned) * std::max(1LL, ngroups * 64), p->stream) != hipSuccess ||
      (weights &&
       hipMemsetAsync(fk.d_w, 0, sizeof(float4) * std::max(1LL, ngroups * 64), p->stream) != hipSuccess))
    return fail("copy failed");
  if (fk_fill(p, hslot, hk) != NUFFT3D_OK || hipStreamSynchronize(p->stream) != hipSuccess)
    return fail("fill failed");
  cleanup();
  fk.ngrp = ngrp;
  fk.nchunk = nchunk;
  fk.ngroups = ngroups;
  if (p->verbose)
    printf("  block-staged forward: %lld groups of %d samples, %lld chunks of %d points, %lld entry "
           "groups (%.2f B/hit)\n", ngrp, G, nchunk, C, ngroups, 4.0 * 64 * ngroups / (double)nh);
  return NUFFT3D_OK;
}

int ensure_fwd_lists(nufft3d_plan* p, bool weights) {
  if (p->fk.d_slot || p->fk.tried || p->prec != 2 || p->K[2] == 1) return NUFFT3D_OK;
  p->fk.tried = true;
  return build_fwd_lists(p, weights);
}

// Pieces of the tiled interpolation: runs of equal sub-bin key in the sorted
// samples, cut into at most kPieceMax samples.
constexpr int kPieceMax = 128;   // one sample per 4-lane group of a 512-thread work-group

static int build_pieces(nufft3d_plan* p, const unsigned* skeys, const BinGeom& g) {
  const long long nnz = p->nnz;
  std::vector<unsigned> k(nnz);
  N3D_HIP(hipMemcpy(k.data(), skeys, sizeof(unsigned) * nnz, hipMemcpyDeviceToHost));
  std::vector<int4> pc;
  pc.reserve(nnz / 32 + 16);
  for (long long s = 0; s < nnz;) {
    long long e = s + 1;
    while (e < nnz && k[e] == k[s]) ++e;
    unsigned tx = 0, ty = 0, tz = 0;
    if (g.mbits > 0) {
      for (int b = 0; b < g.mbits; ++b) {
        tx |= ((k[s] >> (3 * b)) & 1u) << b;
        ty |= ((k[s] >> (3 * b + 1)) & 1u) << b;
        tz |= ((k[s] >> (3 * b + 2)) & 1u) << b;
      }
    } else {
      tx = k[s] % g.nt1;
      ty = (k[s] / g.nt1) % g.nt2;
      tz = k[s] / (g.nt1 * g.nt2);
    }
    for (long long a = s; a < e; a += kPieceMax) {
      const int n = (int)std::min<long long>(kPieceMax, e - a);
      pc.push_back(make_int4((int)(tx | (ty << 10) | (tz << 20)), (int)a, n, 0));
    }
    s = e;
  }
  p->npiece = (long long)pc.size();
  N3D_HIP(dev_malloc(&p->d_piece, sizeof(int4) * (pc.empty() ? 1 : pc.size())));
  if (!pc.empty())
    N3D_HIP(hipMemcpy(p->d_piece, pc.data(), sizeof(int4) * pc.size(), hipMemcpyHostToDevice));
  return NUFFT3D_OK;
}

static int build_plan(nufft3d_plan* p, const double* om_in, int64_t M) {
  const bool dev = is_device_ptr(om_in);
  double* om = nullptr;
  N3D_HIP(dev_malloc(&om, sizeof(double) * 3 * (M > 0 ? M : 1)));
  N3D_HIP(hipMemcpy(om, om_in, sizeof(double) * 3 * M,
                    dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
  // all z == 0 -> 2-D (nufft_3d.m:104-106); decided on the host copy
  {
    std::vector<double> h;
    const double* hz = om_in;
    if (dev) {
      h.resize(3 * M);
      N3D_HIP(hipMemcpy(h.data(), om_in, sizeof(double) * 3 * M, hipMemcpyDeviceToHost));
      hz = h.data();
    }
    bool allz = true;
    for (int64_t m = 0; m < M && allz; ++m) allz = hz[3 * m + 2] == 0.0;
    if (allz) p->N[2] = 1;
    if (p->verbose) {
      double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
      for (int64_t m = 0; m < M; ++m)
        for (int a = 0; a < 3; ++a) { mn[a] = std::min(mn[a], hz[3 * m + a]); mx[a] = std::max(mx[a], hz[3 * m + a]); }
      printf(" Trajectory:     min       max       matrix\n");
      for (int a = 0; a < 3; ++a) printf("  om(%d): %12.3f %9.3f %9d\n", a + 1, mn[a], mx[a], p->N[a]);
    }
  }
  for (int a = 0; a < 3; ++a) p->K[a] = 2 * (int)ceil(p->N[a] * p->u / 2);   // nufft_3d.m:109
  if (p->N[2] == 1) p->K[2] = 1;
  if ((double)p->K[0] * p->K[1] * p->K[2] >= 2147483647.0) {
    set_error("N or u too large");                                          // nufft_3d.m:111
    (void)dev_free(om);
    return NUFFT3D_ERR_INVALID;
  }
  const int Wj = (int)ceil(p->J / 2);                                          // nufft_3d.m:154
  p->W[0] = Wj; p->W[1] = Wj; p->W[2] = p->N[2] == 1 ? 0 : Wj;
  // grids narrower than the footprint 2*ceil(J/2)+1 (N = 2 or 3 on an axis)
  // are fine: wrapped candidates that land on one grid point are separate
  // hits whose sum is the reference's summed duplicate entry
  // (nufft_3d.m:178-180,203,208). The gridding kernels wrap an offset
  // candidate with two corrections of one period, which covers ceil(J/2)
  // up to 2K.
  for (int a = 0; a < 3; ++a)
    if (p->K[a] > 1 && p->W[a] > 2 * p->K[a]) {
      set_error("kernel width J=" + std::to_string(p->J) + " spans more than two periods of the " +
                "oversampled size K=" + std::to_string(p->K[a]));
      (void)dev_free(om);
      return NUFFT3D_ERR_UNSUPPORTED;
    }
  p->M = M;

  // precision policy of fp32 plans: the adjoint ends in the deapodization
  // U, which multiplies the image edge by up to max|U|/min|U| relative to its
  // centre; fp32 rounding anywhere before it (grid, FFT, weights, density, the
  // forward half of A'WA) is amplified by that range. Above 256 (e.g. u=1.25,
  // J=6: 2.5e4..3e5) the arithmetic runs in fp64 behind the fp32 arrays.
  if (p->io_prec == 1 && p->fp64_mode != 0) {
    deap_host(p);
    double umax = 0, umin = 1e300;
    for (double v : p->h_U) { umax = std::max(umax, std::fabs(v)); umin = std::min(umin, std::fabs(v)); }
    const double range = umin > 0 ? umax / umin : 1e300;
    if (p->fp64_mode == 1 || range > 256.0) p->prec = 2;
    if (p->verbose)
      printf("  deapodization range %.3g: %s arithmetic\n", range, p->prec == 2 ? "fp64" : "fp32");
  }

  // tiles: 4x4x4 sample bins where the forward interpolation runs from LDS
  // tiles (3-D, 7-point footprint: the 10^3 neighbourhood of a bin fits in LDS)
  {
    // opt-in (NUFFT3D_TILE_INTERP=1): measured slower than the gather
    // interpolation at 256^3 / J=6 (21.6 against 18.4 ms per 8-coil launch)
    const char* te = getenv("NUFFT3D_TILE_INTERP");
    const bool allow = te && atoi(te) != 0;
    p->tile_interp = allow && !p->tile_user && p->N[2] > 1 && p->W[0] == 3 && p->W[1] == 3 &&
                     p->W[2] == 3 && p->K[0] >= 16 && p->K[1] >= 16 && p->K[2] >= 16;
    if (p->tile_interp) p->T[0] = p->T[1] = p->T[2] = 4;
  }
  // fp64 arithmetic with a wide kernel (J >= 5: ~113 hits per sample of
  // 128-byte fp64 grid rows): samples in the Z-order of their own grid
  // point, not only of their 8^3 bin, so consecutive samples read the same
  // grid rows (256^3, J=6, 8 coils: interpolation 17.7 -> 15.6 ms, spread
  // 23.4 -> 23.0, caller-row permutation 0.60 -> 0.69). With J=4 and fp32
  // rows the caller-layout passes lose more than the gridding gains.
  if (!p->tile_user && !p->tile_interp && p->N[2] > 1 && p->prec == 2 && p->W[0] >= 3)
    p->T[0] = p->T[1] = p->T[2] = 1;
  if (p->N[2] == 1) { p->T[0] = p->T[1] = 16; p->T[2] = 1; }
  for (int a = 0; a < 3; ++a) {
    if (p->T[a] > p->K[a]) p->T[a] = p->K[a];
    p->ntile[a] = (p->K[a] + p->T[a] - 1) / p->T[a];
  }
  p->ntiles = p->ntile[0] * p->ntile[1] * p->ntile[2];
  BinGeom g;
  g.u = p->u;
  g.K1 = p->K[0]; g.K2 = p->K[1]; g.K3 = p->K[2];
  g.T1 = p->T[0]; g.T2 = p->T[1]; g.T3 = p->T[2];
  g.nt1 = p->ntile[0]; g.nt2 = p->ntile[1]; g.nt3 = p->ntile[2];
  g.ntiles = p->ntiles;
  g.W1 = p->W[0]; g.W2 = p->W[1]; g.W3 = p->W[2];
  g.is2d = p->N[2] == 1;
  g.mbits = 0;
  if (!(p->tune_flags & 4)) {   // Morton tile order unless linear order is asked for
    int nt = std::max(p->ntile[0], std::max(p->ntile[1], p->ntile[2]));
    g.mbits = 1;
    while ((1 << g.mbits) < nt) ++g.mbits;
  }
  const long long keymax = g.mbits > 0 ? (1LL << (3 * g.mbits)) : (long long)p->ntiles;

  unsigned *keys = nullptr, *skeys = nullptr;
  int *vals = nullptr, *svals = nullptr, *counts = nullptr;
  const long long Mm = M > 0 ? M : 1;
  N3D_HIP(dev_malloc(&keys, sizeof(unsigned) * Mm));
  N3D_HIP(dev_malloc(&skeys, sizeof(unsigned) * Mm));
  N3D_HIP(dev_malloc(&vals, sizeof(int) * Mm));
  N3D_HIP(dev_malloc(&svals, sizeof(int) * Mm));
  N3D_HIP(dev_malloc(&counts, sizeof(int)));
  N3D_HIP(hipMemsetAsync(counts, 0, sizeof(int), p->stream));
  if (M > 0)
    hipLaunchKernelGGL(k_keys, dim3(nblocks(M)), dim3(256), 0, p->stream, om, (long long)M, g,
                       keys, vals, counts);
  int end_bit = 1;
  while ((1LL << end_bit) <= keymax) ++end_bit;
  size_t tb = 0;
  N3D_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, skeys, vals, svals, (int)M, 0,
                                             end_bit, p->stream));
  void* tmp = nullptr;
  N3D_HIP(dev_malloc(&tmp, tb + 16));
  if (M > 0)
    N3D_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, skeys, vals, svals, (int)M, 0,
                                               end_bit, p->stream));
  int noob = 0;
  N3D_HIP(hipMemcpyAsync(&noob, counts, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(tmp);
  p->nnz = M - noob;
  if (p->verbose)
    printf("  %lld points (out of %lld) are within bounds.\n", (long long)p->nnz, (long long)M);
  const long long nnz = p->nnz;
  N3D_HIP(dev_malloc(&p->d_orig, sizeof(int) * (nnz > 0 ? nnz : 1)));
  N3D_HIP(dev_malloc(&p->d_oob, sizeof(int) * (noob > 0 ? noob : 1)));
  N3D_HIP(dev_malloc(&p->d_base, sizeof(int4) * (nnz > 0 ? nnz : 1)));
  N3D_HIP(dev_malloc(&p->d_off, 3 * sizeof(double) * (nnz > 0 ? nnz : 1)));
  if (M > 0)
    hipLaunchKernelGGL(k_gather, dim3(nblocks(M)), dim3(256), 0, p->stream, om, skeys, svals,
                       (long long)M, nnz, g, p->d_orig, p->d_oob, p->d_base, (double*)p->d_off);
  N3D_HIP(hipGetLastError());
  // inverse of orig: caller row -> sorted position (-1 for out-of-bounds rows)
  N3D_HIP(dev_malloc(&p->d_inv, sizeof(int) * (M > 0 ? M : 1)));
  N3D_HIP(hipMemsetAsync(p->d_inv, 0xff, sizeof(int) * (M > 0 ? M : 1), p->stream));
  if (nnz > 0)
    hipLaunchKernelGGL(k_invert, dim3(nblocks(nnz)), dim3(256), 0, p->stream,
                       (const int*)p->d_orig, nnz, p->d_inv);
  N3D_HIP(hipGetLastError());
  if (p->tile_interp && nnz > 0) N3D_TRY(build_pieces(p, skeys, g));
  (void)dev_free(keys); (void)dev_free(skeys); (void)dev_free(vals); (void)dev_free(svals);
  (void)dev_free(om);

  if (p->prec == 1) {
    N3D_HIP(dev_malloc(&p->d_off32, sizeof(float4) * (nnz > 0 ? nnz : 1)));
    if (nnz > 0)
      hipLaunchKernelGGL(k_off32, dim3(nblocks(nnz)), dim3(256), 0, p->stream,
                         (const double*)p->d_off, nnz, p->d_off32);
  }

  // interpolation pattern: hit lists per sample (forward) and per grid point
  // (adjoint); the adjoint lists are sorted by grid index with a stable sort
  // of the sample-major list, so each point sees its samples in ascending order
  HitGeom hg;
  hg.K1 = p->K[0]; hg.K2 = p->K[1]; hg.K3 = p->K[2];
  hg.W1 = p->W[0]; hg.W2 = p->W[1]; hg.W3 = p->W[2];
  hg.J2 = p->J * p->J;
  hg.radial = p->radial;
  const long long ngrid = (long long)p->K[0] * p->K[1] * p->K[2];
  // 64-bit counts: the scans below run in the input type, and the total hit
  // count exceeds 2^31 for large J=6 plans
  long long* cnt_s = nullptr;
  unsigned long long* cnt_g = nullptr;
  N3D_HIP(dev_malloc(&cnt_s, sizeof(long long) * (nnz + 1)));
  N3D_HIP(dev_malloc(&cnt_g, sizeof(unsigned long long) * (ngrid + 1)));
  N3D_HIP(dev_malloc(&p->d_spos, sizeof(long long) * (nnz + 1)));
  N3D_HIP(dev_malloc(&p->d_rowptr, sizeof(long long) * (ngrid + 1)));
  N3D_HIP(hipMemsetAsync(cnt_s, 0, sizeof(long long) * (nnz + 1), p->stream));
  N3D_HIP(hipMemsetAsync(cnt_g, 0, sizeof(unsigned long long) * (ngrid + 1), p->stream));
  if (nnz > 0)
    hipLaunchKernelGGL(k_hits_count, dim3(nblocks(nnz)), dim3(256), 0, p->stream, p->d_base,
                       (const double*)p->d_off, nnz, hg, cnt_s, cnt_g);
  tb = 0;
  size_t tb2 = 0;
  N3D_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt_s, p->d_spos, (long long)(nnz + 1),
                                           p->stream));
  N3D_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (const long long*)cnt_g, p->d_rowptr,
                                           (long long)(ngrid + 1), p->stream));
  N3D_HIP(dev_malloc(&tmp, std::max(tb, tb2) + 16));
  N3D_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt_s, p->d_spos, (long long)(nnz + 1),
                                           p->stream));
  N3D_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, (const long long*)cnt_g, p->d_rowptr,
                                           (long long)(ngrid + 1), p->stream));
  long long nhits = 0;
  N3D_HIP(hipMemcpyAsync(&nhits, p->d_spos + nnz, sizeof(long long), hipMemcpyDeviceToHost,
                         p->stream));
  std::vector<long long> hcnt(ngrid);
  N3D_HIP(hipMemcpyAsync(hcnt.data(), cnt_g, sizeof(long long) * ngrid, hipMemcpyDeviceToHost,
                         p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(tmp);
  p->nhits = nhits;
  const long long nh = nhits > 0 ? nhits : 1;
  unsigned *hk = nullptr, *hks = nullptr;
  unsigned long long *hv = nullptr, *hvs = nullptr;
  N3D_HIP(dev_malloc(&p->d_fcode, sizeof(unsigned short) * nh));
  N3D_HIP(dev_malloc(&p->d_hit, sizeof(uint2) * nh));
  N3D_HIP(dev_malloc(&hk, sizeof(unsigned) * nh));
  N3D_HIP(dev_malloc(&hks, sizeof(unsigned) * nh));
  N3D_HIP(dev_malloc(&hv, sizeof(unsigned long long) * nh));
  N3D_HIP(dev_malloc(&hvs, sizeof(unsigned long long) * nh));
  if (nnz > 0)
    hipLaunchKernelGGL(k_hits_fill, dim3(nblocks(nnz)), dim3(256), 0, p->stream, p->d_base,
                       (const double*)p->d_off, nnz, hg, (const long long*)p->d_spos, hk, hv,
                       p->d_fcode);
  int key_bits = 1;
  while ((1LL << key_bits) < ngrid) ++key_bits;
  tb = 0;
  N3D_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, hk, hks, hv, hvs, nhits, 0, key_bits,
                                             p->stream));
  N3D_HIP(dev_malloc(&tmp, tb + 16));
  if (nhits > 0) {
    N3D_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, hk, hks, hv, hvs, nhits, 0, key_bits,
                                               p->stream));
    hipLaunchKernelGGL(k_split_vals, dim3(nblocks(nhits)), dim3(256), 0, p->stream,
                       (const unsigned long long*)hvs, nhits, p->d_hit);
  }
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(tmp); (void)dev_free(hk); (void)dev_free(hks); (void)dev_free(hv); (void)dev_free(hvs);
  (void)dev_free(cnt_s); (void)dev_free(cnt_g); (void)dev_free(counts);

  // dense grid points (the k-space centre of radial trajectories) are split
  // into segments of chunk_len hits whose partial sums add in segment order;
  // two segmentations over the same hit lists (see DenseSet)
  {
    int maxc = 0;
    for (long long gi = 0; gi < ngrid; ++gi) maxc = std::max(maxc, (int)hcnt[gi]);
    p->max_hits = maxc;
  }
  for (int set = 0; set < 2; ++set) N3D_TRY(build_dense(p, hcnt, ngrid, p->dense[set]));

  // spread traversal: 8x8x8 bricks of the grid, Z-order unless linear order
  // is asked for, so consecutive work-groups cover compact regions
  {
    const int bx = (p->K[0] + 7) / 8, by = (p->K[1] + 7) / 8, bz = (p->K[2] + 7) / 8;
    int bits = 1;
    while ((1 << bits) < std::max(bx, std::max(by, bz))) ++bits;
    std::vector<std::pair<unsigned, int>> order;
    order.reserve((size_t)bx * by * bz);
    for (int z = 0; z < bz; ++z)
      for (int y = 0; y < by; ++y)
        for (int x = 0; x < bx; ++x) {
          const unsigned key = (p->tune_flags & 4) ? (unsigned)order.size() : morton3(x, y, z, bits);
          order.push_back({key, x | (y << 10) | (z << 20)});
        }
    std::sort(order.begin(), order.end());
    std::vector<int> tab(order.size()), brick_of(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
      tab[i] = order[i].second;
      const int x = tab[i] & 1023, y = (tab[i] >> 10) & 1023, z = tab[i] >> 20;
      brick_of[((size_t)z * by + y) * bx + x] = (int)i;
    }
    p->nbrick = (long long)tab.size();
    N3D_HIP(dev_malloc(&p->d_brick, sizeof(int) * tab.size()));
    N3D_HIP(hipMemcpy(p->d_brick, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    p->h_brick_of.swap(brick_of);
    // the LDS-staged spread's lists are built when that spread is asked for
    // (tune flag 8) or timed by the autotuner
    if (p->tune_flags & 8) N3D_TRY(ensure_brick_lists(p));
  }
  N3D_TRY(build_hit_weights(p));
  if (p->tune_flags & 8) N3D_TRY(ensure_wide_lists(p));
  if (p->tune_flags & 4096) N3D_TRY(ensure_block_lists(p));
  N3D_TRY(build_kb_table(p));   // (the forward lists' stored weights evaluate it)
  if (p->tune_flags & 8192) N3D_TRY(ensure_fwd_lists(p, (p->tune_flags & 16384) != 0));
  // fp64-arithmetic plans read only the weight table in the spread (unless
  // tune flag 128 evaluates the weights) and build no code-based lists, so
  // their (sample, code) copy of the lists goes (8 B per hit)
  if (p->prec == 2 && p->d_hitw && p->d_hit && !(p->tune_flags & 128)) {
    N3D_HIP(hipStreamSynchronize(p->stream));
    N3D_HIP(dev_free(p->d_hit));
    p->d_hit = nullptr;
  }
  N3D_TRY(build_kb_table(p));

  // deapodization (deap.m) and FFT twiddles
  deap_host(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2];
  if (p->prec == 2) {
    N3D_HIP(dev_malloc(&p->d_U, sizeof(double) * nimg));
    N3D_HIP(hipMemcpy(p->d_U, p->h_U.data(), sizeof(double) * nimg, hipMemcpyHostToDevice));
  } else {
    std::vector<float> u32(nimg);
    for (size_t i = 0; i < nimg; ++i) u32[i] = (float)p->h_U[i];
    N3D_HIP(dev_malloc(&p->d_U, sizeof(float) * nimg));
    N3D_HIP(hipMemcpy(p->d_U, u32.data(), sizeof(float) * nimg, hipMemcpyHostToDevice));
  }
  for (int a = 0; a < 3; ++a) N3D_TRY(fft_setup_axis(p, a));
  return NUFFT3D_OK;
}

// Waits for the plan's enqueued work. A stream the plan does not own (the
// caller's, e.g. one shared by several plans of a serving loop) is not
// synchronised: another thread may be capturing a graph on it, and a
// synchronisation would invalidate that capture. The event recorded as the
// plan's last entry point returned marks the end of its work there.
static int wait_plan_work(nufft3d_plan* p) {
  hipError_t e = hipSuccess;
  if (p->own_stream && p->stream) e = hipStreamSynchronize(p->stream);
  else if (p->done_ev) e = hipEventSynchronize(p->done_ev);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return NUFFT3D_ERR_DEVICE;
  }
  return NUFFT3D_OK;
}

static void destroy_plan(nufft3d_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  (void)wait_plan_work(p);
  if (p->done_ev) (void)hipEventDestroy(p->done_ev);
  void* ptrs[] = {p->d_orig, p->d_oob, p->d_inv, p->d_base, p->d_off, p->d_U, p->d_dens,
                  p->d_spos, p->d_fcode, p->d_rowptr, p->d_hit, p->d_hitw,
                  p->d_off32, p->d_piece, p->d_kbtab, p->dense[0].d_hseg, p->dense[0].d_hpt, p->dense[1].d_hseg,
                  p->dense[1].d_hpt, p->d_brick, p->bl.d_ptr, p->bl.d_samp, p->bl.d_ent,
                  p->wl.d_beg, p->wl.d_end, p->wl.d_samp, p->wl.d_went};
  for (void* q : ptrs)
    if (q) (void)dev_free(q);
  drop_block_lists(p);
  drop_fwd_lists(p);
  for (int a = 0; a < 3; ++a)
    for (FftAxis* x : {&p->ax[a], &p->axN[a]})
      for (void* q : {x->d_tw, x->d_chirp, x->d_bhat[0], x->d_bhat[1]})
        if (q) (void)dev_free(q);
  if (p->mr_exec) (void)hipGraphExecDestroy(p->mr_exec);
  for (auto& r : p->prof.recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto& e : p->prof.pool) (void)hipEventDestroy(e);
  release(p->ws_grid); release(p->ws_t1); release(p->ws_t2); release(p->ws_samp);
  release(p->ws_part); release(p->ws_img); release(p->ws_solve); release(p->ws_red);
  release(p->ws_b);
  release(p->ws_host, true);
  if (p->own_stream && p->stream) (void)dev_stream_destroy(p->stream);
  delete p;
}

}  // namespace nufft3d

using namespace nufft3d;

// ================================================================ C ABI
// elementwise precision conversion of caller arrays (fp32 arrays of plans
// that compute in fp64)
template <typename A, typename B>
__global__ void k_convert(const A* __restrict__ in, B* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

// deterministic, non-constant test data for timing runs
template <typename R>
__global__ void k_fill_pattern(R* __restrict__ a, long long n, unsigned seed) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 2654435761u ^ seed;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  a[i] = (R)((double)(h & 0xffffff) / 16777216.0 - 0.5);
}

// ---------------------------------------------------------------- discrep
// Sort keys of the stddev guess of discrep.m:30-32: the real and imaginary
// parts of the nonzero entries; both parts of a zero entry become +inf, which
// sorts after every finite key. cnt counts the nonzero entries.
template <typename R>
__global__ void k_mad_keys(const cplx<R>* __restrict__ d, long long n, R* __restrict__ keys,
                           unsigned long long* __restrict__ cnt) {
  __shared__ unsigned long long sh[256];
  unsigned long long c = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const cplx<R> v = d[i];
    const bool nz = v.x != (R)0 || v.y != (R)0;
    keys[2 * i] = nz ? v.x : (R)INFINITY;
    keys[2 * i + 1] = nz ? v.y : (R)INFINITY;
    c += nz;
  }
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && sh[0]) atomicAdd(cnt, sh[0]);
}

// keys <- |keys - med| for the first m (finite) sorted keys, +inf after them
template <typename R>
__global__ void k_mad_dev(const R* __restrict__ sorted, long long n, long long m, R med,
                          R* __restrict__ keys) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    keys[i] = i < m ? (R)fabs(sorted[i] - med) : (R)INFINITY;
}

// median of the first m entries of a sorted device array (MATLAB median:
// mean of the two middle entries when m is even)
template <typename R>
static int sorted_median(nufft3d_plan* p, const R* sorted, long long m, double* med) {
  R v[2];
  const long long lo = (m - 1) / 2;
  N3D_HIP(hipMemcpyAsync(v, sorted + lo, 2 * sizeof(R), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  *med = (m % 2) ? (double)v[0] : ((double)v[0] + (double)v[1]) / 2.0;
  return NUFFT3D_OK;
}

// stddev = median(|t - median(t)|), t = [real(nonzeros(data)); imag(...)]
// (discrep.m:30-32), with two device radix sorts of the keys.
template <typename R>
static int mad_stddev(nufft3d_plan* p, const void* data, long long n, double* out) {
  Workspace a, b, tmp;
  const long long nk = 2 * n;
  int st = ensure(a, (size_t)nk * sizeof(R) + 64);
  if (!st) st = ensure(b, (size_t)nk * sizeof(R));
  R* keys = (R*)a.p;
  unsigned long long* cnt = (unsigned long long*)((char*)a.p + (size_t)nk * sizeof(R));
  R* sorted = (R*)b.p;
  size_t tb = 0;
  if (!st && nk > 0x7fffffffLL) {
    set_error("discrep: stddev estimate over more than 2^31-1 values is not supported");
    st = NUFFT3D_ERR_UNSUPPORTED;
  }
  if (!st && hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys, sorted, (int)nk, 0,
                                               (int)(8 * sizeof(R)), p->stream) != hipSuccess) {
    set_error("discrep: radix sort set-up failed");
    st = NUFFT3D_ERR_DEVICE;
  }
  if (!st) st = ensure(tmp, tb);
  unsigned long long m2 = 0;
  if (!st) {
    const unsigned nbk = std::min<unsigned>(nblocks(n), 4096);
    N3D_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), p->stream));
    hipLaunchKernelGGL((k_mad_keys<R>), dim3(nbk), dim3(256), 0, p->stream,
                       (const cplx<R>*)data, n, keys, cnt);
    N3D_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys, sorted, (int)nk, 0,
                                              (int)(8 * sizeof(R)), p->stream));
    N3D_HIP(hipMemcpyAsync(&m2, cnt, sizeof(m2), hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
    if (m2 == 0) {
      set_error("discrep: stddev not supplied and data has no nonzero entries");
      st = NUFFT3D_ERR_INVALID;
    }
  }
  double med = 0, mad = 0;
  const long long m = 2 * (long long)m2;
  if (!st) st = sorted_median<R>(p, sorted, m, &med);
  if (!st) {
    const unsigned nbk = std::min<unsigned>(nblocks(nk), 4096);
    hipLaunchKernelGGL((k_mad_dev<R>), dim3(nbk), dim3(256), 0, p->stream, (const R*)sorted, nk, m,
                       (R)med, keys);
    N3D_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys, sorted, (int)nk, 0,
                                              (int)(8 * sizeof(R)), p->stream));
    st = sorted_median<R>(p, sorted, m, &mad);
  }
  release(a);
  release(b);
  release(tmp);
  *out = mad;
  return st;
}

extern "C" {

const char* nufft3d_last_error(void) { return g_last_error.c_str(); }

int nufft3d_plan_warnings(nufft3d_plan* p, char* buf, int64_t len) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  if (buf && len <= 0) { set_error("invalid buffer length"); return NUFFT3D_ERR_INVALID; }
  std::string all;
  for (size_t i = 0; i < p->warnings.size(); ++i) all += (i ? "\n" : "") + p->warnings[i];
  if (buf) {
    const size_t n = std::min<size_t>(all.size(), (size_t)len - 1);
    memcpy(buf, all.data(), n);
    buf[n] = 0;
  }
  return (int)p->warnings.size();
}

int nufft3d_opts_default(nufft3d_opts* o) {
  if (!o) { set_error("opts is NULL"); return NUFFT3D_ERR_INVALID; }
  std::memset(o, 0, sizeof(*o));
  o->J = 4;
  o->u = 2;
  o->alpha = 0;
  o->radial = 1;
  o->precision = 1;
  o->compensated = 1;
  o->fp64_compute = -1;
  return NUFFT3D_OK;
}

int nufft3d_plan_create(const double* om, int64_t M, const int32_t N[3],
                        const nufft3d_opts* opts, nufft3d_plan** out) {
  g_last_error.clear();
  if (!out) { set_error("plan output pointer is NULL"); return NUFFT3D_ERR_INVALID; }
  *out = nullptr;
  nufft3d_opts o;
  if (opts) o = *opts; else nufft3d_opts_default(&o);
  if (!om || M <= 0) {
    set_error("om must be a real array with leading dimension of 3");   // nufft_3d.m:61
    return NUFFT3D_ERR_INVALID;
  }
  if (M >= 2147483647LL) { set_error("more than 2^31-1 samples"); return NUFFT3D_ERR_UNSUPPORTED; }
  if (!N) { set_error("N must be scalar or 3-vector of even integers"); return NUFFT3D_ERR_INVALID; }
  if (o.u < 1 || o.u > 6) {                                               // nufft_3d.m:82-84
    char b[96];
    snprintf(b, sizeof b, "value of u=%f is not recommended", o.u);
    set_error(b);
    return NUFFT3D_ERR_INVALID;
  }
  if (!(o.J > 0)) { set_error("J must be positive"); return NUFFT3D_ERR_INVALID; }
  if ((N[0] == 0 && N[1] == 0 && N[2] == 0) || N[0] % 2 || N[1] % 2 || N[0] <= 0 || N[1] <= 0 ||
      N[2] <= 0) {
    set_error("N must be an even integer");                                // nufft_3d.m:101-103
    return NUFFT3D_ERR_INVALID;
  }
  if (o.precision != 1 && o.precision != 2) {
    set_error("precision must be 1 (fp32) or 2 (fp64)");
    return NUFFT3D_ERR_INVALID;
  }
  if (hipSetDevice(o.device) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipSetDevice failed (no HIP device?)");
    return NUFFT3D_ERR_DEVICE;
  }
  nufft3d_plan* p = new nufft3d_plan();
  p->prec = o.precision;
  p->io_prec = o.precision;
  p->fp64_mode = o.fp64_compute < 0 ? -1 : (o.fp64_compute ? 1 : 0);
  p->J = o.J;
  p->u = o.u;
  p->radial = o.radial ? 1 : 0;
  p->alpha = o.alpha;
  if (p->alpha == 0) {                                                     // nufft_3d.m:85-92
    p->alpha = M_PI * sqrt(p->J * p->J * (1 - 0.5 / p->u) * (1 - 0.5 / p->u) - 0.8);
    if (p->radial)
      p->alpha = M_PI * sqrt(p->J * p->J * (1 - 0.5 / p->u) * (1 - 0.5 / p->u) - 0.6 * p->u);
  }
  if (p->J < 1 || p->J > 7) {                                              // nufft_3d.m:79-81
    char b[96];
    snprintf(b, sizeof b, "value of J=%f is not recommended", p->J);
    p->warnings.push_back(b);
  }
  if (p->radial && (p->J < 3 || p->u < 1.5))                               // nufft_3d.m:93-95
    p->warnings.push_back("radial kernel not recommended with J<3 or u<1.5");
  if (o.verbose)
    for (const std::string& w : p->warnings) printf("Warning: %s\n", w.c_str());
  for (int a = 0; a < 3; ++a) p->N[a] = N[a];
  p->device = o.device;
  p->verbose = o.verbose;
  p->compensated = o.compensated ? 1 : 0;
  if (const char* tv = getenv("NUFFT3D_TUNE")) p->tune_flags = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_FFTB")) p->fft_bundle = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_FFTZT")) p->fft_zthreads = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_FFTXT")) p->fft_xthreads = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_FFTYT")) p->fft_ythreads = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_XCD")) p->xcd_chunk = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_GRAPH")) p->use_graphs = atoi(tv) != 0;
  if (const char* tv = getenv("NUFFT3D_TEST_CAPTURE_FAIL")) p->capture_fail_at = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_TILE_CG")) p->tile_cg = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_TILE_LAYERS")) p->tile_layers = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_BLK_CHUNK")) p->blk_chunk = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_BLK_WIDE")) p->blk_wide = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_BLK_PIECE")) p->blk_piece = atoi(tv);
  if (const char* tv = getenv("NUFFT3D_FWD_GROUP")) p->fwd_group = atoi(tv);
  if (o.tile > 0) {
    p->T[0] = p->T[1] = p->T[2] = o.tile;
    p->tile_user = true;
  }
  // dense-point segments: 64 hits for launches of <= 4 coils, 512 (or the
  // caller's chunk) for larger ones
  p->dense[0].chunk_len = p->dense[0].rmax = o.chunk > 0 ? std::min(64, (int)o.chunk) : 64;
  p->dense[1].chunk_len = p->dense[1].rmax = o.chunk > 0 ? (int)o.chunk : 512;
  if (p->T[0] * p->T[1] * p->T[2] > 512) {
    set_error("tile too large (at most 512 grid points per tile)");
    delete p;
    return NUFFT3D_ERR_INVALID;
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    set_error("hipStreamCreate failed");
    delete p;
    return NUFFT3D_ERR_DEVICE;
  }
  p->own_stream = true;
  if (hipEventCreateWithFlags(&p->done_ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    p->done_ev = nullptr;
    destroy_plan(p);
    set_error("hipEventCreate failed");
    return NUFFT3D_ERR_DEVICE;
  }
  auto t0 = std::chrono::steady_clock::now();
  int st = build_plan(p, om, M);
  if (st == NUFFT3D_OK) {
    if (o.d && o.nd > 0) {
      if (o.nd != M) {
        set_error("supplied density has wrong number of elements (should be " +
                  std::to_string(M) + ")");                             // nufft_3d.m:230-232
        st = NUFFT3D_ERR_INVALID;
      } else {
        st = p->prec == 2 ? user_density_t<double>(p, o.d, o.nd) : user_density_t<float>(p, o.d, o.nd);
      }
    } else {
      st = p->prec == 2 ? density_t<double>(p) : density_t<float>(p);
    }
  }
  if (st == NUFFT3D_OK && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error during plan construction");
    st = NUFFT3D_ERR_DEVICE;
  }
  if (st != NUFFT3D_OK) {
    destroy_plan(p);
    return st;
  }
  if (p->verbose) {
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf(" Created plan in %.3f s: N=[%d %d %d] K=[%d %d %d] J=%g u=%g alpha=%.4f radial=%d\n",
           sec, p->N[0], p->N[1], p->N[2], p->K[0], p->K[1], p->K[2], p->J, p->u, p->alpha,
           p->radial);
    printf("\tbins %d (%dx%dx%d) hits %lld (%.1f/sample) max hits/point %d dense points %d "
           "segments %d\n", p->ntiles, p->T[0], p->T[1], p->T[2], (long long)p->nhits,
           p->nnz ? (double)p->nhits / p->nnz : 0.0, p->max_hits, p->dense[1].nhpt,
           p->dense[1].nhseg);
    if (p->bl.d_ent)
      printf("\tspread bricks %lld: %lld staged samples (%.2f/sample), longest list %d\n",
             p->nbrick, p->bl.total, p->nnz ? (double)p->bl.total / p->nnz : 0.0, p->bl.max_len);
    if (p->wl.d_went)
      printf("\tsplit-adjoint bricks %lld: %lld staged samples (%.2f/sample), longest list %lld\n",
             p->nbrick, p->wl.total, p->nnz ? (double)p->wl.total / p->nnz : 0.0, p->wl.max_len);
  }
  *out = p;
  return NUFFT3D_OK;
}

int nufft3d_plan_destroy(nufft3d_plan* p) {
  destroy_plan(p);
  return NUFFT3D_OK;
}

int64_t nufft3d_nnz(const nufft3d_plan* p) { return p ? p->nnz : -1; }

int nufft3d_device_bytes(const nufft3d_plan* p, int64_t* plan_bytes, int64_t* workspace_bytes) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  const int64_t nnz = std::max<int64_t>(p->nnz, 1), noob = std::max<int64_t>(p->M - p->nnz, 1);
  const int64_t ngrid = (int64_t)p->K[0] * p->K[1] * p->K[2];
  const int64_t nimg = (int64_t)p->N[0] * p->N[1] * p->N[2];
  const int64_t rsz = p->prec == 2 ? 8 : 4;
  int64_t b = 4 * nnz + 4 * noob + 16 * nnz + 24 * nnz;              // orig, oob, base, off
  b += 4 * std::max<int64_t>(p->M, 1);                                // inv
  b += rsz * nimg + rsz * nnz;                                        // U, density
  b += 8 * (nnz + 1) + 8 * (ngrid + 1);                               // spos, rowptr
  if (p->d_fcode) b += 2 * std::max<int64_t>(p->nhits, 1);           // fcode
  if (p->d_hit) b += 8 * std::max<int64_t>(p->nhits, 1);              // (sample, code) hits
  if (p->d_hitw) b += 8 * p->nhits;                                   // (sample, weight) hits
  if (p->d_off32) b += 16 * nnz;
  if (p->d_piece) b += 16 * p->npiece;
  for (const DenseSet& ds : p->dense)
    b += (int64_t)sizeof(HeavySeg) * std::max(1, ds.nhseg) + (int64_t)sizeof(HeavyPt) * std::max(1, ds.nhpt);
  b += 4 * p->nbrick;
  if (p->bl.d_ptr) b += 8 * (p->nbrick + 1) + 4 * std::max<int64_t>(1, p->bl.total);
  if (p->bl.d_ent) b += 4 * std::max<int64_t>(p->nhits, 1);
  if (p->wl.d_went) b += 16 * p->nbrick + 4 * std::max<int64_t>(1, p->wl.total) + 8 * p->nhits;
  if (p->bk.d_piece)
    b += 12 * p->bk.nblk + 8 + (4 * p->bk.C + 8 * p->bk.T / 64) * p->bk.nchunk + 8 +
         20 * 64 * p->bk.ngroups +
         16 * (p->bk.npiece + std::max<int64_t>(1, p->bk.nred));
  if (p->fk.d_slot)
    b += 8 * (p->fk.ngrp + 1) + (4 * p->fk.C + 8 * p->fk.G / 64) * p->fk.nchunk + 8 +
         (p->fk.d_w ? 20 : 4) * 64 * p->fk.ngroups;
  for (int a = 0; a < 3; ++a)
    for (const FftAxis* x : {&p->ax[a], &p->axN[a]})
      if (x->d_tw) b += 2 * rsz * (x->bs ? x->L + 3 * x->L2 : x->L);   // roots (+ chirp, spectra)
  const Workspace* ws[] = {&p->ws_grid, &p->ws_t1, &p->ws_t2, &p->ws_samp, &p->ws_part,
                           &p->ws_img, &p->ws_solve, &p->ws_red, &p->ws_b};
  int64_t w = 0;
  for (const Workspace* x : ws) w += (int64_t)x->bytes;
  if (plan_bytes) *plan_bytes = b;
  if (workspace_bytes) *workspace_bytes = w;
  return NUFFT3D_OK;
}

int nufft3d_release_workspaces(nufft3d_plan* p) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  (void)hipSetDevice(p->device);
  if (wait_plan_work(p) != NUFFT3D_OK) {
    set_error("device error while releasing workspaces");
    return NUFFT3D_ERR_DEVICE;
  }
  Workspace* ws[] = {&p->ws_grid, &p->ws_t1, &p->ws_t2, &p->ws_samp, &p->ws_part,
                     &p->ws_img, &p->ws_solve, &p->ws_red, &p->ws_b};
  for (Workspace* x : ws) release(*x);
  return NUFFT3D_OK;
}

int nufft3d_get_params(const nufft3d_plan* p, int32_t K[3], int32_t N[3], double* J, double* u,
                       double* alpha, int32_t* radial, int32_t* precision) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  for (int a = 0; a < 3; ++a) {
    if (K) K[a] = p->K[a];
    if (N) N[a] = p->N[a];
  }
  if (J) *J = p->J;
  if (u) *u = p->u;
  if (alpha) *alpha = p->alpha;
  if (radial) *radial = p->radial;
  if (precision) *precision = p->io_prec;
  return NUFFT3D_OK;
}

int32_t nufft3d_compute_precision(const nufft3d_plan* p) { return p ? p->prec : -1; }

int nufft3d_set_stream(nufft3d_plan* p, void* stream) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  hipStream_t ns = (hipStream_t)stream;
  if (ns == p->stream) return NUFFT3D_OK;
  if (hipSetDevice(p->device) != hipSuccess) { (void)hipGetLastError(); set_error("hipSetDevice failed"); return NUFFT3D_ERR_DEVICE; }
  if (p->own_stream && p->stream) {
    (void)hipStreamSynchronize(p->stream);
    (void)dev_stream_destroy(p->stream);
  } else if (p->done_ev) {
    // the plan's workspaces are shared by every stream it runs on: work on
    // the new stream starts after the last entry point on the old one ended
    N3D_HIP(hipStreamWaitEvent(ns, p->done_ev, 0));
  }
  p->own_stream = false;
  p->stream = ns;
  return NUFFT3D_OK;
}

int nufft3d_graph_stats(const nufft3d_plan* p, int64_t* captured, int64_t* failed) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  if (captured) *captured = p->graphs_captured;
  if (failed) *failed = p->graphs_failed;
  return NUFFT3D_OK;
}

int nufft3d_set_graphs(nufft3d_plan* p, int32_t enable) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  p->use_graphs = enable != 0;
  if (!p->use_graphs && p->mr_exec) {
    (void)hipGraphExecDestroy(p->mr_exec);
    p->mr_exec = nullptr;
  }
  return NUFFT3D_OK;
}

int nufft3d_profile_enable(nufft3d_plan* p, int32_t enable) {
  if (!p) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  p->prof.on = enable != 0;
  return NUFFT3D_OK;
}

int nufft3d_profile_read(nufft3d_plan* p, char* buf, int64_t len) {
  if (!p || !buf || len <= 0) { set_error("invalid argument"); return NUFFT3D_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(p->mu);
  N3D_HIP(hipStreamSynchronize(p->stream));
  std::vector<std::string> names;
  std::vector<long long> cnt;
  std::vector<double> tot;
  for (auto& r : p->prof.recs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) { (void)hipGetLastError(); ms = 0; }
    size_t i = 0;
    while (i < names.size() && names[i] != r.name) ++i;
    if (i == names.size()) { names.push_back(r.name); cnt.push_back(0); tot.push_back(0); }
    cnt[i] += 1;
    tot[i] += ms;
    p->prof.pool.push_back(r.a);
    p->prof.pool.push_back(r.b);
  }
  p->prof.recs.clear();
  std::string js = "{";
  for (size_t i = 0; i < names.size(); ++i) {
    char b[256];
    snprintf(b, sizeof b, "%s\"%s\": [%lld, %.6f]", i ? ", " : "", names[i].c_str(), cnt[i], tot[i]);
    js += b;
  }
  js += "}";
  if ((int64_t)js.size() + 1 > len) { set_error("profile buffer too small"); return NUFFT3D_ERR_INVALID; }
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return NUFFT3D_OK;
}

int nufft3d_deap(nufft3d_plan* p, double* U) {
  if (!p || !U) { set_error("NULL argument"); return NUFFT3D_ERR_INVALID; }
  std::memcpy(U, p->h_U.data(), sizeof(double) * p->h_U.size());
  return NUFFT3D_OK;
}

// Stage a caller array on the device when it is host memory.
struct Staged {
  void* dev = nullptr;
  bool host = false;
};

static int stage_in(nufft3d_plan* p, const void* src, size_t bytes, Workspace& ws, Staged& s) {
  if (is_device_ptr(src)) { s.dev = const_cast<void*>(src); s.host = false; return NUFFT3D_OK; }
  N3D_TRY(ensure(ws, bytes));
  N3D_HIP(hipMemcpyAsync(ws.p, src, bytes, hipMemcpyHostToDevice, p->stream));
  s.dev = ws.p;
  s.host = true;
  return NUFFT3D_OK;
}

static int stage_out(nufft3d_plan* p, void* dst, size_t bytes, Workspace& ws, Staged& s) {
  if (is_device_ptr(dst)) { s.dev = dst; s.host = false; return NUFFT3D_OK; }
  N3D_TRY(ensure(ws, bytes));
  s.dev = ws.p;
  s.host = true;
  return NUFFT3D_OK;
}

static int finish_out(nufft3d_plan* p, void* dst, size_t bytes, const Staged& s) {
  if (s.host) {
    N3D_HIP(hipMemcpyAsync(dst, s.dev, bytes, hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
  }
  return NUFFT3D_OK;
}

// ---- caller arrays in the array precision (io_prec), device arrays of the
// arithmetic precision (prec): fp32 arrays are widened on the way in and
// rounded on the way out when an fp32 plan computes in fp64.
static size_t cp_esz(const nufft3d_plan* p) { return p->prec == 2 ? sizeof(double2) : sizeof(float2); }

// n complex values (nreal = 2n) or n reals (cplx = false)
static int stage_in_c(nufft3d_plan* p, const void* src, size_t n, Workspace& ws, Workspace& wtmp,
                      Staged& s, bool cplx = true) {
  const size_t k = cplx ? 2 : 1;
  if (p->io_prec == p->prec) return stage_in(p, src, n * k * (p->prec == 2 ? 8 : 4), ws, s);
  const void* d32 = src;
  s.host = !is_device_ptr(src);
  if (s.host) {
    N3D_TRY(ensure(wtmp, n * k * sizeof(float)));
    N3D_HIP(hipMemcpyAsync(wtmp.p, src, n * k * sizeof(float), hipMemcpyHostToDevice, p->stream));
    d32 = wtmp.p;
  }
  N3D_TRY(ensure(ws, n * k * sizeof(double)));
  if (n > 0)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks((long long)(n * k))), dim3(256), 0,
                       p->stream, (const float*)d32, (double*)ws.p, (long long)(n * k));
  N3D_HIP(hipGetLastError());
  s.dev = ws.p;
  return NUFFT3D_OK;
}

static int stage_out_c(nufft3d_plan* p, void* dst, size_t n, Workspace& ws, Staged& s,
                       bool cplx = true) {
  const size_t k = cplx ? 2 : 1;
  if (p->io_prec == p->prec) return stage_out(p, dst, n * k * (p->prec == 2 ? 8 : 4), ws, s);
  N3D_TRY(ensure(ws, n * k * sizeof(double)));
  s.dev = ws.p;
  s.host = !is_device_ptr(dst);
  return NUFFT3D_OK;
}

static int finish_out_c(nufft3d_plan* p, void* dst, size_t n, const Staged& s, Workspace& wtmp,
                        bool cplx = true) {
  const size_t k = cplx ? 2 : 1;
  if (p->io_prec == p->prec) return finish_out(p, dst, n * k * (p->prec == 2 ? 8 : 4), s);
  void* d32 = dst;
  if (s.host) {
    N3D_TRY(ensure(wtmp, n * k * sizeof(float)));
    d32 = wtmp.p;
  }
  if (n > 0)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks((long long)(n * k))), dim3(256), 0,
                       p->stream, (const double*)s.dev, (float*)d32, (long long)(n * k));
  N3D_HIP(hipGetLastError());
  if (s.host) {
    N3D_HIP(hipMemcpyAsync(dst, d32, n * k * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    N3D_HIP(hipStreamSynchronize(p->stream));
  }
  return NUFFT3D_OK;
}

// Records the plan's completion event on its stream when an entry point
// returns (declared after the lock, so it runs while the lock is held).
struct DoneMark {
  nufft3d_plan* p;
  ~DoneMark() {
    if (p->done_ev && p->stream && hipEventRecord(p->done_ev, p->stream) != hipSuccess)
      (void)hipGetLastError();
  }
};

#define N3D_PLAN_GUARD(p)                                          \
  if (!(p)) { set_error("plan is NULL"); return NUFFT3D_ERR_INVALID; } \
  std::lock_guard<std::mutex> lk_((p)->mu);                        \
  g_last_error.clear();                                            \
  if (hipSetDevice((p)->device) != hipSuccess) { (void)hipGetLastError(); set_error("hipSetDevice failed"); return NUFFT3D_ERR_DEVICE; } \
  DoneMark dm_{(p)};

// Split-precision aNUFT: the caller's complex64 samples, float weights,
// double sums into a double grid, fp64 inverse transform, rounded once.
static int adjoint_split(nufft3d_plan* p, const void* k, int nc, void* x) {
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  Workspace tmp_out;
  Staged ki, xo;
  N3D_TRY(stage_in(p, k, nk * sizeof(float2), p->ws_img, ki));
  int st = stage_out(p, x, nimg * sizeof(float2), tmp_out, xo);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * sizeof(float2));
  if (!st) st = ensure(p->ws_grid, ngrid * sizeof(double2));
  if (!st) st = permute_in(p, ki.dev, nc, p->ws_samp.p, nullptr, 1);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr, 1);              // aNUFT.m:5
  // the fp64 inverse transform writes the float image (rounded once)
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, xo.dev, 1.0 / sqrt((double)p->nnz), nullptr, 0.0, 1);
  if (!st) st = finish_out(p, x, nimg * sizeof(float2), xo);
  release(tmp_out);
  if (!st && ki.host) N3D_HIP(hipStreamSynchronize(p->stream));
  return st;
}

// fNUFT of an fp32-array plan computing in fp64: the x pass reads and widens
// the float image, the interpolation writes float samples (each double
// result rounded once), with no conversion passes around them
static int forward_split(nufft3d_plan* p, const void* x, int nc, void* k) {
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  Workspace tmp_out;
  Staged xi, ko;
  N3D_TRY(stage_in(p, x, nimg * sizeof(float2), p->ws_img, xi));
  int st = stage_out(p, k, nk * sizeof(float2), tmp_out, ko);
  if (!st) st = ensure(p->ws_grid, ngrid * sizeof(double2));
  if (!st) st = fft_forward(p, xi.dev, nc, p->ws_grid.p, nullptr, nullptr, 1);          // fNUFT.m:5-6
  if (!st) st = interp(p, p->ws_grid.p, nc, ko.dev, 1, nullptr, 1.0 / sqrt((double)p->nnz), true, 1);
  if (!st) st = zero_oob(p, ko.dev, nc, 1);
  if (!st) st = finish_out(p, k, nk * sizeof(float2), ko);
  release(tmp_out);
  if (!st && xi.host) N3D_HIP(hipStreamSynchronize(p->stream));
  return st;
}

int nufft3d_forward(nufft3d_plan* p, const void* x, int32_t nc, void* k) {
  N3D_PLAN_GUARD(p);
  if (!x || !k || nc <= 0) { set_error("invalid x/k/nc"); return NUFFT3D_ERR_INVALID; }
  if (p->prec == 2 && p->io_prec == 1 && !(p->tune_flags & (16 | 512))) return forward_split(p, x, nc, k);
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  Workspace tmp_out, tmp_io;
  Staged xi, ko;
  N3D_TRY(stage_in_c(p, x, nimg, p->ws_img, tmp_io, xi));
  int st = stage_out_c(p, k, nk, tmp_out, ko);
  if (st) return st;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  N3D_TRY(ensure(p->ws_grid, ngrid * esz));
  N3D_TRY(fft_forward(p, xi.dev, nc, p->ws_grid.p));                                // fNUFT.m:5-6
  N3D_TRY(interp(p, p->ws_grid.p, nc, ko.dev, 1, nullptr, 1.0 / sqrt((double)p->nnz)));  // :8-9
  N3D_TRY(zero_oob(p, ko.dev, nc));
  st = finish_out_c(p, k, nk, ko, tmp_io);
  release(tmp_out);
  release(tmp_io);
  if (xi.host) N3D_HIP(hipStreamSynchronize(p->stream));
  return st;
}

int nufft3d_adjoint(nufft3d_plan* p, const void* k, int32_t nc, void* x) {
  N3D_PLAN_GUARD(p);
  if (!x || !k || nc <= 0) { set_error("invalid x/k/nc"); return NUFFT3D_ERR_INVALID; }
  if (split_precision(p)) return adjoint_split(p, k, nc, x);
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  Workspace tmp_out, tmp_io;
  Staged ki, xo;
  N3D_TRY(stage_in_c(p, k, nk, p->ws_img, tmp_io, ki));
  int st = stage_out_c(p, x, nimg, tmp_out, xo);
  if (st) return st;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  N3D_TRY(ensure(p->ws_grid, ngrid * esz));
  N3D_TRY(spread(p, ki.dev, nc, p->ws_grid.p, 1, nullptr));                          // aNUFT.m:5
  N3D_TRY(fft_adjoint(p, p->ws_grid.p, nc, xo.dev, 1.0 / sqrt((double)p->nnz), nullptr, 0.0));
  st = finish_out_c(p, x, nimg, xo, tmp_io);
  release(tmp_out);
  release(tmp_io);
  if (ki.host) N3D_HIP(hipStreamSynchronize(p->stream));
  return st;
}

// ---- fNUFT that allocates nothing (no counterpart in the reference).
// Its intermediates (the padded and transformed image, the oversampled grid
// and, for fp64 arithmetic behind fp32 arrays, the widened image and the
// unrounded result) live in one file-scope device array of fixed size; coil
// groups are sized to fit it. Each GPU the library runs on holds its own copy
// until the process ends. Calls on one GPU share it: calls that overlap in
// time on one GPU race.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

// The device address of this device's scratch array, looked up once per
// device with that device current (the array's host-side name is not an
// address the device can use).
static int scratch_on_device(int dev, void** out) {
  if (dev < 0 || dev >= kScratchDevices) {
    set_error("device ordinal out of range for the no-allocation scratch");
    return NUFFT3D_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    N3D_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) { set_error("hipGetSymbolAddress returned NULL for the scratch array"); return NUFFT3D_ERR_DEVICE; }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return NUFFT3D_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// scratch layout of one group of g coils: [widened image][T1][T2][grid][unrounded k]
struct NoallocLayout {
  size_t img, t1, t2, grid, kout, total;
};
static NoallocLayout noalloc_layout(const nufft3d_plan* p, int g) {
  const size_t esz = cp_esz(p);
  const bool conv = p->io_prec != p->prec;
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2];
  const size_t n1 = (size_t)p->K[0] * p->N[1] * p->N[2];
  const size_t n2 = p->K[2] > 1 ? (size_t)p->K[0] * p->K[1] * p->N[2] : 0;
  const size_t ng = (size_t)p->K[0] * p->K[1] * p->K[2];
  NoallocLayout L;
  L.img = 0;
  L.t1 = L.img + (conv ? align256(nimg * g * esz) : 0);
  L.t2 = L.t1 + align256(n1 * g * esz);
  L.grid = L.t2 + align256(n2 * g * esz);
  L.kout = L.grid + align256(ng * g * esz);
  L.total = L.kout + (conv ? align256((size_t)p->M * g * esz) : 0);
  return L;
}

int nufft3d_forward_noalloc(nufft3d_plan* p, const void* x, int32_t nc, void* k) {
  N3D_PLAN_GUARD(p);
  if (!x || !k || nc <= 0) { set_error("invalid x/k/nc"); return NUFFT3D_ERR_INVALID; }
  if (!is_device_ptr(x) || !is_device_ptr(k)) {
    set_error("nufft3d_forward_noalloc takes device arrays only (a host array needs a staging copy)");
    return NUFFT3D_ERR_INVALID;
  }
  const NoallocLayout one = noalloc_layout(p, 1);
  if (one.total > kScratchBytes) {
    char b[200];
    snprintf(b, sizeof b, "fNUFT of this plan needs %.1f MiB of scratch per coil; the no-allocation "
             "entry point has %zu MiB", one.total / 1048576.0, kScratchBytes >> 20);
    set_error(b);
    return NUFFT3D_ERR_UNSUPPORTED;
  }
  int g = nc;
  while (g > 1 && noalloc_layout(p, g).total > kScratchBytes) --g;
  const NoallocLayout L = noalloc_layout(p, g);
  unsigned char* base = nullptr;
  N3D_TRY(scratch_on_device(p->device, (void**)&base));
  const bool conv = p->io_prec != p->prec;
  const size_t io_esz = p->io_prec == 2 ? sizeof(double2) : sizeof(float2);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2];
  const size_t M = (size_t)p->M;
  const bool prof = p->prof.on;   // no profiling events here
  p->prof.on = false;
  int st = NUFFT3D_OK;
  for (int c0 = 0; c0 < nc && st == NUFFT3D_OK; c0 += g) {
    const int cg = nc - c0 < g ? nc - c0 : g;
    const unsigned char* xg = (const unsigned char*)x + (size_t)c0 * nimg * io_esz;
    unsigned char* kg = (unsigned char*)k + (size_t)c0 * M * io_esz;
    const void* img = xg;
    if (conv) {
      const long long n = (long long)(nimg * cg * 2);
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(n)), dim3(256), 0, p->stream,
                         (const float*)xg, (double*)(base + L.img), n);
      img = base + L.img;
    }
    void* out = conv ? (void*)(base + L.kout) : (void*)kg;
    st = fft_forward(p, img, cg, base + L.grid, base + L.t1, base + L.t2);             // fNUFT.m:5-6
    if (!st) st = interp(p, base + L.grid, cg, out, 1, nullptr, 1.0 / sqrt((double)p->nnz), false);
    if (!st) st = zero_oob(p, out, cg);
    if (!st && conv && M > 0) {
      const long long n = (long long)(M * cg * 2);
      hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(n)), dim3(256), 0, p->stream,
                         (const double*)out, (float*)kg, n);
    }
    if (!st && hipGetLastError() != hipSuccess) { set_error("kernel launch failed"); st = NUFFT3D_ERR_DEVICE; }
  }
  p->prof.on = prof;
  return st;
}

// W handling of iNUFT.m:55-74 -> per sorted sample weight W(row)*d. With
// maxit <= 1 the reference only warns (iNUFT.m:57-58): W stays the caller's
// length-nrow vector, not multiplied by the density and not validated, and
// W.*raw needs it to have nrow elements.
static int make_weights(nufft3d_plan* p, const double* W, int64_t nW, int maxit, void** out) {
  *out = p->d_dens;
  if (!W || nW <= 1) return NUFFT3D_OK;   // absent, scalar or empty: W = d
  const int64_t M = p->M;
  const bool as_given = maxit <= 1;
  std::vector<double> full;
  if (as_given && nW != M) {
    set_error("W must have one element per row of raw when maxit<=1 (W.*raw, iNUFT.m:122)");
    return NUFFT3D_ERR_INVALID;
  }
  if (nW == M) {
    full.assign(W, W + M);
  } else {
    const int64_t ny = nW;
    if (M % ny) { set_error("W must be a vector of length ny or ny*nr"); return NUFFT3D_ERR_INVALID; }
    const int64_t nr = M / ny;
    full.resize(M);
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nr; ++i) full[i + nr * j] = W[j];
  }
  if (!as_given)
    for (double v : full)
      if (!std::isfinite(v) || v < 0) { set_error("W must be finite and nonnegative"); return NUFFT3D_ERR_INVALID; }
  double* dW = nullptr;
  N3D_HIP(dev_malloc(&dW, sizeof(double) * M));
  N3D_HIP(hipMemcpyAsync(dW, full.data(), sizeof(double) * M, hipMemcpyHostToDevice, p->stream));
  const size_t esz = p->prec == 2 ? sizeof(double) : sizeof(float);
  void* wd = nullptr;
  N3D_HIP(dev_malloc(&wd, esz * (p->nnz > 0 ? p->nnz : 1)));
  if (p->nnz > 0) {
    if (p->prec == 2)
      hipLaunchKernelGGL((k_gather_real<double>), dim3(nblocks(p->nnz)), dim3(256), 0, p->stream,
                         dW, p->d_orig, as_given ? nullptr : (const double*)p->d_dens,
                         (long long)p->nnz, (double*)wd);
    else
      hipLaunchKernelGGL((k_gather_real<float>), dim3(nblocks(p->nnz)), dim3(256), 0, p->stream,
                         dW, p->d_orig, as_given ? nullptr : (const float*)p->d_dens,
                         (long long)p->nnz, (float*)wd);
  }
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(dW);
  *out = wd;
  return NUFFT3D_OK;
}

int nufft3d_inverse(nufft3d_plan* p, const void* raw, int32_t nc, int32_t maxit, double damp,
                    const double* W, int64_t nW, double tol, void* im, nufft3d_info* info) {
  N3D_PLAN_GUARD(p);
  if (!raw || !im || nc <= 0) { set_error("invalid raw/im/nc"); return NUFFT3D_ERR_INVALID; }
  if (maxit < 0) { set_error("maxit must be a nonnegative integer"); return NUFFT3D_ERR_INVALID; }
  if (!(damp >= 0) || !std::isfinite(damp)) { set_error("damp must be finite and nonnegative"); return NUFFT3D_ERR_INVALID; }
  if (!(tol > 0)) tol = 1e-4;                                                  // iNUFT.m:114
  auto t0 = std::chrono::steady_clock::now();
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  void* wsamp = nullptr;
  N3D_TRY(make_weights(p, W, nW, maxit, &wsamp));
  Workspace tmp_out, tmp_io;
  Staged ri, xo;
  int st = stage_in_c(p, raw, nk, p->ws_img, tmp_io, ri);
  if (!st) st = stage_out_c(p, im, nimg, tmp_out, xo);
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz);
  // b = aNUFT(W.*raw) (iNUFT.m:122,130)
  if (!st) st = permute_in(p, ri.dev, nc, p->ws_samp.p, wsamp);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr);
  void* bdst = xo.dev;
  if (!st && maxit > 0) {
    st = ensure(p->ws_b, nimg * esz);    // kept on the plan: no free per call
    bdst = p->ws_b.p;
  }
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, bdst, 1.0 / sqrt((double)p->nnz), nullptr, 0.0);
  if (!st && maxit > 0)                                                        // iNUFT.m:127-136
    st = minres_solve(p, bdst, nc, maxit, damp, wsamp, tol, xo.dev, info);
  else if (!st && info) {
    info->nrhs = nc;
    for (int c = 0; c < nc && c < 64; ++c) { info->iters[c] = 0; info->relres[c] = NAN; }
  }
  if (!st) st = finish_out_c(p, im, nimg, xo, tmp_io);
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in iNUFT");
    st = NUFFT3D_ERR_DEVICE;
  }
  release(tmp_out);
  release(tmp_io);
  if (wsamp != p->d_dens) (void)dev_free(wsamp);
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

int nufft3d_inverse_phase(nufft3d_plan* p, const void* raw, int32_t nc, int32_t maxit, double damp,
                          const double* W, int64_t nW, double lambda, double tol, void* im,
                          nufft3d_info* info) {
  N3D_PLAN_GUARD(p);
  if (!raw || !im || nc <= 0) { set_error("invalid raw/im/nc"); return NUFFT3D_ERR_INVALID; }
  if (maxit < 0) { set_error("maxit must be a nonnegative integer"); return NUFFT3D_ERR_INVALID; }
  if (!(damp >= 0) || !std::isfinite(damp)) { set_error("damp must be finite and nonnegative"); return NUFFT3D_ERR_INVALID; }
  if (!(lambda >= 0) || !std::isfinite(lambda)) { set_error("lambda must be finite and nonnegative"); return NUFFT3D_ERR_INVALID; }
  if (maxit <= 1)                       // iNUFT.m:78-79: the constraint needs maxit > 1
    return nufft3d_inverse(p, raw, nc, maxit, damp, W, nW, tol, im, info);
  if (!(tol > 0)) tol = 1e-4;                                                  // iNUFT.m:114
  auto t0 = std::chrono::steady_clock::now();
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  void* wsamp = nullptr;
  N3D_TRY(make_weights(p, W, nW, maxit, &wsamp));
  Workspace tmp_out, tmp_io, tmp_x, tmp_P, tmp_b;
  Staged ri, xo;
  int st = stage_in_c(p, raw, nk, p->ws_img, tmp_io, ri);
  if (!st) st = stage_out_c(p, im, nimg, tmp_out, xo);
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz);
  if (!st) st = ensure(p->ws_b, nimg * esz);
  if (!st) st = ensure(tmp_x, nimg * esz);
  if (!st) st = ensure(tmp_P, nimg * esz);
  if (!st) st = ensure(tmp_b, nimg * esz);
  // b = aNUFT(W.*raw), x = minres(A'WA + damp^2, b) (iNUFT.m:126-136)
  if (!st) st = permute_in(p, ri.dev, nc, p->ws_samp.p, wsamp);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr);
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, p->ws_b.p, 1.0 / sqrt((double)p->nnz), nullptr, 0.0);
  if (!st) st = minres_solve(p, p->ws_b.p, nc, maxit, damp, wsamp, tol, tmp_x.p, nullptr);
  // smoothed phase, b' = conj(P) b, CG on P'A'WAP + i lambda^2 Im, x = P x' (iNUFT.m:147-165)
  if (!st) st = phase_estimate(p, tmp_x.p, nc, tmp_P.p);
  if (!st) st = phase_mul(p, tmp_P.p, p->ws_b.p, tmp_b.p, nc, true);
  if (!st) st = phase_cg(p, tmp_b.p, nc, maxit, damp, lambda, wsamp, tmp_P.p, tol, tmp_x.p, info);
  if (!st) st = phase_mul(p, tmp_P.p, tmp_x.p, xo.dev, nc, false);
  if (!st) st = finish_out_c(p, im, nimg, xo, tmp_io);
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in iNUFT (phase constraint)");
    st = NUFFT3D_ERR_DEVICE;
  }
  release(tmp_out); release(tmp_io); release(tmp_x); release(tmp_P); release(tmp_b);
  if (wsamp != p->d_dens) (void)dev_free(wsamp);
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

int nufft3d_inverse_cs(nufft3d_plan* p, const void* raw, int32_t nc, int32_t maxit, double damp,
                       const double* W, int64_t nW, double lambda, double tol, void* im,
                       nufft3d_info* info) {
  N3D_PLAN_GUARD(p);
  if (!raw || !im || nc <= 0) { set_error("invalid raw/im/nc"); return NUFFT3D_ERR_INVALID; }
  if (maxit < 0) { set_error("maxit must be a nonnegative integer"); return NUFFT3D_ERR_INVALID; }
  if (!(damp >= 0) || !std::isfinite(damp)) { set_error("damp must be finite and nonnegative"); return NUFFT3D_ERR_INVALID; }
  if (!(lambda >= 0) || !(lambda <= 1)) {
    set_error("lambda (fraction of zero wavelet coefficients) must be in [0, 1]");
    return NUFFT3D_ERR_INVALID;
  }
  if (maxit <= 1)                       // iNUFT.m:78-79: the constraint needs maxit > 1
    return nufft3d_inverse(p, raw, nc, maxit, damp, W, nW, tol, im, info);
  if (!(tol > 0)) tol = 1e-4;                                                  // iNUFT.m:114
  auto t0 = std::chrono::steady_clock::now();
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  void* wsamp = nullptr;
  N3D_TRY(make_weights(p, W, nW, maxit, &wsamp));
  Workspace tmp_out, tmp_io;
  Staged ri, xo;
  int st = stage_in_c(p, raw, nk, p->ws_img, tmp_io, ri);
  if (!st) st = stage_out_c(p, im, nimg, tmp_out, xo);
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz);
  if (!st) st = ensure(p->ws_b, nimg * esz);
  // b = aNUFT(W.*raw) (iNUFT.m:177), x = fistaN(A'WA + damp^2, b, lambda, tol, maxit, Haar)
  if (!st) st = permute_in(p, ri.dev, nc, p->ws_samp.p, wsamp);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr);
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, p->ws_b.p, 1.0 / sqrt((double)p->nnz), nullptr, 0.0);
  if (!st) st = fista_solve(p, p->ws_b.p, nc, maxit, damp, lambda, wsamp, tol, xo.dev, info);
  if (!st) st = finish_out_c(p, im, nimg, xo, tmp_io);
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in iNUFT (compressed sensing)");
    st = NUFFT3D_ERR_DEVICE;
  }
  release(tmp_out); release(tmp_io);
  if (wsamp != p->d_dens) (void)dev_free(wsamp);
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

// discrep.m:2-69 (helpers above the C ABI block)
int nufft3d_discrep(nufft3d_plan* p, const void* data, int32_t nc, double stddev, int32_t iters,
                    double* damp_out, double* stddev_out, double* history) {
  N3D_PLAN_GUARD(p);
  if (!data || !damp_out || nc <= 0) { set_error("invalid data/damp/nc"); return NUFFT3D_ERR_INVALID; }
  if (!std::isfinite(stddev)) {                                                // discrep.m:28
    set_error("Expected input stddev to be finite.");
    return NUFFT3D_ERR_INVALID;
  }
  if (iters <= 0) iters = 20;                                                  // discrep.m:54
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  const int maxit = (int)std::min<long long>(20, (long long)p->N[0] * p->N[1] * p->N[2]);
  Workspace tmp_io, xw, kw;
  Staged di;
  int st = stage_in_c(p, data, nk, p->ws_img, tmp_io, di);
  if (!st && stddev < 0)                                                       // discrep.m:29-33
    st = p->prec == 2 ? mad_stddev<double>(p, di.dev, (long long)nk, &stddev)
                      : mad_stddev<float>(p, di.dev, (long long)nk, &stddev);
  const double discrepancy = stddev * std::sqrt((double)nk);                   // discrep.m:37
  const double sc = 1.0 / std::sqrt((double)p->nnz);
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz);
  if (!st) st = ensure(p->ws_b, nimg * esz);
  if (!st) st = ensure(xw, nimg * esz);
  if (!st) st = ensure(kw, nk * esz);
  // b = aNUFT(d.*data) (discrep.m:40-46)
  if (!st) st = permute_in(p, di.dev, nc, p->ws_samp.p, p->d_dens);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr);
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, p->ws_b.p, sc, nullptr, 0.0);
  double damp = 0.1;                                                           // discrep.m:50
  for (int it = 0; it < iters && !st; ++it) {                                  // discrep.m:54-68
    st = minres_solve(p, p->ws_b.p, nc, maxit, damp, p->d_dens, 1e-4, xw.p, nullptr, it > 0);
    if (!st) st = fft_forward(p, xw.p, nc, p->ws_grid.p);
    if (!st) st = interp(p, p->ws_grid.p, nc, kw.p, 1, nullptr, sc);
    if (!st) st = zero_oob(p, kw.p, nc);
    double r2 = 0;
    if (!st) st = diff_norm2(p, kw.p, di.dev, (int64_t)nk, &r2);
    if (!st) {
      damp = damp * (discrepancy / std::sqrt(r2));
      if (history) history[it] = damp;
    }
  }
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in discrep");
    st = NUFFT3D_ERR_DEVICE;
  }
  release(xw);
  release(kw);
  release(tmp_io);
  if (!st) {
    *damp_out = damp;
    if (stddev_out) *stddev_out = stddev;
  }
  return st;
}

int nufft3d_normal(nufft3d_plan* p, const void* x, int32_t nc, double damp, void* y) {
  N3D_PLAN_GUARD(p);
  if (!x || !y || nc <= 0) { set_error("invalid x/y/nc"); return NUFFT3D_ERR_INVALID; }
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  Workspace tmp_out, tmp_io;
  Staged xi, yo;
  N3D_TRY(stage_in_c(p, x, nimg, p->ws_img, tmp_io, xi));
  int st = stage_out_c(p, y, nimg, tmp_out, yo);
  if (!st) st = apply_normal(p, xi.dev, nc, damp, p->d_dens, yo.dev);
  if (!st) st = finish_out_c(p, y, nimg, yo, tmp_io);
  release(tmp_out);
  release(tmp_io);
  if (xi.host) N3D_HIP(hipStreamSynchronize(p->stream));
  return st;
}

int nufft3d_density(nufft3d_plan* p, void* d) {
  N3D_PLAN_GUARD(p);
  if (!d) { set_error("d is NULL"); return NUFFT3D_ERR_INVALID; }
  Workspace tmp_out, tmp_io;
  Staged o;
  N3D_TRY(stage_out_c(p, d, (size_t)p->M, tmp_out, o, false));
  if (p->prec == 2)
    hipLaunchKernelGGL((k_scatter_dens<double>), dim3(nblocks(p->M)), dim3(256), 0, p->stream,
                       (const double*)p->d_dens, p->d_orig, p->d_oob, (long long)p->nnz,
                       (long long)p->M, (double*)o.dev);
  else
    hipLaunchKernelGGL((k_scatter_dens<float>), dim3(nblocks(p->M)), dim3(256), 0, p->stream,
                       (const float*)p->d_dens, p->d_orig, p->d_oob, (long long)p->nnz,
                       (long long)p->M, (float*)o.dev);
  N3D_HIP(hipGetLastError());
  int st = finish_out_c(p, d, (size_t)p->M, o, tmp_io, false);
  N3D_HIP(hipStreamSynchronize(p->stream));
  release(tmp_out);
  release(tmp_io);
  return st;
}

int nufft3d_autotune(nufft3d_plan* p, int32_t nc, int32_t reps, char* report, int64_t len) {
  N3D_PLAN_GUARD(p);
  if (nc <= 0 || reps <= 0 || (report && len <= 0)) {
    set_error("invalid nc/reps/report");
    return NUFFT3D_ERR_INVALID;
  }
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nks = (size_t)(p->nnz > 0 ? p->nnz : 1) * nc;
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  const double sc = 1.0 / sqrt((double)std::max<int64_t>(p->nnz, 1));
  Workspace xin, xout, ks;
  int st = ensure(xin, nimg * esz);
  if (!st) st = ensure(xout, nimg * esz);
  if (!st) st = ensure(ks, nks * esz);
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (st) { release(xin); release(xout); release(ks); return st; }
  const long long nr = (long long)(nimg * esz / (p->prec == 2 ? 8 : 4));
  if (p->prec == 2)
    hipLaunchKernelGGL((k_fill_pattern<double>), dim3(nblocks(nr)), dim3(256), 0, p->stream,
                       (double*)xin.p, nr, 17u);
  else
    hipLaunchKernelGGL((k_fill_pattern<float>), dim3(nblocks(nr)), dim3(256), 0, p->stream,
                       (float*)xin.p, nr, 17u);
  // one operator pair on device arrays in the internal layouts
  // (split-precision plans: their fNUFT/aNUFT sequence, the adjoint from
  // float samples)
  const bool split = split_precision(p);
  Workspace ks32;
  if (split) {
    if (!st) st = ensure(ks32, nks * sizeof(float2));
    if (!st && hipMemsetAsync(ks32.p, 0, nks * sizeof(float2), p->stream) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMemsetAsync failed");
      st = NUFFT3D_ERR_DEVICE;
    }
  }
  auto pair = [&]() -> int {
    if (split) {   // the adjoint from float samples (ks32)
      N3D_TRY(fft_forward(p, xin.p, nc, p->ws_grid.p));
      N3D_TRY(interp(p, p->ws_grid.p, nc, ks.p, 0, nullptr, sc));
      N3D_TRY(spread(p, ks32.p, nc, p->ws_grid.p, 0, nullptr, 1));
      return fft_adjoint(p, p->ws_grid.p, nc, xout.p, sc, nullptr, 0.0);
    }
    N3D_TRY(fft_forward(p, xin.p, nc, p->ws_grid.p));
    N3D_TRY(interp(p, p->ws_grid.p, nc, ks.p, 0, nullptr, sc));
    N3D_TRY(spread(p, ks.p, nc, p->ws_grid.p, 0, nullptr));
    return fft_adjoint(p, p->ws_grid.p, nc, xout.p, sc, nullptr, 0.0);
  };
  if (!st && !split) st = ensure_brick_lists(p);
  if (!st && p->prec == 2) st = ensure_block_lists(p);
  if (st) { release(xin); release(xout); release(ks); release(ks32); return st; }
  // NUFFT3D_TUNE_BLOCK_ONLY=1 restricts the choice to the block-staged kernels
  // where the plan has their lists
  const char* ob = getenv("NUFFT3D_TUNE_BLOCK_ONLY");
  const bool only_blk = ob && atoi(ob) != 0;
  // brick: 0 linear gather spread, 1 brick-order gather spread, 2 LDS-staged
  // spread (plans with brick lists), 3 block-staged split adjoint (plans
  // with block lists; 8-coil groups)
  struct Cand { int fftb, group, xcd, brick; float ms; };
  std::vector<Cand> cands;
  for (int fftb : {0, 8, 4})
    for (int group : {16, 8})
      for (int xcd : {0, 32, -1})
        for (int brick : {1, 0, 2, 3}) {
          if (group == 8 && nc <= 8) continue;   // same launches as 16
          if (brick == 2 && !(split ? (void*)p->wl.d_went : (void*)p->bl.d_ent)) continue;
          if (brick == 3 && !p->bk.d_piece) continue;   // (8-coil spread groups)
          if (brick != 3 && !p->d_hitw) continue;       // gather lists released
          if (brick != 3 && only_blk && p->bk.d_piece) continue;
          cands.push_back({fftb, group, xcd, brick, 0.f});
        }
  auto flags_for = [&](int brick, int keep) {
    keep &= ~(2 | 8 | 4096);
    return brick == 0 ? keep : brick == 1 ? (keep | 2) : brick == 2 ? (keep | 2 | 8) : (keep | 2 | 4096);
  };
  const int keep_flags = p->tune_flags;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipGetLastError();
    st = NUFFT3D_ERR_DEVICE;
    set_error("hipEventCreate failed");
  }
  int best = -1;
  for (size_t i = 0; i < cands.size() && !st; ++i) {
    Cand& c = cands[i];
    p->fft_bundle = c.fftb;
    p->coil_group = c.group;
    p->xcd_chunk = c.xcd;
    p->tune_flags = flags_for(c.brick, keep_flags);
    st = pair();                                          // warm-up
    if (!st && hipEventRecord(e0, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
    for (int r = 0; r < reps && !st; ++r) st = pair();
    if (!st && hipEventRecord(e1, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
    if (!st && hipEventSynchronize(e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
    float ms = 0;
    if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
    c.ms = ms / reps;
    if (!st && (best < 0 || c.ms < cands[best].ms)) best = (int)i;
  }
  if (best >= 0) {
    const Cand& b = cands[best];
    p->fft_bundle = b.fftb;
    p->coil_group = b.group;
    p->xcd_chunk = b.xcd;
    p->tune_flags = flags_for(b.brick, keep_flags);
  } else {
    p->tune_flags = keep_flags;
  }
  // block lists the choice does not use (and the caller did not ask for) are freed
  if (p->bk.d_piece && !(p->tune_flags & 4096)) {
    (void)hipStreamSynchronize(p->stream);
    drop_block_lists(p);
  }
  // z-pass work-group size, timed on the best setting: 256 threads in both
  // directions, and the automatic per-direction choice (0)
  float zt_ms[2] = {0.f, 0.f};
  int zt_best = p->fft_zthreads;
  if (!st && best >= 0 && p->K[2] > 1) {
    const int keep_zt = p->fft_zthreads;
    for (int zi = 0; zi < 2 && !st; ++zi) {
      p->fft_zthreads = zi ? 0 : 256;
      st = pair();
      if (!st && hipEventRecord(e0, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      for (int r = 0; r < reps && !st; ++r) st = pair();
      if (!st && hipEventRecord(e1, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      if (!st && hipEventSynchronize(e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      float ms = 0;
      if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      zt_ms[zi] = ms / reps;
    }
    zt_best = st ? keep_zt : (zt_ms[1] < zt_ms[0] ? 0 : 256);
    p->fft_zthreads = zt_best;
  }
  // forward kernel, timed likewise on the best setting: the gather, or the
  // block-staged forward where its lists can be built (fp64 arithmetic, 3-D);
  // lists the choice does not use (and the caller did not ask for) are freed
  // (0 gather, 1 block-staged evaluating the weights, 2 block-staged reading
  // the lists' stored weights)
  float fw_ms[3] = {0.f, 0.f, 0.f};
  int fw_best = (p->tune_flags & 8192) ? ((p->tune_flags & 16384) ? 2 : 1) : 0;
  auto fw_flags = [](int keep, int fi) {
    keep &= ~(8192 | 16384);
    return fi == 0 ? keep : fi == 1 ? (keep | 8192) : (keep | 8192 | 16384);
  };
  if (!st && best >= 0 && p->prec == 2 && p->K[2] > 1) {
    st = ensure_fwd_lists(p, true);
    if (!st && p->fk.d_slot) {
      const int keep = p->tune_flags;
      for (int fi = 0; fi < 3 && !st; ++fi) {
        if (fi == 2 && !p->fk.d_w) continue;
        if (fi == 0 && !p->d_fcode) continue;   // gather lists released
        if (fi == 0 && only_blk) continue;
        p->tune_flags = fw_flags(keep, fi);
        st = pair();
        if (!st && hipEventRecord(e0, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
        for (int r = 0; r < reps && !st; ++r) st = pair();
        if (!st && hipEventRecord(e1, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
        if (!st && hipEventSynchronize(e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
        float ms = 0;
        if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
        fw_ms[fi] = ms / reps;
      }
      if (!st) {
        fw_best = -1;
        for (int fi = 0; fi < 3; ++fi)
          if (fw_ms[fi] > 0.f && (fw_best < 0 || fw_ms[fi] < fw_ms[fw_best])) fw_best = fi;
        if (fw_best < 0) fw_best = 1;
      }
      p->tune_flags = fw_flags(keep, fw_best);
      if (fw_best == 0 && !(keep_flags & 8192)) {
        (void)hipStreamSynchronize(p->stream);
        drop_fwd_lists(p);
      } else if (fw_best == 1 && p->fk.d_w && !(keep_flags & 16384)) {
        (void)hipStreamSynchronize(p->stream);
        (void)dev_free(p->fk.d_w);   // the evaluated form keeps no weights
        p->fk.d_w = nullptr;
      }
    }
  }
  // with both block-staged kernels chosen, every gridding launch of the plan
  // runs them (any coil count, data weights, either precision of samples):
  // the gather kernels' per-hit lists (8 B weight table + 2 B codes per hit)
  // go, unless a kernel that reads them was asked for (LDS-staged brick
  // spreads, evaluated spread weights, the tiled interpolation). A later
  // autotune then times the block-staged kernels only.
  if (!st && p->bk.d_piece && p->fk.d_slot && (p->tune_flags & 4096) && (p->tune_flags & 8192) &&
      !(p->tune_flags & (8 | 128)) && !p->tile_interp && (p->d_hitw || p->d_fcode)) {
    (void)hipStreamSynchronize(p->stream);
    for (void** q : {(void**)&p->d_hitw, (void**)&p->d_fcode, (void**)&p->d_hit})
      if (*q) { (void)dev_free(*q); *q = nullptr; }
  }
  // x-pass work-group size, timed likewise: 256 or 512 threads
  float xt_ms[2] = {0.f, 0.f};
  int xt_best = p->fft_xthreads;
  if (!st && best >= 0) {
    const int keep_xt = p->fft_xthreads;
    for (int xi = 0; xi < 2 && !st; ++xi) {
      p->fft_xthreads = xi ? 512 : 256;
      st = pair();
      if (!st && hipEventRecord(e0, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      for (int r = 0; r < reps && !st; ++r) st = pair();
      if (!st && hipEventRecord(e1, p->stream) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      if (!st && hipEventSynchronize(e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      float ms = 0;
      if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = NUFFT3D_ERR_DEVICE;
      xt_ms[xi] = ms / reps;
    }
    xt_best = st ? keep_xt : (xt_ms[1] < xt_ms[0] ? 512 : 256);
    p->fft_xthreads = xt_best;
  }
  if (st == NUFFT3D_ERR_DEVICE) { (void)hipGetLastError(); set_error("device error while autotuning"); }
  if (!st && report) {
    std::string js = "{\"nc\": " + std::to_string(nc) + ", \"candidates\": [";
    for (size_t i = 0; i < cands.size(); ++i) {
      static const char* spread_kind[4] = {"gather", "gather-bricks", "lds-bricks", "block-staged"};
      char b[200];
      snprintf(b, sizeof b, "%s{\"fft_lines\": %d, \"coil_group\": %d, \"xcd_chunk\": %d, "
               "\"brick_order\": %d, \"spread\": \"%s\", \"ms_per_pair\": %.5f}", i ? ", " : "",
               cands[i].fftb, cands[i].group, cands[i].xcd, cands[i].brick != 0,
               spread_kind[cands[i].brick], cands[i].ms);
      js += b;
    }
    char zb[400];
    snprintf(zb, sizeof zb, ", \"z_threads\": %d, \"z_threads_ms\": [%.5f, %.5f], "
             "\"x_threads\": %d, \"x_threads_ms\": [%.5f, %.5f], \"forward\": \"%s\", "
             "\"forward_ms\": [%.5f, %.5f, %.5f]", zt_best, zt_ms[0], zt_ms[1],
             xt_best, xt_ms[0], xt_ms[1],
             fw_best == 0 ? "gather" : fw_best == 1 ? "block-staged" : "block-staged-stored",
             fw_ms[0], fw_ms[1], fw_ms[2]);
    js += "], \"best\": " + std::to_string(best) + zb + "}";
    if ((int64_t)js.size() + 1 > len) { set_error("report buffer too small"); st = NUFFT3D_ERR_INVALID; }
    else std::memcpy(report, js.c_str(), js.size() + 1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(p->stream);
  release(xin); release(xout); release(ks); release(ks32);
  return st;
}

int nufft3d_pruno_opts_default(nufft3d_pruno_opts* o) {
  if (!o) { set_error("opts is NULL"); return NUFFT3D_ERR_INVALID; }
  o->width = 5;     // pruno.m:27-30
  o->radial = 1;
  o->noise = -1.0;
  o->tol = 0.01;
  o->literal_first_nv = 0;
  return NUFFT3D_OK;
}

int nufft3d_gram(nufft3d_plan* p, const void* data, const int32_t dims[4], int32_t width,
                 int32_t radial, void* AA, int32_t* nout) {
  N3D_PLAN_GUARD(p);
  if (!data || !dims || !AA || width < 1) { set_error("invalid argument"); return NUFFT3D_ERR_INVALID; }
  for (int a = 0; a < 4; ++a)
    if (dims[a] < 1) { set_error("dims must be positive"); return NUFFT3D_ERR_INVALID; }
  for (int a = 0; a < (dims[2] == 1 ? 2 : 3); ++a)
    if (dims[a] <= width / 2) {   // shifted indices wrap once
      set_error("data dimensions must exceed the kernel half width");
      return NUFFT3D_ERR_UNSUPPORTED;
    }
  std::vector<int4> taps;
  int w = 0;
  kernel_taps(width, radial, dims[2] == 1, taps, w);
  const int dd[4] = {dims[0], dims[1], dims[2], dims[3]};
  const long long n = (long long)taps.size() * dims[3];
  if (nout) *nout = (int32_t)n;
  const size_t nd = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  Staged di, ao;
  Workspace tmp_in, tmp_out, tmp_io;
  N3D_TRY(stage_in_c(p, data, nd, tmp_in, tmp_io, di));
  int st = stage_out(p, AA, sizeof(double2) * n * n, tmp_out, ao);
  if (!st) st = gram_device(p, di.dev, dd, taps, (double2*)ao.dev);
  if (!st) st = finish_out(p, AA, sizeof(double2) * n * n, ao);
  N3D_HIP(hipStreamSynchronize(p->stream));
  release(tmp_in);
  release(tmp_out);
  release(tmp_io);
  return st;
}

int nufft3d_gram_loraks(nufft3d_plan* p, const void* data, const int32_t dims[4], int32_t width,
                        int32_t radial, const int32_t center[3], void* AA, int32_t* nout) {
  N3D_PLAN_GUARD(p);
  if (!data || !dims || !AA || width < 1) { set_error("invalid argument"); return NUFFT3D_ERR_INVALID; }
  for (int a = 0; a < 4; ++a)
    if (dims[a] < 1) { set_error("dims must be positive"); return NUFFT3D_ERR_INVALID; }
  int cen[3] = {1, 1, 1};                                   // sake.m:79-83
  if (center)
    for (int a = 0; a < 3; ++a) {
      if (center[a] < 1 || center[a] > dims[a]) { set_error("center outside the data"); return NUFFT3D_ERR_INVALID; }
      cen[a] = center[a];
    }
  for (int a = 0; a < (dims[2] == 1 ? 2 : 3); ++a)
    if (dims[a] <= width / 2) {   // shifted indices wrap once
      set_error("data dimensions must exceed the kernel half width");
      return NUFFT3D_ERR_UNSUPPORTED;
    }
  std::vector<int4> taps;
  int w = 0;
  kernel_taps(width, radial, dims[2] == 1, taps, w);
  const int dd[4] = {dims[0], dims[1], dims[2], dims[3]};
  const long long n2 = 2LL * (long long)taps.size() * dims[3];
  if (nout) *nout = (int32_t)n2;
  const size_t nd = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  Staged di, ao;
  Workspace tmp_in, tmp_out, tmp_io;
  N3D_TRY(stage_in_c(p, data, nd, tmp_in, tmp_io, di));
  int st = stage_out(p, AA, sizeof(double2) * n2 * n2, tmp_out, ao);
  if (!st) st = gram_device(p, di.dev, dd, taps, (double2*)ao.dev, true, cen);
  if (!st) st = finish_out(p, AA, sizeof(double2) * n2 * n2, ao);
  N3D_HIP(hipStreamSynchronize(p->stream));
  release(tmp_in);
  release(tmp_out);
  release(tmp_io);
  return st;
}

int nufft3d_pruno(nufft3d_plan* p, const void* data, int32_t nc, const nufft3d_pruno_opts* opts,
                  int32_t* nv, double* sv, double* noise, void* Kt) {
  N3D_PLAN_GUARD(p);
  if (!data || nc < 1) { set_error("invalid argument"); return NUFFT3D_ERR_INVALID; }
  nufft3d_pruno_opts o;
  if (opts) o = *opts; else nufft3d_pruno_opts_default(&o);
  Staged di;
  Workspace tmp_in, tmp_io;
  N3D_TRY(stage_in_c(p, data, (size_t)p->N[0] * p->N[1] * p->N[2] * nc, tmp_in, tmp_io, di));
  std::vector<double> svd;
  int nvv = 0;
  double nz = 0;
  int st = pruno_build(p, di.dev, nc, o, 1.0, nullptr, &nvv, &svd, &nz, Kt);
  if (!st) {
    if (nv) *nv = nvv;
    if (noise) *noise = nz;
    if (sv) std::memcpy(sv, svd.data(), sizeof(double) * svd.size());
  }
  N3D_HIP(hipStreamSynchronize(p->stream));
  release(tmp_in);
  release(tmp_io);
  return st;
}

int nufft3d_inverse_pruno(nufft3d_plan* p, const void* raw, int32_t nc, int32_t maxit, double damp,
                          const double* W, int64_t nW, double lambda,
                          const nufft3d_pruno_opts* opts, int32_t iters, void* im,
                          nufft3d_info* info) {
  N3D_PLAN_GUARD(p);
  if (!raw || !im || nc < 2) {
    set_error("parallel-imaging requires multiple coils");               // iNUFT.m:84
    return NUFFT3D_ERR_INVALID;
  }
  if (maxit < 0 || iters < 0 || !(damp >= 0) || !(lambda >= 0) || !std::isfinite(lambda)) {
    set_error("invalid maxit/iters/damp/lambda");
    return NUFFT3D_ERR_INVALID;
  }
  nufft3d_pruno_opts o;
  if (opts) o = *opts; else nufft3d_pruno_opts_default(&o);
  const double tol = 1e-4;                                                 // iNUFT.m:114
  auto t0 = std::chrono::steady_clock::now();
  const size_t esz = cp_esz(p);
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  void* wsamp = nullptr;
  N3D_TRY(make_weights(p, W, nW, maxit, &wsamp));
  Workspace tmp_out, tmp_b, tmp_x0, tmp_k, tmp_io;
  Staged ri, xo;
  void* dB = nullptr;
  int st = stage_in_c(p, raw, nk, p->ws_img, tmp_io, ri);
  if (!st) st = stage_out_c(p, im, nimg, tmp_out, xo);
  const size_t ngrid = (size_t)p->K[0] * p->K[1] * p->K[2] * nc;
  if (!st) st = ensure(p->ws_grid, ngrid * esz);
  if (!st) st = ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz);
  if (!st) st = ensure(tmp_b, nimg * esz);
  if (!st) st = ensure(tmp_x0, nimg * esz);
  if (!st) st = ensure(tmp_k, nimg * esz);
  // b = aNUFT(W.*raw)
  if (!st) st = permute_in(p, ri.dev, nc, p->ws_samp.p, wsamp);
  if (!st) st = spread(p, p->ws_samp.p, nc, p->ws_grid.p, 0, nullptr);
  if (!st) st = fft_adjoint(p, p->ws_grid.p, nc, tmp_b.p, 1.0 / sqrt((double)p->nnz), nullptr, 0.0);
  // calibration image (iNUFT.m:119-136), then k-space (iNUFT.m:203-204)
  const void* x0 = tmp_b.p;
  if (!st && maxit > 0) {
    st = minres_solve(p, tmp_b.p, nc, maxit, damp, wsamp, tol, tmp_x0.p, nullptr);
    x0 = tmp_x0.p;
  }
  if (!st) st = fft_plain(p, x0, nc, tmp_k.p);
  int nv = 0;
  if (!st) st = pruno_build(p, tmp_k.p, nc, o, lambda, &dB, &nv, nullptr, nullptr);
  // coupled least squares (iNUFT.m:210-218)
  if (!st) st = minres_coupled(p, tmp_b.p, nc, iters, damp, wsamp, dB, tol, xo.dev, info);
  if (!st) st = finish_out_c(p, im, nimg, xo, tmp_io);
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in iNUFT (pruno)");
    st = NUFFT3D_ERR_DEVICE;
  }
  if (dB) (void)dev_free(dB);
  release(tmp_out); release(tmp_b); release(tmp_x0); release(tmp_k); release(tmp_io);
  if (wsamp != p->d_dens) (void)dev_free(wsamp);
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

int nufft3d_sake_opts_default(nufft3d_sake_opts* o) {
  if (!o) { set_error("invalid argument"); return NUFFT3D_ERR_INVALID; }
  o->width = 4;            // sake.m:26
  o->radial = 1;           // sake.m:27
  o->loraks = 0;           // sake.m:28
  o->tol = 1e-6;           // sake.m:29
  o->maxit = 1000;         // sake.m:30
  o->noise = -1.0;         // sake.m:31 (empty: estimate)
  o->center[0] = o->center[1] = o->center[2] = 1;   // sake.m:79-83
  o->damp = 0.0;           // sake.m:34
  o->sparsity = 0.0;       // sake.m:35
  o->dc_iters = 10;        // sake.m:109
  o->output = 0;
  return NUFFT3D_OK;
}

int nufft3d_sake(nufft3d_plan* p, const void* data, int32_t nc, const nufft3d_sake_opts* opts,
                 void* out, double* history, nufft3d_sake_info* info) {
  N3D_PLAN_GUARD(p);
  if (!data || !out || nc < 2) {
    set_error("'data' must be a float array [npts nc].");                 // sake.m:57-59
    return NUFFT3D_ERR_INVALID;
  }
  nufft3d_sake_opts o;
  if (opts) o = *opts; else nufft3d_sake_opts_default(&o);
  if (o.width < 1 || o.maxit < 0 || o.dc_iters < 0 || !(o.damp >= 0) || !std::isfinite(o.damp) ||
      !(o.tol >= 0) || (o.output != 0 && o.output != 1)) {
    set_error("invalid sake option (width, maxit, dc_iters, damp, tol or output)");
    return NUFFT3D_ERR_INVALID;
  }
  if (o.sparsity != 0) {
    set_error("sparsity needs the DWT wavelet transform (sake.m:91-93), which the reference does not ship");
    return NUFFT3D_ERR_UNSUPPORTED;
  }
  for (int a = 0; a < 3; ++a) {
    if (o.center[a] < 1 || o.center[a] > p->N[a]) { set_error("center outside the data"); return NUFFT3D_ERR_INVALID; }
    if (p->N[a] <= o.width / 2) {   // shifted indices wrap at most once
      set_error("image dimensions must exceed the kernel half width");
      return NUFFT3D_ERR_UNSUPPORTED;
    }
  }
  auto t0 = std::chrono::steady_clock::now();
  const size_t nimg = (size_t)p->N[0] * p->N[1] * p->N[2] * nc;
  const size_t nk = (size_t)p->M * nc;
  void* wsamp = p->d_dens;                                                 // iNUFT default W = d
  Workspace tmp_out, tmp_io, tmp_in;
  Staged di, xo;
  int st = stage_in_c(p, data, nk, tmp_in, tmp_io, di);
  if (!st) st = stage_out_c(p, out, nimg, tmp_out, xo);
  if (!st) st = sake_run(p, di.dev, nc, o, wsamp, xo.dev, history, info);
  if (!st) st = finish_out_c(p, out, nimg, xo, tmp_io);
  if (!st && hipStreamSynchronize(p->stream) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device error in sake");
    st = NUFFT3D_ERR_DEVICE;
  }
  release(tmp_out); release(tmp_io); release(tmp_in);
  if (info) info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

}  // extern "C"
