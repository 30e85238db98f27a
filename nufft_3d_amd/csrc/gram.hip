// Hankel (calibration) Gram of SAKE/PRUNO and the PRUNO null operator.
//
// Reference: make_data_matrix (pruno.m:146-179, sake.m:176-224):
//   AA(c,j,c',k) = sum_v conj(row_j[v,c]) * col_k[v,c'],
//   row_j = circshift(data, s_j), col_k = circshift(data, s_k),
//   reshaped to an (nc*nk) x (nc*nk) matrix with the coil index fastest.
// circshift by s gives out[v] = data[(v - s) mod N] per axis.
//
// As a GEMM: X = [V x 2n] real, X[v, p*n + c + nc*j] = Re/Im data_c[v - s_j],
// C = X^T X (2n x 2n, symmetric), AA = (C11 + C22) + i (C12 - C21).
// fp32 plans: split-bf16 (hi/lo) MFMA with fp32 accumulation,
//   x*y ~ xh*yh + xh*yl + xl*yh (the lo*lo term, ~2^-16 relative, dropped);
//   v_mfma_f32_32x32x16_bf16, 128x128 block tiles, upper tiles only, V split
//   across work-groups; partial tiles summed in a fixed order in double.
//   The operands are prepared once per call: the split-bf16 Re/Im planes of
//   the data, pre-shifted for every x-shift of the kernel and padded to rows
//   of a multiple of 32, so a K step (32 consecutive x of one row) reads one
//   aligned 64-byte run per column and array.
// fp64 plans: the same Gram in double with plain FMAs (oracle mode).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nufft3d_internal.h"

namespace nufft3d {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct GramGeom {
  int N1, N2, N3;
  long long V;
  int nc, nk, n, n2;      // n = nc*nk, n2 = 2n
  int ntr;                // tile rows (= tile cols) of C
  int nsplit;             // V chunks (work-groups per tile along V)
  int h;                  // kernel half width (shifts in [-h, h] per axis)
  int Lx;                 // padded row length of the prepared planes (multiple of 32)
  int nxc;                // K steps per row (Lx / 32)
  long long plane;        // N3 * N2 * Lx
  long long nsteps;       // N2 * N3 * nxc
  long long spw;          // K steps per work-group
};

// A column of X: element offset of its hi plane in the prepared operands
// (the lo plane is nc*plane further), its y/z shifts, and whether it exists
// (the last tile is padded past 2n).
struct GramCol {
  long long base;
  int sy, sz;
  int valid, pad;
};

constexpr int GT = 128;   // block tile edge
constexpr int GK = 32;    // voxels per K step
constexpr int GPAD = 8;   // bf16 padding per LDS row

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  // round to nearest even (values are finite)
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f(unsigned short b) {
  return __uint_as_float(((unsigned)b) << 16);
}

// Q[sxi][p][hl][c][z][y][x], x in [0, Lx): the hi (hl=0) and lo (hl=1) bf16
// parts of Re (p=0) / Im (p=1) data_c[(x - sx) mod N1, y, z], sx = sxi - h,
// for x < N1; zero for x >= N1 (so padded voxels add nothing).
__global__ void k_gram_prep(const float2* __restrict__ data, unsigned short* __restrict__ Q,
                            GramGeom g, int nsx) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nsx * g.nc * g.plane;
  if (idx >= total) return;
  const int x = (int)(idx % g.Lx);
  const long long yz = (idx / g.Lx) % ((long long)g.N2 * g.N3);
  const int c = (int)((idx / g.plane) % g.nc);
  const int sxi = (int)(idx / (g.plane * g.nc));
  unsigned short rh = 0, rl = 0, ih = 0, il = 0;
  if (x < g.N1) {
    int xs = x - (sxi - g.h);
    xs = xs < 0 ? xs + g.N1 : (xs >= g.N1 ? xs - g.N1 : xs);
    const float2 d = data[(long long)c * g.V + xs + (long long)g.N1 * yz];
    rh = bf16_bits(d.x);
    rl = bf16_bits(d.x - bf16_to_f(rh));
    ih = bf16_bits(d.y);
    il = bf16_bits(d.y - bf16_to_f(ih));
  }
  const long long pl = (long long)g.nc * g.plane;       // one (sxi, p, hl) block
  const long long o = (long long)sxi * 4 * pl + (long long)c * g.plane + yz * g.Lx + x;
  Q[o] = rh;
  Q[o + pl] = rl;
  Q[o + 2 * pl] = ih;
  Q[o + 3 * pl] = il;
}

// C = X^T Y tiles: X columns from the prepared operand Qa (column table
// colsA), Y columns from Qb (colsB); Y = X for AA, the conjugate reflection
// for the LORAKS block. grid: x = tile pair, y = V chunk. 256 threads = 4
// waves (2x2 of 64x64). Thread t stages column t/2 of both tiles, half (t%2)
// of the 32-voxel run, for the next K step in registers while the MFMAs of
// this one run.
__global__ void __launch_bounds__(256) k_gram_bf16x3(const unsigned short* __restrict__ Qa,
                                                     const GramCol* __restrict__ colsA,
                                                     const unsigned short* __restrict__ Qb,
                                                     const GramCol* __restrict__ colsB,
                                                     const int2* __restrict__ tiles,
                                                     float* __restrict__ partial, GramGeom g) {
  // two stages: the MFMAs read stage (step & 1) while the next step's
  // operands are written to the other, one barrier per step
  __shared__ unsigned short sAh[2][GT][GK + GPAD], sAl[2][GT][GK + GPAD];
  __shared__ unsigned short sBh[2][GT][GK + GPAD], sBl[2][GT][GK + GPAD];
  const int2 tp = tiles[blockIdx.x];
  const int I0 = tp.x * GT, J0 = tp.y * GT;
  const long long s0 = (long long)blockIdx.y * g.spw;
  const long long s1 = min(s0 + g.spw, g.nsteps);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  const int cl = tid >> 1, hh = tid & 1;
  const GramCol ca = colsA[I0 + cl], cb = colsB[J0 + cl];
  const long long lo_off = (long long)g.nc * g.plane;
  uint4 ra[4], rb[4];     // hi[0..1], lo[0..1] chunks of 8 bf16
  auto fetch = [&](long long step) {
    const long long row = step / g.nxc;
    const int xc = (int)(step - row * g.nxc);
    const int y = (int)(row % g.N2), z = (int)(row / g.N2);
    const uint4 zero = make_uint4(0, 0, 0, 0);
    if (ca.valid) {
      int ys = y - ca.sy, zs = z - ca.sz;
      ys = ys < 0 ? ys + g.N2 : (ys >= g.N2 ? ys - g.N2 : ys);
      zs = zs < 0 ? zs + g.N3 : (zs >= g.N3 ? zs - g.N3 : zs);
      const uint4* q = reinterpret_cast<const uint4*>(
          Qa + ca.base + ((long long)zs * g.N2 + ys) * g.Lx + xc * 32 + 16 * hh);
      const uint4* ql = q + lo_off / 8;
      ra[0] = q[0]; ra[1] = q[1]; ra[2] = ql[0]; ra[3] = ql[1];
    } else {
      ra[0] = ra[1] = ra[2] = ra[3] = zero;
    }
    if (cb.valid) {
      int ys = y - cb.sy, zs = z - cb.sz;
      ys = ys < 0 ? ys + g.N2 : (ys >= g.N2 ? ys - g.N2 : ys);
      zs = zs < 0 ? zs + g.N3 : (zs >= g.N3 ? zs - g.N3 : zs);
      const uint4* q = reinterpret_cast<const uint4*>(
          Qb + cb.base + ((long long)zs * g.N2 + ys) * g.Lx + xc * 32 + 16 * hh);
      const uint4* ql = q + lo_off / 8;
      rb[0] = q[0]; rb[1] = q[1]; rb[2] = ql[0]; rb[3] = ql[1];
    } else {
      rb[0] = rb[1] = rb[2] = rb[3] = zero;
    }
  };
  auto stage = [&](int b) {
    uint4* ah = reinterpret_cast<uint4*>(&sAh[b][cl][16 * hh]);
    uint4* al = reinterpret_cast<uint4*>(&sAl[b][cl][16 * hh]);
    uint4* bh = reinterpret_cast<uint4*>(&sBh[b][cl][16 * hh]);
    uint4* bl = reinterpret_cast<uint4*>(&sBl[b][cl][16 * hh]);
    ah[0] = ra[0]; ah[1] = ra[1]; al[0] = ra[2]; al[1] = ra[3];
    bh[0] = rb[0]; bh[1] = rb[1]; bl[0] = rb[2]; bl[1] = rb[3];
  };
  if (s0 < s1) {
    fetch(s0);
    stage(0);
  }
  __syncthreads();
  for (long long step = s0; step < s1; ++step) {
    const int cur = (int)((step - s0) & 1);
    const bool more = step + 1 < s1;
    if (more) fetch(step + 1);   // in flight during this step's MFMAs
#pragma unroll
    for (int kk = 0; kk < GK / 16; ++kk) {
      const int koff = kk * 16 + 8 * (lane >> 5);
      bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const int row = wm * 64 + mb * 32 + (lane & 31);
        ah[mb] = *reinterpret_cast<const bf16x8*>(&sAh[cur][row][koff]);
        al[mb] = *reinterpret_cast<const bf16x8*>(&sAl[cur][row][koff]);
      }
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const int col = wn * 64 + nb * 32 + (lane & 31);
        bh[nb] = *reinterpret_cast<const bf16x8*>(&sBh[cur][col][koff]);
        bl[nb] = *reinterpret_cast<const bf16x8*>(&sBl[cur][col][koff]);
      }
#pragma unroll
      for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mb], bh[nb], acc[mb][nb], 0, 0, 0);
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bl[nb], acc[mb][nb], 0, 0, 0);
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mb], bh[nb], acc[mb][nb], 0, 0, 0);
        }
    }
    // the other stage was last read in the previous step, before the barrier
    // that ended it; the barrier below publishes it for the next step and
    // ends this step's reads of the current stage
    if (more) stage(cur ^ 1);
    __syncthreads();
  }
  // partial[(chunk*ntiles + tile)][128][128], row i (X_I column), col j
  float* out = partial + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * (GT * GT);
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = wm * 64 + mb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int j = wn * 64 + nb * 32 + (lane & 31);
        out[i * GT + j] = acc[mb][nb][r];
      }
}

// red[t][e] = sum over the V chunks k of partial[k][t][e] (fixed order, double)
__global__ void k_gram_reduce(const float* __restrict__ partial, int ntiles, int nsplit,
                              double* __restrict__ red) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)GT * GT;
  if (idx >= (long long)ntiles * per) return;
  double s = 0;
  for (int k = 0; k < nsplit; ++k) s += (double)partial[(long long)k * ntiles * per + idx];
  red[idx] = s;
}

// C[r][c] from the reduced tiles; sym: only the upper tiles were computed
// (C symmetric)
__device__ __forceinline__ double cfetch(const double* __restrict__ red, const int* __restrict__ tid_of,
                                         int ntr, int r, int c, bool sym) {
  int tr = r / GT, tc = c / GT, lr = r % GT, lc = c % GT;
  if (sym && tr > tc) { int t = tr; tr = tc; tc = t; t = lr; lr = lc; lc = t; }
  const int t = tid_of[tr * ntr + tc];
  return red[(long long)t * (GT * GT) + lr * GT + lc];
}

// complex n x n (column-major, double) from the real Gram C = X^T Y:
// sum_v conj(x) y = (C11 + C22) + i (C12 - C21)
__global__ void k_gram_assemble(const double* __restrict__ red, const int* __restrict__ tid_of,
                                GramGeom g, int sym, double2* __restrict__ AA) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)g.n * g.n) return;
  const int a = (int)(idx % g.n), b = (int)(idx / g.n);
  const int n = g.n;
  const bool sy = sym != 0;
  const double c11 = cfetch(red, tid_of, g.ntr, a, b, sy);
  const double c22 = cfetch(red, tid_of, g.ntr, n + a, n + b, sy);
  const double c12 = cfetch(red, tid_of, g.ntr, a, n + b, sy);
  const double c21 = cfetch(red, tid_of, g.ntr, n + a, b, sy);
  double2 o;
  o.x = c11 + c22;
  o.y = c12 - c21;
  AA[idx] = o;
}

// g[x,y,z,c] = conj(data[fx(x), fy(y), fz(z), c]), f(i) = (2 c0 - i) mod N per
// axis: the conjugate reflection about the (0-based) centre c0 (sake.m:86-88)
template <typename R>
__global__ void k_flip_conj(const cplx<R>* __restrict__ d, cplx<R>* __restrict__ gout, int N1,
                            int N2, int N3, int nc, int c1, int c2, int c3) {
  const long long V = (long long)N1 * N2 * N3;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * nc) return;
  const long long v = i % V, c = i / V;
  const int x = (int)(v % N1), y = (int)((v / N1) % N2), z = (int)(v / ((long long)N1 * N2));
  auto fl = [](int a, int c0, int n) { int r = (2 * c0 - a) % n; return r < 0 ? r + n : r; };
  const long long src = c * V + fl(x, c1, N1) + (long long)N1 * (fl(y, c2, N2) + (long long)N2 * fl(z, c3, N3));
  cplx<R> o = d[src];
  o.y = -o.y;
  gout[i] = o;
}

// [AA BB; conj(BB) conj(AA)] (2n x 2n, column-major) from AA and BB (n x n)
__global__ void k_loraks_blocks(const double2* __restrict__ AA, const double2* __restrict__ BB,
                                int n, double2* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int a = (int)(idx % n), b = (int)(idx / n);
  const long long m = 2LL * n;
  const double2 aa = AA[idx], bb = BB[idx];
  out[a + m * b] = aa;
  out[a + m * (n + b)] = bb;
  out[(n + a) + m * b] = make_double2(bb.x, -bb.y);
  out[(n + a) + m * (n + b)] = make_double2(aa.x, -aa.y);
}

// fp64 Gram: one thread per (a,b) upper entry, complex double directly
// sum_v conj(x_a[v]) y_b[v] with x from dataA shifted by s_j and y from dataB
// shifted by bsign*s_k; sym: y = x, upper entries computed and mirrored
__global__ void __launch_bounds__(256) k_gram_fp64(const double2* __restrict__ dataA,
                                                   const double2* __restrict__ dataB,
                                                   const int4* __restrict__ shifts, int bsign,
                                                   int sym, GramGeom g, double2* __restrict__ AA) {
  __shared__ double2 sa[16][64], sb[16][64];
  const int ta = blockIdx.x * 16, tb = blockIdx.y * 16;
  if (sym && ta > tb + 15) return;   // strictly below the diagonal band: filled from the upper part
  const int la = threadIdx.x % 16, lb = threadIdx.x / 16;
  const int a = ta + la, b = tb + lb;
  double re = 0, im = 0;
  for (long long v0 = 0; v0 < g.V; v0 += 64) {
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * 16 * 64; e += 256) {
      const int which = e / (16 * 64), rem = e % (16 * 64);
      const int col = rem / 64, vl = rem % 64;
      const long long v = v0 + vl;
      double2 val; val.x = 0; val.y = 0;
      const int aa = (which ? tb : ta) + col;
      if (v < g.V && aa < g.n) {
        const int c = aa % g.nc, j = aa / g.nc;
        const int4 s = shifts[j];
        const int sg = which ? bsign : 1;
        int x = (int)(v % g.N1);
        const long long r = v / g.N1;
        int y = (int)(r % g.N2), z = (int)(r / g.N2);
        x -= sg * s.x; y -= sg * s.y; z -= sg * s.z;
        x = x < 0 ? x + g.N1 : (x >= g.N1 ? x - g.N1 : x);
        y = y < 0 ? y + g.N2 : (y >= g.N2 ? y - g.N2 : y);
        z = z < 0 ? z + g.N3 : (z >= g.N3 ? z - g.N3 : z);
        const double2* src = which ? dataB : dataA;
        val = src[(long long)c * g.V + x + (long long)g.N1 * (y + (long long)g.N2 * z)];
      }
      if (which) sb[col][vl] = val; else sa[col][vl] = val;
    }
    __syncthreads();
    const int nv = (int)min((long long)64, g.V - v0);
    for (int k = 0; k < nv; ++k) {
      const double2 p = sa[la][k], q = sb[lb][k];
      re += p.x * q.x + p.y * q.y;
      im += p.x * q.y - p.y * q.x;
    }
  }
  if (a < g.n && b < g.n && (!sym || a <= b)) {
    double2 o; o.x = re; o.y = im;
    AA[a + (long long)g.n * b] = o;
    if (sym && a != b) { o.y = -im; AA[b + (long long)g.n * a] = o; }
  }
}

// ---------------------------------------------------------------- null operator
// G[d][c][c'] = sum over tap pairs (t, t') with t' - t = d of P[(t,c),(t',c')],
// P = conj(Kt) Kt^T = conj(Q) with Q = Kt Kt^H; d indexes the (2w-1)^3 offset
// cube. Reads Q and conjugates.
__global__ void k_null_G(const double2* __restrict__ Q, const int4* __restrict__ taps, int nk,
                         int nc, int w, double2* __restrict__ G) {
  const int D = 2 * w - 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)D * D * D * nc * nc;
  if (idx >= total) return;
  const int cc = (int)(idx % (nc * nc));
  const int d = (int)(idx / (nc * nc));
  const int c = cc % nc, c2 = cc / nc;
  const int dx = d % D - (w - 1), dy = (d / D) % D - (w - 1), dz = d / (D * D) - (w - 1);
  const int n = nk * nc;
  double re = 0, im = 0;
  for (int t = 0; t < nk; ++t) {
    const int4 a = taps[t];
    for (int t2 = 0; t2 < nk; ++t2) {
      const int4 b = taps[t2];
      if (b.x - a.x != dx || b.y - a.y != dy || b.z - a.z != dz) continue;
      const double2 v = Q[(c + nc * t) + (long long)n * (c2 + nc * t2)];
      re += v.x;
      im -= v.y;
    }
  }
  double2 o; o.x = re; o.y = im;
  G[idx] = o;
}

// One separable stage of B(x) = sum_d G(d) e^{+2 pi i d.x / N}:
// in [outer][D][inner] -> out [outer][Nax][inner] with inner contiguous.
__global__ void k_null_axis(const double2* __restrict__ in, double2* __restrict__ out, int D, int w,
                            int Nax, long long inner, long long outer) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= outer * Nax * inner) return;
  const long long q = idx % inner;
  const int xx = (int)((idx / inner) % Nax);
  const long long o = idx / (inner * Nax);
  double re = 0, im = 0;
  for (int d = 0; d < D; ++d) {
    const int dd = d - (w - 1);
    long long ph = ((long long)dd * xx) % Nax;
    if (ph < 0) ph += Nax;
    const double ang = 2.0 * M_PI * (double)ph / (double)Nax;
    const double cs = cos(ang), sn = sin(ang);
    const double2 v = in[(o * D + d) * inner + q];
    re += v.x * cs - v.y * sn;
    im += v.x * sn + v.y * cs;
  }
  double2 r; r.x = re; r.y = im;
  out[(o * Nax + xx) * inner + q] = r;
}

// B stored per voxel [V][cc] in the plan precision, scaled
template <typename R>
__global__ void k_null_store(const double2* __restrict__ Bd, long long n, double scale,
                             cplx<R>* __restrict__ B) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  // Bd layout after the three stages: [z][y][x][cc] with x fastest of voxels
  const double2 v = Bd[idx];
  cplx<R> o; o.x = (R)(scale * v.x); o.y = (R)(scale * v.y);
  B[idx] = o;
}

// y[c][v] += sum_c' B[v][c][c'] x[c'][v]  (B[v][c + nc*c'])
template <typename R>
__global__ void k_null_apply(const cplx<R>* __restrict__ B, const cplx<R>* __restrict__ x,
                             cplx<R>* __restrict__ y, long long V, int nc) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const cplx<R>* b = B + v * nc * nc;
  for (int c = 0; c < nc; ++c) {
    R re = 0, im = 0;
    for (int c2 = 0; c2 < nc; ++c2) {
      const cplx<R> m = b[c + nc * c2];
      const cplx<R> xv = x[(long long)c2 * V + v];
      re += m.x * xv.x - m.y * xv.y;
      im += m.x * xv.y + m.y * xv.x;
    }
    cplx<R> yv = y[(long long)c * V + v];
    yv.x += re;
    yv.y += im;
    y[(long long)c * V + v] = yv;
  }
}

// ---------------------------------------------------------------- fp64 Gram by FFTs
// The Gram is a set of circular cross-correlations: with F_c = DFT(D_c),
//   R_cd[e] = sum_u conj(D_c[u]) D_d[u + e] = conj(DFT(F_c conj(F_d)))[e] / V,
// AA[(c,j),(d,k)] = R_cd[s_j - s_k]; the LORAKS block uses g = conj(D(flip)):
// BB[(c,j),(d,k)] = R'_cd[s_j + s_k] with D_d replaced by g_d.
// P[q][v] = F_c[v] conj(G_d[v]) for the coil pairs q = (c + nc d) in [q0, q0 + nq)
__global__ void k_pair_prod(const double2* __restrict__ F, const double2* __restrict__ G, long long V,
                            int nc, int q0, int nq, double2* __restrict__ P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * nq) return;
  const int q = q0 + (int)(i / V);
  const long long v = i % V;
  const int c = q % nc, d = q / nc;
  const double2 a = F[(long long)c * V + v], b = G[(long long)d * V + v];
  double2 o;
  o.x = a.x * b.x + a.y * b.y;
  o.y = a.y * b.x - a.x * b.y;
  P[i] = o;
}

// out[(c + nc j) + n (d + nc k)] = conj(Q[q - q0][e]) / V with the lag
// e = s_j - s_k (sgn = -1) or s_j + s_k (sgn = +1), wrapped per axis
__global__ void k_lag_gather(const double2* __restrict__ Q, int N1, int N2, int N3, const int4* __restrict__ taps,
                             int nk, int nc, int q0, int nq, int sgn, double2* __restrict__ out) {
  const long long per = (long long)nk * nk;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * nq) return;
  const int ql = (int)(i / per);
  const int q = q0 + ql;
  const int j = (int)((i % per) % nk), k = (int)((i % per) / nk);
  const int c = q % nc, d = q / nc;
  const int4 a = taps[j], b = taps[k];
  int ex = (a.x + sgn * b.x) % N1, ey = (a.y + sgn * b.y) % N2, ez = (a.z + sgn * b.z) % N3;
  ex += ex < 0 ? N1 : 0;
  ey += ey < 0 ? N2 : 0;
  ez += ez < 0 ? N3 : 0;
  const long long V = (long long)N1 * N2 * N3;
  const double2 r = Q[(long long)ql * V + ex + (long long)N1 * (ey + (long long)N2 * ez)];
  const double iv = 1.0 / (double)V;
  const long long n = (long long)nk * nc;
  double2 o;
  o.x = r.x * iv;
  o.y = -r.y * iv;
  out[(c + (long long)nc * j) + n * (d + (long long)nc * k)] = o;
}

// ---------------------------------------------------------------- host side
static inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Kernel tap offsets of pruno.m:56-71 / sake.m:61-71 in column-major order
void kernel_taps(int width, int radial, bool is2d, std::vector<int4>& taps, int& w) {
  const int h = width / 2;   // fix(width/2)
  w = 2 * h + 1;
  taps.clear();
  for (int iz = 0; iz < (is2d ? 1 : w); ++iz)
    for (int iy = 0; iy < w; ++iy)
      for (int ix = 0; ix < w; ++ix) {
        const int x = ix - h, y = iy - h, z = is2d ? 0 : iz - h;
        bool keep;
        if (radial) keep = std::sqrt((double)(x * x + y * y + z * z)) <= width / 2.0;
        else keep = std::abs(x) <= width / 2.0 && std::abs(y) <= width / 2.0 && std::abs(z) <= width / 2.0;
        if (keep) taps.push_back(make_int4(x, y, z, ix + w * (iy + w * iz)));
      }
}

// AA = Hankel Gram of data [N1 N2 N3 nc] (device, plan precision) -> dAA
// (device complex double, column-major n x n). With loraks, dAA receives
// the 2n x 2n matrix [AA BB; conj(BB) conj(AA)] (sake.m:186-188,210-214,
// 221-224), BB the Gram of the data against its conjugate reflection about
// the 1-based centre[3]: column (d,k) of that operand is
// conj(data_d[flip(v) - s_k]) = g_d[v + s_k], g = conj(data(flip)).
// fp64 Gram of image-size data by FFTs (needs dims = the plan's N and planned
// N-size axes); returns NUFFT3D_ERR_UNSUPPORTED to fall back to k_gram_fp64
static int gram_fft64(nufft3d_plan* p, const void* data, const void* flip, const int dims[4],
                      const int4* dtaps, int nk, double2* outA, double2* outB) {
  if (dims[0] != p->N[0] || dims[1] != p->N[1] || dims[2] != p->N[2]) return NUFFT3D_ERR_UNSUPPORTED;
  const int nc = dims[3];
  const long long V = (long long)dims[0] * dims[1] * dims[2];
  const size_t vb = sizeof(double2) * (size_t)V;
  double2 *F = nullptr, *G = nullptr, *P = nullptr, *Q = nullptr;
  struct Free {
    double2** q[4];
    ~Free() { for (double2** a : q) if (*a) (void)dev_free(*a); }
  } fr{{&F, &G, &P, &Q}};
  const int npair = nc * nc;
  const int chunk = (int)std::max<long long>(1, std::min<long long>(npair, (1LL << 31) / (long long)vb));
  if (dev_malloc(&F, vb * nc) != hipSuccess || (flip && dev_malloc(&G, vb * nc) != hipSuccess) ||
      dev_malloc(&P, vb * chunk) != hipSuccess || dev_malloc(&Q, vb * chunk) != hipSuccess) {
    (void)hipGetLastError();
    set_error("device allocation for the FFT Gram failed");
    return NUFFT3D_ERR_NOMEM;
  }
  N3D_TRY(fft_plain(p, data, nc, F));
  if (flip) N3D_TRY(fft_plain(p, flip, nc, G));
  for (int pass = 0; pass < (flip ? 2 : 1); ++pass) {
    const double2* Gd = pass == 0 ? F : G;
    double2* out = pass == 0 ? outA : outB;
    for (int q0 = 0; q0 < npair; q0 += chunk) {
      const int nq = std::min(chunk, npair - q0);
      ProfScope ps(p, "k_gram_fft64");
      hipLaunchKernelGGL(k_pair_prod, dim3(nblk(V * nq)), dim3(256), 0, p->stream, (const double2*)F, Gd, V,
                         nc, q0, nq, P);
      N3D_TRY(fft_plain(p, P, nq, Q));
      hipLaunchKernelGGL(k_lag_gather, dim3(nblk((long long)nk * nk * nq)), dim3(256), 0, p->stream,
                         (const double2*)Q, dims[0], dims[1], dims[2], dtaps, nk, nc, q0, nq,
                         pass == 0 ? -1 : 1, out);
      N3D_HIP(hipGetLastError());
    }
  }
  N3D_HIP(hipStreamSynchronize(p->stream));
  return NUFFT3D_OK;
}

int gram_device(nufft3d_plan* p, const void* data, const int dims[4], const std::vector<int4>& taps,
                double2* dAA, bool loraks, const int* center) {
  GramGeom g;
  g.N1 = dims[0]; g.N2 = dims[1]; g.N3 = dims[2];
  g.V = (long long)g.N1 * g.N2 * g.N3;
  g.nc = dims[3];
  g.nk = (int)taps.size();
  g.n = g.nc * g.nk;
  g.n2 = 2 * g.n;
  const size_t nn = (size_t)g.n * g.n;
  const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
  int4* dtaps = nullptr;
  void* flip = nullptr;
  double2 *tmpA = nullptr, *tmpB = nullptr;
  N3D_HIP(dev_malloc(&dtaps, sizeof(int4) * taps.size()));
  N3D_HIP(hipMemcpyAsync(dtaps, taps.data(), sizeof(int4) * taps.size(), hipMemcpyHostToDevice, p->stream));
  if (loraks) {
    N3D_HIP(dev_malloc(&flip, esz * g.V * g.nc));
    N3D_HIP(dev_malloc(&tmpA, sizeof(double2) * nn));
    N3D_HIP(dev_malloc(&tmpB, sizeof(double2) * nn));
    const int c1 = center ? center[0] - 1 : 0, c2 = center ? center[1] - 1 : 0,
              c3 = center ? center[2] - 1 : 0;
    ProfScope ps(p, "k_flip_conj");
    if (p->prec == 2)
      hipLaunchKernelGGL((k_flip_conj<double>), dim3(nblk(g.V * g.nc)), dim3(256), 0, p->stream,
                         (const double2*)data, (double2*)flip, g.N1, g.N2, g.N3, g.nc, c1, c2, c3);
    else
      hipLaunchKernelGGL((k_flip_conj<float>), dim3(nblk(g.V * g.nc)), dim3(256), 0, p->stream,
                         (const float2*)data, (float2*)flip, g.N1, g.N2, g.N3, g.nc, c1, c2, c3);
  }
  double2* outA = loraks ? tmpA : dAA;
  int st = NUFFT3D_OK;
  if (p->prec == 2) {
    st = getenv("NUFFT3D_GRAM64_DIRECT") ? NUFFT3D_ERR_UNSUPPORTED
                                         : gram_fft64(p, data, loraks ? flip : nullptr, dims, dtaps,
                                                      g.nk, outA, tmpB);
    if (st == NUFFT3D_ERR_UNSUPPORTED) {   // other sizes: direct double sums
      (void)hipGetLastError();
      st = NUFFT3D_OK;
      dim3 grid((g.n + 15) / 16, (g.n + 15) / 16);
      ProfScope ps(p, "k_gram_fp64");
      hipLaunchKernelGGL(k_gram_fp64, grid, dim3(256), 0, p->stream, (const double2*)data,
                         (const double2*)data, (const int4*)dtaps, 1, 1, g, outA);
      if (loraks)
        hipLaunchKernelGGL(k_gram_fp64, grid, dim3(256), 0, p->stream, (const double2*)data,
                           (const double2*)flip, (const int4*)dtaps, -1, 0, g, tmpB);
    }
  } else {
    g.ntr = (g.n2 + GT - 1) / GT;
    int hmax = 0;
    for (const int4& t : taps) hmax = std::max(hmax, std::max(std::abs(t.x), std::max(std::abs(t.y), std::abs(t.z))));
    g.h = hmax;
    g.Lx = ((g.N1 + 31) / 32) * 32;
    g.nxc = g.Lx / 32;
    g.plane = (long long)g.N3 * g.N2 * g.Lx;
    g.nsteps = (long long)g.N2 * g.N3 * g.nxc;
    const int nsx = 2 * g.h + 1;
    // columns (p, c, j) of an operand: hi-plane offsets of the pre-shifted
    // planes; shift sign -1 for the reflected operand (g_d[v + s_k])
    auto make_cols = [&](int sgn) {
      std::vector<GramCol> cols((size_t)g.ntr * GT);
      for (int col = 0; col < g.ntr * GT; ++col) {
        GramCol& cc = cols[col];
        cc.valid = col < g.n2;
        cc.base = 0; cc.sy = 0; cc.sz = 0; cc.pad = 0;
        if (!cc.valid) continue;
        const int pp = col >= g.n;
        const int a = pp ? col - g.n : col;
        const int c = a % g.nc, j = a / g.nc;
        const int4 t = taps[j];
        const int sxi = sgn * t.x + g.h;
        cc.base = ((long long)(sxi * 4 + 2 * pp) * g.nc + c) * g.plane;
        cc.sy = sgn * t.y;
        cc.sz = sgn * t.z;
      }
      return cols;
    };
    const long long nq = (long long)nsx * 4 * g.nc * g.plane;
    unsigned short *Q = nullptr, *Q2 = nullptr;
    GramCol *dcols = nullptr, *dcols2 = nullptr;
    N3D_HIP(dev_malloc(&Q, sizeof(unsigned short) * nq));
    {
      const std::vector<GramCol> cols = make_cols(1);
      N3D_HIP(dev_malloc(&dcols, sizeof(GramCol) * cols.size()));
      N3D_HIP(hipMemcpy(dcols, cols.data(), sizeof(GramCol) * cols.size(), hipMemcpyHostToDevice));
    }
    {
      ProfScope ps(p, "k_gram_prep");
      hipLaunchKernelGGL(k_gram_prep, dim3(nblk((long long)nsx * g.nc * g.plane)), dim3(256), 0,
                         p->stream, (const float2*)data, Q, g, nsx);
    }
    if (loraks) {
      N3D_HIP(dev_malloc(&Q2, sizeof(unsigned short) * nq));
      const std::vector<GramCol> cols2 = make_cols(-1);
      N3D_HIP(dev_malloc(&dcols2, sizeof(GramCol) * cols2.size()));
      N3D_HIP(hipMemcpy(dcols2, cols2.data(), sizeof(GramCol) * cols2.size(), hipMemcpyHostToDevice));
      ProfScope ps(p, "k_gram_prep");
      hipLaunchKernelGGL(k_gram_prep, dim3(nblk((long long)nsx * g.nc * g.plane)), dim3(256), 0,
                         p->stream, (const float2*)flip, Q2, g, nsx);
    }
    // C = X^T Y over the upper tiles (sym, Y = X) or all tiles, then the
    // complex n x n result
    auto gram_pair = [&](const unsigned short* Qb, const GramCol* colsB, bool sym,
                         double2* out) -> int {
      std::vector<int2> tl;
      std::vector<int> tid_of(g.ntr * g.ntr, -1);
      for (int i = 0; i < g.ntr; ++i)
        for (int j = sym ? i : 0; j < g.ntr; ++j) {
          tid_of[i * g.ntr + j] = (int)tl.size();
          tl.push_back(make_int2(i, j));
        }
      const int ntiles = (int)tl.size();
      GramGeom gg = g;
      // ~24 work-groups per CU (measured at 96^3, n = 1000: 2048 -> 16.0 ms,
      // 4096 -> 15.3, 6144 -> 14.8, 8192 -> 15.1); NUFFT3D_GRAM_BLOCKS overrides
      const char* eb = getenv("NUFFT3D_GRAM_BLOCKS");
      const int target = eb ? std::max(64, atoi(eb)) : 6144;
      int nsplit = (target + ntiles - 1) / ntiles;
      if (nsplit > gg.nsteps) nsplit = (int)gg.nsteps;
      gg.spw = (gg.nsteps + nsplit - 1) / nsplit;
      nsplit = (int)((gg.nsteps + gg.spw - 1) / gg.spw);
      gg.nsplit = nsplit;
      int2* dtl = nullptr;
      int* dtid = nullptr;
      float* part = nullptr;
      double* red = nullptr;
      int rs = NUFFT3D_OK;
      if (dev_malloc(&dtl, sizeof(int2) * ntiles) != hipSuccess ||
          dev_malloc(&dtid, sizeof(int) * tid_of.size()) != hipSuccess ||
          dev_malloc(&part, sizeof(float) * (size_t)nsplit * ntiles * GT * GT) != hipSuccess ||
          dev_malloc(&red, sizeof(double) * (size_t)ntiles * GT * GT) != hipSuccess) {
        set_error("device allocation for the Gram tiles failed");
        rs = NUFFT3D_ERR_NOMEM;
      }
      if (!rs && (hipMemcpyAsync(dtl, tl.data(), sizeof(int2) * ntiles, hipMemcpyHostToDevice,
                                 p->stream) != hipSuccess ||
                  hipMemcpyAsync(dtid, tid_of.data(), sizeof(int) * tid_of.size(),
                                 hipMemcpyHostToDevice, p->stream) != hipSuccess))
        rs = NUFFT3D_ERR_DEVICE;
      if (!rs) {
        {
          ProfScope ps(p, "k_gram_bf16x3");
          hipLaunchKernelGGL(k_gram_bf16x3, dim3(ntiles, nsplit), dim3(256), 0, p->stream,
                             (const unsigned short*)Q, (const GramCol*)dcols, Qb, colsB,
                             (const int2*)dtl, part, gg);
        }
        ProfScope ps(p, "k_gram_assemble");
        hipLaunchKernelGGL(k_gram_reduce, dim3(nblk((long long)ntiles * GT * GT)), dim3(256), 0,
                           p->stream, (const float*)part, ntiles, nsplit, red);
        hipLaunchKernelGGL(k_gram_assemble, dim3(nblk((long long)g.n * g.n)), dim3(256), 0,
                           p->stream, (const double*)red, (const int*)dtid, gg, sym ? 1 : 0, out);
        if (hipStreamSynchronize(p->stream) != hipSuccess) rs = NUFFT3D_ERR_DEVICE;
      }
      if (dtl) (void)dev_free(dtl);
      if (dtid) (void)dev_free(dtid);
      if (part) (void)dev_free(part);
      if (red) (void)dev_free(red);
      return rs;
    };
    st = gram_pair(Q, dcols, true, outA);
    if (!st && loraks) st = gram_pair(Q2, dcols2, false, tmpB);
    (void)dev_free(Q); (void)dev_free(dcols);
    if (Q2) (void)dev_free(Q2);
    if (dcols2) (void)dev_free(dcols2);
  }
  if (!st && loraks) {
    ProfScope ps(p, "k_loraks_blocks");
    hipLaunchKernelGGL(k_loraks_blocks, dim3(nblk((long long)nn)), dim3(256), 0, p->stream,
                       (const double2*)tmpA, (const double2*)tmpB, g.n, dAA);
  }
  N3D_HIP(hipGetLastError());
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(dtaps);
  if (flip) (void)dev_free(flip);
  if (tmpA) (void)dev_free(tmpA);
  if (tmpB) (void)dev_free(tmpB);
  if (st == NUFFT3D_ERR_DEVICE) set_error("device error in the Gram kernel");
  return st;
}

// Hermitian eigendecomposition in FP64 on the device (rocSOLVER zheevd):
// eigenvalues ascending in dW, eigenvectors overwrite dA (column-major)
int heev_device(nufft3d_plan* p, double2* dA, int n, double* dW) {
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    set_error("rocblas_create_handle failed");
    return NUFFT3D_ERR_DEVICE;
  }
  rocblas_set_stream(h, p->stream);
  double* E = nullptr;
  rocblas_int* info = nullptr;
  N3D_HIP(dev_malloc(&E, sizeof(double) * (n > 0 ? n : 1)));
  N3D_HIP(dev_malloc(&info, sizeof(rocblas_int)));
  rocblas_status rs = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, n,
                                       (rocblas_double_complex*)dA, n, dW, E, info);
  rocblas_int hinfo = 0;
  (void)hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, p->stream);
  (void)hipStreamSynchronize(p->stream);
  (void)dev_free(E);
  (void)dev_free(info);
  rocblas_destroy_handle(h);
  if (rs != rocblas_status_success || hinfo != 0) {
    set_error("rocsolver_zheevd failed (status " + std::to_string((int)rs) + ", info " +
              std::to_string(hinfo) + ")");
    return NUFFT3D_ERR_DEVICE;
  }
  return NUFFT3D_OK;
}

// MATLAB hist(y, nbins) counts with equally spaced bins over [min, max]
// MATLAB hist(y, x) with a scalar bin count x: floor(x) bins (0:x stops at
// floor(x)) whose edges are miny + k*(maxy-miny)/x, k = 0..floor(x), the last
// edge replaced by maxy and every edge raised by eps(edge), so a bin is
// (a, b]; values at or below the first edge count in bin 1. A constant y is
// spread over [y - floor(x/2) - 0.5, y + ceil(x/2) - 0.5]. floor(x) < 1
// gives no bins.
static double eps_of(double v) {
  const double a = std::fabs(v);
  return std::nextafter(a, INFINITY) - a;
}

static std::vector<int> hist_counts(const std::vector<double>& y, double x) {
  const int n = x >= 1 ? (int)std::floor(x) : 0;
  std::vector<int> h(n, 0);
  if (n < 1 || y.empty()) return h;
  double lo = *std::min_element(y.begin(), y.end()), hi = *std::max_element(y.begin(), y.end());
  if (lo == hi) {
    lo = lo - std::floor(x / 2) - 0.5;
    hi = hi + std::ceil(x / 2) - 0.5;
  }
  const double bw = (hi - lo) / x;
  std::vector<double> e(n + 1);
  for (int k = 0; k <= n; ++k) e[k] = lo + bw * k;
  e[n] = hi;
  for (double& v : e) v += eps_of(v);
  for (double v : y) {
    if (!(v <= e[n])) continue;   // above the last edge (or NaN): not counted
    const int c = (int)(std::upper_bound(e.begin(), e.end(), v) - e.begin());   // edges <= v
    h[std::min(std::max(c, 1), n) - 1]++;
  }
  return h;
}

static double median_of(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t m = v.size();
  if (m == 0) return 0;
  return m % 2 ? v[m / 2] : 0.5 * (v[m / 2 - 1] + v[m / 2]);
}

// pruno.m:93-138 on the host: eigen-pairs (descending), noise, filter, nv and
// the null kernels Kt[(t,c), v] = (1 - f_v) V[c + nc*t, v]
int pruno_select(const std::vector<double>& evals_desc, double noise_in, double density,
                 int nx, int ny, double tol, double& noise_out, std::vector<double>& f, int& nv) {
  const int n = (int)evals_desc.size();
  double noise = noise_in;
  if (!(noise >= 0)) {   // estimate (pruno.m:97-107)
    const double W1 = evals_desc[0];
    const double thr = std::nextafter(n * W1, INFINITY) - n * W1;   // eps(numel(W)*W(1))
    int hi = 0;
    for (double v : evals_desc) hi += v > thr;
    int lo = 1;
    for (lo = 1; lo <= hi; ++lo) {
      std::vector<double> seg(evals_desc.begin() + (lo - 1), evals_desc.begin() + hi);
      std::vector<int> h = hist_counts(seg, std::sqrt((double)(hi - lo)));
      // [~,k] = max(h): the first maximum; no bins leaves k empty (no break)
      const int k = h.empty() ? 0 : (int)(std::max_element(h.begin(), h.end()) - h.begin()) + 1;
      if (k > 1) break;
    }
    if (lo > hi) lo = hi;   // the for loop's variable keeps its last value
    double sigma = 0;       // hi = 0: every eigenvalue is zero
    if (hi > 0) {
      std::vector<double> seg(evals_desc.begin() + (lo - 1), evals_desc.begin() + hi);
      sigma = std::sqrt(median_of(seg));
    }
    noise = sigma / std::sqrt(density * nx * ny);
  }
  noise_out = noise;
  const double sigma = noise * std::sqrt(density * nx * ny);     // pruno.m:108
  f.assign(n, 0.0);
  nv = 0;
  for (int i = 0; i < n; ++i) {
    const double Wv = std::sqrt(std::max(evals_desc[i], 0.0));   // pruno.m:111
    const double fv = Wv > 0 ? std::max(0.0, 1 - sigma * sigma / (Wv * Wv)) : 0.0;
    f[i] = fv;                                                   // pruno.m:114
    nv += fv < 1 - tol;                                          // pruno.m:115
  }
  return NUFFT3D_OK;
}

// Per-voxel coupling operator B(x) = lambda^2 sum_v conj(fnull_v) fnull_v^T
// from the tap-domain null kernels Kt (device complex double [(t,c) x nv]).
int null_operator(nufft3d_plan* p, const double2* dKt, int nk, int nc, int nv, int w,
                  const std::vector<int4>& taps, const int N[3], double lambda, void* dB) {
  const int n = nk * nc;
  double2 *P = nullptr, *G = nullptr;
  int4* dtaps = nullptr;
  N3D_HIP(dev_malloc(&P, sizeof(double2) * (size_t)n * n));
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    set_error("rocblas_create_handle failed");
    return NUFFT3D_ERR_DEVICE;
  }
  rocblas_set_stream(h, p->stream);
  // Q = Kt Kt^H (n x n); P = conj(Kt) Kt^T = conj(Q)
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  rocblas_status rs = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose,
                                    n, n, nv, &one, (const rocblas_double_complex*)dKt, n,
                                    (const rocblas_double_complex*)dKt, n, &zero,
                                    (rocblas_double_complex*)P, n);
  rocblas_destroy_handle(h);
  if (rs != rocblas_status_success) {
    (void)dev_free(P);
    set_error("rocblas_zgemm failed");
    return NUFFT3D_ERR_DEVICE;
  }
  const int D = 2 * w - 1;
  const long long nG = (long long)D * D * D * nc * nc;
  N3D_HIP(dev_malloc(&G, sizeof(double2) * nG));
  N3D_HIP(dev_malloc(&dtaps, sizeof(int4) * taps.size()));
  N3D_HIP(hipMemcpyAsync(dtaps, taps.data(), sizeof(int4) * taps.size(), hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_null_G, dim3(nblk(nG)), dim3(256), 0, p->stream, (const double2*)P,
                     (const int4*)dtaps, nk, nc, w, G);
  const int st = lag_field(p, G, w, (long long)nc * nc, N, lambda * lambda, dB);
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(P); (void)dev_free(G); (void)dev_free(dtaps);
  return st;
}

// Separable inverse DFT of lag-domain matrices: G [dz][dy][dx][cc] (offsets
// d - (w-1), D = 2w-1 per axis) -> B [z][y][x][cc] = scale * sum_d G(d)
// e^{+2 pi i d.x/N}, stored in the plan precision. Temporaries in double.
int lag_field(nufft3d_plan* p, const double2* G, int w, long long cc, const int N[3], double scale,
              void* dB) {
  const int D = 2 * w - 1;
  const int N1 = N[0], N2 = N[1], N3 = N[2];
  double2 *T1 = nullptr, *T2 = nullptr, *T3 = nullptr;
  N3D_HIP(dev_malloc(&T1, sizeof(double2) * (size_t)D * D * N1 * cc));
  N3D_HIP(dev_malloc(&T2, sizeof(double2) * (size_t)D * N2 * N1 * cc));
  N3D_HIP(dev_malloc(&T3, sizeof(double2) * (size_t)N3 * N2 * N1 * cc));
  // x: in [dz*dy][dx][cc] -> [dz*dy][x][cc]
  hipLaunchKernelGGL(k_null_axis, dim3(nblk((long long)D * D * N1 * cc)), dim3(256), 0, p->stream,
                     G, T1, D, w, N1, cc, (long long)D * D);
  // y: in [dz][dy][x*cc] -> [dz][y][x*cc]
  hipLaunchKernelGGL(k_null_axis, dim3(nblk((long long)D * N2 * N1 * cc)), dim3(256), 0, p->stream,
                     (const double2*)T1, T2, D, w, N2, (long long)N1 * cc, (long long)D);
  // z: in [dz][y*x*cc] -> [z][y*x*cc]
  hipLaunchKernelGGL(k_null_axis, dim3(nblk((long long)N3 * N2 * N1 * cc)), dim3(256), 0, p->stream,
                     (const double2*)T2, T3, D, w, N3, (long long)N2 * N1 * cc, 1LL);
  const long long n = (long long)N1 * N2 * N3 * cc;
  if (p->prec == 2)
    hipLaunchKernelGGL((k_null_store<double>), dim3(nblk(n)), dim3(256), 0, p->stream,
                       (const double2*)T3, n, scale, (double2*)dB);
  else
    hipLaunchKernelGGL((k_null_store<float>), dim3(nblk(n)), dim3(256), 0, p->stream,
                       (const double2*)T3, n, scale, (float2*)dB);
  N3D_HIP(hipGetLastError());
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(T1); (void)dev_free(T2); (void)dev_free(T3);
  return NUFFT3D_OK;
}

int null_apply(nufft3d_plan* p, const void* B, const void* x, void* y, int nc) {
  const long long V = (long long)p->N[0] * p->N[1] * p->N[2];
  ProfScope ps(p, "k_null_apply");
  if (p->prec == 2)
    hipLaunchKernelGGL((k_null_apply<double>), dim3(nblk(V)), dim3(256), 0, p->stream,
                       (const double2*)B, (const double2*)x, (double2*)y, V, nc);
  else
    hipLaunchKernelGGL((k_null_apply<float>), dim3(nblk(V)), dim3(256), 0, p->stream,
                       (const float2*)B, (const float2*)x, (float2*)y, V, nc);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

template <typename R>
__global__ void k_count_nonzero(const cplx<R>* __restrict__ d, long long n, unsigned long long* cnt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long mine = 0;
  if (i < n) { const cplx<R> v = d[i]; mine = (v.x != 0 || v.y != 0); }
  // wave-level ballot then one atomic per wave (integer: order-independent)
  const unsigned long long b = __ballot(mine);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(cnt, (unsigned long long)__popcll(b));
}

// Kt[:, v] = scale_v * E[:, col_v] (eigenvector columns of the ascending
// decomposition, complex double, column-major n x n)
__global__ void k_pick_columns(const double2* __restrict__ E, int n, const int* __restrict__ cols,
                               const double* __restrict__ scale, int nv, double2* __restrict__ Kt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * nv) return;
  const int r = (int)(i % n), v = (int)(i / n);
  const double2 e = E[r + (long long)n * cols[v]];
  double2 o; o.x = scale[v] * e.x; o.y = scale[v] * e.y;
  Kt[i] = o;
}

// pruno.m:43-143 for k-space data [N1 N2 N3 nc] on the device: Gram, FP64
// eigendecomposition, noise floor, filter, nv, null kernels; then the
// per-voxel coupling operator of iprojection (nufft_3d.m:317-320) with
// fnull = K*lambda (iNUFT.m:207-208). dB receives [V][nc][nc].
int pruno_build(nufft3d_plan* p, const void* kdata, int nc, const nufft3d_pruno_opts& o,
                double lambda, void** dB, int* nv_out, std::vector<double>* sv_desc,
                double* noise_out, void* kt_host) {
  const int dims[4] = {p->N[0], p->N[1], p->N[2], nc};
  const long long V = (long long)dims[0] * dims[1] * dims[2];
  // fully sampled check (pruno.m:77-81)
  unsigned long long* dcnt = nullptr;
  N3D_HIP(dev_malloc(&dcnt, sizeof(unsigned long long)));
  N3D_HIP(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), p->stream));
  if (p->prec == 2)
    hipLaunchKernelGGL((k_count_nonzero<double>), dim3(nblk(V * nc)), dim3(256), 0, p->stream,
                       (const double2*)kdata, V * nc, dcnt);
  else
    hipLaunchKernelGGL((k_count_nonzero<float>), dim3(nblk(V * nc)), dim3(256), 0, p->stream,
                       (const float2*)kdata, V * nc, dcnt);
  unsigned long long cnt = 0;
  N3D_HIP(hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, p->stream));
  N3D_HIP(hipStreamSynchronize(p->stream));
  (void)dev_free(dcnt);
  const double density = (double)cnt / (double)(V * nc);
  if (density < 0.99) { set_error("data must be fully sampled"); return NUFFT3D_ERR_INVALID; }

  std::vector<int4> taps;
  int w = 0;
  kernel_taps(o.width, o.radial, dims[2] == 1, taps, w);
  const int nk = (int)taps.size();
  const int n = nk * nc;
  double2* dAA = nullptr;
  double* dW = nullptr;
  N3D_HIP(dev_malloc(&dAA, sizeof(double2) * (size_t)n * n));
  N3D_HIP(dev_malloc(&dW, sizeof(double) * n));
  int st = gram_device(p, kdata, dims, taps, dAA);
  if (!st) st = heev_device(p, dAA, n, dW);
  std::vector<double> wasc(n);
  if (!st) {
    N3D_HIP(hipMemcpy(wasc.data(), dW, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  std::vector<double> wdesc(wasc.rbegin(), wasc.rend());
  std::vector<double> f;
  int nv = 0;
  double noise = 0;
  if (!st) st = pruno_select(wdesc, o.noise, density, dims[0], dims[1], o.tol, noise, f, nv);
  if (sv_desc) {
    sv_desc->resize(n);
    for (int i = 0; i < n; ++i) (*sv_desc)[i] = std::sqrt(std::max(wdesc[i], 0.0));
  }
  if (noise_out) *noise_out = noise;
  if (nv_out) *nv_out = nv;
  if (!st && (dB || kt_host)) {
    // the nv vectors: literal = first nv (largest, pruno.m:128-130),
    // otherwise the nv smallest (those with f < 1 - tol)
    std::vector<int> cols(std::max(nv, 1));
    std::vector<double> scale(std::max(nv, 1));
    for (int v = 0; v < nv; ++v) {
      const int desc = o.literal_first_nv ? v : n - nv + v;
      cols[v] = n - 1 - desc;
      scale[v] = 1.0 - f[desc];
    }
    int* dcols = nullptr;
    double* dscale = nullptr;
    double2* dKt = nullptr;
    N3D_HIP(dev_malloc(&dcols, sizeof(int) * cols.size()));
    N3D_HIP(dev_malloc(&dscale, sizeof(double) * scale.size()));
    N3D_HIP(dev_malloc(&dKt, sizeof(double2) * (size_t)n * std::max(nv, 1)));
    N3D_HIP(hipMemcpy(dcols, cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice));
    N3D_HIP(hipMemcpy(dscale, scale.data(), sizeof(double) * scale.size(), hipMemcpyHostToDevice));
    if (nv > 0)
      hipLaunchKernelGGL(k_pick_columns, dim3(nblk((long long)n * nv)), dim3(256), 0, p->stream,
                         (const double2*)dAA, n, (const int*)dcols, (const double*)dscale, nv, dKt);
    if (kt_host && nv > 0)
      N3D_HIP(hipMemcpyAsync(kt_host, dKt, sizeof(double2) * (size_t)n * nv, hipMemcpyDeviceToHost,
                             p->stream));
    if (dB) {
      const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
      N3D_HIP(dev_malloc(dB, esz * (size_t)V * nc * nc));
      if (nv > 0) st = null_operator(p, dKt, nk, nc, nv, w, taps, p->N, lambda, *dB);
      else N3D_HIP(hipMemset(*dB, 0, esz * (size_t)V * nc * nc));
    }
    N3D_HIP(hipStreamSynchronize(p->stream));
    (void)dev_free(dcols); (void)dev_free(dscale); (void)dev_free(dKt);
  }
  (void)dev_free(dAA);
  (void)dev_free(dW);
  return st;
}

}  // namespace nufft3d
