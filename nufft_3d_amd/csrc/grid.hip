// Gridding kernels: forward interpolation (H*g) and deterministic adjoint
// spread (H'*k), both with Kaiser-Bessel weights computed on the fly from
// the bin-sorted samples instead of a stored sparse H.
//
// Reference: spmv / spmv_t (nufft_3d.m:259-282) over the H assembled at
// nufft_3d.m:153-212. Candidate neighbours of a sample are
// round(u*om)+d, d in -ceil(J/2)..ceil(J/2) per axis (nufft_3d.m:154,168-170),
// wrapped mod K (nufft_3d.m:178-180), kept when inside the radial or
// separable support with <= (nufft_3d.m:182-191).
#include <algorithm>
#include <type_traits>

#include "kb.h"
#include "nufft3d_internal.h"

namespace nufft3d {

struct GridGeom {
  int flags;  // bit 1: brick-order spread
  int xchunk; // XCD block renumbering run length (0 = off)
  int K1, K2, K3;
  int W1, W2, W3;
  int nc;     // coils in the grid / sample arrays
  int c0;     // first coil of this launch's coil group
  int cg;     // coils in this launch's group
  int nnz;
  long long M;
};

// Work-groups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2. With a run length C > 0, blocks are
// renumbered so that XCD x receives runs of C consecutive logical blocks
// (runs x, x+8, x+16, ...): the blocks an XCD holds at once cover one compact
// region of samples / grid points, which share grid rows and samples in its
// L2, while every XCD still gets a share of the dense k-space centre. Blocks
// past the last whole round of 8C keep their number. C < 0: C = grid / 8,
// each XCD walks one contiguous eighth of the (Z-ordered) work.
__device__ __forceinline__ long long xcd_block_of(int chunk, unsigned b, unsigned nb) {
  if (chunk < 0) chunk = (int)(nb / 8);          // one contiguous eighth per XCD
  if (chunk <= 0) return b;
  const unsigned span = 8u * (unsigned)chunk;
  if (b >= nb / span * span) return b;
  const unsigned x = b % 8, k = b / 8;
  return (long long)((k / chunk) * 8 + x) * chunk + k % chunk;
}

__device__ __forceinline__ long long xcd_block(int chunk) {
  return xcd_block_of(chunk, blockIdx.x, gridDim.x);
}

// a base in [0, K) plus an offset of at most 2K, wrapped into [0, K)
__device__ __forceinline__ int wrapk(int g, int K) {
  g = g < 0 ? g + K : (g >= K ? g - K : g);
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

// Weight of a hit whose support membership was decided at plan time in
// double (nufft_3d.m:182-191). fp64: the reference value rounded to float
// and widened (nufft_3d.m:160,205); fp32: evaluated in float.
__device__ __forceinline__ void decode_hit(unsigned short code, const GridGeom& g, int& dx, int& dy,
                                           int& dz) {
  dx = (int)(code & 31u) - g.W1;
  dy = (int)((code >> 5) & 31u) - g.W2;
  dz = (int)(code >> 10) - g.W3;
}

__device__ __forceinline__ double hit_weight_d(int dx, int dy, int dz, const double* f,
                                               const KBParams& p) {
#pragma clang fp contract(off)
  const double ex = (double)dx - f[0], ey = (double)dy - f[1], ez = (double)dz - f[2];
  if (p.radial) return (double)(float)kb_d(ex * ex + ey * ey + ez * ez, p);
  return (double)(float)(kb_d(ex * ex, p) * kb_d(ey * ey, p) * kb_d(ez * ez, p));
}

// fp64 arithmetic behind fp32 arrays: the tabulated kernel, rounded to float
__device__ __forceinline__ double hit_weight_tab(int dx, int dy, int dz, double fx, double fy,
                                                 double fz, const double* tab, const KBParams& p) {
  const double ex = (double)dx - fx, ey = (double)dy - fy, ez = (double)dz - fz;
  if (p.radial) return (double)(float)kb_tab_d(ex * ex + ey * ey + ez * ez, tab, p);
  return (double)(float)(kb_tab_d(ex * ex, tab, p) * kb_tab_d(ey * ey, tab, p) *
                         kb_tab_d(ez * ez, tab, p));
}

__device__ __forceinline__ float hit_weight_f(int dx, int dy, int dz, const float4 f,
                                              const KBParams& p) {
  const float ex = (float)dx - f.x, ey = (float)dy - f.y, ez = (float)dz - f.z;
  if (p.radial) return kb_f(ex * ex + ey * ey + ez * ez, p);
  return kb_f(ex * ex, p) * kb_f(ey * ey, p) * kb_f(ez * ez, p);
}

// Stage the kernel table (kb_tab_d) in LDS; every thread of the work-group
// must call it (it synchronises). tab == nullptr: nothing to stage.
__device__ __forceinline__ const double* stage_kb_table(const double* __restrict__ tab, double* lds) {
  if (!tab) return nullptr;
  for (int i = threadIdx.x; i < kKbIntervals * kKbCoef; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  return lds;
}

// accumulator type: fp64 plans and compensated fp32 plans accumulate in
// double (an fp32 x fp32 product is exact in double, so the sum carries one
// rounding per term at 2^-53 instead of 2^-24); plain fp32 plans in float
template <typename R, bool COMP> struct AccT { using type = double; };
template <> struct AccT<float, false> { using type = float; };

// ---------------------------------------------------------------------------
// Lane groups. LPS consecutive lanes (1, 2 or 4; always inside one DPP quad)
// share a sample (interp) or grid point (spread), each owning CPL coils.
// They walk the hit list LPS hits at a time: lane q evaluates the weight and
// address of hit e+q only, and quad-permute broadcasts hand every lane of the
// group all LPS weights and addresses. Each lane then accumulates its coils
// over the hits in list order, so a coil's sum is the same whatever LPS,
// CPL or the batch around it. All lanes of a group take the same branches
// (same hit list), so a broadcast never reads an inactive lane.
template <int LPS, int J>
__device__ __forceinline__ int grp_bcast_i(int v) {
  if constexpr (LPS == 1) {
    return v;
  } else if constexpr (LPS == 2) {
    return __builtin_amdgcn_mov_dpp(v, J | (J << 2) | ((J + 2) << 4) | ((J + 2) << 6), 0xF, 0xF,
                                    false);
  } else {
    static_assert(LPS == 4, "lane groups are 1, 2 or 4 lanes");
    return __builtin_amdgcn_mov_dpp(v, J * 0x55, 0xF, 0xF, false);
  }
}

template <int LPS, int J, typename T>
__device__ __forceinline__ T grp_bcast(T v) {
  if constexpr (LPS == 1) {
    return v;
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, grp_bcast_i<LPS, J>(__builtin_bit_cast(int, v)));
  } else {
    static_assert(sizeof(T) == 8, "4- or 8-byte values");
    int2 h = __builtin_bit_cast(int2, v);
    h.x = grp_bcast_i<LPS, J>(h.x);
    h.y = grp_bcast_i<LPS, J>(h.y);
    return __builtin_bit_cast(T, h);
  }
}

// Coils of a lane. Lane q of an LPS-lane group owns CPL of the group's coils:
// with FULL groups of float data and an even CPL, coil pairs are dealt
// round-robin over the lanes (lane q: pairs q, q+LPS, ...) so that one
// 16-byte load per lane makes the group read one contiguous run of 16*LPS
// bytes of a row; otherwise lane q owns coils q*CPL .. q*CPL+CPL-1.
template <typename R, int CPL, int LPS, bool FULL>
__device__ __forceinline__ int coil_of(int q, int c) {
  if constexpr (FULL && sizeof(R) == 4 && CPL % 2 == 0) return ((c / 2) * LPS + q) * 2 + (c & 1);
  else return q * CPL + c;
}

// val[c] = row[coil_of(q, c)] (load_coils), then ar/ai[c] += w * val[c]
// (fma_coils). FULL: the group's coils fill every lane and the row start is
// 16-byte aligned (even coil count and offset), so the coils come in as
// 16-byte loads. Otherwise coil indices past the group are clamped to its
// last coil (the row stays in bounds; those accumulators are never stored).
template <typename R, int CPL, int LPS, bool FULL>
__device__ __forceinline__ void load_coils(cplx<R>* val, const cplx<R>* __restrict__ row, int q,
                                           int cg) {
  if constexpr (FULL && sizeof(R) == 4 && CPL % 2 == 0) {
    const float4* p4 = reinterpret_cast<const float4*>(row);
#pragma unroll
    for (int c = 0; c < CPL / 2; ++c) {
      const float4 t = p4[c * LPS + q];
      val[2 * c].x = t.x; val[2 * c].y = t.y;
      val[2 * c + 1].x = t.z; val[2 * c + 1].y = t.w;
    }
  } else if constexpr (FULL) {
#pragma unroll
    for (int c = 0; c < CPL; ++c) val[c] = row[q * CPL + c];
  } else {
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = q * CPL + c;
      val[c] = row[k < cg ? k : cg - 1];
    }
  }
}

template <typename R, int CPL, typename A>
__device__ __forceinline__ void fma_coils(A* ar, A* ai, R w, const cplx<R>* val) {
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    ar[c] += (A)w * (A)val[c].x;
    ai[c] += (A)w * (A)val[c].y;
  }
}

// ---------------------------------------------------------------------------
// Forward: k[s,c] = scale * w_s * sum_{hits g of s} H(s,g) G[g,c]. Hits come
// in z-y-x candidate order. The result goes to the sorted coil-interleaved
// layout out[s*nc + c], or with USER straight to the caller's [M nc] layout
// out[orig[s] + c*M] (no separate permutation pass).
template <typename R, int CPL, int LPS, bool COMP, bool FULL, bool USER, typename RO = R>
__global__ void __launch_bounds__(256) k_interp(const long long* __restrict__ spos,
                                                const unsigned short* __restrict__ fcode,
                                                const int4* __restrict__ base,
                                                const double* __restrict__ off,
                                                const float4* __restrict__ off32,
                                                const cplx<R>* __restrict__ grid,
                                                cplx<RO>* __restrict__ out,
                                                const R* __restrict__ wsamp, R scale,
                                                const int* __restrict__ orig,
                                                const double* __restrict__ kbtab,
                                                GridGeom g, KBParams kp) {
  using A = typename AccT<R, COMP>::type;
  __shared__ double stab[sizeof(R) == 8 ? kKbIntervals * kKbCoef : 1];
  const double* tab = nullptr;
  if constexpr (sizeof(R) == 8) tab = stage_kb_table(kbtab, stab);
  const long long gt = xcd_block(g.xchunk) * blockDim.x + threadIdx.x;
  const int s = (int)(gt / LPS);
  const int q = (int)(gt % LPS);
  if (s >= g.nnz) return;
  const int4 b = base[s];
  float4 f32;
  if constexpr (sizeof(R) == 4) f32 = off32[s];
  const double* f64 = off + 3LL * s;
  A ar[CPL], ai[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) { ar[c] = 0; ai[c] = 0; }
  const long long e0 = spos[s], e1 = spos[s + 1];
  // lane q takes hit e+q of each step; past the end of the list it repeats
  // the step's first hit with weight 0 (a point the sample does touch).
  // Codes are loaded one step ahead.
  // (the prefetch is issued on every path, index clamped into the list, so
  // the wait counts do not depend on the branch taken)
  unsigned short code = e0 < e1 ? fcode[e0 + q < e1 ? e0 + q : e0] : 0;
  for (long long e = e0; e < e1; e += LPS) {
    const unsigned short cur = code;
    {
      const long long en = e + LPS < e1 ? e + LPS : e;
      code = fcode[en + q < e1 ? en + q : en];
    }
    int dx, dy, dz;
    decode_hit(cur, g, dx, dy, dz);
    R w;
    if constexpr (sizeof(R) == 8)
      w = tab ? hit_weight_tab(dx, dy, dz, f64[0], f64[1], f64[2], tab, kp)
              : hit_weight_d(dx, dy, dz, f64, kp);
    else
      w = hit_weight_f(dx, dy, dz, f32, kp);
    if (e + q >= e1) w = 0;
    const int gx = wrapk(b.x + dx, g.K1), gy = wrapk(b.y + dy, g.K2), gz = wrapk(b.z + dz, g.K3);
    const int gi = (gz * g.K2 + gy) * g.K1 + gx;
    R wj[LPS];
    cplx<R> val[LPS][CPL];
#define N3D_STEP(J)                                                                   \
    if constexpr (J < LPS) {                                                          \
      wj[J] = grp_bcast<LPS, J>(w);                                                   \
      const int gj = grp_bcast_i<LPS, J>(gi);                                         \
      load_coils<R, CPL, LPS, FULL>(val[J], grid + (long long)gj * g.nc + g.c0, q, g.cg); \
    }
    N3D_STEP(0) N3D_STEP(1) N3D_STEP(2) N3D_STEP(3)
#pragma unroll
    for (int j = 0; j < LPS; ++j) fma_coils<R, CPL>(ar, ai, wj[j], val[j]);
  }
  R sc = scale;
  if (wsamp) sc *= wsamp[s];
  if constexpr (USER) {
    cplx<RO>* op = out + orig[s] + (long long)g.c0 * g.M;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = coil_of<R, CPL, LPS, FULL>(q, c);
      if (k < g.cg) {
        cplx<RO> v; v.x = (RO)((R)ar[c] * sc); v.y = (RO)((R)ai[c] * sc);
        op[(long long)k * g.M] = v;
      }
    }
  } else {
    cplx<RO>* op = out + (long long)s * g.nc + g.c0;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int k = coil_of<R, CPL, LPS, FULL>(q, c);
      if (k < g.cg) {
        cplx<RO> v; v.x = (RO)((R)ar[c] * sc); v.y = (RO)((R)ai[c] * sc);
        op[k] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward from LDS tiles (3-D plans with a 7-point footprint, ceil(J/2) = 3).
// Samples are sorted by 4x4x4 sub-bins of their rounded position; a
// work-group takes one piece (up to 128 samples of one sub-bin, one per
// 4-lane group) whose hits all lie in the sub-bin's 10x10x10 grid
// neighbourhood (wrapped). That neighbourhood is staged in LDS NL z-layers
// at a time (coalesced row reads), and each group walks its sample's hit
// list through the passes: the codes are in z-y-x order, so the hits of a
// pass are a run of the list, and a pass boundary can fall inside a step of
// LPS hits without changing the order of the sums. Same hits, weights and
// summation order as k_interp; a few KB of LDS per work-group, so several
// work-groups share a CU and one's staging overlaps another's arithmetic.
constexpr int kTileEdge = 10;
constexpr int kTileThreads = 512;

// Lane q of a 4-lane group owns CPL coils of the group: UPL 16-byte units of
// UB coils each. Every staged grid point holds CG*sizeof(C) = 128 bytes and so
// starts on the same LDS bank; with two units per lane, odd groups own their
// units in the opposite order (unit_of), so in each read instruction the 8
// lanes of two neighbouring groups read 8 different units - all 32 banks -
// whatever points they read.
template <typename R, int CPL>
struct TileLane {
  static constexpr int UB = 16 / (int)sizeof(cplx<R>);   // coils per 16-byte unit
  static constexpr int UPL = CPL / UB > 0 ? CPL / UB : 1;
  // the coil (within the group) of the lane's c-th accumulator
  static __device__ __forceinline__ int coil(int q, int c, int par) {
    if constexpr (UPL == 2) return q * CPL + ((c / UB) ^ par) * UB + c % UB;
    else return q * CPL + c;
  }
};

template <typename R, int CG, int LPS, bool COMP, bool USER, bool FASTW>
__global__ void __launch_bounds__(kTileThreads) k_interp_tile(
    const int4* __restrict__ pieces, const long long* __restrict__ spos,
    const unsigned short* __restrict__ fcode, const int4* __restrict__ base,
    const double* __restrict__ off, const float4* __restrict__ off32,
    const cplx<R>* __restrict__ grid, cplx<R>* __restrict__ out, const R* __restrict__ wsamp,
    R scale, const int* __restrict__ orig, const double* __restrict__ kbtab, GridGeom g,
    KBParams kp, int NL) {
  using A = typename AccT<R, COMP>::type;
  using C = cplx<R>;
  using TL = TileLane<R, CG / LPS>;
  constexpr int CPL = CG / LPS;
  constexpr int P2 = kTileEdge * kTileEdge;   // points of one z-layer
  constexpr int TAB = FASTW ? kKbIntervals * kKbCoef : 0;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* stab = reinterpret_cast<double*>(smem_raw);
  C* tile = reinterpret_cast<C*>(smem_raw + TAB * sizeof(double));
  if constexpr (FASTW)
    for (int i = threadIdx.x; i < TAB; i += kTileThreads) stab[i] = kbtab[i];
  const int4 pc = pieces[xcd_block(g.xchunk)];
  const int x0 = (pc.x & 1023) * 4 - 3, y0 = ((pc.x >> 10) & 1023) * 4 - 3, z0 = (pc.x >> 20) * 4 - 3;
  const int q = threadIdx.x % LPS;
  const int lane = threadIdx.x & 63;
  const int gsh = lane & ~(LPS - 1);
  const int par = (lane / LPS) & 1;
  const int j = threadIdx.x / LPS;
  const bool act = j < pc.z;
  const int s = pc.y + (act ? j : 0);
  const int4 b = base[s];
  const int lbz = b.z - z0;
  const int lb = (lbz * kTileEdge + (b.y - y0)) * kTileEdge + (b.x - x0);
  float4 f32;
  if constexpr (sizeof(R) == 4) f32 = off32[s];
  const double fx = off[3LL * s], fy = off[3LL * s + 1], fz = off[3LL * s + 2];
  A ar[CPL], ai[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) { ar[c] = 0; ai[c] = 0; }
  const long long e0 = spos[s];
  const unsigned short* codes = fcode + e0;          // this sample's hit codes
  const int nh = act ? (int)(spos[s + 1] - e0) : 0;  // >= 1 for an in-bounds sample
  const int hl = nh > 0 ? nh - 1 : 0;
  int e = 0;
  unsigned short cur = codes[min(e + q, hl)];
  for (int zl = 0; zl < kTileEdge; zl += NL) {
    const int nl = kTileEdge - zl < NL ? kTileEdge - zl : NL;
    __syncthreads();   // the previous layers are consumed (and the table staged)
    for (int i = threadIdx.x; i < nl * P2 * CG; i += kTileThreads) {
      const int pt = i / CG, c = i - pt * CG;
      const int lx = pt % kTileEdge, lyz = pt / kTileEdge;
      const int ly = lyz % kTileEdge, lz = zl + lyz / kTileEdge;
      const int gx = wrapk(x0 + lx, g.K1), gy = wrapk(y0 + ly, g.K2), gz = wrapk(z0 + lz, g.K3);
      C v;
      v.x = 0;
      v.y = 0;
      if (c < g.cg) v = grid[(((long long)gz * g.K2 + gy) * g.K1 + gx) * g.nc + g.c0 + c];
      tile[pt * CG + c] = v;
    }
    __syncthreads();
    const int zh = zl + nl;
    while (e < nh) {
      const unsigned short nxt = codes[min(e + LPS + q, hl)];
      int dx, dy, dz;
      decode_hit(cur, g, dx, dy, dz);
      // hits of this pass: a run at the head of the step (dz never decreases)
      const bool mine = e + q < nh && lbz + dz < zh;
      const unsigned long long bal = __ballot(mine);
      const int n = __popcll((bal >> gsh) & ((1ull << LPS) - 1));
      R w = 0;
      if (mine) {
        if constexpr (sizeof(R) == 8) {
          if constexpr (FASTW) w = hit_weight_tab(dx, dy, dz, fx, fy, fz, stab, kp);
          else w = hit_weight_d(dx, dy, dz, off + 3LL * s, kp);
        } else {
          w = hit_weight_f(dx, dy, dz, f32, kp);
        }
      }
      const int li = lb + (dz * kTileEdge + dy) * kTileEdge + dx - zl * P2;
#define N3D_TSTEP(J)                                                        \
      if constexpr (J < LPS) {                                              \
        if (J < n) {                                                        \
          const R wj = grp_bcast<LPS, J>(w);                                \
          const C* row = tile + grp_bcast_i<LPS, J>(li) * CG;               \
          _Pragma("unroll") for (int c = 0; c < CPL; ++c) {                 \
            const C v = row[TL::coil(q, c, par)];                           \
            ar[c] += (A)wj * (A)v.x;                                        \
            ai[c] += (A)wj * (A)v.y;                                        \
          }                                                                 \
        }                                                                   \
      }
      N3D_TSTEP(0) N3D_TSTEP(1) N3D_TSTEP(2) N3D_TSTEP(3)
#undef N3D_TSTEP
      e += n;
      if (n < LPS) {   // the rest of the list belongs to later passes
        cur = codes[min(e + q, hl)];
        break;
      }
      cur = nxt;
    }
  }
  if (!act) return;
  R sc = scale;
  if (wsamp) sc *= wsamp[s];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int k = TL::coil(q, c, par);
    if (k < g.cg) {
      C v;
      v.x = (R)ar[c] * sc;
      v.y = (R)ai[c] * sc;
      if constexpr (USER) out[orig[s] + (long long)(g.c0 + k) * g.M] = v;
      else out[(long long)s * g.nc + g.c0 + k] = v;
    }
  }
}

// Sum over hits [e0, e1) of one grid point's list (sorted sample ids, codes
// alongside, or with pw the fp32 weights of the plan's table), weights times
// wsamp when given, into ar/ai (LPS-lane group).
// Software pipeline over steps of LPS hits: the list entries are loaded two
// steps ahead and the sample offsets one step ahead, so a step waits for one
// round trip (its sample values) instead of three dependent ones.
template <typename R, int CPL, int LPS, bool FULL, typename A>
__device__ __forceinline__ void spread_hits(long long e0, long long e1, int q,
                                            const uint2* __restrict__ hits, bool pw,
                                            const double* __restrict__ off,
                                            const float4* __restrict__ off32,
                                            const cplx<R>* __restrict__ ks,
                                            const R* __restrict__ wsamp, const GridGeom& g,
                                            const KBParams& kp, const double* tab, A* ar,
                                            A* ai) {
  if (e0 >= e1) return;
  // lane q's entry of the step starting at e (the step's first entry past the end)
  auto ent = [&](long long e) { return e + q < e1 ? e + q : e; };
  uint2 h0 = hits[ent(e0)];
  int s0 = (int)h0.x;
  unsigned c0 = h0.y;                  // code, or weight bits when pw
  float4 f0;
  if constexpr (sizeof(R) == 4) {
    if (!pw) f0 = off32[s0];
  }
  R m0 = wsamp ? wsamp[s0] : (R)1;
  const long long e1st = e0 + LPS < e1 ? e0 + LPS : e0;
  uint2 h1 = hits[ent(e1st)];
  for (long long e = e0; e < e1; e += LPS) {
    const int s = s0;
    R w;
    if (pw) {
      w = (R)__uint_as_float(c0);        // exact: fp64 weights are float-rounded too
    } else if constexpr (sizeof(R) == 8) {
      int dx, dy, dz;
      decode_hit((unsigned short)c0, g, dx, dy, dz);
      const double* f = off + 3LL * s;
      w = tab ? hit_weight_tab(dx, dy, dz, f[0], f[1], f[2], tab, kp)
              : hit_weight_d(dx, dy, dz, f, kp);
    } else {
      int dx, dy, dz;
      decode_hit((unsigned short)c0, g, dx, dy, dz);
      w = hit_weight_f(dx, dy, dz, f0, kp);
    }
    w *= m0;
    if (e + q >= e1) w = 0;
    // this step's sample values
    R wj[LPS];
    cplx<R> val[LPS][CPL];
#define N3D_SLOAD(J)                                                                 \
    if constexpr (J < LPS) {                                                         \
      wj[J] = grp_bcast<LPS, J>(w);                                                  \
      const int sj = grp_bcast_i<LPS, J>(s);                                         \
      load_coils<R, CPL, LPS, FULL>(val[J], ks + (long long)sj * g.nc + g.c0, q, g.cg); \
    }
    N3D_SLOAD(0) N3D_SLOAD(1) N3D_SLOAD(2) N3D_SLOAD(3)
#undef N3D_SLOAD
    // advance the pipeline while they are in flight; the loads are issued
    // on every path (indices clamped into the list) so the wait counts do not
    // depend on the branch taken
    {
      const long long en = e + LPS < e1 ? e + LPS : e;
      const long long enn = e + 2 * LPS < e1 ? e + 2 * LPS : en;
      s0 = (int)h1.x;
      c0 = h1.y;
      if constexpr (sizeof(R) == 4) {
        if (!pw) f0 = off32[s0];
      }
      m0 = wsamp ? wsamp[s0] : (R)1;
      h1 = hits[ent(enn)];
    }
#pragma unroll
    for (int j = 0; j < LPS; ++j) fma_coils<R, CPL>(ar, ai, wj[j], val[j]);
  }
}
#undef N3D_STEP

// ---------------------------------------------------------------------------
// One segment of a dense point's hit list -> partial[slot][cg] (work-group b
// of the segment blocks)
template <typename R, int CPL, int LPS, bool COMP, bool FULL>
__device__ __forceinline__ void spread_heavy_block(long long b, const HeavySeg* __restrict__ segs, int nseg,
                                                   const uint2* __restrict__ hits, bool pw,
                                                   const double* __restrict__ off,
                                                   const float4* __restrict__ off32,
                                                   const cplx<R>* __restrict__ ks,
                                                   const R* __restrict__ wsamp,
                                                   cplx<R>* __restrict__ partial, const GridGeom& g,
                                                   const KBParams& kp, const double* tab) {
  using A = typename AccT<R, COMP>::type;
  const long long gt = b * blockDim.x + threadIdx.x;
  const int si = (int)(gt / LPS);
  const int q = (int)(gt % LPS);
  if (si >= nseg) return;
  const HeavySeg sg = segs[si];
  A ar[CPL], ai[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) { ar[c] = 0; ai[c] = 0; }
  spread_hits<R, CPL, LPS, FULL>(sg.start, sg.end, q, hits, pw, off, off32, ks, wsamp, g, kp,
                                 tab, ar, ai);
  // partial sums keep the accumulator precision until the final reduction
  A* pp = reinterpret_cast<A*>(partial) + (long long)sg.slot * g.cg * 2;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int k = coil_of<R, CPL, LPS, FULL>(q, c);
    if (k < g.cg) { pp[2 * k] = ar[c]; pp[2 * k + 1] = ai[c]; }
  }
}

// Adjoint, output driven: an LPS-lane group per grid point walks its hit list
// (sorted sample ids, fixed at plan time) and sums in registers. No atomics:
// the result is bitwise reproducible. Points with more than rmax hits (the
// dense k-space centre of radial trajectories) are skipped here and summed
// by k_spread_heavy + k_spread_hreduce; the plan holds two segmentations
// (short segments for launches of at most 4 coils), so a coil's sum is
// grouped the same way within either size class.
template <typename R, int CPL, int LPS, bool COMP, bool FULL, typename GO = cplx<R>>
__global__ void __launch_bounds__(256) k_spread(const long long* __restrict__ rowptr,
                                                const uint2* __restrict__ hits, bool pw,
                                                const double* __restrict__ off,
                                                const float4* __restrict__ off32,
                                                const cplx<R>* __restrict__ ks,
                                                const R* __restrict__ wsamp,
                                                GO* __restrict__ grid, long long ngrid,
                                                const int* __restrict__ bricks,
                                                int rmax, GridGeom g, KBParams kp,
                                                const HeavySeg* __restrict__ segs, int nseg,
                                                int nhb, cplx<R>* __restrict__ partial,
                                                const double* __restrict__ kbtab) {
  using A = typename AccT<R, COMP>::type;
  __shared__ double stab[sizeof(R) == 8 ? kKbIntervals * kKbCoef : 1];
  const double* tab = nullptr;
  if constexpr (sizeof(R) == 8) tab = stage_kb_table(kbtab, stab);
  // the first nhb work-groups (a multiple of 8, so the XCD of every later
  // work-group is unchanged) sum the dense points' segments; the rest are
  // the light points, renumbered for the XCDs among themselves
  if ((int)blockIdx.x < nhb) {
    spread_heavy_block<R, CPL, LPS, COMP, FULL>(blockIdx.x, segs, nseg, hits, pw, off, off32, ks,
                                                wsamp, partial, g, kp, tab);
    return;
  }
  const long long gt = xcd_block_of(g.xchunk, blockIdx.x - nhb, gridDim.x - nhb) * blockDim.x +
                       threadIdx.x;
  const long long lin = gt / LPS;
  const int q = (int)(gt % LPS);
  if (lin >= ngrid) return;
  long long gpt = lin;                     // linear order (x fastest)
  if (g.flags & 2) {                       // 8x8x8 bricks in the plan's brick order
    const int t = bricks[lin / 512];
    const int l = (int)(lin % 512);
    const int tx = t & 1023, ty = (t >> 10) & 1023, tz = t >> 20;
    // Z-order within the brick, so a wave's points form a compact 4x4x2
    // block that shares more sample rows (256^3, J=6: 23.1 -> 22.7 ms; equal
    // at 128^3, J=4); tune flag 2048: x fastest
    int lx = (l & 1) | ((l >> 2) & 2) | ((l >> 4) & 4);
    int ly = ((l >> 1) & 1) | ((l >> 3) & 2) | ((l >> 5) & 4);
    int lz = ((l >> 2) & 1) | ((l >> 4) & 2) | ((l >> 6) & 4);
    if (g.flags & 2048) { lx = l & 7; ly = (l >> 3) & 7; lz = l >> 6; }
    const int x = tx * 8 + lx, y = ty * 8 + ly, z = tz * 8 + lz;
    if (x >= g.K1 || y >= g.K2 || z >= g.K3) return;
    gpt = ((long long)z * g.K2 + y) * g.K1 + x;
  }
  const long long e0 = rowptr[gpt], e1 = rowptr[gpt + 1];
  if (e1 - e0 > rmax) return;
  A ar[CPL], ai[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) { ar[c] = 0; ai[c] = 0; }
  spread_hits<R, CPL, LPS, FULL>(e0, e1, q, hits, pw, off, off32, ks, wsamp, g, kp, tab, ar, ai);
  GO* gp = grid + gpt * g.nc + g.c0;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int k = coil_of<R, CPL, LPS, FULL>(q, c);
    if (k < g.cg) {
      GO o;
      o.x = (decltype(o.x))ar[c];
      o.y = (decltype(o.y))ai[c];
      gp[k] = o;
    }
  }
}

// Adjoint, output driven with the samples staged in LDS (fp32 arithmetic,
// CG coils per launch in the aligned layout). One 512-lane work-group per
// 8x8x8 brick, one lane per grid point. The samples that touch the brick's
// points (its BrickLists list, ascending ids) pass through LDS in chunks of
// S: each chunk's coil rows and offsets are read once with coalesced 16-byte
// loads, then every lane walks its point's hits whose slot falls in the
// chunk. A point's hits are in ascending sample order, so it meets them
// chunk by chunk in list order and sums them in the same order as k_spread:
// the two kernels give bitwise the same grid. Points with more than rmax
// hits are left to k_spread_heavy.
template <int CG> struct TileSlots { static constexpr int S = CG == 16 ? 384 : 512; };

template <int CG, bool COMP>
__global__ void __launch_bounds__(512) k_spread_tile(const long long* __restrict__ rowptr,
                                                     const unsigned* __restrict__ ent,
                                                     const long long* __restrict__ bl_ptr,
                                                     const int* __restrict__ bl_samp,
                                                     const float4* __restrict__ off32,
                                                     const float2* __restrict__ ks,
                                                     float2* __restrict__ grid,
                                                     const int* __restrict__ bricks, int rmax,
                                                     GridGeom g, KBParams kp) {
  using A = typename AccT<float, COMP>::type;
  constexpr int S = TileSlots<CG>::S;
  constexpr int V4 = CG / 2;        // 16-byte pieces of a staged row
  constexpr int ROW4 = V4 + 1;      // row stride with 16 bytes of padding (LDS banks)
  __shared__ float4 s_val[S * ROW4];
  __shared__ float4 s_off[S];
  const long long bi = xcd_block(g.xchunk);
  const int t = bricks[bi];
  const int l = threadIdx.x;
  const int x = (t & 1023) * 8 + (l & 7), y = ((t >> 10) & 1023) * 8 + ((l >> 3) & 7),
            z = (t >> 20) * 8 + (l >> 6);
  const bool inside = x < g.K1 && y < g.K2 && z < g.K3;
  const long long gpt = ((long long)z * g.K2 + y) * g.K1 + x;
  long long e = 0, e1 = 0;
  bool store = false;
  if (inside) {
    e = rowptr[gpt];
    e1 = rowptr[gpt + 1];
    store = e1 - e <= rmax;
    if (!store) e1 = e;
  }
  A ar[CG], ai[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) { ar[c] = 0; ai[c] = 0; }
  // entries are loaded four ahead; indices past the list are clamped to its
  // last entry (or entry 0 for an empty list), so every load is in bounds
  const long long ecl = e1 > e ? e1 - 1 : 0;
  unsigned pf0 = ent[e < ecl ? e : ecl], pf1 = ent[e + 1 < ecl ? e + 1 : ecl];
  unsigned pf2 = ent[e + 2 < ecl ? e + 2 : ecl], pf3 = ent[e + 3 < ecl ? e + 3 : ecl];
  const long long base = bl_ptr[bi];
  const int nsl = (int)(bl_ptr[bi + 1] - base);
  const float4* ks4 = reinterpret_cast<const float4*>(ks);
  const long long rs4 = g.nc / 2;
  const int co4 = g.c0 / 2;
  for (int c0 = 0; c0 < nsl; c0 += S) {
    const int cn = nsl - c0 < S ? nsl - c0 : S;
    __syncthreads();   // the previous chunk is consumed
    for (int i = l; i < cn * ROW4; i += blockDim.x) {
      const int r = i / ROW4, k = i - r * ROW4;
      const int sm = bl_samp[base + c0 + r];
      if (k < V4) s_val[r * ROW4 + k] = ks4[(long long)sm * rs4 + co4 + k];
      else s_off[r] = off32[sm];
    }
    __syncthreads();
    const int cend = c0 + cn;
    while (e < e1) {
      const unsigned cur = pf0;
      const int sl = (int)(cur >> 15);
      if (sl >= cend) break;
      pf0 = pf1;
      pf1 = pf2;
      pf2 = pf3;
      pf3 = ent[e + 4 < ecl ? e + 4 : ecl];
      ++e;
      int dx, dy, dz;
      decode_hit((unsigned short)(cur & 0x7fffu), g, dx, dy, dz);
      const int r = sl - c0;
      const float w = hit_weight_f(dx, dy, dz, s_off[r], kp);
      const float4* row = s_val + r * ROW4;
#pragma unroll
      for (int k = 0; k < V4; ++k) {
        const float4 v = row[k];
        ar[2 * k] += (A)w * (A)v.x;
        ai[2 * k] += (A)w * (A)v.y;
        ar[2 * k + 1] += (A)w * (A)v.z;
        ai[2 * k + 1] += (A)w * (A)v.w;
      }
    }
  }
  if (store) {
    float4* gp = reinterpret_cast<float4*>(grid + gpt * g.nc + g.c0);
#pragma unroll
    for (int k = 0; k < V4; ++k)
      gp[k] = make_float4((float)ar[2 * k], (float)ai[2 * k], (float)ar[2 * k + 1],
                          (float)ai[2 * k + 1]);
  }
}

// Split-precision adjoint (fp32 samples, fp64 grid) with the samples staged
// in LDS: one 512-lane work-group per 8x8x8 brick, one lane per grid point,
// CG coils. Each hit's entry holds the slot of its sample in the brick's list
// (ascending ids) and its fp32 weight from the plan's table; the brick's
// sample rows pass through LDS in chunks of S, read once with 16-byte loads.
// A point meets its hits chunk by chunk in list order and sums them in
// double with the same products and order as k_spread, so the grids are
// bitwise equal. Points with more than rmax hits are summed by the first nhb
// work-groups (a multiple of 8, so the XCD of every later work-group is
// unchanged) as k_spread_heavy does, overlapping the bricks.
template <int CPL, int LPS>
__global__ void __launch_bounds__(512) k_spread_tile_w(const long long* __restrict__ rowptr,
                                                       const uint2* __restrict__ went,
                                                       const long long* __restrict__ wbeg,
                                                       const long long* __restrict__ wend,
                                                       const int* __restrict__ wsamp_ids,
                                                       const float2* __restrict__ ks,
                                                       double2* __restrict__ grid,
                                                       const int* __restrict__ bricks, int rmax,
                                                       GridGeom g, const uint2* __restrict__ hitw,
                                                       const HeavySeg* __restrict__ segs, int nseg,
                                                       int nhb, double2* __restrict__ partial,
                                                       KBParams kp) {
  constexpr int CG = CPL * LPS;
  constexpr int S = 512;
  constexpr int V4 = CG / 2;        // 16-byte pieces of a staged row
  constexpr int ROW4 = V4 + 1;      // row stride with 16 bytes of padding (LDS banks)
  __shared__ float4 s_val[S * ROW4];
  if ((int)blockIdx.x < nhb) {
    spread_heavy_block<float, CPL, LPS, true, true>(blockIdx.x, segs, nseg, hitw, true, nullptr,
                                                    nullptr, reinterpret_cast<const cplx<float>*>(ks),
                                                    nullptr, reinterpret_cast<cplx<float>*>(partial),
                                                    g, kp, nullptr);
    return;
  }
  const long long bi = xcd_block_of(g.xchunk, blockIdx.x - nhb, gridDim.x - nhb);
  const int t = bricks[bi];
  const int l = threadIdx.x;
  const int x = (t & 1023) * 8 + (l & 7), y = ((t >> 10) & 1023) * 8 + ((l >> 3) & 7),
            z = (t >> 20) * 8 + (l >> 6);
  const bool inside = x < g.K1 && y < g.K2 && z < g.K3;
  const long long gpt = ((long long)z * g.K2 + y) * g.K1 + x;
  long long e = 0, e1 = 0;
  bool store = false;
  if (inside) {
    e = rowptr[gpt];
    e1 = rowptr[gpt + 1];
    store = e1 - e <= rmax;
    if (!store) e1 = e;
  }
  double ar[CG], ai[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) { ar[c] = 0; ai[c] = 0; }
  // entries are loaded four ahead; indices past the list are clamped to its
  // last entry (or entry 0 for an empty list), so every load is in bounds
  const long long ecl = e1 > e ? e1 - 1 : 0;
  uint2 pf0 = went[e < ecl ? e : ecl], pf1 = went[e + 1 < ecl ? e + 1 : ecl];
  uint2 pf2 = went[e + 2 < ecl ? e + 2 : ecl], pf3 = went[e + 3 < ecl ? e + 3 : ecl];
  const long long base = wbeg[bi];
  const int nsl = (int)(wend[bi] - base);
  const float4* ks4 = reinterpret_cast<const float4*>(ks);
  const long long rs4 = g.nc / 2;
  const int co4 = g.c0 / 2;
  for (int c0 = 0; c0 < nsl; c0 += S) {
    const int cn = nsl - c0 < S ? nsl - c0 : S;
    __syncthreads();   // the previous chunk is consumed
    for (int i = l; i < cn * V4; i += blockDim.x) {
      const int r = i / V4, k = i - r * V4;
      const int sm = wsamp_ids[base + c0 + r];
      s_val[r * ROW4 + k] = ks4[(long long)sm * rs4 + co4 + k];
    }
    __syncthreads();
    const unsigned cend = (unsigned)(c0 + cn);
    while (e < e1) {
      const uint2 cur = pf0;
      if (cur.x >= cend) break;
      pf0 = pf1;
      pf1 = pf2;
      pf2 = pf3;
      pf3 = went[e + 4 < ecl ? e + 4 : ecl];
      ++e;
      const float w = __uint_as_float(cur.y);
      const float4* row = s_val + (int)(cur.x - (unsigned)c0) * ROW4;
#pragma unroll
      for (int k = 0; k < V4; ++k) {
        const float4 v = row[k];
        ar[2 * k] += (double)w * (double)v.x;
        ai[2 * k] += (double)w * (double)v.y;
        ar[2 * k + 1] += (double)w * (double)v.z;
        ai[2 * k + 1] += (double)w * (double)v.w;
      }
    }
  }
  if (store) {
    double2* gp = grid + gpt * g.nc + g.c0;
#pragma unroll
    for (int c = 0; c < CG; ++c) gp[c] = make_double2(ar[c], ai[c]);
  }
}

// Adjoint of fp64-arithmetic plans from the block-staged lists (BlockLists):
// one wavefront per piece (a run of chunks of one 4x4x4 block), one lane per
// point of the block, 8 coils. A chunk's 64 sample rows are staged in the
// wave's LDS slice, one row per lane: float coils (the split-precision
// aNUFT, 64 bytes) or double coils times the sample's data weight (iNUFT's
// internal samples, 128 bytes). Every lane then walks its entries of the
// chunk, 4 steps per load pair, and sums weight x value in double (for float
// rows the exact products of fp32 weights and fp32 samples). A point's sum
// runs over its chunks in order and within a chunk in list order, fixed at
// plan time: deterministic (not bitwise the gather's order). The four waves
// of a work-group run independent pieces, with no work-group barrier.
template <typename S> struct BlkRow { static constexpr int U = sizeof(S) == 8 ? 5 : 9; };  // 16-byte units (data + 1 of padding)

// entry groups the wide block-staged kernels keep in flight ahead of use:
// at the 256^3 precision workload the forward measured 6.18 / 6.08 / 6.11 /
// 6.35 ms with 1 / 2 / 3 / 5 groups, and the adjoint 6.70 ms with 1 and
// 6.98-7.17 ms with more (the registers cost it a wave per SIMD)
constexpr int kEntryAheadAdj = 1;
constexpr int kEntryAheadFwd = 2;

// LDS writes of a wavefront visible to its own later reads
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename S, int C>
__global__ void __launch_bounds__(256) k_spread_blk(const int4* __restrict__ pieces,
                                                    long long npiece, const int* __restrict__ blk,
                                                    const long long* __restrict__ cbase,
                                                    const int* __restrict__ samp,
                                                    const long long* __restrict__ eoff,
                                                    const unsigned* __restrict__ eslot,
                                                    const float4* __restrict__ ew,
                                                    const S* __restrict__ ks,
                                                    const double* __restrict__ wsamp,
                                                    double2* __restrict__ grid,
                                                    double2* __restrict__ partial, GridGeom g) {
  constexpr int RU = BlkRow<S>::U;
  constexpr int DU = RU - 1;          // 16-byte units of data per row (4 float, 8 double)
  __shared__ float4 s_row[4 * C * RU];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long pi = xcd_block(g.xchunk) * 4 + wv;
  if (pi >= npiece) return;
  const int4 pc = pieces[pi];
  const int bc = blk[pc.x];
  const int x = (bc & 1023) * 4 + (lane & 3), y = ((bc >> 10) & 1023) * 4 + ((lane >> 2) & 3),
            z = (bc >> 20) * 4 + (lane >> 4);
  double ar[8], ai[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) { ar[c] = 0; ai[c] = 0; }
  const long long cb = cbase[pc.x];
  float4* srow = s_row + wv * (C * RU);
  const float4* ks4 = reinterpret_cast<const float4*>(ks);
  const long long rs4 = (long long)g.nc * sizeof(S) / 16;   // 16-byte units per sample row
  const long long co4 = (long long)g.c0 * sizeof(S) / 16;
  // a whole 8-coil group at a 16-byte aligned row offset takes 16-byte loads
  const bool full = g.cg == 8 && (sizeof(S) == 16 || ((g.nc & 1) == 0 && (g.c0 & 1) == 0));
  for (int c = pc.y; c < pc.z; ++c) {
    const long long ch = cb + c;
    wave_lds_sync();   // the previous chunk's rows are consumed
#pragma unroll
    for (int h = 0; h < C / 64; ++h) {   // rows lane, lane + 64 of the chunk
      const int sm = samp[ch * C + h * 64 + lane];
      const float4* src = ks4 + (long long)sm * rs4 + co4;
      const S* srcc = ks + (long long)sm * g.nc + g.c0;   // coil by coil (partial groups)
      double m = 1.0;   // iNUFT's data weight of the sample, applied as it is staged
      if constexpr (sizeof(S) == 16) {
        if (wsamp) m = wsamp[sm];
      }
      float4* dst = srow + (h * 64 + lane) * RU;
      float4 v[DU];
      if (full) {   // all loads issued before any is used
#pragma unroll
        for (int k = 0; k < DU; ++k) v[k] = src[k];
      } else {
#pragma unroll
        for (int k = 0; k < DU; ++k) {
          if constexpr (sizeof(S) == 8) {   // coils 2k, 2k+1 (zero past the group)
            const float2 a = 2 * k < g.cg ? srcc[2 * k] : make_float2(0.f, 0.f);
            const float2 b = 2 * k + 1 < g.cg ? srcc[2 * k + 1] : make_float2(0.f, 0.f);
            v[k] = make_float4(a.x, a.y, b.x, b.y);
          } else {                          // coil k
            v[k] = k < g.cg ? __builtin_bit_cast(float4, srcc[k]) : make_float4(0.f, 0.f, 0.f, 0.f);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < DU; ++k) {
        if constexpr (sizeof(S) == 16) {
          if (wsamp) {
            double2 d = __builtin_bit_cast(double2, v[k]);
            d.x *= m;
            d.y *= m;
            v[k] = __builtin_bit_cast(float4, d);
          }
        }
        dst[k] = v[k];
      }
    }
    wave_lds_sync();
    const long long q0 = eoff[ch], q1 = eoff[ch + 1];
    // entry groups loaded one group ahead (index clamped into the chunk)
    unsigned sl4 = 0xffffffffu;
    float4 w4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q0 < q1) { sl4 = eslot[q0 * 64 + lane]; w4 = ew[q0 * 64 + lane]; }
    for (long long q = q0; q < q1; ++q) {
      const long long qn = q + 1 < q1 ? q + 1 : q;
      const unsigned nsl4 = eslot[qn * 64 + lane];
      const float4 nw4 = ew[qn * 64 + lane];
      const float wk[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned sl = (sl4 >> (8 * k)) & 255u;
        if (sl != 255u) {
          const float4* row = srow + sl * RU;
          const double w = (double)wk[k];
          if constexpr (sizeof(S) == 8) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const float4 v = row[j];
              ar[2 * j] += w * (double)v.x;
              ai[2 * j] += w * (double)v.y;
              ar[2 * j + 1] += w * (double)v.z;
              ai[2 * j + 1] += w * (double)v.w;
            }
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const double2 v = __builtin_bit_cast(double2, row[j]);
              ar[j] += w * v.x;
              ai[j] += w * v.y;
            }
          }
        }
      }
      sl4 = nsl4;
      w4 = nw4;
    }
  }
  if (x >= g.K1 || y >= g.K2 || z >= g.K3) return;
  if (pc.w < 0) {
    double2* o = grid + (((long long)z * g.K2 + y) * g.K1 + x) * g.nc + g.c0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (c < g.cg) o[c] = make_double2(ar[c], ai[c]);
  } else {
    double2* o = partial + ((long long)pc.w * 64 + lane) * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c) o[c] = make_double2(ar[c], ai[c]);
  }
}

// The same adjoint from the wide lists: one work-group per piece of an
// 8x8x4 block, four wavefronts of 4x4x4 points (wavefront ((y>>2)&1,
// (x>>2)&1), lane (x&3, y&3, z&3)). A chunk's kBlkWideRows sample rows are
// staged in LDS by the whole work-group (one row per thread) and each
// wavefront walks its own entries of the chunk; the next chunk's row is
// loaded into registers while the current chunk is consumed.
template <typename S>
__global__ void __launch_bounds__(256) k_spread_blkw(const int4* __restrict__ pieces,
                                                     long long npiece, const int* __restrict__ blk,
                                                     const long long* __restrict__ cbase,
                                                     const int* __restrict__ samp,
                                                     const long long* __restrict__ eoff,
                                                     const unsigned* __restrict__ eslot,
                                                     const float4* __restrict__ ew,
                                                     const S* __restrict__ ks,
                                                     const double* __restrict__ wsamp,
                                                     double2* __restrict__ grid,
                                                     double2* __restrict__ partial, GridGeom g) {
  constexpr int C = kBlkWideRows;
  constexpr int RU = BlkRow<S>::U;
  constexpr int DU = RU - 1;          // 16-byte units of data per row (4 float, 8 double)
  __shared__ float4 s_row[C * RU];
  const long long pi = xcd_block(g.xchunk);
  if (pi >= npiece) return;   // the whole work-group
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const int4 pc = pieces[pi];
  const int bc = blk[pc.x];
  const int x = (bc & 1023) * 8 + (wv & 1) * 4 + (lane & 3);
  const int y = ((bc >> 10) & 1023) * 8 + (wv >> 1) * 4 + ((lane >> 2) & 3);
  const int z = (bc >> 20) * 4 + (lane >> 4);
  double ar[8], ai[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) { ar[c] = 0; ai[c] = 0; }
  const long long cb = cbase[pc.x];
  const float4* ks4 = reinterpret_cast<const float4*>(ks);
  const long long rs4 = (long long)g.nc * sizeof(S) / 16;   // 16-byte units per sample row
  const long long co4 = (long long)g.c0 * sizeof(S) / 16;
  // a whole 8-coil group at a 16-byte aligned row offset takes 16-byte loads
  const bool full = g.cg == 8 && (sizeof(S) == 16 || ((g.nc & 1) == 0 && (g.c0 & 1) == 0));
  float4 v[DU];
  // row t of chunk ch into v (threads past the chunk's rows load nothing)
  auto fetch = [&](long long ch) {
    if (t >= C) return;
    const int sm = samp[ch * C + t];
    const float4* src = ks4 + (long long)sm * rs4 + co4;
    const S* srcc = ks + (long long)sm * g.nc + g.c0;   // coil by coil (partial groups)
    if (full) {   // all loads issued before any is used
#pragma unroll
      for (int k = 0; k < DU; ++k) v[k] = src[k];
    } else {
#pragma unroll
      for (int k = 0; k < DU; ++k) {
        if constexpr (sizeof(S) == 8) {   // coils 2k, 2k+1 (zero past the group)
          const float2 a = 2 * k < g.cg ? srcc[2 * k] : make_float2(0.f, 0.f);
          const float2 b = 2 * k + 1 < g.cg ? srcc[2 * k + 1] : make_float2(0.f, 0.f);
          v[k] = make_float4(a.x, a.y, b.x, b.y);
        } else {                          // coil k
          v[k] = k < g.cg ? __builtin_bit_cast(float4, srcc[k]) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
    }
    if constexpr (sizeof(S) == 16) {   // iNUFT's data weight of the sample, applied as it is staged
      if (wsamp) {
        const double m = wsamp[sm];
#pragma unroll
        for (int k = 0; k < DU; ++k) {
          double2 d = __builtin_bit_cast(double2, v[k]);
          d.x *= m;
          d.y *= m;
          v[k] = __builtin_bit_cast(float4, d);
        }
      }
    }
  };
  if (pc.y < pc.z) fetch(cb + pc.y);
  for (int c = pc.y; c < pc.z; ++c) {
    const long long ch = cb + c;
    __syncthreads();   // the previous chunk's rows are consumed
    if (t < C) {
      float4* dst = s_row + t * RU;
#pragma unroll
      for (int k = 0; k < DU; ++k) dst[k] = v[k];
    }
    __syncthreads();
    const long long q0 = eoff[ch * 4 + wv], q1 = eoff[ch * 4 + wv + 1];
    // entry groups loaded kEntryAheadAdj groups ahead (indices clamped into the range)
    unsigned slq[kEntryAheadAdj];
    float4 wq[kEntryAheadAdj];
#pragma unroll
    for (int d = 0; d < kEntryAheadAdj; ++d) {
      const long long qd = q0 + d < q1 ? q0 + d : q1 - 1;
      slq[d] = 0xffffffffu;
      wq[d] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (q0 < q1) { slq[d] = eslot[qd * 64 + lane]; wq[d] = ew[qd * 64 + lane]; }
    }
    if (c + 1 < pc.z) fetch(ch + 1);
    for (long long q = q0; q < q1; ++q) {
      const unsigned sl4 = slq[0];
      const float4 w4 = wq[0];
#pragma unroll
      for (int d = 0; d + 1 < kEntryAheadAdj; ++d) { slq[d] = slq[d + 1]; wq[d] = wq[d + 1]; }
      const long long qn = q + kEntryAheadAdj < q1 ? q + kEntryAheadAdj : q1 - 1;
      slq[kEntryAheadAdj - 1] = eslot[qn * 64 + lane];
      wq[kEntryAheadAdj - 1] = ew[qn * 64 + lane];
      const float wk[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned sl = (sl4 >> (8 * k)) & 255u;
        if (sl != 255u) {
          const float4* row = s_row + sl * RU;
          const double w = (double)wk[k];
          if constexpr (sizeof(S) == 8) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const float4 r = row[j];
              ar[2 * j] += w * (double)r.x;
              ai[2 * j] += w * (double)r.y;
              ar[2 * j + 1] += w * (double)r.z;
              ai[2 * j + 1] += w * (double)r.w;
            }
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              const double2 r = __builtin_bit_cast(double2, row[j]);
              ar[j] += w * r.x;
              ai[j] += w * r.y;
            }
          }
        }
      }
    }
  }
  if (x >= g.K1 || y >= g.K2 || z >= g.K3) return;
  if (pc.w < 0) {
    double2* o = grid + (((long long)z * g.K2 + y) * g.K1 + x) * g.nc + g.c0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (c < g.cg) o[c] = make_double2(ar[c], ai[c]);
  } else {
    double2* o = partial + ((long long)pc.w * 256 + t) * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c) o[c] = make_double2(ar[c], ai[c]);
  }
}

// Blocks in several pieces: each point's partial sums added in piece order.
// One work-group per block, one thread per (point, coil): 8 consecutive
// threads read a point's 128-byte partial row.
// T points per block (64: 4x4x4 blocks; 256: wide 8x8x4 blocks, a point's
// thread as in k_spread_blkw): work-group (red entry, wavefront) covers the
// 64 points of one wavefront.
__global__ void __launch_bounds__(512) k_spread_blk_reduce(const int4* __restrict__ red,
                                                           const int* __restrict__ blk, int T,
                                                           const double2* __restrict__ partial,
                                                           double2* __restrict__ grid, GridGeom g) {
  const int4 r = red[blockIdx.x];   // (block, first partial slot, pieces, 0)
  const int wv = blockIdx.y, lane = threadIdx.x >> 3, c = threadIdx.x & 7;
  if (c >= g.cg) return;
  const int bc = blk[r.x];
  const int bs = T == 256 ? 8 : 4;
  const int x = (bc & 1023) * bs + (wv & 1) * 4 + (lane & 3);
  const int y = ((bc >> 10) & 1023) * bs + (wv >> 1) * 4 + ((lane >> 2) & 3);
  const int z = (bc >> 20) * 4 + (lane >> 4);
  if (x >= g.K1 || y >= g.K2 || z >= g.K3) return;
  const long long ps = (long long)T * 8;   // one piece's partial rows
  const double2* pp = partial + ((long long)r.y * T + wv * 64 + lane) * 8 + c;
  double ax = 0, ay = 0;
  int j = 0;
  for (; j + 4 <= r.z; j += 4) {   // four pieces' loads in flight, added in piece order
    const double2 v0 = pp[0], v1 = pp[ps], v2 = pp[2 * ps], v3 = pp[3 * ps];
    ax += v0.x; ay += v0.y;
    ax += v1.x; ay += v1.y;
    ax += v2.x; ay += v2.y;
    ax += v3.x; ay += v3.y;
    pp += 4 * ps;
  }
  for (; j < r.z; ++j, pp += ps) {
    const double2 v = pp[0];
    ax += v.x; ay += v.y;
  }
  grid[(((long long)z * g.K2 + y) * g.K1 + x) * g.nc + g.c0 + c] = make_double2(ax, ay);
}

// Forward (fp64 arithmetic) from the block-staged lists (FwdLists): one
// wavefront per group of 64 sorted samples, one lane per sample, 8 coils. A
// chunk's 64 grid rows (128 bytes of fp64 coils each, and the point's
// coordinates) are staged in the wave's LDS slice, one row per lane; each
// lane walks its entries of the chunk and evaluates each hit's weight from
// its own sample's offsets and the staged point's coordinates, with the same
// function and arguments as k_interp (the same weights), and sums the
// products in double. A sample's sum runs over its chunks in order, fixed at
// plan time: deterministic (not bitwise the gather's z-y-x order). The four
// waves of a work-group run independent groups after staging the kernel
// table together.
constexpr int kFwdRow = 9;   // 16-byte units per staged row: 8 of data, 1 of padding (LDS banks)

template <typename RO, bool USER, bool TAB, bool STW>
__global__ void __launch_bounds__(256) k_interp_blk(const long long* __restrict__ cbase,
                                                    long long ngrp, const int* __restrict__ gpt,
                                                    const long long* __restrict__ eoff,
                                                    const unsigned* __restrict__ eslot,
                                                    const float4* __restrict__ ew,
                                                    const int4* __restrict__ base,
                                                    const double* __restrict__ off,
                                                    const double2* __restrict__ grid,
                                                    cplx<RO>* __restrict__ out,
                                                    const double* __restrict__ wsamp, double scale,
                                                    const int* __restrict__ orig,
                                                    const double* __restrict__ kbtab, GridGeom g,
                                                    KBParams kp) {
  __shared__ double2 s_row[4 * 64 * kFwdRow];
  __shared__ int s_pc[4 * 64];
  __shared__ double stab[TAB && !STW ? kKbIntervals * kKbCoef : 1];
  const double* tab = nullptr;
  if constexpr (TAB && !STW) tab = stage_kb_table(kbtab, stab);   // every thread (it synchronises)
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long j = xcd_block(g.xchunk) * 4 + wv;
  if (j >= ngrp) return;
  const long long s = j * 64 + lane;
  const bool act = s < g.nnz;
  const long long sv = act ? s : (long long)g.nnz - 1;   // a lane past the end has no entries
  const int4 b = base[sv];
  const double* f64 = off + 3 * sv;
  const double fx = f64[0], fy = f64[1], fz = f64[2];
  double ar[8], ai[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) { ar[c] = 0; ai[c] = 0; }
  double2* srow = s_row + wv * 64 * kFwdRow;
  int* spc = s_pc + wv * 64;
  const long long c0 = cbase[j], c1 = cbase[j + 1];
  for (long long ch = c0; ch < c1; ++ch) {
    const int gp = gpt[ch * 64 + lane];
    const double2* srcc = grid + (long long)gp * g.nc + g.c0;
    const int px = gp % g.K1, py = (gp / g.K1) % g.K2, pz = gp / (g.K1 * g.K2);
    wave_lds_sync();   // the previous chunk's rows are consumed
    double2* dst = srow + lane * kFwdRow;
    if (g.cg == 8) {
      const double4* src = reinterpret_cast<const double4*>(srcc);
      const double4 r0 = src[0], r1 = src[1], r2 = src[2], r3 = src[3];
      dst[0] = make_double2(r0.x, r0.y); dst[1] = make_double2(r0.z, r0.w);
      dst[2] = make_double2(r1.x, r1.y); dst[3] = make_double2(r1.z, r1.w);
      dst[4] = make_double2(r2.x, r2.y); dst[5] = make_double2(r2.z, r2.w);
      dst[6] = make_double2(r3.x, r3.y); dst[7] = make_double2(r3.z, r3.w);
    } else {   // a partial group: its coils, zero past them
#pragma unroll
      for (int k = 0; k < 8; ++k) dst[k] = k < g.cg ? srcc[k] : make_double2(0.0, 0.0);
    }
    spc[lane] = px | (py << 10) | (pz << 20);
    wave_lds_sync();
    const long long q0 = eoff[ch], q1 = eoff[ch + 1];
    unsigned sl4 = q0 < q1 ? eslot[q0 * 64 + lane] : 0xffffffffu;
    float4 w4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (STW) if (q0 < q1) w4 = ew[q0 * 64 + lane];
    for (long long q = q0; q < q1; ++q) {
      const long long qn = q + 1 < q1 ? q + 1 : q;
      const unsigned nsl4 = eslot[qn * 64 + lane];
      float4 nw4 = w4;
      if constexpr (STW) nw4 = ew[qn * 64 + lane];
      const float wk[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned sl = (sl4 >> (8 * k)) & 255u;
        if (sl != 255u) {
          double w;
          if constexpr (STW) {
            w = (double)wk[k];   // the plan's fp32 weight of this hit
          } else {
          const int pc = spc[sl];
          // candidate offset of the staged point from the sample's base
          // (K >= 2W + 2 per axis: one candidate reaches it)
          int dx = (pc & 1023) - b.x, dy = ((pc >> 10) & 1023) - b.y, dz = (pc >> 20) - b.z;
          dx = dx > g.W1 ? dx - g.K1 : (dx < -g.W1 ? dx + g.K1 : dx);
          dy = dy > g.W2 ? dy - g.K2 : (dy < -g.W2 ? dy + g.K2 : dy);
          dz = dz > g.W3 ? dz - g.K3 : (dz < -g.W3 ? dz + g.K3 : dz);
          if constexpr (TAB) w = hit_weight_tab(dx, dy, dz, fx, fy, fz, tab, kp);
          else w = hit_weight_d(dx, dy, dz, f64, kp);
          }
          const double2* row = srow + sl * kFwdRow;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const double2 v = row[c];
            ar[c] += w * v.x;
            ai[c] += w * v.y;
          }
        }
      }
      sl4 = nsl4;
      w4 = nw4;
    }
  }
  if (!act) return;
  double sc = scale;
  if (wsamp) sc *= wsamp[s];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c >= g.cg) break;
    cplx<RO> v;
    v.x = (RO)(ar[c] * sc);
    v.y = (RO)(ai[c] * sc);
    if constexpr (USER) out[orig[s] + (long long)(g.c0 + c) * g.M] = v;
    else out[s * g.nc + g.c0 + c] = v;
  }
}

// The same forward from lists built with groups of 256 sorted samples: one
// work-group per group, one thread per sample. The group's chunks of 255
// grid rows (slot 255 stays "none") are staged in LDS by the whole
// work-group (one row per thread) and each wavefront walks its own entries
// of the chunk; a larger group shares each staged row among more samples,
// and a larger chunk evens out the lanes' step counts (6.25 -> 6.00 ms at
// the 256^3 precision workload against chunks of 128). The next chunk's
// rows are loaded into registers while the current chunk is consumed.
constexpr int kFwdRowsW = 255;

template <typename RO, bool USER, bool TAB, bool STW>
__global__ void __launch_bounds__(256) k_interp_blkw(const long long* __restrict__ cbase,
                                                     long long ngrp, const int* __restrict__ gpt,
                                                     const long long* __restrict__ eoff,
                                                     const unsigned* __restrict__ eslot,
                                                     const float4* __restrict__ ew,
                                                     const int4* __restrict__ base,
                                                     const double* __restrict__ off,
                                                     const double2* __restrict__ grid,
                                                     cplx<RO>* __restrict__ out,
                                                     const double* __restrict__ wsamp, double scale,
                                                     const int* __restrict__ orig,
                                                     const double* __restrict__ kbtab, GridGeom g,
                                                     KBParams kp) {
  __shared__ double2 s_row[kFwdRowsW * kFwdRow];
  __shared__ int s_pc[kFwdRowsW];
  __shared__ double stab[TAB && !STW ? kKbIntervals * kKbCoef : 1];
  const double* tab = nullptr;
  if constexpr (TAB && !STW) tab = stage_kb_table(kbtab, stab);   // every thread (it synchronises)
  const long long j = xcd_block(g.xchunk);
  if (j >= ngrp) return;   // the whole work-group
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const long long s = j * 256 + t;
  const bool act = s < g.nnz;
  const long long sv = act ? s : (long long)g.nnz - 1;   // a thread past the end has no entries
  const int4 b = base[sv];
  const double* f64 = off + 3 * sv;
  const double fx = f64[0], fy = f64[1], fz = f64[2];
  double ar[8], ai[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) { ar[c] = 0; ai[c] = 0; }
  // staging role: row t of the chunk, all 8 coils (threads past the rows idle)
  const int r = t;
  double4 v0 = make_double4(0.0, 0.0, 0.0, 0.0), v1 = v0, v2 = v0, v3 = v0;
  int pcv = 0;
  auto fetch = [&](long long ch) {
    if (r >= kFwdRowsW) return;
    const int gp = gpt[ch * kFwdRowsW + r];
    const double2* src = grid + (long long)gp * g.nc + g.c0;
    if (g.cg == 8) {
      v0 = reinterpret_cast<const double4*>(src)[0];
      v1 = reinterpret_cast<const double4*>(src)[1];
      v2 = reinterpret_cast<const double4*>(src)[2];
      v3 = reinterpret_cast<const double4*>(src)[3];
    } else {   // a partial group: its coils, zero past them
      const double2 z = make_double2(0.0, 0.0);
      double2 a[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = k < g.cg ? src[k] : z;
      v0 = make_double4(a[0].x, a[0].y, a[1].x, a[1].y);
      v1 = make_double4(a[2].x, a[2].y, a[3].x, a[3].y);
      v2 = make_double4(a[4].x, a[4].y, a[5].x, a[5].y);
      v3 = make_double4(a[6].x, a[6].y, a[7].x, a[7].y);
    }
    const int px = gp % g.K1, py = (gp / g.K1) % g.K2, pz = gp / (g.K1 * g.K2);
    pcv = px | (py << 10) | (pz << 20);
  };
  const long long c0 = cbase[j], c1 = cbase[j + 1];
  if (c0 < c1) fetch(c0);
  for (long long ch = c0; ch < c1; ++ch) {
    __syncthreads();   // the previous chunk's rows are consumed
    if (r < kFwdRowsW) {
      double2* dst = s_row + r * kFwdRow;
      dst[0] = make_double2(v0.x, v0.y);
      dst[1] = make_double2(v0.z, v0.w);
      dst[2] = make_double2(v1.x, v1.y);
      dst[3] = make_double2(v1.z, v1.w);
      dst[4] = make_double2(v2.x, v2.y);
      dst[5] = make_double2(v2.z, v2.w);
      dst[6] = make_double2(v3.x, v3.y);
      dst[7] = make_double2(v3.z, v3.w);
      s_pc[r] = pcv;
    }
    __syncthreads();
    const long long q0 = eoff[ch * 4 + wv], q1 = eoff[ch * 4 + wv + 1];
    // entry groups loaded kEntryAheadFwd groups ahead (indices clamped into the range)
    unsigned slq[kEntryAheadFwd];
    float4 wq[kEntryAheadFwd];
#pragma unroll
    for (int d = 0; d < kEntryAheadFwd; ++d) {
      const long long qd = q0 + d < q1 ? q0 + d : q1 - 1;
      slq[d] = q0 < q1 ? eslot[qd * 64 + lane] : 0xffffffffu;
      wq[d] = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (STW) if (q0 < q1) wq[d] = ew[qd * 64 + lane];
    }
    if (ch + 1 < c1) fetch(ch + 1);
    for (long long q = q0; q < q1; ++q) {
      const unsigned sl4 = slq[0];
      const float4 w4 = wq[0];
#pragma unroll
      for (int d = 0; d + 1 < kEntryAheadFwd; ++d) { slq[d] = slq[d + 1]; wq[d] = wq[d + 1]; }
      const long long qn = q + kEntryAheadFwd < q1 ? q + kEntryAheadFwd : q1 - 1;
      slq[kEntryAheadFwd - 1] = eslot[qn * 64 + lane];
      if constexpr (STW) wq[kEntryAheadFwd - 1] = ew[qn * 64 + lane];
      const float wk[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned sl = (sl4 >> (8 * k)) & 255u;
        if (sl != 255u) {
          double w;
          if constexpr (STW) {
            w = (double)wk[k];   // the plan's fp32 weight of this hit
          } else {
          const int pc = s_pc[sl];
          // candidate offset of the staged point from the sample's base
          // (K >= 2W + 2 per axis: one candidate reaches it)
          int dx = (pc & 1023) - b.x, dy = ((pc >> 10) & 1023) - b.y, dz = (pc >> 20) - b.z;
          dx = dx > g.W1 ? dx - g.K1 : (dx < -g.W1 ? dx + g.K1 : dx);
          dy = dy > g.W2 ? dy - g.K2 : (dy < -g.W2 ? dy + g.K2 : dy);
          dz = dz > g.W3 ? dz - g.K3 : (dz < -g.W3 ? dz + g.K3 : dz);
          if constexpr (TAB) w = hit_weight_tab(dx, dy, dz, fx, fy, fz, tab, kp);
          else w = hit_weight_d(dx, dy, dz, f64, kp);
          }
          const double2* row = s_row + sl * kFwdRow;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const double2 v = row[c];
            ar[c] += w * v.x;
            ai[c] += w * v.y;
          }
        }
      }
    }
  }
  if (!act) return;
  double sc = scale;
  if (wsamp) sc *= wsamp[s];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c >= g.cg) break;
    cplx<RO> v;
    v.x = (RO)(ar[c] * sc);
    v.y = (RO)(ai[c] * sc);
    if constexpr (USER) out[orig[s] + (long long)(g.c0 + c) * g.M] = v;
    else out[s * g.nc + g.c0 + c] = v;
  }
}

template <typename R, int CPL, int LPS, bool COMP, bool FULL>
__global__ void __launch_bounds__(256) k_spread_heavy(const HeavySeg* __restrict__ segs, int nseg,
                                                      const uint2* __restrict__ hits, bool pw,
                                                      const double* __restrict__ off,
                                                      const float4* __restrict__ off32,
                                                      const cplx<R>* __restrict__ ks,
                                                      const R* __restrict__ wsamp,
                                                      cplx<R>* __restrict__ partial, GridGeom g,
                                                      KBParams kp, const double* __restrict__ kbtab) {
  __shared__ double stab[sizeof(R) == 8 ? kKbIntervals * kKbCoef : 1];
  const double* tab = nullptr;
  if constexpr (sizeof(R) == 8) tab = stage_kb_table(kbtab, stab);
  spread_heavy_block<R, CPL, LPS, COMP, FULL>(blockIdx.x, segs, nseg, hits, pw, off, off32, ks,
                                              wsamp, partial, g, kp, tab);
}

// Sum of a dense point's segment partials, one work-group per point: the R
// = blockDim/cgp threads of coil c (cgp = cg rounded up to a power of two)
// each add the slots r, r+R, r+2R, ... in order, then a fixed-order tree over
// the R partial sums. Deterministic, and a coil's value does not depend on
// the other coils of the launch (R = 32, or 16 for 16-coil groups).
template <typename R, bool COMP>
__global__ void __launch_bounds__(256) k_spread_hreduce(const HeavyPt* __restrict__ pts,
                                                        const cplx<R>* __restrict__ partial,
                                                        cplx<R>* __restrict__ grid, int cgp,
                                                        GridGeom g) {
  using A = typename AccT<R, COMP>::type;
  __shared__ A sre[256], sim[256];
  const HeavyPt hp = pts[blockIdx.x];
  const int nr = blockDim.x / cgp;
  const int c = threadIdx.x / nr, r = threadIdx.x % nr;
  const A* pp = reinterpret_cast<const A*>(partial);
  A re = 0, im = 0;
  if (c < g.cg)
    for (int j = r; j < hp.nslots; j += nr) {
      const long long o = ((long long)(hp.first_slot + j) * g.cg + c) * 2;
      re += pp[o];
      im += pp[o + 1];
    }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int w = nr / 2; w > 0; w >>= 1) {
    if (r < w) {
      sre[threadIdx.x] += sre[threadIdx.x + w];
      sim[threadIdx.x] += sim[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (r == 0 && c < g.cg) {
    cplx<R> o; o.x = (R)sre[threadIdx.x]; o.y = (R)sim[threadIdx.x];
    grid[(long long)hp.g * g.nc + g.c0 + c] = o;
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

// k_user[orig[s] + c*M] = ks[s*nc + c]
template <typename R>
__global__ void k_permute_out(const int* __restrict__ orig, int nnz, long long M, int nc,
                              const cplx<R>* __restrict__ ks, cplx<R>* __restrict__ ku) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nnz * nc) return;
  const long long s = i / nc;
  const int c = (int)(i % nc);
  ku[orig[s] + (long long)c * M] = ks[i];
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

// The same gather through LDS, one work-group per 64 sorted samples: coil by
// coil, 64 lanes read the rows orig[s] of one [M] plane (runs of consecutive
// rows within a bin share cache lines), then the 64 x nc tile is written as
// 64 contiguous sample rows. nc is taken in chunks of at most 32 coils.
template <typename R>
__global__ void __launch_bounds__(256) k_permute_in_tile(const int* __restrict__ orig, int nnz,
                                                         long long M, int nc,
                                                         const cplx<R>* __restrict__ ku,
                                                         const R* __restrict__ wsamp,
                                                         cplx<R>* __restrict__ ks) {
  constexpr int TS = 64, CC = 32;
  __shared__ cplx<R> t[TS][CC + 1];
  __shared__ int so[TS];
  __shared__ R sw[TS];
  const long long s0 = (long long)blockIdx.x * TS;
  const int ns = nnz - s0 < TS ? (int)(nnz - s0) : TS;
  if (threadIdx.x < ns) {
    so[threadIdx.x] = orig[s0 + threadIdx.x];
    sw[threadIdx.x] = wsamp ? wsamp[s0 + threadIdx.x] : (R)1;
  }
  for (int c0 = 0; c0 < nc; c0 += CC) {
    const int cc = nc - c0 < CC ? nc - c0 : CC;
    __syncthreads();   // so/sw written; the previous chunk's tile is consumed
    for (int i = threadIdx.x; i < TS * cc; i += blockDim.x) {
      const int c = i / TS, j = i % TS;
      if (j < ns) {
        cplx<R> v = ku[so[j] + (long long)(c0 + c) * M];
        if (wsamp) { v.x *= sw[j]; v.y *= sw[j]; }
        t[j][c] = v;
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ns * cc; i += blockDim.x) {
      const int j = i / cc, c = i % cc;
      ks[(s0 + j) * nc + c0 + c] = t[j][c];
    }
  }
}

// The same permutation driven by the caller's rows, one work-group per 64
// consecutive rows: coil by coil, 64 lanes read one coalesced run of a [M]
// plane, then each in-bounds row's nc values are written as one contiguous
// sample row at its sorted position inv[row] (out-of-bounds rows have -1).
// Every sorted position is written once; values and weights as k_permute_in.
constexpr int kPermRows = 256;   // caller rows per k_permute_in_rows work-group (= its threads)

template <typename R>
__global__ void __launch_bounds__(256) k_permute_in_rows(const int* __restrict__ inv, long long M,
                                                         int nc, const cplx<R>* __restrict__ ku,
                                                         const R* __restrict__ wsamp,
                                                         cplx<R>* __restrict__ ks) {
  // 256 caller rows per work-group, 8 coils per pass: each thread's loads of
  // a pass are issued together, and a row's 8 coils leave as one 64-byte
  // (float) or 128-byte (double) write
  constexpr int TS = kPermRows, CC = 8;
  __shared__ cplx<R> t[TS][CC + 1];
  __shared__ int si[TS];
  __shared__ R sw[TS];
  const long long r0 = (long long)blockIdx.x * TS;
  const int nr = M - r0 < TS ? (int)(M - r0) : TS;
  if (threadIdx.x < nr) {
    const int s = inv[r0 + threadIdx.x];
    si[threadIdx.x] = s;
    sw[threadIdx.x] = (wsamp && s >= 0) ? wsamp[s] : (R)1;
  }
  for (int c0 = 0; c0 < nc; c0 += CC) {
    const int cc = nc - c0 < CC ? nc - c0 : CC;
    __syncthreads();   // si/sw written; the previous pass's tile is consumed
    cplx<R> v[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) {   // row threadIdx.x of coil c0 + c
      const int j = threadIdx.x;
      if (c < cc && j < nr) v[c] = ku[r0 + j + (long long)(c0 + c) * M];
    }
#pragma unroll
    for (int c = 0; c < CC; ++c)
      if (c < cc && (int)threadIdx.x < nr) t[threadIdx.x][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CC; ++k) {   // rows 32k .. 32k + 31, one coil per thread
      const int i = threadIdx.x + k * TS;
      const int j = i / CC, c = i % CC;
      if (j >= nr || c >= cc) continue;
      const int s = si[j];
      if (s < 0) continue;
      cplx<R> w = t[j][c];
      if (wsamp) { w.x *= sw[j]; w.y *= sw[j]; }
      ks[(long long)s * nc + c0 + c] = w;
    }
  }
}

// The inverse permutation driven by the caller's rows: each work-group reads
// the sorted sample rows inv[row] of 64 consecutive caller rows (contiguous
// nc values each; zeros for out-of-bounds rows), then writes every coil
// plane as one coalesced run of 64 rows.
template <typename R>
__global__ void __launch_bounds__(256) k_permute_out_rows(const int* __restrict__ inv, long long M,
                                                          int nc, const cplx<R>* __restrict__ ks,
                                                          cplx<R>* __restrict__ ku) {
  constexpr int TS = 64, CC = 32;
  __shared__ cplx<R> t[TS][CC + 1];
  __shared__ int si[TS];
  const long long r0 = (long long)blockIdx.x * TS;
  const int nr = M - r0 < TS ? (int)(M - r0) : TS;
  if (threadIdx.x < nr) si[threadIdx.x] = inv[r0 + threadIdx.x];
  for (int c0 = 0; c0 < nc; c0 += CC) {
    const int cc = nc - c0 < CC ? nc - c0 : CC;
    __syncthreads();   // si written; the previous chunk's tile is consumed
    for (int i = threadIdx.x; i < nr * cc; i += blockDim.x) {
      const int j = i / cc, c = i % cc;
      const int s = si[j];
      cplx<R> v;
      v.x = 0;
      v.y = 0;
      if (s >= 0) v = ks[(long long)s * nc + c0 + c];
      t[j][c] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TS * cc; i += blockDim.x) {
      const int c = i / TS, j = i % TS;
      if (j < nr) ku[r0 + j + (long long)(c0 + c) * M] = t[j][c];
    }
  }
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
  k.large_b0 = a >= 3.75;
  {
    const double y = 3.75 / a;
    const double P0 = 0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 + y * (-0.157565e-2 +
                      y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
                      y * (-0.1647633e-1 + y * 0.392377e-2)))))));
    k.inv_P0f = (float)(1.0 / P0);
    k.inv_P0 = 1.0 / P0;
  }
  k.inv_half2 = 1.0 / k.half2;
  k.inv_b0 = 1.0 / b0;
  k.c375a = 3.75 / a;
  k.qb = a >= 3.75 ? 1.0 - (3.75 / a) * (3.75 / a) : -1.0;
  k.tscale = a >= 3.75 ? kKbIntervals / k.qb : 0.0;
  k.cy = (a / 3.75) * (a / 3.75);
  k.radial = p->radial;
  return k;
}

// The kernel table replaces the literal evaluation in plans computing in
// fp64 behind fp32 arrays (their accuracy bar is 1e-4); fp64 plans keep
// kernel.m's arithmetic for the 1e-12 bar.
static const double* kb_table_of(const nufft3d_plan* p) {
  return (p->prec == 2 && p->io_prec == 1 && !(p->tune_flags & 256)) ? p->d_kbtab : nullptr;
}

static GridGeom make_geom(const nufft3d_plan* p, int nc, int c0, int cg) {
  GridGeom g;
  g.K1 = p->K[0]; g.K2 = p->K[1]; g.K3 = p->K[2];
  g.W1 = p->W[0]; g.W2 = p->W[1]; g.W3 = p->W[2];
  g.nc = nc; g.c0 = c0; g.cg = cg;
  g.nnz = (int)p->nnz;
  g.M = p->M;
  g.flags = p->tune_flags;
  g.xchunk = p->xcd_chunk;
  return g;
}

// the block-staged forward of one coil group of at most 8 coils
static void launch_interp_blk(nufft3d_plan* p, const void* grid, int nc, int c0, int cg, void* out,
                              const void* wsamp, double scale, bool user, bool narrow) {
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  const bool wide = p->fk.G == 256;
  ProfScope ps(p, wide ? "k_interp_blkw" : "k_interp_blk");
  const dim3 grid_((unsigned)(wide ? p->fk.ngrp : (p->fk.ngrp + 3) / 4)), blk_(256);
  const double* tb = kb_table_of(p);
  const bool stw = p->fk.d_w && (p->tune_flags & 16384);
#define N3D_IBLK(RO, USR, TB, SW)                                                                 \
  do {                                                                                        \
    if (wide)                                                                                 \
      hipLaunchKernelGGL((k_interp_blkw<RO, USR, TB, SW>), grid_, blk_, 0, p->stream,         \
                         p->fk.d_cbase, p->fk.ngrp, p->fk.d_gpt, p->fk.d_eoff, p->fk.d_slot,  \
                         p->fk.d_w, p->d_base, (const double*)p->d_off, (const double2*)grid, \
                         (cplx<RO>*)out, (const double*)wsamp, scale, p->d_orig, tb, g, kp);  \
    else                                                                                      \
      hipLaunchKernelGGL((k_interp_blk<RO, USR, TB, SW>), grid_, blk_, 0, p->stream,          \
                         p->fk.d_cbase, p->fk.ngrp, p->fk.d_gpt, p->fk.d_eoff, p->fk.d_slot,  \
                         p->fk.d_w, p->d_base, (const double*)p->d_off, (const double2*)grid, \
                         (cplx<RO>*)out, (const double*)wsamp, scale, p->d_orig, tb, g, kp);  \
  } while (0)
#define N3D_IBLK2(RO, USR)                                                                        \
  do {                                                                                        \
    if (stw) N3D_IBLK(RO, USR, true, true);                                                   \
    else if (tb) N3D_IBLK(RO, USR, true, false);                                              \
    else N3D_IBLK(RO, USR, false, false);                                                     \
  } while (0)
  if (narrow) N3D_IBLK2(float, true);
  else if (user) N3D_IBLK2(double, true);
  else N3D_IBLK2(double, false);
#undef N3D_IBLK2
#undef N3D_IBLK
}

template <typename R, int CPL, int LPS, bool COMP, bool FULL>
static void launch_interp(nufft3d_plan* p, const void* grid, int nc, int c0, int cg, void* out,
                          const void* wsamp, double scale, bool user, bool narrow) {
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  const long long nthreads = p->nnz * (long long)LPS;
  const int nb = (int)((nthreads + 255) / 256);
  ProfScope ps(p, "k_interp");
  if (narrow)   // float samples in the caller's layout out of a double grid
    hipLaunchKernelGGL((k_interp<R, CPL, LPS, COMP, FULL, true, float>), dim3(nb), dim3(256), 0,
                       p->stream, p->d_spos, p->d_fcode, p->d_base, (const double*)p->d_off,
                       p->d_off32, (const cplx<R>*)grid, (cplx<float>*)out, (const R*)wsamp,
                       (R)scale, p->d_orig, kb_table_of(p), g, kp);
  else if (user)
    hipLaunchKernelGGL((k_interp<R, CPL, LPS, COMP, FULL, true>), dim3(nb), dim3(256), 0, p->stream,
                       p->d_spos, p->d_fcode, p->d_base, (const double*)p->d_off, p->d_off32,
                       (const cplx<R>*)grid, (cplx<R>*)out, (const R*)wsamp, (R)scale, p->d_orig,
                       kb_table_of(p), g, kp);
  else
    hipLaunchKernelGGL((k_interp<R, CPL, LPS, COMP, FULL, false>), dim3(nb), dim3(256), 0, p->stream,
                       p->d_spos, p->d_fcode, p->d_base, (const double*)p->d_off, p->d_off32,
                       (const cplx<R>*)grid, (cplx<R>*)out, (const R*)wsamp, (R)scale, p->d_orig,
                       kb_table_of(p), g, kp);
}

// The exponential branch of kernel.m (bessi0(ax)/bessi0(alpha) with ax =
// alpha*sqrt(1-q), both >= 3.75) in extended precision: exp(ax - alpha) *
// sqrt(alpha/ax) * P(3.75/ax) / P(3.75/alpha), P the polynomial of
// kernel.m:28-30.
static long double kb_large_ld(long double q, long double alpha) {
  auto P = [](long double z) {
    const long double c[9] = {0.39894228L, 0.1328592e-1L, 0.225319e-2L, -0.157565e-2L,
                              0.916281e-2L, -0.2057706e-1L, 0.2635537e-1L, -0.1647633e-1L,
                              0.392377e-2L};
    long double r = 0;
    for (int k = 8; k >= 0; --k) r = r * z + c[k];
    return r;
  };
  const long double ax = alpha * sqrtl(1 - q);
  return expl(ax - alpha) * sqrtl(alpha / ax) * P(3.75L / ax) / P(3.75L / alpha);
}

// kb_tab_d's table: on each interval the polynomial through the values at
// kKbCoef Chebyshev nodes (Vandermonde system in t, solved in long double),
// stored coefficient-major (coefficient j of interval i at j*kKbIntervals + i).
int build_kb_table(nufft3d_plan* p) {
  if (p->d_kbtab) return NUFFT3D_OK;
  std::vector<double> tab((size_t)kKbIntervals * kKbCoef, 0.0);
  const long double a = p->alpha;
  if (a >= 3.75L) {
    const long double qb = 1 - (3.75L / a) * (3.75L / a);
    const long double h = qb / kKbIntervals;
    for (int i = 0; i < kKbIntervals; ++i) {
      const long double mid = (i + 0.5L) * h;
      long double V[kKbCoef][kKbCoef + 1];
      for (int k = 0; k < kKbCoef; ++k) {
        const long double t = cosl(3.141592653589793238462643383279502884L * (k + 0.5L) / kKbCoef);
        long double tp = 1;
        for (int j = 0; j < kKbCoef; ++j) { V[k][j] = tp; tp *= t; }
        V[k][kKbCoef] = kb_large_ld(mid + t * h / 2, a);
      }
      for (int c = 0; c < kKbCoef; ++c) {   // Gaussian elimination, partial pivoting
        int piv = c;
        for (int r = c + 1; r < kKbCoef; ++r)
          if (fabsl(V[r][c]) > fabsl(V[piv][c])) piv = r;
        for (int j = 0; j <= kKbCoef; ++j) std::swap(V[c][j], V[piv][j]);
        for (int r = 0; r < kKbCoef; ++r) {
          if (r == c) continue;
          const long double f = V[r][c] / V[c][c];
          for (int j = c; j <= kKbCoef; ++j) V[r][j] -= f * V[c][j];
        }
      }
      for (int j = 0; j < kKbCoef; ++j) tab[(size_t)j * kKbIntervals + i] = (double)(V[j][kKbCoef] / V[j][j]);
    }
  }
  N3D_HIP(dev_malloc(&p->d_kbtab, sizeof(double) * tab.size()));
  N3D_HIP(hipMemcpy(p->d_kbtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  return NUFFT3D_OK;
}

// raise a kernel's dynamic LDS limit once (launches above 64 KiB)
// (one flag per kernel instantiation: instantiations share a function type)
template <typename K>
static void allow_tile_lds(K* kernel, size_t bytes) {
  static std::mutex mu;
  static std::vector<const void*> granted;
  if (bytes <= 65536) return;
  const void* f = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(mu);
  if (std::find(granted.begin(), granted.end(), f) != granted.end()) return;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess)
    granted.push_back(f);
  (void)hipGetLastError();
}

// Tiled forward of one coil group (c0, cg): launches of CG staged coils
// (8 or 4 for fp64 arithmetic, 16 or 8 for fp32; NUFFT3D_TILE_CG picks the
// smaller), 4 lanes per sample.
template <typename R, int CG, bool COMP, bool FASTW>
static void launch_interp_tile_cg(nufft3d_plan* p, const void* grid, int nc, int c0, int cg,
                                  void* out, const void* wsamp, double scale, bool user) {
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  const int NL = p->tile_layers > 0 && p->tile_layers <= kTileEdge ? p->tile_layers : 2;
  const size_t lds = (size_t)NL * kTileEdge * kTileEdge * CG * sizeof(cplx<R>) +
                     (FASTW ? sizeof(double) * kKbIntervals * kKbCoef : 0);
  ProfScope ps(p, "k_interp_tile");
  if (user) {
    auto* k = k_interp_tile<R, CG, 4, COMP, true, FASTW>;
    allow_tile_lds(k, lds);
    hipLaunchKernelGGL(k, dim3((unsigned)p->npiece), dim3(kTileThreads), lds, p->stream,
                       p->d_piece, p->d_spos, p->d_fcode, p->d_base, (const double*)p->d_off,
                       p->d_off32, (const cplx<R>*)grid, (cplx<R>*)out, (const R*)wsamp, (R)scale,
                       p->d_orig, (const double*)p->d_kbtab, g, kp, NL);
  } else {
    auto* k = k_interp_tile<R, CG, 4, COMP, false, FASTW>;
    allow_tile_lds(k, lds);
    hipLaunchKernelGGL(k, dim3((unsigned)p->npiece), dim3(kTileThreads), lds, p->stream,
                       p->d_piece, p->d_spos, p->d_fcode, p->d_base, (const double*)p->d_off,
                       p->d_off32, (const cplx<R>*)grid, (cplx<R>*)out, (const R*)wsamp, (R)scale,
                       p->d_orig, (const double*)p->d_kbtab, g, kp, NL);
  }
}

template <typename R, bool COMP>
static void launch_interp_tile(nufft3d_plan* p, const void* grid, int nc, int c0, int cg, void* out,
                               const void* wsamp, double scale, bool user) {
  constexpr int CGMAX = sizeof(R) == 8 ? 8 : 16;
  const int cgmax = p->tile_cg == CGMAX / 2 ? CGMAX / 2 : CGMAX;
  // fp64 arithmetic behind fp32 arrays: the tabulated weights
  const bool fast = sizeof(R) == 8 && kb_table_of(p) != nullptr;
  for (int a = 0; a < cg; a += cgmax) {
    const int n = cg - a < cgmax ? cg - a : cgmax;
    if (n > CGMAX / 2) {
      if (fast) launch_interp_tile_cg<R, CGMAX, COMP, true>(p, grid, nc, c0 + a, n, out, wsamp, scale, user);
      else launch_interp_tile_cg<R, CGMAX, COMP, false>(p, grid, nc, c0 + a, n, out, wsamp, scale, user);
    } else {
      if (fast) launch_interp_tile_cg<R, CGMAX / 2, COMP, true>(p, grid, nc, c0 + a, n, out, wsamp, scale, user);
      else launch_interp_tile_cg<R, CGMAX / 2, COMP, false>(p, grid, nc, c0 + a, n, out, wsamp, scale, user);
    }
  }
}

// The weight of every hit of the grid-major lists, stored as float bits
// beside its sample id: the value the spread itself would evaluate
// (hit_weight_f in fp32 arithmetic; hit_weight_d in fp64, which is the
// reference's float-rounded weight, so float holds it exactly), computed once
// per plan. The spread then reads one 8-byte entry per hit and neither the
// sample's offsets (16 or 24 bytes) nor the kernel evaluation: bitwise the
// same sums.
template <bool D>
__global__ void k_hit_weights(const uint2* __restrict__ hits, long long nhits,
                              const float4* __restrict__ off32, const double* __restrict__ off,
                              uint2* __restrict__ hw, GridGeom g, KBParams kp) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nhits) return;
  const uint2 h = hits[e];
  int dx, dy, dz;
  decode_hit((unsigned short)h.y, g, dx, dy, dz);
  float w;
  if constexpr (D) w = (float)hit_weight_d(dx, dy, dz, off + 3LL * h.x, kp);
  else w = hit_weight_f(dx, dy, dz, off32[h.x], kp);
  hw[e] = make_uint2(h.x, __float_as_uint(w));
}

// Built with the plan's hit lists (8 bytes per hit).
int build_hit_weights(nufft3d_plan* p) {
  if (p->d_hitw || p->nhits <= 0) return NUFFT3D_OK;
  if (p->prec == 1 ? !p->d_off32 : !p->d_off) return NUFFT3D_OK;
  N3D_HIP(dev_malloc(&p->d_hitw, sizeof(uint2) * p->nhits));
  const GridGeom g = make_geom(p, 1, 0, 1);
  const dim3 grid((unsigned)((p->nhits + 255) / 256));
  if (p->prec == 1)
    hipLaunchKernelGGL(k_hit_weights<false>, grid, dim3(256), 0, p->stream, p->d_hit,
                       (long long)p->nhits, p->d_off32, (const double*)nullptr, p->d_hitw, g,
                       make_kb(p));
  else
    hipLaunchKernelGGL(k_hit_weights<true>, grid, dim3(256), 0, p->stream, p->d_hit,
                       (long long)p->nhits, (const float4*)nullptr, (const double*)p->d_off,
                       p->d_hitw, g, make_kb(p));
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

// The block-staged forward's entries: per hit of sample s (sample-major,
// z-y-x codes), its slot byte in its chunk and, with w, its fp32 weight from
// the function the forward kernels evaluate (so stored and evaluated weights
// are the same floats)
template <bool TB>
__global__ void k_fk_fill(const long long* __restrict__ spos, const unsigned short* __restrict__ fcode,
                          long long nnz, const long long* __restrict__ cbase,
                          const long long* __restrict__ eoff, const unsigned short* __restrict__ hslot,
                          const unsigned short* __restrict__ hk, unsigned char* __restrict__ slot,
                          float* __restrict__ w, const double* __restrict__ off,
                          const double* __restrict__ tab, int G, GridGeom g, KBParams kp) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const long long j = s / G;
  const int t = (int)(s % G), nw = G / 64, wv = t / 64, lane = t % 64;
  const long long c0 = cbase[j], nch = cbase[j + 1] - c0;
  const double* f = off + 3 * s;
  for (long long e = spos[s]; e < spos[s + 1]; ++e) {
    const long long sl = hslot[e];
    const long long c = sl % nch;
    const int k = hk[e];
    const long long q = ((eoff[(c0 + c) * nw + wv] + k / 4) * 64 + lane) * 4 + (k & 3);
    slot[q] = (unsigned char)(sl / nch);
    if (w) {
      int dx, dy, dz;
      decode_hit(fcode[e], g, dx, dy, dz);
      double v;
      if constexpr (TB) v = hit_weight_tab(dx, dy, dz, f[0], f[1], f[2], tab, kp);
      else v = hit_weight_d(dx, dy, dz, f, kp);
      w[q] = (float)v;
    }
  }
}

int fk_fill(nufft3d_plan* p, const unsigned short* hslot, const unsigned short* hk) {
  const long long nnz = p->nnz;
  if (nnz <= 0) return NUFFT3D_OK;
  const GridGeom g = make_geom(p, 1, 0, 1);
  const double* tb = kb_table_of(p);
  const dim3 grid((unsigned)((nnz + 255) / 256));
  if (tb)
    hipLaunchKernelGGL(k_fk_fill<true>, grid, dim3(256), 0, p->stream, p->d_spos, p->d_fcode, nnz,
                       p->fk.d_cbase, p->fk.d_eoff, hslot, hk, (unsigned char*)p->fk.d_slot,
                       (float*)p->fk.d_w, (const double*)p->d_off, tb, p->fk.G, g, make_kb(p));
  else
    hipLaunchKernelGGL(k_fk_fill<false>, grid, dim3(256), 0, p->stream, p->d_spos, p->d_fcode, nnz,
                       p->fk.d_cbase, p->fk.d_eoff, hslot, hk, (unsigned char*)p->fk.d_slot,
                       (float*)p->fk.d_w, (const double*)p->d_off, tb, p->fk.G, g, make_kb(p));
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

// the block-staged adjoint of one coil group of at most 8 coils: float
// samples of the split aNUFT, or double samples (with data weights) of an
// fp64 plan or of iNUFT; its partials take the workspace spread() sized
template <typename R>
static void launch_spread_blk(nufft3d_plan* p, const void* samp, int nc, int c0, int cg, void* grid,
                              void* partial, const void* wsamp) {
  GridGeom g = make_geom(p, nc, c0, cg);
  if (p->bk.T == 256) {
    ProfScope ps(p, "k_spread_blkw");
    hipLaunchKernelGGL((k_spread_blkw<cplx<R>>), dim3((unsigned)p->bk.npiece), dim3(256), 0,
                       p->stream, (const int4*)p->bk.d_piece, (long long)p->bk.npiece,
                       (const int*)p->bk.d_blk, (const long long*)p->bk.d_cbase,
                       (const int*)p->bk.d_samp, (const long long*)p->bk.d_eoff,
                       (const unsigned*)p->bk.d_slot, (const float4*)p->bk.d_w,
                       (const cplx<R>*)samp, (const double*)wsamp, (double2*)grid,
                       (double2*)partial, g);
  } else {
    ProfScope ps(p, "k_spread_blk");
#define N3D_SBLK(CC)                                                                              \
    hipLaunchKernelGGL((k_spread_blk<cplx<R>, CC>), dim3((unsigned)((p->bk.npiece + 3) / 4)),   \
                       dim3(256), 0, p->stream, (const int4*)p->bk.d_piece,                     \
                       (long long)p->bk.npiece, (const int*)p->bk.d_blk,                        \
                       (const long long*)p->bk.d_cbase, (const int*)p->bk.d_samp,               \
                       (const long long*)p->bk.d_eoff, (const unsigned*)p->bk.d_slot,           \
                       (const float4*)p->bk.d_w, (const cplx<R>*)samp, (const double*)wsamp,    \
                       (double2*)grid, (double2*)partial, g)
    if (p->bk.C == 128) N3D_SBLK(128);
    else N3D_SBLK(64);
#undef N3D_SBLK
  }
  if (p->bk.nred > 0) {
    ProfScope ps(p, "k_spread_blk_reduce");
    hipLaunchKernelGGL(k_spread_blk_reduce, dim3((unsigned)p->bk.nred, (unsigned)(p->bk.T / 64)),
                       dim3(512), 0, p->stream,
                       (const int4*)p->bk.d_red, (const int*)p->bk.d_blk, p->bk.T,
                       (const double2*)partial, (double2*)grid, g);
  }
}

template <typename R, int CPL, int LPS, bool COMP, bool FULL>
static void launch_spread(nufft3d_plan* p, const void* samp, int nc, int c0, int cg, void* grid,
                          void* partial, const void* wsamp, bool wide) {
  // wide: float samples into a double grid (the split adjoint), the double
  // accumulators stored unrounded
  using G = typename std::conditional<sizeof(R) == 4, double2, cplx<R>>::type;
  GridGeom g = make_geom(p, nc, c0, cg);
  KBParams kp = make_kb(p);
  const long long ngrid = (long long)p->K[0] * p->K[1] * p->K[2];
  // one lane per point (small fp32 groups, every fp64 group): short dense
  // segments; lane groups: long ones
  const DenseSet& ds = p->dense[LPS == 1 ? 0 : 1];
  constexpr bool kTileable = sizeof(R) == 4 && FULL && LPS * CPL >= 8;
  bool merged = false;   // dense segments summed inside the k_spread launch
  // the plan's per-hit weights, unless tune flag 128
  const uint2* hw = !(p->tune_flags & 128) ? p->d_hitw : nullptr;
  const uint2* hits = hw ? hw : p->d_hit;
  const bool pw = hw != nullptr;
  constexpr bool kTileW = sizeof(R) == 4 && COMP && FULL && (LPS * CPL == 8 || LPS * CPL == 4);
  if (kTileW && wide && p->wl.d_went && hw && (p->tune_flags & 8) && !wsamp && cg == CPL * LPS &&
      p->wl.rmax == ds.rmax) {
    // the dense points' segments in the same launch (512 lanes: LPS per
    // segment), unless tune flag 32 asks for their own launch
    merged = ds.nhseg > 0 && !(p->tune_flags & 32);
    const int nhb = merged ? (int)((((long long)ds.nhseg * LPS + 511) / 512 + 7) / 8 * 8) : 0;
    ProfScope ps(p, "k_spread_tile_w");
    hipLaunchKernelGGL((k_spread_tile_w<(kTileW ? CPL : 4), (kTileW ? LPS : 2)>),
                       dim3((unsigned)(nhb + p->nbrick)), dim3(512), 0, p->stream, p->d_rowptr,
                       p->wl.d_went, p->wl.d_beg, p->wl.d_end, p->wl.d_samp, (const float2*)samp,
                       (double2*)grid, p->d_brick, ds.rmax, g, hw, ds.d_hseg, ds.nhseg, nhb,
                       (double2*)partial, kp);
  } else if (kTileable && !wide && p->bl.d_ent && (p->tune_flags & 8) && !wsamp && cg == CPL * LPS &&
      p->bl.rmax == ds.rmax) {
    ProfScope ps(p, "k_spread_tile");
    hipLaunchKernelGGL((k_spread_tile<(kTileable ? CPL * LPS : 8), COMP>), dim3((unsigned)p->nbrick),
                       dim3(512), 0, p->stream, p->d_rowptr, p->bl.d_ent, p->bl.d_ptr,
                       p->bl.d_samp, p->d_off32, (const float2*)samp, (float2*)grid, p->d_brick,
                       ds.rmax, g, kp);
  } else {
    // brick order covers whole 8x8x8 bricks (points outside the grid exit);
    // the dense points' segments run in the same launch, ahead of the light
    // points, unless tune flag 32 asks for their own launch
    const long long npts = (p->tune_flags & 2) ? p->nbrick * 512 : ngrid;
    const long long nthreads = npts * LPS;
    merged = ds.nhseg > 0 && !(p->tune_flags & 32);
    const int nhb = merged ? (int)((((long long)ds.nhseg * LPS + 255) / 256 + 7) / 8 * 8) : 0;
    ProfScope ps(p, "k_spread");
    if (wide)
      hipLaunchKernelGGL((k_spread<R, CPL, LPS, COMP, FULL, G>),
                         dim3((unsigned)(nhb + (nthreads + 255) / 256)), dim3(256), 0, p->stream,
                         p->d_rowptr, hits, pw, (const double*)p->d_off, p->d_off32,
                         (const cplx<R>*)samp, (const R*)wsamp, (G*)grid, npts, p->d_brick,
                         ds.rmax, g, kp, ds.d_hseg, ds.nhseg, nhb, (cplx<R>*)partial,
                         kb_table_of(p));
    else
      hipLaunchKernelGGL((k_spread<R, CPL, LPS, COMP, FULL>),
                         dim3((unsigned)(nhb + (nthreads + 255) / 256)), dim3(256), 0, p->stream,
                         p->d_rowptr, hits, pw, (const double*)p->d_off, p->d_off32,
                         (const cplx<R>*)samp, (const R*)wsamp, (cplx<R>*)grid, npts, p->d_brick,
                         ds.rmax, g, kp, ds.d_hseg, ds.nhseg, nhb, (cplx<R>*)partial,
                         kb_table_of(p));
  }
  if (ds.nhseg > 0) {
    if (!merged) {
      ProfScope ps(p, "k_spread_heavy");
      const long long nthreads = (long long)ds.nhseg * LPS;
      hipLaunchKernelGGL((k_spread_heavy<R, CPL, LPS, COMP, FULL>),
                         dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, p->stream,
                         ds.d_hseg, ds.nhseg, hits, pw, (const double*)p->d_off,
                         p->d_off32, (const cplx<R>*)samp, (const R*)wsamp, (cplx<R>*)partial, g,
                         kp, kb_table_of(p));
    }
    ProfScope ps(p, "k_spread_hreduce");
    int cgp = 1;
    while (cgp < cg) cgp *= 2;
    const int nthr = cgp * 32 < 256 ? cgp * 32 : 256;
    if (wide)   // double partials (COMP) summed into the double grid
      hipLaunchKernelGGL((k_spread_hreduce<double, false>), dim3((unsigned)ds.nhpt), dim3(nthr), 0,
                         p->stream, ds.d_hpt, (const double2*)partial, (double2*)grid, cgp, g);
    else
      hipLaunchKernelGGL((k_spread_hreduce<R, COMP>), dim3((unsigned)ds.nhpt), dim3(nthr), 0,
                         p->stream, ds.d_hpt, (const cplx<R>*)partial, (cplx<R>*)grid, cgp, g);
  }
}

// coil groups of up to coil_group (16 or 8); fp32 groups of more than 4
// coils put 2 or 4 lanes on one sample / grid point (4 coils each), fp64
// groups of more than 2 coils put 4 lanes on it (1, 2 or 4 coils each).
// split: float samples with a double grid (spread: read, interp: written)
template <typename R, bool COMP, bool SPREAD>
static void run_groups(nufft3d_plan* p, const void* in, int nc, void* out, void* partial,
                       const void* wsamp, double scale, bool user = false, bool split = false) {
  const bool wide = split;
  // the block-staged split adjoint takes 8-coil groups whatever the group size
  // (and the block-staged forward 8-coil groups of fp64 arithmetic)
  const int gmax = (SPREAD && p->bk.d_piece && (p->tune_flags & 4096) && p->prec == 2) ||
                           (!SPREAD && sizeof(R) == 8 && p->fk.d_slot && (p->tune_flags & 8192))
                       ? 8
                       : p->coil_group;
  for (int c0 = 0; c0 < nc; c0 += gmax) {
    const int cg = nc - c0 < gmax ? nc - c0 : gmax;
    // full groups with an even coil stride and offset take aligned vector loads
    const bool even = (nc % 2 == 0) && (c0 % 2 == 0);
#define N3D_GO2(CPL, LPS, FULL)                                                                  \
    do {                                                                                         \
      if constexpr (SPREAD) launch_spread<R, CPL, LPS, COMP, FULL>(p, in, nc, c0, cg, out, partial, wsamp, wide); \
      else launch_interp<R, CPL, LPS, COMP, FULL>(p, in, nc, c0, cg, out, wsamp, scale, user, \
                                                  split && sizeof(R) == 8);                     \
    } while (0)
#define N3D_GO(CPL, LPS)                                                                         \
    do {                                                                                         \
      if (cg == (CPL) * (LPS) && ((CPL) == 1 || even)) N3D_GO2(CPL, LPS, true);                  \
      else N3D_GO2(CPL, LPS, false);                                                             \
    } while (0)
    if constexpr (SPREAD) {
      if (block_staged(p, cg) && (sizeof(R) == 8 || (wide && COMP && !wsamp))) {
        launch_spread_blk<R>(p, in, nc, c0, cg, out, partial, wsamp);
        continue;
      }
    } else {
      if (sizeof(R) == 8 && fwd_block_staged(p, cg)) {
        launch_interp_blk(p, in, nc, c0, cg, out, wsamp, scale, user, split);
        continue;
      }
      if (p->tile_interp && p->d_piece && p->npiece > 0 && !split) {
        launch_interp_tile<R, COMP>(p, in, nc, c0, cg, out, wsamp, scale, user);
        continue;
      }
    }
    if constexpr (sizeof(R) == 8) {
      if (cg <= 1) N3D_GO(1, 1);
      else if (cg <= 2) N3D_GO(2, 1);
      else if (cg <= 4) N3D_GO(1, 4);
      else if (cg <= 8) N3D_GO(2, 4);
      else N3D_GO(4, 4);
    } else {
      if (cg <= 1) N3D_GO(1, 1);
      else if (cg <= 2) N3D_GO(2, 1);
      else if (cg <= 4) N3D_GO(4, 1);
      else if (cg <= 8) N3D_GO(4, 2);
      else N3D_GO(4, 4);
    }
#undef N3D_GO
#undef N3D_GO2
  }
}

template <typename R>
static int interp_t(nufft3d_plan* p, const void* grid, int nc, void* out, const void* wsamp,
                    double scale, bool user, bool narrow = false) {
  if (p->nnz == 0) return NUFFT3D_OK;
  if (!p->d_fcode && p->nhits > 0 && !(sizeof(R) == 8 && p->fk.d_slot && (p->tune_flags & 8192))) {
    set_error("the plan's gather lists were released for the block-staged kernels");
    return NUFFT3D_ERR_INVALID;
  }
  if (sizeof(R) == 4 && p->compensated)
    run_groups<R, true, false>(p, grid, nc, out, nullptr, wsamp, scale, user);
  else
    run_groups<R, false, false>(p, grid, nc, out, nullptr, wsamp, scale, user, narrow);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int permute_out(nufft3d_plan* p, const void* ks, int nc, void* k_user);

// user_layout: the caller's [M nc] array (rows of out-of-bounds samples are
// left to zero_oob); otherwise the sorted coil-interleaved layout
int interp(nufft3d_plan* p, const void* grid, int nc, void* out, int user_layout,
           const void* wsamp, double scale, bool allow_ws, int ep) {
  if (ep == 1 && p->prec == 2) {
    if (!user_layout) {
      set_error("float samples from an fp64 plan are written in the caller's layout only");
      return NUFFT3D_ERR_INVALID;
    }
    return interp_t<double>(p, grid, nc, out, wsamp, scale, true, true);
  }
  if (user_layout && (p->tune_flags & 16) && allow_ws) {
    // interpolate into the sorted layout, then scatter rows (separate pass)
    const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
    N3D_TRY(ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz));
    N3D_TRY(p->prec == 2 ? interp_t<double>(p, grid, nc, p->ws_samp.p, wsamp, scale, false)
                         : interp_t<float>(p, grid, nc, p->ws_samp.p, wsamp, scale, false));
    return permute_out(p, p->ws_samp.p, nc, out);
  }
  return p->prec == 2 ? interp_t<double>(p, grid, nc, out, wsamp, scale, user_layout != 0)
                      : interp_t<float>(p, grid, nc, out, wsamp, scale, user_layout != 0);
}

// samples in the internal layout ks[s*nc + c] (sorted order)
template <typename R>
static int spread_t(nufft3d_plan* p, const void* samp, int nc, void* grid, const void* wsamp) {
  if (!p->d_hitw && !p->d_hit && p->nhits > 0 &&
      !(sizeof(R) == 8 && p->bk.d_piece && (p->tune_flags & 4096))) {
    set_error("the plan's gather lists were released for the block-staged kernels");
    return NUFFT3D_ERR_INVALID;
  }
  const int cgmax = nc < p->coil_group ? nc : p->coil_group;
  void* partial = nullptr;
  const int nseg = std::max(p->dense[0].nhseg, p->dense[1].nhseg);
  size_t pbytes = (size_t)nseg * cgmax * 2 * sizeof(double);
  if (sizeof(R) == 8 && p->bk.d_piece && (p->tune_flags & 4096))   // block-staged pieces
    pbytes = std::max(pbytes, (size_t)p->bk.npslot * p->bk.T * 8 * sizeof(double2));
  if (pbytes > 0) {
    N3D_TRY(ensure(p->ws_part, pbytes));
    partial = p->ws_part.p;
  }
  if (sizeof(R) == 4 && p->compensated) run_groups<R, true, true>(p, samp, nc, grid, partial, wsamp, 1.0);
  else run_groups<R, false, true>(p, samp, nc, grid, partial, wsamp, 1.0);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int spread(nufft3d_plan* p, const void* samp, int nc, void* grid, int user_layout,
           const void* wsamp, int ep) {
  if (ep == 1 && p->prec == 2) {   // split adjoint: float samples, double grid
    if (user_layout) {
      N3D_TRY(ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * sizeof(float2)));
      N3D_TRY(permute_in(p, samp, nc, p->ws_samp.p, wsamp, 1));
      samp = p->ws_samp.p;
      wsamp = nullptr;
    }
    const bool blk = p->bk.d_piece && (p->tune_flags & 4096);
    if (!blk && (!p->d_hitw || (p->tune_flags & 128))) {   // the float samples carry no offsets
      set_error("the split-precision adjoint needs the plan's weight table");
      return NUFFT3D_ERR_INVALID;
    }
    const int cgmax = nc < p->coil_group ? nc : p->coil_group;
    void* partial = nullptr;
    const int nseg = std::max(p->dense[0].nhseg, p->dense[1].nhseg);
    size_t pbytes = (size_t)nseg * cgmax * 2 * sizeof(double);
    if (p->bk.d_piece && (p->tune_flags & 4096))   // block-staged pieces: 64 points x 8 coils
      pbytes = std::max(pbytes, (size_t)p->bk.npslot * p->bk.T * 8 * sizeof(double2));
    if (pbytes > 0) {
      N3D_TRY(ensure(p->ws_part, pbytes));
      partial = p->ws_part.p;
    }
    run_groups<float, true, true>(p, samp, nc, grid, partial, wsamp, 1.0, false, true);
    N3D_HIP(hipGetLastError());
    return NUFFT3D_OK;
  }
  if (user_layout) {
    // gather the caller's [M nc] samples into the sorted coil-interleaved layout
    const size_t esz = p->prec == 2 ? sizeof(double2) : sizeof(float2);
    N3D_TRY(ensure(p->ws_samp, (size_t)(p->nnz > 0 ? p->nnz : 1) * nc * esz));
    N3D_TRY(permute_in(p, samp, nc, p->ws_samp.p, wsamp));
    samp = p->ws_samp.p;
    wsamp = nullptr;
  }
  return p->prec == 2 ? spread_t<double>(p, samp, nc, grid, wsamp)
                      : spread_t<float>(p, samp, nc, grid, wsamp);
}

int zero_oob(nufft3d_plan* p, void* k_user, int nc, int ep) {
  const long long noob = p->M - p->nnz;
  if (noob <= 0) return NUFFT3D_OK;
  const long long n = noob * nc;
  const int nb = (int)((n + 255) / 256);
  ProfScope ps(p, "k_zero_oob");
  if ((ep ? ep : p->prec) == 2)
    hipLaunchKernelGGL((k_zero_oob<double>), dim3(nb), dim3(256), 0, p->stream, p->d_oob, noob,
                       (long long)p->M, nc, (double2*)k_user);
  else
    hipLaunchKernelGGL((k_zero_oob<float>), dim3(nb), dim3(256), 0, p->stream, p->d_oob, noob,
                       (long long)p->M, nc, (float2*)k_user);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int permute_out(nufft3d_plan* p, const void* ks, int nc, void* k_user) {
  const long long n = p->nnz * (long long)nc;
  if (n == 0) return NUFFT3D_OK;
  if (p->d_inv && !(p->tune_flags & 64)) {   // caller rows written in order (zeros out of bounds)
    const unsigned nb = (unsigned)((p->M + 63) / 64);
    ProfScope ps(p, "k_permute_out");
    if (p->prec == 2)
      hipLaunchKernelGGL((k_permute_out_rows<double>), dim3(nb), dim3(256), 0, p->stream, p->d_inv,
                         (long long)p->M, nc, (const double2*)ks, (double2*)k_user);
    else
      hipLaunchKernelGGL((k_permute_out_rows<float>), dim3(nb), dim3(256), 0, p->stream, p->d_inv,
                         (long long)p->M, nc, (const float2*)ks, (float2*)k_user);
    N3D_HIP(hipGetLastError());
    return NUFFT3D_OK;
  }
  const int nb = (int)((n + 255) / 256);
  ProfScope ps(p, "k_permute_out");
  if (p->prec == 2)
    hipLaunchKernelGGL((k_permute_out<double>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                       (int)p->nnz, (long long)p->M, nc, (const double2*)ks, (double2*)k_user);
  else
    hipLaunchKernelGGL((k_permute_out<float>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                       (int)p->nnz, (long long)p->M, nc, (const float2*)ks, (float2*)k_user);
  N3D_HIP(hipGetLastError());
  return NUFFT3D_OK;
}

int permute_in(nufft3d_plan* p, const void* k_user, int nc, void* ks, const void* wsamp, int ep) {
  const long long n = p->nnz * (long long)nc;
  if (n == 0) return NUFFT3D_OK;
  ProfScope ps(p, "k_permute_in");
  const int pr = ep ? ep : p->prec;
  if (nc >= 4 && p->d_inv && !(p->tune_flags & 64)) {   // caller rows read in order
    const unsigned nb = (unsigned)((p->M + kPermRows - 1) / kPermRows);
    if (pr == 2)
      hipLaunchKernelGGL((k_permute_in_rows<double>), dim3(nb), dim3(256), 0, p->stream, p->d_inv,
                         (long long)p->M, nc, (const double2*)k_user, (const double*)wsamp,
                         (double2*)ks);
    else
      hipLaunchKernelGGL((k_permute_in_rows<float>), dim3(nb), dim3(256), 0, p->stream, p->d_inv,
                         (long long)p->M, nc, (const float2*)k_user, (const float*)wsamp,
                         (float2*)ks);
    N3D_HIP(hipGetLastError());
    return NUFFT3D_OK;
  }
  if (nc >= 4) {   // tiled: coalesced rows on both sides
    const unsigned nb = (unsigned)((p->nnz + 63) / 64);
    if (pr == 2)
      hipLaunchKernelGGL((k_permute_in_tile<double>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                         (int)p->nnz, (long long)p->M, nc, (const double2*)k_user,
                         (const double*)wsamp, (double2*)ks);
    else
      hipLaunchKernelGGL((k_permute_in_tile<float>), dim3(nb), dim3(256), 0, p->stream, p->d_orig,
                         (int)p->nnz, (long long)p->M, nc, (const float2*)k_user,
                         (const float*)wsamp, (float2*)ks);
    N3D_HIP(hipGetLastError());
    return NUFFT3D_OK;
  }
  const int nb = (int)((n + 255) / 256);
  if (pr == 2)
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
