// Internal plan state and helpers shared by the HIP sources.
//
// Device layouts (all device-resident, owned by the plan):
//   image   : [nc][N3][N2][N1] complex, x fastest (MATLAB column-major
//             [N1 N2 N3 nc], the reference's layout, nufft_3d.m:186)
//   grid    : [K3][K2][K1][nc] complex, coil fastest, so one grid point of
//             every coil is one contiguous run (nc*8 B) for the gather/spread
//   samples : sorted by spread tile; ks[s*nc + c] internally, user arrays are
//             [nc][M] (sample fastest) and reached through orig[s]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/nufft3d.h"

namespace nufft3d {

// Every device allocation and release of the library goes through these.
// HIP synchronises the device in hipFree, which invalidates a stream capture
// in progress on another thread, so allocations and releases wait while any
// thread captures a graph (capture_scope holds the lock shared), and a
// capture waits for them. A thread that is itself capturing allocates
// without the lock (its own capture then fails and falls back).
std::shared_mutex& alloc_capture_mutex();
extern thread_local bool t_capturing;
template <class T>
inline hipError_t dev_malloc(T** ptr, size_t bytes) {
  if (t_capturing) return hipMalloc(reinterpret_cast<void**>(ptr), bytes);
  std::unique_lock<std::shared_mutex> lk(alloc_capture_mutex());
  return hipMalloc(reinterpret_cast<void**>(ptr), bytes);
}
inline hipError_t dev_free(void* ptr) {
  if (t_capturing) return hipFree(ptr);
  std::unique_lock<std::shared_mutex> lk(alloc_capture_mutex());
  return hipFree(ptr);
}
inline hipError_t dev_stream_destroy(hipStream_t st) {
  if (t_capturing) return hipStreamDestroy(st);
  std::unique_lock<std::shared_mutex> lk(alloc_capture_mutex());
  return hipStreamDestroy(st);
}
struct capture_scope {
  std::shared_lock<std::shared_mutex> lk;
  capture_scope() : lk(alloc_capture_mutex()) { t_capturing = true; }
  ~capture_scope() { t_capturing = false; }
};

void set_error(const std::string& msg);

#define N3D_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::nufft3d::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return NUFFT3D_ERR_DEVICE;                                               \
    }                                                                          \
  } while (0)

#define N3D_TRY(call)          \
  do {                         \
    int s_ = (call);           \
    if (s_ != NUFFT3D_OK) return s_; \
  } while (0)

template <typename R> struct cplx_t;
template <> struct cplx_t<float> { using type = float2; };
template <> struct cplx_t<double> { using type = double2; };
template <typename R> using cplx = typename cplx_t<R>::type;

// 1-D FFT description for one axis (length L = product of radices)
struct FftAxis {
  int L = 1;
  int nstage = 0;
  int radix[24] = {0};
  int plan_id = 0;        // compiled stage plan (0 = generic stage loop)
  void* d_tw = nullptr;   // L roots of unity exp(-2 pi i t / L), precision of the plan
  // Bluestein (chirp-z) transform for a length with a prime factor above 13:
  // the stages above describe the smooth length L2 >= 2L - 1 of the
  // convolution, d_tw holds L2 roots; d_chirp the L values exp(-i pi n^2/L),
  // d_bhat[0] / [1] the L2-point DFT of the conjugate chirp, wrapped, over L2
  // for the forward (-) and inverse (+) transform.
  int bs = 0;
  int L2 = 0;
  void* d_chirp = nullptr;
  void* d_bhat[2] = {nullptr, nullptr};
};

// A dense grid point (k-space centre) whose hit list is split into segments;
// the segments' partial sums are added in segment order
struct HeavySeg {
  long long start, end;   // hit range of this segment
  int g;                  // linear grid index
  int slot;               // partial slot
};
struct HeavyPt {
  int g;
  int first_slot;
  int nslots;
  int pad;
};

// Grid points with more than rmax hits are summed as segments of chunk_len
// hits (k_spread_heavy) whose partials add in segment order (k_spread_hreduce)
struct DenseSet {
  int rmax = 512;
  int chunk_len = 512;
  HeavySeg* d_hseg = nullptr;
  int nhseg = 0;
  HeavyPt* d_hpt = nullptr;
  int nhpt = 0;
};

// LDS-staged spread of fp32-arithmetic 3-D plans: for each 8x8x8 brick of
// the traversal order, the ascending ids of the samples that touch its
// points with at most rmax hits, and for each hit of those points one word
// (slot of the sample in its brick's list << 15 | candidate code). Built
// for the large-coil-group dense set (its rmax); absent (d_ent null) when a
// brick would need more than 2^17 slots or the plan has 2^31 hits or more.
struct BrickLists {
  int rmax = 0;
  long long* d_ptr = nullptr;   // [nbrick + 1] offsets into d_samp
  int* d_samp = nullptr;        // [total] sample ids, ascending per brick
  unsigned* d_ent = nullptr;    // [nhits] packed slot/code (entries of dense points unused)
  long long total = 0;
  int max_len = 0;
};

// Lists of the LDS-staged split-precision adjoint (fp32 samples into an fp64
// grid), built one z-slab of bricks at a time so that any hit count works:
// brick b stages samples samp[beg[b] .. end[b]) (ascending ids), and every
// hit of a point with at most rmax hits holds (slot in its brick's list,
// fp32 weight bits) in the plan's hit order.
struct WideLists {
  int rmax = 0;
  long long* d_beg = nullptr;   // [nbrick]
  long long* d_end = nullptr;   // [nbrick]
  int* d_samp = nullptr;        // [total]
  uint2* d_went = nullptr;      // [nhits] (entries of dense points unused)
  long long total = 0;
  long long max_len = 0;
  bool tried = false;
};

// Block-staged adjoint (k_spread_blk): the grid in 4x4x4 blocks, one
// wavefront per block and one lane per point. A block's samples (the
// distinct samples that touch any of its points, ascending ids) are dealt
// round-robin into chunks of C samples: sample i of the list goes to chunk
// i % nch at slot i / nch, so each chunk is an even subsample of the block's
// neighbourhood and a point meets about the same share of its hits in every
// chunk (contiguous chunks leave most lanes idle: a run of sorted samples is
// one small region). A chunk's sample rows pass through the wave's LDS; its
// entries are groups of 4 steps x 64 lanes (4 slot bytes per lane in one
// word, 4 fp32 weights per lane in one float4), lanes with fewer hits
// padded with slot 255 and weight 0. Blocks with more entry rows than a
// piece holds are split into pieces whose partial sums add in piece order.
// Wide lists (k_spread_blkw): 8x8x4 blocks of four 4x4x4 wavefronts that
// share each staged chunk of 255 samples (slot 255 stays "none"); entries
// per (chunk, wavefront), a chunk's rows counted as its largest wavefront's.
constexpr int kBlkPieceRows = 256;  // entry rows (steps) per piece, at least one chunk
constexpr int kBlkWideRows = 255;   // samples per chunk of the wide lists
struct BlockLists {
  int C = 64;                       // samples staged per chunk (64 or 128; kBlkWideRows wide)
  int T = 64;                       // points per block: 64 (4x4x4) or 256 (8x8x4, wide)
  long long nblk = 0;               // blocks, in launch order (Z-order of block coordinates)
  int* d_blk = nullptr;             // [nblk] x | y << 10 | z << 20 (block coordinates)
  long long* d_cbase = nullptr;     // [nblk + 1] first chunk of each block
  long long nchunk = 0;
  int* d_samp = nullptr;            // [nchunk * C] sample ids (padding: the chunk's first)
  long long* d_eoff = nullptr;      // [nchunk * T/64 + 1] first entry group of each (chunk, wavefront)
  unsigned* d_slot = nullptr;       // [groups * 64] 4 slots per lane (byte k: step 4j+k; 255 none)
  float4* d_w = nullptr;            // [groups * 64] 4 weights per lane (0 for padding)
  int4* d_piece = nullptr;          // [npiece] (block, first chunk, end chunk, partial slot or -1)
  long long npiece = 0;
  int4* d_red = nullptr;            // [nred] (block, first partial slot, pieces, 0)
  long long nred = 0;
  long long npslot = 0;
  long long ngroups = 0;
  bool tried = false;
};

// Block-staged forward (k_interp_blkw / k_interp_blk) of fp64-arithmetic 3-D
// plans: sorted samples in groups of 256 (four wavefronts; 64, one
// wavefront, where a group's union could pass 16-bit slots), one lane per
// sample. A group's grid points (the distinct points its samples touch,
// ascending index) are dealt round-robin into chunks of 255 (64) as the
// adjoint's samples are; a chunk's grid rows (8 coils of fp64) pass through
// LDS and each lane walks its entries of the chunk (4 slot bytes per lane per
// word, 255 = none), per (chunk, wavefront). The weights are evaluated from
// the lane's own sample offsets and the staged point's coordinates, as the
// gather kernel evaluates them, or read as stored at plan time.
struct FwdLists {
  int G = 64;                       // sorted samples per group (64: one wavefront; 256: four
                                    // wavefronts sharing each staged chunk)
  int C = 64;                       // grid rows per chunk (64; 255 for groups of 256)
  long long ngrp = 0;               // sample groups
  long long* d_cbase = nullptr;     // [ngrp + 1] first chunk of each group
  long long nchunk = 0;
  int* d_gpt = nullptr;             // [nchunk * 64] grid point indices (padding: the chunk's first)
  long long* d_eoff = nullptr;      // [nchunk * G/64 + 1] first entry group of each (chunk, wavefront)
  unsigned* d_slot = nullptr;       // [groups * 64] 4 slots per lane (byte k: step 4j+k; 255 none)
  float4* d_w = nullptr;            // [groups * 64] the hits' fp32 weights (optional: tune flag 16384)
  long long ngroups = 0;
  bool tried = false;
};

struct Workspace {
  void* p = nullptr;
  size_t bytes = 0;
};

// Optional per-kernel timing: HIP events recorded on the plan's stream
// around each launch while profiling is enabled.
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
struct Profiler {
  bool on = false;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
};

}  // namespace nufft3d

struct nufft3d_plan {
  int prec = 1;             // arithmetic: 1 fp32, 2 fp64
  int io_prec = 1;          // caller arrays: 1 complex64, 2 complex128
  int fp64_mode = -1;       // precision-1 plans: -1 auto, 0 never, 1 always fp64 arithmetic
  double J = 4, u = 2, alpha = 0;
  int radial = 1;
  int N[3] = {0, 0, 0};
  int K[3] = {0, 0, 0};
  int W[3] = {0, 0, 0};     // candidate half range per axis (ceil(J/2); 0 on a 2-D z axis)
  int64_t M = 0;            // samples passed (rows of H)
  int64_t nnz = 0;          // in-bounds samples
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;     // recorded on the plan's stream as each entry point returns
  bool own_stream = false;
  int verbose = 0;
  int compensated = 1;      // fp32 plans: products accumulated in double
  int tune_flags = 2;       // gridding launch layout (see GridGeom::flags; 4 = linear tile order,
                            // 8 = LDS-staged spread where the plan has brick lists,
                            // 16 = forward result through a separate permutation pass,
                            // 32 = dense-point segments in their own launch,
                            // 64 = aNUFT input gathered by sorted sample, not scattered
                            // by caller row, 128 = the spread evaluates each hit's
                            // weight instead of reading the plan's weight table,
                            // 256 = kernel.m's literal arithmetic for the weights of
                            // fp64 arithmetic behind fp32 arrays, not the kernel table,
                            // 512 = no split-precision adjoint (see split_precision),
                            // 1024 = dense-point segments in (segment index, brick)
                            // order instead of by their first sample, 2048 = the
                            // spread's points x-fastest within a brick, not Z-order,
                            // 4096 = the block-staged split adjoint where the plan
                            // has its lists, 8192 = the block-staged forward where
                            // the plan has its lists, 16384 = that forward reads the
                            // lists' stored weights instead of evaluating them)
  int xcd_chunk = 0;        // blocks per XCD run in the gridding block renumbering (0 = off)
  int coil_group = 16;      // coils per gridding launch (16 or 8)
  int fft_bundle = 0;       // lines per FFT work-group (0 = automatic)
  int fft_xthreads = 0;     // threads per x-pass FFT work-group (256, 512; 0 = automatic)
  int fft_ythreads = 256;   // threads per y-pass FFT work-group (256 or 512)
  int fft_zthreads = 0;     // threads per z-pass FFT work-group (256, 512; 0 = automatic)

  // sorted, in-bounds samples
  int* d_orig = nullptr;    // [nnz] original row of sorted sample s
  int* d_oob = nullptr;     // [M-nnz] rows that are out of bounds
  int* d_inv = nullptr;     // [M] sorted position of caller row i (-1: out of bounds)
  int4* d_base = nullptr;   // [nnz] wrapped base (round(u*om) mod K) + tile id
  void* d_off = nullptr;    // double [3*nnz]: u*om - round(u*om) (exact)

  // deapodization and density
  void* d_U = nullptr;      // [N1*N2*N3] real
  void* d_dens = nullptr;   // [nnz] real, sorted order
  std::vector<double> h_U;  // fp64 copy of U (host)

  // sample bins (samples are sorted by bin for locality)
  int T[3] = {8, 8, 8};
  int ntile[3] = {1, 1, 1};
  int ntiles = 0;

  // interpolation structure (the sparsity pattern of H, weights on the fly):
  //   forward : per sorted sample s, hits [spos[s], spos[s+1]) as candidate
  //             codes (dx,dy,dz biased by W, 5 bits each) in z-y-x order
  //   adjoint : per grid point g, hits [rowptr[g], rowptr[g+1]) as sorted
  //             sample ids (ascending) with the same codes alongside
  int64_t nhits = 0;
  long long* d_spos = nullptr;      // [nnz + 1]
  unsigned short* d_fcode = nullptr;  // [nhits] forward codes (sample-major)
  long long* d_rowptr = nullptr;    // [K1*K2*K3 + 1]
  uint2* d_hit = nullptr;           // [nhits] (sample id, code) of each hit (grid-major):
                                    // one 8-byte load per hit in the spread
  uint2* d_hitw = nullptr;          // [nhits] (sample id, weight as float bits) of each hit,
                                    // in d_hit's order
  float4* d_off32 = nullptr;        // [nnz] (fx, fy, fz, 0)
  // dense-point segmentations of the same hit lists: [0] for launches with
  // one lane per point (fp32 groups of at most 4 coils, every fp64 group:
  // short walks), [1] for lane groups (fp32 groups of more than 4 coils)
  nufft3d::DenseSet dense[2];
  int* d_brick = nullptr;           // [nbrick] 8x8x8 spread bricks in traversal order (tx|ty<<10|tz<<20)
  long long nbrick = 0;
  nufft3d::BrickLists bl;           // LDS-staged spread lists over the same bricks
  std::vector<int> h_brick_of;      // brick (x,y,z) -> traversal index, for building bl
  bool bl_tried = false;
  nufft3d::WideLists wl;            // LDS-staged lists of the split-precision adjoint
  nufft3d::BlockLists bk;           // block-staged lists of the split-precision adjoint
  int blk_chunk = 64;               // samples per chunk of the block-staged adjoint's lists
  int blk_wide = 1;                 // block-staged adjoint over 8x8x4 blocks (0: 4x4x4)
  int blk_piece = 0;                // entry rows per piece of the block-staged adjoint (0: automatic)
  int fwd_group = 256;              // samples per group of the block-staged forward's lists
                                    // (64 where a group's union could pass 16-bit slots)
  nufft3d::FwdLists fk;             // block-staged lists of the fp64-arithmetic forward
  int max_hits = 0;

  // LDS-tiled forward interpolation (3-D plans with a 7-point footprint,
  // ceil(J/2) = 3): samples sorted by 4x4x4 sub-bins; each piece is up to
  // 128 samples of one sub-bin, whose 10x10x10 grid neighbourhood a
  // work-group stages in LDS. x = sub-bin origin (tx | ty << 10 | tz << 20, in
  // sub-bins), y = first sorted sample, z = sample count.
  int tile_interp = 0;
  int tile_cg = 0;                  // coils staged per tiled launch (0 = 8 fp64 / 16 fp32)
  int tile_layers = 0;              // z-layers staged per pass (0 = 2)
  bool tile_user = false;           // the caller fixed the sample bin size (opts.tile)
  int4* d_piece = nullptr;
  long long npiece = 0;
  double* d_kbtab = nullptr;        // piecewise-polynomial kernel table (fp64 behind fp32 arrays)

  nufft3d::FftAxis ax[3];
  nufft3d::FftAxis axN[3];  // image-size axes (unshifted FFT of the PRUNO calibration)
  nufft3d::Profiler prof;

  // cached HIP graph of one block of MINRES iterations and what it captured
  int use_graphs = 1;
  hipGraphExec_t mr_exec = nullptr;
  std::vector<uint64_t> mr_key;
  long long graphs_captured = 0, graphs_failed = 0;   // MINRES block captures
  int capture_fail_at = 0;  // test hook: a MINRES block capture fails after this many
                            // iterations (NUFFT3D_TEST_CAPTURE_FAIL; 0 = never)

  // scratch, grown on demand
  nufft3d::Workspace ws_grid, ws_t1, ws_t2, ws_samp, ws_part, ws_img, ws_host, ws_solve, ws_red;
  nufft3d::Workspace ws_b;  // iNUFT right-hand side A'W raw
  std::vector<std::string> warnings;   // the reference's constructor warnings, in order
  std::mutex mu;
};

namespace nufft3d {

// Split precision: a plan with fp32 arrays that computes in fp64 (the
// deapodization range makes fp32 rounding before the final U too large)
// runs its standalone aNUFT on the caller's fp32 samples as given: the
// reference's fp32-rounded weights times fp32 data are exact in double, so
// the spread accumulates them in double into an fp64 grid and the inverse
// transform runs in fp64, with no widened copy of the samples. (The forward
// stays fp64: U multiplies the image before the transform, and with energy
// near the field-of-view edge an fp32 transform or grid loses up to the
// deapodization range.) Tune flag 512 turns it off.
inline bool split_precision(const nufft3d_plan* p) {
  return p->prec == 2 && p->io_prec == 1 && !(p->tune_flags & (128 | 512));
}

// the block-staged adjoint (k_spread_blk) of an fp64-arithmetic plan runs
// for this launch: its lists are built, tune flag 4096 asks for it, a coil
// group of at most 8
inline bool block_staged(const nufft3d_plan* p, int cg) {
  return p->bk.d_piece && (p->tune_flags & 4096) && cg >= 1 && cg <= 8 && p->prec == 2;
}
int ensure_block_lists(nufft3d_plan* p);   // builds them once if the plan qualifies
// the block-staged forward (k_interp_blk) runs for this launch: its lists
// are built, tune flag 8192 asks for it, a coil group of at most 8
inline bool fwd_block_staged(const nufft3d_plan* p, int cg) {
  return p->fk.d_slot && (p->tune_flags & 8192) && cg >= 1 && cg <= 8 && p->prec == 2;
}
int ensure_fwd_lists(nufft3d_plan* p, bool weights);
int fk_fill(nufft3d_plan* p, const unsigned short* hslot, const unsigned short* hk);   // grid.hip

int ensure(Workspace& w, size_t bytes);
void prof_begin(nufft3d_plan* p, const char* name);
void prof_end(nufft3d_plan* p);

// RAII bracket of one kernel launch for the profiler
struct ProfScope {
  nufft3d_plan* p;
  ProfScope(nufft3d_plan* q, const char* name) : p(q) { prof_begin(p, name); }
  ~ProfScope() { prof_end(p); }
};
int ensure_host(Workspace& w, size_t bytes);

// grid.hip
// allow_ws = false never takes the plan's workspaces (the direct caller-row write)
// ep = 1 on an fp64-arithmetic plan: the samples are written as float
// (caller layout only), each the double result rounded once
int interp(nufft3d_plan* p, const void* grid, int nc, void* out, int user_layout,
           const void* wsamp, double scale, bool allow_ws = true, int ep = 0);
// ep = 1 on an fp64-arithmetic plan: float samples, the plan's float
// weight table, double accumulation into a double grid (the split adjoint)
int spread(nufft3d_plan* p, const void* samp, int nc, void* grid, int user_layout,
           const void* wsamp, int ep = 0);
int zero_oob(nufft3d_plan* p, void* k_user, int nc, int ep = 0);
int build_hit_weights(nufft3d_plan* p);   // per-hit fp32 weights of the spread lists
int build_kb_table(nufft3d_plan* p);      // kb_tab_d's table for the plan's J and alpha
int permute_in(nufft3d_plan* p, const void* k_user, int nc, void* ks, const void* wsamp,
               int ep = 0);

// fft.hip
int fft_setup_axis(nufft3d_plan* p, int axis);
// deap, pad, FFT; t1/t2 = caller-given intermediates ([K1 N2 N3 nc], [K1 K2 N3 nc]
// complex of the arithmetic precision) or NULL for the plan's workspaces
// ep = 1 on an fp64-arithmetic plan: the image is float (read and widened,
// or rounded once on the way out; no add term)
int fft_forward(nufft3d_plan* p, const void* img, int nc, void* grid, void* t1 = nullptr,
                void* t2 = nullptr, int ep = 0);
int fft_adjoint(nufft3d_plan* p, const void* grid, int nc, void* img, double scale,
                const void* img_add, double add_coef, int ep = 0);        // IFFT, crop, deap
int fft_plain(nufft3d_plan* p, const void* in, int nc, void* out);       // unshifted N-size FFT

// solve.hip
int minres_solve(nufft3d_plan* p, const void* b, int nc, int maxit, double damp,
                 const void* wsamp, double tol, void* x, nufft3d_info* info, bool warm = false);
int diff_norm2(nufft3d_plan* p, const void* a, const void* c, int64_t n, double* out);
int apply_normal(nufft3d_plan* p, const void* x, int nc, double damp, const void* wsamp,
                 void* y);
int sum_real(nufft3d_plan* p, const void* v, int64_t n, double* out);
int minres_coupled(nufft3d_plan* p, const void* b, int nc, int maxit, double damp,
                   const void* wsamp, const void* B, double tol, void* x, nufft3d_info* info);
// phase constraint (iNUFT.m:144-166): smoothed phase P of x, y = P x or conj(P) x,
// and CG with real inner products on conj(P)(A'WA + damp^2)P + i lambda^2 Im
int phase_estimate(nufft3d_plan* p, const void* x, int nc, void* P);
int phase_mul(nufft3d_plan* p, const void* P, const void* x, void* y, int nc, bool conj);
int phase_cg(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, double lambda,
             const void* wsamp, const void* P, double tol, void* x, nufft3d_info* info);
// compressed sensing (iNUFT.m:170-189): FISTA with a Haar sparsity prox that
// zeroes a fraction `frac` of the coefficients; all coils as one vector
int fista_solve(nufft3d_plan* p, const void* b, int nc, int maxit, double damp, double frac,
                const void* wsamp, double tol, void* x, nufft3d_info* info);

// gram.hip
int pruno_build(nufft3d_plan* p, const void* kdata, int nc, const nufft3d_pruno_opts& o,
                double lambda, void** dB, int* nv_out, std::vector<double>* sv_desc,
                double* noise_out, void* kt_host = nullptr);
int gram_device(nufft3d_plan* p, const void* data, const int dims[4], const std::vector<int4>& taps,
                double2* dAA, bool loraks = false, const int* center = nullptr);
void kernel_taps(int width, int radial, bool is2d, std::vector<int4>& taps, int& w);
int null_apply(nufft3d_plan* p, const void* B, const void* x, void* y, int nc);
int heev_device(nufft3d_plan* p, double2* dA, int n, double* dW);   // ascending, vectors in dA
int pruno_select(const std::vector<double>& evals_desc, double noise_in, double density,
                 int nx, int ny, double tol, double& noise_out, std::vector<double>& f, int& nv);
int lag_field(nufft3d_plan* p, const double2* G, int w, long long cc, const int N[3], double scale,
              void* dB);

// sake.hip: the Cadzow loop on device-resident data [M nc] (caller layout,
// plan precision) with sample weights wsamp; out [nc][N3][N2][N1].
int sake_run(nufft3d_plan* p, const void* data, int nc, const nufft3d_sake_opts& o,
             const void* wsamp, void* out, double* history, nufft3d_sake_info* info);

}  // namespace nufft3d
