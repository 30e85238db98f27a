/* nufft3d.h -- C ABI of the MI355X-native (gfx950 HIP) 3-D NUFFT.
 *
 * Drop-in boundary for the reference class @nufft_3d (MATLAB). A MEX gateway
 * binds these entry points one to one; each declaration names the reference
 * interface it replaces.
 *
 * Conventions
 *   - om is [3, M] double, column-major (x,y,z of one sample adjacent), in
 *     cycles/FOV, exactly the reference trajectory argument (nufft_3d.m:11-12).
 *   - Images are [N1 N2 N3 nc] and k-space is [M nc], column-major,
 *     interleaved complex: complex64 for a precision-1 plan, complex128 for a
 *     precision-2 plan (MATLAB R2018a interleaved complex API).
 *   - Array pointers may be host or device (HIP) memory; the library detects
 *     which. Device pointers avoid the PCIe copies.
 *   - Every call returns NUFFT3D_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is nufft3d_last_error().
 *   - A plan is used by one host thread at a time (calls on one plan are
 *     serialised by an internal mutex).
 */
#ifndef NUFFT3D_H
#define NUFFT3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUFFT3D_OK 0
#define NUFFT3D_ERR_INVALID (-1)     /* bad argument: reference error() */
#define NUFFT3D_ERR_NOMEM (-2)       /* device or host allocation failed */
#define NUFFT3D_ERR_DEVICE (-3)      /* HIP runtime error */
#define NUFFT3D_ERR_UNSUPPORTED (-4) /* valid in the reference, not here */

typedef struct nufft3d_plan nufft3d_plan;

/* Constructor name/value pairs of nufft_3d.m:71-76 ('J','u','alpha',
 * 'radial','gpu','d'). precision: 1 = FP32 (reference gpu=1, gpuSparse
 * single), 2 = FP64 (reference gpu=2, gpuArray double). */
typedef struct nufft3d_opts {
  double J;          /* kernel width, default 4 (nufft_3d.m:27) */
  double u;          /* oversampling, default 2 (nufft_3d.m:28) */
  double alpha;      /* Kaiser-Bessel alpha, 0 = Beatty (nufft_3d.m:30,85-92) */
  int32_t radial;    /* 1 = radial kernel (nufft_3d.m:31) */
  int32_t precision; /* 1 = fp32, 2 = fp64 */
  const double* d;   /* optional density [M] (host), nufft_3d.m:227-238 */
  int64_t nd;        /* number of elements of d (0 = compute Pipe density) */
  int32_t tile;      /* spread tile edge in grid points (0 = auto) */
  int32_t chunk;     /* max samples per spread work chunk (0 = auto) */
  int32_t device;    /* HIP device ordinal */
  int32_t verbose;   /* print the reference's plan report to stdout */
  int32_t compensated; /* fp32 plans: 1 = gridding sums of the (exact) fp32
                          weight x data products kept in double, one rounding
                          per output (default); 0 = plain fp32 sums */
  int32_t fp64_compute; /* precision-1 plans: compute in fp64 behind fp32 arrays
                           (-1 auto: when the deapodization range max|U|/min|U|
                           exceeds 256, where fp32 rounding anywhere before the
                           final deapodization is amplified past 1e-4; 0 never;
                           1 always). No counterpart in the reference. */
} nufft3d_opts;

/* Per-right-hand-side result of nufft3d_inverse (MATLAB minres flag/relres). */
typedef struct nufft3d_info {
  int32_t nrhs;
  int32_t iters[64];   /* iterations taken, per coil (first 64 coils) */
  double relres[64];   /* recurrence ||b-Ax||/||b||, per coil */
  double seconds;      /* wall time of the solve */
} nufft3d_info;

/* Fill opts with the reference defaults. */
int nufft3d_opts_default(nufft3d_opts* opts);

/* obj = nufft_3d(om,N,varargin): validation, alpha, K, scaling, bounds
 * (nufft_3d.m:54-127), interpolation set-up (nufft_3d.m:129-212: on-the-fly
 * weights from bin-sorted samples instead of the stored H), deap
 * (nufft_3d.m:224, deap.m) and density (nufft_3d.m:227-238, density.m). */
int nufft3d_plan_create(const double* om, int64_t M, const int32_t N[3],
                        const nufft3d_opts* opts, nufft3d_plan** plan);

/* The warnings the reference constructor issues with warning(...), in order:
 * "value of J=... is not recommended" (nufft_3d.m:79-81), "radial kernel not
 * recommended with J<3 or u<1.5" (nufft_3d.m:93-95), "mean(density) should
 * be 1 (...)" for a user density (nufft_3d.m:235-237). Writes them to buf
 * (may be NULL) separated by newlines, truncated to len-1 bytes, and returns
 * their number. */
int nufft3d_plan_warnings(nufft3d_plan* plan, char* buf, int64_t len);

/* Frees the plan and its device memory (MATLAB clearing the object). */
int nufft3d_plan_destroy(nufft3d_plan* plan);

/* k = obj.fNUFT(x) (fNUFT.m:2-9): x [N1 N2 N3 nc] -> k [M nc]. */
int nufft3d_forward(nufft3d_plan* plan, const void* x, int32_t nc, void* k);

/* k = obj.fNUFT(x) as nufft3d_forward, allocating nothing: no counterpart in
 * the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * x and k must be device arrays. The intermediates live in a fixed 64 MiB
 * device array per GPU, held until the process ends; coils run in groups
 * that fit it, and a plan whose single coil needs more (roughly K1*K2*K3 *
 * 1.75 * 8 bytes for fp32 arithmetic) returns NUFFT3D_ERR_UNSUPPORTED.
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * race. Results equal nufft3d_forward's. */
int nufft3d_forward_noalloc(nufft3d_plan* plan, const void* x, int32_t nc, void* k);

/* x = obj.aNUFT(k) (aNUFT.m:2-9): k [M nc] -> x [N1 N2 N3 nc]. */
int nufft3d_adjoint(nufft3d_plan* plan, const void* k, int32_t nc, void* x);

/* im = obj.iNUFT(raw,maxit,damp,W) (iNUFT.m:2-138,225): raw [M nc];
 * W NULL (density only) or nW = M or nW = ny real weights (multiplied by
 * the density, iNUFT.m:55-74); tol 1e-4 as iNUFT.m:114 (pass <=0 for it).
 * maxit = 0 is regridding (iNUFT.m:119-122). info may be NULL. */
int nufft3d_inverse(nufft3d_plan* plan, const void* raw, int32_t nc, int32_t maxit,
                    double damp, const double* W, int64_t nW, double tol, void* im,
                    nufft3d_info* info);

/* im = obj.iNUFT(raw,maxit,damp,W,'phase-constraint',lambda) (iNUFT.m:144-166,
 * pprojection nufft_3d.m:325-334): the plain MINRES solution x (as
 * nufft3d_inverse), its smoothed phase P = exp(i angle(convn(x,ones(3),'same'))),
 * then maxit iterations (tol) of conjugate gradients with real inner products
 * (the method of the external pcgpc) on conj(P) (A'WA + damp^2 I)(P x') +
 * i lambda^2 Im(x') = conj(P) A'W raw, and im = P x'. One independent solve
 * per coil, each coil with the phase of its own solution (the reference
 * reshapes all coils into one image, which MATLAB rejects for nc > 1).
 * maxit <= 1 runs nufft3d_inverse (iNUFT.m:78-79). info as nufft3d_inverse,
 * for the CG stage. */
int nufft3d_inverse_phase(nufft3d_plan* plan, const void* raw, int32_t nc, int32_t maxit,
                          double damp, const double* W, int64_t nW, double lambda, double tol,
                          void* im, nufft3d_info* info);

/* im = obj.iNUFT(raw,maxit,damp,W,'compressed-sensing',lambda) (iNUFT.m:170-189):
 * b = A'W raw, then maxit iterations (stop when ||x_k - x_{k-1}|| <= tol ||x_k||)
 * of FISTA on (A'WA + damp^2 I) x = b with an orthonormal 3-D Haar sparsity
 * prox, restating the external fistaN and HWT: gradient step 1/L (L = 1.05 x
 * 20 power iterations), soft threshold of the Haar coefficients at the
 * magnitude that zeroes a fraction lambda (0..1) of them (the sparsity
 * convention of sake.m:35), momentum (t_k - 1)/t_{k+1}, zero start. All
 * coils form one vector (iNUFT.m:180). Haar levels: while every
 * non-singleton active length is even, at most 3. maxit <= 1 runs
 * nufft3d_inverse. info->iters[0] / relres[0]: iterations and the last
 * relative change. */
int nufft3d_inverse_cs(nufft3d_plan* plan, const void* raw, int32_t nc, int32_t maxit,
                       double damp, const double* W, int64_t nW, double lambda, double tol,
                       void* im, nufft3d_info* info);

/* damp = obj.discrep(data,stddev) (discrep.m:2-69): discrepancy-principle
 * estimate of the iNUFT damping. data [M nc]; stddev < 0 = not supplied:
 * guessed as median(|t - median(t)|) over t = [real; imag] of the nonzero
 * entries (discrep.m:29-33; an error if every entry is zero). Starting from
 * damp = 0.1, `iters` rounds (<= 0: the reference's 20) of MINRES on
 * A'dA + damp^2 I (tol 1e-4, maxit min(prod(N),20), warm-started from the
 * previous round) each followed by damp *= stddev*sqrt(M*nc)/||fNUFT(x)-data||.
 * stddev_out (may be NULL) receives the stddev used, history (may be NULL,
 * else [iters]) damp after each round. */
int nufft3d_discrep(nufft3d_plan* plan, const void* data, int32_t nc, double stddev,
                    int32_t iters, double* damp, double* stddev_out, double* history);

/* y = iprojection(x,damp,W) with W = density (nufft_3d.m:305-313):
 * A'WA x + damp^2 x on [N1 N2 N3 nc]. */
int nufft3d_normal(nufft3d_plan* plan, const void* x, int32_t nc, double damp, void* y);

/* obj.d (density.m:1-42): d [M] real (float or double per precision). */
int nufft3d_density(nufft3d_plan* plan, void* d);

/* obj.U (deap.m): U [N1 N2 N3] real, always double, host pointer. */
int nufft3d_deap(nufft3d_plan* plan, double* U);

/* obj.nnz (nufft_3d.m:126); obj.K (nufft_3d.m:109); obj.alpha. */
int64_t nufft3d_nnz(const nufft3d_plan* plan);
int nufft3d_get_params(const nufft3d_plan* plan, int32_t K[3], int32_t N[3], double* J,
                       double* u, double* alpha, int32_t* radial, int32_t* precision);

/* Arithmetic precision of the plan's kernels (1 fp32, 2 fp64): differs from
 * the array precision of nufft3d_get_params when fp64_compute applies. No
 * counterpart in the reference. */
int32_t nufft3d_compute_precision(const nufft3d_plan* plan);

/* Device memory held by a plan: the resident plan arrays (sorted samples,
 * hit lists, density, U, twiddles) and the workspaces grown by calls so far
 * (they scale with the largest coil count seen). No counterpart in the
 * reference: a serving loop admits work against an HBM budget with it. */
int nufft3d_device_bytes(const nufft3d_plan* plan, int64_t* plan_bytes, int64_t* workspace_bytes);

/* Free the plan's workspaces (they regrow on the next call). No counterpart
 * in the reference: lets a serving loop give back memory after a large batch. */
int nufft3d_release_workspaces(nufft3d_plan* plan);

/* Bind subsequent work of this plan to a HIP stream (hipStream_t as void*). */
int nufft3d_set_stream(nufft3d_plan* plan, void* stream);

/* Run iNUFT's MINRES iterations as captured HIP graphs (enable = 1, the
 * default) or launch by launch (0). No counterpart in the reference. The
 * library's own device allocations and releases wait while any thread
 * captures, so they do not invalidate a capture; a process whose other
 * code frees device memory (or synchronises this plan's stream) meanwhile
 * may still invalidate one, and that block then runs launch by launch. The
 * results are the same either way. (The library's own destroy and release
 * wait on the plan's completion event, not on a stream it does not own.) */
int nufft3d_set_graphs(nufft3d_plan* plan, int32_t enable);

/* MINRES graph blocks of this plan since its creation: captured (and
 * instantiated) and failed captures (run launch by launch instead). No
 * counterpart in the reference: measurement. */
int nufft3d_graph_stats(const nufft3d_plan* plan, int64_t* captured, int64_t* failed);

/* Options of pruno(obj,data,varargin) (pruno.m:26-41). */
typedef struct nufft3d_pruno_opts {
  int32_t width;           /* kernel width, default 5 (pruno.m:27) */
  int32_t radial;          /* 1 = spherical kernel support (pruno.m:28) */
  double noise;            /* noise std; < 0 = estimate (pruno.m:29,97-107) */
  double tol;              /* singular value filter threshold, 0.01 (pruno.m:30) */
  int32_t literal_first_nv;/* 1 = the reference's selection of the first nv
                              (largest) vectors (pruno.m:128-130); 0 = the nv
                              null-space (smallest) vectors */
} nufft3d_pruno_opts;

int nufft3d_pruno_opts_default(nufft3d_pruno_opts* opts);

/* AA = make_data_matrix(data,opts) (pruno.m:146-179, sake.m:176-208):
 * Hankel Gram of k-space data [dims0 dims1 dims2 dims3=nc] (complex64 for a
 * precision-1 plan, complex128 for precision 2) with a width/radial kernel;
 * AA is complex128 [n n] column-major, n = nc*nk. FP32 plans use split-BF16
 * MFMA with FP32 accumulation; FP64 plans compute in double. */
int nufft3d_gram(nufft3d_plan* plan, const void* data, const int32_t dims[4], int32_t width,
                 int32_t radial, void* AA, int32_t* n);

/* make_data_matrix(data,opts) with opts.loraks = 1 (sake.m:176-224): the
 * 2n x 2n matrix [AA BB; conj(BB) conj(AA)], BB(:,j,:,k) = row_j' *
 * conj(col_k(flip.x,flip.y,flip.z,:)) with the conjugate reflection about the
 * 1-based k-space centre (sake.m:79-88; NULL = [1 1 1]). Same data, width,
 * radial and precision rules as nufft3d_gram; AA is complex128 [2n 2n]
 * column-major and *n receives 2n. */
int nufft3d_gram_loraks(nufft3d_plan* plan, const void* data, const int32_t dims[4], int32_t width,
                        int32_t radial, const int32_t center[3], void* AA, int32_t* n);

/* K = pruno(obj,data) (pruno.m:2-143) on k-space data [N1 N2 N3 nc] of the
 * plan's image size: returns nv, the singular values sv [n] (descending,
 * may be NULL), the noise std used and, in Kt (may be NULL; complex128,
 * column-major, room for n x n), the tap-domain null kernels of
 * pruno.m:124-130: column v < nv is (1 - f(v)) times the selected right
 * singular vector, row c + nc*t for coil c and kernel tap t (taps in
 * column-major order of the width^3 (or width^2) mask, x fastest). The
 * image-space K [N1 N2 N3 nc nv] of pruno.m:132-143 is the zero-padded
 * inverse DFT of these taps placed in the mask cube, times N per axis; the
 * Python and MATLAB bindings build it on the host. */
int nufft3d_pruno(nufft3d_plan* plan, const void* data, int32_t nc,
                  const nufft3d_pruno_opts* opts, int32_t* nv, double* sv, double* noise,
                  void* Kt);

/* iNUFT(raw,maxit,damp,W,'parallel-imaging-pruno',lambda) (iNUFT.m:200-218):
 * MINRES (maxit) or regridding (maxit=0) for calibration, k-space by an
 * unshifted FFT, pruno null kernels scaled by lambda, then `iters` coupled
 * MINRES iterations of A'WA + damp^2 I + sum_v conj(fnull) fnull (the
 * reference passes an undefined preconditioner M, iNUFT.m:218; none is used). */
int nufft3d_inverse_pruno(nufft3d_plan* plan, const void* raw, int32_t nc, int32_t maxit,
                          double damp, const double* W, int64_t nW, double lambda,
                          const nufft3d_pruno_opts* opts, int32_t iters, void* im,
                          nufft3d_info* info);

/* Options of sake(obj,data,varargin) (sake.m:25-35). */
typedef struct nufft3d_sake_opts {
  int32_t width;           /* kernel width, default 4 (sake.m:26) */
  int32_t radial;          /* 1 = spherical kernel support (sake.m:27,63-67) */
  int32_t loraks;          /* 1 = add the conjugate-reflection (phase) block (sake.m:28) */
  double tol;              /* stop when the largest singular value changes by a
                              fraction below tol, default 1e-6 (sake.m:29,155-161) */
  int32_t maxit;           /* maximum Cadzow iterations, default 1000 (sake.m:30) */
  double noise;            /* noise std; < 0 = estimate on iteration 1 (sake.m:31,128-138) */
  int32_t center[3];       /* 1-based k-space centre, default [1 1 1] (sake.m:32,79-88) */
  double damp;             /* Tikhonov damping of the data-consistency solve (sake.m:34,109) */
  double sparsity;         /* wavelet thresholding (sake.m:35,91-93,114-118): needs the
                              external DWT, so > 0 returns NUFFT3D_ERR_UNSUPPORTED */
  int32_t dc_iters;        /* MINRES iterations of the data-consistency iNUFT, 10 (sake.m:109) */
  int32_t output;          /* 0 = k-space ksp (sake.m:1); 1 = the image ifft3(ksp), the
                              result iNUFT's 'parallel-imaging-sake' branch means to
                              return (iNUFT.m:195 returns k-space as the image) */
} nufft3d_sake_opts;

/* Summary of a sake call. */
typedef struct nufft3d_sake_info {
  int32_t iters;           /* Cadzow iterations run */
  int32_t rank;            /* nnz(f) of the last filter (sake.m:290) */
  double noise;            /* noise std used (given or estimated) */
  double sigma;            /* noise floor in singular-value units (sake.m:139) */
  double seconds;
} nufft3d_sake_info;

int nufft3d_sake_opts_default(nufft3d_sake_opts* opts);

/* ksp = sake(obj,data,varargin) (sake.m:1-171,176-284): Cadzow matrix
 * completion from data [M nc] (nc >= 2, sake.m:57-59). Each iteration:
 * data consistency ksp -= fft3(iNUFT(fNUFT(ifft3(ksp)) - data, dc_iters, damp))
 * with W = density; the Hankel Gram (the split-BF16 MFMA Gram for fp32 plans)
 * and its eigen-decomposition (= svd of the PSD Gram); noise floor and the
 * minimum-variance filter f = max(0, 1 - sigma^2/W^2); F = V diag(f) V';
 * anti-diagonal averaging of F applied to ksp. fft3/ifft3 are the unshifted
 * unnormalised 3-D DFT and its inverse; the reference's matrix_density =
 * obj.sd (a property the class lacks, sake.m:96) is min(1, nnz/prod(N)).
 * out [N1 N2 N3 nc] (k-space or image, per opts->output; same precision as
 * data). history (may be NULL, else [maxit][5]) receives per iteration the
 * largest, sum and 2-norm of the singular values (sake.m:152-154), the
 * convergence measure (sake.m:155-160) and nnz(f). */
int nufft3d_sake(nufft3d_plan* plan, const void* data, int32_t nc, const nufft3d_sake_opts* opts,
                 void* out, double* history, nufft3d_sake_info* info);

/* Autotuner, no counterpart in the reference: times `reps` fwd+adj operator
 * pairs of nc coils for each combination of FFT lines per work-group (auto,
 * 8, 4), coils per gridding launch (16, 8), XCD block remap (off, runs of 32,
 * one contiguous eighth of the launch per XCD)
 * and spread kernel (gather in linear or brick order, or the LDS-staged brick
 * spread where the plan has brick lists), then on the fastest of those the
 * z-pass work-group size (256 threads, or the automatic 1024 forward / 512
 * adjoint for planned axes), keeps the fastest on the plan, and
 * writes the timings as JSON to report (may be NULL). Results are unchanged
 * by every choice except the coil grouping, which can move the dense-point
 * segmentation class (rounding). */
int nufft3d_autotune(nufft3d_plan* plan, int32_t nc, int32_t reps, char* report, int64_t len);

/* Measurement hook, no counterpart in the reference (which only prints
 * tic/toc wall times, nufft_3d.m:130,213): while enabled, every kernel launch
 * of the plan is bracketed by HIP events on the plan's stream. */
int nufft3d_profile_enable(nufft3d_plan* plan, int32_t enable);

/* Synchronises the plan's stream and writes a JSON object
 * {"kernel": [launches, total_ms], ...} of the launches since the last read. */
int nufft3d_profile_read(nufft3d_plan* plan, char* buf, int64_t len);

/* Text of the last failure on the calling thread ("" if none). */
const char* nufft3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NUFFT3D_H */
