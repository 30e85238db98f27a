/* nufft3d_mex.c -- MEX gateway from the MATLAB class shim (@nufft_3d) to the
 * C ABI of include/nufft3d.h.
 *
 * Build (MATLAB R2018a or later, interleaved complex API):
 *   mex -R2018a nufft3d_mex.c -I<repo>/include -L<repo>/nufft_3d_amd -lnufft3d
 *
 * Not compiled in this repository's environments (neither has MATLAB's
 * mex.h); the same entry points are exercised by the Python binding's tests.
 *
 * Calling convention: nufft3d_mex(cmd, handle, ...). The handle is a uint64
 * scalar holding the plan pointer. Arrays cross as host arrays: single
 * complex for precision-1 plans (the reference's gpu=1), double complex for
 * precision-2 plans (gpu=2); the class shim casts before calling. A nonzero
 * status becomes an error with identifier nufft3d:e<code> and the library's
 * message (nufft3d_last_error).
 *
 *   h = nufft3d_mex('create', om, N, J, u, alpha, radial, gpu, d)      nufft_3d.m:54-238
 *       nufft3d_mex('destroy', h)
 *   [K, N, J, u, alpha, radial, gpu, nnz] = nufft3d_mex('params', h)
 *   k  = nufft3d_mex('fNUFT', h, x, M)                                  fNUFT.m:2-9
 *   x  = nufft3d_mex('aNUFT', h, k)                                     aNUFT.m:2-9
 *   y  = nufft3d_mex('iprojection', h, x, damp)                         nufft_3d.m:305-313
 *   [im, iters, relres] = nufft3d_mex('iNUFT', h, raw, maxit, damp, W, constraint, lambda)
 *                                                                       iNUFT.m:2-227
 *   ksp = nufft3d_mex('sake', h, data, opts)                            sake.m:1-171
 *   [nv, sv, noise] = nufft3d_mex('pruno', h, data, opts)               pruno.m:2-143
 *   AA = nufft3d_mex('gram', h, data, width, radial, loraks, center)    sake.m:176-224
 *   [damp, stddev] = nufft3d_mex('discrep', h, data, stddev)            discrep.m:2-69
 *   d  = nufft3d_mex('density', h, M)                                   density.m
 *   U  = nufft3d_mex('deap', h)                                         deap.m
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "mex.h"
#include "nufft3d.h"

static void check(int st) {
  if (st != NUFFT3D_OK) {
    char id[32];
    snprintf(id, sizeof id, "nufft3d:e%d", -st);
    mexErrMsgIdAndTxt(id, "%s", nufft3d_last_error());
  }
}

static nufft3d_plan* handle_of(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1)
    mexErrMsgIdAndTxt("nufft3d:handle", "invalid plan handle");
  return (nufft3d_plan*)(uintptr_t)(*mxGetUint64s(a));
}

static double scalar_or(const mxArray* a, double dflt) {
  return (a && !mxIsEmpty(a)) ? mxGetScalar(a) : dflt;
}

typedef struct {
  int32_t K[3], N[3], radial, prec;
  double J, u, alpha;
} params_t;

static params_t params_of(nufft3d_plan* p) {
  params_t q;
  check(nufft3d_get_params(p, q.K, q.N, &q.J, &q.u, &q.alpha, &q.radial, &q.prec));
  return q;
}

/* complex data pointer of an array in the plan's precision (interleaved) */
static void* cdata(const mxArray* a, int prec, const char* what) {
  if (!mxIsComplex(a) || (prec == 1 ? !mxIsSingle(a) : !mxIsDouble(a)))
    mexErrMsgIdAndTxt("nufft3d:type", "'%s' must be %s complex", what,
                      prec == 1 ? "single" : "double");
  return prec == 1 ? (void*)mxGetComplexSingles(a) : (void*)mxGetComplexDoubles(a);
}

static mxArray* new_complex(mwSize nd, const mwSize* dims, int prec) {
  return mxCreateNumericArray(nd, dims, prec == 1 ? mxSINGLE_CLASS : mxDOUBLE_CLASS, mxCOMPLEX);
}

static mxArray* new_image(const params_t* q, int32_t nc) {
  const mwSize d[4] = {(mwSize)q->N[0], (mwSize)q->N[1], (mwSize)q->N[2], (mwSize)nc};
  return new_complex(4, d, q->prec);
}

static int64_t image_elems(const params_t* q) { return (int64_t)q->N[0] * q->N[1] * q->N[2]; }

static double field_or(const mxArray* s, const char* name, double dflt) {
  const mxArray* f = s && mxIsStruct(s) ? mxGetField(s, 0, name) : NULL;
  return scalar_or(f, dflt);
}

/* number of kernel taps: offsets in [-fix(w/2), fix(w/2)]^3 (^2 for 2-D data)
 * within radius w/2 (radial) or the cube (pruno.m:56-71, sake.m:62-68) */
static int32_t count_taps(int32_t width, int32_t radial, int is2d) {
  const int h = width / 2;
  int32_t n = 0;
  for (int z = is2d ? 0 : -h; z <= (is2d ? 0 : h); ++z)
    for (int y = -h; y <= h; ++y)
      for (int x = -h; x <= h; ++x)
        if (!radial || 4.0 * (x * x + y * y + z * z) <= (double)width * width) ++n;
  return n;
}

static void fill_info(const nufft3d_info* info, int nlhs, mxArray* plhs[]) {
  const int n = info->nrhs < 64 ? info->nrhs : 64;
  if (nlhs > 1) {
    plhs[1] = mxCreateDoubleMatrix(1, n, mxREAL);
    for (int i = 0; i < n; ++i) mxGetDoubles(plhs[1])[i] = info->iters[i];
  }
  if (nlhs > 2) {
    plhs[2] = mxCreateDoubleMatrix(1, n, mxREAL);
    for (int i = 0; i < n; ++i) mxGetDoubles(plhs[2])[i] = info->relres[i];
  }
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd))
    mexErrMsgIdAndTxt("nufft3d:cmd", "first argument must be a command string");

  if (!strcmp(cmd, "create")) {                       /* create(om, N, J, u, alpha, radial, gpu, d) */
    if (nrhs < 8) mexErrMsgIdAndTxt("nufft3d:args", "create needs om, N, J, u, alpha, radial, gpu");
    nufft3d_opts o;
    check(nufft3d_opts_default(&o));
    o.J = mxGetScalar(prhs[3]);
    o.u = mxGetScalar(prhs[4]);
    o.alpha = mxGetScalar(prhs[5]);
    o.radial = (int32_t)mxGetScalar(prhs[6]);
    o.precision = (int32_t)mxGetScalar(prhs[7]);
    if (nrhs > 8 && !mxIsEmpty(prhs[8])) {
      o.d = mxGetDoubles(prhs[8]);
      o.nd = (int64_t)mxGetNumberOfElements(prhs[8]);
    }
    if (!mxIsDouble(prhs[1]) || mxIsComplex(prhs[1]) || mxGetM(prhs[1]) != 3)
      mexErrMsgIdAndTxt("nufft3d:om", "om must be a real double array [3 M]");
    const double* Nd = mxGetDoubles(prhs[2]);
    const size_t nN = mxGetNumberOfElements(prhs[2]);
    int32_t N[3];
    for (int a = 0; a < 3; ++a) N[a] = (int32_t)Nd[nN == 1 ? 0 : a];
    nufft3d_plan* p = NULL;
    check(nufft3d_plan_create(mxGetDoubles(prhs[1]), (int64_t)mxGetN(prhs[1]), N, &o, &p));
    char w[4096];
    if (nufft3d_plan_warnings(p, w, sizeof w) > 0)    /* nufft_3d.m:79-81,93-95,235-237 */
      for (char* t = strtok(w, "\n"); t; t = strtok(NULL, "\n")) mexWarnMsgIdAndTxt("nufft3d:warn", "%s", t);
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    *mxGetUint64s(plhs[0]) = (uint64_t)(uintptr_t)p;
    return;
  }

  if (nrhs < 2) mexErrMsgIdAndTxt("nufft3d:args", "missing plan handle");
  nufft3d_plan* p = handle_of(prhs[1]);

  if (!strcmp(cmd, "destroy")) {
    check(nufft3d_plan_destroy(p));
    return;
  }

  const params_t q = params_of(p);

  if (!strcmp(cmd, "params")) {                        /* [K, N, J, u, alpha, radial, gpu, nnz] */
    mxArray* K = mxCreateDoubleMatrix(1, 3, mxREAL);
    mxArray* N = mxCreateDoubleMatrix(1, 3, mxREAL);
    for (int a = 0; a < 3; ++a) {
      mxGetDoubles(K)[a] = q.K[a];
      mxGetDoubles(N)[a] = q.N[a];
    }
    mxArray* out[8] = {K, N, mxCreateDoubleScalar(q.J), mxCreateDoubleScalar(q.u),
                       mxCreateDoubleScalar(q.alpha), mxCreateDoubleScalar(q.radial),
                       mxCreateDoubleScalar(q.prec), mxCreateDoubleScalar((double)nufft3d_nnz(p))};
    for (int i = 0; i < 8; ++i) {
      if (i < (nlhs > 0 ? nlhs : 1)) plhs[i] = out[i];
      else mxDestroyArray(out[i]);
    }
    return;
  }

  if (!strcmp(cmd, "fNUFT")) {                         /* k = fNUFT(h, x, M) */
    const int64_t ni = image_elems(&q);
    const int32_t nc = (int32_t)(mxGetNumberOfElements(prhs[2]) / ni);
    if ((int64_t)nc * ni != (int64_t)mxGetNumberOfElements(prhs[2]) || nc < 1)
      mexErrMsgIdAndTxt("nufft3d:size", "x must have prod(N)*nc elements");
    const mwSize d[2] = {(mwSize)mxGetScalar(prhs[3]), (mwSize)nc};
    plhs[0] = new_complex(2, d, q.prec);
    check(nufft3d_forward(p, cdata(prhs[2], q.prec, "x"), nc, mxGetData(plhs[0])));
    return;
  }

  if (!strcmp(cmd, "aNUFT")) {                         /* x = aNUFT(h, k) with k [M nc] */
    const int32_t nc = (int32_t)mxGetN(prhs[2]);
    plhs[0] = new_image(&q, nc);
    check(nufft3d_adjoint(p, cdata(prhs[2], q.prec, "k"), nc, mxGetData(plhs[0])));
    return;
  }

  if (!strcmp(cmd, "iprojection")) {                   /* y = iprojection(h, x, damp) */
    const int32_t nc = (int32_t)(mxGetNumberOfElements(prhs[2]) / image_elems(&q));
    plhs[0] = new_image(&q, nc);
    check(nufft3d_normal(p, cdata(prhs[2], q.prec, "x"), nc, scalar_or(nrhs > 3 ? prhs[3] : NULL, 0.0),
                         mxGetData(plhs[0])));
    return;
  }

  if (!strcmp(cmd, "iNUFT")) {       /* [im, iters, relres] = iNUFT(h, raw[M nc], maxit, damp, W, constraint, lambda) */
    const int32_t nc = (int32_t)mxGetN(prhs[2]);
    const int32_t maxit = (int32_t)scalar_or(nrhs > 3 ? prhs[3] : NULL, 20);
    const double damp = scalar_or(nrhs > 4 ? prhs[4] : NULL, 0.0);
    const double* W = NULL;
    int64_t nW = 0;
    if (nrhs > 5 && !mxIsEmpty(prhs[5])) {
      W = mxGetDoubles(prhs[5]);
      nW = (int64_t)mxGetNumberOfElements(prhs[5]);
    }
    char con[40] = "";
    if (nrhs > 6 && !mxIsEmpty(prhs[6])) mxGetString(prhs[6], con, sizeof con);
    const double lambda = scalar_or(nrhs > 7 ? prhs[7] : NULL, 0.0);
    const void* raw = cdata(prhs[2], q.prec, "raw");
    nufft3d_info info;
    memset(&info, 0, sizeof info);
    if (!strcmp(con, "parallel-imaging-sake")) {       /* iNUFT.m:191-196 (image, not k-space) */
      nufft3d_sake_opts so;
      check(nufft3d_sake_opts_default(&so));
      so.damp = damp;
      so.loraks = lambda > 0;
      so.output = 1;
      nufft3d_sake_info si;
      plhs[0] = new_image(&q, nc);
      check(nufft3d_sake(p, raw, nc, &so, mxGetData(plhs[0]), NULL, &si));
      info.nrhs = 1;
      info.iters[0] = si.iters;
    } else {
      plhs[0] = new_image(&q, nc);
      void* im = mxGetData(plhs[0]);
      if (!strcmp(con, "phase-constraint"))            /* iNUFT.m:144-166 */
        check(nufft3d_inverse_phase(p, raw, nc, maxit, damp, W, nW, lambda, 1e-4, im, &info));
      else if (!strcmp(con, "compressed-sensing"))     /* iNUFT.m:170-189 */
        check(nufft3d_inverse_cs(p, raw, nc, maxit, damp, W, nW, lambda, 1e-4, im, &info));
      else if (!strcmp(con, "parallel-imaging-pruno")) /* iNUFT.m:200-218 */
        check(nufft3d_inverse_pruno(p, raw, nc, maxit, damp, W, nW, lambda, NULL, 100, im, &info));
      else if (con[0] == 0)
        check(nufft3d_inverse(p, raw, nc, maxit, damp, W, nW, 1e-4, im, &info));
      else
        mexErrMsgIdAndTxt("nufft3d:constraint", "unknown constraint");   /* iNUFT.m:86 */
    }
    fill_info(&info, nlhs, plhs);
    return;
  }

  if (!strcmp(cmd, "sake")) {                          /* ksp = sake(h, data, opts) */
    const mxArray* s = nrhs > 3 ? prhs[3] : NULL;
    nufft3d_sake_opts o;
    check(nufft3d_sake_opts_default(&o));
    o.width = (int32_t)field_or(s, "width", o.width);
    o.radial = (int32_t)field_or(s, "radial", o.radial);
    o.loraks = (int32_t)field_or(s, "loraks", o.loraks);
    o.tol = field_or(s, "tol", o.tol);
    o.maxit = (int32_t)field_or(s, "maxit", o.maxit);
    o.noise = field_or(s, "noise", -1.0);
    o.damp = field_or(s, "damp", o.damp);
    o.sparsity = field_or(s, "sparsity", o.sparsity);
    const mxArray* c = s && mxIsStruct(s) ? mxGetField(s, 0, "center") : NULL;
    if (c && mxGetNumberOfElements(c) == 3)
      for (int a = 0; a < 3; ++a) o.center[a] = (int32_t)mxGetDoubles(c)[a];
    const int32_t nc = (int32_t)mxGetN(prhs[2]);
    plhs[0] = new_image(&q, nc);
    check(nufft3d_sake(p, cdata(prhs[2], q.prec, "data"), nc, &o, mxGetData(plhs[0]), NULL, NULL));
    return;
  }

  if (!strcmp(cmd, "pruno")) {                         /* [nv, sv, noise, Kt] = pruno(h, data, opts) */
    const mxArray* s = nrhs > 3 ? prhs[3] : NULL;
    nufft3d_pruno_opts o;
    check(nufft3d_pruno_opts_default(&o));
    o.width = (int32_t)field_or(s, "width", o.width);
    o.radial = (int32_t)field_or(s, "radial", o.radial);
    o.noise = field_or(s, "noise", -1.0);
    o.tol = field_or(s, "tol", o.tol);
    const int32_t nc = (int32_t)(mxGetNumberOfElements(prhs[2]) / image_elems(&q));
    int32_t nv = 0;
    double noise = 0;
    const mwSize n = (mwSize)nc * count_taps(o.width, o.radial, q.N[2] == 1);
    mxArray* sv = mxCreateDoubleMatrix(n, 1, mxREAL);      /* singular values, descending */
    mxArray* kt = mxCreateDoubleMatrix(n, n, mxCOMPLEX);   /* tap-domain kernels, n x nv used */
    check(nufft3d_pruno(p, cdata(prhs[2], q.prec, "data"), nc, &o, &nv, mxGetDoubles(sv), &noise,
                        mxGetComplexDoubles(kt)));
    mxSetN(kt, (mwSize)nv);
    plhs[0] = mxCreateDoubleScalar(nv);
    if (nlhs > 1) plhs[1] = sv; else mxDestroyArray(sv);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(noise);
    if (nlhs > 3) plhs[3] = kt; else mxDestroyArray(kt);
    return;
  }

  if (!strcmp(cmd, "gram")) {                          /* AA = gram(h, data, width, radial, loraks, center) */
    const mwSize* dd = mxGetDimensions(prhs[2]);
    const mwSize nd = mxGetNumberOfDimensions(prhs[2]);
    int32_t dims[4] = {(int32_t)dd[0], nd > 1 ? (int32_t)dd[1] : 1, nd > 2 ? (int32_t)dd[2] : 1,
                       nd > 3 ? (int32_t)dd[3] : 1};
    const int32_t width = (int32_t)scalar_or(nrhs > 3 ? prhs[3] : NULL, 5);
    const int32_t radial = (int32_t)scalar_or(nrhs > 4 ? prhs[4] : NULL, 1);
    const int loraks = (int)scalar_or(nrhs > 5 ? prhs[5] : NULL, 0);
    const mwSize nmax = (mwSize)dims[3] * count_taps(width, radial, dims[2] == 1) * (loraks ? 2 : 1);
    mxArray* AA = mxCreateDoubleMatrix(nmax, nmax, mxCOMPLEX);
    int32_t n = 0;
    const void* d = cdata(prhs[2], q.prec, "data");
    if (loraks) {
      int32_t cen[3] = {1, 1, 1};
      if (nrhs > 6 && mxGetNumberOfElements(prhs[6]) == 3)
        for (int a = 0; a < 3; ++a) cen[a] = (int32_t)mxGetDoubles(prhs[6])[a];
      check(nufft3d_gram_loraks(p, d, dims, width, radial, cen, mxGetComplexDoubles(AA), &n));
    } else {
      check(nufft3d_gram(p, d, dims, width, radial, mxGetComplexDoubles(AA), &n));
    }
    if ((mwSize)n != nmax) mexErrMsgIdAndTxt("nufft3d:gram", "unexpected Gram size");
    plhs[0] = AA;
    return;
  }

  if (!strcmp(cmd, "discrep")) {                       /* [damp, stddev] = discrep(h, data, stddev) */
    const int32_t nc = (int32_t)mxGetN(prhs[2]);
    const double sd = (nrhs > 3 && !mxIsEmpty(prhs[3])) ? mxGetScalar(prhs[3]) : -1.0;
    double damp = 0, used = 0;
    check(nufft3d_discrep(p, cdata(prhs[2], q.prec, "data"), nc, sd, 20, &damp, &used, NULL));
    if (sd < 0) mexWarnMsgIdAndTxt("nufft3d:discrep", "stddev not supplied. Guessing %.2e.", used);
    plhs[0] = mxCreateDoubleScalar(damp);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(used);
    return;
  }

  if (!strcmp(cmd, "density")) {                       /* d = density(h, M) */
    const mwSize M = (mwSize)mxGetScalar(prhs[2]);
    plhs[0] = mxCreateNumericMatrix(M, 1, q.prec == 1 ? mxSINGLE_CLASS : mxDOUBLE_CLASS, mxREAL);
    check(nufft3d_density(p, mxGetData(plhs[0])));
    return;
  }

  if (!strcmp(cmd, "deap")) {                          /* U = deap(h), double [N1 N2 N3] */
    const mwSize d[3] = {(mwSize)q.N[0], (mwSize)q.N[1], (mwSize)q.N[2]};
    plhs[0] = mxCreateNumericArray(3, d, mxDOUBLE_CLASS, mxREAL);
    check(nufft3d_deap(p, mxGetDoubles(plhs[0])));
    return;
  }

  mexErrMsgIdAndTxt("nufft3d:cmd", "unknown command '%s'", cmd);
}
