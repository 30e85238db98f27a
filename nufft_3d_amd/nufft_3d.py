"""``nufft_3d``: the reference ``@nufft_3d`` operator API over the HIP library.

Mirrors the MATLAB class (``@nufft_3d/nufft_3d.m:1-50``) method for method:

* ``nufft_3d(om, N, J=4, u=2, alpha=0, radial=1, gpu=1, d=None)``  constructor
  (``nufft_3d.m:54``; name/value pairs become keyword arguments)
* ``fNUFT(x)``  (``fNUFT.m:2``), ``aNUFT(k)`` (``aNUFT.m:2``)
* ``iNUFT(raw, maxit=20, damp=0, W=None, constraint=None, lambda_=None)``
  (``iNUFT.m:2``)
* properties ``J u N alpha radial gpu K U d nnz`` (``nufft_3d.m:26-43``)

``gpu`` keeps its reference meaning for the device paths: 1 = single
precision (the gpuSparse path), 2 = double precision (the gpuArray path).
``gpu=0`` (MATLAB's CPU sparse path) is not offered: every call runs the HIP
kernels, and a missing library is an error, never a silent fallback.

Arrays: NumPy arrays are host arrays in MATLAB (Fortran) order; torch tensors
on a HIP device are used in place (no copy), in C order with the reversed
shape, i.e. an image is (nc, N3, N2, N1) and k-space is (nc, M). Results come
back in the same kind of array as the input.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings

import numpy as np

from . import _lib

_CTYPE = {1: np.complex64, 2: np.complex128}
_RTYPE = {1: np.float32, 2: np.float64}


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class nufft_3d:
    """Non-uniform 3-D FFT operator (Fessler-style, sparse formulation)."""

    def __init__(self, om, N=None, J=4, u=2, alpha=0, radial=1, gpu=1, d=None,
                 tile=0, chunk=0, device=0, verbose=False, compensated=True, fp64_compute=None):
        if gpu not in (1, 2):
            raise ValueError("gpu must be 1 (single) or 2 (double): the HIP library has no "
                             "CPU path")
        lib = _lib.load()
        om_np = om.detach().cpu().numpy() if _is_torch(om) else np.asarray(om)
        if om_np.ndim < 1 or om_np.shape[0] != 3 or om_np.size == 0 or np.iscomplexobj(om_np):
            raise ValueError("om must be a real array with leading dimension of 3")
        om_np = np.reshape(om_np, (3, -1), order="F")
        if N is None:                                           # nufft_3d.m:65-68
            N = 2 * np.ceil(np.max(np.abs(om_np), axis=1))
            warnings.warn("N argument not supplied - using [%d %d %d]" % tuple(N))
        Nv = np.ones(3) * np.reshape(np.asarray(N, dtype=np.float64), (-1,))
        if Nv.size != 3:
            raise ValueError("N must be scalar or 3-vector of even integers")
        om_c = np.ascontiguousarray(om_np.T.astype(np.float64))  # [M,3] == column-major [3,M]
        M = om_c.shape[0]
        opts = _lib.Opts()
        _lib.check(lib.nufft3d_opts_default(C.byref(opts)))
        opts.J = float(J)
        opts.u = float(u)
        opts.alpha = float(alpha)
        opts.radial = int(bool(radial))
        opts.precision = int(gpu)
        opts.tile = int(tile)
        opts.chunk = int(chunk)
        opts.device = int(device)
        opts.verbose = int(bool(verbose))
        opts.compensated = int(bool(compensated))
        # gpu=1 plans: None = automatic (fp64 arithmetic when the deapodization
        # range exceeds 256), False = always fp32, True = always fp64
        opts.fp64_compute = -1 if fp64_compute is None else int(bool(fp64_compute))
        self._dkeep = None
        if d is not None:
            dd = np.ascontiguousarray(np.reshape(np.asarray(d, dtype=np.float64), (-1,)))
            self._dkeep = dd
            opts.d = dd.ctypes.data_as(C.POINTER(C.c_double))
            opts.nd = dd.size
        Ni = (C.c_int32 * 3)(*[int(v) for v in Nv])
        h = C.c_void_p()
        _lib.check(lib.nufft3d_plan_create(om_c.ctypes.data, M, Ni, C.byref(opts), C.byref(h)))
        self._h = h
        self._lib = lib
        wbuf = C.create_string_buffer(4096)
        nw = lib.nufft3d_plan_warnings(h, wbuf, len(wbuf))   # nufft_3d.m:79-81,93-95,235-237
        if nw > 0:
            for msg in wbuf.value.decode(errors="replace").split("\n"):
                warnings.warn(msg)
        K = (C.c_int32 * 3)()
        Nout = (C.c_int32 * 3)()
        Jd, ud, ad = C.c_double(), C.c_double(), C.c_double()
        rad, prec = C.c_int32(), C.c_int32()
        _lib.check(lib.nufft3d_get_params(h, K, Nout, C.byref(Jd), C.byref(ud), C.byref(ad),
                                          C.byref(rad), C.byref(prec)))
        self.K = tuple(K)
        self.N = tuple(Nout)
        self.J, self.u, self.alpha = Jd.value, ud.value, ad.value
        self.radial = bool(rad.value)
        self.gpu = prec.value
        self.M = M
        self.nnz = int(lib.nufft3d_nnz(h))
        self.compute_precision = int(lib.nufft3d_compute_precision(h))
        self._U = None
        self._d = None

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.nufft3d_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        """(plan bytes, workspace bytes) of device memory held by this plan."""
        pb, wb = C.c_int64(), C.c_int64()
        _lib.check(self._lib.nufft3d_device_bytes(self._h, C.byref(pb), C.byref(wb)))
        return pb.value, wb.value

    def autotune(self, nc=1, reps=3):
        """Pick the fastest launch settings for nc coils on this plan
        (nufft3d_autotune); returns the timings as a dict."""
        import json
        buf = C.create_string_buffer(1 << 16)
        _lib.check(self._lib.nufft3d_autotune(self._h, int(nc), int(reps), buf, len(buf)))
        return json.loads(buf.value.decode())

    def release_workspaces(self):
        """Free the per-call workspaces (they regrow on the next call)."""
        _lib.check(self._lib.nufft3d_release_workspaces(self._h))

    def set_graphs(self, enable):
        """Run iNUFT's MINRES iterations as captured HIP graphs (default) or
        launch by launch, for processes whose other threads free device
        memory while this plan may be capturing (HIP then invalidates the
        capture). Results are the same."""
        _lib.check(self._lib.nufft3d_set_graphs(self._h, 1 if enable else 0))

    def graph_stats(self):
        """(captured, failed) MINRES graph blocks of this plan so far."""
        a, b = C.c_int64(), C.c_int64()
        _lib.check(self._lib.nufft3d_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ------------------------------------------------------------ helpers
    @property
    def cdtype(self):
        return _CTYPE[self.gpu]

    def _bind_stream(self, t):
        import torch
        s = torch.cuda.current_stream(t.device)
        _lib.check(self._lib.nufft3d_set_stream(self._h, C.c_void_p(s.cuda_stream)))

    def _torch_in(self, t, n_inner, what):
        import torch
        want = torch.complex64 if self.gpu == 1 else torch.complex128
        if t.dtype != want:
            t = t.to(want)
        if t.numel() % n_inner:
            raise ValueError("%s has %d elements, not a multiple of %d" % (what, t.numel(), n_inner))
        return t.contiguous()

    # ------------------------------------------------------------ operators
    def fNUFT(self, x):
        """k = A*x (fNUFT.m:2-9): image [N1 N2 N3 nc] -> k [M nc]."""
        nimg = self.N[0] * self.N[1] * self.N[2]
        if _is_torch(x):
            import torch
            xt = self._torch_in(x, nimg, "x")
            nc = xt.numel() // nimg
            k = torch.empty((nc, self.M), dtype=xt.dtype, device=xt.device)
            self._bind_stream(xt)
            _lib.check(self._lib.nufft3d_forward(self._h, C.c_void_p(xt.data_ptr()), nc,
                                                 C.c_void_p(k.data_ptr())))
            return k
        xa = np.asarray(x)
        if xa.size % nimg:
            raise ValueError("x has %d elements, expected a multiple of %d" % (xa.size, nimg))
        nc = xa.size // nimg
        xb = np.asfortranarray(np.reshape(xa, (self.N[0], self.N[1], self.N[2], nc), order="F"),
                               dtype=self.cdtype)
        k = np.empty((self.M, nc), dtype=self.cdtype, order="F")
        _lib.check(self._lib.nufft3d_forward(self._h, xb.ctypes.data, nc, k.ctypes.data))
        return k

    def aNUFT(self, k):
        """x = A'*k (aNUFT.m:2-9): k [M nc] -> image [N1 N2 N3 nc]."""
        if _is_torch(k):
            import torch
            kt = self._torch_in(k, self.M, "k")
            nc = kt.numel() // self.M
            x = torch.empty((nc, self.N[2], self.N[1], self.N[0]), dtype=kt.dtype, device=kt.device)
            self._bind_stream(kt)
            _lib.check(self._lib.nufft3d_adjoint(self._h, C.c_void_p(kt.data_ptr()), nc,
                                                 C.c_void_p(x.data_ptr())))
            return x
        ka = np.asarray(k)
        if ka.size % self.M:
            raise ValueError("k has %d elements, expected a multiple of %d" % (ka.size, self.M))
        nc = ka.size // self.M
        kb = np.asfortranarray(np.reshape(ka, (self.M, nc), order="F"), dtype=self.cdtype)
        x = np.empty((self.N[0], self.N[1], self.N[2], nc), dtype=self.cdtype, order="F")
        _lib.check(self._lib.nufft3d_adjoint(self._h, kb.ctypes.data, nc, x.ctypes.data))
        return x

    def normal(self, x, damp=0.0):
        """A'DA x + damp^2 x with D = density (iprojection, nufft_3d.m:305-313)."""
        nimg = self.N[0] * self.N[1] * self.N[2]
        if _is_torch(x):
            import torch
            xt = self._torch_in(x, nimg, "x")
            y = torch.empty_like(xt)
            self._bind_stream(xt)
            _lib.check(self._lib.nufft3d_normal(self._h, C.c_void_p(xt.data_ptr()),
                                                xt.numel() // nimg, float(damp),
                                                C.c_void_p(y.data_ptr())))
            return y
        xa = np.asarray(x)
        nc = xa.size // nimg
        xb = np.asfortranarray(np.reshape(xa, (self.N[0], self.N[1], self.N[2], nc), order="F"),
                               dtype=self.cdtype)
        y = np.empty_like(xb, order="F")
        _lib.check(self._lib.nufft3d_normal(self._h, xb.ctypes.data, nc, float(damp), y.ctypes.data))
        return y

    def iNUFT(self, raw, maxit=20, damp=0.0, W=None, constraint=None, lambda_=None,
              tol=1e-4, return_info=False, pruno_opts=None, pruno_iters=100, sake_opts=None):
        """Inverse by MINRES on A'WA + damp^2 I (iNUFT.m:2-138,225).

        constraint 'phase-constraint' runs iNUFT.m:144-166 (each coil with the
        phase of its own solution); 'compressed-sensing' runs iNUFT.m:170-189
        (FISTA with a Haar prox, lambda = fraction of zero coefficients);
        'parallel-imaging-pruno' runs iNUFT.m:200-218;
        'parallel-imaging-sake' runs sake with damp and loraks = lambda_ > 0
        (iNUFT.m:191-196) and returns the image ifft3(ksp) (the reference
        returns the k-space itself as the image, iNUFT.m:195,225);
        sake_opts passes further sake options (e.g. maxit)."""
        if maxit is None:
            maxit = 20
        if damp is None:
            damp = 0.0
        if not (isinstance(maxit, (int, np.integer)) and maxit >= 0):
            raise ValueError("maxit must be a nonnegative integer")
        if not (np.isfinite(damp) and damp >= 0):
            raise ValueError("damp must be finite and nonnegative")
        pruno_mode = sake_mode = phase_mode = cs_mode = False
        if constraint:                                          # iNUFT.m:75-91
            if maxit <= 1:
                warnings.warn("phase constraint is only effective when maxit>1")
                constraint = None
            elif constraint not in ("phase-constraint", "compressed-sensing",
                                    "parallel-imaging-sake", "parallel-imaging-pruno"):
                raise ValueError("unknown constraint")
            elif lambda_ is None:
                raise ValueError("lambda must be supplied with %s" % constraint)
            elif constraint == "parallel-imaging-pruno":
                pruno_mode = True
            elif constraint == "parallel-imaging-sake":
                sake_mode = True
            elif constraint == "phase-constraint":
                if not (np.isscalar(lambda_) and np.isfinite(lambda_) and lambda_ >= 0):
                    raise ValueError("lambda must be a finite nonnegative scalar")   # iNUFT.m:90
                phase_mode = True
            elif constraint == "compressed-sensing":
                if not (np.isscalar(lambda_) and 0 <= lambda_ <= 1):
                    raise ValueError("lambda (fraction of zero wavelet coefficients) must be in "
                                     "[0, 1]")
                cs_mode = True
            else:
                raise NotImplementedError(
                    "constraint %r is not implemented on the HIP path" % constraint)
        Wp, nW = None, 0
        if W is not None and np.size(W) > 1:
            if maxit <= 1:
                warnings.warn("weighting is only effective when maxit>1")
            Wa = np.ascontiguousarray(np.reshape(np.asarray(W, dtype=np.float64), (-1,)))
            Wp, nW = Wa, Wa.size
        info = _lib.Info()
        wptr = Wp.ctypes.data if Wp is not None else None
        if sake_mode:
            if _is_torch(raw):
                rs = self._torch_in(raw, self.M, "raw").reshape(-1, self.M)
            else:
                ra = np.asarray(raw)
                if ra.size % self.M or (ra.shape[0] != self.M and (
                        ra.ndim < 2 or ra.shape[0] * ra.shape[1] != self.M)):
                    raise ValueError("raw data leading dimension(s) must be length %d" % self.M)
                rs = np.reshape(ra, (self.M, ra.size // self.M), order="F")
            if (rs.shape[0] if _is_torch(raw) else rs.shape[1]) == 1:
                raise ValueError("sake-low-rank requires multiple coils")     # iNUFT.m:84
            so = dict(sake_opts or {})
            so.update(damp=float(damp), loraks=lambda_ > 0, output="image")
            res = self.sake(rs, return_info=return_info, **so)
            return res
        if pruno_mode:
            return self._inufft_pruno(raw, maxit, damp, wptr, nW, lambda_, pruno_opts or {},
                                      pruno_iters, return_info)
        def solve(rptr, nc, xptr):
            if phase_mode:                                      # iNUFT.m:144-166
                return self._lib.nufft3d_inverse_phase(self._h, rptr, nc, int(maxit), float(damp),
                                                       wptr, nW, float(lambda_), float(tol), xptr,
                                                       C.byref(info))
            if cs_mode:                                         # iNUFT.m:170-189
                return self._lib.nufft3d_inverse_cs(self._h, rptr, nc, int(maxit), float(damp),
                                                    wptr, nW, float(lambda_), float(tol), xptr,
                                                    C.byref(info))
            return self._lib.nufft3d_inverse(self._h, rptr, nc, int(maxit), float(damp), wptr, nW,
                                             float(tol), xptr, C.byref(info))

        if _is_torch(raw):
            import torch
            rt = self._torch_in(raw, self.M, "raw")
            nc = rt.numel() // self.M
            x = torch.empty((nc, self.N[2], self.N[1], self.N[0]), dtype=rt.dtype, device=rt.device)
            self._bind_stream(rt)
            _lib.check(solve(C.c_void_p(rt.data_ptr()), nc, C.c_void_p(x.data_ptr())))
        else:
            ra = np.asarray(raw)
            if ra.shape[0] == self.M:
                nc = ra.size // self.M
            else:
                if ra.ndim < 2 or ra.shape[0] * ra.shape[1] != self.M:
                    raise ValueError("raw data leading dimension(s) must be length %d" % self.M)
                nc = ra.size // self.M
            rb = np.asfortranarray(np.reshape(ra, (self.M, nc), order="F"), dtype=self.cdtype)
            x = np.empty((self.N[0], self.N[1], self.N[2], nc), dtype=self.cdtype, order="F")
            _lib.check(solve(rb.ctypes.data, nc, x.ctypes.data))
        if return_info:
            n = min(info.nrhs if info.nrhs > 0 else nc, 64)
            return x, {"iters": list(info.iters[:n]), "relres": list(info.relres[:n]),
                       "seconds": info.seconds}
        return x

    def discrep(self, data, stddev=None, iters=20, return_info=False):
        """damp by the discrepancy principle (discrep.m:2-69).

        data is [M nc] or [nr ny nc] with nr*ny = M. Without stddev it is
        guessed from the data (median absolute deviation of the real and
        imaginary parts of the nonzero entries) with the reference's warning.
        """
        if stddev is not None:
            if not (np.isscalar(stddev) and np.isfinite(stddev) and stddev >= 0):
                raise ValueError("stddev must be a finite nonnegative scalar")   # discrep.m:28
        iters = int(iters)
        if iters <= 0:
            iters = 20                                                          # discrep.m:54
        hist = np.zeros(iters, dtype=np.float64)
        damp, sd = C.c_double(), C.c_double()
        sarg = -1.0 if stddev is None else float(stddev)
        if _is_torch(data):
            dt = self._torch_in(data, self.M, "data")
            nc = dt.numel() // self.M
            self._bind_stream(dt)
            _lib.check(self._lib.nufft3d_discrep(self._h, C.c_void_p(dt.data_ptr()), nc, sarg,
                                                 iters, C.byref(damp), C.byref(sd),
                                                 hist.ctypes.data))
        else:
            da = np.asarray(data)
            if da.shape[0] != self.M and (da.ndim < 2 or da.shape[0] * da.shape[1] != self.M):
                raise ValueError("data leading dimension(s) must be length %d" % self.M)
            if da.size % self.M:
                raise ValueError("data size %s seems to have too many dimensions" % (da.shape,))
            nc = da.size // self.M
            db = np.asfortranarray(np.reshape(da, (self.M, nc), order="F"), dtype=self.cdtype)
            _lib.check(self._lib.nufft3d_discrep(self._h, db.ctypes.data, nc, sarg, iters,
                                                 C.byref(damp), C.byref(sd), hist.ctypes.data))
        if stddev is None:
            warnings.warn("stddev not supplied. Guessing %.2e." % sd.value)     # discrep.m:33
        if return_info:
            return damp.value, {"stddev": sd.value, "damp_history": hist[:iters].tolist()}
        return damp.value

    # ------------------------------------------------------------ SAKE/PRUNO
    def _popts(self, opts):
        o = _lib.PrunoOpts()
        _lib.check(self._lib.nufft3d_pruno_opts_default(C.byref(o)))
        for k, v in (opts or {}).items():
            if k == "noise":
                v = -1.0 if v is None else float(v)
            if not hasattr(o, k):
                warnings.warn("'%s' is not a valid option." % k)   # pruno.m:37-39
                continue
            setattr(o, k, v)
        return o

    def _host_c(self, a, shape):
        return np.asfortranarray(np.reshape(np.asarray(a), shape, order="F"), dtype=self.cdtype)

    def gram(self, data, width=5, radial=1, loraks=False, center=None):
        """AA = make_data_matrix(data) (pruno.m:146-179, sake.m:176-224):
        complex128 [n n]; with loraks the [2n 2n] matrix
        [AA BB; conj(BB) conj(AA)] with the conjugate reflection about the
        1-based k-space centre (default [1 1 1], sake.m:79-88)."""
        data = np.asarray(data)
        if data.ndim != 4:
            raise ValueError("'data' must be a 4d float array")
        db = self._host_c(data, data.shape)
        dims = (C.c_int32 * 4)(*data.shape)
        n = C.c_int32()
        # size query from the tap count
        from . import taps as _taps
        nk = len(_taps.kernel_taps(width, radial, data.shape[2] == 1))
        m = nk * data.shape[3] * (2 if loraks else 1)
        AA = np.empty((m, m), dtype=np.complex128, order="F")
        if loraks:
            cen = None if center is None else (C.c_int32 * 3)(*[int(c) for c in center])
            _lib.check(self._lib.nufft3d_gram_loraks(self._h, db.ctypes.data, dims, int(width),
                                                     int(radial), cen, AA.ctypes.data, C.byref(n)))
        else:
            _lib.check(self._lib.nufft3d_gram(self._h, db.ctypes.data, dims, int(width),
                                              int(radial), AA.ctypes.data, C.byref(n)))
        assert n.value == AA.shape[0]
        return AA

    def pruno(self, data, return_info=False, output="image", **opts):
        """K = pruno(obj,data,...) (pruno.m:2-143): the image-space null
        kernels K [N1 N2 N3 nc nv] (complex128). output="taps" returns the
        tap-domain kernels Kt [n nv] instead (row c + nc*t, taps in
        ``taps.kernel_taps`` order), which skips the N1*N2*N3*nc*nv
        expansion. return_info=True also returns dict(nv, sv, noise, Kt,
        taps): the null-space size, the singular values (descending) and
        the noise std used."""
        if output not in ("image", "taps"):
            raise ValueError("output must be 'image' or 'taps'")
        data = np.asarray(data)
        nc = data.size // (self.N[0] * self.N[1] * self.N[2])
        db = self._host_c(data, self.N + (nc,))
        o = self._popts(opts)
        from . import taps as _taps
        taps = _taps.kernel_taps(o.width, o.radial, self.N[2] == 1)
        n = len(taps) * nc
        sv = np.empty(n, dtype=np.float64)
        Kt = np.zeros((n, n), dtype=np.complex128, order="F")
        nv = C.c_int32()
        noise = C.c_double()
        _lib.check(self._lib.nufft3d_pruno(self._h, db.ctypes.data, nc, C.byref(o), C.byref(nv),
                                           sv.ctypes.data, C.byref(noise), Kt.ctypes.data))
        Kt = np.ascontiguousarray(Kt[:, :nv.value])
        K = Kt if output == "taps" else _taps.null_kernels_image(Kt, taps, o.width, self.N, nc)
        if not return_info:
            return K
        return K, {"nv": nv.value, "sv": sv, "noise": noise.value, "Kt": Kt, "taps": taps}

    _SAKE_KEYS = ("width", "radial", "loraks", "tol", "maxit", "noise", "center", "W", "damp",
                  "sparsity", "dc_iters", "output")

    def sake(self, data, return_info=False, **opts):
        """ksp = sake(obj,data,...) (sake.m:1-171): Cadzow matrix completion.

        data [M nc] (nc >= 2). Options as sake.m:26-35 (width 4, radial 1,
        loraks 0, tol 1e-6, maxit 1000, noise None = estimate, center [1 1 1],
        damp 0, sparsity 0) plus dc_iters (10, the data-consistency MINRES
        iterations, sake.m:109) and output ("kspace" or "image"). ``W`` is
        accepted and, as in the reference, unused. Returns k-space
        [N1 N2 N3 nc] (or the image), in the array kind of the input.
        """
        o = _lib.SakeOpts()
        _lib.check(self._lib.nufft3d_sake_opts_default(C.byref(o)))
        for k, v in opts.items():
            if k not in self._SAKE_KEYS:
                warnings.warn("'%s' is not a valid option." % k)              # sake.m:42-44
                continue
            if k == "W":
                continue                                                      # never read by sake.m
            if k == "noise":
                v = -1.0 if v is None else float(v)
            elif k == "center":
                if v is None:
                    continue
                o.center = (C.c_int32 * 3)(*[int(c) for c in v])
                continue
            elif k == "output":
                v = {"kspace": 0, "image": 1}.get(v, v)
            elif k == "loraks":
                v = int(bool(v))
            setattr(o, k, v)
        hist = np.zeros((max(int(o.maxit), 1), 5), dtype=np.float64)
        info = _lib.SakeInfo()
        if _is_torch(data):
            import torch
            dt = self._torch_in(data, self.M, "data")
            nc = dt.numel() // self.M
            if nc < 2:
                raise ValueError("'data' must be a float array [npts nc].")
            out = torch.empty((nc, self.N[2], self.N[1], self.N[0]), dtype=dt.dtype, device=dt.device)
            self._bind_stream(dt)
            _lib.check(self._lib.nufft3d_sake(self._h, C.c_void_p(dt.data_ptr()), nc, C.byref(o),
                                              C.c_void_p(out.data_ptr()), hist.ctypes.data,
                                              C.byref(info)))
        else:
            da = np.asarray(data)
            if da.ndim != 2 or da.shape[0] != self.M or da.shape[1] < 2:
                raise ValueError("'data' must be a float array [npts nc].")   # sake.m:57-59
            nc = da.shape[1]
            db = self._host_c(da, (self.M, nc))
            out = np.empty(self.N + (nc,), dtype=self.cdtype, order="F")
            _lib.check(self._lib.nufft3d_sake(self._h, db.ctypes.data, nc, C.byref(o),
                                              out.ctypes.data, hist.ctypes.data, C.byref(info)))
        if return_info:
            h = [dict(norm_inf=r[0], norm_1=r[1], norm_2=r[2], tol=r[3], rank=int(r[4]))
                 for r in hist[:info.iters]]
            return out, {"iters": info.iters, "rank": info.rank, "noise": info.noise,
                         "sigma": info.sigma, "seconds": info.seconds, "history": h}
        return out

    def _inufft_pruno(self, raw, maxit, damp, wptr, nW, lam, opts, iters, return_info):
        ra = np.asarray(raw)
        nc = ra.size // self.M
        rb = np.asfortranarray(np.reshape(ra, (self.M, nc), order="F"), dtype=self.cdtype)
        x = np.empty((self.N[0], self.N[1], self.N[2], nc), dtype=self.cdtype, order="F")
        info = _lib.Info()
        o = self._popts(opts)
        _lib.check(self._lib.nufft3d_inverse_pruno(self._h, rb.ctypes.data, nc, int(maxit),
                                                   float(damp), wptr, nW, float(lam), C.byref(o),
                                                   int(iters), x.ctypes.data, C.byref(info)))
        if return_info:
            return x, {"iters": [info.iters[0]], "relres": [info.relres[0]], "seconds": info.seconds}
        return x

    # ------------------------------------------------------------ properties
    @property
    def U(self):
        """Deapodization matrix [N1 N2 N3] (deap.m), double."""
        if self._U is None:
            U = np.empty(self.N, dtype=np.float64, order="F")
            _lib.check(self._lib.nufft3d_deap(self._h, U.ctypes.data))
            self._U = U
        return self._U

    @property
    def d(self):
        """Density weights [M] (density.m), in the plan precision."""
        if self._d is None:
            d = np.empty(self.M, dtype=_RTYPE[self.gpu])
            _lib.check(self._lib.nufft3d_density(self._h, d.ctypes.data))
            self._d = d
        return self._d
