"""FP64 CPU oracle of the reference sparse-matrix NUFFT (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it. The shipped path (``nufft_3d_amd``) runs HIP kernels and fails loudly when
its library is missing; it never routes through this file.

It restates, in NumPy/SciPy double precision, the reference ``gpu=0`` path of
``@nufft_3d`` (stored CSR interpolation matrix ``H``, FFT with explicit
pad/shift/crop, Pipe density, MINRES):

* ``kernel`` / ``bessi0``      -> ``@nufft_3d/kernel.m:2-34``
* ``deap``                     -> ``@nufft_3d/deap.m:1-36,97-104``
* ``NufftOracle.__init__``     -> ``@nufft_3d/nufft_3d.m:54-127`` (checks,
  alpha, K, scaling, bounds) and ``:153-212`` (H assembly)
* ``fft3_pad`` / ``ifft3_crop``-> ``nufft_3d.m:285-302``
* ``fNUFT`` / ``aNUFT``        -> ``fNUFT.m:2-9`` / ``aNUFT.m:2-9``
* ``density``                  -> ``density.m:1-42``
* ``iprojection``              -> ``nufft_3d.m:305-322``
* ``minres``                   -> MATLAB ``minres`` as called at ``iNUFT.m:136``
  (Paige-Saunders; per-column independent solves for a matrix RHS)
* ``iNUFT``                    -> ``iNUFT.m:22-138,225``

How each function is pinned (tests/test_oracle.py):
  bessi0/kernel : scipy.special.i0 within the NR polynomial error, exact
                  identities kernel(0)=1, kernel((J/2)^2)=1/bessi0(alpha),
                  imaginary-argument branch against J0 (asymptotic accuracy).
  H             : brute force over every grid point of tiny grids.
  deap          : numpy irfft (the library routine that equals MATLAB's
                  ``ifft(.., 'symmetric')``) of the padded table.
  fft3_pad/crop : explicit DFT sums; exact adjoint pair.
  fNUFT/aNUFT   : adjointness, linearity, NUDFT closed form on tiny N.
  density       : invariants d>=0, d(~ok)=0, sum(d)=nnz, fixed-point property.
  minres        : Krylov-space least squares by QR on tiny systems.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

EPS = np.finfo(np.float64).eps  # MATLAB eps (density.m:151)


# ---------------------------------------------------------------- kernel.m
def bessi0(ax):
    """Numerical-Recipes polynomial I0 with the imaginary-argument x2 hack.

    Restates ``@nufft_3d/kernel.m:14-34``. ``ax`` may be real or complex
    (pure imaginary arguments occur in deap beyond the kernel support).
    """
    ax = np.asarray(ax)
    out = np.zeros(ax.shape, dtype=ax.dtype)
    k = np.abs(ax) < 3.75                                   # kernel.m:19
    y = ax[k] / 3.75                                        # kernel.m:20
    y = y * y                                               # kernel.m:21
    out[k] = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (
        0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))))  # kernel.m:22-23
    kk = ~k                                                 # kernel.m:26
    a = ax[kk]
    y = 3.75 / a                                            # kernel.m:27
    out[kk] = (np.exp(a) / np.sqrt(a)) * (0.39894228 + y * (0.1328592e-1 + y * (
        0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (
            0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))))  # kernel.m:28-30
    k2 = kk & (np.real(ax) == 0)                            # kernel.m:33
    out[k2] = 2 * out[k2]                                   # kernel.m:34
    return out


def kernel(dist2, J, alpha):
    """Kaiser-Bessel weight of a squared distance (``kernel.m:2-11``)."""
    dist2 = np.asarray(dist2, dtype=np.float64)
    t = 1.0 - dist2 / (J / 2.0) ** 2                        # kernel.m:4
    if np.any(t < 0):   # MATLAB sqrt of a negative real returns complex
        t = t.astype(np.complex128)
    arg = alpha * np.sqrt(t)
    s = bessi0(arg)
    b0 = bessi0(np.asarray([alpha], dtype=np.float64))[0]
    return np.real(s) / b0                                  # kernel.m:9,11


def beatty_alpha(J, u, radial):
    """Default alpha (``nufft_3d.m:85-92``)."""
    if radial:
        return math.pi * math.sqrt(J ** 2 * (1 - 0.5 / u) ** 2 - 0.6 * u)
    return math.pi * math.sqrt(J ** 2 * (1 - 0.5 / u) ** 2 - 0.8)


def matlab_round(x):
    """MATLAB ``round``: half away from zero, exact for doubles."""
    x = np.asarray(x, dtype=np.float64)
    r = np.trunc(x)
    f = x - r
    return r + np.where(np.abs(f) >= 0.5, np.sign(x), 0.0)


# ---------------------------------------------------------------- deap.m
def deap_table(J, alpha, radial, is2d=False):
    """Kernel at half-integer offsets, not truncated (``deap.m:5-27``)."""
    nJ = int(math.floor(J))
    nz = 1 if is2d else nJ + 1
    T = np.zeros((nJ + 1, nJ + 1, nz))
    for dx in range(nJ + 1):
        for dy in range(nJ + 1):
            for dz in range(nz):
                dx2 = (dx / 2) ** 2
                dy2 = (dy / 2) ** 2
                dz2 = (dz / 2) ** 2
                if radial:
                    T[dx, dy, dz] = kernel(np.array([dx2 + dy2 + dz2]), J, alpha)[0]
                else:
                    T[dx, dy, dz] = (kernel(np.array([dx2]), J, alpha)[0]
                                     * kernel(np.array([dy2]), J, alpha)[0]
                                     * kernel(np.array([dz2]), J, alpha)[0])
    return T


def deap_axis_matrix(n_img, K, nterms):
    """1-D map table->image of ``ifft(T*K,2K,'symmetric')``+crop+fftshift.

    Row p (image index, x = p - n_img/2) holds 0.5*c_m*cos(pi*m*x/K) with
    c_0 = 1, c_m = 2: the closed form of MATLAB's symmetric inverse FFT of a
    real table padded to 2K (``deap.m:29-36``).
    """
    x = np.arange(n_img) - n_img // 2
    m = np.arange(nterms)
    c = np.where(m == 0, 1.0, 2.0)
    return 0.5 * c[None, :] * np.cos(np.pi * np.outer(x, m) / K)


def deap(N, K, J, alpha, radial):
    """Deapodization ``U = 1/Phi`` on the image grid (``deap.m``)."""
    is2d = N[2] == 1
    T = deap_table(J, alpha, radial, is2d)
    A = deap_axis_matrix(N[0], K[0], T.shape[0])
    B = deap_axis_matrix(N[1], K[1], T.shape[1])
    U = np.einsum('pa,abc->pbc', A, T)
    U = np.einsum('qb,pbc->pqc', B, U)
    if not is2d:
        C = deap_axis_matrix(N[2], K[2], T.shape[2])
        U = np.einsum('rc,pqc->pqr', C, U)
    return 1.0 / U                                          # deap.m:98


# ---------------------------------------------------------------- FFT helpers
def fft3_pad(x, N, K):
    """Zero-pad, fftshift, fft per axis (``nufft_3d.m:285-291``)."""
    for j in range(3):
        if K[j] == 1 and N[j] == 1:
            continue
        pad = (K[j] - N[j]) // 2
        widths = [(0, 0)] * x.ndim
        widths[j] = (pad, pad)
        x = np.pad(x, widths)
        x = np.fft.fft(np.fft.fftshift(x, axes=j), axis=j)
    return x


def ifft3_crop(k, N, K):
    """ifft, ifftshift, xK, crop per axis (``nufft_3d.m:294-302``)."""
    for j in range(3):
        if K[j] == 1 and N[j] == 1:
            continue
        mid = (K[j] - N[j]) // 2
        k = np.fft.ifftshift(np.fft.ifft(k, axis=j), axes=j) * K[j]
        sl = [slice(None)] * k.ndim
        sl[j] = slice(mid, K[j] - mid)
        k = k[tuple(sl)]
    return k


# ---------------------------------------------------------------- the class
class NufftOracle:
    """Restatement of ``@nufft_3d`` with ``gpu=0`` (double CPU, stored H).

    ``om`` is [3, M] in cycles/FOV; ``N`` scalar or 3-vector of even ints.
    """

    def __init__(self, om, N, J=4, u=2, alpha=0, radial=1, d=None, build_H=True):
        om = np.asarray(om)
        if om.ndim < 1 or om.shape[0] != 3 or om.size == 0 or np.iscomplexobj(om):
            raise ValueError('om must be a real array with leading dimension of 3')
        om = np.reshape(om, (3, -1), order='F')
        self.J = float(J)
        self.u = float(u)
        self.radial = bool(radial)
        if self.u < 1 or self.u > 6:                        # nufft_3d.m:82-84
            raise ValueError('value of u=%f is not recommended' % self.u)
        self.alpha = float(alpha) if alpha else beatty_alpha(self.J, self.u, self.radial)
        Nv = np.ones(3) * np.reshape(np.asarray(N, dtype=np.float64), (-1,))
        if Nv.size != 3:
            raise ValueError('N must be scalar or 3-vector of even integers')
        if np.all(Nv == 0) or Nv[0] % 2 or Nv[1] % 2:       # nufft_3d.m:101-103
            raise ValueError('N must be an even integer')
        if np.all(om[2] == 0):                              # nufft_3d.m:104-106
            Nv[2] = 1
        self.N = tuple(int(v) for v in Nv)
        K = 2 * np.ceil(Nv * self.u / 2)                    # nufft_3d.m:109
        if self.N[2] == 1:
            K[2] = 1
        if np.prod(K) >= 2 ** 31 - 1:
            raise ValueError('N or u too large')
        self.K = tuple(int(v) for v in K)
        oms = self.u * om.astype(np.float64)                # nufft_3d.m:120
        K1, K2, K3 = self.K
        ok = (oms[0] >= -K1 / 2) & (oms[0] <= K1 / 2 - 1)    # nufft_3d.m:123-125
        ok = (oms[1] >= -K2 / 2) & (oms[1] <= K2 / 2 - 1) & ok
        ok = ((oms[2] >= -K3 / 2) & (oms[2] <= K3 / 2 - 1) & ok) | (self.N[2] == 1)
        self.ok = ok
        self.nnz = int(np.count_nonzero(ok))
        self.M = om.shape[1]
        self.om_s = oms
        self.H = self._assemble_H() if build_H else None
        self.U = deap(self.N, self.K, self.J, self.alpha, self.radial)
        if d is None:
            self.d = self.density(ok)
        else:
            d = np.reshape(np.asarray(d, dtype=np.float64), (-1,)).copy()
            if d.size != self.M:
                raise ValueError('supplied density has wrong number of elements')
            d[~ok] = 0
            self.d = d

    # nufft_3d.m:153-212. The per-axis squares (x-om)^2 and wrapped indices
    # are formed once per axis offset and combined per (dx,dy,dz): the same
    # values and the same sum order dx2 + dy2 + dz2 as the reference's loop.
    def _assemble_H(self):
        J, K = self.J, self.K
        oms, ok = self.om_s, self.ok
        h = int(math.ceil(J / 2))
        rng = range(-h, h + 1)
        rows, cols, vals = [], [], []
        base = [matlab_round(oms[a]) for a in range(3)]
        stride = (1, K[0], K[0] * K[1])
        d2, idx = [], []
        for a in range(3):
            d2.append({})
            idx.append({})
            for d in rng:
                x = base[a] + d
                d2[a][d] = (x - oms[a]) * (x - oms[a])
                idx[a][d] = np.mod(x, K[a]).astype(np.int64) * stride[a]
        J2 = J ** 2
        for dx in rng:
            for dy in rng:
                sxy = d2[0][dx] + d2[1][dy]
                cxy = idx[0][dx] + idx[1][dy]
                for dz in rng:
                    if self.N[2] == 1 and dz != 0:
                        continue
                    if self.radial:
                        dist2 = sxy + d2[2][dz]
                        i = ok & (4 * dist2 <= J2)
                        r = np.nonzero(i)[0]
                        s = kernel(dist2[r], J, self.alpha)
                    else:
                        i = ok & (4 * d2[0][dx] <= J2) & (4 * d2[1][dy] <= J2) & (4 * d2[2][dz] <= J2)
                        r = np.nonzero(i)[0]
                        s = (kernel(d2[0][dx][r], J, self.alpha) * kernel(d2[1][dy][r], J, self.alpha)
                             * kernel(d2[2][dz][r], J, self.alpha))
                    rows.append(r)
                    cols.append(cxy[r] + idx[2][dz][r])
                    vals.append(s.astype(np.float32).astype(np.float64))  # nufft_3d.m:160,205
        rows = np.concatenate(rows)
        cols = np.concatenate(cols)
        vals = np.concatenate(vals)
        H = sp.csr_matrix((vals, (rows, cols)), shape=(self.M, int(np.prod(K))))
        H.sum_duplicates()
        return H

    # nufft_3d.m:259-282
    def spmv(self, g):
        return self.H @ g

    def spmv_t(self, k):
        return self.H.T @ k

    # fNUFT.m:2-9
    def fNUFT(self, x):
        N, K = self.N, self.K
        x = np.reshape(np.asarray(x, dtype=np.complex128), (N[0], N[1], N[2], -1), order='F')
        g = fft3_pad(x * self.U[..., None], N, K)
        g = np.reshape(g, (int(np.prod(K)), -1), order='F')
        return self.spmv(g) / math.sqrt(self.nnz)

    # aNUFT.m:2-9
    def aNUFT(self, k):
        N, K = self.N, self.K
        k = np.reshape(np.asarray(k, dtype=np.complex128), (self.M, -1), order='F')
        g = self.spmv_t(k)
        g = np.reshape(g, (K[0], K[1], K[2], -1), order='F')
        x = ifft3_crop(g, N, K)
        return x * self.U[..., None] / math.sqrt(self.nnz)

    # density.m:1-42
    def density(self, ok, maxit=20):
        d = np.reshape(ok, (-1,)).astype(np.float64)
        for _ in range(maxit):
            HHd = self.spmv(self.spmv_t(d))
            d = d / np.maximum(HHd, EPS)
        return d / (np.sum(d) / self.nnz)

    # nufft_3d.m:305-313 (coils independent)
    def iprojection(self, x, damp, W):
        N = self.N
        x = np.reshape(x, (N[0], N[1], N[2], -1), order='F')
        y = self.fNUFT(x)
        y = W[:, None] * y
        y = self.aNUFT(y)
        y = y + damp ** 2 * x
        return np.reshape(y, (int(np.prod(N)), -1), order='F')

    # iNUFT.m:22-138,225
    def iNUFT(self, raw, maxit=20, damp=0.0, W=None, tol=1e-4, return_info=False):
        raw = np.asarray(raw, dtype=np.complex128)
        if raw.shape[0] == self.M:
            raw = np.reshape(raw, (self.M, -1), order='F')
        else:
            nr, ny = raw.shape[0], raw.shape[1]
            if nr * ny != self.M:
                raise ValueError('raw data leading dimension(s) must be length %d' % self.M)
            raw = np.reshape(raw, (self.M, -1), order='F')
        nc = raw.shape[1]
        if W is None or np.isscalar(W) or np.size(W) == 0:
            Wv = self.d
        elif maxit <= 1:
            # iNUFT.m:57-58 only warns: W is used as given (no density, no
            # expansion), and W.*raw needs one weight per row
            Wv = np.reshape(np.asarray(W, dtype=np.float64), (-1,))
            if Wv.size != self.M:
                raise ValueError('W must have one element per row of raw when maxit<=1')
        else:
            Wv = np.reshape(np.asarray(W, dtype=np.float64), (-1,))
            if Wv.size != self.M:
                ny = Wv.size
                nr = self.M / ny
                if nr % 1 or nr * ny != self.M:
                    raise ValueError('W must be a vector of length ny or ny*nr')
                Wv = np.tile(Wv[None, :], (int(nr), 1)).reshape(-1, order='F')
            Wv = Wv * self.d
        N = self.N
        if maxit == 0:
            x = self.aNUFT(Wv[:, None] * raw)
            info = {'iters': [0] * nc, 'relres': [float('nan')] * nc}
        else:
            b = self.aNUFT(Wv[:, None] * raw)
            b = np.reshape(b, (int(np.prod(N)), nc), order='F')
            x, info = minres(lambda v: self.iprojection(v, damp, Wv), b, tol, maxit)
        im = np.reshape(x, (N[0], N[1], N[2], nc), order='F')
        return (im, info) if return_info else im

    # discrep.m:2-69: discrepancy-principle estimate of damp
    def discrep(self, data, stddev=None, iters=20, return_info=False, maxit=None):
        data = np.asarray(data, dtype=np.complex128)
        if data.shape[0] == self.M:
            data = np.reshape(data, (self.M, -1), order='F')
        else:
            nr, ny = data.shape[0], data.shape[1]
            if nr * ny != self.M:
                raise ValueError('data leading dimension(s) must be length %d' % self.M)
            data = np.reshape(data, (self.M, -1), order='F')
        if stddev is None:                                  # discrep.m:29-33
            tmp = data[data != 0]
            tmp = np.concatenate([tmp.real, tmp.imag])
            stddev = float(np.median(np.abs(tmp - np.median(tmp))))
        discrepancy = stddev * np.sqrt(data.size)           # discrep.m:37
        W = self.d
        N = self.N
        b = np.reshape(self.aNUFT(W[:, None] * data), (int(np.prod(N)), -1), order='F')
        if maxit is None:                                   # MATLAB minres default
            maxit = min(int(np.prod(N)), 20)
        x = None
        damp = 0.1
        hist = []
        for _ in range(iters):                              # discrep.m:54-68
            x, _ = minres(lambda v: self.iprojection(v, damp, W), b, 1e-4, maxit, x0=x)
            k = self.fNUFT(np.reshape(x, (N[0], N[1], N[2], -1), order='F'))
            resnorm = float(np.linalg.norm(k - data))
            damp = damp * (discrepancy / resnorm)
            hist.append(damp)
        info = {'stddev': stddev, 'discrepancy': discrepancy, 'damp_history': hist}
        return (damp, info) if return_info else damp


# ---------------------------------------------------------------- MINRES
def minres(A, b, tol=1e-4, maxit=20, x0=None):
    """Paige-Saunders MINRES for a Hermitian operator, from x0 (default 0).

    ``b`` is [n, nrhs]; each column is an independent solve with its own
    Lanczos scalars (the matrix-RHS semantics ``iNUFT.m:133-136`` implies).
    ``A`` maps an [n, nrhs] block to an [n, nrhs] block column by column.
    A column stops updating when its recurrence residual drops to
    ``tol*||b||`` (the MATLAB stopping test ``||b-Ax||/||b|| <= tol``). With
    an initial guess (as ``minres(A,b,tol,maxit,[],[],x0)`` in
    ``discrep.m:60``) the recurrence runs on r0 = b - A x0 from x0, and a
    column whose r0 already meets the test takes no iteration. Returns
    (x, info) with per-column iteration counts and relres ||b-Ax||/||b||.
    """
    b = np.asarray(b, dtype=np.complex128)
    n, nr = b.shape
    bnorm = np.sqrt(np.real(np.sum(np.conj(b) * b, axis=0)))
    if x0 is None:
        x = np.zeros_like(b)
    else:
        x = np.array(x0, dtype=np.complex128).reshape(n, nr)
        b = b - A(x)
    beta1 = np.sqrt(np.real(np.sum(np.conj(b) * b, axis=0)))
    bden = np.where(bnorm > 0, bnorm, 1.0)
    x[:, bnorm == 0] = 0              # zero right-hand side: x = 0, no iteration
    active = (bnorm > 0) & (beta1 > 0) & (beta1 / bden > tol)
    r1 = b.copy()
    r2 = b.copy()
    y = b.copy()
    w = np.zeros_like(b)
    w2 = np.zeros_like(b)
    oldb = np.zeros(nr)
    beta = beta1.copy()
    dbar = np.zeros(nr)
    epsln = np.zeros(nr)
    phibar = beta1.copy()
    cs = -np.ones(nr)
    sn = np.zeros(nr)
    iters = np.zeros(nr, dtype=np.int64)
    relres = np.where(bnorm > 0, beta1 / bden, 0.0)
    for itn in range(1, maxit + 1):
        if not np.any(active):
            break
        s = np.where(active, 1.0 / np.where(beta > 0, beta, 1.0), 0.0)
        v = s[None, :] * y
        y = A(v)
        if itn >= 2:
            y = y - (beta / np.where(oldb > 0, oldb, 1.0))[None, :] * r1
        alfa = np.real(np.sum(np.conj(v) * y, axis=0))
        y = y - (alfa / np.where(beta > 0, beta, 1.0))[None, :] * r2
        r1 = r2
        r2 = y
        oldb = beta
        beta = np.sqrt(np.real(np.sum(np.conj(r2) * y, axis=0)))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = np.maximum(np.hypot(gbar, beta), EPS)
        cs = gbar / gamma
        sn = beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1 = w2
        w2 = w
        w = (v - oldeps[None, :] * w1 - delta[None, :] * w2) / gamma[None, :]
        upd = active.astype(np.float64)
        x = x + (upd * phi)[None, :] * w
        iters = iters + active
        rr = np.where(bnorm > 0, phibar / bden, 0.0)
        relres = np.where(active, rr, relres)
        active = active & (rr > tol)
    return x, {'iters': iters.tolist(), 'relres': relres.tolist()}
