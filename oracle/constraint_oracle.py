"""FP64 oracle of iNUFT's phase-constraint and compressed-sensing branches (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.

Restates, in NumPy double precision:
* ``phase_estimate`` -> ``iNUFT.m:147-149``: ``P = exp(i angle(convn(x,
  ones(3), 'same')))`` (3x3x3 box sum with zero padding; angle(0) = 0)
* ``pprojection``    -> ``nufft_3d.m:325-334``:
  ``y = conj(P) .* iprojection(P .* x) + i lambda^2 imag(x)``
* ``pcgpc``          -> the external ``pcgpc`` (marcsous/parallel, not
  shipped with the reference; ``README.md:8``) called at ``iNUFT.m:163``.
  Restated from its published method (Bydder & Robson, MRM 2005;53:1393):
  conjugate gradients on the real-linear operator with the real inner
  product Re<a, b>, zero start, no preconditioner, one independent solve
  per column, stop when ||r|| <= tol ||b|| or after maxit iterations.
* ``inufft_phase``   -> ``iNUFT.m:119-136`` then ``:144-166``.
* ``haar_fwd`` / ``haar_inv``, ``sparsity_thresh`` -> the external ``HWT``
  (``iNUFT.m:174``) restated as the orthonormal 3-D Haar transform, one
  level = the 2-point butterfly ``(a +- b)/sqrt(2)`` along x, then y, then z
  of the current low-pass corner (low half first), levels while every
  non-singleton active length is even, at most 3; thresholding keeps the
  sparsity convention of the same helper family (``sake.m:35``: 0.1 = 10%
  zeros): soft threshold at the magnitude that zeroes that fraction of all
  coefficients.
* ``fista_n``        -> the external ``fistaN`` (``iNUFT.m:186-187``)
  restated as FISTA (Beck & Teboulle, SIAM J Imaging Sci 2009;2:183) on the
  normal equations: gradient step ``y - (A y - b)/L`` with L from 20 power
  iterations on A (x 1.05), prox = Haar threshold, momentum
  ``(t_k - 1)/t_{k+1}``, zero start, stop when ``||x_k - x_{k-1}|| <= tol ||x_k||``
  or after maxit iterations. All coils form one vector, as the reference
  reshapes b to one column (``iNUFT.m:180``).
* ``inufft_cs``      -> ``iNUFT.m:170-189``.

The reference reshapes the whole solution to one image (``iNUFT.m:147``),
which MATLAB rejects for more than one coil; here each coil gets the phase
of its own solution (identical for one coil).

Pinning (tests/test_constraint_oracle.py): ``pcgpc`` against CG's defining
property on small real-linear systems (iterate k minimises the energy norm
of the error over the Krylov space of the equivalent real symmetric system,
found by least squares) and exact convergence at k = 2n; ``phase_estimate``
against ``scipy.signal.convolve`` in 'same' mode; ``pprojection``
self-adjoint and positive definite under Re<.,.>; ``inufft_phase`` with
lambda = 0 equals the plain least-squares solution (the operator is then
unitarily similar to A'WA).
``haar_fwd`` against the explicit Kronecker product of 2-point Haar
matrices (one level) and orthonormality (multi-level); ``sparsity_thresh``
zero count and shrinkage; ``fista_n`` with a zero fraction converging to the
least-squares solution.
**Parity unpinned** against the external ``pcgpc``, ``fistaN`` and ``HWT``
themselves (absent): their stopping rules, breakdown handling, wavelet
levels and threshold convention.
"""
from __future__ import annotations

import numpy as np


def phase_estimate(x):
    """iNUFT.m:147-149 per coil: x [N1 N2 N3 nc] -> P (unit modulus)."""
    x = np.asarray(x, dtype=np.complex128)
    s = np.zeros_like(x)
    N = x.shape[:3]
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                src = [slice(max(0, -d), n - max(0, d)) for d, n in zip((dx, dy, dz), N)]
                dst = [slice(max(0, d), n - max(0, -d)) for d, n in zip((dx, dy, dz), N)]
                s[dst[0], dst[1], dst[2]] += x[src[0], src[1], src[2]]
    return np.exp(1j * np.angle(s))


def pprojection(orc, x, damp, lam, W, P):
    """nufft_3d.m:325-334 on x [N1 N2 N3 nc]."""
    N = orc.N
    x = np.reshape(x, N + (-1,), order='F')
    y = orc.iprojection(P * x, damp, W).reshape(x.shape, order='F')
    return np.conj(P) * y + 1j * lam ** 2 * np.imag(x)


def pcgpc(A, b, tol=1e-4, maxit=20):
    """CG with real inner products on each column of b [n nc]; A maps
    [n nc] -> [n nc]. Returns x and info (iters, relres per column)."""
    b = np.asarray(b, dtype=np.complex128)
    n, nc = b.shape
    x = np.zeros_like(b)
    iters, relres = [0] * nc, [0.0] * nc
    for c in range(nc):
        bc = b[:, c]
        bn = np.linalg.norm(bc)
        if bn == 0:
            continue
        xc = np.zeros(n, dtype=np.complex128)
        r = bc.copy()
        p = None
        rho_old = None
        rel = 1.0
        k = 0
        for k in range(1, maxit + 1):
            rho = float(np.real(np.vdot(r, r)))
            p = r.copy() if k == 1 else r + (rho / rho_old) * p
            col = np.zeros((n, nc), dtype=np.complex128)
            col[:, c] = p
            q = A(col)[:, c]
            pq = float(np.real(np.vdot(p, q)))
            alpha = rho / pq
            xc = xc + alpha * p
            r = r - alpha * q
            rho_old = rho
            rel = np.linalg.norm(r) / bn
            if rel <= tol:
                break
        x[:, c] = xc
        iters[c], relres[c] = k, float(rel)
    return x, dict(iters=iters, relres=relres)


def inufft_phase(orc, raw, maxit, damp, lam, W=None, tol=1e-4, return_info=False):
    """iNUFT(raw,maxit,damp,W,'phase-constraint',lambda): the plain MINRES
    solution (iNUFT.m:126-136), its smoothed phase, then pcgpc on
    P'A'WAP + i lambda^2 Im (iNUFT.m:144-166)."""
    N = orc.N
    raw = np.reshape(np.asarray(raw, dtype=np.complex128), (orc.M, -1), order='F')
    nc = raw.shape[1]
    Wv = orc.d if W is None else np.reshape(np.asarray(W, dtype=np.float64), (-1,)) * orc.d
    x = orc.iNUFT(raw, maxit, damp, W)
    P = phase_estimate(x)
    b = np.conj(P) * orc.aNUFT(Wv[:, None] * raw)
    V = int(np.prod(N))

    def A(v):
        return pprojection(orc, np.reshape(v, N + (nc,), order='F'), damp, lam, Wv, P).reshape(
            V, nc, order='F')

    xs, info = pcgpc(A, b.reshape(V, nc, order='F'), tol, maxit)
    im = P * xs.reshape(N + (nc,), order='F')
    return (im, info) if return_info else im


def haar_levels(N, cap=3):
    """Levels while every non-singleton active length is even (at most cap)."""
    act = [n for n in N if n > 1]
    lv = 0
    while lv < cap and act and all(a % 2 == 0 for a in act):
        lv += 1
        act = [a // 2 for a in act]
    return lv


def _haar_axis(x, ax, n, inverse):
    """One butterfly pass along axis ax over the first n entries (low half first)."""
    y = x.copy()
    h = n // 2
    sl = [slice(None)] * x.ndim
    r2 = np.sqrt(0.5)

    def take(i0, step):
        s = list(sl)
        s[ax] = slice(i0, i0 + step * h, step) if step else slice(i0, i0 + h)
        return tuple(s)
    if not inverse:
        a, b = x[take(0, 2)], x[take(1, 2)]
        y[take(0, 0)] = (a + b) * r2
        y[take(h, 0)] = (a - b) * r2
    else:
        lo, hi = x[take(0, 0)], x[take(h, 0)]
        y[take(0, 2)] = (lo + hi) * r2
        y[take(1, 2)] = (lo - hi) * r2
    return y


def haar_fwd(x, levels=None):
    """x [N1 N2 N3 nc] -> Haar coefficients of each coil (same shape)."""
    x = np.asarray(x, dtype=np.complex128)
    N = x.shape[:3]
    L = haar_levels(N) if levels is None else levels
    for lv in range(L):
        act = [n >> lv if n > 1 else 1 for n in N]
        blk = tuple(slice(0, a) for a in act)
        sub = x[blk]
        for ax in range(3):
            if N[ax] > 1:
                sub = _haar_axis(sub, ax, act[ax], False)
        x = x.copy()
        x[blk] = sub
    return x


def haar_inv(w, levels=None):
    w = np.asarray(w, dtype=np.complex128)
    N = w.shape[:3]
    L = haar_levels(N) if levels is None else levels
    for lv in reversed(range(L)):
        act = [n >> lv if n > 1 else 1 for n in N]
        blk = tuple(slice(0, a) for a in act)
        sub = w[blk]
        for ax in (2, 1, 0):
            if N[ax] > 1:
                sub = _haar_axis(sub, ax, act[ax], True)
        w = w.copy()
        w[blk] = sub
    return w


def sparsity_thresh(w, frac):
    """Soft threshold at the magnitude that zeroes round(frac * w.size) entries."""
    w = np.asarray(w, dtype=np.complex128)
    nz = int(np.floor(frac * w.size + 0.5))      # half away from zero
    if nz <= 0:
        return w.copy(), 0.0
    mags = np.abs(w)
    t = float(np.sort(mags.ravel())[min(nz, w.size) - 1])
    with np.errstate(divide='ignore', invalid='ignore'):
        s = np.where(mags > t, 1 - t / np.where(mags > 0, mags, 1), 0.0)
    return w * s, t


def fista_n(A, b, frac, tol=1e-4, maxit=20, power_iters=20):
    """FISTA on (1/2) x'Ax - Re b'x + Haar sparsity; b, x [N1 N2 N3 nc]."""
    b = np.asarray(b, dtype=np.complex128)
    N = b.shape[:3]
    v = np.ones(b.shape, dtype=np.complex128) / np.sqrt(b.size)
    L = 0.0
    for _ in range(power_iters):
        w = A(v)
        L = float(np.linalg.norm(w))
        v = w / L
    L *= 1.05
    x = np.zeros_like(b)
    y = x.copy()
    t = 1.0
    k = 0
    rel = float('nan')
    if np.linalg.norm(b) == 0:
        return x, dict(iters=0, L=L, relchange=0.0)
    for k in range(1, maxit + 1):
        z = y - (A(y) - b) / L
        wz, _ = sparsity_thresh(haar_fwd(z), frac)
        xn = haar_inv(wz)
        tn = (1 + np.sqrt(1 + 4 * t * t)) / 2
        y = xn + ((t - 1) / tn) * (xn - x)
        nx = np.linalg.norm(xn)
        rel = np.linalg.norm(xn - x) / nx if nx > 0 else 0.0
        x, t = xn, tn
        if rel <= tol:
            break
    return x, dict(iters=k, L=L, relchange=float(rel))


def inufft_cs(orc, raw, maxit, damp, lam, W=None, tol=1e-4, return_info=False):
    """iNUFT(raw,maxit,damp,W,'compressed-sensing',lambda) (iNUFT.m:170-189)
    with lambda the fraction of zero Haar coefficients."""
    N = orc.N
    raw = np.reshape(np.asarray(raw, dtype=np.complex128), (orc.M, -1), order='F')
    Wv = orc.d if W is None else np.reshape(np.asarray(W, dtype=np.float64), (-1,)) * orc.d
    b = orc.aNUFT(Wv[:, None] * raw)

    def A(v):
        return orc.iprojection(v, damp, Wv).reshape(v.shape, order='F')

    x, info = fista_n(A, b, lam, tol, maxit)
    return (x, info) if return_info else x
