"""FP64 oracle of iNUFT's phase-constrained branch (test infrastructure only).

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
**Parity unpinned** against the external ``pcgpc`` itself (absent): its
exact stopping rule and breakdown handling.
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
