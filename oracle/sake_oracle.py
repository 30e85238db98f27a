"""FP64 oracle of the SAKE Cadzow loop (test infrastructure only).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.

Restates, in NumPy double precision:
* ``undo_data_matrix``  -> ``sake.m:227-284`` literally (circshift copies,
  nk^2 products [V x nc][nc x nc], conjugate-reflection terms with loraks,
  average over nk or 2 nk)
* ``lag_kernels``       -> the same operator as a correlation over tap
  differences (and, with loraks, tap sums) with nc x nc matrix weights
* ``undo_image``        -> that correlation applied in the image domain as a
  per-voxel coil mixing (the form the device uses); ``ifftn(undo_data_matrix
  (F, fftn(x))) == undo_image(x)``
* ``sake``              -> ``sake.m:48-171`` (data consistency with
  ``iNUFT(..,10,damp)``, Hankel Gram, SVD, noise floor, minimum-variance
  filter, anti-diagonal averaging, convergence on the largest singular value)

Choices where the reference is not self-contained (its helpers are absent):
* ``fft3``/``ifft3`` (``sake.m:108-110``) are the unshifted, unnormalised
  ``fftn`` and its inverse ``ifftn`` over the three image axes: k-space DC at
  index 1, which is the reference's default ``center = [1 1 1]``
  (``sake.m:78-83``), and the same transform ``iNUFT.m:204`` uses to form
  PRUNO's k-space.
* ``matrix_density = obj.sd`` (``sake.m:96``) names a property the class does
  not have (``nufft_3d.m:26-49``). It is taken as the sampled fraction
  ``min(1, nnz / prod(N))``. It only scales a user-supplied ``noise``: an
  estimated noise is divided by and multiplied by the same factor
  (``sake.m:136,139``).
* ``svd(AA)`` of the Hermitian PSD Gram: eigen-decomposition, singular values
  ``|lambda|`` sorted descending (``sake.m:124-125``).
* ``opts.sparsity`` needs the external ``DWT`` (``sake.m:91-93,114-118``):
  not restated; ``sparsity > 0`` raises.

Pinning (tests/test_sake_oracle.py): ``undo_data_matrix`` is the literal loop;
it is checked against the identity (F = I returns the data), a hand-derived
width-1 closed form (per-voxel coil mixing, with loraks the conjugate
reflection), and the independent image-domain form. The loop is checked
against the closed form of one iteration with a zero noise floor (F = I:
``fftn(iNUFT(data, 10, damp))``), against a convergence property (the
returned k-space's Hankel matrix has the filtered spectrum), and the singular
values against ``numpy.linalg.svd`` of the explicit Hankel matrix.
"""
from __future__ import annotations

import math

import numpy as np

from .pruno_oracle import estimate_noise, flip_index, gram_fft, gram_fft_loraks, kernel_taps


def fft3(x):
    return np.fft.fftn(x, axes=(0, 1, 2))


def ifft3(k):
    return np.fft.ifftn(k, axes=(0, 1, 2))


def undo_data_matrix(F, data, taps, loraks=False, center=(1, 1, 1)):
    """sake.m:227-284: ksp = sum_{j,k} circshift(colA_j * F(:,j,:,k), -s_k) / nk,
    with loraks the conjugate-reflection columns colZ and the 2 nk blocks."""
    nx, ny, nz, nc = data.shape
    nk = len(taps)
    m = 2 if loraks else 1
    F4 = np.reshape(F, (nc, m * nk, nc, m * nk), order='F')
    out = np.zeros_like(data, dtype=np.complex128)
    shp = data.shape
    if loraks:
        fl = np.ix_(flip_index(nx, center[0]), flip_index(ny, center[1]), flip_index(nz, center[2]))
    for j in range(nk):
        colA = np.roll(data, tuple(taps[j]), axis=(0, 1, 2))
        A2 = colA.reshape(-1, nc, order='F')
        if loraks:
            Z2 = np.conj(colA[fl]).reshape(-1, nc, order='F')
        for k in range(nk):
            chunk = A2 @ F4[:, j, :, k]
            if loraks:
                chunk = chunk + Z2 @ F4[:, j + nk, :, k]
                cz = A2 @ F4[:, j, :, k + nk] + Z2 @ F4[:, j + nk, :, k + nk]
                cz = np.conj(cz.reshape(shp, order='F')[fl])
                chunk = chunk.reshape(shp, order='F') + cz
            else:
                chunk = chunk.reshape(shp, order='F')
            out += np.roll(chunk, tuple(-taps[k]), axis=(0, 1, 2))
    return out / (m * nk)


def lag_kernels(F, taps, nc, loraks=False):
    """The operator of undo_data_matrix as correlations:
    ksp'(v,d) = 1/(m nk) sum_e sum_c [D(v+e,c) G_e[c,d] + g(v+e,c) H_e[c,d]],
    g(u) = conj(D(flip(u))), with
    G_e = sum_{s_k - s_j = e} F(c,j,d,k) + conj(F(c,j+nk,d,k+nk))   (loraks)
    H_e = sum_{s_j + s_k = e} F(c,j+nk,d,k) + conj(F(c,j,d,k+nk)).
    Returns dicts e -> [nc, nc] (H empty without loraks)."""
    nk = len(taps)
    m = 2 if loraks else 1
    F4 = np.reshape(F, (nc, m * nk, nc, m * nk), order='F')
    G, H = {}, {}
    for j in range(nk):
        for k in range(nk):
            e = tuple(taps[k] - taps[j])
            blk = F4[:, j, :, k].copy()
            if loraks:
                blk = blk + np.conj(F4[:, j + nk, :, k + nk])
                e2 = tuple(taps[j] + taps[k])
                H[e2] = H.get(e2, 0) + F4[:, j + nk, :, k] + np.conj(F4[:, j, :, k + nk])
            G[e] = G.get(e, 0) + blk
    return G, H


def undo_image(F, x, taps, loraks=False, center=(1, 1, 1)):
    """ifft3(undo_data_matrix(F, fft3(x))) computed in the image domain:
    a k-space correlation sum_e D(v+e) G_e is the per-voxel product
    x(r) B_G(r) with B_G(r) = sum_e G_e exp(-2 pi i e.r/N), and g = fft3(xt)
    with xt(r) = conj(x(r)) exp(+2 pi i (2 c0).r / N), c0 = center - 1."""
    N = x.shape[:3]
    nc = x.shape[3]
    nk = len(taps)
    m = 2 if loraks else 1
    G, H = lag_kernels(F, taps, nc, loraks)
    r = np.meshgrid(*[np.arange(n) for n in N], indexing='ij')

    def field(K):
        B = np.zeros(N + (nc, nc), dtype=np.complex128)
        for e, M in K.items():
            ph = np.exp(-2j * np.pi * sum(e[a] * r[a] / N[a] for a in range(3)))
            B += ph[..., None, None] * M[None, None, None]
        return B

    BG = field(G)
    out = np.einsum('xyzc,xyzcd->xyzd', x, BG)
    if loraks:
        c0 = [c - 1 for c in center]
        ph = np.exp(2j * np.pi * sum(2 * c0[a] * r[a] / N[a] for a in range(3)))
        xt = np.conj(x) * ph[..., None]
        out = out + np.einsum('xyzc,xyzcd->xyzd', xt, field(H))
    return out / (m * nk)


def hankel_matrix(data, taps, loraks=False, center=(1, 1, 1)):
    """The explicit data matrix A (rows = voxels, columns (c, j) c fastest) whose
    Gram A^H A is make_data_matrix; with loraks [A, Z] and Z the conjugate
    reflections, whose Gram is [AA BB; conj(BB) conj(AA)] up to the
    column order the reference uses. For tiny inputs only."""
    nx, ny, nz, nc = data.shape
    cols = []
    for j in range(len(taps)):
        cols.append(np.roll(data, tuple(taps[j]), axis=(0, 1, 2)).reshape(-1, nc, order='F'))
    A = np.concatenate(cols, axis=1)
    if not loraks:
        return A
    fl = np.ix_(flip_index(nx, center[0]), flip_index(ny, center[1]), flip_index(nz, center[2]))
    Z = [np.conj(np.roll(data, tuple(t), axis=(0, 1, 2))[fl]).reshape(-1, nc, order='F') for t in taps]
    return np.concatenate([A] + Z, axis=1)


def sake(orc, data, width=4, radial=1, loraks=False, tol=1e-6, maxit=1000, noise=None,
         center=None, damp=0.0, sparsity=0, dc_iters=10, return_info=False, literal=False):
    """sake.m:48-171 (options sake.m:26-35). Returns ksp [nx ny nz nc]."""
    data = np.asarray(data, dtype=np.complex128)
    N = tuple(orc.N)
    nx, ny, nz = N
    if data.ndim != 2 or data.shape[0] != orc.M or data.shape[1] < 2:
        raise ValueError("'data' must be a float array [npts nc].")       # sake.m:57-59
    if sparsity:
        raise NotImplementedError("sparsity needs the external DWT (sake.m:91-93)")
    nc = data.shape[1]
    taps, _ = kernel_taps(width, radial)                                  # sake.m:62-71
    nk = len(taps)
    if center is None:
        center = (1, 1, 1)                                                # sake.m:79-83
    density = min(1.0, orc.nnz / float(np.prod(N)))                      # sake.m:96 (obj.sd)
    ksp = np.zeros(N + (nc,), dtype=np.complex128)
    hist = []
    sigma = None
    for it in range(1, int(maxit) + 1):
        tmp = ifft3(ksp)                                                  # sake.m:108-111
        tmp = orc.iNUFT(orc.fNUFT(tmp) - data, dc_iters, damp)
        ksp = ksp - fft3(tmp)
        if literal:                                                       # sake.m:121
            from .pruno_oracle import make_data_matrix
            AA = make_data_matrix(ksp, taps, loraks, center)
        else:
            AA = gram_fft_loraks(ksp, taps, center) if loraks else gram_fft(ksp, taps)
        lam, V = np.linalg.eigh(AA)                                       # sake.m:124-125
        W = np.abs(lam)
        order = np.argsort(-W, kind='stable')
        W, V = W[order], V[:, order]
        if noise is None:                                                 # sake.m:128-138
            noise = estimate_noise(W, density, nx, ny)
        sigma = noise * math.sqrt(density * nx * ny)                      # sake.m:139
        Ws = np.sqrt(W)                                                   # sake.m:142
        with np.errstate(divide='ignore', invalid='ignore'):
            f = np.where(Ws > 0, np.maximum(0.0, 1 - sigma ** 2 / np.where(Ws > 0, Ws, 1) ** 2),
                         0.0)                                             # sake.m:145
        F = (V * f[None, :]) @ V.conj().T                                 # sake.m:146
        if literal:
            ksp = undo_data_matrix(F, ksp, taps, loraks, center)          # sake.m:149
        else:
            ksp = fft3(undo_image(F, ifft3(ksp), taps, loraks, center))
        n_inf, n_1, n_2 = float(Ws.max()), float(Ws.sum()), float(np.sqrt(np.sum(Ws ** 2)))
        if it == 1:
            t = tol                                                       # sake.m:155-160
        else:
            t = abs(n_inf - hist[-1]['norm_inf']) / n_inf
        hist.append(dict(norm_inf=n_inf, norm_1=n_1, norm_2=n_2, tol=t, rank=int(np.count_nonzero(f))))
        if t < tol:                                                       # sake.m:161,169
            break
    if return_info:
        return ksp, dict(iters=len(hist), noise=noise, sigma=sigma, history=hist, taps=taps)
    return ksp
