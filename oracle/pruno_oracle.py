"""FP64 oracle of the SAKE/PRUNO calibration path (test infrastructure only).

Restates, in NumPy double precision:
* ``kernel_taps``       -> kernel offsets of ``pruno.m:56-71`` / ``sake.m:61-71``
  (``ndgrid(-fix(w/2):fix(w/2))`` masked, column-major order)
* ``make_data_matrix``  -> ``pruno.m:146-179`` literally (circshift copies,
  j<=k products, conjugate-symmetric fill)
* ``gram_fft``          -> the same matrix from circular cross-correlations
  (pinned against ``make_data_matrix`` on small inputs)
* ``pruno``             -> ``pruno.m:43-143`` (svd of AA, noise floor,
  filter, nv, null kernels, image-space transform)
* ``null_term``         -> the coupled term of ``nufft_3d.m:317-320``
* ``inufft_pruno``      -> ``iNUFT.m:200-218`` with no preconditioner (the
  reference passes an undefined ``M``)

Pinning: ``make_data_matrix`` is the literal loop; ``gram_fft`` and the
device Gram are checked against it; the null projector is compared as the
phase-invariant ``sum_v (1-f_v)^2 v v^H``; the null operator is checked
against its direct definition with image-space kernels.
"""
from __future__ import annotations

import math

import numpy as np

from .nufft_oracle import minres


def kernel_taps(width=5, radial=1, is2d=False):
    h = int(width) // 2                          # fix(width/2)
    r = np.arange(-h, h + 1)
    if is2d:
        x, y = np.meshgrid(r, r, indexing='ij')
        z = np.zeros_like(x)
    else:
        x, y, z = np.meshgrid(r, r, r, indexing='ij')
    if radial:
        k = np.sqrt(x ** 2 + y ** 2 + z ** 2) <= width / 2
    else:
        k = (np.abs(x) <= width / 2) & (np.abs(y) <= width / 2) & (np.abs(z) <= width / 2)
    kf = k.ravel(order='F')
    taps = np.stack([x.ravel(order='F')[kf], y.ravel(order='F')[kf], z.ravel(order='F')[kf]], 1)
    return taps, k


def flip_index(n, center=1):
    """sake.m:86-88: circshift(n:-1:1, [0 2*center-1]) as 0-based indices
    (reflection about the 1-based k-space centre)."""
    return np.roll(np.arange(n, 0, -1), 2 * center - 1) - 1


def make_data_matrix(data, taps, loraks=False, center=(1, 1, 1), jmax=None):
    """pruno.m:146-179 / sake.m:176-224: AA(:,j,:,k) = row'*col with
    circshift copies; with loraks the conjugate-reflection block
    BB(:,j,:,k) = row'*conj(col(flip.x,flip.y,flip.z,:)) (sake.m:210-214)
    and the 2n x 2n matrix [AA BB; conj(BB) conj(AA)] (sake.m:221-224).
    jmax: only the outer-loop taps j < jmax (a timed sample of the loop;
    the blocks of the other taps stay zero)."""
    nx, ny, nz, nc = data.shape
    nk = len(taps)
    jend = nk if jmax is None else min(nk, int(jmax))
    AA = np.zeros((nc, nk, nc, nk), dtype=np.complex128)
    BB = np.zeros((nc, nk, nc, nk), dtype=np.complex128) if loraks else None
    if loraks:
        fl = np.ix_(flip_index(nx, center[0]), flip_index(ny, center[1]), flip_index(nz, center[2]))
    for j in range(jend):
        row = np.roll(data, tuple(taps[j]), axis=(0, 1, 2)).reshape(-1, nc, order='F')
        for k in range(j, nk):
            colv = np.roll(data, tuple(taps[k]), axis=(0, 1, 2))
            col = colv.reshape(-1, nc, order='F')
            blk = row.conj().T @ col
            AA[:, j, :, k] = blk
            AA[:, k, :, j] = blk.conj().T
            if loraks:
                colz = np.conj(colv[fl]).reshape(-1, nc, order='F')
                blk = row.conj().T @ colz
                BB[:, j, :, k] = blk
                BB[:, k, :, j] = blk.T
    AA = AA.reshape(nc * nk, nc * nk, order='F')
    if not loraks:
        return AA
    BB = BB.reshape(nc * nk, nc * nk, order='F')
    return np.block([[AA, BB], [BB.conj(), AA.conj()]])


def gram_fft(data, taps):
    """AA from R_cc'(d) = sum_v conj(D_c[v]) D_c'[v+d], AA = R(s_j - s_k)."""
    nx, ny, nz, nc = data.shape
    F = np.fft.fftn(data, axes=(0, 1, 2))
    nk = len(taps)
    AA = np.zeros((nc, nk, nc, nk), dtype=np.complex128)
    for c in range(nc):
        for c2 in range(nc):
            R = np.fft.ifftn(np.conj(F[..., c]) * F[..., c2])   # R[d] = sum conj(D_c[v]) D_c2[v+d]
            for j in range(nk):
                d = taps[j] - taps       # s_j - s_k for all k
                AA[c, j, c2, :] = R[d[:, 0] % nx, d[:, 1] % ny, d[:, 2] % nz]
    return AA.reshape(nc * nk, nc * nk, order='F')


def gram_fft_loraks(data, taps, center=(1, 1, 1)):
    """The loraks matrix of make_data_matrix by FFT: with g = conj(data(flip)),
    BB[(c,j),(d,k)] = sum_u conj(D_c[u]) g_d[u + s_j + s_k]."""
    nx, ny, nz, nc = data.shape
    nk = len(taps)
    fl = np.ix_(flip_index(nx, center[0]), flip_index(ny, center[1]), flip_index(nz, center[2]))
    g = np.conj(data[fl])
    F = np.fft.fftn(data, axes=(0, 1, 2))
    G = np.fft.fftn(g, axes=(0, 1, 2))
    BB = np.zeros((nc, nk, nc, nk), dtype=np.complex128)
    for c in range(nc):
        for d in range(nc):
            R = np.fft.ifftn(np.conj(F[..., c]) * G[..., d])   # R[e] = sum conj(D_c[u]) g_d[u+e]
            for j in range(nk):
                e = taps[j] + taps
                BB[c, j, d, :] = R[e[:, 0] % nx, e[:, 1] % ny, e[:, 2] % nz]
    AA = gram_fft(data, taps)
    BB = BB.reshape(nc * nk, nc * nk, order='F')
    return np.block([[AA, BB], [BB.conj(), AA.conj()]])


def _eps(v):
    a = abs(float(v))
    return float(np.nextafter(a, np.inf) - a)


def _hist(y, x):
    """MATLAB hist(y, x) with a scalar bin count x (hist.m): floor(x) bins
    (0:x stops at floor(x)) with edges miny + k*(maxy-miny)/x, k=0..floor(x),
    the last edge replaced by maxy and every edge raised by eps(edge), so the
    bins are (a, b]; values at or below the first edge fall in bin 1. A
    constant y is spread over [y - floor(x/2) - 0.5, y + ceil(x/2) - 0.5].
    floor(x) < 1 gives no bins."""
    n = int(math.floor(x)) if x >= 1 else 0
    h = np.zeros(n, int)
    y = np.asarray(y, dtype=np.float64)
    if n < 1 or y.size == 0:
        return h
    lo, hi = float(np.min(y)), float(np.max(y))
    if lo == hi:
        lo = lo - math.floor(x / 2) - 0.5
        hi = hi + math.ceil(x / 2) - 0.5
    bw = (hi - lo) / x
    e = [lo + bw * k for k in range(n + 1)]
    e[n] = hi
    e = np.array([v + _eps(v) for v in e])
    for v in y:
        if not v <= e[n]:
            continue
        c = int(np.searchsorted(e, v, side='right'))     # edges <= v
        h[min(max(c, 1), n) - 1] += 1
    return h


def estimate_noise(W, density, nx, ny):
    """pruno.m:97-106 = sake.m:128-137: noise floor from the squared singular
    values W (descending); skips true zeros, then the leading values while the
    histogram's first bin holds the maximum, and takes the median."""
    n = len(W)
    thr = np.spacing(n * W[0])
    hi = int(np.count_nonzero(W > thr))
    if hi == 0:
        return 0.0
    lo = 1
    while lo <= hi:
        h = _hist(W[lo - 1:hi], math.sqrt(hi - lo))
        # [~,k] = max(h): the first maximum; no bins leaves k empty (no break)
        if h.size and int(np.argmax(h)) + 1 > 1:
            break
        lo += 1
    lo = min(lo, hi)
    sigma = math.sqrt(float(np.median(W[lo - 1:hi])))
    return sigma / math.sqrt(density * nx * ny)


def select(W, density, nx, ny, noise=None, tol=0.01):
    """pruno.m:96-115 on the squared singular values W (descending): the noise
    std (estimated when None), the singular values sqrt(W), the
    minimum-variance filter f = max(0, 1 - sigma^2/W) and nv = #(f < 1-tol)."""
    if noise is None:
        noise = estimate_noise(W, density, nx, ny)
    sigma = noise * math.sqrt(density * nx * ny)              # pruno.m:108
    Ws = np.sqrt(np.maximum(W, 0))                            # pruno.m:111
    with np.errstate(divide='ignore'):
        f = np.where(Ws > 0, np.maximum(0, 1 - sigma ** 2 / np.where(Ws > 0, Ws, 1) ** 2), 0.0)
    nv = int(np.count_nonzero(f < 1 - tol))                   # pruno.m:115
    return noise, Ws, f, nv


def pruno(data, N, width=5, radial=1, noise=None, tol=0.01, literal_first_nv=False, AA=None):
    """pruno.m:43-143. Returns dict(K, Kt, nv, sv, f, noise, taps, mask)."""
    nx, ny, nz, nc = data.shape
    taps, mask = kernel_taps(width, radial, nz == 1)
    nk = len(taps)
    density = np.count_nonzero(data) / data.size
    if density < 0.99:
        raise ValueError('data must be fully sampled')
    if AA is None:
        AA = gram_fft(data, taps)
    w, V = np.linalg.eigh(AA)
    order = np.argsort(w)[::-1]
    W = w[order]
    V = V[:, order]
    noise, Ws, f, nv = select(W, density, nx, ny, noise, tol)
    n = nc * nk
    sel = np.arange(nv) if literal_first_nv else np.arange(n - nv, n)
    Kt = V[:, sel] * (1 - f[sel])[None, :]                    # [(c + nc*t), v]
    K = kernels_image(Kt, mask, N, nc)
    return dict(K=K, Kt=Kt, nv=nv, sv=Ws, f=f, noise=noise, taps=taps, mask=mask)


def kernels_image(Kt, mask, N, nc):
    """pruno.m:124-143: the tap-domain kernels Kt [(c + nc*t), v] into the
    mask cube (K(mask,:,v) = V(:,:,v) after the [nc nk] -> [nk nc] permute),
    reshaped to [size(mask) (1) nc nv], then ifft(K, N(a), a) * N(a) per axis."""
    nv = Kt.shape[1]
    nk = int(np.count_nonzero(mask))
    wdt = mask.shape[0]
    shape = (wdt, wdt, mask.shape[2] if mask.ndim == 3 else 1, nc, nv)
    flat = np.zeros((mask.size, nc, nv), dtype=np.complex128)
    flat[mask.ravel(order='F')] = Kt.reshape(nc, nk, nv, order='F').transpose(1, 0, 2)
    K = flat.reshape(shape, order='F')
    for ax in range(3):
        K = np.fft.ifft(K, n=N[ax], axis=ax) * N[ax]
    return K


def null_term(fnull, x):
    """nufft_3d.m:317-320 with anull = conj(fnull): sum_v anull*sum_c fnull*x."""
    r = np.sum(fnull * x[..., None], axis=3)                  # [N1 N2 N3 nv]
    return np.sum(np.conj(fnull) * r[:, :, :, None, :], axis=4)


def inufft_pruno(orc, raw, maxit, damp, lam, width=5, radial=1, noise=None, tol=0.01,
                 literal_first_nv=False, iters=100):
    """iNUFT.m:119-136 then 200-218 (coupled MINRES, no preconditioner)."""
    N = orc.N
    raw = np.reshape(np.asarray(raw, dtype=np.complex128), (orc.M, -1), order='F')
    nc = raw.shape[1]
    W = orc.d
    b = orc.aNUFT(W[:, None] * raw)
    if maxit == 0:
        x = b
    else:
        x, _ = minres(lambda v: orc.iprojection(v, damp, W), b.reshape(-1, nc, order='F'),
                      1e-4, maxit)
    x = np.reshape(x, N + (nc,), order='F')
    xk = np.fft.fft(np.fft.fft(np.fft.fft(x, axis=0), axis=1), axis=2)
    pr = pruno(xk, N, width, radial, noise, tol, literal_first_nv)
    fnull = pr['K'] * lam

    def A(v):
        v4 = np.reshape(v, N + (nc,), order='F')
        y = orc.iprojection(v4, damp, W).reshape(N + (nc,), order='F')
        y = y + null_term(fnull, v4)
        return y.reshape(-1, 1, order='F')

    xs, info = minres(A, b.reshape(-1, 1, order='F'), 1e-4, iters)
    return xs.reshape(N + (nc,), order='F'), pr, info


def null_B(Kt, taps, nc, N, lam):
    """Per-voxel B(x)[c,c'] = lam^2 sum_v conj(fnull_c,v(x)) fnull_c',v(x) built
    from the tap-domain kernels via G(d) = sum_{t'-t=d} conj(K_t) K_t'^T and a
    separable inverse DFT (the reformulation the device path uses)."""
    nk = len(taps)
    w = int(taps.max() - taps.min()) + 1 if len(taps) else 1
    h = (w - 1) // 2
    D = 2 * w - 1
    P = np.conj(Kt) @ Kt.T                       # [(t,c),(t',c')]
    G = np.zeros((D, D, D, nc, nc), dtype=np.complex128)
    for t in range(nk):
        for t2 in range(nk):
            d = taps[t2] - taps[t] + (w - 1)
            G[d[0], d[1], d[2]] += P[t * nc:(t + 1) * nc, t2 * nc:(t2 + 1) * nc]
    # ifft kernels sit at cube index (tap + h): d = (t2+h) - (t+h) = t2 - t
    B = G
    for ax in range(3):
        n = N[ax]
        dd = np.arange(D) - (w - 1)
        E = np.exp(2j * np.pi * np.outer(np.arange(n), dd) / n)     # [x, d]
        B = np.moveaxis(np.tensordot(E, np.moveaxis(B, ax, 0), axes=(1, 0)), 0, ax)
    return lam ** 2 * B                          # [N1 N2 N3 nc nc], B[..., c, c']
