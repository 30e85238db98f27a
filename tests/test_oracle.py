"""Pins the FP64 oracle against closed forms, library routines and brute force.

The reference ships no golden vectors or known-answer tests (its only test,
``test_nufft3d.m``, displays images), so every oracle function is pinned here
by mathematics that a plausible mistake (wrong sign, dropped term, swapped
index, off-by-one wrap) would break.
"""
import math

import numpy as np
import pytest
import scipy.special as ss

from oracle import nufft_oracle as O
from nufft_3d_amd import synth


# ------------------------------------------------------------------ kernel.m
def test_bessi0_matches_i0_within_nr_error():
    x = np.linspace(0, 30, 3001)
    rel = np.abs(O.bessi0(x) / ss.i0(x) - 1)
    assert rel.max() < 2.0e-7          # NR 6.6: |eps| < 1.6e-7 / 1.9e-7


def test_bessi0_imaginary_branch_tracks_j0():
    # I0(i t) = J0(t); the x2 hack (kernel.m:32-34) turns the large-argument
    # asymptotic form into J0's; without it the value would be half.
    t = np.linspace(0.01, 40, 4000)
    r = np.real(O.bessi0(1j * t))
    small = t < 3.75
    assert np.max(np.abs(r[small] - ss.j0(t[small]))) < 5e-3
    assert np.max(np.abs(r[~small] - ss.j0(t[~small]))) < 2e-2


@pytest.mark.parametrize("J,u,radial", [(4, 2, 1), (4, 2, 0), (6, 1.25, 1), (5, 1.5, 0)])
def test_kernel_identities(J, u, radial):
    a = O.beatty_alpha(J, u, radial)
    assert O.kernel(np.array([0.0]), J, a)[0] == 1.0
    edge = O.kernel(np.array([(J / 2) ** 2]), J, a)[0]
    assert edge == pytest.approx(1.0 / O.bessi0(np.array([a]))[0], rel=1e-15)
    d2 = np.linspace(0, (J / 2) ** 2, 101)
    exact = ss.i0(a * np.sqrt(1 - d2 / (J / 2) ** 2)) / ss.i0(a)
    assert np.max(np.abs(O.kernel(d2, J, a) - exact)) < 1e-6


def test_beatty_alpha_values():
    # nufft_3d.m:85-92 at the benchmark settings
    assert O.beatty_alpha(4, 2, 1) == pytest.approx(8.773990786893783, rel=1e-14)
    assert O.beatty_alpha(4, 2, 0) == pytest.approx(8.996152293560439, rel=1e-14)
    assert O.beatty_alpha(6, 1.25, 1) == pytest.approx(10.977607650909238, rel=1e-14)


def test_matlab_round_half_away():
    x = np.array([-2.5, -1.5, -0.5, -0.49999999999999994, 0.5, 1.5, 2.5, 0.49999999999999994])
    assert O.matlab_round(x).tolist() == [-3, -2, -1, -0.0, 1, 2, 3, 0]


# ------------------------------------------------------------------ H
def _brute_H(om_s, ok, K, J, alpha, radial):
    """Enumerate every lattice point within J of each sample (superset)."""
    rows = {}
    for m in range(om_s.shape[1]):
        if not ok[m]:
            continue
        p = om_s[:, m]
        lo = np.floor(p - J).astype(int)
        hi = np.ceil(p + J).astype(int)
        ent = {}
        for gx in range(lo[0], hi[0] + 1):
            for gy in range(lo[1], hi[1] + 1):
                for gz in range(lo[2], hi[2] + 1):
                    d = (gx - p[0]) ** 2, (gy - p[1]) ** 2, (gz - p[2]) ** 2
                    if radial:
                        inc = 4 * (d[0] + d[1] + d[2]) <= J * J
                    else:
                        inc = all(4 * v <= J * J for v in d)
                    if not inc:
                        continue
                    col = (gx % K[0]) + K[0] * (gy % K[1]) + K[0] * K[1] * (gz % K[2])
                    if radial:
                        w = ss.i0(alpha * math.sqrt(max(0.0, 1 - (d[0] + d[1] + d[2]) / (J / 2) ** 2)))
                    else:
                        w = np.prod([ss.i0(alpha * math.sqrt(max(0.0, 1 - v / (J / 2) ** 2))) for v in d])
                        w /= ss.i0(alpha) ** 2
                    ent[col] = ent.get(col, 0.0) + w / ss.i0(alpha)
        rows[m] = ent
    return rows


@pytest.mark.parametrize("radial", [1, 0])
def test_H_against_brute_force(radial):
    rng = np.random.default_rng(5)
    N = 8
    om = rng.uniform(-N / 2 - 1, N / 2 + 0.5, size=(3, 60))
    om[:, 0] = [0.25, -0.25, 0.5]       # exact half: rounding rule matters
    om[:, 1] = [-4.0, 3.5, -3.75]       # on and beyond the bound
    o = O.NufftOracle(om, N, J=4, u=2, radial=radial)
    bf = _brute_H(o.om_s, o.ok, o.K, o.J, o.alpha, radial)
    H = o.H.tocsr()
    for m in range(om.shape[1]):
        row = H.getrow(m)
        got = dict(zip(row.indices.tolist(), row.data.tolist()))
        want = bf.get(m, {})
        assert set(got) == set(want), m
        for c, v in want.items():
            assert got[c] == pytest.approx(v, rel=2e-6, abs=1e-7)
            assert np.float32(got[c]) == got[c]          # fp32-rounded (nufft_3d.m:160)


def test_bounds_and_nnz():
    om = np.array([[-4, 3.5, 3.51, -4.01, 0], [0, 0, 0, 0, 0], [0, 0, 0, 0, 1]], dtype=float)
    o = O.NufftOracle(om, 8, J=4, u=2)
    # u*om in [-K/2, K/2-1] = [-8, 7] inclusive (nufft_3d.m:123-125)
    assert o.ok.tolist() == [True, True, False, False, True]
    assert o.nnz == 3
    assert o.H.getrow(2).nnz == 0 and o.H.getrow(3).nnz == 0


# ------------------------------------------------------------------ deap.m
@pytest.mark.parametrize("radial", [1, 0])
def test_deap_equals_symmetric_irfft(radial):
    N, u, J = (8, 10, 12), 2, 4
    K = tuple(int(2 * math.ceil(n * u / 2)) for n in N)
    a = O.beatty_alpha(J, u, radial)
    T = O.deap_table(J, a, radial)
    V = T.copy()
    for j in range(3):
        # MATLAB ifft(V*K, 2K, j, 'symmetric') == irfft of the first K+1 bins
        X = np.moveaxis(V, j, -1) * K[j]
        pad = np.zeros(X.shape[:-1] + (K[j] + 1,))
        pad[..., :X.shape[-1]] = X
        y = np.fft.irfft(pad, n=2 * K[j], axis=-1)
        keep = np.r_[0:N[j] // 2, 2 * K[j] - N[j] // 2:2 * K[j]]
        V = np.moveaxis(y[..., keep], -1, j)
    V = np.fft.fftshift(V)
    U = O.deap(N, K, J, a, radial)
    assert np.allclose(1.0 / U, V, rtol=1e-13, atol=1e-15)


def test_deap_table_uses_complex_continuation():
    # beyond J/2 the radial table evaluates the kernel with imaginary args
    a = O.beatty_alpha(4, 2, 1)
    T = O.deap_table(4, a, 1)
    assert T[4, 4, 4] != 0.0 and np.isfinite(T[4, 4, 4])
    assert T[0, 0, 0] == 1.0


# ------------------------------------------------------------------ FFT
def test_fft3_pad_is_centered_dft():
    rng = np.random.default_rng(1)
    N, K = (4, 6, 2), (8, 10, 4)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    G = O.fft3_pad(x[..., None], N, K)[..., 0]
    xs = [np.arange(n) - n // 2 for n in N]
    want = np.zeros(K, complex)
    for a in range(K[0]):
        for b in range(K[1]):
            for c in range(K[2]):
                ph = (np.exp(-2j * np.pi * a * xs[0] / K[0])[:, None, None]
                      * np.exp(-2j * np.pi * b * xs[1] / K[1])[None, :, None]
                      * np.exp(-2j * np.pi * c * xs[2] / K[2])[None, None, :])
                want[a, b, c] = np.sum(x * ph)
    assert np.allclose(G, want, rtol=1e-12, atol=1e-12)


def test_fft3_pad_crop_adjoint():
    rng = np.random.default_rng(2)
    N, K = (6, 4, 8), (12, 8, 16)
    x = rng.normal(size=N + (2,)) + 1j * rng.normal(size=N + (2,))
    g = rng.normal(size=K + (2,)) + 1j * rng.normal(size=K + (2,))
    lhs = np.vdot(O.fft3_pad(x, N, K), g)
    rhs = np.vdot(x, O.ifft3_crop(g, N, K))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


# ------------------------------------------------------------------ fNUFT/aNUFT
def _small_plan(N=12, nspokes=40, nr=16, radial=1, J=4, u=2):
    om = synth.koosh_ball(N, nspokes, nr)
    return O.NufftOracle(om, N, J=J, u=u, radial=radial), om


@pytest.mark.parametrize("radial", [1, 0])
def test_forward_adjoint_pair(radial):
    o, _ = _small_plan(radial=radial)
    rng = np.random.default_rng(3)
    x = rng.normal(size=o.N + (2,)) + 1j * rng.normal(size=o.N + (2,))
    k = rng.normal(size=(o.M, 2)) + 1j * rng.normal(size=(o.M, 2))
    lhs = np.vdot(o.fNUFT(x), k)
    rhs = np.vdot(x, o.aNUFT(k))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # linearity
    y = rng.normal(size=o.N + (2,))
    assert np.allclose(o.fNUFT(x + 2 * y), o.fNUFT(x) + 2 * o.fNUFT(y), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("radial,J,u,tol", [(1, 4, 2, 3e-3), (0, 4, 2, 3e-3), (1, 6, 2, 2e-4)])
def test_forward_approximates_nudft(radial, J, u, tol):
    N = 10
    rng = np.random.default_rng(4)
    om = rng.uniform(-N / 2, N / 2 - 0.6, size=(3, 300))
    o = O.NufftOracle(om, N, J=J, u=u, radial=radial)
    x = rng.normal(size=(N, N, N)) + 1j * rng.normal(size=(N, N, N))
    k = o.fNUFT(x)[:, 0]
    xs = np.arange(N) - N // 2
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing='ij')
    E = np.exp(-2j * np.pi * (np.outer(om[0], X.ravel()) + np.outer(om[1], Y.ravel())
                              + np.outer(om[2], Z.ravel())) / N)
    want = E @ x.ravel() / math.sqrt(o.nnz)
    err = np.linalg.norm(k - want) / np.linalg.norm(want)
    assert err < tol


def test_out_of_bounds_samples_are_zero():
    om = synth.koosh_ball(8, 10, 8)
    om[:, 3] = [10.0, 0, 0]
    o = O.NufftOracle(om, 8)
    x = np.ones((8, 8, 8), complex)
    assert o.fNUFT(x)[3, 0] == 0
    k = np.zeros((o.M, 1), complex)
    k[3] = 1.0
    assert np.all(o.aNUFT(k) == 0)


# ------------------------------------------------------------------ density
def test_density_invariants():
    om = synth.koosh_ball(12, 30, 16)
    om[:, 5] = [50, 0, 0]
    o = O.NufftOracle(om, 12)
    d = o.d
    assert np.all(d >= 0)
    assert d[5] == 0
    assert np.sum(d) == pytest.approx(o.nnz, rel=1e-12)
    # one more Pipe step changes d by little once converged (density.m:149-152)
    d1 = d / np.maximum(o.H @ (o.H.T @ d), O.EPS)
    d1 = d1 / (np.sum(d1) / o.nnz)
    assert np.linalg.norm(d1 - d) / np.linalg.norm(d) < 0.1


@pytest.mark.parametrize("radial", [1, 0])
def test_density_full_cartesian_coverage_is_one(radial):
    # every integer point of [-N/2, N/2)^3 sampled once: u*om covers every
    # even point of the periodic K-grid, so H H' 1 is constant by translation
    # symmetry and the Pipe weights are exactly 1 after any iteration
    N = 6
    g = np.arange(-N // 2, N // 2, dtype=float)
    om = np.stack([a.ravel() for a in np.meshgrid(g, g, g, indexing='ij')])
    o = O.NufftOracle(om, N, J=4, u=2, radial=radial)
    assert o.nnz == N ** 3
    assert np.max(np.abs(o.d - 1)) < 1e-12


def test_density_takes_twenty_pipe_iterations():
    # density.m:5,11-14: 20 steps d = d./max(H H' d, eps), then mean 1. A
    # small cluster whose iteration is still moving at step 20, with H
    # assembled by brute force from scipy's I0 and rounded to fp32
    # (nufft_3d.m:160): 19 or 21 steps differ from 20 by ~3e-3
    om = np.array([[2.0, -0.5786, -0.1867, -0.8895, -0.4275, -0.1419],
                   [0.0, 0.2397, -0.2152, 0.4055, 0.2770, -0.1238],
                   [0.0, 0.2378, 0.5585, -0.2848, 0.0786, -0.5467]])
    o = O.NufftOracle(om, 8, J=4, u=2, radial=1)
    bf = _brute_H(o.om_s, o.ok, o.K, o.J, o.alpha, 1)
    H = np.zeros((om.shape[1], int(np.prod(o.K))))
    for m, ent in bf.items():
        for c, v in ent.items():
            H[m, c] = np.float32(v)
    A = H @ H.T

    def pipe(k):
        d = o.ok.astype(float)
        for _ in range(k):
            d = d / np.maximum(A @ d, O.EPS)
        return d / (np.sum(d) / o.nnz)

    def r(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    assert r(o.d, pipe(20)) < 1e-6
    assert r(pipe(19), pipe(20)) > 1e-3 and r(pipe(21), pipe(20)) > 1e-3


# ------------------------------------------------------------------ MINRES
def _krylov_min(A, b, k):
    n = b.size
    Q = np.zeros((n, k), complex)
    q = b / np.linalg.norm(b)
    for j in range(k):
        Q[:, j] = q
        q = A @ q
        q = q - Q[:, :j + 1] @ (Q[:, :j + 1].conj().T @ q)
        q = q - Q[:, :j + 1] @ (Q[:, :j + 1].conj().T @ q)
        q /= np.linalg.norm(q)
    c, *_ = np.linalg.lstsq(A @ Q, b, rcond=None)
    return Q @ c


@pytest.mark.parametrize("indefinite", [False, True])
def test_minres_is_krylov_minimiser(indefinite):
    rng = np.random.default_rng(6)
    n = 14
    B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = B @ B.conj().T / n + 0.3 * np.eye(n)
    if indefinite:
        A = A - 1.5 * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    for k in (1, 2, 4, 7):
        x, _ = O.minres(lambda v: A @ v, b[:, None], tol=0.0, maxit=k)
        want = _krylov_min(A, b, k)
        assert np.linalg.norm(x[:, 0] - want) < 1e-8 * np.linalg.norm(want)
    x, info = O.minres(lambda v: A @ v, b[:, None], tol=1e-12, maxit=n + 4)
    assert np.linalg.norm(x[:, 0] - np.linalg.solve(A, b)) < 1e-7 * np.linalg.norm(x)
    assert info['relres'][0] <= 1e-12


def test_minres_columns_independent():
    rng = np.random.default_rng(7)
    n = 10
    B = rng.normal(size=(n, n))
    A = B @ B.T + np.eye(n)
    b = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    X, _ = O.minres(lambda v: A @ v, b, tol=0.0, maxit=5)
    for c in range(3):
        x1, _ = O.minres(lambda v: A @ v, b[:, c:c + 1], tol=0.0, maxit=5)
        assert np.allclose(X[:, c], x1[:, 0], rtol=1e-13, atol=1e-14)


def test_inufft_weights_with_maxit_at_most_1_are_used_as_given():
    # iNUFT.m:55-58: maxit <= 1 with a vector W only warns, so W is not
    # multiplied by the density: regridding is aNUFT(W.*raw)
    o, om = _small_plan(N=10, nspokes=60, nr=20)
    raw = np.random.default_rng(3).normal(size=(o.M, 2)) + 0j
    W = np.random.default_rng(4).uniform(0.2, 1, size=o.M)
    assert np.allclose(o.iNUFT(raw, 0, W=W), o.aNUFT(W[:, None] * raw), rtol=0, atol=1e-15)
    # maxit = 1: one MINRES step, x1 = b <Ab,b>/||Ab||^2 per coil, with the
    # operator A'WA of the same (undensified) W
    x1 = o.iNUFT(raw, 1, W=W, tol=0.0).reshape(-1, 2, order='F')
    b = o.aNUFT(W[:, None] * raw).reshape(-1, 2, order='F')
    Ab = o.iprojection(b, 0.0, W)
    for c in range(2):
        coef = np.vdot(Ab[:, c], b[:, c]).real / np.vdot(Ab[:, c], Ab[:, c]).real
        assert np.linalg.norm(x1[:, c] - coef * b[:, c]) < 1e-12 * np.linalg.norm(x1[:, c])
    with pytest.raises(ValueError):
        o.iNUFT(raw, 0, W=W[:20])


def test_inufft_regridding_and_iterations():
    o, om = _small_plan(N=10, nspokes=60, nr=20)
    img = synth.phantom_complex(10)
    raw = o.fNUFT(img)
    x0 = o.iNUFT(raw, 0)
    assert np.allclose(x0[..., 0], o.aNUFT(o.d[:, None] * raw)[..., 0])
    x5, info = o.iNUFT(raw, 5, 0.0, return_info=True)
    assert info['iters'] == [5] or info['relres'][0] <= 1e-4
    # residual of the normal equations decreases
    A = lambda v: o.iprojection(v, 0.0, o.d)
    b = o.aNUFT(o.d[:, None] * raw).reshape(-1, 1, order='F')
    r0 = np.linalg.norm(b)
    r5 = np.linalg.norm(b - A(x5.reshape(-1, 1, order='F')))
    assert r5 < 0.5 * r0


# ------------------------------------------------------------------ synth
def test_phantom_spot_values():
    p = synth.phantom3d(32)
    # centre voxel lies in ellipsoids 1 and 2 only: 1 - 0.8
    assert p[16, 16, 16] == pytest.approx(0.2)
    assert p[0, 0, 0] == 0.0
    assert set(np.unique(np.round(p, 6))) <= {0.0, 0.1, 0.2, 0.3, 0.4, 1.0, -0.0}


def test_koosh_ball_geometry():
    om = synth.koosh_ball(64, 5, 8)
    assert om.shape == (3, 40)
    # first spoke along +z (polar angle 0), readout from -N/2
    assert np.allclose(om[:, 0], [0, 0, -32])
    assert np.allclose(om[:, 1], [0, 0, -24])
    r = np.linalg.norm(om, axis=0)
    assert np.allclose(r.reshape(8, 5, order='F')[:, 0], np.abs(-32 + 8 * np.arange(8)))


def test_minres_warm_start():
    rng = np.random.default_rng(8)
    n = 12
    B = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = B @ B.conj().T / n + 0.5 * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    xs = np.linalg.solve(A, b)
    # the exact solution as initial guess: no iteration, returned unchanged
    x, info = O.minres(lambda v: A @ v, b[:, None], tol=1e-8, maxit=5, x0=xs[:, None])
    assert info['iters'] == [0] and np.array_equal(x[:, 0], xs)
    # otherwise x0 + the Krylov least-squares minimiser over K_k(A, b - A x0)
    x0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    r0 = b - A @ x0
    for k in (1, 3, 6):
        x, info = O.minres(lambda v: A @ v, b[:, None], tol=0.0, maxit=k, x0=x0[:, None])
        want = x0 + _krylov_min(A, r0, k)
        assert np.linalg.norm(x[:, 0] - want) < 1e-8 * np.linalg.norm(want)
        # relres is measured against ||b|| (the MATLAB convention)
        assert abs(info['relres'][0] - np.linalg.norm(b - A @ x[:, 0]) / np.linalg.norm(b)) < 1e-8


def test_discrep_stddev_estimate_is_the_mad_of_nonzeros():
    o, om = _small_plan(N=8, nspokes=30, nr=16)
    rng = np.random.default_rng(9)
    data = rng.normal(size=(o.M, 2)) + 1j * rng.normal(size=(o.M, 2))
    data[::7] = 0
    _, info = o.discrep(data, iters=1, return_info=True)
    v = data[data != 0]
    t = np.concatenate([v.real, v.imag])
    assert info['stddev'] == pytest.approx(np.median(np.abs(t - np.median(t))), rel=1e-15)
    assert info['discrepancy'] == pytest.approx(info['stddev'] * np.sqrt(data.size), rel=1e-15)


def test_discrep_matches_a_dense_restatement():
    # the same fixed point damp <- damp * discrepancy / ||A x(damp) - data||
    # with x(damp) from a dense solve of (A'WA + damp^2 I) x = A'W data
    o, om = _small_plan(N=6, nspokes=24, nr=12)
    rng = np.random.default_rng(10)
    img = synth.phantom_complex(6)
    data = o.fNUFT(img) + 0.05 * (rng.normal(size=(o.M, 1)) + 1j * rng.normal(size=(o.M, 1)))
    n = 6 ** 3
    E = np.eye(n, dtype=np.complex128)
    Acols = np.stack([o.fNUFT(E[:, j].reshape(6, 6, 6, order='F'))[:, 0] for j in range(n)], axis=1)
    W = o.d
    AWA = Acols.conj().T @ (W[:, None] * Acols)
    b = Acols.conj().T @ (W * data[:, 0])
    damp = 0.1
    stddev = 0.05
    disc = stddev * np.sqrt(data.size)
    for _ in range(8):
        x = np.linalg.solve(AWA + damp ** 2 * np.eye(n), b)
        damp = damp * disc / np.linalg.norm(Acols @ x - data[:, 0])
    got = o.discrep(data, stddev=stddev, iters=8, maxit=4 * n)
    assert got == pytest.approx(damp, rel=2e-3)


@pytest.mark.parametrize("N,u,J,radial", [((8, 8, 2), 1.25, 6, 1), ((8, 2, 2), 1, 4, 0), ((6, 6, 2), 1.5, 6, 1)])
def test_H_on_grids_narrower_than_the_footprint(N, u, J, radial):
    # K < 2*ceil(J/2)+1 on an axis: candidates wrap onto the same grid point
    # and the reference's sparse() sums those entries (nufft_3d.m:178-180,
    # 203,208); the brute force sums every lattice point within J mapped
    # mod K, so a dropped or doubled duplicate fails
    rng = np.random.default_rng(8)
    om = rng.uniform(-min(N) / 2, min(N) / 2 - 0.5, size=(3, 40))
    o = O.NufftOracle(om, N, J=J, u=u, radial=radial)
    assert min(o.K) < 2 * math.ceil(J / 2) + 1
    bf = _brute_H(o.om_s, o.ok, o.K, o.J, o.alpha, radial)
    H = o.H.tocsr()
    dup = 0
    for m in range(om.shape[1]):
        row = H.getrow(m)
        got = dict(zip(row.indices.tolist(), row.data.tolist()))
        want = bf.get(m, {})
        assert set(got) == set(want), m
        for c, v in want.items():
            assert got[c] == pytest.approx(v, rel=4e-6, abs=1e-7)
        dup += sum(1 for c in got if float(np.float32(got[c])) != got[c])
    assert dup > 0          # some entries are sums of several float weights
