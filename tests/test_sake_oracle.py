"""Pins the SAKE Cadzow oracle (sake.m:100-171,227-284) on the CPU."""
import numpy as np
import pytest

from oracle import sake_oracle as S
from oracle import pruno_oracle as P
from oracle.nufft_oracle import NufftOracle
from nufft_3d_amd import synth


def _c(shape, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=shape) + 1j * r.normal(size=shape)


def _herm(n, seed):
    a = _c((n, n), seed)
    return a + a.conj().T


@pytest.mark.parametrize("loraks", [False, True])
def test_undo_with_identity_filter_returns_the_data(loraks):
    # F = I: every tap pair j == k contributes the data once (sake.m:281-283)
    d = _c((5, 4, 6, 2), 1)
    taps, _ = P.kernel_taps(3, 1)
    n = len(taps) * 2 * (2 if loraks else 1)
    out = S.undo_data_matrix(np.eye(n), d, taps, loraks, (2, 1, 3))
    assert np.allclose(out, d, rtol=0, atol=1e-12)


def test_undo_width1_is_coil_mixing():
    # one tap (s = 0): ksp(v,:) = ksp(v,:) F
    d = _c((4, 5, 3, 3), 2)
    taps, _ = P.kernel_taps(1, 1)
    assert len(taps) == 1
    F = _c((3, 3), 3)
    out = S.undo_data_matrix(F, d, taps)
    assert np.allclose(out, np.einsum('xyzc,cd->xyzd', d, F), atol=1e-12)


@pytest.mark.parametrize("center", [(1, 1, 1), (3, 2, 4)])
def test_undo_width1_loraks_closed_form(center):
    # one tap with loraks, F = [[F11 F12],[F21 F22]] (nc x nc blocks):
    # ksp' = (D (F11 + conj F22) + g (F21 + conj F12)) / 2, g = conj(D(flip))
    nc = 2
    d = _c((5, 4, 6, nc), 4)
    taps, _ = P.kernel_taps(1, 1)
    F = _c((2 * nc, 2 * nc), 5)
    F11, F12, F21, F22 = F[:nc, :nc], F[:nc, nc:], F[nc:, :nc], F[nc:, nc:]
    fl = np.ix_(*[P.flip_index(n, c) for n, c in zip(d.shape[:3], center)])
    g = np.conj(d[fl])
    want = (np.einsum('xyzc,cd->xyzd', d, F11 + F22.conj())
            + np.einsum('xyzc,cd->xyzd', g, F21 + F12.conj())) / 2
    got = S.undo_data_matrix(F, d, taps, True, center)
    assert np.allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("loraks,center", [(False, (1, 1, 1)), (True, (1, 1, 1)), (True, (2, 3, 4))])
def test_image_domain_form_equals_literal_loop(loraks, center):
    x = _c((6, 5, 7, 2), 6)
    taps, _ = P.kernel_taps(3, 1)
    n = len(taps) * 2 * (2 if loraks else 1)
    F = _herm(n, 7)
    want = S.undo_data_matrix(F, S.fft3(x), taps, loraks, center)
    got = S.fft3(S.undo_image(F, x, taps, loraks, center))
    assert np.allclose(got, want, rtol=1e-11, atol=1e-9 * np.abs(want).max())


def test_singular_values_match_explicit_hankel_svd():
    d = _c((5, 5, 4, 2), 8)
    taps, _ = P.kernel_taps(3, 1)
    A = S.hankel_matrix(d, taps)
    sv = np.linalg.svd(A, compute_uv=False)
    lam = np.linalg.eigvalsh(P.gram_fft(d, taps))
    assert np.allclose(np.sort(np.sqrt(np.abs(lam)))[::-1], sv, rtol=1e-10)
    Al = S.hankel_matrix(d, taps, True, (2, 1, 3))
    svl = np.linalg.svd(Al, compute_uv=False)
    laml = np.linalg.eigvalsh(P.gram_fft_loraks(d, taps, (2, 1, 3)))
    assert np.allclose(np.sort(np.sqrt(np.abs(laml)))[::-1], svl, rtol=1e-9, atol=1e-9 * svl[0])


@pytest.fixture(scope="module")
def small():
    N = 8
    om = synth.koosh_ball(N, 30, 2 * N)
    orc = NufftOracle(om, N, J=4, u=2, radial=1)
    img = np.stack([synth.phantom_complex(N) * (1 + 0.4 * c) for c in range(2)], axis=-1)
    data = orc.fNUFT(img)
    data = data + 0.01 * _c(data.shape, 9)
    return orc, data


def test_zero_noise_floor_one_iteration_is_data_consistency(small):
    # noise = 0: sigma = 0, f = 1 on the (full) range, F = I, so iteration 1
    # is ksp = 0 - fft3(iNUFT(fNUFT(0) - data, 10, damp)) (sake.m:108-111)
    orc, data = small
    got = S.sake(orc, data, noise=0.0, maxit=1, damp=0.1)
    want = S.fft3(orc.iNUFT(data, 10, 0.1))
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("loraks", [False, True])
def test_fast_forms_equal_literal_loop(small, loraks):
    orc, data = small
    a, ia = S.sake(orc, data, maxit=2, loraks=loraks, return_info=True)
    b, ib = S.sake(orc, data, maxit=2, loraks=loraks, return_info=True, literal=True)
    assert ia['noise'] == pytest.approx(ib['noise'], rel=1e-9)
    assert np.allclose(a, b, rtol=1e-9, atol=1e-9 * np.abs(b).max())


def test_history_and_convergence_rule(small):
    orc, data = small
    ksp, info = S.sake(orc, data, maxit=4, tol=1e-12, return_info=True)
    h = info['history']
    assert len(h) == 4 and h[0]['tol'] == 1e-12              # sake.m:155-156
    for a, b in zip(h, h[1:]):
        assert b['tol'] == pytest.approx(abs(b['norm_inf'] - a['norm_inf']) / b['norm_inf'])
    # a loose tolerance stops at iteration 2 (the first with a real change)
    _, info2 = S.sake(orc, data, maxit=10, tol=1.0, return_info=True)
    assert info2['iters'] == 2
    assert info['noise'] > 0 and info['sigma'] > 0


def test_filtered_spectrum_shrinks_the_tail(small):
    # the minimum-variance filter removes the noise floor: the Hankel matrix
    # of the returned k-space has a smaller tail than the data-consistent one
    orc, data = small
    taps, _ = P.kernel_taps(4, 1)
    ksp = S.sake(orc, data, maxit=1)
    dc = S.fft3(orc.iNUFT(data, 10, 0.0))
    s_out = np.sqrt(np.abs(np.linalg.eigvalsh(P.gram_fft(ksp, taps))))[::-1]
    s_dc = np.sqrt(np.abs(np.linalg.eigvalsh(P.gram_fft(dc, taps))))[::-1]
    half = len(s_out) // 2
    assert np.sum(s_out[half:]) < 0.7 * np.sum(s_dc[half:])


def test_argument_checks(small):
    orc, data = small
    with pytest.raises(ValueError):
        S.sake(orc, data[:, :1], maxit=1)
    with pytest.raises(ValueError):
        S.sake(orc, data[:-1], maxit=1)
    with pytest.raises(NotImplementedError):
        S.sake(orc, data, maxit=1, sparsity=0.1)
