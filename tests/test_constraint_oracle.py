"""Pins the phase-constraint oracle (iNUFT.m:144-166, nufft_3d.m:325-334) on the CPU."""
import numpy as np
import pytest
from scipy import signal

from oracle import constraint_oracle as PC
from oracle.nufft_oracle import NufftOracle
from nufft_3d_amd import synth


def _c(shape, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=shape) + 1j * r.normal(size=shape)


def _small_operator(n, lam, seed):
    """A(x) = conj(P) M (P x) + i lam^2 Im(x), M Hermitian PD, |P| = 1."""
    a = _c((n, n), seed)
    M = a @ a.conj().T + n * np.eye(n)
    P = np.exp(1j * np.random.default_rng(seed + 1).uniform(0, 2 * np.pi, n))

    def A(X):
        return np.conj(P)[:, None] * (M @ (P[:, None] * X)) + 1j * lam ** 2 * np.imag(X)
    return A


def _real_matrix(A, n):
    """The 2n x 2n real matrix of a real-linear map on C^n ([Re; Im] stacking)."""
    cols = []
    for k in range(2 * n):
        e = np.zeros(n, dtype=np.complex128)
        e[k % n] = 1.0 if k < n else 1j
        y = A(e[:, None])[:, 0]
        cols.append(np.concatenate([y.real, y.imag]))
    return np.array(cols).T


@pytest.mark.parametrize("lam", [0.0, 1.5])
def test_pcgpc_iterates_minimise_the_energy_error_over_krylov_spaces(lam):
    n = 6
    A = _small_operator(n, lam, 3)
    AR = _real_matrix(A, n)
    assert np.allclose(AR, AR.T, atol=1e-10)                    # self-adjoint under Re<.,.>
    assert np.min(np.linalg.eigvalsh(0.5 * (AR + AR.T))) > 0   # positive definite
    b = _c((n, 1), 4)
    bR = np.concatenate([b[:, 0].real, b[:, 0].imag])
    xs = np.linalg.solve(AR, bR)
    for k in range(1, 2 * n + 1):
        x, info = PC.pcgpc(A, b, tol=0.0, maxit=k)
        K = [bR]
        for _ in range(k - 1):
            K.append(AR @ K[-1])
        K = np.linalg.qr(np.array(K).T)[0]
        c = np.linalg.solve(K.T @ AR @ K, K.T @ bR)
        want = K @ c
        got = np.concatenate([x[:, 0].real, x[:, 0].imag])
        assert np.allclose(got, want, rtol=1e-7, atol=1e-9 * np.linalg.norm(xs)), k
    assert np.allclose(got, xs, rtol=1e-8)                      # exact at k = 2n
    assert info["iters"] == [2 * n]


def test_pcgpc_columns_are_independent_and_zero_rhs():
    n = 5
    A = _small_operator(n, 0.7, 5)
    b = _c((n, 3), 6)
    b[:, 1] = 0
    x, info = PC.pcgpc(A, b, tol=1e-12, maxit=40)
    assert np.all(x[:, 1] == 0) and info["iters"][1] == 0
    for c in (0, 2):
        xc, _ = PC.pcgpc(A, b[:, [c]], tol=1e-12, maxit=40)
        assert np.allclose(x[:, c], xc[:, 0], rtol=1e-13, atol=0)
        assert np.linalg.norm(A(x[:, [c]])[:, 0] - b[:, c]) <= 1e-10 * np.linalg.norm(b[:, c])


def test_phase_estimate_is_the_zero_padded_box_sum_phase():
    x = _c((5, 6, 4, 2), 7)
    x[0, 0, 0, 0] = 0
    P = PC.phase_estimate(x)
    for c in range(2):
        s = signal.convolve(x[..., c], np.ones((3, 3, 3)), mode='same')
        assert np.allclose(P[..., c], np.exp(1j * np.angle(s)), atol=1e-12)
    assert np.allclose(np.abs(P), 1)
    z = np.zeros((3, 3, 3, 1), dtype=complex)
    assert np.all(PC.phase_estimate(z) == 1)                     # angle(0) = 0


@pytest.fixture(scope="module")
def small():
    N = 8
    om = synth.koosh_ball(N, 30, 2 * N)
    orc = NufftOracle(om, N, J=4, u=2, radial=1)
    img = synth.phantom_complex(N)[..., None] * np.array([1.0, 0.6 + 0.3j])
    data = orc.fNUFT(img) + 0.01 * _c((om.shape[1], 2), 8)
    return orc, data


def test_pprojection_self_adjoint_and_definite(small):
    orc, data = small
    N = orc.N
    P = PC.phase_estimate(_c(N + (2,), 9))
    x, y = _c(N + (2,), 10), _c(N + (2,), 11)
    Ax = PC.pprojection(orc, x, 0.1, 0.8, orc.d, P)
    Ay = PC.pprojection(orc, y, 0.1, 0.8, orc.d, P)
    assert np.real(np.vdot(y, Ax)) == pytest.approx(np.real(np.vdot(Ay, x)), rel=1e-10)
    assert np.real(np.vdot(x, Ax)) > 0


def test_lambda_zero_recovers_the_least_squares_solution(small):
    # lambda = 0: P'A'WAP is unitarily similar to A'WA, so the converged
    # constrained solve returns the converged plain solution
    orc, data = small
    im = PC.inufft_phase(orc, data, 200, 0.3, 0.0, tol=1e-12)
    ls = orc.iNUFT(data, 400, 0.3, tol=1e-13)
    assert np.allclose(im, ls, rtol=1e-7, atol=1e-8 * np.abs(ls).max())


def test_large_lambda_pulls_the_solution_onto_the_phase(small):
    orc, data = small
    x0 = orc.iNUFT(data, 20, 0.0)
    P = PC.phase_estimate(x0)
    weak = PC.inufft_phase(orc, data, 20, 0.0, 0.0)
    strong = PC.inufft_phase(orc, data, 20, 0.0, 10.0)
    off = lambda im: np.linalg.norm(np.imag(np.conj(P) * im)) / np.linalg.norm(im)   # noqa: E731
    assert off(strong) < 0.1 * off(weak)


# ---------------------------------------------------------------- compressed sensing

def _haar_matrix(n):
    """One-level analysis matrix of length n: low rows (e_2i + e_2i+1)/sqrt2, then high rows."""
    H = np.zeros((n, n))
    for i in range(n // 2):
        H[i, 2 * i] = H[i, 2 * i + 1] = np.sqrt(0.5)
        H[n // 2 + i, 2 * i] = np.sqrt(0.5)
        H[n // 2 + i, 2 * i + 1] = -np.sqrt(0.5)
    return H


def test_haar_one_level_is_the_kronecker_butterfly():
    x = _c((4, 6, 2, 1), 12)
    w = PC.haar_fwd(x, levels=1)
    # column-major vec: x fastest, so the 3-D transform is Hz (x) Hy (x) Hx
    K = np.kron(_haar_matrix(2), np.kron(_haar_matrix(6), _haar_matrix(4)))
    want = K @ x[..., 0].ravel(order='F')
    assert np.allclose(w[..., 0].ravel(order='F'), want, atol=1e-13)


def test_haar_multilevel_orthonormal_and_nested():
    x = _c((8, 16, 4, 2), 13)
    assert PC.haar_levels((8, 16, 4)) == 2 and PC.haar_levels((16, 16, 16)) == 3
    assert PC.haar_levels((12, 12, 1)) == 2 and PC.haar_levels((7, 8, 8)) == 0
    w = PC.haar_fwd(x)
    assert np.linalg.norm(w) == pytest.approx(np.linalg.norm(x), rel=1e-13)
    assert np.allclose(PC.haar_inv(w), x, atol=1e-13)
    w1 = PC.haar_fwd(x, levels=1)
    # level 2 only transforms the low corner of level 1
    keep = np.ones(x.shape[:3], bool)
    keep[:4, :8, :2] = False
    assert np.array_equal(w[keep], w1[keep])
    assert np.allclose(w[:4, :8, :2], PC.haar_fwd(w1[:4, :8, :2], levels=1), atol=1e-13)


def test_sparsity_thresh_zero_count_and_shrinkage():
    w = _c((6, 5, 4, 2), 14)
    out, t = PC.sparsity_thresh(w, 0.3)
    nz = int(round(0.3 * w.size))
    assert np.count_nonzero(out == 0) == nz
    m = np.abs(w) > t
    assert np.allclose(np.abs(out[m]), np.abs(w[m]) - t)
    assert np.allclose(np.angle(out[m]), np.angle(w[m]))
    same, t0 = PC.sparsity_thresh(w, 0.0)
    assert t0 == 0.0 and np.array_equal(same, w)


def test_fista_without_sparsity_converges_to_least_squares(small):
    orc, data = small
    # damp = 1 makes the problem strongly convex (linear rate)
    x, info = PC.inufft_cs(orc, data, 400, 1.0, 0.0, tol=1e-12, return_info=True)
    ls = orc.iNUFT(data, 400, 1.0, tol=1e-13)
    assert info['iters'] < 400
    assert np.allclose(x, ls, rtol=1e-9, atol=1e-10 * np.abs(ls).max())


def test_fista_sparsity_zeroes_the_requested_fraction(small):
    orc, data = small
    x = PC.inufft_cs(orc, data, 30, 0.0, 0.5)
    w = PC.haar_fwd(x)
    frac = np.count_nonzero(np.abs(w) < 1e-12 * np.abs(w).max()) / w.size
    assert 0.35 < frac <= 0.5 + 1e-9
