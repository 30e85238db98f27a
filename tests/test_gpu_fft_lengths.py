"""Grid and image lengths with prime factors 11 and 13 (direct DFT
butterflies in the stage loop) and above 13 (Bluestein chirp-z transform)
vs the FP64 oracle: fNUFT, aNUFT, iNUFT, operator adjointness, and the
unshifted image-size transform inside the FFT Hankel Gram.

Tolerances (relative L2) as the smooth-length parity tests: FP64 plans
1e-12 per operator (1e-10 through MINRES), FP32 plans 5e-5 per operator
and 1e-4 end to end.
"""
import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import nufft_oracle as O
from oracle import pruno_oracle as P

pytestmark = pytest.mark.gpu

TOL32_OP = 5e-5
TOL32 = 1e-4
TOL64 = 1e-12


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rand(shape, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=shape) + 1j * r.normal(size=shape)


def _box(N, M, seed):
    # samples over the whole k-space box (grid units, |om| <= N/2)
    half = np.asarray(N if np.ndim(N) else (N,) * 3, dtype=float)[:, None] / 2
    return np.random.default_rng(seed).uniform(-1, 1, size=(3, M)) * half


CASES = {
    # name: (om builder, N, kwargs); K = 2*ceil(N*u/2) per axis
    "K44_radix11": (lambda: synth.koosh_ball(22, 60, 32), 22, dict(J=4, u=2, radial=1)),
    "K52_radix13": (lambda: synth.koosh_ball(26, 60, 32), 26, dict(J=4, u=2, radial=0)),
    "K68_bluestein17": (lambda: synth.koosh_ball(34, 50, 40), 34, dict(J=4, u=2, radial=1)),
    "K58_bluestein29_u125": (lambda: synth.golden_means(46, 60, 48), 46,
                             dict(J=6, u=1.25, radial=1)),
    "mixed_34_22_26": (lambda: _box((34, 22, 26), 4000, 5), (34, 22, 26),
                       dict(J=4, u=2, radial=1)),
    "K220_radix11_u125": (lambda: _box((176, 20, 12), 6000, 6), (176, 20, 12),
                          dict(J=6, u=1.25, radial=0)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    build, N, kw = CASES[request.param]
    om = build()
    orc = O.NufftOracle(om, N, **kw)
    print("case", request.param, "M", om.shape[1], "K", orc.K, flush=True)
    g32 = nufft_3d(om, N, gpu=1, **kw)
    g64 = nufft_3d(om, N, gpu=2, **kw)
    assert tuple(g64.K) == tuple(orc.K)
    return request.param, om, orc, g32, g64


@pytest.mark.parametrize("nc", [1, 3])
def test_forward(case, nc):
    name, om, orc, g32, g64 = case
    x = _rand(orc.N + (nc,), 10 + nc)
    want = orc.fNUFT(x)
    e64, e32 = rel(g64.fNUFT(x), want), rel(g32.fNUFT(x), want)
    print("%s fNUFT nc=%d fp64 %.2e fp32 %.2e" % (name, nc, e64, e32), flush=True)
    assert e64 < TOL64
    assert e32 < TOL32_OP


@pytest.mark.parametrize("nc", [1, 3])
def test_adjoint(case, nc):
    name, om, orc, g32, g64 = case
    k = _rand((orc.M, nc), 20 + nc)
    want = orc.aNUFT(k)
    e64, e32 = rel(g64.aNUFT(k), want), rel(g32.aNUFT(k), want)
    print("%s aNUFT nc=%d fp64 %.2e fp32 %.2e" % (name, nc, e64, e32), flush=True)
    assert e64 < TOL64
    assert e32 < TOL32_OP


def test_inverse(case):
    name, om, orc, g32, g64 = case
    x = _rand(orc.N + (2,), 31)
    raw = orc.fNUFT(x) + 0.01 * _rand((orc.M, 2), 32)
    want, winfo = orc.iNUFT(raw, 6, 0.05, return_info=True)
    got64, info64 = g64.iNUFT(raw, 6, 0.05, return_info=True)
    e32 = rel(g32.iNUFT(raw, 6, 0.05), want)
    print("%s iNUFT fp32 %.2e" % (name, e32), flush=True)
    assert rel(got64, want) < 1e-10
    assert info64["iters"] == winfo["iters"]
    assert e32 < TOL32


def test_operator_adjointness_fp64(case):
    name, om, orc, g32, g64 = case
    x = _rand(orc.N + (1,), 41)
    k = _rand((orc.M, 1), 42)
    fx = g64.fNUFT(x)
    lhs = np.vdot(fx, k)
    rhs = np.vdot(x, g64.aNUFT(k))
    assert abs(lhs - rhs) < 1e-12 * np.linalg.norm(fx) * np.linalg.norm(k)


@pytest.mark.parametrize("N", [(34, 22, 26), (46, 12, 10)])
def test_fft_gram_on_image_lengths_with_large_prime_factors(N):
    # The FP64 Gram of image-size data runs through the unshifted N-point
    # transforms of the plan (lengths 34 = 2*17, 46 = 2*23: Bluestein;
    # 22, 26: radices 11, 13).
    om = _box(N, 500, 7)
    g64 = nufft_3d(om, N, gpu=2)
    d = _rand(N + (2,), 8)
    taps, _ = P.kernel_taps(3, 1)
    want = P.make_data_matrix(d, taps)
    e = rel(g64.gram(d, 3, 1), want)
    print("gram N=%s rel err %.2e" % (N, e), flush=True)
    assert e < 1e-12
