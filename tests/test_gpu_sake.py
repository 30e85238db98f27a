"""SAKE Cadzow loop (sake.m:100-171,227-284) on the HIP path vs the FP64 oracle (GPU).

FP64 plans agree with the oracle to 1e-9 relative L2 after several
iterations (MINRES, Gram and eigen-decomposition in a different summation
order); FP32 plans (fp32 MINRES, split-BF16 MFMA Gram) to the 1e-4 gate. The
image-domain averaging the device uses is pinned against the literal
circshift loop of sake.m:227-284 on the CPU (tests/test_sake_oracle.py).
"""
import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import nufft_oracle as O
from oracle import sake_oracle as S

pytestmark = pytest.mark.gpu

N = 12


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def setup():
    om = synth.koosh_ball(N, 40, 2 * N)
    orc = O.NufftOracle(om, N, J=4, u=2, radial=1)
    rng = np.random.default_rng(3)
    img = np.stack([synth.phantom_complex(N) * (1 + 0.3 * c) * np.exp(0.4j * c) for c in range(3)],
                   axis=-1)
    data = orc.fNUFT(img)
    data = data + 0.02 * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))
    print("sake case N", N, "M", om.shape[1], "nc", data.shape[1], flush=True)
    return om, orc, data


@pytest.mark.parametrize("loraks,center", [(False, None), (True, None), (True, (3, 2, 5))])
def test_sake_fp64_matches_oracle(setup, loraks, center):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    print("sake fp64 loraks", loraks, "center", center, flush=True)
    want, wi = S.sake(orc, data, maxit=3, loraks=loraks, center=center, return_info=True)
    got, gi = g.sake(data, maxit=3, loraks=loraks, center=center, return_info=True)
    assert gi["iters"] == wi["iters"] == 3
    assert gi["noise"] == pytest.approx(wi["noise"], rel=1e-9)
    for a, b in zip(gi["history"], wi["history"]):
        assert a["norm_inf"] == pytest.approx(b["norm_inf"], rel=1e-9)
        assert a["norm_1"] == pytest.approx(b["norm_1"], rel=1e-9)
        assert a["rank"] == b["rank"]
    err = _rel(got, want)
    print("rel err vs FP64 oracle", err, flush=True)
    assert err < 1e-9


@pytest.mark.parametrize("loraks", [False, True])
def test_sake_fp32_matches_oracle(setup, loraks):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1, fp64_compute=False)
    want = S.sake(orc, data.astype(np.complex64), maxit=2, loraks=loraks)
    got = g.sake(data, maxit=2, loraks=loraks)
    assert got.dtype == np.complex64
    err = _rel(got.astype(np.complex128), want)
    print("fp32 rel err vs FP64 oracle (loraks=%s)" % loraks, err, flush=True)
    assert err < 1e-4


def test_zero_noise_first_iteration_is_data_consistency(setup):
    # noise = 0 gives F = I, so one iteration returns fft3(iNUFT(data, 10, damp))
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    got = g.sake(data, maxit=1, noise=0.0, damp=0.05)
    want = S.fft3(g.iNUFT(data, 10, 0.05))
    assert _rel(got, want) < 1e-10


def test_given_noise_and_convergence_stop(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    # the oracle stops at iteration 4 (changes 5.4e-3, 2.15e-3, 1.90e-3 of the
    # largest singular value against tol 2e-3)
    want, wi = S.sake(orc, data, maxit=20, noise=1.5, tol=2e-3, return_info=True)
    got, gi = g.sake(data, maxit=20, noise=1.5, tol=2e-3, return_info=True)
    assert gi["iters"] == wi["iters"] == 4
    assert gi["noise"] == 1.5
    assert gi["sigma"] == pytest.approx(wi["sigma"], rel=1e-12)
    assert _rel(got, want) < 1e-9


def test_image_output_and_inufft_branch(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    ksp = g.sake(data, maxit=2)
    img = g.sake(data, maxit=2, output="image")
    assert _rel(S.fft3(img), ksp) < 1e-12
    # iNUFT(raw,maxit,damp,W,'parallel-imaging-sake',lambda) (iNUFT.m:191-196)
    im2 = g.iNUFT(data, 20, 0.0, None, "parallel-imaging-sake", 0.0, sake_opts={"maxit": 2})
    assert _rel(im2, img) < 1e-12
    iml = g.iNUFT(data, 20, 0.0, None, "parallel-imaging-sake", 0.5, sake_opts={"maxit": 2})
    assert _rel(iml, g.sake(data, maxit=2, loraks=True, output="image")) < 1e-12


def test_torch_input_matches_host(setup):
    torch = pytest.importorskip("torch")
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1)
    host = g.sake(data, maxit=2)
    dt = torch.from_numpy(np.ascontiguousarray(data.T.astype(np.complex64))).cuda()
    dev = g.sake(dt, maxit=2)
    assert tuple(dev.shape) == (data.shape[1], N, N, N)
    d = dev.cpu().numpy().transpose(3, 2, 1, 0)
    assert _rel(d, host) < 1e-5


def test_argument_errors(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1)
    with pytest.raises(ValueError):
        g.sake(data[:, :1], maxit=1)                     # sake.m:57-59
    with pytest.raises(Exception, match="DWT"):
        g.sake(data, maxit=1, sparsity=0.1)
    with pytest.raises(Exception, match="center"):
        g.sake(data, maxit=1, center=(N + 1, 1, 1))
    with pytest.warns(UserWarning, match="not a valid option"):
        g.sake(data, maxit=1, bogus=3)
    with pytest.raises(ValueError, match="multiple coils"):
        g.iNUFT(data[:, :1], 20, 0.0, None, "parallel-imaging-sake", 0.0)


def test_2d_plan_is_reported_unsupported():
    # sake.m builds 3-D taps always (sake.m:62); a 2-D plan (N3 = 1) would
    # wrap every z tap onto the same plane: refused with a clear status
    r = np.random.default_rng(18)
    om = np.vstack([r.uniform(-8, 8, size=(2, 800)), np.zeros((1, 800))])
    g = nufft_3d(om, 16, gpu=2)
    data = r.normal(size=(800, 2)) + 1j * r.normal(size=(800, 2))
    with pytest.raises(Exception, match="kernel half width"):
        g.sake(data, maxit=1)
