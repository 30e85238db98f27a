"""iNUFT 'compressed-sensing' (iNUFT.m:170-189: FISTA with a Haar sparsity
prox) on the HIP path vs the FP64 oracle (GPU).

The external fistaN/HWT are restated (oracle/constraint_oracle.py); parity
here is against that restatement: FP64 plans to 1e-9 relative L2 with the
same iteration count, FP32 plans to the 1e-4 gate.
"""
import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import constraint_oracle as PC
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu

N = 12


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def setup():
    om = synth.koosh_ball(N, 40, 2 * N)
    orc = O.NufftOracle(om, N, J=4, u=2, radial=1)
    rng = np.random.default_rng(6)
    img = np.stack([synth.phantom_complex(N) * (1 + 0.25 * c) for c in range(2)], axis=-1)
    data = orc.fNUFT(img)
    data = data + 0.02 * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))
    print("cs case N", N, "M", om.shape[1], "nc", data.shape[1], flush=True)
    return om, orc, data


@pytest.mark.parametrize("nc,lam", [(1, 0.3), (2, 0.3), (2, 0.7)])
def test_cs_fp64_matches_oracle(setup, nc, lam):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    d = data[:, :nc]
    want, wi = PC.inufft_cs(orc, d, 15, 0.05, lam, return_info=True)
    got, gi = g.iNUFT(d, 15, 0.05, None, "compressed-sensing", lam, return_info=True)
    assert gi["iters"] == [wi["iters"]]
    err = _rel(got, want)
    print("cs fp64 nc", nc, "lambda", lam, "iters", wi["iters"], "rel err", err, flush=True)
    assert err < 1e-9


def test_cs_fp32_matches_oracle(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1, fp64_compute=False)
    want = PC.inufft_cs(orc, data.astype(np.complex64), 15, 0.05, 0.3)
    got = g.iNUFT(data, 15, 0.05, None, "compressed-sensing", 0.3)
    assert got.dtype == np.complex64
    err = _rel(got.astype(np.complex128), want)
    print("cs fp32 rel err vs FP64 oracle", err, flush=True)
    assert err < 1e-4


def test_no_sparsity_converges_to_least_squares(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    x = g.iNUFT(data, 400, 1.0, None, "compressed-sensing", 0.0, tol=1e-12)
    ls = g.iNUFT(data, 400, 1.0, tol=1e-13)
    assert _rel(x, ls) < 1e-9


def test_result_has_the_requested_sparsity(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    x = g.iNUFT(data, 20, 0.0, None, "compressed-sensing", 0.5)
    w = PC.haar_fwd(x)
    frac = np.count_nonzero(np.abs(w) < 1e-12 * np.abs(w).max()) / w.size
    assert 0.35 < frac <= 0.5 + 1e-9


def test_lambda_out_of_range(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1)
    with pytest.raises(ValueError):
        g.iNUFT(data, 10, 0.0, None, "compressed-sensing", 1.5)


def test_cs_on_a_2d_plan_matches_oracle():
    # 2-D plan (N3 = 1): the Haar transform runs on x and y only
    r = np.random.default_rng(17)
    om = np.vstack([r.uniform(-8, 8, size=(2, 800)), np.zeros((1, 800))])
    orc = O.NufftOracle(om, 16)
    g = nufft_3d(om, 16, gpu=2)
    x = r.normal(size=(16, 16, 1, 2)) + 1j * r.normal(size=(16, 16, 1, 2))
    data = orc.fNUFT(x)
    want, wi = PC.inufft_cs(orc, data, 12, 0.05, 0.4, return_info=True)
    got, gi = g.iNUFT(data, 12, 0.05, None, "compressed-sensing", 0.4, return_info=True)
    assert gi["iters"] == [wi["iters"]]
    assert _rel(got, want) < 1e-9
