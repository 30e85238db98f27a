"""iNUFT 'phase-constraint' (iNUFT.m:144-166, nufft_3d.m:325-334) on the HIP path
vs the FP64 oracle (GPU).

FP64 plans agree with the oracle to 1e-9 relative L2 and take the same CG
iteration counts; FP32 plans to the 1e-4 gate (fp32 MINRES start, fp32 CG).
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
    rng = np.random.default_rng(5)
    img = np.stack([synth.phantom_complex(N) * np.exp(0.5j * c) * (1 - 0.2 * c) for c in range(2)],
                   axis=-1)
    data = orc.fNUFT(img)
    data = data + 0.02 * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))
    print("phase case N", N, "M", om.shape[1], "nc", data.shape[1], flush=True)
    return om, orc, data


@pytest.mark.parametrize("nc,lam", [(1, 0.5), (2, 0.5), (2, 3.0)])
def test_phase_fp64_matches_oracle(setup, nc, lam):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    d = data[:, :nc]
    want, wi = PC.inufft_phase(orc, d, 10, 0.05, lam, return_info=True)
    got, gi = g.iNUFT(d, 10, 0.05, None, "phase-constraint", lam, return_info=True)
    assert gi["iters"] == wi["iters"]
    err = _rel(got, want)
    print("phase fp64 nc", nc, "lambda", lam, "rel err", err, flush=True)
    assert err < 1e-9


def test_phase_fp32_matches_oracle(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1, fp64_compute=False)
    want = PC.inufft_phase(orc, data.astype(np.complex64), 10, 0.05, 0.5)
    got = g.iNUFT(data, 10, 0.05, None, "phase-constraint", 0.5)
    assert got.dtype == np.complex64
    err = _rel(got.astype(np.complex128), want)
    print("phase fp32 rel err vs FP64 oracle", err, flush=True)
    assert err < 1e-4


def test_lambda_zero_converges_to_least_squares(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    im = g.iNUFT(data, 300, 0.3, None, "phase-constraint", 0.0, tol=1e-11)
    ls = g.iNUFT(data, 300, 0.3, tol=1e-12)
    assert _rel(im, ls) < 1e-8


def test_maxit_one_is_the_plain_solve(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    with pytest.warns(UserWarning, match="maxit>1"):
        a = g.iNUFT(data, 1, 0.0, None, "phase-constraint", 1.0)
    b = g.iNUFT(data, 1, 0.0)
    assert np.array_equal(a, b)


def test_torch_input_matches_host(setup):
    torch = pytest.importorskip("torch")
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1)
    host = g.iNUFT(data, 8, 0.05, None, "phase-constraint", 0.5)
    dt = torch.from_numpy(np.ascontiguousarray(data.T.astype(np.complex64))).cuda()
    dev = g.iNUFT(dt, 8, 0.05, None, "phase-constraint", 0.5)
    d = dev.cpu().numpy().transpose(3, 2, 1, 0)
    assert _rel(d, host) < 1e-5


def test_phase_on_a_2d_plan_matches_oracle():
    # all om(3,:)==0 selects the 2-D path (nufft_3d.m:104-106); the phase box
    # sum then covers the 3x3 in-plane neighbours
    r = np.random.default_rng(16)
    om = np.vstack([r.uniform(-8, 8, size=(2, 800)), np.zeros((1, 800))])
    orc = O.NufftOracle(om, 16)
    g = nufft_3d(om, 16, gpu=2)
    assert g.N == (16, 16, 1)
    x = r.normal(size=(16, 16, 1, 2)) + 1j * r.normal(size=(16, 16, 1, 2))
    data = orc.fNUFT(x)
    want, wi = PC.inufft_phase(orc, data, 8, 0.05, 0.5, return_info=True)
    got, gi = g.iNUFT(data, 8, 0.05, None, "phase-constraint", 0.5, return_info=True)
    assert gi["iters"] == wi["iters"]
    assert _rel(got, want) < 1e-9
