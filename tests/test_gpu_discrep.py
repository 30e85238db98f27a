"""discrep (discrep.m:2-69) on the HIP path vs the FP64 oracle (GPU).

Every round of the damp update is compared (the per-round history), so the
cold first solve and the warm-started later solves (minres with x0 =
the previous round's x, discrep.m:60) are each checked. FP64 plans agree
with the oracle to 1e-9 relative (MINRES in a different summation order);
FP32 plans to 1e-3 (fp32 solves, 20 rounds of a contracting fixed point).
"""
import warnings

import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu

N = 12


@pytest.fixture(scope="module")
def setup():
    om = synth.koosh_ball(N, 40, 2 * N)
    orc = O.NufftOracle(om, N, J=4, u=2, radial=1)
    rng = np.random.default_rng(1)
    img = np.stack([synth.phantom_complex(N) * (1 + 0.3 * c) for c in range(2)], axis=-1)
    data = orc.fNUFT(img)
    data = data + 0.02 * (rng.normal(size=data.shape) + 1j * rng.normal(size=data.shape))
    print("discrep case N", N, "M", om.shape[1], "nc", data.shape[1], flush=True)
    return om, orc, data


@pytest.mark.parametrize("gpu", [2, 1])
def test_discrep_given_stddev_matches_oracle(setup, gpu):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=gpu)
    want, winfo = orc.discrep(data, stddev=0.02, return_info=True)
    got, ginfo = g.discrep(data, stddev=0.02, return_info=True)
    tol = 1e-9 if gpu == 2 else 1e-3
    assert ginfo["stddev"] == 0.02
    h_want = np.array(winfo["damp_history"])
    h_got = np.array(ginfo["damp_history"])
    assert h_got.shape == (20,)
    assert np.max(np.abs(h_got - h_want) / h_want) < tol
    assert got == pytest.approx(want, rel=tol)


@pytest.mark.parametrize("gpu", [2, 1])
def test_discrep_guessed_stddev(setup, gpu):
    om, orc, data = setup
    d = data.copy()
    d[::5] = 0                           # zero entries are left out of the guess
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=gpu)
    want, winfo = orc.discrep(d.astype(g.cdtype), iters=3, return_info=True)
    with pytest.warns(UserWarning, match="stddev not supplied"):
        got, ginfo = g.discrep(d, iters=3, return_info=True)
    if gpu == 2:
        assert ginfo["stddev"] == winfo["stddev"]
        assert got == pytest.approx(want, rel=1e-9)
    else:
        assert ginfo["stddev"] == pytest.approx(winfo["stddev"], rel=1e-6)
        assert got == pytest.approx(want, rel=1e-3)


def test_discrep_torch_input_matches_host(setup):
    torch = pytest.importorskip("torch")
    om, orc, data = setup
    d = data[:, :1].copy()
    d[1::2] = 0
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    want = orc.discrep(d, iters=2, return_info=True)[1]["stddev"]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        host = g.discrep(d, iters=2, return_info=True)
        dt = torch.from_numpy(np.ascontiguousarray(d.T)).cuda()
        dev = g.discrep(dt, iters=2, return_info=True)
    assert host[1]["stddev"] == want
    assert dev[0] == host[0] and dev[1]["stddev"] == host[1]["stddev"]


def test_discrep_all_zero_data_without_stddev_is_an_error(setup):
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=1)
    with pytest.raises(Exception, match="no nonzero entries"):
        g.discrep(np.zeros_like(data), iters=1)


def test_discrep_accepts_readout_by_spoke_layout(setup):
    # data as [nr ny nc] with nr*ny = M (discrep.m:14-20) gives the same damp
    om, orc, data = setup
    g = nufft_3d(om, N, J=4, u=2, radial=1, gpu=2)
    nr = 2 * N
    d3 = np.reshape(data, (nr, om.shape[1] // nr, data.shape[1]), order="F")
    a = g.discrep(data, stddev=0.02, iters=3)
    b = g.discrep(d3, stddev=0.02, iters=3)
    assert a == b
    with pytest.raises(ValueError):
        g.discrep(data[:-1], stddev=0.02, iters=1)
