"""SAKE/PRUNO calibration on the GPU vs the FP64 oracle: Hankel Gram (FP64
path at 1e-12, split-BF16 MFMA path at 1e-4), PRUNO spectrum and nv, and the
coupled PRUNO reconstruction."""
import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import pruno_oracle as P
from oracle.nufft_oracle import NufftOracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _data(shape, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=shape) + 1j * r.normal(size=shape)


@pytest.fixture(scope="module")
def plans():
    om = synth.koosh_ball(8, 40, 16)
    return om, nufft_3d(om, 8, gpu=1), nufft_3d(om, 8, gpu=2)


@pytest.mark.parametrize("width,radial", [(3, 1), (3, 0), (4, 1)])
def test_gram_small(plans, width, radial):
    om, g32, g64 = plans
    d = _data((8, 6, 5, 2), 1)
    taps, _ = P.kernel_taps(width, radial)
    want = P.make_data_matrix(d, taps)
    assert rel(g64.gram(d, width, radial), want) < 1e-12
    got = g32.gram(d.astype(np.complex64), width, radial)
    assert rel(got, want.astype(np.complex64).astype(np.complex128)) < 1e-4
    assert np.allclose(got, got.conj().T, rtol=0, atol=1e-6 * np.abs(want).max())


def test_gram_bf16x3_many_tiles(plans):
    # n = 4 coils x 81 taps = 324 -> a 648x648 real Gram: 6 tile rows, V split
    om, g32, g64 = plans
    d = _data((24, 24, 24, 4), 2).astype(np.complex64)
    taps, _ = P.kernel_taps(5, 1)
    want = P.gram_fft(d.astype(np.complex128), taps)
    got = g32.gram(d, 5, 1)
    e = rel(got, want)
    print("bf16x3 gram rel err %.3e" % e, flush=True)
    assert e < 1e-4


def test_pruno_spectrum_and_nv(plans):
    om, g32, g64 = plans
    d = _data((8, 8, 8, 3), 3)
    ref = P.pruno(d, (8, 8, 8), width=3, noise=0.05)
    _, got = g64.pruno(d, width=3, noise=0.05, output="taps", return_info=True)
    assert got["nv"] == ref["nv"]
    assert rel(got["sv"], ref["sv"]) < 1e-10
    est_ref = P.pruno(d, (8, 8, 8), width=3)
    _, est = g64.pruno(d, width=3, output="taps", return_info=True)
    assert est["noise"] == pytest.approx(est_ref["noise"], rel=1e-9)
    _, got32 = g32.pruno(d.astype(np.complex64), width=3, noise=0.05, output="taps",
                         return_info=True)
    assert rel(got32["sv"], ref["sv"]) < 1e-4


def _voxel_gram(K):
    # sum_v K_v K_v^H per voxel: invariant to the phase of each singular
    # vector (and to rotations inside equal-filter subspaces)
    return np.einsum('xyzcv,xyzdv->xyzcd', K, K.conj())


@pytest.mark.parametrize("gpu,literal,radial,N", [(2, 0, 1, (8, 8, 8)), (2, 1, 1, (8, 8, 8)),
                                                  (2, 0, 0, (10, 8, 6)), (1, 0, 1, (8, 8, 8))])
def test_pruno_returns_the_image_space_null_kernels(gpu, literal, radial, N):
    # K = pruno(obj,data) is [N1 N2 N3 nc nv] (pruno.m:2,124-143): compared
    # with the oracle through the tap-domain projector Kt Kt' and the
    # per-voxel coil Gram of K, which do not depend on the eigenvectors'
    # phases
    om = synth.koosh_ball(8, 40, 16)
    g = nufft_3d(om, N, gpu=gpu)
    d = _data(N + (3,), 5)
    ref = P.pruno(d, N, width=3, radial=radial, noise=0.35, literal_first_nv=bool(literal))
    if gpu == 1:
        d = d.astype(np.complex64)
    K, info = g.pruno(d, width=3, radial=radial, noise=0.35, literal_first_nv=literal,
                      return_info=True)
    assert info["nv"] == ref["nv"] and 0 < ref["nv"] < len(ref["sv"])
    assert K.shape == ref["K"].shape == N + (3, ref["nv"])
    Kt = info["Kt"]
    tol = 1e-10 if gpu == 2 else 1e-4
    e_taps = rel(Kt @ Kt.conj().T, ref["Kt"] @ ref["Kt"].conj().T)
    e_img = rel(_voxel_gram(K), _voxel_gram(ref["K"]))
    print("pruno K gpu=%d literal=%d radial=%d nv=%d taps %.2e image %.2e"
          % (gpu, literal, radial, ref["nv"], e_taps, e_img), flush=True)
    assert e_taps < tol and e_img < tol
    # the taps output is the same kernels before the expansion
    Kt2 = g.pruno(d, width=3, radial=radial, noise=0.35, literal_first_nv=literal, output="taps")
    assert np.array_equal(Kt2, Kt)


@pytest.mark.parametrize("gpu,tol", [(2, 1e-8), (1, 1e-4)])
def test_inverse_pruno(gpu, tol):
    N, nc = 8, 3
    om = synth.koosh_ball(N, 60, 16)
    orc = NufftOracle(om, N)
    maps = synth.coil_maps(N, nc, seed=1)
    img = synth.phantom_complex(N)[..., None] * maps
    raw = orc.fNUFT(img) + 1e-3 * _data((orc.M, nc), 4)
    want, pr, winfo = P.inufft_pruno(orc, raw, 4, 0.01, 0.3, width=3, noise=1e-3, iters=12)
    g = nufft_3d(om, N, gpu=gpu)
    got, info = g.iNUFT(raw, 4, 0.01, constraint="parallel-imaging-pruno", lambda_=0.3,
                        pruno_opts=dict(width=3, noise=1e-3), pruno_iters=12, return_info=True)
    e = rel(got, want)
    print("pruno inverse gpu=%d rel err %.3e nv=%d" % (gpu, e, pr["nv"]), flush=True)
    assert e < tol


@pytest.mark.parametrize("width,radial,center", [(3, 1, None), (4, 1, (2, 3, 1)), (3, 0, (4, 1, 2))])
def test_gram_loraks_small(plans, width, radial, center):
    # [AA BB; conj(BB) conj(AA)] of sake.m:210-224 against the literal oracle
    om, g32, g64 = plans
    d = _data((8, 6, 5, 2), 5)
    taps, _ = P.kernel_taps(width, radial)
    cen = center or (1, 1, 1)
    want = P.make_data_matrix(d, taps, loraks=True, center=cen)
    got64 = g64.gram(d, width, radial, loraks=True, center=center)
    assert got64.shape == want.shape
    assert rel(got64, want) < 1e-12
    got = g32.gram(d.astype(np.complex64), width, radial, loraks=True, center=center)
    assert rel(got, P.make_data_matrix(d.astype(np.complex64).astype(np.complex128), taps,
                                       loraks=True, center=cen)) < 1e-4


def test_gram_loraks_many_tiles(plans):
    # SAKE's default kernel (width 4, radial: 33 taps) on 4 coils: 2n = 264,
    # a 528-column real Gram for each block, all BB tiles computed
    om, g32, g64 = plans
    d = _data((24, 20, 16, 4), 6).astype(np.complex64)
    taps, _ = P.kernel_taps(4, 1)
    want = P.gram_fft_loraks(d.astype(np.complex128), taps, (5, 1, 9))
    got = g32.gram(d, 4, 1, loraks=True, center=(5, 1, 9))
    e = rel(got, want)
    print("bf16x3 loraks gram rel err %.3e" % e, flush=True)
    assert e < 1e-4


def test_gram_loraks_rejects_a_centre_outside_the_data(plans):
    om, g32, g64 = plans
    d = _data((8, 6, 5, 2), 7)
    with pytest.raises(Exception, match="center outside"):
        g64.gram(d, 3, 1, loraks=True, center=(9, 1, 1))


@pytest.mark.parametrize("loraks,center", [(False, None), (True, (2, 3, 1))])
def test_fp64_gram_by_ffts_at_the_plan_size(plans, loraks, center, monkeypatch):
    # data of the plan's image size takes the cross-correlation (FFT) form;
    # NUFFT3D_GRAM64_DIRECT forces the direct double sums for comparison
    om, g32, g64 = plans
    d = _data((8, 8, 8, 3), 7)
    taps, _ = P.kernel_taps(4, 1)
    want = P.make_data_matrix(d, taps, loraks, center or (1, 1, 1))
    fft = g64.gram(d, 4, 1, loraks=loraks, center=center)
    assert rel(fft, want) < 1e-12
    monkeypatch.setenv("NUFFT3D_GRAM64_DIRECT", "1")
    direct = g64.gram(d, 4, 1, loraks=loraks, center=center)
    assert rel(direct, want) < 1e-12
    assert rel(fft, direct) < 1e-12
