"""Block-staged gridding kernels (GPU): the split-precision adjoint and the
fp64-arithmetic forward.

The aNUFT of a plan with fp32 arrays that computes in fp64 (u=1.25, J=6)
sums float weight x float sample products in double into an fp64 grid
(spmv_t, nufft_3d.m:274-282). With tune flag 4096 the grid is walked in
4x4x4 blocks whose samples pass through LDS in round-robin chunks; a point
adds the same exact products as the gather kernel in a different fixed
order. Each image value is that double sum, transformed in fp64 and
rounded once to float, so the two images differ only where the reordering
moves a value across a float rounding boundary (a few elements by one
float ulp: 1e-8 relative L2 bounds it, against the 6e-8 resolution of a
float); the block kernel is bitwise repeatable. Against the FP64 oracle the
fp32 gate per operator (5e-5) applies.

The forward of an fp64-arithmetic plan (spmv, nufft_3d.m:259-271) with tune
flag 8192 runs over groups of 64 sorted samples whose grid rows pass through
LDS in round-robin chunks; each hit's weight is evaluated as the gather
kernel evaluates it, and the sum over a sample's hits runs in another fixed
order: fp64 plans agree with the gather and with the FP64 oracle within the
FP64 gate (1e-12); fp32-array plans as above (1e-8 against the
gather, 5e-5 against the oracle).
"""
import ctypes as C
import json

import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, synth
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu

TOL32_OP = 5e-5


def rel(a, b):
    a = np.asarray(a).ravel().astype(np.complex128)
    b = np.asarray(b).ravel().astype(np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rand(shape, seed):
    r = np.random.default_rng(seed)
    return r.normal(size=shape) + 1j * r.normal(size=shape)


def _plan_with_env(env, *a, **kw):
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return nufft_3d(*a, **kw)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _kernels_run(p, fn):
    p._lib.nufft3d_profile_enable(p._h, 1)
    fn()
    buf = C.create_string_buffer(1 << 16)
    p._lib.nufft3d_profile_read(p._h, buf, len(buf))
    p._lib.nufft3d_profile_enable(p._h, 0)
    return set(json.loads(buf.value.decode()))


# the block-staged adjoint kernels: 4x4x4 blocks (one wavefront) or wide
# 8x8x4 blocks (four wavefronts sharing each staged chunk)
ADJ_BLK = {"k_spread_blk", "k_spread_blkw"}


def _adj_kernel(wide):
    return "k_spread_blkw" if wide else "k_spread_blk"


CASES = [
    # N, trajectory, J, radial, coils
    (24, "koosh", 1, 8),
    (24, "koosh", 0, 8),
    (24, "golden", 1, 16),
    (26, "golden", 1, 8),     # K = 34: blocks cut by the grid edge
    (16, "dense", 1, 8),      # 2,000 spokes on a 20^3 grid: blocks in many pieces
    (24, "koosh", 1, 3),      # a partial coil group (coils loaded one by one, the rest zero)
    (24, "golden", 1, 11),    # an 8-coil group and a partial one
]


def _traj(N, kind):
    if kind == "koosh":
        return synth.koosh_ball(N, 300, 2 * N)
    if kind == "golden":
        return synth.golden_means(N, 300, 2 * N)
    return synth.koosh_ball(N, 2000, 2 * N)


@pytest.mark.parametrize("wide", [0, 1])
@pytest.mark.parametrize("N,kind,radial,nc", CASES)
def test_block_staged_adjoint_matches_gather_and_oracle(N, kind, radial, nc, wide):
    om = _traj(N, kind)
    kw = dict(J=6, u=1.25, radial=radial, gpu=1)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096), "NUFFT3D_BLK_WIDE": str(wide)}, om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    assert blk.compute_precision == 2
    k = _rand((om.shape[1], nc), 71).astype(np.complex64)
    got = {}
    ran = _kernels_run(blk, lambda: got.setdefault("b", blk.aNUFT(k)))
    assert ADJ_BLK & ran == {_adj_kernel(wide)} and "k_spread" not in ran, ran
    if kind == "dense":
        assert "k_spread_blk_reduce" in ran, ran
    want = gat.aNUFT(k)
    assert rel(got["b"], want) < 1e-8
    assert np.array_equal(blk.aNUFT(k), got["b"])       # deterministic
    orc = O.NufftOracle(om, N, J=6, u=1.25, radial=radial)
    assert rel(got["b"], orc.aNUFT(k.astype(np.complex128))) < TOL32_OP


@pytest.mark.parametrize("piece", [16, 64, 4096])
def test_block_staged_adjoint_piece_sizes(piece):
    # pieces of 16 / 64 / 4096 entry rows split the dense trajectory's blocks
    # into many, some or few partials; each split is its own fixed order of
    # the same products (deterministic), all within the float rounding of
    # the gather's image
    N = 16
    om = _traj(N, "dense")
    kw = dict(J=6, u=1.25, gpu=1)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096), "NUFFT3D_BLK_PIECE": str(piece)},
                         om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    k = _rand((om.shape[1], 8), 79).astype(np.complex64)
    got = {}
    ran = _kernels_run(blk, lambda: got.setdefault("b", blk.aNUFT(k)))
    assert "k_spread_blkw" in ran, ran
    if piece < 4096:
        assert "k_spread_blk_reduce" in ran, ran
    assert rel(got["b"], gat.aNUFT(k)) < 1e-8
    assert np.array_equal(blk.aNUFT(k), got["b"])


@pytest.mark.parametrize("wide", [0, 1])
@pytest.mark.parametrize("kind,nc", [("golden", 8), ("dense", 16)])
def test_block_staged_adjoint_of_an_fp64_plan(kind, nc, wide):
    # double samples into the double grid (fp64 plans, and iNUFT's internal
    # samples on any fp64-arithmetic plan): the same products summed in
    # another fixed order. The deapodization (its range is ~1e7 at u=1.25,
    # J=6) multiplies double rounding of the grid at the field-of-view edge,
    # so both the gather and the FP64 oracle are held to the FP64 gate, 1e-12
    N = 16 if kind == "dense" else 24
    om = _traj(N, kind)
    kw = dict(J=6, u=1.25, gpu=2)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096), "NUFFT3D_BLK_WIDE": str(wide)}, om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    k = _rand((om.shape[1], nc), 78)
    got = {}
    ran = _kernels_run(blk, lambda: got.setdefault("b", blk.aNUFT(k)))
    assert ADJ_BLK & ran == {_adj_kernel(wide)}, ran
    assert rel(got["b"], gat.aNUFT(k)) < 1e-12
    assert np.array_equal(blk.aNUFT(k), got["b"])
    orc = O.NufftOracle(om, N, J=6, u=1.25)
    assert rel(got["b"], orc.aNUFT(k)) < 1e-12


def test_block_staged_adjoint_on_a_grid_narrower_than_the_footprint():
    # K3 = 4 < 7: a sample reaches some points through two wrapped
    # candidates; each is its own entry of the same staged row
    rng = np.random.default_rng(5)
    N = (16, 16, 2)
    om = rng.uniform(-8, 7.5, size=(3, 1500))
    om[2] = rng.uniform(-1, 0.5, size=1500)
    kw = dict(J=6, u=1.25, radial=1, gpu=1)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096)}, om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    assert blk.compute_precision == 2 and min(blk.K) < 7
    k = _rand((om.shape[1], 8), 72).astype(np.complex64)
    ran = _kernels_run(blk, lambda: blk.aNUFT(k))
    assert ADJ_BLK & ran
    assert rel(blk.aNUFT(k), gat.aNUFT(k)) < 1e-8
    orc = O.NufftOracle(om, N, J=6, u=1.25, radial=1)
    assert rel(blk.aNUFT(k), orc.aNUFT(k.astype(np.complex128))) < TOL32_OP


def test_autotuner_times_the_block_staged_adjoint():
    om = synth.golden_means(24, 300, 48)
    g = nufft_3d(om, 24, J=6, u=1.25, gpu=1)
    rep = g.autotune(nc=8, reps=1)
    kinds = {c["spread"] for c in rep["candidates"]}
    assert "block-staged" in kinds, kinds
    assert rep["forward"] in ("block-staged", "block-staged-stored", "gather")
    assert len(rep["forward_ms"]) == 3
    k = _rand((om.shape[1], 8), 73).astype(np.complex64)
    orc = O.NufftOracle(om, 24, J=6, u=1.25)
    assert rel(g.aNUFT(k), orc.aNUFT(k.astype(np.complex128))) < TOL32_OP


# the block-staged forward kernels: groups of 64 samples (one wavefront) or of
# 256 (a work-group sharing each staged chunk)
FWD_BLK = {"k_interp_blk", "k_interp_blkw"}

FWD_CASES = [
    # N, trajectory, gpu, coils
    (24, "golden", 1, 8),
    (24, "koosh", 1, 16),
    (26, "golden", 1, 8),
    (16, "dense", 1, 8),
    (24, "golden", 2, 8),     # fp64 plan: kernel.m's literal arithmetic, double outputs
    (24, "koosh", 1, 3),      # a partial coil group
    (24, "koosh", 2, 5),      # a partial coil group of an fp64 plan
]


@pytest.mark.parametrize("group", [64, 256])
@pytest.mark.parametrize("stored", [False, True])
@pytest.mark.parametrize("N,kind,gpu,nc", FWD_CASES)
def test_block_staged_forward_matches_gather_and_oracle(N, kind, gpu, nc, stored, group):
    # stored: the lists hold each hit's fp32 weight, evaluated at plan time
    # by the same function the kernels evaluate; group: samples per group
    om = _traj(N, kind)
    kw = dict(J=6, u=1.25, gpu=gpu)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 8192 | (16384 if stored else 0)),
                          "NUFFT3D_FWD_GROUP": str(group)}, om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    assert blk.compute_precision == 2
    x = _rand((N, N, N, nc), 74)
    if gpu == 1:
        x = x.astype(np.complex64)
    got = {}
    ran = _kernels_run(blk, lambda: got.setdefault("b", blk.fNUFT(x)))
    assert FWD_BLK & ran == {"k_interp_blkw" if group == 256 else "k_interp_blk"}, ran
    assert "k_interp" not in ran, ran
    want = gat.fNUFT(x)
    assert rel(got["b"], want) < (1e-12 if gpu == 2 else 1e-8)
    assert np.array_equal(blk.fNUFT(x), got["b"])       # deterministic
    orc = O.NufftOracle(om, N, J=6, u=1.25)
    assert rel(got["b"], orc.fNUFT(x.astype(np.complex128))) < (1e-12 if gpu == 2 else TOL32_OP)


@pytest.mark.parametrize("gpu", [1, 2])
def test_block_staged_forward_and_adjoint_in_the_inverse(gpu):
    # iNUFT keeps its samples internal (fp64 rows, data weights): the forward
    # writes the sorted layout with the density weights applied
    N = 20
    om = synth.golden_means(N, 250, 2 * N)
    kw = dict(J=6, u=1.25, gpu=gpu)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096 | 8192 | 16384)}, om, N, **kw)
    gat = _plan_with_env({"NUFFT3D_TUNE": "2"}, om, N, **kw)
    orc = O.NufftOracle(om, N, J=6, u=1.25)
    x = _rand((N, N, N, 8), 75)
    raw = orc.fNUFT(x) + 0.01 * _rand((om.shape[1], 8), 76)
    if gpu == 1:
        raw = raw.astype(np.complex64)
    got = {}
    ran = _kernels_run(blk, lambda: got.setdefault("b", blk.iNUFT(raw, 6, 0.02)))
    assert FWD_BLK & ran and ADJ_BLK & ran, ran
    want = gat.iNUFT(raw, 6, 0.02)
    assert rel(got["b"], want) < (1e-11 if gpu == 2 else 1e-7)
    ref = orc.iNUFT(raw.astype(np.complex128), 6, 0.02)
    assert rel(got["b"], ref) < (1e-10 if gpu == 2 else 1e-4)


def test_block_staged_forward_not_built_on_a_narrow_grid():
    # K3 = 4 < 2*ceil(J/2) + 2: a grid point can be reached by two candidates
    # of one sample, so the forward keeps the gather kernel
    rng = np.random.default_rng(6)
    N = (16, 16, 2)
    om = rng.uniform(-8, 7.5, size=(3, 1500))
    om[2] = rng.uniform(-1, 0.5, size=1500)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 8192)}, om, N, J=6, u=1.25, gpu=1)
    x = _rand(N + (8,), 77).astype(np.complex64)
    ran = _kernels_run(blk, lambda: blk.fNUFT(x))
    assert not FWD_BLK & ran and "k_interp" in ran, ran
    orc = O.NufftOracle(om, N, J=6, u=1.25)
    assert rel(blk.fNUFT(x), orc.fNUFT(x.astype(np.complex128))) < TOL32_OP


def test_gather_lists_released_after_the_autotuner_picks_block_kernels():
    # with both block-staged kernels chosen, the plan frees the gather
    # kernels' per-hit lists (fp32 weight table, codes); every later launch
    # (any coil count: 8-coil groups and a partial group, data weights,
    # float or double samples) runs the block-staged kernels
    import os
    N = 24
    om = synth.golden_means(N, 600, 2 * N)
    os.environ["NUFFT3D_TUNE_BLOCK_ONLY"] = "1"
    try:
        g = nufft_3d(om, N, J=6, u=1.25, gpu=1)
        rep = g.autotune(nc=8, reps=1)
    finally:
        os.environ.pop("NUFFT3D_TUNE_BLOCK_ONLY", None)
    assert rep["candidates"][rep["best"]]["spread"] == "block-staged", rep
    assert rep["forward"].startswith("block-staged"), rep
    # the same lists built at creation keep the gather lists beside them
    kept = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096 | 8192 | 16384)}, om, N, J=6, u=1.25, gpu=1)
    assert g.device_bytes()[0] < kept.device_bytes()[0], (g.device_bytes(), kept.device_bytes())
    orc = O.NufftOracle(om, N, J=6, u=1.25)
    for nc in (3, 8, 11):
        x = _rand((N, N, N, nc), 80 + nc).astype(np.complex64)
        k = _rand((om.shape[1], nc), 90 + nc).astype(np.complex64)
        ran = _kernels_run(g, lambda: (g.fNUFT(x), g.aNUFT(k)))
        assert FWD_BLK & ran and ADJ_BLK & ran, ran
        assert "k_interp" not in ran and "k_spread" not in ran, ran
        assert rel(g.fNUFT(x), orc.fNUFT(x.astype(np.complex128))) < TOL32_OP
        assert rel(g.aNUFT(k), orc.aNUFT(k.astype(np.complex128))) < TOL32_OP
    raw = orc.fNUFT(_rand((N, N, N, 3), 99)).astype(np.complex64)
    assert rel(g.iNUFT(raw, 5, 0.05), orc.iNUFT(raw.astype(np.complex128), 5, 0.05)) < 1e-4


@pytest.mark.parametrize("N,J,u,radial,gpu", [((24, 20, 16), 6, 1.25, 1, 1),
                                              ((20, 20, 20), 6, 1.25, 0, 1),
                                              ((20, 16, 24), 5, 1.25, 0, 2),
                                              ((16, 16, 16), 7, 1.5, 1, 2)])
def test_block_staged_kernels_on_other_shapes_and_kernels(N, J, u, radial, gpu):
    # non-cubic grids, the separable kernel, odd J and a wider footprint (J=7:
    # 9 candidates per axis), both kernels, fp32 arrays (u = 1.25: fp64
    # arithmetic behind them) and fp64 plans (any u)
    rng = np.random.default_rng(J + 3 * radial)
    om = np.stack([rng.uniform(-n / 2, n / 2 - 0.5, size=4000) for n in N])
    kw = dict(J=J, u=u, radial=radial, gpu=gpu)
    blk = _plan_with_env({"NUFFT3D_TUNE": str(2 | 4096 | 8192 | 16384)}, om, N, **kw)
    assert blk.compute_precision == 2
    x = _rand(tuple(N) + (8,), 101)
    k = _rand((om.shape[1], 8), 102)
    if gpu == 1:
        x, k = x.astype(np.complex64), k.astype(np.complex64)
    got = {}
    ran = _kernels_run(blk, lambda: (got.setdefault("f", blk.fNUFT(x)), got.setdefault("a", blk.aNUFT(k))))
    assert FWD_BLK & ran and ADJ_BLK & ran, ran
    orc = O.NufftOracle(om, N, J=J, u=u, radial=radial)
    tol = 1e-12 if gpu == 2 else TOL32_OP
    assert rel(got["f"], orc.fNUFT(x.astype(np.complex128))) < tol
    assert rel(got["a"], orc.aNUFT(k.astype(np.complex128))) < tol
