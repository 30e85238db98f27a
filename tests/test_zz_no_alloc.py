"""nufft3d_forward_noalloc (the fNUFT that allocates nothing) through the C
ABI, against nufft3d_forward on the same plan (GPU).

Tolerances (relative L2): FP32 plans 1e-6 (the accuracy gate is 1e-4; the
two entry points run the same kernels and differ only in coil grouping),
FP64 plans 1e-12. Two calls in a row, the second with fewer coils, so the
second reads a scratch array the first left full.
"""
import ctypes as C

import numpy as np
import pytest

from nufft_3d_amd import _lib, nufft_3d, synth

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _image(N, nc, seed, dtype):
    import torch
    r = np.random.default_rng(seed)
    x = (r.normal(size=(nc, N, N, N)) + 1j * r.normal(size=(nc, N, N, N))).astype(dtype)
    return torch.from_numpy(x).cuda()


def _noalloc(g, x):
    import torch
    nc = x.shape[0]
    k = torch.full((nc, g.M), float("nan"), dtype=x.dtype, device=x.device)
    lib = g._lib
    _lib.check(lib.nufft3d_set_stream(g._h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    _lib.check(lib.nufft3d_forward_noalloc(g._h, C.c_void_p(x.data_ptr()), nc,
                                           C.c_void_p(k.data_ptr())))
    torch.cuda.synchronize()
    return k.cpu().numpy()


@pytest.mark.parametrize("N,gpu,fp64c,ncs,tol", [
    (64, 1, False, (5, 2), 1e-6),     # 29 MiB per coil: groups of 2 coils
    (32, 2, False, (3, 1), 1e-12),
    (48, 1, True, (3, 1), 1e-6),      # fp64 arithmetic behind fp32 arrays: widen/round in scratch
])
def test_noalloc_forward_equals_forward_twice(N, gpu, fp64c, ncs, tol):
    import torch
    om = synth.koosh_ball(N, 300, 2 * N)
    om[:, 7] = [N, 0, 0]                  # an out-of-bounds row (zeroed)
    kw = dict(fp64_compute=True) if fp64c else {}
    g = nufft_3d(om, N, gpu=gpu, **kw)
    dtype = np.complex64 if gpu == 1 else np.complex128
    for i, nc in enumerate(ncs):
        print("noalloc N=%d gpu=%d fp64_compute=%s nc=%d" % (N, gpu, fp64c, nc), flush=True)
        x = _image(N, nc, 10 + i, dtype)
        got = _noalloc(g, x)
        want = g.fNUFT(x)
        torch.cuda.synchronize()
        want = want.cpu().numpy()
        assert np.all(np.isfinite(got))
        assert np.all(got[:, 7] == 0)
        e = _rel(got, want)
        print("  rel err vs nufft3d_forward %.3e" % e, flush=True)
        assert e <= tol


def test_noalloc_rejects_plans_beyond_the_scratch_and_host_arrays():
    import torch
    om = synth.koosh_ball(128, 20, 64)
    g = nufft_3d(om, 128, gpu=1)            # K = 256: 235 MiB per coil
    x = _image(128, 1, 3, np.complex64)
    k = torch.empty((1, g.M), dtype=torch.complex64, device="cuda")
    st = g._lib.nufft3d_forward_noalloc(g._h, C.c_void_p(x.data_ptr()), 1, C.c_void_p(k.data_ptr()))
    assert st == _lib.NUFFT3D_ERR_UNSUPPORTED
    assert b"MiB" in g._lib.nufft3d_last_error()
    small = nufft_3d(synth.koosh_ball(16, 20, 32), 16, gpu=1)
    xh = np.zeros((16, 16, 16), np.complex64)
    kh = np.zeros(small.M, np.complex64)
    st = small._lib.nufft3d_forward_noalloc(small._h, xh.ctypes.data, 1, kh.ctypes.data)
    assert st == _lib.NUFFT3D_ERR_INVALID
    assert b"device arrays" in small._lib.nufft3d_last_error()
