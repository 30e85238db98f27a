"""Full-size parity on the five benchmark configurations (GPU).

Tolerances (relative L2): fp32-array plans 1e-4 (the accuracy gate), FP64
plans 1e-12 against the FP64 oracle (1e-10 for 20 MINRES iterations; 2e-11
for the 256^3 adjoint of random data, whose conditioning is stated there).

* 64^3 plain (2,000 x 128 koosh spokes, 1 coil, J=4 u=2): fNUFT, aNUFT,
  density and a 20-iteration iNUFT against the oracle (fNUFT.m:2-9,
  aNUFT.m:2-9, density.m, iNUFT.m:127-136).
* 128^3 batched coils (12,000 x 256 spokes, 16 coils): all 16 coils fp32
  against FP64; coils 0-1 against the oracle.
* 256^3 precision (40,000 x 512 golden-means spokes, 8 coils, J=6 u=1.25,
  2.32G hits: the >2^31-hit plan path): FP64 adjointness, fp32 policy
  against FP64 (fNUFT, aNUFT, 30-iteration iNUFT), bitwise repeatability,
  and 51,200 rows of whole spokes recomputed by the oracle on the same
  padded grid (fNUFT rows; the adjoint of data kept on those rows), which
  checks the plan's hit lists independently of the device's plan path.
* 96^3 low rank (Cartesian R~4 + 24^3 centre, 8 coils, 5^3 cube kernel,
  n = 1000): the Hankel Gram (split-BF16 MFMA / FP64) against the oracle's,
  PRUNO's noise estimate, singular values and nv.
* serving at the real size classes (64/128/192^3, 1-32 coils): each result
  the same alone and in the mix.
"""
import math

import numpy as np
import pytest

from nufft_3d_amd import nufft_3d, serve, synth
from oracle import nufft_oracle as O
from oracle import pruno_oracle as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a).ravel().astype(np.complex128)
    b = np.asarray(b).ravel().astype(np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def trel(a, b):
    import torch
    a = a.to(torch.complex128)
    b = b.to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def _noisy(k, seed):
    r = np.random.default_rng(seed)
    s = 0.01 * np.abs(k).max()
    return k + s * (r.normal(size=k.shape) + 1j * r.normal(size=k.shape))


def _coil_images(N, nc, seed=2):
    return synth.phantom_complex(N)[..., None] * synth.coil_maps(N, nc, seed=seed)


def _torch_img(x, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.transpose(x, (3, 2, 1, 0))).astype(dtype)).cuda()


def _to_fortran(t):
    """torch (nc, N3, N2, N1) -> numpy [N1 N2 N3 nc]."""
    return np.transpose(t.cpu().numpy(), (3, 2, 1, 0))


def test_config1_plain_64_full_size_against_oracle():
    N = 64
    om = synth.koosh_ball(N, 2000, 128)
    print("config 1: M=%d" % om.shape[1], flush=True)
    orc = O.NufftOracle(om, N)                # density: 20 Pipe iterations
    g32 = nufft_3d(om, N, gpu=1)
    g64 = nufft_3d(om, N, gpu=2)
    img = synth.phantom_complex(N)
    k_ref = orc.fNUFT(img)
    assert rel(g64.fNUFT(img), k_ref) < 1e-12
    assert rel(g32.fNUFT(img), k_ref) < 1e-4
    raw = _noisy(k_ref, 1)
    a_ref = orc.aNUFT(raw)
    assert rel(g64.aNUFT(raw), a_ref) < 1e-12
    assert rel(g32.aNUFT(raw), a_ref) < 1e-4
    assert rel(g64.d, orc.d) < 1e-12
    assert rel(g32.d, orc.d) < 1e-4
    x_ref, info = orc.iNUFT(raw, 20, 0.0, return_info=True)
    x64, i64 = g64.iNUFT(raw, 20, 0.0, return_info=True)
    x32 = g32.iNUFT(raw, 20, 0.0)
    e64, e32 = rel(x64, x_ref), rel(x32, x_ref)
    print("config 1 iNUFT(20) rel err fp64 %.2e fp32 %.2e" % (e64, e32), flush=True)
    assert i64["iters"][0] == info["iters"][0]
    assert e64 < 1e-10 and e32 < 1e-4


def test_config2_batched_coils_128_full_size():
    import torch
    N, nc = 128, 16
    om = synth.koosh_ball(N, 12000, 256)
    M = om.shape[1]
    print("config 2: M=%d" % M, flush=True)
    g32 = nufft_3d(om, N, gpu=1)
    g64 = nufft_3d(om, N, gpu=2)
    ximg = _coil_images(N, nc)
    x32 = _torch_img(ximg, np.complex64)
    k32 = g32.fNUFT(x32)
    k64 = g64.fNUFT(x32.to(torch.complex128))
    assert trel(k32, k64) < 1e-4
    data = k64 + 0.01 * float(k64.abs().max()) * torch.randn(
        k64.shape, dtype=torch.complex128, device=k64.device, generator=torch.Generator("cuda").manual_seed(3))
    a32 = g32.aNUFT(data.to(torch.complex64))
    a64 = g64.aNUFT(data)
    assert trel(a32, a64) < 1e-4
    torch.cuda.synchronize()
    # coils 0-1 against the oracle (the density does not enter fNUFT/aNUFT)
    orc = O.NufftOracle(om, N, d=np.ones(M))
    k_ref = orc.fNUFT(ximg[..., :2].astype(np.complex64))    # the image the GPU was given
    kk64 = k64[:2].cpu().numpy().T
    assert rel(kk64, k_ref) < 1e-12
    assert rel(k32[:2].cpu().numpy().T, k_ref) < 1e-4
    d2 = data[:2].cpu().numpy().T
    a_ref = orc.aNUFT(d2)
    assert rel(_to_fortran(a64[:2]), a_ref) < 1e-12
    assert rel(_to_fortran(a32[:2]), a_ref) < 1e-4


def test_config3_precision_256_full_size():
    import torch
    N, nc, J, u = 256, 8, 6, 1.25
    om = synth.golden_means(N, 40000, 512)
    M = om.shape[1]
    print("config 3: M=%d" % M, flush=True)
    g32 = nufft_3d(om, N, J=J, u=u, gpu=1)
    g64 = nufft_3d(om, N, J=J, u=u, gpu=2)
    print("config 3 plans built, compute precision %d" % g32.compute_precision, flush=True)
    ximg = _coil_images(N, nc)
    x32 = _torch_img(ximg, np.complex64)
    del ximg
    x64 = x32.to(torch.complex128)
    k32 = g32.fNUFT(x32)
    k32b = g32.fNUFT(x32)
    k64 = g64.fNUFT(x64)
    torch.cuda.synchronize()
    assert torch.equal(k32, k32b)
    del k32b
    e = trel(k32, k64)
    print("config 3 fNUFT fp32 vs fp64 %.2e" % e, flush=True)
    assert e < 1e-4
    data = k64 + 0.01 * float(k64.abs().max()) * torch.randn(
        k64.shape, dtype=torch.complex128, device=k64.device, generator=torch.Generator("cuda").manual_seed(5))
    d32 = data.to(torch.complex64)
    a32 = g32.aNUFT(d32)
    a32b = g32.aNUFT(d32)
    a64 = g64.aNUFT(data)
    torch.cuda.synchronize()
    assert torch.equal(a32, a32b)
    del a32b
    e = trel(a32, a64)
    print("config 3 aNUFT fp32 vs fp64 %.2e" % e, flush=True)
    assert e < 1e-4
    # adjointness of the FP64 operator pair: <A x, y> = <x, A' y>
    lhs = complex(torch.vdot(k64.flatten(), data.flatten()))
    rhs = complex(torch.vdot(x64.flatten(), a64.flatten()))
    print("config 3 adjointness %.2e" % (abs(lhs - rhs) / abs(lhs)), flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # 30 MINRES iterations, fp32 policy against FP64
    i32 = g32.iNUFT(d32, 30, 0.0)
    i64 = g64.iNUFT(data, 30, 0.0)
    e = trel(i32, i64)
    print("config 3 iNUFT(30) fp32 vs fp64 %.2e" % e, flush=True)
    assert e < 1e-4
    del i32, i64
    # whole spokes (every 400th: 100 x 512 rows) recomputed by the oracle
    rows = (np.arange(0, 40000, 400)[:, None] * 512 + np.arange(512)[None, :]).ravel()
    orc = O.NufftOracle(om[:, rows], N, J=J, u=u, d=np.ones(len(rows)))
    sc = math.sqrt(orc.nnz / g64.nnz)
    x0 = _to_fortran(x64[:1])[..., 0]
    k_ref = orc.fNUFT(x0)[:, 0] * sc
    e64 = rel(k64[0].cpu().numpy()[rows], k_ref)
    e32 = rel(k32[0].cpu().numpy()[rows], k_ref)
    print("config 3 oracle rows fNUFT fp64 %.2e fp32 %.2e" % (e64, e32), flush=True)
    assert e64 < 1e-12 and e32 < 1e-4
    mask = torch.zeros(M, dtype=torch.bool, device=data.device)
    mask[torch.from_numpy(rows).to(data.device)] = True
    dsub = data[:1] * mask[None, :]
    a_ref = orc.aNUFT(dsub[0].cpu().numpy()[rows][:, None])[..., 0] * sc
    a64s = _to_fortran(g64.aNUFT(dsub))[..., 0]
    a32s = _to_fortran(g32.aNUFT(dsub.to(torch.complex64)))[..., 0]
    e64, e32 = rel(a64s, a_ref), rel(a32s, a_ref)
    # The deapodization U spans max|U|/min|U| ~ 2.3e7 here, and with random
    # data on a subset of rows the field-of-view corners, where U is largest,
    # carry most of the adjoint's L2 norm: the rounding of the inverse FFT is
    # amplified by up to that range. The oracle itself (NumPy FFT) is within
    # 8.3e-13 of the same sum evaluated in long double; the device's fp64
    # transform measures ~7e-12. So this check holds the fp64 path to 2e-11,
    # not 1e-12 (every other fp64 comparison in the suite is at 1e-12).
    Ug = g64.U
    eU = float(np.max(np.abs(Ug - orc.U) / np.abs(orc.U)))
    print("config 3 oracle rows aNUFT fp64 %.2e (U max rel diff %.2e, range %.2e) fp32 %.2e"
          % (e64, eU, np.abs(orc.U).max() / np.abs(orc.U).min(), e32), flush=True)
    assert e64 < 2e-11 and e32 < 1e-4
    g32.close()
    g64.close()
    torch.cuda.empty_cache()


def test_config3_block_staged_full_size():
    # the block-staged kernels at the 2.32G-hit size (the autotuner limited
    # to them, as the bench's autotune picks them here; the gather kernels'
    # lists are then released): fp32 policy against the gather results and
    # FP64, FP64-plan adjointness, and whole spokes against the oracle
    import os
    import torch
    N, nc, J, u = 256, 8, 6, 1.25
    om = synth.golden_means(N, 40000, 512)
    M = om.shape[1]
    print("config 3 block-staged: M=%d" % M, flush=True)
    g32 = nufft_3d(om, N, J=J, u=u, gpu=1)
    ximg = _coil_images(N, nc)
    x32 = _torch_img(ximg, np.complex64)
    del ximg
    x64 = x32.to(torch.complex128)
    kg = g32.fNUFT(x32)                          # gather kernels (no autotune yet)
    dat = (kg + 0.01 * float(kg.abs().max()) * torch.randn(
        kg.shape, dtype=torch.complex64, device=kg.device,
        generator=torch.Generator("cuda").manual_seed(6))).contiguous()
    ag = g32.aNUFT(dat)
    os.environ["NUFFT3D_TUNE_BLOCK_ONLY"] = "1"
    try:
        rep = g32.autotune(nc, reps=1)
        g64 = nufft_3d(om, N, J=J, u=u, gpu=2)
        rep64 = g64.autotune(nc, reps=1)
    finally:
        os.environ.pop("NUFFT3D_TUNE_BLOCK_ONLY", None)
    print("config 3 block-staged autotune:", rep["candidates"][rep["best"]]["spread"], rep["forward"],
          "plan GB %.1f" % (g32.device_bytes()[0] / 1e9), flush=True)
    assert rep["candidates"][rep["best"]]["spread"] == "block-staged"
    assert rep["forward"].startswith("block-staged") and rep64["forward"].startswith("block-staged")
    kb = g32.fNUFT(x32)
    ab = g32.aNUFT(dat)
    torch.cuda.synchronize()
    e_k, e_a = trel(kb, kg), trel(ab, ag)
    print("config 3 block vs gather (fp32 arrays): fNUFT %.2e aNUFT %.2e" % (e_k, e_a), flush=True)
    assert e_k < 1e-8 and e_a < 1e-8
    assert torch.equal(g32.aNUFT(dat), ab)
    k64 = g64.fNUFT(x64)
    d64 = dat.to(torch.complex128)
    a64 = g64.aNUFT(d64)
    e_k, e_a = trel(kb, k64), trel(ab, a64)
    print("config 3 block fp32 policy vs fp64: fNUFT %.2e aNUFT %.2e" % (e_k, e_a), flush=True)
    assert e_k < 1e-4 and e_a < 1e-4
    lhs = complex(torch.vdot(k64.flatten(), d64.flatten()))
    rhs = complex(torch.vdot(x64.flatten(), a64.flatten()))
    print("config 3 block-staged FP64 adjointness %.2e" % (abs(lhs - rhs) / abs(lhs)), flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    i32 = g32.iNUFT(dat, 10, 0.0)
    i64 = g64.iNUFT(d64, 10, 0.0)
    e = trel(i32, i64)
    print("config 3 block-staged iNUFT(10) fp32 vs fp64 %.2e" % e, flush=True)
    assert e < 1e-4
    del i32, i64
    rows = (np.arange(0, 40000, 400)[:, None] * 512 + np.arange(512)[None, :]).ravel()
    orc = O.NufftOracle(om[:, rows], N, J=J, u=u, d=np.ones(len(rows)))
    sc = math.sqrt(orc.nnz / g64.nnz)
    x0 = _to_fortran(x64[:1])[..., 0]
    k_ref = orc.fNUFT(x0)[:, 0] * sc
    e64 = rel(k64[0].cpu().numpy()[rows], k_ref)
    e32 = rel(kb[0].cpu().numpy()[rows], k_ref)
    print("config 3 block-staged oracle rows fNUFT fp64 %.2e fp32 %.2e" % (e64, e32), flush=True)
    assert e64 < 1e-12 and e32 < 1e-4
    g32.close()
    g64.close()
    torch.cuda.empty_cache()


def test_config4_lowrank_96_full_size_gram_and_pruno():
    N, nc, width, radial = 96, 8, 5, 0
    om = synth.cartesian_jittered(N, centre=24, seed=4)
    print("config 4: M=%d" % om.shape[1], flush=True)
    data = np.fft.fftn(_coil_images(N, nc, seed=3), axes=(0, 1, 2))
    taps, _ = P.kernel_taps(width, radial)
    want = P.gram_fft(data, taps)
    g32 = nufft_3d(om, N, gpu=1)
    g64 = nufft_3d(om, N, gpu=2)
    e64 = rel(g64.gram(data, width, radial), want)
    want32 = P.gram_fft(data.astype(np.complex64).astype(np.complex128), taps)
    G32 = g32.gram(data.astype(np.complex64), width, radial)
    e32 = rel(G32, want32)
    print("config 4 gram n=%d rel err fp64 %.2e fp32 (split-BF16 MFMA) %.2e"
          % (want.shape[0], e64, e32), flush=True)
    assert want.shape[0] == 1000
    assert e64 < 1e-12 and e32 < 1e-4
    w = np.linalg.eigvalsh(want)[::-1]
    noise, sv, f, nv = P.select(w, 1.0, N, N)
    _, got = g64.pruno(data, width=width, radial=radial, output="taps", return_info=True)
    print("config 4 pruno noise %.4e (oracle %.4e) nv %d (oracle %d)"
          % (got["noise"], noise, got["nv"], nv), flush=True)
    # eigenvalues agree to the eigen-solver's backward error (n eps lambda_max,
    # which is also where two fp64 Grams 4e-16 apart differ), so a singular
    # value s is good to ~1e-12 (s_max/s)^2 relative: the 100 largest to
    # 1e-13, every one above 1e-3 of the largest to 1e-9
    lam_d, lam_o = got["sv"] ** 2, sv ** 2
    assert np.max(np.abs(lam_d - lam_o)) <= 1e-12 * lam_o[0]
    assert np.max(np.abs(got["sv"][:100] - sv[:100]) / sv[:100]) < 1e-13
    big = sv > 1e-3 * sv[0]
    assert np.max(np.abs(got["sv"][big] - sv[big]) / sv[big]) < 1e-9
    # the noise floor is a histogram/median heuristic, discontinuous in the
    # eigenvalues (a value crossing a bin edge moves the median by one
    # element): it is pinned on the device's own eigenvalues, and against
    # the oracle's within that step
    assert got["noise"] == pytest.approx(P.estimate_noise(got["sv"] ** 2, 1.0, N, N), rel=1e-12)
    assert got["noise"] == pytest.approx(noise, rel=1e-4)
    assert got["nv"] == P.select(got["sv"] ** 2, 1.0, N, N, got["noise"])[3]
    assert abs(got["nv"] - nv) <= 1
    # fp32 plan (split-BF16 Gram): by Weyl's inequality each eigenvalue moves
    # by at most the Gram's error in the 2-norm (<= its Frobenius norm)
    _, got32 = g32.pruno(data.astype(np.complex64), width=width, radial=radial, noise=noise,
                         output="taps", return_info=True)
    lam32 = np.linalg.eigvalsh(want32)[::-1]
    bound = np.linalg.norm(G32 - want32) + 1e-12 * lam32[0]
    assert np.max(np.abs(got32["sv"] ** 2 - lam32)) <= bound
    assert np.max(np.abs(got32["sv"][:100] - np.sqrt(lam32[:100])) / np.sqrt(lam32[:100])) < 1e-4


def test_config5_serving_size_classes_alone_equals_mixed():
    import torch
    sizes = (64, 128, 192)
    reqs = serve.make_stream(10, 1e4, seed=8, sizes=sizes, coils=(1, 8, 16, 32),
                             samples=(100_000, 600_000), traj_per_size=1, maxit=20)
    print("config 5: %d requests %s" % (len(reqs), [(r.N, r.nc, r.M, r.kind) for r in reqs]),
          flush=True)
    be = serve.GpuBackend(sizes, max_coils=32)
    eng = serve.Engine(be, slots=3, max_batch_coils=64, keep_results=True)
    eng.serve(reqs)
    mixed = {r.rid: r.result.clone() for r in reqs}
    eng.close()
    serve.reset(reqs)
    serve.serve_one_at_a_time(be, reqs, keep_results=True)
    torch.cuda.synchronize()
    worst = max(trel(mixed[r.rid], r.result) for r in reqs)
    print("config 5 worst relative L2 difference alone vs mixed: %.3e" % worst, flush=True)
    assert worst <= 1e-6
