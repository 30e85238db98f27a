"""Write the small golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no golden vectors (its only test displays images), so
these fixtures freeze the oracle's outputs on seeded inputs; the oracle
itself is pinned by closed forms and brute force in tests/test_oracle.py.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from nufft_3d_amd import synth  # noqa: E402
from oracle.nufft_oracle import NufftOracle  # noqa: E402

CASES = {
    "radial_J4_N8": dict(om=lambda: synth.koosh_ball(8, 24, 16), N=8, kw=dict(J=4, u=2, radial=1)),
    "separable_J4_N8": dict(om=lambda: synth.koosh_ball(8, 24, 16), N=8, kw=dict(J=4, u=2, radial=0)),
    "radial_J6_u125_N8": dict(om=lambda: synth.golden_means(8, 30, 16), N=8,
                              kw=dict(J=6, u=1.25, radial=1)),
    "oob_N10x8x12": dict(om=lambda: np.random.default_rng(12).uniform(-7, 7, size=(3, 400)),
                         N=(10, 8, 12), kw=dict(J=4, u=2, radial=1)),
}


def make(name):
    c = CASES[name]
    om = c["om"]()
    o = NufftOracle(om, c["N"], **c["kw"])
    r = np.random.default_rng(100)
    x = r.normal(size=o.N + (2,)) + 1j * r.normal(size=o.N + (2,))
    kin = r.normal(size=(o.M, 2)) + 1j * r.normal(size=(o.M, 2))
    raw = o.fNUFT(x) + 0.01 * (r.normal(size=(o.M, 2)) + 1j * r.normal(size=(o.M, 2)))
    im5, info = o.iNUFT(raw, 5, 0.02, return_info=True)
    return dict(om=om, N=np.array(o.N), K=np.array(o.K), J=o.J, u=o.u, alpha=o.alpha,
                radial=int(o.radial), nnz=o.nnz, U=o.U, d=o.d, x=x, k_fwd=o.fNUFT(x), k_in=kin,
                x_adj=o.aNUFT(kin), raw=raw, im_regrid=o.iNUFT(raw, 0), im_minres5=im5,
                minres5_iters=np.array(info["iters"]), damp=0.02)


if __name__ == "__main__":
    for n in CASES:
        np.savez_compressed(os.path.join(HERE, n + ".npz"), **make(n))
        print("wrote", n)
