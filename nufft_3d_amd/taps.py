"""Kernel tap offsets of SAKE/PRUNO (``pruno.m:56-71``, ``sake.m:61-71``).

``ndgrid(-fix(w/2):fix(w/2))`` masked by the radius ``w/2`` (radial) or the
cube, listed in MATLAB column-major order (x fastest). The HIP library builds
the same list in C++; this copy sizes output arrays on the host.
"""
import math

import numpy as np


def kernel_taps(width=5, radial=1, is2d=False):
    h = int(width) // 2
    out = []
    zs = [0] if is2d else range(-h, h + 1)
    for z in zs:
        for y in range(-h, h + 1):
            for x in range(-h, h + 1):
                if radial:
                    keep = math.sqrt(x * x + y * y + z * z) <= width / 2
                else:
                    keep = abs(x) <= width / 2 and abs(y) <= width / 2 and abs(z) <= width / 2
                if keep:
                    out.append((x, y, z))
    return out


def null_kernels_image(Kt, taps, width, N, nc):
    """Image-space null kernels K [N1 N2 N3 nc nv] of pruno.m:124-143 from
    the tap-domain kernels Kt [n nv] (row c + nc*t): each tap's coil vector
    goes to its place (offset + fix(width/2)) in the width^3 (width^2 x 1 for
    2d) mask cube, then every axis gets the zero-padded inverse DFT times its
    length, sum_k K[k] exp(+2 pi i j k / N) (truncated to N when the cube is
    longer, as MATLAB's ifft(K, N) does)."""
    Kt = np.asarray(Kt)
    nv = Kt.shape[1]
    h = int(width) // 2
    w = 2 * h + 1
    is2d = N[2] == 1
    cube = np.zeros((w, w, 1 if is2d else w, nc, nv), dtype=np.complex128)
    Kr = Kt.reshape(len(taps), nc, nv)
    for t, (x, y, z) in enumerate(taps):
        cube[x + h, y + h, 0 if is2d else z + h] = Kr[t]
    K = cube
    for ax in range(3):
        n = int(N[ax])
        if K.shape[ax] > n:
            K = K.take(np.arange(n), axis=ax)
        L = K.shape[ax]
        jk = np.outer(np.arange(n), np.arange(L)) % n
        E = np.exp(2j * np.pi * jk / n)
        K = np.moveaxis(np.tensordot(E, np.moveaxis(K, ax, 0), axes=(1, 0)), 0, ax)
    return K
