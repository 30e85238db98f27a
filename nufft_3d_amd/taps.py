"""Kernel tap offsets of SAKE/PRUNO (``pruno.m:56-71``, ``sake.m:61-71``).

``ndgrid(-fix(w/2):fix(w/2))`` masked by the radius ``w/2`` (radial) or the
cube, listed in MATLAB column-major order (x fastest). The HIP library builds
the same list in C++; this copy sizes output arrays on the host.
"""
import math


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
