"""Seeded synthetic inputs: phantom, coil maps and 3-D radial trajectories.

These generate the workloads the tests and the benchmark run on; they mirror
the reference demo's inputs:

* ``phantom3d``   -> ``phantom3d.m:72-128`` with the modified Shepp-Logan
  table of ``phantom3d.m:199-218`` (MATLAB ``meshgrid`` index convention:
  p[i1,i2,i3] sits at y=r[i1], x=r[i2], z=r[i3]).
* ``koosh_ball``  -> ``test_nufft3d.m:13-45``: full-diameter spokes whose
  directions lie on a Fibonacci sphere (``cos(polar)=1-2(k-1)/(n-1)``,
  ``azimuth=(k-1)*pi*(3-sqrt(5))``) and readout points equispaced over
  [-N/2, N/2) cycles.
* ``golden_means``-> 3-D golden-means radial ordering (Chan et al. 2009,
  phi1=0.4656, phi2=0.6823) with the same readout.

Arrays are returned in the reference layout: om is [3, nr*nspokes] with the
readout index fastest; images are [N1, N2, N3] (x fastest when flattened in
Fortran order).
"""
from __future__ import annotations

import math

import numpy as np

_MODIFIED_SHEPP_LOGAN = np.array([
    [1, .6900, .920, .810, 0, 0, 0, 0, 0, 0],
    [-.8, .6624, .874, .780, 0, -.0184, 0, 0, 0, 0],
    [-.2, .1100, .310, .220, .22, 0, 0, -18, 0, 10],
    [-.2, .1600, .410, .280, -.22, 0, 0, 18, 0, 10],
    [.1, .2100, .250, .410, 0, .35, -.15, 0, 0, 0],
    [.1, .0460, .046, .050, 0, .1, .25, 0, 0, 0],
    [.1, .0460, .046, .050, 0, -.1, .25, 0, 0, 0],
    [.1, .0460, .023, .050, -.08, -.605, 0, 0, 0, 0],
    [.1, .0230, .023, .020, 0, -.606, 0, 0, 0, 0],
    [.1, .0230, .046, .020, .06, -.605, 0, 0, 0, 0]])


def phantom3d(n, table=None):
    """Sum-of-ellipsoids phantom of size n^3 (double)."""
    e = _MODIFIED_SHEPP_LOGAN if table is None else np.asarray(table, dtype=np.float64)
    r = (np.arange(n) - (n - 1) / 2) / ((n - 1) / 2)
    Y, X, Z = np.meshgrid(r, r, r, indexing='ij')
    p = np.zeros((n, n, n))
    for A, a, b, c, x0, y0, z0, phi, theta, psi in e:
        phi, theta, psi = np.deg2rad([phi, theta, psi])
        cphi, sphi = np.cos(phi), np.sin(phi)
        cth, sth = np.cos(theta), np.sin(theta)
        cpsi, spsi = np.cos(psi), np.sin(psi)
        R = np.array([
            [cpsi * cphi - cth * sphi * spsi, cpsi * sphi + cth * cphi * spsi, spsi * sth],
            [-spsi * cphi - cth * sphi * cpsi, -spsi * sphi + cth * cphi * cpsi, cpsi * sth],
            [sth * sphi, -sth * cphi, cth]])
        xp = R[0, 0] * X + R[0, 1] * Y + R[0, 2] * Z
        yp = R[1, 0] * X + R[1, 1] * Y + R[1, 2] * Z
        zp = R[2, 0] * X + R[2, 1] * Y + R[2, 2] * Z
        dist = (xp - x0) ** 2 / a ** 2 + (yp - y0) ** 2 / b ** 2 + (zp - z0) ** 2 / c ** 2
        p[dist <= 1] += A
    return p


def phantom_complex(n):
    """Phantom with phase added where it equals 1 (``test_nufft3d.m:6-7``)."""
    p = phantom3d(n).astype(np.complex128)
    p[p == 1] = 1j
    return p


def _readout(N, nr):
    return -N / 2 + np.arange(nr) * (N / nr)


def koosh_ball(N, nspokes, nr):
    """Fibonacci-sphere koosh-ball trajectory, [3, nr*nspokes] double."""
    k = np.arange(nspokes, dtype=np.float64)
    dH = 1 - 2 * k / max(nspokes - 1, 1)
    pol = np.arccos(np.clip(dH, -1, 1))
    az = k * math.pi * (3 - math.sqrt(5))
    dirs = np.stack([np.cos(az) * np.sin(pol), np.sin(az) * np.sin(pol), np.cos(pol)])
    r = _readout(N, nr)
    om = dirs[:, None, :] * r[None, :, None]           # [3, nr, nspokes]
    return om.reshape(3, -1, order='F')


def golden_means(N, nspokes, nr, phi1=0.4656, phi2=0.6823):
    """3-D golden-means radial (hemisphere directions, full-diameter spokes)."""
    k = np.arange(1, nspokes + 1, dtype=np.float64)
    z = np.mod(k * phi1, 1.0)
    beta = 2 * math.pi * np.mod(k * phi2, 1.0)
    s = np.sqrt(1 - z * z)
    dirs = np.stack([s * np.cos(beta), s * np.sin(beta), z])
    r = _readout(N, nr)
    om = dirs[:, None, :] * r[None, :, None]
    return om.reshape(3, -1, order='F')


def coil_maps(N, nc, seed=2, radius=0.5, width=0.35):
    """Smooth complex Gaussian-bump coil sensitivities, [N, N, N, nc].

    Centres lie on a sphere of ``radius`` FOV (seeded directions), each bump
    has standard deviation ``width`` FOV and a seeded constant phase.
    """
    rng = np.random.default_rng(seed)
    r = (np.arange(N) - N / 2) / N
    X, Y, Z = np.meshgrid(r, r, r, indexing='ij')
    maps = np.empty((N, N, N, nc), dtype=np.complex128)
    for c in range(nc):
        v = rng.normal(size=3)
        v = radius * v / np.linalg.norm(v)
        ph = rng.uniform(0, 2 * math.pi)
        g = np.exp(-((X - v[0]) ** 2 + (Y - v[1]) ** 2 + (Z - v[2]) ** 2) / (2 * width ** 2))
        maps[..., c] = g * np.exp(1j * ph)
    return maps


def complex_noise(shape, sigma, seed):
    rng = np.random.default_rng(seed)
    return sigma * (rng.normal(size=shape) + 1j * rng.normal(size=shape))


def cartesian_jittered(N, centre=24, block=(2, 2, 1), seed=4):
    """3-D Cartesian sampling on the integer k-space grid [-N/2, N/2)^3
    (cycles/FOV): the central ``centre``^3 points fully sampled, elsewhere one
    random point per ``block`` of grid points (R = prod(block) outside the
    centre; a jittered, blue-noise stand-in for Poisson-disc sampling).
    Returns om [3, M] double with integer entries."""
    rng = np.random.default_rng(seed)
    bx, by, bz = block
    g = np.arange(N) - N // 2
    picks = []
    for z0 in range(0, N, bz):
        for y0 in range(0, N, by):
            # one point per block along x for this (y0, z0) slab, vectorised
            x0 = np.arange(0, N, bx)
            dx = rng.integers(0, bx, size=x0.size)
            dy = rng.integers(0, by, size=x0.size)
            dz = rng.integers(0, bz, size=x0.size)
            picks.append(np.stack([x0 + dx, y0 + dy, z0 + dz]))
    p = np.concatenate(picks, axis=1)
    h = centre // 2
    c = np.arange(N // 2 - h, N // 2 - h + centre)
    C = np.stack(np.meshgrid(c, c, c, indexing='ij')).reshape(3, -1)
    inside = np.all((p >= N // 2 - h) & (p < N // 2 - h + centre), axis=0)
    idx = np.concatenate([p[:, ~inside], C], axis=1)
    return g[idx].astype(np.float64)
