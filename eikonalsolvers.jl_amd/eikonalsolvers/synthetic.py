"""Deterministic synthetic workloads with the shapes of the benchmark configurations.

Random numbers come from a counter-based generator (splitmix64 of (seed, index), Box-Muller for
normals) so any machine regenerates bit-identical inputs without shipping the fields.

* ``c1_2d_forward``   -- 2D 300x220, h=0.5, v = 2.5 + 0.034*j, 4 sources, 10 receivers each
* ``c2_3d``           -- 3D 201x201x101, h=1, v = 3 + 0.02*z + Gaussian random field (std 0.2,
                         correlation 15 cells), 16 sources at depth 60-90, 64 top-face receivers,
                         observed picks from a second perturbed model plus N(0, 0.01^2) noise
* ``c3_2d_batched``   -- 2D 600x400, h=0.25, 8 random layers + smooth field, 256 sources, 128
                         surface receivers each
* ``c4_3d_sphere``    -- 3D spherical 120x150x150 (dr = 2 km, dtheta = dphi = 0.1 deg, outer radius
                         6371 km), v(r) = 5.8 -> 8.0 km/s with depth times (1 + 0.03 GRF), 32
                         sources at 10-150 km depth, 200 surface receivers
* ``c5_request``      -- one request of the serving mix (seeded by its id): 70 % forward / 30 %
                         gradient, 2D 100^2-400^2 or 3D 64^3-160^3 (h=1), 1-64 sources, 16-128
                         receivers, smooth random velocity in [1.5, 6]
* ``c5_arrivals``     -- Poisson arrival times (seed 50501)
"""
from __future__ import annotations

import numpy as np

from .grids import Grid2DCart, Grid3DCart, Grid3DSphere

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """splitmix64 of the counter seed*2^32 + idx (vectorised, wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) << np.uint64(32)) + idx.astype(np.uint64)
        z = z * np.uint64(0x9E3779B97F4A7C15) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, np.arange(n)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = np.maximum(u[:m], 1e-300)
    u2 = u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    return np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]


def gaussian_random_field(shape, seed: int, corr_cells: float, std: float) -> np.ndarray:
    """White N(0,1) filtered by a separable periodic Gaussian of sigma=corr_cells (FFT), rescaled."""
    w = normal(seed, int(np.prod(shape))).reshape(shape, order="F")
    spec = np.fft.rfftn(w, axes=list(range(len(shape))))
    for ax, n in enumerate(shape):
        k = np.fft.fftfreq(n) if ax < len(shape) - 1 else np.fft.rfftfreq(n)
        g = np.exp(-2.0 * (np.pi * k * corr_cells) ** 2)
        sh = [1] * len(shape)
        sh[ax] = g.size
        spec = spec * g.reshape(sh)
    f = np.fft.irfftn(spec, s=shape, axes=list(range(len(shape))))
    f = (f - f.mean()) / (f.std() + 1e-300)
    return std * f


def c1_2d_forward():
    grd = Grid2DCart(hgrid=0.5, cooinit=(0.0, 0.0), grsize=(300, 220))
    nx, ny = grd.grsize
    vel = np.empty((nx, ny))
    vel[:, :] = 2.5 + 0.034 * np.arange(1, ny + 1)[None, :]
    src = np.column_stack([grd.hgrid * np.linspace(10, 290, 4), grd.hgrid * 200 * np.ones(4)])
    rec = np.column_stack([grd.hgrid * np.linspace(8, 294, 10), grd.hgrid * 20 * np.ones(10)])
    return grd, vel, src, [rec.copy() for _ in range(4)]


def c2_3d(nsrc: int = 16, grsize=(201, 201, 101)):
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=tuple(grsize))
    nx, ny, nz = grd.grsize
    z = grd.z[None, None, :]
    vel = np.clip(3.0 + 0.02 * z + gaussian_random_field(grd.grsize, 20261, 15.0, 0.2), 1.5, 6.0)
    u = uniform(20262, 3 * nsrc).reshape(nsrc, 3)
    src = np.column_stack([u[:, 0] * grd.x[-1], u[:, 1] * grd.y[-1], 60.0 + 30.0 * u[:, 2]])
    lat = np.linspace(10.0, 190.0, 8) * (grd.x[-1] / 200.0)
    rec = np.array([[x, y, 0.0] for y in lat for x in lat])
    velobs = np.clip(3.0 + 0.02 * z + gaussian_random_field(grd.grsize, 20263, 15.0, 0.2), 1.5, 6.0)
    return grd, vel, src, [rec.copy() for _ in range(nsrc)], velobs


def c2_observed(picks_obs_model, nsrc: int):
    """Observed picks = forward on the second model + N(0, 0.01^2) noise (seed 20264)."""
    noise = normal(20264, sum(len(p) for p in picks_obs_model)) * 0.01
    out, k = [], 0
    for p in picks_obs_model:
        out.append(np.asarray(p) + noise[k:k + len(p)])
        k += len(p)
    std = [np.full(len(p), 0.01) for p in picks_obs_model]
    return out, std


def c3_2d_batched(nsrc: int = 256, nrec: int = 128):
    grd = Grid2DCart(hgrid=0.25, cooinit=(0.0, 0.0), grsize=(600, 400))
    nx, ny = grd.grsize
    depths = np.sort(uniform(30301, 8)) * grd.y[-1]
    vlay = 1.5 + 4.0 * uniform(30302, 9)
    y = grd.y
    layer = np.searchsorted(depths, y)
    vel = np.empty((nx, ny))
    vel[:, :] = vlay[layer][None, :]
    vel = vel + gaussian_random_field(grd.grsize, 30303, 20.0, 0.1)
    vel = np.clip(vel, 1.2, 6.0)
    u = uniform(30304, 2 * nsrc).reshape(nsrc, 2)
    src = np.column_stack([grd.hgrid + u[:, 0] * (grd.x[-1] - 2 * grd.hgrid),
                           grd.hgrid + u[:, 1] * (grd.y[-1] - 2 * grd.hgrid)])
    rec = np.column_stack([np.linspace(0.5, 149.25, nrec), np.zeros(nrec)])
    return grd, vel, src, [rec.copy() for _ in range(nsrc)]


def smooth_velocity(shape, seed: int, lo: float = 1.5, hi: float = 6.0, nmodes: int = 6) -> np.ndarray:
    """Sum of a few random low-wavenumber cosine modes, mapped to [lo, hi]; cheap on large grids."""
    u = uniform(seed, nmodes * (2 * len(shape) + 1)).reshape(nmodes, -1)
    f = np.zeros(shape)
    axes = [np.arange(n, dtype=np.float64) / n for n in shape]
    for m in range(nmodes):
        arg = None
        for d, a in enumerate(axes):
            term = (1.0 + 3.0 * u[m, d]) * 2.0 * np.pi * a
            sh = [1] * len(shape)
            sh[d] = shape[d]
            term = term.reshape(sh)
            arg = term if arg is None else arg + term
        f = f + np.cos(arg + 2.0 * np.pi * u[m, -1])
    f = (f - f.min()) / max(f.max() - f.min(), 1e-300)
    return lo + (hi - lo) * (0.15 + 0.7 * f)


def c5_request(rid: int, small: bool = False) -> dict:
    """Request `rid` of the serving mix.  small=True shrinks grids and source counts (tests)."""
    u = uniform(50500 + rid, 8)
    gradient = u[0] < 0.3
    three_d = u[1] < 0.5
    if three_d:
        n = int(24 + 16 * u[2]) if small else int(64 + 96 * u[2])
        grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(n, n, n))
    else:
        n = int(100 + 60 * u[2]) if small else int(100 + 300 * u[2])
        grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(n, n))
    nsrc = 1 + int((7 if small else 63) * u[3])
    nrec = 16 + int(112 * u[4])
    vel = smooth_velocity(grd.grsize, 50600 + rid)
    L = np.array([(k - 1) * grd.hgrid for k in grd.grsize])
    ps = uniform(50700 + rid, nsrc * grd.ndim).reshape(nsrc, grd.ndim)
    src = (0.05 + 0.9 * ps) * L
    pr = uniform(50800 + rid, nrec * grd.ndim).reshape(nrec, grd.ndim)
    rec = pr * L
    rec[:, -1] = 0.0  # top face / surface
    out = {"id": rid, "kind": "gradient" if gradient else "traveltime", "grd": grd, "vel": vel, "src": src,
           "recs": [rec.copy() for _ in range(nsrc)]}
    if gradient:
        # observations: straight-ray times in a uniform 3.5 km/s medium (a plausible misfit)
        obs = [np.sqrt(((rec - s) ** 2).sum(1)) / 3.5 for s in src]
        out["pickobs"] = obs
        out["stdobs"] = [np.full(nrec, 0.01) for _ in range(nsrc)]
    return out


def c5_arrivals(n: int, rate: float, seed: int = 50501) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson stream with the given rate (1/s)."""
    gaps = -np.log(np.maximum(uniform(seed, n), 1e-300)) / rate
    return np.cumsum(gaps) - gaps[0]


def c4_3d_sphere(nsrc: int = 32, nrec: int = 200, grsize=(120, 150, 150)):
    nr, nt, npp = grsize
    grd = Grid3DSphere(Δr=2.0, Δθ=0.1, Δφ=0.1, grsize=tuple(grsize), cooinit=(6371.0 - 2.0 * (nr - 1), 82.55, 10.0))
    r = grd.r[:, None, None]
    v1 = 5.8 + 2.2 * (6371.0 - r) / 238.0
    vel = v1 * (1.0 + 0.03 * gaussian_random_field(grd.grsize, 40401, 10.0, 1.0))
    u = uniform(40402, 3 * nsrc).reshape(nsrc, 3)
    th0, th1 = grd.θ[0], grd.θ[-1]
    ph0, ph1 = grd.φ[0], grd.φ[-1]
    depth = np.minimum(10.0 + 140.0 * u[:, 0], 6371.0 - grd.r[0])
    src = np.column_stack([6371.0 - depth, th0 + (0.05 + 0.9 * u[:, 1]) * (th1 - th0),
                           ph0 + (0.05 + 0.9 * u[:, 2]) * (ph1 - ph0)])
    w = uniform(40403, 2 * nrec).reshape(nrec, 2)
    rec = np.column_stack([np.full(nrec, 6371.0), th0 + w[:, 0] * (th1 - th0), ph0 + w[:, 1] * (ph1 - ph0)])
    return grd, vel, np.ascontiguousarray(src), [rec.copy() for _ in range(nsrc)]
