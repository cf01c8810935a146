"""Deterministic test models shared by the CPU and GPU tests (seeded numpy generators)."""
import numpy as np

from eikonalsolvers.grids import Grid3DSphere, Grid2DCart, Grid2DSphere, Grid3DCart


def smooth_field(shape, seed, passes=8, lo=1.5, hi=6.0, base=3.0, amp=1.0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape)
    for _ in range(passes):
        acc = v.copy()
        for ax in range(v.ndim):
            acc += np.roll(v, 1, axis=ax) + np.roll(v, -1, axis=ax)
        v = acc / (1 + 2 * v.ndim)
    v = (v - v.mean()) / (v.std() + 1e-12)
    return np.clip(base + amp * v, lo, hi)


def c1_model():
    """The reference README's 2D forward case: 300x220, h=0.5, v = 2.5 + 0.034*j, 4 sources."""
    grd = Grid2DCart(hgrid=0.5, cooinit=(0.0, 0.0), grsize=(300, 220))
    nx, ny = grd.grsize
    vel = np.zeros((nx, ny))
    for j in range(ny):
        vel[:, j] = 2.5 + 0.034 * (j + 1)
    src = np.column_stack([grd.hgrid * np.linspace(10, 290, 4), grd.hgrid * 200 * np.ones(4)])
    rec1 = np.column_stack([grd.hgrid * np.linspace(8, 294, 10), grd.hgrid * 20 * np.ones(10)])
    return grd, vel, src, [rec1.copy() for _ in range(4)]


def random2d(seed=1, n=(96, 80), h=1.0, nsrc=3, nrec=12):
    grd = Grid2DCart(hgrid=h, cooinit=(0.0, 0.0), grsize=n)
    vel = smooth_field(n, seed)
    rng = np.random.default_rng(seed + 100)
    L = np.array([(n[0] - 1) * h, (n[1] - 1) * h])
    src = rng.uniform(0.05, 0.95, size=(nsrc, 2)) * L
    recs = [rng.uniform(0.0, 1.0, size=(nrec, 2)) * L for _ in range(nsrc)]
    return grd, vel, src, recs


def random3d(seed=2, n=(40, 36, 28), h=1.0, nsrc=2, nrec=10):
    grd = Grid3DCart(hgrid=h, cooinit=(0.0, 0.0, 0.0), grsize=n)
    vel = smooth_field(n, seed, passes=4)
    rng = np.random.default_rng(seed + 100)
    L = np.array([(n[0] - 1) * h, (n[1] - 1) * h, (n[2] - 1) * h])
    src = rng.uniform(0.1, 0.9, size=(nsrc, 3)) * L
    recs = []
    for _ in range(nsrc):
        r = rng.uniform(0.0, 1.0, size=(nrec, 3)) * L
        r[:, 2] = 0.0  # top face
        recs.append(r)
    return grd, vel, src, recs


def sphere3d(seed=11):
    """Small 3D spherical case far from the equator (theta 30-34.6 deg), so the sin(theta) of the
    phi metric is ~0.5 and matters."""
    grd = Grid3DSphere(Δr=2.0, Δθ=0.2, Δφ=0.2, grsize=(20, 24, 22), cooinit=(6300.0, 30.0, 10.0))
    vel = smooth_field(grd.grsize, seed, lo=5.0, hi=8.0, base=6.5, amp=0.4)
    src = np.array([[6310.3, 31.13, 11.07], [6322.9, 33.41, 12.93]])
    th, ph = np.meshgrid(np.linspace(30.4, 34.2, 4), np.linspace(10.4, 13.8, 3), indexing="ij")
    rec = np.column_stack([np.full(th.size, 6336.0), th.ravel(), ph.ravel()])
    return grd, vel, src, [rec, rec.copy()]


def sphere2d(seed=3):
    grd = Grid2DSphere(Δr=2.0, Δθ=0.2, grsize=(60, 90), cooinit=(6251.0, 40.0))
    vel = smooth_field(grd.grsize, seed, lo=4.0, hi=9.0, base=6.5, amp=0.5)
    src = np.array([[6300.0, 47.3], [6280.5, 52.05]])
    recs = [np.column_stack([6369.0 * np.ones(8), np.linspace(41.0, 57.0, 8)]) for _ in range(2)]
    return grd, vel, src, recs
