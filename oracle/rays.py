"""CPU restatement of the reference's ray tracing -- TEST INFRASTRUCTURE ONLY.

tracerays / tracerays_singlesrc (src/EikSolv/eikutils.jl:201-328) interpolate each traveltime
grid with Interpolations.jl ``scale(interpolate(ttime, BSpline(Cubic())), axes...)`` and descend
its gradient.  Interpolations.jl is not vendored (not in /root/reference); its published cubic
B-spline algorithm is restated here:

* coefficients: for every grid line of n values, n + 2 coefficients c solve the banded system
  whose inner rows are the interpolation conditions (c[i-1] + 4 c[i] + c[i+1]) / 6 = y[i-1] and
  whose first and last rows are the default ``Line(OnGrid())`` boundary, a zero second difference
  c[0] - 2 c[1] + c[2] = 0 (resp. at the end); the tensor product is built axis by axis;
* evaluation: sum over the 4^nd neighbouring coefficients of c times the product of the centred
  cubic B-spline B(t) (2/3 - t^2 + |t|^3/2 for |t| <= 1, (2 - |t|)^3 / 6 for |t| <= 2) at
  t = u - (p - 1), u the position in node units; the gradient uses B'(t) divided by the spacing.
  The cell is clamped to [0, n-2] so the last node is evaluated with t = 1;
* the ray loop, distances and the Grid2DSphere step follow eikutils.jl:241-320 (polar step through
  Cartesian coordinates, x = r cos(theta), y = r sin(theta), utils.jl:351-367).

Pins (tests/test_oracle_rays.py): the interpolant reproduces the data at the nodes and linear
functions exactly, the boundary rows hold, and rays in a constant-velocity model are straight
segments to the source.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import solve_banded


def _line_system(n):
    """Banded form (l = u = 2) of the (n+2) x (n+2) prefilter matrix."""
    ab = np.zeros((5, n + 2))
    # row 0: c0 - 2 c1 + c2 = 0
    ab[2, 0] = 1.0
    ab[1, 1] = -2.0
    ab[0, 2] = 1.0
    for i in range(1, n + 1):
        ab[3, i - 1] = 1.0 / 6.0
        ab[2, i] = 4.0 / 6.0
        ab[1, i + 1] = 1.0 / 6.0
    # last row: c[n-1] - 2 c[n] + c[n+1] = 0
    ab[4, n - 1] = 1.0
    ab[3, n] = -2.0
    ab[2, n + 1] = 1.0
    return ab


def bspline_coefs(y: np.ndarray) -> np.ndarray:
    """Padded coefficient array (each extent + 2) of an nd array of node values."""
    c = np.asarray(y, dtype=np.float64)
    for ax in range(c.ndim):
        n = c.shape[ax]
        ab = _line_system(n)
        moved = np.moveaxis(c, ax, 0)
        rhs = np.zeros((n + 2,) + moved.shape[1:])
        rhs[1:n + 1] = moved
        sol = solve_banded((2, 2), ab, rhs.reshape(n + 2, -1)).reshape(rhs.shape)
        c = np.moveaxis(sol, 0, ax)
    return c


def _B(t):
    t = abs(t)
    if t <= 1.0:
        return 2.0 / 3.0 - t * t + 0.5 * t ** 3
    if t <= 2.0:
        return (2.0 - t) ** 3 / 6.0
    return 0.0


def _dB(t):
    a = abs(t)
    sg = 1.0 if t >= 0 else -1.0
    if a <= 1.0:
        return sg * (-2.0 * a + 1.5 * a * a)
    if a <= 2.0:
        return sg * (-0.5 * (2.0 - a) ** 2)
    return 0.0


def _cell(x, o, h, n):
    u = (x - o) / h
    i = min(max(int(math.floor(u)), 0), n - 2)
    return u, i


def evaluate(c, o, h, x, deriv=None):
    """Interpolant (deriv None) or its derivative along axis `deriv` at point x."""
    nd = c.ndim
    n = [s - 2 for s in c.shape]
    cells = [_cell(x[d], o[d], h[d], n[d]) for d in range(nd)]
    total = 0.0
    for off in np.ndindex(*([4] * nd)):
        w = 1.0
        idx = []
        for d in range(nd):
            u, i = cells[d]
            p = i + off[d]            # padded index; the spline is centred at node p - 1
            t = u - (p - 1)
            w *= (_dB(t) / h[d]) if d == deriv else _B(t)
            idx.append(p)
        total += w * c[tuple(idx)]
    return total


def _dist(kind, a, b):
    if kind == "sphere2d":
        return math.sqrt(a[0] ** 2 + b[0] ** 2 - 2.0 * a[0] * b[0] * math.cos(math.radians(a[1] - b[1])))
    return math.sqrt(sum((ai - bi) ** 2 for ai, bi in zip(a, b)))


def trace(grd, ttime, src, recs, steplen=0.01, max_steps=10 ** 7):
    """Rays of one source (tracerays_singlesrc)."""
    from eikonalsolvers.grids import Grid2DCart, Grid2DSphere, Grid3DCart

    if isinstance(grd, (Grid2DCart, Grid3DCart)):
        kind = "cart"
        h = [grd.hgrid] * grd.ndim
        rstep = steplen * grd.hgrid
        thr = grd.hgrid * math.sqrt(2.0)
    elif isinstance(grd, Grid2DSphere):
        kind = "sphere2d"
        h = [grd.Δr, grd.Δθ]
        rstep = steplen
        thr = grd.Δr * math.sqrt(2.0)
    else:
        raise ValueError("tracerays_singlesrc(): ray tracing for spherical grids not yet implemented.")
    nd = grd.ndim
    o = list(grd.cooinit)
    end = [o[d] + (grd.grsize[d] - 1) * h[d] for d in range(nd)]
    c = bspline_coefs(ttime)
    rays = []
    for rec in np.asarray(recs, dtype=np.float64).reshape(-1, nd):
        path = [list(rec)]
        p = list(rec)
        dist = _dist(kind, p, src)
        steps = 0
        while dist >= thr:
            if steps >= max_steps:
                raise RuntimeError("ray did not reach the source")
            g = [evaluate(c, o, h, p, d) for d in range(nd)]
            if kind == "cart":
                nrm = math.sqrt(sum(v * v for v in g))
                newpt = [p[d] - rstep * (g[d] / nrm) for d in range(nd)]
            else:
                rc, th = p[0], math.radians(p[1])
                gth = math.degrees(g[1])
                dx = math.cos(th) * g[0] - 1.0 / rc * math.sin(th) * gth
                dy = math.sin(th) * g[0] + 1.0 / rc * math.cos(th) * gth
                nrm = math.sqrt(dx * dx + dy * dy)
                x = rc * math.cos(math.radians(p[1])) - rstep * (dx / nrm)
                y = rc * math.sin(math.radians(p[1])) - rstep * (dy / nrm)
                newpt = [math.sqrt(x * x + y * y), math.degrees(math.atan2(y, x))]
            for d in range(nd):
                if newpt[d] < o[d]:
                    newpt[d] = o[d]
                elif newpt[d] > end[d]:
                    newpt[d] = end[d]
            p = newpt
            path.append(list(p))
            dist = _dist(kind, p, src)
            steps += 1
        path[-1] = list(np.asarray(src, dtype=np.float64))
        rays.append(np.array(path))
    return rays
