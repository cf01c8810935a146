"""CPU-side checks of the drop-in boundary: libeik.so loads and exports every symbol that
include/eik.h declares (no device call is made), and the host mirror validates its inputs the way
the reference does before anything reaches the device."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "eik.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(eik_\w+)\s*\(", src, flags=re.M)))


def test_header_and_binding_agree():
    from eikonalsolvers import _lib
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from eikonalsolvers import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("libeik.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIBPATH)
    for name in declared_symbols():
        assert hasattr(L, name), name
    _lib.load()


def test_extraparams_validation_matches_reference():
    from eikonalsolvers.grids import ExtraParams
    with pytest.raises(ValueError):
        ExtraParams(smoothgradkern=4)
    with pytest.raises(ValueError):
        ExtraParams(parallelkind="gpu")
    with pytest.raises(AssertionError):
        ExtraParams(radiussmoothgradsrc=-1)
    assert ExtraParams().refinearoundsrc and ExtraParams().radiussmoothgradsrc == 3


def test_grid_validation_matches_reference():
    from eikonalsolvers.grids import Grid2DCart, Grid2DSphere, Grid3DSphere
    with pytest.raises(AssertionError):
        Grid2DCart(hgrid=-1.0, cooinit=(0.0, 0.0), grsize=(3, 3))
    with pytest.raises(AssertionError):
        Grid2DSphere(Δr=1.0, Δθ=10.0, grsize=(3, 30), cooinit=(10.0, 0.0))  # theta beyond 180
    with pytest.raises(AssertionError):
        Grid3DSphere(Δr=1.0, Δθ=1.0, Δφ=1.0, grsize=(3, 3, 3), cooinit=(10.0, 2.0, 0.0))  # theta < 5
    g = Grid2DCart(hgrid=0.5, cooinit=(1.0, 2.0), grsize=(4, 3))
    np.testing.assert_array_equal(g.x, [1.0, 1.5, 2.0, 2.5])
    assert g.cooend == (2.5, 3.0)


def test_ray_step_cap_is_a_length_on_wide_spheres():
    """The ray step cap counts 20 domain lengths of steps on every grid kind: on a Grid2DSphere the
    length is the outer diameter (a length in km, not its square root).  A 40 degree sector at
    r = 6371 km with 0.5 km steps must allow the ~4500 km chord between its ends many times over."""
    import ctypes
    from eikonalsolvers import _lib
    from eikonalsolvers.api import _grid_struct
    from eikonalsolvers.grids import Grid2DCart, Grid2DSphere
    if not os.path.exists(_lib.LIBPATH):
        pytest.skip("libeik.so not built (run __graft_entry__.build())")
    L = _lib.load()
    out = ctypes.c_int64(0)
    grd = Grid2DSphere(Δr=2.0, Δθ=0.1, grsize=(120, 401), cooinit=(6133.0, 40.0))
    assert L.eik_tracerays_max_steps(ctypes.byref(_grid_struct(grd)), 0.5, ctypes.byref(out)) == 0
    rmax = 6133.0 + 119 * 2.0
    chord = 2 * rmax * np.sin(np.deg2rad(40.0) / 2)  # the longest straight path in the sector
    assert out.value == int(20.0 * 2 * rmax / 0.5 + 1000.0)
    assert out.value * 0.5 > 20 * chord
    # Cartesian: 20 box diagonals of steplen*hgrid
    g2 = Grid2DCart(hgrid=0.25, cooinit=(0.0, 0.0), grsize=(600, 400))
    assert L.eik_tracerays_max_steps(ctypes.byref(_grid_struct(g2)), 0.1, ctypes.byref(out)) == 0
    diag = np.hypot(599 * 0.25, 399 * 0.25)
    assert out.value == int(20.0 * diag / (0.1 * 0.25) + 1000.0)
