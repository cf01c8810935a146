"""Reverse-mode rules for the misfit (no counterpart in the reference, which has no ChainRulesCore
dependency, Project.toml:6-13).

``rrule_eikttimemisfit`` follows ChainRulesCore's ``rrule`` contract for ``eikttimemisfit``
(eikmaingrad.jl:1302-1328): it returns the misfit and a pullback mapping a cotangent ``dmisf`` to
the cotangents of ``vel`` and ``coordsrc``.  One ``eik_gradient`` call (``:gradvelandsrcloc``)
computes the misfit and both gradients, so the pullback only scales them.  The gradients are the
ones ``eikgradient`` returns: with the default ``ExtraParams`` they include the source-region taper
(``radiussmoothgradsrc``) and, when ``smoothgradkern > 0``, the smoothing; set
``radiussmoothgradsrc=0`` for the exact derivative of the discrete misfit.

``misfit_torch`` wraps the same rule as a ``torch.autograd.Function`` so the misfit composes with
PyTorch autograd (velocity and source coordinates as float64 tensors of any device; the solve
runs on the model's GPU through libeik).
"""
from __future__ import annotations

import numpy as np

from .api import EikModel


def rrule_eikttimemisfit(vel, grd, coordsrc, coordrec, pickobs, stdobs, *, extraparams=None, model=None):
    """(misf, pullback) with pullback(dmisf) -> (dvel, dcoordsrc).

    ``model``: an ``EikModel`` on ``grd`` to reuse (its velocity is replaced by ``vel``); otherwise
    a temporary one is created.
    """
    own = model is None
    m = EikModel(vel, grd) if own else model
    try:
        if not own:
            m.update_vel(vel)
        gv, gs, misf = m.gradient(coordsrc, coordrec, pickobs, stdobs, "gradvelandsrcloc", extraparams=extraparams)
    finally:
        if own:
            m.close()

    def pullback(dmisf):
        d = float(dmisf)
        return d * gv, d * gs

    return misf, pullback


def misfit_torch(vel, coordsrc, model, coordrec, pickobs, stdobs, *, extraparams=None):
    """The misfit as a differentiable torch scalar of ``vel`` (grid-shaped, column-major index
    meaning as everywhere in this package) and ``coordsrc`` (nsrc x ndim)."""
    import torch

    class _EikMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v, cs):
            vn = v.detach().to("cpu", torch.float64).numpy()
            cn = cs.detach().to("cpu", torch.float64).numpy()
            want_v, want_s = ctx.needs_input_grad
            model.update_vel(vn)
            if want_v and want_s:
                gv, gs, misf = model.gradient(cn, coordrec, pickobs, stdobs, "gradvelandsrcloc", extraparams=extraparams)
            elif want_v:
                gv, misf = model.gradient(cn, coordrec, pickobs, stdobs, "gradvel", extraparams=extraparams)
                gs = np.zeros_like(cn)
            elif want_s:
                gs, misf = model.gradient(cn, coordrec, pickobs, stdobs, "gradsrcloc", extraparams=extraparams)
                gv = np.zeros_like(vn)
            else:
                misf = model.misfit(cn, coordrec, pickobs, stdobs, extraparams=extraparams)
                gv, gs = np.zeros_like(vn), np.zeros_like(cn)
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(gv)).to(v.device),
                                  torch.from_numpy(np.ascontiguousarray(gs)).to(cs.device))
            return torch.tensor(misf, dtype=torch.float64, device=v.device)

        @staticmethod
        def backward(ctx, g):
            gv, gs = ctx.saved_tensors
            return g.to(gv.device) * gv, g.to(gs.device) * gs

    return _EikMisfit.apply(vel, coordsrc)
