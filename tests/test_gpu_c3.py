"""2D batched gradients at the C3 configuration: 600x400 grid (h = 0.25), 8 random layers plus a
smooth random field, 256 sources with continuous positions, 128 surface receivers each; all 256
sources run in shared launches.

* FP64 :gradvelandsrcloc against the oracle (1e-7 relative L2, rounding level);
* mixed precision (FP32 stencil increments on FP64 traveltimes, FP64 adjoint) against the FP64
  oracle within the stated 1e-3 relative L2 (the oracle's restatement of it measures :gradvel
  3.6e-5, :gradsrcloc 6.8e-6);
* the FP64 and the mixed adjoint against central finite differences of the FP64 misfit on 20 random cells
  (seed 30305, dh = 1e-3 km/s) and on all 512 source coordinates (dh = 1e-4 km), the steps the
  reference's own tests use (test_cartesian_2D.jl:226,323).  Finite differences of a heap FMM are
  noisy (acceptance order can change under the perturbation), so the comparison is in relative L2
  over the sampled entries; a source coordinate whose differences at dh and dh/2 disagree sits on a
  jump of the misfit and is counted, not compared (28 of 512 here).

Observed picks: FP64 forward on the model scaled by 1 + 0.02 * smooth field (seed 30306), plus
N(0, 0.01^2) noise (seed 30307); std 0.01.
"""
import numpy as np
import pytest

from eikonalsolvers.grids import ExtraParams
from eikonalsolvers.synthetic import c3_2d_batched, gaussian_random_field, normal, uniform

pytestmark = pytest.mark.gpu



@pytest.fixture(scope="module")
def c3(eik):
    grd, vel, src, recs = c3_2d_batched()
    velobs = vel * (1.0 + 0.02 * gaussian_random_field(grd.grsize, 30306, 20.0, 1.0))
    model_obs = eik.EikModel(velobs, grd)
    picks = model_obs.traveltime(src, recs, extraparams=ExtraParams())
    model_obs.close()
    noise = normal(30307, sum(len(p) for p in picks)) * 0.01
    obs, k = [], 0
    for p in picks:
        obs.append(p + noise[k:k + len(p)])
        k += len(p)
    std = [np.full(len(p), 0.01) for p in picks]
    model = eik.EikModel(vel, grd)
    yield grd, vel, src, recs, obs, std, model
    model.close()


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def oracle_grad(c3, oracle):
    # one thread = the reference's serial source loop: its per-worker gradient buffer carries the
    # previous source's value at source nodes (calcgradsomesrc), so the chunking changes results
    grd, vel, src, recs, obs, std, _ = c3
    return oracle.gradient(vel, grd, src, recs, obs, std, "gradvelandsrcloc", extraparams=ExtraParams(),
                           nthreads=1)


def test_c3_fp64_gradient_matches_oracle(c3, oracle_grad):
    grd, vel, src, recs, obs, std, model = c3
    gv, gs, mg = model.gradient(src, recs, obs, std, "gradvelandsrcloc", extraparams=ExtraParams())
    ov, os_, mo = oracle_grad
    print(f"C3 fp64: gradvel rel L2 {_rel_l2(gv, ov):.3e}, gradsrc rel L2 {_rel_l2(gs, os_):.3e}, "
          f"misfit {mg:.10e} vs {mo:.10e}", flush=True)
    assert mg == pytest.approx(mo, rel=1e-10)
    assert _rel_l2(gv, ov) <= 1e-7
    assert _rel_l2(gs, os_) <= 1e-7


def test_c3_mixed_gradsrc_within_tolerance(c3, oracle_grad):
    grd, vel, src, recs, obs, std, model = c3
    gv, gs, mg = model.gradient(src, recs, obs, std, "gradvelandsrcloc", extraparams=ExtraParams(precision="mixed"))
    ov, os_, mo = oracle_grad
    print(f"C3 mixed: gradvel rel L2 {_rel_l2(gv, ov):.3e}, gradsrc rel L2 {_rel_l2(gs, os_):.3e}, "
          f"misfit rel {abs(mg - mo) / mo:.3e}", flush=True)
    assert _rel_l2(gs, os_) <= 1e-3
    assert mg == pytest.approx(mo, rel=1e-5)


def test_c3_mixed_gradvel_within_tolerance(c3, oracle_grad):
    grd, vel, src, recs, obs, std, model = c3
    gv, mg = model.gradient(src, recs, obs, std, "gradvel", extraparams=ExtraParams(precision="mixed"))
    ov, os_, mo = oracle_grad
    print(f"C3 mixed :gradvel rel L2 {_rel_l2(gv, ov):.3e}, misfit rel {abs(mg - mo) / mo:.3e}", flush=True)
    assert np.all(np.isfinite(gv))
    assert _rel_l2(gv, ov) <= 1e-3


FD_TOL_VEL = 2e-3  # measured 4.4e-4 (FP64 adjoint, 20 cells)
FD_TOL_SRC = 5e-3  # measured 1.6e-3 over the 484 resolved source coordinates (oracle, FP64)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_adjoint_matches_finite_differences(c3, precision):
    """The adjoint of either precision against central differences of the FP64 misfit (an FP32
    misfit's differences would be rounding noise)."""
    grd, vel, src, recs, obs, std, model = c3
    ep = ExtraParams(radiussmoothgradsrc=0)
    gv, gs, _ = model.gradient(src, recs, obs, std, "gradvelandsrcloc",
                               extraparams=ExtraParams(radiussmoothgradsrc=0, precision=precision))
    nx, ny = grd.grsize
    u = uniform(30305, 40).reshape(20, 2)
    cells = [(int(a * (nx - 1)), int(b * (ny - 1))) for a, b in u]
    dh = 1e-3
    fd = []
    for (i, j) in cells:
        vp, vm = vel.copy(), vel.copy()
        vp[i, j] += dh
        vm[i, j] -= dh
        model.update_vel(vp)
        mp = model.misfit(src, recs, obs, std, extraparams=ep)
        model.update_vel(vm)
        mm = model.misfit(src, recs, obs, std, extraparams=ep)
        fd.append((mp - mm) / (2 * dh))
    model.update_vel(vel)
    adj = np.array([gv[i, j] for (i, j) in cells])
    fd = np.array(fd)
    ev = _rel_l2(adj, fd)
    # every coordinate of every source (512): a perturbed source only changes its own misfit term,
    # so the 2 x 512 perturbed sources run as one batched forward call and each term is summed here
    ns = len(src)

    def fd_sources(dhs):
        pert, prec = [], []
        for s in range(ns):
            for d in range(2):
                for sg in (1.0, -1.0):
                    q = src[s].copy()
                    q[d] += sg * dhs
                    pert.append(q)
                    prec.append(recs[s])
        picks = model.traveltime(np.array(pert), prec, extraparams=ep)
        k, out = 0, np.zeros((ns, 2))
        for s in range(ns):
            for d in range(2):
                mp = 0.5 * float(np.sum((picks[k] - obs[s]) ** 2 / std[s] ** 2))
                mm = 0.5 * float(np.sum((picks[k + 1] - obs[s]) ** 2 / std[s] ** 2))
                out[s, d] = (mp - mm) / (2 * dhs)
                k += 2
        return out

    # The heap FMM's misfit is only piecewise smooth in the source position: where the source sits
    # within dh of a point at which the heap's order flips (e.g. halfway between two refined-grid
    # nodes, where two source-box corners are equidistant) the misfit jumps and a central
    # difference measures the jump, not a derivative.  A coordinate counts as resolved when the
    # differences with dh = 1e-4 and 5e-5 agree to 1 %; the adjoint is compared on those.
    fds, fds2 = fd_sources(1e-4), fd_sources(5e-5)
    resolved = np.abs(fds - fds2) <= 1e-2 * np.maximum(np.abs(fds), np.abs(fds2))
    es = _rel_l2(gs[resolved], fds[resolved])
    print(f"C3 FD check ({precision} adjoint): gradvel rel L2 over 20 cells {ev:.3e} (|g| ~ {np.abs(adj).mean():.3e}); "
          f"gradsrc rel L2 over the {int(resolved.sum())} of {2 * ns} source coordinates whose differences are "
          f"resolved {es:.3e} (all {2 * ns}: {_rel_l2(gs, fds):.3e})", flush=True)
    # a few times the measured errors: 4.4e-4 (cells, FP64) and 1.6e-3 over 484 resolved coordinates
    assert ev <= FD_TOL_VEL
    assert resolved.sum() >= 0.9 * resolved.size
    assert es <= FD_TOL_SRC
