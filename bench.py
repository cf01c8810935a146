#!/usr/bin/env python3
"""Benchmark of the eikonal hot path on one MI355X.

Workload: 3D Cartesian grid 201x201x101 (h = 1), v = 3 + 0.02 z + Gaussian random field
(std 0.2 km/s, correlation 15 cells, clipped to [1.5, 6]), 16 sources at depth 60-90, 64 receivers
on the top face, default ExtraParams (refinement around the source, df = 5, noderadius = 3).
A step = one eiktraveltime call over the 16 sources with the model resident in HBM.
value = source-solves/s in the reference's FP64 arithmetic.  The mixed-precision rate (FP64 values
and keys, FP32 quadratic increments), the FP32 throughput mode, the gradient-evaluation rates,
the rooflines of the FP64 and FP32 round kernels, the serving mix, and CPU baselines for the
forward and the gradient (the oracle restating the reference's heap FMM and adjoint, on bounded
samples of sources) ride on the same line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "eikonalsolvers.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "source-solves/s + gradient evals/s, 3D 201²×101 grid; serving p50/p99 latency"
HBM_PEAK_GBS = 8000.0
TILE_NODES_3D = 64  # forward tiles are 4x4x4 (one wavefront each)
# algorithmic bytes per node-visit: slowness + traveltime read + traveltime write (FP64 pools for
# fp64 and mixed, FP32 pools for fp32)
BYTES_PER_VISIT = {"fp64": 24, "mixed": 24, "fp32": 12}
KERNELS = {"fp64": "k_fim_round<3,double,false,false,false>", "mixed": "k_fim_round<3,double,false,true,false>",
           "fp32": "k_fim_round<3,float,false,false,false>"}


def run_steps(model, src, recs, ep, steps, warmup):
    for _ in range(warmup):
        model.traveltime(src, recs, extraparams=ep)
    tot = {"_kernel_ms": 0.0, "tile_activations": 0.0, "rounds": 0.0, "node_updates": 0.0, "changed_updates": 0.0,
           "tile_sweeps": 0.0, "launches": 0.0, "fine_kernel_ms": 0.0, "fine_launches": 0.0,
           "fine_tile_activations": 0.0, "kernel_busy_ms": 0.0, "all_kernel_busy_ms": 0.0}
    t0 = time.perf_counter()
    for _ in range(steps):
        model.traveltime(src, recs, extraparams=ep)  # synchronous: returns after the device is idle
        st = model.ctx.last_stats()
        for k in tot:
            tot[k] += st[k]
    dt = time.perf_counter() - t0
    return dt, tot


def serving_leg(n_req: int, nslots: int) -> dict:
    """Serving mix (70 % forward / 30 % gradient, 2D 100^2-400^2 and 3D 64^3-160^3, 1-64 sources,
    16-128 receivers, FP32 stencil / FP64 adjoint) through the continuous-batching server: first
    all requests at once (capacity), then Poisson arrivals at 50 % and 80 % of that capacity.
    Latency = completion - arrival; models are resident before the clock starts."""
    import eikonalsolvers as eik
    from eikonalsolvers.grids import ExtraParams
    from eikonalsolvers.synthetic import c5_arrivals, c5_request

    ep = ExtraParams(precision="mixed")
    reqs = [c5_request(i) for i in range(n_req)]
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    nsolves = sum(len(r["src"]) for r in reqs)

    def run(arrivals):
        with eik.Server(nslots=nslots) as srv:
            t0 = time.perf_counter()
            handles = []
            for r, m, ta in zip(reqs, models, arrivals):
                dt = t0 + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                if r["kind"] == "traveltime":
                    handles.append(srv.submit_traveltime(m, r["src"], r["recs"], extraparams=ep))
                else:
                    handles.append(srv.submit_gradient(m, r["src"], r["recs"], r["pickobs"], r["stdobs"],
                                                       "gradvelandsrcloc", extraparams=ep))
            for h in handles:
                h.result()
            makespan = time.perf_counter() - t0
        lat = np.array([h.latency_ms for h in handles])
        # the slowest requests, with what their calls did (rounds, settling stage, shared launches)
        slow = []
        for i in np.argsort(-lat)[:3]:
            h, r = handles[i], reqs[i]
            slow.append({"kind": r["kind"], "grid": list(r["grd"].grsize), "nsrc": len(r["src"]),
                         "latency_ms": round(h.latency_ms, 1), "queue_ms": round(h.queue_ms, 1),
                         "run_ms": round(h.run_ms, 1), "rounds": h.stats["rounds"],
                         "settle_stage": h.stats["settle_stage"], "verify_rounds": h.stats["verify_rounds"],
                         "requests_in_call": h.stats["requests_in_call"]})
        return {"p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "makespan_s": round(makespan, 3), "source_solves_per_s": round(nsolves / makespan, 2),
                "requests_per_s": round(n_req / makespan, 3), "slowest": slow,
                "max_settle_stage": max(h.stats["settle_stage"] for h in handles)}

    run(np.zeros(n_req))  # warm-up: pools of every slot sized
    sat = run(np.zeros(n_req))
    cap = sat["requests_per_s"]
    out = {"mode": "f32 stencil / f64 near-source, adjoint and reductions", "slots": nslots, "requests": n_req,
           "gradient_requests": sum(r["kind"] == "gradient" for r in reqs), "source_solves": nsolves,
           "saturated": {"offered_rate": "all at t=0", **sat}, "poisson": []}
    for frac in (0.5, 0.8):
        res = run(c5_arrivals(n_req, frac * cap))
        out["poisson"].append({"offered_req_per_s": round(frac * cap, 3), "load": frac, **res})
    for m in models:
        m.close()
    return out


def _traffic_from_profile(prec: str):
    """HBM bytes per round-kernel launch of one precision, measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by the same counters on a plain copy) and committed under profiles/ as
    k_fim_round_3d_<prec>_traffic.json; None if absent."""
    p = os.path.join(ROOT, "profiles", f"k_fim_round_3d_{prec}_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def _roofline(prec: str, f: dict) -> dict:
    """Roofline object of the round kernel of one precision from its measured leg `f`."""
    launches = f["launches_per_step"]
    alg_per_launch = f["kernel_visits_per_step"] * BYTES_PER_VISIT[prec] / max(launches, 1.0)
    out = {
        "bound": "hbm",
        "kernel": KERNELS[prec],
        "achieved": round(f["per_launch_gbs"], 2) if f["per_launch_gbs"] else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(f["per_launch_gbs"] / HBM_PEAK_GBS, 5) if f["per_launch_gbs"] else None,
        "traffic": None,
        "traffic_scope": "HBM bytes per launch of this kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain "
                         "copy of known size)",
        "achieved_scope": "algorithmic bytes per launch / average launch duration (HIP events on the stream each "
                          "launch runs on); the call's sources are split over two streams whose launches overlap, "
                          "so achieved_device = algorithmic bytes of the step's launches / time in which at least "
                          "one of them ran",
        "achieved_device": round(f["achieved_gbs"], 2) if f["achieved_gbs"] else None,
        "frac_device": round(f["achieved_gbs"] / HBM_PEAK_GBS, 5) if f["achieved_gbs"] else None,
        "kernel_busy_ms_per_step": round(f["kernel_busy_ms_per_step"], 3),
        "bytes_per_node_visit": BYTES_PER_VISIT[prec],
        "algorithmic_bytes_per_launch": round(alg_per_launch),
        "launches_per_step": launches,
        "avg_launch_us": round(1e3 * f["kernel_all_ms_per_step"] / max(launches, 1.0), 2),
        "launch_scope": "every launch of this instantiation in a step (FP64: the refined boxes' solve too)",
        "node_visits_per_step": f["node_visits_per_step"],
        "visits_per_node": round(f["visits_per_node"], 3),
        "updates_per_node": round(f["updates_per_node"], 3),
        "changes_per_node": round(f["changes_per_node"], 3),
        "sweeps_per_activation": round(f["sweeps_per_activation"], 3),
        "rounds_per_step": f["rounds_per_step"],
        "kernel_ms_per_step": round(f["kernel_ms_per_step"], 3),
    }
    tr = _traffic_from_profile(prec)
    if tr is not None:
        out["traffic"] = tr["hbm_bytes_per_launch"]
        out["traffic_bytes_per_step"] = round(tr["hbm_bytes_per_launch"] * launches)
        out["traffic_over_algorithmic"] = round(tr["hbm_bytes_per_launch"] / max(alg_per_launch, 1.0), 3)
        out["traffic_measured_on"] = tr.get("measured_on", "")
    return out


def _host_cpu() -> dict:
    """The host CPU this process runs on: model name, logical CPUs, CPUs this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "cpus_available": avail}


def config_leg(name: str, steps: int) -> dict:
    """Forward (and gradient) timings of one of the other workloads, FP64 and mixed, through the
    public API: c1 = the README's 2D 300x220 case (4 sources), c3 = 2D 600x400 with 256 sources
    (:gradvel), c4 = 3D spherical 120x150x150 with 32 sources (:gradvel, no refinement)."""
    import eikonalsolvers as eik
    from eikonalsolvers.grids import ExtraParams
    from eikonalsolvers import synthetic as syn

    def timed(f):
        f()
        t0 = time.perf_counter()
        for _ in range(steps):
            f()
        return (time.perf_counter() - t0) / steps

    if name == "c1":
        grd, vel, src, recs = syn.c1_2d_forward()
        grad, extra = False, {}
    elif name == "c3":
        grd, vel, src, recs = syn.c3_2d_batched()
        grad, extra = True, {}
    else:
        grd, vel, src, recs = syn.c4_3d_sphere()
        grad, extra = True, {"refinearoundsrc": False}
    m = eik.EikModel(vel, grd)
    out = {"config": name, "grid": list(grd.grsize), "nsrc": len(src), "steps": steps}
    picks = m.traveltime(src, recs, extraparams=ExtraParams(**extra))
    obs = [p + 0.01 * np.cos(np.arange(len(p))) for p in picks]
    std = [np.full(len(p), 0.01) for p in picks]
    for prec in ("fp64", "mixed"):
        ep = ExtraParams(precision=prec, **extra)
        dt = timed(lambda: m.traveltime(src, recs, extraparams=ep))
        out[prec] = {"forward_ms": round(1e3 * dt, 2), "source_solves_per_s": round(len(src) / dt, 1)}
        if grad:
            dg = timed(lambda: m.gradient(src, recs, obs, std, "gradvel", extraparams=ep))
            out[prec].update({"gradvel_ms": round(1e3 * dg, 2), "evals_per_s": round(len(src) / dg, 1)})
    m.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0, help="sources timed on the CPU (0 = threads used)")
    ap.add_argument("--legs", default="all", choices=["all", "fp64", "mixed", "fp32"],
                    help="one forward leg only, printed alone (for counter-collection runs)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (request mix) leg")
    ap.add_argument("--serving-requests", type=int, default=32)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of 512 MiB (calibrates HBM byte counters)")
    ap.add_argument("--config", choices=["c1", "c3", "c4"], default=None,
                    help="print the timings of another workload instead of the benchmark line")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": x.numel(), "copy_probe_bytes_written": y.numel(), "launches": 5}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of it "
                                   "(the sources of a call would shard across GPUs by chunk)", "n_gpus": args.gpus}))
        return 2
    if args.config:
        print(json.dumps(config_leg(args.config, max(1, args.steps))), flush=True)
        return 0

    import eikonalsolvers as eik
    from eikonalsolvers.grids import ExtraParams
    from eikonalsolvers.synthetic import c2_3d, c2_observed

    grd, vel, src, recs, velobs = c2_3d()
    nsrc = src.shape[0]
    model = eik.EikModel(vel, grd)

    res = {}
    for prec in (("fp64", "mixed", "fp32", "fp64_fast") if args.legs == "all" else (args.legs,)):
        ep = ExtraParams(precision=prec.replace("_fast", ""), fast=prec.endswith("_fast"))
        dt, tot = run_steps(model, src, recs, ep, args.steps, args.warmup)
        kern_ms = tot["_kernel_ms"]
        visits = tot["tile_activations"] * TILE_NODES_3D
        nn = args.steps * nsrc * grd.npts
        # the round kernel as a profiler sees it: in FP64 the refined boxes around the sources are
        # solved by the same kernel instantiation as the coarse grid, so their launches count too
        k_ms, k_visits, k_launches, k_busy = kern_ms, visits, tot["launches"], tot["kernel_busy_ms"]
        if prec == "fp64":
            k_ms += tot["fine_kernel_ms"]
            k_visits += tot["fine_tile_activations"] * TILE_NODES_3D
            k_launches += tot["fine_launches"]
            k_busy = tot["all_kernel_busy_ms"]
        res[prec] = {
            "solves_per_s": nsrc * args.steps / dt,
            "ms_per_step": 1e3 * dt / args.steps,
            "kernel_ms_per_step": kern_ms / args.steps,
            "rounds_per_step": tot["rounds"] / args.steps,
            "launches_per_step": k_launches / args.steps,
            "kernel_visits_per_step": k_visits / args.steps,
            "kernel_all_ms_per_step": k_ms / args.steps,
            "kernel_busy_ms_per_step": k_busy / args.steps,
            "node_visits_per_step": visits / args.steps,
            "visits_per_node": visits / nn,
            "updates_per_node": tot["node_updates"] / nn,
            "changes_per_node": tot["changed_updates"] / nn,
            "sweeps_per_activation": tot["tile_sweeps"] / max(tot["tile_activations"], 1.0),
            "achieved_gbs": (k_visits * BYTES_PER_VISIT[prec.replace("_fast", "")] / 1e9) / (k_busy / 1e3)
            if k_busy > 0 else None,
            "per_launch_gbs": (k_visits * BYTES_PER_VISIT[prec.replace("_fast", "")] / 1e9) / (k_ms / 1e3)
            if k_ms > 0 else None,
        }
    if args.legs != "all":
        print(json.dumps({"legs": args.legs, **res[args.legs]}), flush=True)
        return 0

    # gradient evaluations: observed picks from the second model plus N(0, 0.01^2) noise
    obsmodel = eik.EikModel(velobs, grd)
    pobs, sobs = c2_observed(obsmodel.traveltime(src, recs, extraparams=ExtraParams()), nsrc)
    obsmodel.close()
    grad = {}
    for prec in ("fp64", "mixed", "fp32"):
        ep = ExtraParams(precision=prec)
        for _ in range(args.warmup):
            model.gradient(src, recs, pobs, sobs, "gradvelandsrcloc", extraparams=ep)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gv, gs, misf = model.gradient(src, recs, pobs, sobs, "gradvelandsrcloc", extraparams=ep)
        dt = time.perf_counter() - t0
        grad[prec] = {"evals_per_s": nsrc * args.steps / dt, "ms_per_call": 1e3 * dt / args.steps, "gv": gv, "gs": gs,
                      "misf": misf}
    g64, gmx, g32 = grad["fp64"], grad["mixed"], grad["fp32"]
    grad_rel_l2 = float(np.linalg.norm(gmx["gv"] - g64["gv"]) / np.linalg.norm(g64["gv"]))
    grad_rel_l2_32 = float(np.linalg.norm(g32["gv"] - g64["gv"]) / np.linalg.norm(g64["gv"]))

    out = {
        "metric": METRIC,
        "value": round(res["fp64"]["solves_per_s"], 3),
        "unit": "source-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp64"]["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "fp64_mode": "FP64 stencil replaying the heap FMM bit for bit (picks parity vs the oracle below)",
        "data": "synthetic (counter-based RNG, regenerated identically on every machine)",
        "config": {"workload": "3D Cartesian 201x201x101 h=1, GRF velocity, 16 sources, 64 top-face receivers, "
                               "refinearoundsrc=true (df=5, noderadius=3), eiktraveltime",
                   "nsrc": nsrc, "grid": list(grd.grsize), "nrec_per_src": 64},
        "mixed_precision": {"solves_per_s": round(res["mixed"]["solves_per_s"], 3),
                            "ms_per_step": round(res["mixed"]["ms_per_step"], 3),
                            "dtype": "f64 traveltimes, keys and comparisons; f32 quadratic increment (f64 stencil "
                                     "near the source when unrefined)",
                            "updates_per_node": round(res["mixed"]["updates_per_node"], 3),
                            "rounds_per_step": res["mixed"]["rounds_per_step"],
                            "kernel_ms_per_step": round(res["mixed"]["kernel_ms_per_step"], 3)},
        "fp32_mode": {"solves_per_s": round(res["fp32"]["solves_per_s"], 3),
                      "ms_per_step": round(res["fp32"]["ms_per_step"], 3),
                      "dtype": "f32 traveltimes and stencil (f64 refined box around the source)",
                      "updates_per_node": round(res["fp32"]["updates_per_node"], 3),
                      "rounds_per_step": res["fp32"]["rounds_per_step"],
                      "kernel_ms_per_step": round(res["fp32"]["kernel_ms_per_step"], 3)},
        "fp64_fast": {"solves_per_s": round(res["fp64_fast"]["solves_per_s"], 3),
                      "ms_per_step": round(res["fp64_fast"]["ms_per_step"], 3),
                      "note": "ExtraParams.fast=True: changes below 1e-12 relative are not propagated; within the "
                              "stated tolerances, not bit-identical",
                      "updates_per_node": round(res["fp64_fast"]["updates_per_node"], 3)},
        "gradient": {"evals_per_s_fp64": round(g64["evals_per_s"], 3), "ms_per_call_fp64": round(g64["ms_per_call"], 3),
                     "evals_per_s_mixed": round(gmx["evals_per_s"], 3), "ms_per_call_mixed": round(gmx["ms_per_call"], 3),
                     "evals_per_s_fp32": round(g32["evals_per_s"], 3), "ms_per_call_fp32": round(g32["ms_per_call"], 3),
                     "whichgrad": "gradvelandsrcloc", "misfit_fp64": g64["misf"],
                     "mixed_vs_fp64_gradvel_rel_l2": grad_rel_l2, "fp32_vs_fp64_gradvel_rel_l2": grad_rel_l2_32},
    }
    out["roofline"] = _roofline("fp64", res["fp64"])
    out["roofline_fp32"] = _roofline("fp32", res["fp32"])

    if not args.no_serving:
        out["serving"] = serving_leg(args.serving_requests, args.serving_slots)

    if not args.no_cpu_baseline:
        from oracle import oracle as O

        host = _host_cpu()
        # at most 16 threads: the host CPU share that goes with one GPU on the machines this runs on
        nthreads = max(1, min(host["cpus_available"], 16))
        ns = args.cpu_sample or nthreads
        ns = min(ns, nsrc)
        t0 = time.perf_counter()
        picks_cpu = O.traveltime(vel, grd, src[:ns], recs[:ns], extraparams=ExtraParams(), nthreads=nthreads)
        tcpu = time.perf_counter() - t0
        picks_gpu = model.traveltime(src[:ns], recs[:ns], extraparams=ExtraParams())
        rel = max(float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(picks_gpu, picks_cpu))
        out["cpu_baseline"] = {"value": round(ns / tcpu, 4), "unit": "source-solves/s", "cores": min(nthreads, ns),
                               "kind": "port", **host,
                               "sample": f"{ns} of the 16 sources of the same workload, heap FMM with refinement "
                                         f"(C restatement of the reference), OpenMP by source chunk, {tcpu:.2f} s",
                               "gpu_fp64_over_cpu": round(out["value"] / (ns / tcpu), 2)}
        # the gradient the same way: the restated adjoint (eikgradient, threaded by source chunk)
        t0 = time.perf_counter()
        _, _, misf_cpu = O.gradient(vel, grd, src[:ns], recs[:ns], pobs[:ns], sobs[:ns], "gradvelandsrcloc",
                                             extraparams=ExtraParams(), nthreads=nthreads)
        tg = time.perf_counter() - t0
        _, _, misf_gpu = model.gradient(src[:ns], recs[:ns], pobs[:ns], sobs[:ns], "gradvelandsrcloc",
                                                  extraparams=ExtraParams())
        out["cpu_baseline_gradient"] = {
            "value": round(ns / tg, 4), "unit": "gradient evals/s (source-evaluations of :gradvelandsrcloc)",
            "cores": min(nthreads, ns), "kind": "port", **host,
            "sample": f"{ns} of the 16 sources of the same workload, forward + adjoint with refinement (C restatement "
                      f"of eikgradient), OpenMP by source chunk, {tg:.2f} s",
            "gpu_fp64_over_cpu": round(g64["evals_per_s"] / (ns / tg), 2),
            "parity_misfit_rel_vs_oracle": float(abs(misf_gpu - misf_cpu) / abs(misf_cpu)),
            "note": "the reference's multi-source gradient keeps one stale buffer per worker across the sources "
                    "of its chunk (eikmaingrad.jl:112-133), so its :gradvel depends on the worker count; the GPU "
                    "follows the one-worker loop, which tests/test_gpu_c2.py compares with the oracle"}
        out["parity_fp64_picks_max_rel_vs_oracle"] = rel
        # the mixed form's error against the same oracle picks (north-star receiver tolerance 1e-5)
        picks_mx = model.traveltime(src[:ns], recs[:ns], extraparams=ExtraParams(precision="mixed"))
        out["mixed_precision"]["picks_max_rel_vs_oracle"] = max(
            float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(picks_mx, picks_cpu))
        picks_32 = model.traveltime(src[:ns], recs[:ns], extraparams=ExtraParams(precision="fp32"))
        out["fp32_mode"]["picks_max_rel_vs_oracle"] = max(
            float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(picks_32, picks_cpu))
        picks_fast = model.traveltime(src[:ns], recs[:ns], extraparams=ExtraParams(fast=True))
        out["fp64_fast"]["picks_max_rel_vs_oracle"] = max(
            float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(picks_fast, picks_cpu))
        # the same restatement on one core, two sources (the single-thread rate of the reference path)
        t0 = time.perf_counter()
        O.traveltime(vel, grd, src[:2], recs[:2], extraparams=ExtraParams(), nthreads=1)
        t1 = time.perf_counter() - t0
        out["cpu_baseline_1core"] = {"value": round(2 / t1, 4), "unit": "source-solves/s", "cores": 1, **host,
                                     "kind": "port", "sample": f"2 of the 16 sources, one thread, {t1:.2f} s",
                                     "gpu_fp64_over_1core": round(out["value"] / (2 / t1), 1)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
