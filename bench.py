#!/usr/bin/env python3
"""DOF-updates/s per RK stage of the CHORUS-MHD residual + SSP-RK time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ot512]

One "step" is one SSP(5,3) time step (five residual evaluations fused with
their stage updates, then the B := curl A_z correction) of the whole mesh,
whose state is resident in HBM before timing starts.  The timed steps are
chorus_time_steps: the kernels of chorus_step with HIP events around every
launch.  DOF = one solution
point (9 FP64 state values); value = NCELL*N^2 * 5 * K / wall time of the K
timed steps.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "chorus-mhd_amd"))

from chorus_mhd import QuadInp, quad_grid          # noqa: E402
from chorus_mhd import ic as ics                    # noqa: E402
from chorus_mhd.solver import Batch, Instance, Solver, topology   # noqa: E402
from chorus_mhd.diagnostics import magnetic_energy, point_volumes   # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes of one stage (all kernels of the stage), rocprofv3 FETCH_SIZE and WRITE_SIZE per
# dispatch scaled by kernels of known bytes (FETCH x 2.00, WRITE x 1.00); summaries under profiles/
MEASURED_STAGE_TRAFFIC = {("ot512", "fp64"): 8.07e9, ("ot512", "mixed"): 5.73e9,
                          ("ot512ideal", "fp64"): 2.76e9, ("ot512ideal", "mixed"): 2.47e9}
METRIC = "DOF-updates/s per RK stage"

# workloads (n, m, domain, N, jitter, seed, bc, ic, gamma, AV)
WORKLOADS = {
    "ot512": dict(desc="Orszag-Tang 512x512 quads, interior nodes jittered +-10% (seed 7), SD order 4 "
                       "(N=5), gamma=5/3, SSP(5,3), CFL 0.2",
                  n=512, m=512, box=(0.0, 2 * math.pi, 0.0, 2 * math.pi), N=5, jitter=0.10, seed=7,
                  bc="CYCL", ic="ot", gam=5.0 / 3, artiv=1),
    "ot64": dict(desc="Orszag-Tang 64x64 uniform quads, SD order 3 (N=4), gamma=5/3, SSP(5,3), CFL 0.2, AV off",
                 n=64, m=64, box=(0.0, 2 * math.pi, 0.0, 2 * math.pi), N=4, jitter=0.0, seed=0,
                 bc="CYCL", ic="ot", gam=5.0 / 3, artiv=0),
    "ot512ideal": dict(desc="Orszag-Tang 512x512 quads, interior nodes jittered +-10% (seed 7), SD order 4 "
                            "(N=5), gamma=5/3, SSP(5,3), CFL 0.2",
                       n=512, m=512, box=(0.0, 2 * math.pi, 0.0, 2 * math.pi), N=5, jitter=0.10, seed=7,
                       bc="CYCL", ic="ot", gam=5.0 / 3, artiv=0),
    "loop2048": dict(desc="field loop 2048x1024 jittered (seed 7), order 4 (N=5), u=(2,1), AV off",
                     n=2048, m=1024, box=(-1.0, 1.0, -0.5, 0.5), N=5, jitter=0.10, seed=7,
                     bc="CYCL", ic="loop", gam=5.0 / 3, artiv=0),
    "rotor1024": dict(desc="MHD rotor 1024x1024 quads on [0,1]^2, interior nodes jittered +-10% (seed 7), "
                           "FREE (outflow) boundaries, SD order 5 (N=6), gamma=1.4, SSP(5,3), CFL 0.2, AV on",
                      n=1024, m=1024, box=(0.0, 1.0, 0.0, 1.0), N=6, jitter=0.10, seed=7,
                      bc="FREE", ic="rotor", gam=1.4, artiv=1, steps=50),
    "loop_crossings": dict(desc="field loop 256x128 jittered (seed 7), order 4 (N=5), u=(2,1), AV off, "
                                "two domain crossings (t=2)",
                           n=256, m=128, box=(-1.0, 1.0, -0.5, 0.5), N=5, jitter=0.10, seed=7,
                           bc="CYCL", ic="loop", gam=5.0 / 3, artiv=0, t_end=2.0),
}


ENSEMBLE = dict(
    desc="batched ensemble: 256 independent Orszag-Tang / rotor instances (seed 7 draws the type, "
         "B0 ~ U[0.5,1.5] scaling the problem's field, interior node jitter ~ U[0,15%]), meshes uniform "
         "in {32^2,64^2,128^2,256^2}, SD orders uniform in {2..5} (N=3..6), AV on, SSP(5,3); "
         "instances of one order and problem type share one CFL-0.2 time step",
    ninst=256, seed=7)


def ensemble_members(ninst, seed):
    """Meshes, raw fields and per-group time steps of the ensemble workload."""
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(ninst):
        kind = "ot" if rng.random() < 0.5 else "rotor"
        b0 = float(rng.uniform(0.5, 1.5))
        jit = float(rng.uniform(0.0, 0.15))
        n = int(rng.choice([32, 64, 128, 256]))
        N = int(rng.integers(2, 6)) + 1
        spec.append((kind, b0, jit, n, N))
    raw, dts = [], {}
    for i, (kind, b0, jit, n, N) in enumerate(spec):
        if kind == "ot":
            mesh = quad_grid(n, n, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=jit, seed=1000 + i, bc="CYCL")
            gam = 5.0 / 3
            Q, p = ics.orszag_tang_fields(mesh, N, gam, b0)
        else:
            mesh = quad_grid(n, n, 0.0, 1.0, 0.0, 1.0, jitter=jit, seed=1000 + i, bc="FREE")
            gam = 1.4
            Q, p = ics.rotor_fields(mesh, N, gam, b0 * 5.0 / math.sqrt(4 * math.pi))
        dt = ics.cfl_dt(mesh, N, Q, gam, 0.2)
        key = (N, kind)
        dts[key] = min(dts.get(key, dt), dt)
        raw.append(dict(mesh=mesh, N=N, gam=gam, Q=Q, p=p, key=key, topo=topology(mesh)))
    for r in raw:
        r["dt"] = dts[r["key"]]
    return raw


def run_ensemble(args):
    w = ENSEMBLE
    steps = 10 if args.steps == 200 else args.steps
    t0 = time.perf_counter()
    raw = ensemble_members(w["ninst"], w["seed"])
    setup_s = time.perf_counter() - t0
    results, finals = {}, {}
    for precision in (["mixed"] if args.no_mixed else ["mixed", "fp64"]):
        members = [Instance(r["mesh"], r["N"], r["dt"], gam=r["gam"], artiv=1, precision=precision, topo=r["topo"])
                   for r in raw]
        b = Batch(members)
        for i, r in enumerate(raw):
            b.set_state(i, r["Q"])
        b.correct_b()
        for i, r in enumerate(raw):
            q = b.get_state(i)
            q[..., 4] = ics.energy(q, r["p"], r["gam"])
            b.set_state(i, q)
        t1 = time.perf_counter()
        tiles = b.autotune()
        tune_s = time.perf_counter() - t1
        if args.warmup:
            b.step(args.warmup)
        ms = b.time_steps(steps)
        dof = sum(m.mesh.ncell * m.N * m.N for m in members)
        finals[precision] = [b.get_state(i) for i in range(len(members))]
        results[precision] = dict(value=dof * 5 * steps / (ms * 1e-3), ms_per_step=ms / steps, dof=dof,
                                  groups=b.groups(), tiles=tiles, autotune_s=tune_s,
                                  divb_max=float(np.nanmax(b.divb())))
        b.close()
    r = results["mixed"]
    out = {
        "metric": METRIC, "value": r["value"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": w["desc"], "form": "mixed, batched", "instances": w["ninst"], "dof": r["dof"],
                   "groups": [dict(g, tile_shape=["256/128", "128/64", "512/256"][g["tile"]]) for g in r["groups"]]},
        "setup_s": setup_s, "autotune_s": r["autotune_s"], "divb_max": r["divb_max"],
    }
    if "fp64" in results:
        f = results["fp64"]
        rel = []
        for a, bq in zip(finals["fp64"], finals["mixed"]):
            if np.isfinite(a).all() and np.isfinite(bq).all():
                rel.append(float(np.linalg.norm(bq - a) / np.linalg.norm(a)))
        out["fp64_batched"] = {"value": f["value"], "ms_per_step": f["ms_per_step"], "divb_max": f["divb_max"]}
        out["speedup_vs_fp64"] = r["value"] / f["value"]
        out["rel_l2_vs_fp64_max"] = max(rel) if rel else None
        out["finite_instances"] = len(rel)
        out["finite_fp64"] = sum(bool(np.isfinite(a).all()) for a in finals["fp64"])
        out["finite_mixed"] = sum(bool(np.isfinite(a).all()) for a in finals["mixed"])
    print(json.dumps(out), flush=True)


def run_crossings(w, args):
    """Field loop to t_end in both forms: magnetic-energy decay, div B over time,
    and the mixed form's deviation from FP64 at t_end."""
    N = w["N"]
    mesh = quad_grid(w["n"], w["m"], *w["box"], jitter=w["jitter"], seed=w["seed"], bc=w["bc"])
    topo = topology(mesh)
    vol = point_volumes(mesh, N)
    s = Solver(mesh, N, 1e-3, gam=w["gam"], precision="fp64", topo=topo)
    Q0 = initial_state(w, mesh, s)
    s.close()
    dt0 = ics.cfl_dt(mesh, N, Q0, w["gam"], 0.2)
    nsteps = int(math.ceil(w["t_end"] / dt0))
    dt = w["t_end"] / nsteps
    e0 = magnetic_energy(mesh, N, Q0, vol)
    chunk = max(1, nsteps // 20)
    res, finals = {}, {}
    for precision in ("fp64", "mixed"):
        s = Solver(mesh, N, dt, gam=w["gam"], artiv=w["artiv"], precision=precision, topo=topo)
        s.set_state(Q0)
        done, ms, hist = 0, 0.0, []
        while done < nsteps:
            k = min(chunk, nsteps - done)
            ms += s.time_steps(k)["ms_total"]
            done += k
            d = s.diag()
            Q = s.get_state()
            hist.append((done * dt, magnetic_energy(mesh, N, Q, vol) / e0, d["divbmax"], d["divbglob"]))
        finals[precision] = s.get_state()
        s.close()
        res[precision] = dict(value=mesh.ncell * N * N * 5 * nsteps / (ms * 1e-3), ms_per_step=ms / nsteps,
                              energy_ratio=hist[-1][1], divb_max=max(h[2] for h in hist),
                              divb_rms_end=hist[-1][3], history=[[round(h[0], 6), h[1], h[2]] for h in hist])
    a, b = finals["fp64"], finals["mixed"]
    out = {
        "metric": METRIC, "value": res["fp64"]["value"], "unit": "DOF-updates/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 0, "ms_per_step": res["fp64"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"], "ncell": mesh.ncell, "N": N, "dt": dt, "t_end": w["t_end"]},
        "fp64": res["fp64"], "mixed": res["mixed"],
        "mixed_rel_l2_vs_fp64": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
        "mixed_energy_ratio_minus_fp64": res["mixed"]["energy_ratio"] - res["fp64"]["energy_ratio"],
    }
    print(json.dumps(out), flush=True)


def run_output(w, args):
    """Output path at the workload (FP64 drop-in state after one step): the
    restart dump (chorus_write_restart) and the Tecplot file of NRE x NRE plot
    cells per cell (chorus_write_tecplot), each the mean wall time of a few
    calls including the file write, beside the host route the driver would
    otherwise take for the dump (download + transpose + host pack and write)."""
    from chorus_mhd import write_restart
    mesh, topo, s, artiv = setup(w, "fp64", True)
    Q0 = initial_state(w, mesh, s)
    s.close()
    N, nre, reps = w["N"], args.nre, 3
    s = Solver(mesh, N, 1e-4, gam=w["gam"], artiv=artiv, topo=topo)
    t0 = time.perf_counter()
    s.plot_setup(nre)
    t_setup = time.perf_counter() - t0
    s.set_state(Q0)
    s.step(1)
    res = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out")
        for name, fn in (("restart", lambda: s.write_restart(path, 1, 1e-4)),
                         ("tecplot", lambda: s.write_tecplot(path, 1e-4)),
                         ("restart_host_route", lambda: write_restart(path, 1, 1e-4, s.get_state()))):
            fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            res[name] = ((time.perf_counter() - t0) / reps, os.path.getsize(path))
            os.remove(path)
    s.close()
    nnode = (w["n"] * nre + 1) * (w["m"] * nre + 1) if w["bc"] == "CYCL" else None
    base = None if args.no_cpu else cpu_output_baseline(w, nre)
    out = {
        "metric": "output wall time per dump", "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": w["desc"] + (", AV on" if artiv else ", AV off"), "ncell": mesh.ncell, "N": N,
                   "nre": nre, "plot_nodes": nnode, "calls_timed": reps},
        "restart_s": res["restart"][0], "restart_bytes": res["restart"][1],
        "restart_host_route_s": res["restart_host_route"][0],
        "tecplot_s": res["tecplot"][0], "tecplot_bytes": res["tecplot"][1],
        "plot_setup_s": t_setup,
        "tecplot_nodes_per_s": nnode / res["tecplot"][0] if nnode else None,
    }
    if base:
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


def cpu_output_baseline(w, nre, n=128):
    """The reference's own Tecplot chain (tecplotter2dsetup, then tecplotter2d +
    tecpost_parallel) timed by the compiled reference harness on one host core,
    on an n x n sample of the workload after one step; plot nodes per second of
    the output call."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    if not os.path.exists(orc.REF_HARNESS) or abs(w["gam"] - 5.0 / 3) > 1e-15 or w["bc"] != "CYCL":
        return None
    mesh = quad_grid(n, n, *w["box"], jitter=w["jitter"], seed=w["seed"], bc=w["bc"])
    with tempfile.TemporaryDirectory() as d:
        mesh.write(d)
        os.makedirs(os.path.join(d, "TEC"))
        QuadInp(N=w["N"], dt=1e-4, MAXITER=1, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl, ARTIV=1, NRE=nre,
                restart_ord=w["N"]).write(os.path.join(d, "QUAD.INP"))
        r = subprocess.run([orc.REF_HARNESS, "plot"], cwd=d, capture_output=True, text=True, timeout=600,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        if r.returncode != 0:
            return None
        setup_s, out_s = (float(x) for x in open(os.path.join(d, "plot_timing.txt")).read().split())
    nodes = (n * nre + 1) ** 2
    return dict(value=nodes / out_s, unit="plot nodes/s", cores=1, kind="reference", output_s=out_s,
                setup_s=setup_s, sample="%dx%d %s mesh (jitter %.2f, seed %d), N=%d, NRE=%d, AV on, after one step "
                "(reference Fortran tecplotter2d + tecpost_parallel, flang -O3)" % (
                    n, n, w["ic"], w["jitter"], w["seed"], w["N"], nre))


def b_stage(N: int) -> float:
    """Algorithmic HBM bytes per DOF per stage: FP64 state + RK registers + metrics
    (224 + 32(N+1)/N)."""
    return 224.0 + 32.0 * (N + 1) / N


def setup(w, precision, av_supported):
    mesh = quad_grid(w["n"], w["m"], *w["box"], jitter=w["jitter"], seed=w["seed"], bc=w["bc"])
    topo = topology(mesh)
    artiv = w["artiv"] if av_supported else 0
    s = Solver(mesh, w["N"], 1e-3, gam=w["gam"], artiv=artiv, precision=precision, topo=topo)
    return mesh, topo, s, artiv


def initial_state(w, mesh, s):
    if w["ic"] == "ot":
        return ics.orszag_tang(s, mesh, w["N"], w["gam"])
    if w["ic"] == "loop":
        return ics.field_loop(s, mesh, w["N"], w["gam"])
    return ics.rotor(s, mesh, w["N"], w["gam"])


def cpu_baseline(w, artiv, budget_s=30.0):
    """The compiled reference (oracle/_ref/ref_harness) on a bounded sample of
    the workload on one host core; falls back to the C oracle port."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    n = 96 if w["N"] >= 5 else 128
    steps = 2
    N = w["N"]
    sample = "%dx%d %s mesh (jitter %.2f, seed %d), N=%d, AV %s, %d SSP(5,3) steps" % (
        n, n, w["ic"], w["jitter"], w["seed"], N, "on" if artiv else "off", steps)
    mesh = quad_grid(n, n, *w["box"], jitter=w["jitter"], seed=w["seed"], bc=w["bc"])
    dofs = mesh.ncell * N * N * 5 * steps
    if os.path.exists(orc.REF_HARNESS) and abs(w["gam"] - 5.0 / 3) < 1e-15:
        with tempfile.TemporaryDirectory() as d:
            mesh.write(d)
            QuadInp(N=N, dt=1e-4, MAXITER=steps, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl,
                    ARTIV=artiv, restart_ord=N).write(os.path.join(d, "QUAD.INP"))
            r = subprocess.run([orc.REF_HARNESS, "run"], cwd=d, capture_output=True, text=True,
                               timeout=600, env=dict(os.environ, OMP_NUM_THREADS="1"))
            if r.returncode == 0:
                t = open(os.path.join(d, "timing.txt")).read().split()
                sec = float(t[1])
                return dict(value=dofs / sec, unit="DOF-updates/s", cores=1, kind="reference",
                            sample=sample + " (reference Fortran, flang -O3, time loop only)")
    orc.build()
    o = orc.Oracle(mesh, N, dt=1e-4, artiv=artiv, gam=w["gam"])
    o.init_ot()
    t0 = time.perf_counter()
    o.steps(steps)
    sec = time.perf_counter() - t0
    return dict(value=dofs / sec, unit="DOF-updates/s", cores=1, kind="port",
                sample=sample + " (C oracle port)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ot512", choices=sorted(WORKLOADS) + ["ensemble"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision form")
    ap.add_argument("--tile", type=int, default=None,
                    help="fixed tile shape of the mixed form (0: 256/128, 1: 128/64, 2: 512/256) instead of "
                         "autotuning, e.g. for profiles without the tuner's launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run two kernels of known HBM bytes (a 256 MiB fill: 256 MiB written; an "
                         "in-place add of two 256 MiB arrays: 512 MiB read, 256 MiB written) so that "
                         "hardware byte counters collected over this run can be scaled")
    ap.add_argument("--output", action="store_true",
                    help="time the output path (restart dump, Tecplot file) at the workload instead of steps")
    ap.add_argument("--nre", type=int, default=4, help="plot cells per cell edge of the Tecplot file (--output)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build holds the whole mesh as one "
              "partition on one MI355X (multi-GPU halo exchange is not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        run_ensemble(args)
        return
    w = WORKLOADS[args.workload]
    if args.output:
        run_output(w, args)
        return
    if "t_end" in w:
        run_crossings(w, args)
        return
    if "steps" in w and args.steps == 200:
        args.steps = w["steps"]
    av_supported = True
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.full_like(src, 2.0)
        dst.add_(src)
        torch.cuda.synchronize()
        del src, dst
    N = w["N"]

    def run(precision, Q0=None):
        mesh, topo, s, artiv = setup(w, precision, av_supported)
        if Q0 is None:
            Q0 = initial_state(w, mesh, s)
        else:
            s.set_state(Q0)
        dt = ics.cfl_dt(mesh, N, Q0, w["gam"], 0.2)
        s.close()
        s = Solver(mesh, N, dt, gam=w["gam"], artiv=artiv, precision=precision, topo=topo)
        s.set_state(Q0)
        tile = None
        if precision == "mixed":
            if args.tile is None:
                tile = s.autotune()
            else:
                s.set_tile(args.tile)
                tile = args.tile
        if args.warmup:
            s.step(args.warmup)
            s.set_state(Q0)
        # the timed steps carry HIP events around every stage and CT launch
        # (chorus_time_steps), which give the per-launch kernel times
        t0 = time.perf_counter()
        tm = s.time_steps(args.steps)
        wall = time.perf_counter() - t0
        Q = s.get_state()
        d = s.diag()
        s.close()
        return dict(mesh=mesh, dt=dt, wall=wall, tm=tm, Q=Q, Q0=Q0, diag=d, artiv=artiv, tile=tile)

    r64 = run("fp64")
    mesh = r64["mesh"]
    ndof = mesh.ncell * N * N
    value = ndof * 5 * args.steps / r64["wall"]
    stage_s = r64["tm"]["ms_stage_avg"] * 1e-3
    achieved = ndof * b_stage(N) / stage_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w["desc"] + (", AV on" if r64["artiv"] else ", AV off"),
            "form": "FP64 drop-in",
            "ncell": mesh.ncell, "N": N, "dof": ndof, "dt": r64["dt"],
            "av_requested": bool(w["artiv"]), "av_applied": bool(r64["artiv"]),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_STAGE_TRAFFIC.get((args.workload, "fp64")),
            "kernel": ("one stage: k_face_inv + k_grad + k_face_visc + k_stage_visc <%d,double>" % N)
                      if r64["artiv"] else "one stage: k_face_inv + k_stage_ideal <%d,double>" % N,
            "stage_kernel_ms": r64["tm"]["ms_stage_avg"], "ct_kernel_ms": r64["tm"]["ms_ct_avg"],
            "bytes_per_dof_stage": b_stage(N), "algorithmic_bytes_per_stage": ndof * b_stage(N),
        },
        "divb": {"rms": r64["diag"]["divbglob"], "max": r64["diag"]["divbmax"]},
    }
    if not args.no_mixed:
        rm = run("mixed", Q0=r64["Q0"])
        vm = ndof * 5 * args.steps / rm["wall"]
        a, b = r64["Q"], rm["Q"]
        out["mixed"] = {
            "value": vm, "speedup_vs_fp64": vm / value,
            "rel_l2_vs_fp64": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
            "linf_vs_fp64": float(np.abs(b - a).max()),
            "stage_kernel_ms": rm["tm"]["ms_stage_avg"],
            "roofline_frac": ndof * b_stage(N) / (rm["tm"]["ms_stage_avg"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": MEASURED_STAGE_TRAFFIC.get((args.workload, "mixed")),
            "divb_max": rm["diag"]["divbmax"],
            "tile_shape": ["256/128", "128/64", "512/256"][rm["tile"]],
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, r64["artiv"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
