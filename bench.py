#!/usr/bin/env python3
"""DOF-updates/s per RK stage of the CHORUS-MHD residual + SSP-RK time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ensemble|ot512|...]

Default workload: the batched ensemble of 256 independent Orszag-Tang / rotor
instances (N = 3..6, meshes 32^2..256^2, AV on), the largest configuration that
fits one GPU, in three forms: the mixed-precision batched form (the headline
value), the FP64 batched form, and the FP64 drop-in run one instance at a time.
One "step" is one SSP(5,3) time step (five residual evaluations fused with
their stage updates, then the B := curl A_z correction) of every instance,
whose states are resident in HBM before timing starts; the timed steps are the
product path (chorus_batch_step / chorus_step: one hipGraph launch per step).
DOF = one solution point (9 FP64 state values); value = sum over instances of
NCELL*N^2 * 5 * K / wall time of the K timed steps.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "chorus-mhd_amd"))

from chorus_mhd import QuadInp, quad_grid, write_restart   # noqa: E402
from chorus_mhd import ic as ics                    # noqa: E402
from chorus_mhd.solver import Batch, Instance, Solver, topology   # noqa: E402
from chorus_mhd.diagnostics import magnetic_energy, point_volumes   # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "DOF-updates/s per RK stage"
# host cores of one GPU's share of the GPU box (the CPU baseline's process count)
CPU_SHARE = 16
PROFILES = os.path.join(REPO, "profiles")
# tile index -> (stage shape, gradient shape) of the mixed cell kernels; a shape
# is the threads-per-block target of the ideal / viscous kernels
_SHAPES = ["256/128", "128/64", "512/256"]
_TILE_PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)]
TILES = [_SHAPES[s] if s == g else "%s, gradient %s" % (_SHAPES[s], _SHAPES[g].split("/")[1])
         for s, g in _TILE_PAIRS]


def measured_traffic(key: str):
    """HBM bytes of one unit of work (a stage or a batched step) from the
    calibrated rocprofv3 counter summaries under profiles/: the line
    'traffic[<key>] = <bytes> bytes' of the summary that measured it."""
    pat = re.compile(r"traffic\[%s\] = ([0-9.eE+]+) bytes" % re.escape(key))
    for f in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if f.endswith(".md"):
            m = pat.search(open(os.path.join(PROFILES, f)).read())
            if m:
                return float(m.group(1))
    return None


def host_info() -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(cpu_model=model, nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)))


def b_stage(N: int) -> float:
    """Algorithmic HBM bytes per DOF per stage: FP64 state + RK registers + metrics
    (224 + 32(N+1)/N)."""
    return 224.0 + 32.0 * (N + 1) / N


def b_ct(N: int) -> float:
    """Algorithmic HBM bytes per DOF of the CT correction: A_z read, Bx and By
    written, metrics (32 + 32(N+1)/N)."""
    return 32.0 + 32.0 * (N + 1) / N


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
         "instances of one order and problem type share one time step (Orszag-Tang CFL 0.2, "
         "rotor CFL 0.1)",
    ninst=256, seed=7, cfl={"ot": 0.2, "rotor": 0.1})


def _progress(msg):
    """One line per finished phase on stderr (the JSON line stays alone on stdout)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ensemble_members(ninst, seed, cfl=None):
    """Meshes, raw fields and per-group time steps of the ensemble workload."""
    cfl = cfl or ENSEMBLE["cfl"]
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
        dt = ics.cfl_dt(mesh, N, Q, gam, cfl[kind])
        key = (N, kind)
        dts[key] = min(dts.get(key, dt), dt)
        raw.append(dict(mesh=mesh, N=N, gam=gam, Q=Q, p=p, key=key, kind=kind, n=n, topo=topology(mesh)))
    for r in raw:
        r["dt"] = dts[r["key"]]
    return raw


def _members(raw, precision):
    return [Instance(r["mesh"], r["N"], r["dt"], gam=r["gam"], artiv=1, precision=precision, topo=r["topo"])
            for r in raw]


def _initial_states(raw):
    """B := curl A_z of every instance on the device (FP64 in both forms), then
    the energy from the corrected B: the states every form starts from."""
    b = Batch(_members(raw, "fp64"))
    for i, r in enumerate(raw):
        b.set_state(i, r["Q"])
    b.correct_b()
    q0 = []
    for i, r in enumerate(raw):
        q = b.get_state(i)
        q[..., 4] = ics.energy(q, r["p"], r["gam"])
        q0.append(q)
        r["Q"] = None          # the raw fields are no longer needed
    b.close()
    return q0


def _timed_batch(raw, q0, precision, args):
    b = Batch(_members(raw, precision))
    for i, q in enumerate(q0):
        b.set_state(i, q)
    tiles, tune_s = None, 0.0
    if precision == "mixed":
        t0 = time.perf_counter()
        tiles = b.autotune()
        tune_s = time.perf_counter() - t0
    if args.warmup:
        b.step(args.warmup)           # captures the step graphs of every buffer rotation
    t0 = time.perf_counter()
    b.step(args.steps)                # synchronous: returns when the last step has finished
    wall = time.perf_counter() - t0
    states = [b.get_state(i) for i in range(len(raw))]
    divb = b.divb()
    groups = b.groups()
    b.close()
    return dict(wall=wall, states=states, divb=divb, groups=groups, tiles=tiles, autotune_s=tune_s)


def _dropin_sequential(raw, q0, args):
    """The FP64 drop-in as a user of chorus_step runs the ensemble: one context
    per instance, one after the other, K timed steps each (after W warmup
    steps that capture its step graphs)."""
    wall = 0.0
    for i, r in enumerate(raw):
        s = Solver(r["mesh"], r["N"], r["dt"], gam=r["gam"], artiv=1, precision="fp64", topo=r["topo"])
        s.set_state(q0[i])
        if args.warmup:
            s.step(args.warmup)
        t0 = time.perf_counter()
        s.step(args.steps)
        wall += time.perf_counter() - t0
        s.close()
    return wall


def _write_instance(d, mesh, N, dt, steps, artiv, q0=None, mu=0.0, eta0=0.0, metis=None):
    mesh.write(d)
    q = QuadInp(N=N, dt=dt, MAXITER=steps, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl, ARTIV=artiv,
                restart_ord=N, mu=mu, eta0=eta0)
    if q0 is not None:
        write_restart(os.path.join(d, "restart.in"), 0, 0.0, q0)
        q.restart = 1
        q.NAMERESTART = "restart.in"
    if metis is not None:
        with open(os.path.join(d, "metis.dat"), "w") as f:
            f.write("\n".join(str(int(p)) for p in metis) + "\n")
        q.METIS_CELL = "metis.dat"
    q.write(os.path.join(d, "QUAD.INP"))


def _harness(gam):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    exe = orc.REF_HARNESS if gam == 5.0 / 3 else (orc.REF_HARNESS_G14 if gam == 1.4 else None)
    return exe if exe and os.path.exists(exe) else None


def cpu_ensemble_baseline(raw, q0, P=CPU_SHARE, steps=1):
    """The compiled reference (oracle/_ref: unmodified sources; the gamma = 1.4
    build for rotors) on the first P instances of the ensemble, P single-rank
    processes at once, one per host core of the GPU's share: the aggregate
    DOF-updates/s of the P concurrent runs (sum of each run's DOF x stages over
    its own time-loop seconds).  Each instance starts from the GPU runs'
    initial state (restart file).  The reference loop evaluates a diagnostic
    flux and gradient every iteration (iterations.f90:35-36); the same runs
    without it give the overhead."""
    P = min(P, len(raw), os.cpu_count() or 1)
    sample = list(range(P))
    if any(_harness(raw[i]["gam"]) is None for i in sample):
        return None
    res = {}
    with tempfile.TemporaryDirectory() as top:
        dirs = []
        for i in sample:
            d = os.path.join(top, "i%03d" % i)
            os.makedirs(d)
            r = raw[i]
            _write_instance(d, r["mesh"], r["N"], r["dt"], steps, 1, q0=q0[i])
            dirs.append(d)
        for mode in ("run", "runnd"):
            env = dict(os.environ, OMP_NUM_THREADS="1")
            procs = [subprocess.Popen([_harness(raw[i]["gam"]), mode], cwd=d, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL, env=env) for i, d in zip(sample, dirs)]
            t0 = time.perf_counter()
            rc = [p.wait(timeout=900) for p in procs]
            wall = time.perf_counter() - t0
            if any(rc):
                return None
            secs = [float(open(os.path.join(d, "timing.txt")).read().split()[1]) for d in dirs]
            dofs = [raw[i]["mesh"].ncell * raw[i]["N"] ** 2 * 5 * steps for i in sample]
            res[mode] = dict(value=sum(n / s for n, s in zip(dofs, secs)), loop_s=sum(secs), wall=wall)
    h = host_info()
    return dict(value=res["run"]["value"], unit="DOF-updates/s", cores=P, kind="reference",
                cpu_model=h["cpu_model"], nproc=h["nproc"], processes=P,
                value_without_diag_flux=res["runnd"]["value"],
                diag_flux_overhead=res["run"]["loop_s"] / res["runnd"]["loop_s"] - 1.0,
                sample="the first %d ensemble instances (%s), %d SSP(5,3) step each, as %d concurrent "
                       "single-rank runs of the reference Fortran (flang -O3, unmodified sources; gamma=1.4 "
                       "build for rotors), time loop only (setup excluded), including the diagnostic flux "
                       "the reference loop evaluates every iteration (iterations.f90:35-36)" % (
                           P, ", ".join("%s %d^2 N=%d" % (raw[i]["kind"], raw[i]["n"], raw[i]["N"]) for i in sample),
                           steps, P))


def run_ensemble(args):
    w = ENSEMBLE
    t0 = time.perf_counter()
    raw = ensemble_members(w["ninst"], w["seed"])
    q0 = _initial_states(raw)
    setup_s = time.perf_counter() - t0
    _progress("ensemble set up: %d instances, %.1f s" % (len(raw), setup_s))
    ndof = np.array([r["mesh"].ncell * r["N"] ** 2 for r in raw], dtype=np.float64)
    res = {}
    for precision in ("mixed", "fp64"):
        res[precision] = _timed_batch(raw, q0, precision, args)
        _progress("%s batched: %d steps in %.3f s" % (precision, args.steps, res[precision]["wall"]))
    finite = np.array([bool(np.isfinite(a).all() and np.isfinite(b).all())
                       for a, b in zip(res["fp64"]["states"], res["mixed"]["states"])])
    rel = [float(np.linalg.norm(b - a) / np.linalg.norm(a))
           for a, b, f in zip(res["fp64"]["states"], res["mixed"]["states"], finite) if f]
    dof_all, dof_fin = float(ndof.sum()), float(ndof[finite].sum())
    K = args.steps
    mx, f64 = res["mixed"], res["fp64"]
    value = dof_fin * 5 * K / mx["wall"]
    ms_step = mx["wall"] * 1e3 / K
    step_bytes = float(sum(n * (5 * b_stage(r["N"]) + b_ct(r["N"])) for n, r in zip(ndof, raw)))
    achieved = step_bytes / (ms_step * 1e-3) / 1e9
    f64_ms = f64["wall"] * 1e3 / K
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": w["desc"], "form": "mixed precision (FP32 flux arithmetic, FP64 state, RK, CT), "
                   "batched: one hipGraph launch per step", "instances": w["ninst"],
                   "dof": dof_all, "dof_finite": dof_fin, "finite_instances": int(finite.sum()),
                   "groups": [dict(g, tile_shape=TILES[g["tile"]]) for g in mx["groups"]]},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("ensemble256/mixed/step"),
            "kernel": "one batched time step (one hipGraph launch: 8 groups' stage kernels x 5 stages + CT, "
                      "groups concurrent); dominant kernels k_stage_visc_mx<N=4..6> and k_grad_mx",
            "algorithmic_bytes_per_step": step_bytes,
            "bytes_per_dof": "5 x (224 + 32(N+1)/N) + 32 + 32(N+1)/N",
        },
        "fp64_batched": {"value": dof_fin * 5 * K / f64["wall"], "ms_per_step": f64_ms,
                         "roofline_frac": step_bytes / (f64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": measured_traffic("ensemble256/fp64/step")},
        "speedup_vs_fp64_batched": f64["wall"] / mx["wall"],
        "mixed_rel_l2_vs_fp64_max": max(rel) if rel else None,
        "mixed_rel_l2_vs_fp64_median": float(np.median(rel)) if rel else None,
        "divb_max": {"fp64_finite": float(np.max(f64["divb"][finite])),
                     "mixed_finite": float(np.max(mx["divb"][finite])),
                     "fp64_ot": float(np.max([d for d, r, f in zip(f64["divb"], raw, finite) if f and r["kind"] == "ot"])),
                     "fp64_rotor": float(np.max([d for d, r, f in zip(f64["divb"], raw, finite) if f and r["kind"] == "rotor"]))},
        "setup_s": setup_s, "autotune_s": mx["autotune_s"],
    }
    if not args.no_dropin:
        t = _dropin_sequential(raw, q0, args)
        _progress("FP64 drop-in one instance at a time: %.3f s" % t)
        out["fp64_dropin_sequential"] = {"value": dof_fin * 5 * K / t, "ms_per_step": t * 1e3 / K}
        out["speedup_vs_fp64_dropin"] = t / mx["wall"]
    if not args.no_cpu:
        _progress("CPU baseline (reference Fortran on the host cores)")
        out["cpu_baseline"] = cpu_ensemble_baseline(raw, q0)
    if args.divb_report:
        out["divb_per_instance"] = [dict(kind=r["kind"], n=r["n"], N=r["N"], finite=bool(f),
                                         fp64=float(a), mixed=float(b))
                                    for r, f, a, b in zip(raw, finite, f64["divb"], mx["divb"])]
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


def cpu_reference_mpi(w, artiv, P=CPU_SHARE, steps=2, max_dof=8.0e6):
    """The compiled reference (oracle/_ref: unmodified sources; the gamma = 1.4
    build for the rotor) with P MPI ranks (mpirun, a strip partition of the
    cell rows as METIS_CELL, readGRIDinfo.f90:100-107), one per host core of
    the GPU's share: the full mesh when it has at most max_dof points, else
    the same workload on a coarser mesh of about that size.  Orszag-Tang starts
    from the reference's own initial condition, the others from the same
    initial-condition routine as the GPU run (restart file).  The timed part
    is the reference's time loop (barrier to barrier), once as the reference
    runs it (with the diagnostic flux of iterations.f90:35-36 every
    iteration) and once without that flux."""
    exe = _harness(w["gam"])
    mpirun = "/opt/conda/bin/mpirun"
    if exe is None or not os.path.exists(mpirun):
        return None
    P = min(P, os.cpu_count() or 1)
    N = w["N"]
    n, m = w["n"], w["m"]
    scale = 1.0
    while n * m * N * N * scale * scale > max_dof:
        scale *= 0.5
    n, m = max(8, int(round(n * scale))), max(8, int(round(m * scale)))
    mesh = quad_grid(n, m, *w["box"], jitter=w["jitter"], seed=w["seed"], bc=w["bc"])
    q0 = None
    if w["ic"] != "ot":
        s = Solver(mesh, N, 1e-3, gam=w["gam"], artiv=artiv, topo=topology(mesh))
        q0 = initial_state(w, mesh, s)
        s.close()
        dt = ics.cfl_dt(mesh, N, q0, w["gam"], 0.2)
    else:
        dt = ics.cfl_dt(mesh, N, ics.orszag_tang_fields(mesh, N, w["gam"])[0], w["gam"], 0.2)
    part = ((np.arange(mesh.ncell) // n) * P) // m          # strips of cell rows, rank 0..P-1
    res = {}
    with tempfile.TemporaryDirectory() as d:
        _write_instance(d, mesh, N, dt, steps, artiv, q0=q0, metis=part if P > 1 else None)
        for mode in ("run", "runnd"):
            r = subprocess.run([mpirun, "-np", str(P), exe, mode], cwd=d, capture_output=True, text=True,
                               timeout=900, env=dict(os.environ, OMP_NUM_THREADS="1"))
            if r.returncode != 0:
                return None
            t = open(os.path.join(d, "timing.txt")).read().split()
            res[mode] = float(t[1])
    dofs = mesh.ncell * N * N * 5 * steps
    h = host_info()
    return dict(value=dofs / res["run"], unit="DOF-updates/s", cores=P, kind="reference",
                cpu_model=h["cpu_model"], nproc=h["nproc"], mpi_ranks=P,
                value_without_diag_flux=dofs / res["runnd"], diag_flux_overhead=res["run"] / res["runnd"] - 1.0,
                sample="%dx%d %s mesh (jitter %.2f, seed %d), N=%d, AV %s, %d SSP(5,3) steps, %d MPI ranks "
                       "(strip partition), reference Fortran (flang -O3, unmodified sources), time loop only, "
                       "including the diagnostic flux the reference loop evaluates every iteration "
                       "(iterations.f90:35-36)" % (n, m, w["ic"], w["jitter"], w["seed"], N,
                                                    "on" if artiv else "off", steps, P))


def run_single(w, args):
    """One instance (config 1, 2, 3 or 5): the FP64 drop-in, then the mixed
    form from the same initial state."""
    if "steps" in w and args.steps == 200:
        args.steps = w["steps"]
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
        dst = torch.full_like(src, 2.0)
        dst.add_(src)
        torch.cuda.synchronize()
        del src, dst
    N = w["N"]
    K = args.steps

    def run(precision, Q0=None):
        mesh, topo, s, artiv = setup(w, precision, True)
        if Q0 is None:
            Q0 = initial_state(w, mesh, s)
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
            s.step(args.warmup)         # captures the step graphs of every buffer rotation
            s.set_state(Q0)
        # the timed steps: chorus_step, one hipGraph launch per step (the product path)
        t0 = time.perf_counter()
        s.step(K)
        wall = time.perf_counter() - t0
        Q = s.get_state()
        d = s.diag()
        # a separate pass with HIP events around every stage and CT launch
        # (chorus_time_steps, plain launches) gives the per-launch kernel times
        tm = None
        if not args.no_event_pass:
            s.set_state(Q0)
            tm = s.time_steps(min(K, 20))
        s.close()
        return dict(mesh=mesh, dt=dt, wall=wall, tm=tm, Q=Q, Q0=Q0, diag=d, artiv=artiv, tile=tile)

    r64 = run("fp64")
    _progress("FP64 drop-in: %d steps in %.3f s" % (K, r64["wall"]))
    mesh = r64["mesh"]
    ndof = mesh.ncell * N * N
    value = ndof * 5 * K / r64["wall"]
    key = args.workload
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": r64["wall"] * 1e3 / K, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": w["desc"] + (", AV on" if r64["artiv"] else ", AV off"),
            "form": "FP64 drop-in (chorus_step: one hipGraph launch per step)",
            "ncell": mesh.ncell, "N": N, "dof": ndof, "dt": r64["dt"],
        },
        "divb": {"rms": r64["diag"]["divbglob"], "max": r64["diag"]["divbmax"]},
    }
    if r64["tm"]:
        stage_s = r64["tm"]["ms_stage_avg"] * 1e-3
        achieved = ndof * b_stage(N) / stage_s / 1e9
        out["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("%s/fp64/stage" % key),
            "kernel": ("one stage: k_face_inv + k_grad + k_face_visc + k_stage_visc <%d,double>" % N)
                      if r64["artiv"] else "one stage: k_face_inv + k_stage_ideal <%d,double>" % N,
            "stage_kernel_ms": r64["tm"]["ms_stage_avg"], "ct_kernel_ms": r64["tm"]["ms_ct_avg"],
            "bytes_per_dof_stage": b_stage(N), "algorithmic_bytes_per_stage": ndof * b_stage(N),
        }
    if not args.no_mixed:
        rm = run("mixed", Q0=r64["Q0"])
        _progress("mixed: %d steps in %.3f s" % (K, rm["wall"]))
        vm = ndof * 5 * K / rm["wall"]
        a, b = r64["Q"], rm["Q"]
        out["mixed"] = {
            "value": vm, "speedup_vs_fp64": vm / value, "ms_per_step": rm["wall"] * 1e3 / K,
            "rel_l2_vs_fp64": float(np.linalg.norm(b - a) / np.linalg.norm(a)),
            "linf_vs_fp64": float(np.abs(b - a).max()),
            "traffic": measured_traffic("%s/mixed/stage" % key),
            "divb_max": rm["diag"]["divbmax"],
            "tile_shape": TILES[rm["tile"]],
        }
        if rm["tm"]:
            out["mixed"]["stage_kernel_ms"] = rm["tm"]["ms_stage_avg"]
            out["mixed"]["roofline_frac"] = ndof * b_stage(N) / (rm["tm"]["ms_stage_avg"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    if not args.no_cpu:
        _progress("CPU baseline (reference Fortran, MPI on the host cores)")
        out["cpu_baseline"] = cpu_reference_mpi(w, r64["artiv"])
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--workload", default="ensemble", choices=sorted(WORKLOADS) + ["ensemble"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="single-instance workloads: skip the mixed form")
    ap.add_argument("--no-dropin", action="store_true",
                    help="ensemble: skip the FP64 drop-in run one instance at a time")
    ap.add_argument("--no-event-pass", action="store_true",
                    help="single-instance workloads: skip the per-launch event pass (e.g. under a profiler)")
    ap.add_argument("--divb-report", action="store_true", help="ensemble: max |div B| of every instance")
    ap.add_argument("--tile", type=int, default=None,
                    help="single-instance workloads: fixed tile of the mixed form (0-8, see TILES; 0: 256/128, "
                         "1: 128/64, 2: 512/256) instead of autotuning, e.g. for profiles without the tuner's launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run two kernels of known HBM bytes (a 256 MiB fill: 256 MiB written; an "
                         "in-place add of two 256 MiB arrays: 512 MiB read, 256 MiB written) so that "
                         "hardware byte counters collected over this run can be scaled")
    ap.add_argument("--output", action="store_true",
                    help="time the output path (restart dump, Tecplot file) at the workload instead of steps")
    ap.add_argument("--nre", type=int, default=4, help="plot cells per cell edge of the Tecplot file (--output)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build holds a whole mesh (or a whole ensemble) on one "
              "MI355X; multi-GPU halo exchange is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        if args.copy_calibration:
            import torch
            src = torch.ones(1 << 25, dtype=torch.float64, device="cuda")
            dst = torch.full_like(src, 2.0)
            dst.add_(src)
            torch.cuda.synchronize()
            del src, dst
        run_ensemble(args)
        return
    w = WORKLOADS[args.workload]
    if args.output:
        run_output(w, args)
        return
    if "t_end" in w:
        run_crossings(w, args)
        return
    run_single(w, args)


if __name__ == "__main__":
    main()
