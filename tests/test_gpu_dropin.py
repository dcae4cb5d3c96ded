"""The Fortran drop-in end to end: the reference's own set-up chain, restart
reader/writer and loop structure, with the time step on the GPU through the
iso_c_binding module (chorus-mhd_amd/fortran).  Its restart dump after k steps
must equal the compiled reference's state (golden fixtures) and the file the
reference's own writer makes of the same state, and the Tecplot
file it writes at the last step through the library's device-side output must
equal the reference's own file byte for byte (tests/golden/plot_*)."""
import math
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, golden_names, load_golden

from chorus_mhd import QuadInp, quad_grid, read_restart, write_restart

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(REPO, "chorus-mhd_amd", "fortran", "_build", "sd2d_mhd_hip")
# the driver linked with the reference objects built with gam = 1.4
DROPIN_G14 = os.path.join(REPO, "chorus-mhd_amd", "fortran", "_build", "g14", "sd2d_mhd_hip")


def _driver(g):
    return DROPIN_G14 if float(g["gam"]) == 1.4 else DROPIN


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(DROPIN_G14)),
                    reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("name", golden_names())
def test_fortran_dropin_matches_reference(name, tmp_path):
    print("case dropin", name, flush=True)
    g = load_golden(name)
    mesh, N, steps = g["mesh"], int(g["N"]), int(g["steps"])
    d = str(tmp_path)
    mesh.write(d)
    os.makedirs(os.path.join(d, "TEC"))
    write_restart(os.path.join(d, "restart.in"), 0, 0.0, g["q0"])
    QuadInp(N=N, dt=float(g["dt"]), MAXITER=steps, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl,
            ARTIV=int(g["artiv"]), k_stage=int(g["k_stage"]), restart=1, restart_ord=N,
            NAMERESTART="restart.in", nwrite=1000000, mu=float(g["mu"]),
            eta0=float(g["eta0"])).write(os.path.join(d, "QUAD.INP"))
    r = subprocess.run([_driver(g)], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # the same run with the reference's own restart writer on the copied-back
    # state: the device-packed dumps must equal its files byte for byte
    d2 = os.path.join(d, "refio")
    os.makedirs(os.path.join(d2, "TEC"))
    for f in os.listdir(d):
        if f in ("restart.in", "QUAD.INP") or f.endswith(".dat") and not f.startswith("restart."):
            shutil.copy(os.path.join(d, f), d2)
    r = subprocess.run([_driver(g), "refio"], cwd=d2, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    dumps = sorted(f for f in os.listdir(d2) if f.startswith("restart.0"))
    assert dumps == sorted(f for f in os.listdir(d) if f.startswith("restart.0")) and len(dumps) == 2
    for f in dumps:
        assert open(os.path.join(d, f), "rb").read() == open(os.path.join(d2, f), "rb").read(), f
    it, t, Q = read_restart(os.path.join(d, "restart.%08d.dat" % steps), N)
    assert it == steps
    if int(g["artiv"]) == 0:
        assert np.array_equal(Q, g["qend"])
    else:
        assert np.abs(Q - g["qend"]).max() <= 1e-13 * np.abs(g["qend"]).max()
    lines = open(os.path.join(d, "resid.out")).read().split("\n")
    row = lines[1].split()
    assert int(row[0]) == 1


PLOT_CASES = ["plot_ot_8x6_N4_nre3", "plot_ot_6x6_N3_free_nre2"]


@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("name", PLOT_CASES)
def test_fortran_dropin_tecplot_matches_reference(name, tmp_path):
    """Same steps from the same state as the reference run that wrote the golden
    .plt (restart from its initial state; the reference started from its own
    initial condition, which is that state)."""
    print("case dropin tecplot", name, flush=True)
    g = np.load(os.path.join(REPO, "tests", "golden", name + ".npz"))
    mesh = quad_grid(int(g["nx"]), int(g["ny"]), 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=float(g["jitter"]),
                     seed=int(g["seed"]), bc=str(g["bc"]))
    N, steps = int(g["N"]), int(g["steps"])
    d = str(tmp_path)
    mesh.write(d)
    os.makedirs(os.path.join(d, "TEC"))
    write_restart(os.path.join(d, "restart.in"), 0, 0.0, g["q0"])
    QuadInp(N=N, dt=float(g["dt"]), MAXITER=steps, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl, ARTIV=1,
            NRE=int(g["nre"]), restart=1, restart_ord=N, NAMERESTART="restart.in",
            nwrite=1000000).write(os.path.join(d, "QUAD.INP"))
    r = subprocess.run([DROPIN], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = os.path.join(d, "TEC", str(g["plt_name"]))
    assert os.path.exists(out), sorted(os.listdir(os.path.join(d, "TEC")))
    ref = os.path.join(REPO, "tests", "golden", name + ".plt")
    assert open(out, "rb").read() == open(ref, "rb").read()


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(DROPIN_G14)),
                    reason="drop-in driver not built (needs the reference sources)")
@pytest.mark.parametrize("name", ["ot_n8_N4_jitter_av", "smooth_n6_N5_free_av", "visc_n6_N5_free_av",
                                  "rotor_n16_N6_free_av_g14"])
def test_fortran_dropin_mixed_within_budget(name, tmp_path):
    """The driver's mixed form (argument "mixed") against the reference's FP64
    state: relative L2 within the 1e-6 budget of the mixed form."""
    print("case dropin mixed", name, flush=True)
    g = load_golden(name)
    mesh, N, steps = g["mesh"], int(g["N"]), int(g["steps"])
    d = str(tmp_path)
    mesh.write(d)
    os.makedirs(os.path.join(d, "TEC"))
    write_restart(os.path.join(d, "restart.in"), 0, 0.0, g["q0"])
    QuadInp(N=N, dt=float(g["dt"]), MAXITER=steps, DXCYCL=mesh.dxcycl, DYCYCL=mesh.dycycl,
            ARTIV=int(g["artiv"]), k_stage=int(g["k_stage"]), restart=1, restart_ord=N,
            NAMERESTART="restart.in", nwrite=1000000, mu=float(g["mu"]),
            eta0=float(g["eta0"])).write(os.path.join(d, "QUAD.INP"))
    r = subprocess.run([_driver(g), "mixed"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    it, t, Q = read_restart(os.path.join(d, "restart.%08d.dat" % steps), N)
    assert it == steps
    rel = np.linalg.norm(Q - g["qend"]) / np.linalg.norm(g["qend"])
    assert rel < 1e-6, rel
