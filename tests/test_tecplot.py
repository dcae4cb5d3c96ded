"""Tecplot output (chorus_plot_setup / chorus_write_tecplot) against the
reference's own files: tests/golden/plot_*.plt were written by the unmodified
reference (tecplotter2dsetup, tecplotter2d, tecpost_parallel) after the same
steps from the same state (tests/golden/make_golden.py).

CPU: the plot-node topology (node numbering, FP32 coordinates, FP64 ranges,
connectivity) of the library's host builder equals the reference's.
GPU: the FP64 drop-in writes the reference's file byte for byte (node sums in
the reference's order, muAV and etaAV of the last stage's input); the mixed
form's variables agree to its error budget.
"""
import math
import os

import numpy as np
import pytest

from chorus_mhd import quad_grid
from chorus_mhd.solver import Solver, plot_topology
from chorus_mhd.tecplot import read_plt

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["plot_ot_8x6_N4_nre3", "plot_ot_6x6_N3_free_nre2"]


def _case(name):
    g = np.load(os.path.join(HERE, "golden", name + ".npz"))
    mesh = quad_grid(int(g["nx"]), int(g["ny"]), 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=float(g["jitter"]),
                     seed=int(g["seed"]), bc=str(g["bc"]))
    return g, mesh, os.path.join(HERE, "golden", name + ".plt")


@pytest.mark.parametrize("name", CASES)
def test_plot_topology_matches_reference(name):
    g, mesh, plt = _case(name)
    f = read_plt(plt)
    xy, c2n = plot_topology(mesh, int(g["nre"]))
    assert f["names"] == ["X", "Y", "RHO", "U", "V", "W", "P", "Bx", "By", "Bz", "Az", "muAV", "etaAV"]
    assert xy.shape[1] == f["nnode"] and c2n.shape[0] == f["ncell"] == mesh.ncell * int(g["nre"]) ** 2
    assert np.array_equal(c2n - 1, f["conn"])
    assert np.array_equal(xy.astype(np.float32), f["data"][:2])
    assert np.array_equal([xy[0].min(), xy[0].max()], f["minmax"][0])
    assert np.array_equal([xy[1].min(), xy[1].max()], f["minmax"][1])


def test_plot_topology_counts():
    """Closed form on an n x m periodic grid: (n*NRE+1)(m*NRE+1) nodes (periodic
    sides are distinct boundary faces, so their nodes are not merged)."""
    mesh = quad_grid(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.05, seed=3)
    for nre in (1, 2, 4):
        xy, c2n = plot_topology(mesh, nre)
        assert xy.shape[1] == (5 * nre + 1) * (4 * nre + 1)
        assert c2n.shape == (20 * nre * nre, 4) and c2n.min() == 1 and c2n.max() == xy.shape[1]


def _gpu_file(name, precision, tmp_path):
    g, mesh, plt = _case(name)
    s = Solver(mesh, int(g["N"]), float(g["dt"]), artiv=1, precision=precision)
    s.plot_setup(int(g["nre"]))
    s.set_state(g["q0"])
    s.step(int(g["steps"]))
    ctime = 0.0
    for _ in range(int(g["steps"])):
        ctime = ctime + float(g["dt"])     # ctime = ctime0 + dt per step (iterations.f90:184-215)
    assert Solver.tecplot_name(ctime) == str(g["plt_name"])
    out = str(tmp_path / Solver.tecplot_name(ctime))
    s.write_tecplot(out, ctime)
    s.close()
    return out, plt


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_tecplot_fp64_byte_identical(name, tmp_path):
    print("tecplot fp64", name, flush=True)
    out, plt = _gpu_file(name, "fp64", tmp_path)
    a, b = open(out, "rb").read(), open(plt, "rb").read()
    if a != b:
        fa, fb = read_plt(out), read_plt(plt)
        bad = [fb["names"][v] for v in range(13) if not np.array_equal(fa["data"][v], fb["data"][v])]
        raise AssertionError("file differs from the reference's; variables: %s" % bad)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_tecplot_mixed_within_budget(name, tmp_path):
    print("tecplot mixed", name, flush=True)
    out, plt = _gpu_file(name, "mixed", tmp_path)
    fa, fb = read_plt(out), read_plt(plt)
    assert np.array_equal(fa["conn"], fb["conn"]) and np.array_equal(fa["data"][:2], fb["data"][:2])
    for v in range(2, 11):
        ref = fb["data"][v].astype(np.float64)
        err = np.abs(fa["data"][v] - ref).max() / max(np.abs(ref).max(), 1e-30)
        assert err < 1e-6, (fb["names"][v], err)


@pytest.mark.gpu
@pytest.mark.parametrize("N,nre,bc", [(2, 1, "FREE"), (5, 3, "CYCL"), (6, 4, "FREE"), (8, 2, "CYCL")])
def test_tecplot_linear_state_exact(N, nre, bc, tmp_path):
    """Orders and plot resolutions the reference files do not cover: a state
    linear in x and y is bilinear in each cell's reference coordinates, so the
    interpolation to every plot node, the averaging over the cells sharing it
    and the primitive variables reproduce the closed form at the node's
    coordinates (to FP32 rounding); the AV columns are zero without AV."""
    print("tecplot linear", N, nre, bc, flush=True)
    from chorus_mhd.ic import solution_xy
    mesh = quad_grid(7, 5, 0.0, 2.0, 0.0, 1.0, jitter=0.12, seed=4, bc=bc)
    gam = 5.0 / 3
    xy = solution_xy(mesh, N)
    coef = np.array([[2.0, 0.1, 0.05], [0.3, 0.2, -0.1], [-0.2, 0.05, 0.1], [0.1, -0.03, 0.02], [9.0, 0.4, 0.3],
                     [0.5, 0.1, 0.0], [-0.4, 0.0, 0.2], [0.2, 0.05, 0.05], [0.1, 0.3, -0.2]])
    Q = coef[:, 0] + xy[..., 0:1] * coef[:, 1] + xy[..., 1:2] * coef[:, 2]
    s = Solver(mesh, N, 1e-4, gam=gam)
    s.plot_setup(nre)
    s.set_state(Q)
    out = str(tmp_path / "lin.plt")
    s.write_tecplot(out, 0.5)
    s.close()
    f = read_plt(out)
    pxy, c2n = plot_topology(mesh, nre)
    assert np.array_equal(f["data"][:2], pxy.astype(np.float32)) and np.array_equal(f["conn"], c2n - 1)
    q = coef[:, 0:1] + pxy[0][None] * coef[:, 1:2] + pxy[1][None] * coef[:, 2:3]
    u = q[1:4] / q[0]
    p = (q[4] - 0.5 * q[0] * (u * u).sum(0) - 0.5 * (q[5:8] ** 2).sum(0)) * (gam - 1.0)
    want = np.concatenate([q[0:1], u, p[None], q[5:9]])
    for k in range(9):
        err = np.abs(f["data"][2 + k] - want[k]).max() / np.abs(want[k]).max()
        assert err < 2e-6, (f["names"][2 + k], err)
    assert not f["data"][11:].any()


@pytest.mark.gpu
def test_av_columns_reset_by_set_state(tmp_path):
    """A newly loaded state has been evaluated by no stage: its muAV/etaAV
    columns are zeros and chorus_diag's wave speed is NaN, as before the first
    stage; the next step fills them again.  In a batch only the reloaded
    member loses its columns, not the other member of its group."""
    print("AV columns after set_state", flush=True)
    from conftest import golden_params, load_golden
    from chorus_mhd.solver import Batch, Instance
    g = load_golden("rough_n6_N4_jitter_av")       # AV ramp active in every cell
    mesh, N, dt = g["mesh"], int(g["N"]), float(g["dt"])
    s = Solver(mesh, N, dt, **golden_params(g))
    s.plot_setup(2)
    s.set_state(g["q0"])
    out = str(tmp_path / "a.plt")
    s.step(1)
    s.write_tecplot(out, 0.0)
    assert read_plt(out)["data"][11].any()
    assert math.isfinite(s.diag()["eigvmax"])
    s.set_state(g["q0"])
    s.write_tecplot(out, 0.0)
    assert not read_plt(out)["data"][11:].any()
    assert math.isnan(s.diag()["eigvmax"])
    s.step(1)
    s.write_tecplot(out, 0.0)
    assert read_plt(out)["data"][11].any()
    s.close()
    b = Batch([Instance(mesh, N, dt, precision="fp64", **golden_params(g)) for _ in range(2)])
    assert len(b.groups()) == 1
    for i in range(2):
        b.plot_setup(i, 2)
        b.set_state(i, g["q0"])
    b.step(1)
    b.set_state(0, g["q0"])
    b.write_tecplot(0, out, 0.0)
    assert not read_plt(out)["data"][11:].any()
    b.write_tecplot(1, out, 0.0)
    assert read_plt(out)["data"][11].any()
    b.close()
