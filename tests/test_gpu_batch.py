"""Batched ensembles (chorus_batch_*): every instance of a batch must evolve
exactly as the same instance run alone through chorus_init/chorus_step, in
both precision forms, through several hipGraph replays (the buffer rotation
of SSP(5,3) cycles through six graphs), with per-instance AV quantities, and
whatever tile shape the autotuner picks (the tile only decides which block
evaluates a cell, never the arithmetic of a cell).
"""
import math
import sys

import numpy as np
import pytest

from conftest import golden_params, load_golden

import oracle as orc
from chorus_mhd import quad_grid, write_restart
from chorus_mhd import ic as ics
from chorus_mhd.solver import Batch, Instance, Solver

pytestmark = pytest.mark.gpu

# tile indices of the mixed form (chorus_set_tile): three shapes shared by the
# stage and gradient kernels, and the six pairs of different ones
NTILES = 9


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _members(precision, artiv):
    """Uneven sizes and orders; two Orszag-Tang instances share (N, gamma, dt)
    and therefore one group; the last instance has gamma 1.4 and a FREE
    boundary (the Orszag-Tang fields on the unit square: the rotor's sub-cell
    taper would make the reference's residual NaN on a mesh this coarse)."""
    two_pi = 2 * math.pi
    spec = [
        dict(n=8, m=8, box=(0, two_pi, 0, two_pi), N=4, jit=0.10, seed=1, bc="CYCL", ic="ot", gam=5 / 3),
        dict(n=10, m=6, box=(0, two_pi, 0, two_pi), N=4, jit=0.05, seed=2, bc="CYCL", ic="ot", gam=5 / 3),
        dict(n=6, m=6, box=(0, two_pi, 0, two_pi), N=3, jit=0.12, seed=3, bc="CYCL", ic="ot", gam=5 / 3),
        dict(n=8, m=8, box=(0, 1, 0, 1), N=5, jit=0.08, seed=4, bc="FREE", ic="ot", gam=1.4),
    ]
    out = []
    for s in spec:
        mesh = quad_grid(s["n"], s["m"], *s["box"], jitter=s["jit"], seed=s["seed"], bc=s["bc"])
        dt = 2e-4 if s["gam"] > 1.5 else 2e-5
        out.append((Instance(mesh, s["N"], dt, gam=s["gam"], artiv=artiv, precision=precision), s["ic"]))
    return out


def _initial(inst, kind):
    s = Solver(inst.mesh, inst.N, inst.dt, gam=inst.gam, artiv=inst.artiv, precision=inst.precision)
    q = ics.orszag_tang(s, inst.mesh, inst.N, inst.gam) if kind == "ot" else ics.rotor(s, inst.mesh, inst.N, inst.gam)
    s.close()
    return q


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("artiv", [0, 1])
def test_batch_equals_single_instances(precision, artiv, tmp_path):
    _say("batch", precision, "artiv", artiv)
    mem = _members(precision, artiv)
    q0 = [_initial(m, k) for m, k in mem]
    b = Batch([m for m, _ in mem])
    groups = b.groups()
    assert sum(g["ninst"] for g in groups) == len(mem)
    assert len(groups) == 3          # the two N=4 Orszag-Tang instances share a group
    for i, q in enumerate(q0):
        b.set_state(i, q)
    for i, (m, _) in enumerate(mem):
        assert np.array_equal(b.compresid(i), _single(m, q0[i], 0, resid=True)), "residual of instance %d" % i
    b.step(7)
    for i, (m, _) in enumerate(mem):
        ref = _single(m, q0[i], 7)
        got = b.get_state(i)
        assert np.array_equal(got, ref), "instance %d differs (max %g)" % (i, np.abs(got - ref).max())
        # the instance's restart dump, packed on the device, equals the host-written one
        b.write_restart(i, str(tmp_path / "dev.dat"), 7, 0.5)
        write_restart(str(tmp_path / "host.dat"), 7, 0.5, got)
        assert (tmp_path / "dev.dat").read_bytes() == (tmp_path / "host.dat").read_bytes(), "dump of instance %d" % i
    b.close()


def _single(m, q0, nsteps, resid=False):
    s = Solver(m.mesh, m.N, m.dt, gam=m.gam, artiv=m.artiv, precision=m.precision)
    s.set_state(q0)
    if resid:
        r = s.compresid()
        s.close()
        return r
    s.step(nsteps)
    q = s.get_state()
    s.close()
    return q


def test_batch_autotune_and_divb():
    _say("batch autotune")
    mem = _members("mixed", 1)
    q0 = [_initial(m, k) for m, k in mem]
    b = Batch([m for m, _ in mem])
    for i, q in enumerate(q0):
        b.set_state(i, q)
    tiles = b.autotune()
    assert len(tiles) == len(b.groups()) and all(0 <= t < NTILES for t in tiles)
    assert [g["tile"] for g in b.groups()] == tiles
    b.step(3)
    dv = b.divb()
    for i, (m, _) in enumerate(mem):
        s = Solver(m.mesh, m.N, m.dt, gam=m.gam, artiv=m.artiv, precision=m.precision)
        s.set_state(q0[i])
        s.step(3)
        assert np.array_equal(b.get_state(i), s.get_state()), "instance %d after autotune" % i
        assert dv[i] == s.diag()["divbmax"]
        s.close()
    b.close()


# reference fixtures run as members of one batch (all SSP(5,3)): N = 3..6,
# periodic and FREE faces, AV on and off, mu/eta0 != 0, the gam = 1.4 rotor
# and the field loop; the three N=4 AV members with dt = 1e-3 share one group
BATCH_GOLDENS = ["ot_n6_N3_free_av", "ot_n5_N6_free", "rotor_n16_N6_free_av_g14", "ot_n8_N4_jitter_av",
                 "rough_n6_N4_jitter_av", "loop_n16x8_N5", "visc_n6_N4_jitter"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_members_vs_reference(precision):
    """Every member of a batch against the reference's own fixture (or, for
    the one member made here, against the oracle), not against another HIP
    run: FP64 bitwise without AV and within 1e-13 of each field's magnitude
    with AV; mixed within 1e-6 relative L2 (residuals within 1e-5)."""
    _say("batch vs reference", precision)
    gs = [load_golden(n) for n in BATCH_GOLDENS]
    mem = [Instance(g["mesh"], int(g["N"]), float(g["dt"]), precision=precision, **golden_params(g))
           for g in gs]
    # an oracle-made Orszag-Tang member in the group of the N=4 AV fixtures
    om = quad_grid(10, 6, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=0.12, seed=21)
    o = orc.Oracle(om, 4, dt=1e-3, artiv=1)
    o.init_ot()
    oq0 = o.get_q()
    o.steps(3)
    oq3 = o.get_q()
    mem.append(Instance(om, 4, 1e-3, artiv=1, precision=precision))
    b = Batch(mem)
    groups = b.groups()
    assert sum(x["ninst"] for x in groups) == len(mem)
    assert max(x["ninst"] for x in groups) == 3
    assert sorted({x["N"] for x in groups}) == [3, 4, 5, 6]
    for i, g in enumerate(gs):
        b.set_state(i, g["q0"])
    b.set_state(len(gs), oq0)

    def check(i, got, ref, artiv, what, budget=1e-6):
        if precision == "mixed":
            rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            _say(what, "mixed rel L2", rel)
            assert rel <= budget, what
        elif artiv == 0:
            assert np.array_equal(got, ref), what
        else:
            scale = np.abs(ref).max(axis=(0, 1, 2))
            assert (np.abs(got - ref).max(axis=(0, 1, 2)) <= 1e-13 * scale).all(), what

    for i, g in enumerate(gs):
        # the mixed form's residual budget is 1e-5 (its state budget 1e-6)
        check(i, b.compresid(i), g["resid"], int(g["artiv"]), "residual of " + BATCH_GOLDENS[i], 1e-5)
    done = 0
    for target in (2, 3):
        b.step(target - done)
        done = target
        for i, g in enumerate(gs):
            if int(g["steps"]) == target:
                check(i, b.get_state(i), g["qend"], int(g["artiv"]), "%s after %d steps" % (BATCH_GOLDENS[i], target))
    check(len(gs), b.get_state(len(gs)), oq3, 1, "oracle member after 3 steps")
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_ssp54_viscous_and_av_vs_reference(precision):
    """FiveStageRk batches with physical viscosity, resistivity and AV together
    (VM=3) and AV alone (VM=2), N=5 on FREE faces, against the reference's
    fixtures; an oracle-made Orszag-Tang member shares the VM=3 group, so that
    group's stage maximum of the wave speed is formed per instance."""
    _say("batch SSP(5,4) vs reference", precision)
    names = ["visc_n6_N5_free_av", "smooth_n6_N5_free_av"]
    gs = [load_golden(n) for n in names]
    mem = [Instance(g["mesh"], int(g["N"]), float(g["dt"]), precision=precision, **golden_params(g))
           for g in gs]
    kw = golden_params(gs[0])
    assert kw["k_stage"] == 5 and kw["artiv"] == 1 and kw["mu"] > 0 and kw["eta0"] > 0
    om = quad_grid(7, 5, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=0.1, seed=33, bc="FREE")
    o = orc.Oracle(om, 5, dt=float(gs[0]["dt"]), **kw)
    o.init_ot()
    oq0 = o.get_q()
    o.steps(2)
    oq2 = o.get_q()
    mem.append(Instance(om, 5, float(gs[0]["dt"]), precision=precision, **kw))
    b = Batch(mem)
    groups = b.groups()
    assert sorted(x["ninst"] for x in groups) == [1, 2]
    for i, g in enumerate(gs):
        b.set_state(i, g["q0"])
    b.set_state(2, oq0)
    assert all(int(g["steps"]) == 2 for g in gs)
    b.step(2)
    for i, (got, ref, what) in enumerate(((b.get_state(0), gs[0]["qend"], names[0]),
                                          (b.get_state(1), gs[1]["qend"], names[1]),
                                          (b.get_state(2), oq2, "oracle member"))):
        if precision == "mixed":
            rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            _say(what, "mixed rel L2", rel)
            assert rel <= 1e-6, what
        else:
            scale = np.abs(ref).max(axis=(0, 1, 2))
            err = np.abs(got - ref).max(axis=(0, 1, 2))
            _say(what, "fp64 max err / scale", float((err / np.where(scale > 0, scale, 1.0)).max()))
            assert (err <= 1e-13 * scale).all(), what
    b.close()


def test_batch_rejects_mixed_steppers():
    mesh = quad_grid(4, 4, 0, 1, 0, 1)
    with pytest.raises(Exception):
        Batch([Instance(mesh, 3, 1e-4, k_stage=3), Instance(mesh, 3, 1e-4, k_stage=5)])


@pytest.mark.parametrize("artiv", [0, 1])
def test_single_context_tiles_agree(artiv):
    """chorus_set_tile / chorus_autotune on one context: every tile shape gives
    the same bits (a tile decides which block evaluates a cell, not how)."""
    _say("single-context tiles, artiv", artiv)
    mesh = quad_grid(12, 10, 0, 2 * math.pi, 0, 2 * math.pi, jitter=0.1, seed=9)
    N = 5
    s0 = Solver(mesh, N, 2e-4, artiv=artiv, precision="mixed")
    q0 = ics.orszag_tang(s0, mesh, N)
    s0.close()
    outs = []
    for tile in range(NTILES):
        s = Solver(mesh, N, 2e-4, artiv=artiv, precision="mixed")
        s.set_tile(tile)
        s.set_state(q0)
        s.step(3)
        outs.append(s.get_state())
        s.close()
    for t in range(1, NTILES):
        assert np.array_equal(outs[0], outs[t]), "tile %d differs" % t
    s = Solver(mesh, N, 2e-4, artiv=artiv, precision="mixed")
    s.set_state(q0)
    assert 0 <= s.autotune() < NTILES
    s.step(3)
    assert np.array_equal(s.get_state(), outs[0])
    with pytest.raises(Exception):
        s.set_tile(NTILES)
    s.close()
    f = Solver(mesh, N, 2e-4, artiv=artiv)
    assert f.autotune() == 0
    with pytest.raises(Exception):
        f.set_tile(1)
    f.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_tecplot_equals_single_instances(precision, tmp_path):
    """Each instance's Tecplot file, written from the batch, is the file the
    same instance writes alone (AV on: muAV and etaAV of the last stage)."""
    _say("batch tecplot", precision)
    mem = _members(precision, 1)[:4]
    q0 = [_initial(m, k) for m, k in mem]
    b = Batch([m for m, _ in mem])
    for i, q in enumerate(q0):
        b.set_state(i, q)
        b.plot_setup(i, 2 + i % 2)
    b.step(3)
    for i, (m, _) in enumerate(mem):
        b.write_tecplot(i, str(tmp_path / "b.plt"), 0.125)
        s = Solver(m.mesh, m.N, m.dt, gam=m.gam, artiv=m.artiv, precision=m.precision)
        s.plot_setup(2 + i % 2)
        s.set_state(q0[i])
        s.step(3)
        s.write_tecplot(str(tmp_path / "s.plt"), 0.125)
        s.close()
        assert (tmp_path / "b.plt").read_bytes() == (tmp_path / "s.plt").read_bytes(), "instance %d" % i
    b.close()
