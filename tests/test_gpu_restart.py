"""Restart dump written from the device (chorus_write_restart, replacing
WRITE_ALL_DATA_PARALLEL, write_data.f90:5-83).

The file must be byte-identical to the dump format restated in
chorus_mhd.formats.write_restart (whose layout the reference's own reader
accepts: every golden run starts from such a file) for the same state, and the
global cell numbers passed in must land in each record's first word.  The
Fortran drop-in test compares the same writer with the reference's own
WRITE_ALL_DATA_PARALLEL byte for byte."""
import os
import struct

import numpy as np
import pytest

from conftest import golden_names, golden_params, load_golden

from chorus_mhd import quad_grid, read_restart, write_restart
from chorus_mhd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", golden_names()[:3])
def test_restart_dump_matches_format(name, precision, tmp_path):
    print("restart", name, precision, flush=True)
    g = load_golden(name)
    N = int(g["N"])
    s = Solver(g["mesh"], N, float(g["dt"]), precision=precision, **golden_params(g))
    s.set_state(g["q0"])
    s.step(2)
    ctime = float(g["dt"]) + float(g["dt"])
    out = str(tmp_path / Solver.restart_name(2))
    s.write_restart(out, 2, ctime)
    ref = str(tmp_path / "host.dat")
    write_restart(ref, 2, ctime, s.get_state())
    assert open(out, "rb").read() == open(ref, "rb").read()
    # global cell numbers: a permutation lands in the records, the payload does not move
    nc = g["mesh"].ncell
    icg = np.random.default_rng(5).permutation(nc).astype(np.int32) + 1
    s.write_restart(out, 7, 0.25, icg)
    s.close()
    raw = open(out, "rb").read()
    assert struct.unpack_from("<id", raw, 0) == (7, 0.25)
    rec = np.dtype([("icg", "<i4"), ("q", "<f8", (9 * N * N,))])
    body = np.frombuffer(raw, dtype=rec, offset=12)
    assert np.array_equal(body["icg"], icg)
    host = np.frombuffer(open(ref, "rb").read(), dtype=rec, offset=12)
    assert np.array_equal(body["q"], host["q"])


def test_restart_dump_large_mesh(tmp_path):
    """A 96 x 80 mesh at N = 7 (many blocks per record run): the file read back
    equals the state."""
    print("restart large", flush=True)
    mesh = quad_grid(96, 80, 0.0, 2 * np.pi, 0.0, 2 * np.pi, jitter=0.1, seed=11)
    N = 7
    rng = np.random.default_rng(3)
    q = rng.standard_normal((mesh.ncell, N, N, 9))
    q[..., 0] = 1.0 + 0.1 * np.abs(q[..., 0])
    q[..., 4] = 5.0 + np.abs(q[..., 4])
    s = Solver(mesh, N, 1e-4)
    s.set_state(q)
    out = str(tmp_path / "r.dat")
    s.write_restart(out, 123, 4.5)
    s.close()
    assert os.path.getsize(out) == 12 + mesh.ncell * (4 + 8 * 9 * N * N)
    it, t, Q = read_restart(out, N)
    assert (it, t) == (123, 4.5)
    assert np.array_equal(Q, q)


def test_restart_dump_rejects_bad_path(tmp_path):
    mesh = quad_grid(4, 4, 0.0, 1.0, 0.0, 1.0)
    s = Solver(mesh, 3, 1e-3)
    with pytest.raises(Exception, match="cannot open"):
        s.write_restart(str(tmp_path / "missing" / "r.dat"), 1, 0.0)
    s.close()
