"""Host-side logic (no GPU): mesh generator, reference file formats, the C-ABI
library's exports and its host-side topology helper against the reference."""
import math
import os
import re

import numpy as np
import pytest

from conftest import REPO, golden_names, load_golden, split_topo

from chorus_mhd import QuadInp, quad_grid, read_mesh, read_restart, write_restart, fmt_d20


def test_quad_grid_counts_and_orientation():
    m = quad_grid(7, 5, 0.0, 2.0, -1.0, 0.5, jitter=0.15, seed=11)
    assert m.ncell == 35 and m.nvert == 48 and m.bnd.shape[0] == 2 * (7 + 5)
    x = m.xv[m.cells - 1]
    y = m.yv[m.cells - 1]
    # counter-clockwise convex quads (CHECK_MESH, connectivity.f90:255-263)
    for k in range(4):
        a, b, c = k, (k + 1) % 4, (k + 2) % 4
        cross = (x[:, b] - x[:, a]) * (y[:, c] - y[:, b]) - (y[:, b] - y[:, a]) * (x[:, c] - x[:, b])
        assert (cross > 0).all()
    # boundary vertices stay on the grid: periodic midpoints match within tolCYC
    hx, hy = 2.0 / 7, 1.5 / 5
    for iv in np.unique(m.bnd):
        xx, yy = m.xv[iv - 1], m.yv[iv - 1]
        assert abs(xx / hx - round(xx / hx)) < 1e-12 and abs((yy + 1.0) / hy - round((yy + 1.0) / hy)) < 1e-12


def test_jitter_bounds_and_seed():
    a = quad_grid(6, 6, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=5)
    b = quad_grid(6, 6, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=5)
    u = quad_grid(6, 6, 0.0, 1.0, 0.0, 1.0)
    assert np.array_equal(a.xv, b.xv)
    d = np.abs(a.xv - u.xv)
    assert d.max() <= 0.1 / 6 + 1e-15 and d.max() > 0


def test_files_round_trip(tmp_path):
    m = quad_grid(4, 3, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=0.1, seed=1)
    m.write(str(tmp_path))
    q = QuadInp(N=5, dt=1.2345678901234567e-3, MAXITER=7, DXCYCL=m.dxcycl, DYCYCL=m.dycycl, ARTIV=1)
    q.write(str(tmp_path / "QUAD.INP"))
    q2 = QuadInp.read(str(tmp_path / "QUAD.INP"))
    assert q2 == q
    m2 = read_mesh(str(tmp_path), q2)
    assert np.array_equal(m2.xv, m.xv) and np.array_equal(m2.cells, m.cells)
    assert m2.bnd_type == m.bnd_type
    lines = open(tmp_path / "verts.dat").read().splitlines()
    assert all(len(l) == 9 + 6 + 3 * 21 for l in lines)     # (I9,6X,3(D20.12,1X))
    with pytest.raises(ValueError):
        fmt_d20(1.0e30)


def test_restart_round_trip(tmp_path):
    Q = np.random.default_rng(0).standard_normal((5, 3, 3, 9))
    p = str(tmp_path / "restart.00000012.dat")
    write_restart(p, 12, 0.5, Q)
    it, t, Q2 = read_restart(p, 3)
    assert it == 12 and t == 0.5 and np.array_equal(Q, Q2)
    assert os.path.getsize(p) == 4 + 8 + 5 * (4 + 81 * 8)


# ------------------------------------------------------------------ C-ABI library
def _header_symbols():
    txt = open(os.path.join(REPO, "include", "chorus_mhd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(chorus_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from chorus_mhd import lib
    L = lib.load()
    syms = _header_symbols()
    assert len(syms) >= 11
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(lib.SIGNATURES)


@pytest.mark.parametrize("name", golden_names())
def test_library_topology_matches_reference(name):
    """chorus_topology (host code of the library) == the reference's connectivity."""
    from chorus_mhd.solver import topology
    g = load_golden(name)
    t = split_topo(g["topo"])
    mine = topology(g["mesh"])
    assert mine.nface == t["nface"]
    assert np.array_equal(mine.if2c, t["if2c"])
    assert np.array_equal(mine.ibfinter, t["ibfinter"])
    assert np.array_equal(mine.cyc2fpair, t["cyc2fpair"])
    assert np.array_equal(mine.ibfree, t["ibfree"])


def test_topology_rejects_unpaired_periodic_face():
    from chorus_mhd.solver import ChorusError, topology
    m = quad_grid(4, 4, 0.0, 1.0, 0.0, 1.0)
    m.dxcycl = 0.9            # shifts no longer match: map_cyc_loc.f90:37-40 stops
    with pytest.raises(ChorusError):
        topology(m)


def test_product_path_has_no_cpu_fallback():
    """Without a HIP device chorus_init fails loudly instead of computing on the CPU."""
    from chorus_mhd import lib
    from chorus_mhd.solver import ChorusError, Solver
    import ctypes
    L = lib.load()
    n = ctypes.c_int(0)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a HIP device is present")
    m = quad_grid(4, 4, 0.0, 1.0, 0.0, 1.0)
    with pytest.raises(ChorusError, match="no HIP device"):
        Solver(m, 4, 1e-3)
    del L, n


def test_quadrature_and_energy_diagnostics():
    """Solution-point quadrature is exact for degree N-1 polynomials on [0,1]
    and the cell volumes of a jittered mesh sum to the domain area, so the
    magnetic energy of a uniform field is |B|^2/2 times the area."""
    from chorus_mhd.diagnostics import magnetic_energy, point_volumes, quadrature_weights
    from chorus_mhd.ic import solution_points
    for N in range(2, 9):
        w, x = quadrature_weights(N), solution_points(N)
        for k in range(N):
            assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-13
    mesh = quad_grid(9, 6, -1.0, 1.0, -0.5, 0.5, jitter=0.12, seed=3)
    N = 4
    assert abs(point_volumes(mesh, N).sum() - 2.0) < 1e-13
    Q = np.zeros((mesh.ncell, N, N, 9))
    Q[..., 5], Q[..., 6], Q[..., 7] = 0.3, -0.4, 1.2
    assert abs(magnetic_energy(mesh, N, Q) - 0.5 * (0.09 + 0.16 + 1.44) * 2.0) < 1e-12


def test_glob2local_renumbers_by_first_appearance():
    """GLOB2LOCAL (readGRIDinfo.f90:124-199) on one rank: vertices numbered by
    first appearance in the cells, geometry and boundary edges unchanged."""
    mesh = quad_grid(5, 4, 0.0, 1.0, 0.0, 1.0, jitter=0.1, seed=2, bc="FREE")
    rng = np.random.default_rng(0)
    perm = rng.permutation(mesh.nvert)                 # scramble the file numbering first
    inv = np.empty_like(perm)
    inv[perm] = np.arange(mesh.nvert)
    from chorus_mhd.formats import Mesh
    scr = Mesh(mesh.xv[perm], mesh.yv[perm], (inv[mesh.cells - 1] + 1).astype(np.int32),
               (inv[mesh.bnd - 1] + 1).astype(np.int32), list(mesh.bnd_type), mesh.dxcycl, mesh.dycycl)
    loc = scr.glob2local()
    seen = []
    for c in loc.cells:
        for v in c:
            if v not in seen:
                seen.append(int(v))
    assert seen == list(range(1, loc.nvert + 1))
    assert np.array_equal(loc.xv[loc.cells - 1], scr.xv[scr.cells - 1])
    assert np.array_equal(loc.yv[loc.cells - 1], scr.yv[scr.cells - 1])
    assert np.array_equal(loc.xv[loc.bnd - 1], scr.xv[scr.bnd - 1])


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_text_round_trip_vectorised():
    """The vectorised round trip equals formatting every value and parsing it
    back, over magnitudes from 1e-20 to 1e3 (both signs, and zero)."""
    from chorus_mhd.formats import text_round_trip
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-1, 1, 4000) * 10.0 ** rng.integers(-20, 4, 4000),
                        [0.0, -0.0, 0.0625, -0.0625, np.nextafter(0.0625, 0), 2 * math.pi, 1.0 - 1e-17]])
    want = np.array([float(fmt_d20(v)) for v in x])
    got = text_round_trip(x)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
