"""8-node (curved) cells: the reference computes the 8-node shape-function
derivatives (MAPDER_8_NODE, mapder.f90:117-253) and stops before using them
(initsetup.f90:71-78).  The oracle sums them over a cell's 8 nodes as
calcjacob sums the 4-node ones (calcjacob.f90:1-86).  CPU pins:

* the oracle's tables equal the ones the reference computed (harness mode
  mapder8, tests/golden/mapder8_tables.npz), bit for bit, N = 2..8;
* straight cells with their edge midpoints as nodes 5..8: the 8-node metric
  and solution points equal the 4-node ones to round-off;
* a map in the 8-node span (x = psi + a psi^2 eta, y = eta + b psi eta^2):
  the Jacobian and metric rows equal the closed form at every point;
* free stream on a curved periodic mesh: resid(1:8) vanishes to round-off
  (the discrete metric identity of the curved cells).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

import oracle as orc
from chorus_mhd import Mesh, curved_grid, quad_grid

GAM = 5.0 / 3


@pytest.mark.parametrize("N", range(2, 9))
def test_mapder8_tables_equal_reference(N):
    g = np.load(os.path.join(GOLDEN, "mapder8_tables.npz"))
    a, b, c = orc.mapder8(N)
    assert np.array_equal(np.concatenate([a, b, c]), g["N%d" % N])


def _midpoints(mesh):
    X, Y = mesh.xv[mesh.cells - 1], mesh.yv[mesh.cells - 1]
    xm = np.stack([0.5 * (X[:, k] + X[:, (k + 1) % 4]) for k in range(4)], axis=1)
    ym = np.stack([0.5 * (Y[:, k] + Y[:, (k + 1) % 4]) for k in range(4)], axis=1)
    return xm, ym


@pytest.mark.parametrize("N", [3, 5])
def test_straight_cells_reduce_to_bilinear(N):
    mesh = quad_grid(5, 4, 0.0, 1.3, 0.0, 1.0, jitter=0.15, seed=2, bc="FREE")
    o4, o8 = orc.Oracle(mesh, N), orc.Oracle(mesh, N)
    o8.set_nodes8(*_midpoints(mesh))
    for name in ("Jac", "S1", "S2", "XXsolu"):
        a, b = o4.array(name), o8.array(name)
        assert np.allclose(a, b, rtol=0, atol=1e-14 * np.abs(a).max()), name
    # a wrong node order (mid nodes rotated by one edge) is not the bilinear map
    xm, ym = _midpoints(mesh)
    o8.set_nodes8(np.roll(xm, 1, axis=1), np.roll(ym, 1, axis=1))
    assert not np.allclose(o4.array("Jac"), o8.array("Jac"), rtol=0, atol=1e-6)


def test_quadratic_map_closed_form():
    a, b = 0.1, 0.08
    ref = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0], [1, 0.5], [0.5, 1], [0, 0.5]], dtype=float)
    phys = np.stack([ref[:, 0] + a * ref[:, 0] ** 2 * ref[:, 1], ref[:, 1] + b * ref[:, 0] * ref[:, 1] ** 2], 1)
    mesh = Mesh(phys[:4, 0].copy(), phys[:4, 1].copy(), np.array([[1, 2, 3, 4]], dtype=np.int32),
                np.array([[1, 2], [2, 3], [3, 4], [4, 1]], dtype=np.int32), ["FREE"] * 4, 1.0, 1.0)
    N = 4
    o = orc.Oracle(mesh, N)
    o.set_nodes8(phys[4:, 0][None, :], phys[4:, 1][None, :])
    xs, xf = o.array("Xs"), o.array("Xf")
    jac = o.array("Jac").reshape(N, N)              # [js][is]
    for js in range(N):
        for i in range(N):
            p, e = xs[i], xs[js]
            xp, xe, yp, ye = 1 + 2 * a * p * e, a * p * p, b * e * e, 1 + 2 * b * p * e
            assert jac[js, i] == pytest.approx(xp * ye - xe * yp, rel=1e-14, abs=0)
    s1 = o.array("S1").reshape(N, N + 1, 2, 2)      # [jfp][ifp][b][a]
    for jfp in range(N):
        for ifp in range(N + 1):
            p, e = xf[ifp], xs[jfp]
            xe, ye = a * p * p, 1 + 2 * b * p * e
            assert s1[jfp, ifp, 0, 0] == pytest.approx(ye, rel=1e-14, abs=1e-15)
            assert s1[jfp, ifp, 1, 0] == pytest.approx(-xe, rel=1e-14, abs=1e-15)
    xy = o.array("XXsolu").reshape(N, N, 2)
    p, e = xs[None, :], xs[:, None]
    assert np.allclose(xy[..., 0], p + a * p * p * e, rtol=0, atol=1e-14)
    assert np.allclose(xy[..., 1], e + b * p * e * e, rtol=0, atol=1e-14)


@pytest.mark.parametrize("N", [2, 3, 5, 6])
def test_free_stream_on_curved_cells(N):
    mesh, xm, ym = curved_grid(6, 5, 0.0, 1.2, 0.0, 1.0, amp=0.15)
    o = orc.Oracle(mesh, N)
    o.set_nodes8(xm, ym)
    rho, u, v, w, p, bx, by, bz = 1.3, 0.7, -0.4, 0.2, 0.9, 0.5, -0.3, 0.25
    q = np.empty((mesh.ncell, N, N, 9))
    q[..., 0], q[..., 1], q[..., 2], q[..., 3] = rho, rho * u, rho * v, rho * w
    q[..., 4] = p / (GAM - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    q[..., 5], q[..., 6], q[..., 7], q[..., 8] = bx, by, bz, 0.1
    o.set_q(q)
    r = o.compresid()
    assert np.abs(r[..., :8]).max() < 1e-11
    # the cells are curved: the 8-node metric differs from the corners' bilinear one
    o4 = orc.Oracle(mesh, N)
    assert np.abs(o4.array("Jac") - o.array("Jac")).max() > 1e-4
