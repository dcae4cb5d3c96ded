"""Boundary faces of the kinds the reference lists but never gives a flux
(INLE, OUTL, SYMP, WALL: connect_bdry.f90:88-110, compflux.f90:33-49), as the
oracle states them (the right state of the Rusanov flux: OUTL the cell's own
trace, SYMP/WALL the mirror of the normal momentum and normal B, INLE a
prescribed state).  CPU, the oracle only; pinned without its boundary code:

* ghost column: on a uniform mesh, the residual next to a left boundary of a
  kind equals the residual of the same cells in the mesh extended by one
  column of interior cells holding that kind's right state of a uniform
  interior state (the boundary face becomes an interior face);
* conservation: with WALL on every side the discrete integrals of density,
  energy and Bz have zero rate (the normal mass and energy fluxes of the
  mirror state vanish), while with FREE sides they do not;
* free stream: a uniform flow parallel to WALL faces is preserved.
"""
import math

import numpy as np
import pytest

import oracle as orc
from chorus_mhd import quad_grid
from chorus_mhd.diagnostics import quadrature_weights

GAM = 5.0 / 3


def _uniform(rho, u, v, w, p, bx, by, bz, az):
    e = p / (GAM - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    return np.array([rho, rho * u, rho * v, rho * w, e, bx, by, bz, az])


U = _uniform(1.2, 0.6, -0.25, 0.15, 0.9, 0.4, -0.3, 0.2, 0.7)
QIN = _uniform(1.5, 0.9, 0.1, -0.05, 1.1, 0.35, 0.25, -0.1, 0.0)


def _ghost_state(kind):
    if kind == "OUTL":
        return U.copy()
    if kind in ("SYMP", "WALL"):
        g = U.copy()        # left boundary: unit normal (-1, 0)
        g[1] = -g[1]
        g[5] = -g[5]
        return g
    g = QIN.copy()
    g[8] = U[8]
    return g


@pytest.mark.parametrize("kind", ["OUTL", "SYMP", "WALL", "INLE"])
@pytest.mark.parametrize("N", [3, 5])
def test_boundary_flux_equals_ghost_column(kind, N):
    n, m, lx, ly = 5, 4, 1.0, 0.8
    hx = lx / n
    bnd = quad_grid(n, m, 0.0, lx, 0.0, ly, bc=("CYCL", "FREE", "CYCL", kind))
    ext = quad_grid(n + 1, m, -hx, lx, 0.0, ly, bc=("CYCL", "FREE", "CYCL", "FREE"))
    ob = orc.Oracle(bnd, N, qin=QIN)
    oe = orc.Oracle(ext, N)
    qb = np.broadcast_to(U, (bnd.ncell, N, N, 9)).copy()
    qe = np.broadcast_to(U, (ext.ncell, N, N, 9)).copy()
    ghost = np.arange(ext.ncell) % (n + 1) == 0
    qe[ghost] = _ghost_state(kind)
    ob.set_q(qb)
    oe.set_q(qe)
    rb, re = ob.compresid(), oe.compresid()
    inner = re[~ghost]
    # relative to the residual of the boundary column; for OUTL (zero
    # gradient on a uniform state) both are the free stream's round-off zero
    scale = max(np.abs(inner).max(), 1.0)
    err = np.abs(rb - inner).max() / scale
    assert err < 1e-12, err
    if kind != "OUTL":
        # the boundary matters: the first column's residual is not the free stream's zero
        assert np.abs(rb[np.arange(bnd.ncell) % n == 0]).max() > 1e-3


def _integrals(o, mesh, N, r):
    """sum over cells and points of w_i w_j J resid_v (v = all nine)."""
    J = o.array("Jac").reshape(mesh.ncell, N, N)
    w = quadrature_weights(N)
    ww = w[None, :, None] * w[None, None, :]
    return (r * (J * ww)[..., None]).sum(axis=(0, 1, 2)), (np.abs(r) * np.abs(J * ww)[..., None]).sum(axis=(0, 1, 2))


@pytest.mark.parametrize("N", [3, 4, 6])
def test_wall_box_conserves_mass_energy_bz(N):
    mesh = quad_grid(6, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.12, seed=3, bc="WALL")
    free = quad_grid(6, 5, 0.0, 1.0, 0.0, 1.0, jitter=0.12, seed=3, bc="FREE")
    ow, of = orc.Oracle(mesh, N), orc.Oracle(free, N)
    ow.init_ot()                       # only for the solution points
    xy = ow.array("XXsolu").reshape(mesh.ncell, N, N, 2)
    x, y = xy[..., 0], xy[..., 1]
    q = np.empty((mesh.ncell, N, N, 9))
    rho = 1.0 + 0.3 * np.sin(3 * x) * np.cos(2 * y)
    u, v, w = 0.4 * np.cos(2 * y) + 0.2, -0.3 * np.sin(3 * x) + 0.1, 0.2 * np.cos(x + y)
    bx, by, bz = 0.3 + 0.1 * np.sin(y), -0.2 + 0.1 * np.cos(x), 0.25 * np.sin(x - y)
    p = 1.0 + 0.2 * np.cos(x) * np.sin(y)
    q[..., 0], q[..., 1], q[..., 2], q[..., 3] = rho, rho * u, rho * v, rho * w
    q[..., 4] = p / (GAM - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bx * bx + by * by + bz * bz)
    q[..., 5], q[..., 6], q[..., 7], q[..., 8] = bx, by, bz, 0.2 * x * y
    ow.set_q(q)
    of.set_q(q)
    sw, aw = _integrals(ow, mesh, N, ow.compresid())
    sf, af = _integrals(of, free, N, of.compresid())
    for v in (0, 4, 7):                # density, energy, Bz: no normal flux through a wall
        assert abs(sw[v]) < 1e-12 * aw[v], (v, sw[v], aw[v])
        assert abs(sf[v]) > 1e-4 * af[v], (v, sf[v], af[v])
    # the walls push: the momentum integrals are not zero
    assert abs(sw[1]) > 1e-4 * aw[1] and abs(sw[2]) > 1e-4 * aw[2]


def test_free_stream_along_walls():
    """Periodic in x, WALL at the bottom and top: a uniform flow along the
    walls with B along them, A_z linear with curl A_z = B: resid(1:8) = 0."""
    mesh = quad_grid(6, 5, 0.0, 1.2, 0.0, 1.0, jitter=0.15, seed=9, bc=("WALL", "CYCL", "WALL", "CYCL"))
    N = 4
    o = orc.Oracle(mesh, N)
    o.init_ot()
    xy = o.array("XXsolu").reshape(mesh.ncell, N, N, 2)
    q = np.broadcast_to(_uniform(1.3, 0.7, 0.0, 0.2, 0.9, 0.5, 0.0, 0.25, 0.0), (mesh.ncell, N, N, 9)).copy()
    q[..., 8] = 0.5 * xy[..., 1]
    o.set_q(q)
    r = o.compresid()
    assert np.abs(r[..., :8]).max() < 1e-12
