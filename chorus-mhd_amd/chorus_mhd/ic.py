"""Initial conditions and the CFL time-step rule of the workloads.

Orszag-Tang follows SETINITIALCOND (setinitialcond.f90:1-80): analytic state at
the solution points, B := curl A_z on the device (correct_B_via_curl_Az), then
the energy recomputed with the corrected B.  Rotor and field loop are new
workloads (the reference hard-codes only Orszag-Tang) built the same way.

The reference has no CFL logic (dt is an input, readinput.f90:24); this
package fixes dt = CFL * h_min / (lambda_max * N) with h_min the smallest
sqrt(cell area) and lambda_max = max(|u| + c_f) over the initial state.
"""
from __future__ import annotations

import math

import numpy as np


def solution_points(N: int) -> np.ndarray:
    """Chebyshev-Gauss points on [0,1] (setsandpoints, initsetup.f90:174-176)."""
    i = np.arange(1, N + 1, dtype=np.float64)
    return 0.5 * (1.0 - np.cos((2 * i - 1) * (math.atan(1.0) * 4.0) / (2 * float(N))))


def solution_xy(mesh, N: int) -> np.ndarray:
    """Physical solution-point coordinates (xyCoor_atGps + MapBaseFunc,
    mapder.f90:79-87, 255-273), shape (NCELL, N(js), N(is), 2)."""
    xs = solution_points(N)
    xi = xs[None, :]            # is along the last axis
    eta = xs[:, None]           # js
    f = [(1.0 - xi) * (1.0 - eta), (1.0 - eta) * xi, xi * eta, (1.0 - xi) * eta]
    X = mesh.xv[mesh.cells - 1]
    Y = mesh.yv[mesh.cells - 1]
    out = np.empty((mesh.ncell, N, N, 2))
    x = np.zeros((mesh.ncell, N, N))
    y = np.zeros((mesh.ncell, N, N))
    for k in range(4):
        x = x + X[:, k, None, None] * f[k][None]
        y = y + Y[:, k, None, None] * f[k][None]
    out[..., 0] = x
    out[..., 1] = y
    return out


def _energy(Q, p, gam):
    rho = Q[..., 0]
    ke = 0.5 * rho * ((Q[..., 1] / rho) ** 2 + (Q[..., 2] / rho) ** 2 + (Q[..., 3] / rho) ** 2)
    me = 0.5 * (Q[..., 5] ** 2 + Q[..., 6] ** 2 + Q[..., 7] ** 2)
    return p / (gam - 1) + ke + me


def energy(Q, p, gam):
    """Total energy of Q's density, momentum and B with pressure p."""
    return _energy(Q, p, gam)


def _with_ct(solver, Q, p, gam):
    """B := curl A_z on the device, then E from the corrected B."""
    solver.set_state(Q)
    solver.correct_b()
    Q = solver.get_state()
    Q[..., 4] = _energy(Q, p, gam)
    solver.set_state(Q)
    return Q


def orszag_tang(solver, mesh, N, gam=5.0 / 3):
    """Orszag-Tang vortex (setinitialcond.f90:16-35), B0 = 1 unless scaled."""
    return orszag_tang_scaled(solver, mesh, N, gam, 1.0)


def orszag_tang_scaled(solver, mesh, N, gam=5.0 / 3, b0=1.0):
    return _with_ct(solver, *orszag_tang_fields(mesh, N, gam, b0), gam)


def orszag_tang_fields(mesh, N, gam=5.0 / 3, b0=1.0):
    """(Q before the B := curl A_z correction, p) of the Orszag-Tang vortex."""
    xy = solution_xy(mesh, N)
    x, y = xy[..., 0], xy[..., 1]
    rho = np.full_like(x, gam * gam)
    p = np.full_like(x, gam)
    u, v = -np.sin(y), np.sin(x)
    Q = np.zeros(x.shape + (9,))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 5] = -b0 * np.sin(y)
    Q[..., 6] = b0 * np.sin(2 * x)
    Q[..., 8] = b0 * (0.5 * np.cos(2 * x) + np.cos(y))
    Q[..., 4] = _energy(Q, p, gam)
    return Q, p


def rotor(solver, mesh, N, gam=1.4, b0=5.0 / math.sqrt(4 * math.pi)):
    """MHD rotor on [0,1]^2 (Balsara-Spicer form): rho=10, u = 20(-(y-1/2), x-1/2)
    for r<0.1; in the ring 0.1<r<0.115 with f = (0.115-r)/0.015, rho = 1+9f and
    u = f*2*(-(y-1/2), x-1/2)/r (the disk's edge speed 2 tapered by f); outside
    rho=1, u=0.  p=1, Bx=B0 (A_z = B0*y), By=0."""
    return _with_ct(solver, *rotor_fields(mesh, N, gam, b0), gam)


def rotor_fields(mesh, N, gam=1.4, b0=5.0 / math.sqrt(4 * math.pi)):
    """(Q before the B := curl A_z correction, p) of the MHD rotor."""
    xy = solution_xy(mesh, N)
    x, y = xy[..., 0] - 0.5, xy[..., 1] - 0.5
    r = np.sqrt(x * x + y * y)
    r0, r1, u0 = 0.1, 0.115, 2.0
    fr = np.clip((r1 - r) / (r1 - r0), 0.0, 1.0)
    rho = 1.0 + 9.0 * fr
    # angular speed: u0/r0 inside the disk, f*u0/r in the ring, 0 outside
    omega = np.where(r < r0, u0 / r0, fr * u0 / np.maximum(r, r0))
    u, v = -omega * y, omega * x
    p = np.ones_like(x)
    Q = np.zeros(x.shape + (9,))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 5] = b0
    Q[..., 8] = b0 * xy[..., 1]
    Q[..., 4] = _energy(Q, p, gam)
    return Q, p


def field_loop(solver, mesh, N, gam=5.0 / 3, a0=1e-3, R=0.3, vel=(2.0, 1.0)):
    """Field-loop advection: rho=1, p=1, u=(2,1), A_z = a0 (R - r) for r < R."""
    return _with_ct(solver, *field_loop_fields(mesh, N, gam, a0, R, vel), gam)


def field_loop_fields(mesh, N, gam=5.0 / 3, a0=1e-3, R=0.3, vel=(2.0, 1.0)):
    """(Q before the B := curl A_z correction, p) of the field loop."""
    xy = solution_xy(mesh, N)
    x, y = xy[..., 0], xy[..., 1]
    r = np.sqrt(x * x + y * y)
    Q = np.zeros(x.shape + (9,))
    Q[..., 0] = 1.0
    Q[..., 1] = vel[0]
    Q[..., 2] = vel[1]
    Q[..., 8] = np.where(r < R, a0 * (R - r), 0.0)
    p = np.ones_like(x)
    Q[..., 4] = _energy(Q, p, gam)
    return Q, p


def cfl_dt(mesh, N, Q, gam, cfl=0.2) -> float:
    """dt = cfl * h_min / (lambda_max * N)."""
    X = mesh.xv[mesh.cells - 1]
    Y = mesh.yv[mesh.cells - 1]
    area = 0.5 * np.abs((X[:, 0] - X[:, 2]) * (Y[:, 1] - Y[:, 3]) - (X[:, 1] - X[:, 3]) * (Y[:, 0] - Y[:, 2]))
    h = float(np.sqrt(area.min()))
    rho = Q[..., 0]
    u = Q[..., 1] / rho
    v = Q[..., 2] / rho
    w = Q[..., 3] / rho
    B2 = Q[..., 5] ** 2 + Q[..., 6] ** 2 + Q[..., 7] ** 2
    p = (gam - 1) * (Q[..., 4] - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * B2)
    cf = np.sqrt((gam * p + B2) / rho)
    lam = float((np.sqrt(u * u + v * v) + cf).max())
    return cfl * h / (lam * N)
