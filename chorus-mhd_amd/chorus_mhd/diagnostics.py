"""Integral diagnostics of a state on the host (field-loop energy decay).

The reference writes no integral diagnostics; these are workload reports.
Integrals use the interpolatory quadrature of the solution points (the
Chebyshev-Gauss points of setsandpoints, initsetup.f90:174-176) in each
direction and the Jacobian of the bilinear cell map (mapder.f90:14-26).
"""
from __future__ import annotations

import numpy as np

from .ic import solution_points


def quadrature_weights(N: int) -> np.ndarray:
    """Weights w on [0,1] with sum_i w_i x_i^k = 1/(k+1) for k < N (exact for
    polynomials of degree N-1 on the N solution points)."""
    x = solution_points(N)
    V = np.vander(x, N, increasing=True).T
    return np.linalg.solve(V, 1.0 / np.arange(1, N + 1))


def point_volumes(mesh, N: int) -> np.ndarray:
    """w_i w_j |J| at every solution point, shape (NCELL, N(js), N(is))."""
    xs = solution_points(N)
    xi = xs[None, :]
    eta = xs[:, None]
    X = mesh.xv[mesh.cells - 1]
    Y = mesh.yv[mesh.cells - 1]
    dxi = [-(1.0 - eta), (1.0 - eta), eta, -eta]          # d/dxi of the four shape functions
    deta = [-(1.0 - xi), -xi, xi, (1.0 - xi)]             # d/deta
    xx = sum(X[:, k, None, None] * dxi[k][None] for k in range(4))
    yx = sum(Y[:, k, None, None] * dxi[k][None] for k in range(4))
    xe = sum(X[:, k, None, None] * deta[k][None] for k in range(4))
    ye = sum(Y[:, k, None, None] * deta[k][None] for k in range(4))
    J = np.abs(xx * ye - xe * yx)
    w = quadrature_weights(N)
    return J * w[None, :, None] * w[None, None, :]


def magnetic_energy(mesh, N: int, Q: np.ndarray, vol: np.ndarray = None) -> float:
    """Integral of |B|^2 / 2 over the domain."""
    if vol is None:
        vol = point_volumes(mesh, N)
    b2 = Q[..., 5] ** 2 + Q[..., 6] ** 2 + Q[..., 7] ** 2
    return float(0.5 * np.sum(b2 * vol))
