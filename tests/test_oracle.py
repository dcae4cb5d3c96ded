"""Pin the C oracle (oracle/chorus_oracle.c) before trusting it.

* bitwise against fixtures written by the compiled, unmodified reference
  (tests/golden/make_golden.py): topology, initial condition, residual, AV
  sensor, multi-step state (both SSP steppers), diagnostics;
* closed forms: Chebyshev / Gauss-Legendre points, polynomial exactness of the
  extrapolation and derivative operators, free-stream preservation on a
  jittered mesh, round-off div B after the CT correction on a uniform mesh.
"""
import math

import numpy as np
import pytest

from conftest import golden_names, golden_params, load_golden, split_topo

import oracle as orc

NAMES = golden_names()


@pytest.fixture(scope="module", autouse=True)
def _build():
    orc.build()


def _oracle(g):
    return orc.Oracle(g["mesh"], int(g["N"]), dt=float(g["dt"]), **golden_params(g))


@pytest.mark.parametrize("name", NAMES)
def test_topology_bitwise(name):
    g = load_golden(name)
    t = split_topo(g["topo"])
    o = orc.Topology(g["mesh"])
    assert o.nface == t["nface"]
    assert np.array_equal(o.if2c, t["if2c"])
    assert np.array_equal(o.ibfinter, t["ibfinter"])
    assert np.array_equal(o.cyc2fpair, t["cyc2fpair"])
    assert np.array_equal(o.ibfree, t["ibfree"])


@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("ot_")])
def test_orszag_tang_ic_bitwise(name):
    g = load_golden(name)
    o = _oracle(g)
    o.init_ot()
    assert np.array_equal(o.get_q(), g["q0"])


@pytest.mark.parametrize("name", NAMES)
def test_residual_bitwise(name):
    g = load_golden(name)
    o = _oracle(g)
    o.set_q(g["q0"])
    r = o.compresid()
    assert np.array_equal(r, g["resid"])
    if int(g["artiv"]):
        assert o.scalar("eigvmax") == float(g["eigvmax"])
        assert o.scalar("epsilon0") == float(g["epsilon0"])
        nc = g["mesh"].ncell
        sm = o.array("SM_IND").reshape(nc, 9)
        mu = o.array("muAV").reshape(nc, 9)
        # NaN where the sensor divides 0 by 0 (AV.f90:64-70): compare with NaNs equal
        assert np.array_equal(sm, g["sm_ind"], equal_nan=True)
        assert np.array_equal(mu, g["muav"], equal_nan=True)


@pytest.mark.parametrize("name", NAMES)
def test_steps_bitwise(name):
    g = load_golden(name)
    o = _oracle(g)
    o.set_q(g["q0"])
    o.steps(int(g["steps"]))
    assert np.array_equal(o.get_q(), g["qend"])


@pytest.mark.parametrize("name", NAMES)
def test_diagnostics_first_iteration(name):
    """iterations.f90:66-67 norms printed at iter 1, bit for bit."""
    g = load_golden(name)
    o = _oracle(g)
    o.set_q(g["q0"])
    r, d, _ = o.steps(1)
    assert r == float(g["diag_resid"])
    assert d == float(g["diag_divb"])


# ---------------------------------------------------------------- closed forms
def _ops(N):
    from chorus_mhd import quad_grid
    m = quad_grid(2, 2, 0.0, 1.0, 0.0, 1.0, bc="FREE")
    o = orc.Oracle(m, N)
    return (o.array("Xs"), o.array("Xf"), o.array("Lmat").reshape(N, N + 1).T,
            o.array("Mmat").reshape(N, N + 1).T)


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7, 8])
def test_points_closed_form(N):
    Xs, Xf, _, _ = _ops(N)
    k = np.arange(1, N + 1)
    assert np.allclose(Xs, 0.5 * (1 - np.cos((2 * k - 1) * np.pi / (2 * N))), atol=0, rtol=1e-15)
    if N >= 2:
        gl = np.sort((np.polynomial.legendre.leggauss(N - 1)[0] + 1) / 2)
        # closed forms up to N=6, 12-digit tables from N=7 (initsetup.f90:211-252)
        tol = 1e-15 if N <= 6 else 1e-12
        assert np.allclose(Xf[1:-1], gl, atol=tol, rtol=0)
    assert Xf[0] == 0.0 and Xf[-1] == 1.0


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7, 8])
def test_operators_polynomial_exactness(N):
    Xs, Xf, L, M = _ops(N)
    for p in range(N):          # extrapolation exact up to degree N-1
        assert np.allclose(L @ Xs ** p, Xf ** p, atol=2e-14)
    for p in range(N + 1):      # derivative of flux-point data exact up to degree N
        d = p * Xs ** (p - 1) if p else np.zeros(N)
        assert np.allclose(M.T @ Xf ** p, d, atol=1e-12 * N ** 2)
    # a dropped term or transposed operand breaks these
    assert not np.allclose(L.T[:, :N] @ Xs ** 1, Xf[:N])


def test_free_stream_preservation_jittered():
    """Uniform flow, uniform B and A_z linear with curl A_z = B on a jittered FREE
    mesh: resid(1:8) vanishes to round-off, resid(9) = u*By - v*Bx exactly."""
    from chorus_mhd import quad_grid
    m = quad_grid(6, 5, 0.0, 1.0, 0.0, 0.8, jitter=0.15, seed=3, bc="FREE")
    N = 4
    o = orc.Oracle(m, N)
    o.init_ot()                  # only to obtain the solution points below
    xy = o.array("XXsolu").reshape(m.ncell, N, N, 2)
    gam = 5.0 / 3
    rho, u, v, w, p, Bx, By, Bz = 1.3, 0.7, -0.4, 0.2, 0.9, 0.5, -0.3, 0.25
    Q = np.empty((m.ncell, N, N, 9))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 3] = rho * w
    Q[..., 4] = p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (Bx * Bx + By * By + Bz * Bz)
    Q[..., 5], Q[..., 6], Q[..., 7] = Bx, By, Bz
    Q[..., 8] = Bx * xy[..., 1] - By * xy[..., 0]
    o.set_q(Q)
    R = o.compresid()
    assert np.abs(R[..., :8]).max() < 1e-12
    assert np.allclose(R[..., 8], u * By - v * Bx, atol=1e-12, rtol=0)
    # the CT correction reproduces B from the linear A_z
    o.correct_b()
    Qc = o.get_q()
    assert np.allclose(Qc[..., 5], Bx, atol=1e-12) and np.allclose(Qc[..., 6], By, atol=1e-12)


def test_divb_roundoff_uniform_mesh():
    """On an affine mesh B = curl_h A_z is discretely divergence-free."""
    g = load_golden("ot_n8_N4_uniform")
    o = _oracle(g)
    o.set_q(g["q0"])
    _, divb, divbmax = o.steps(2)
    assert divb < 1e-13 and divbmax < 1e-12
