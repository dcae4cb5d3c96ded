import glob
import math
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "chorus-mhd_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_names():
    """The solver fixtures (the Tecplot fixtures plot_* and the metric tables
    mapder8_* have their own tests)."""
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("plot_", "mapder8_")))


def load_golden(name):
    """Fixture arrays plus the mesh rebuilt from its recipe (checked against the stored coordinates).
    Fixtures without the keys nx/ny/dom/gam/mu/eta0 are n x n on [0,2pi]^2 with gam=5/3, mu=eta0=0."""
    from chorus_mhd import quad_grid
    g = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    n = int(g["n"])
    nx, ny = int(g.get("nx", n)), int(g.get("ny", n))
    x0, x1, y0, y1 = g.get("dom", (0.0, 2 * math.pi, 0.0, 2 * math.pi))
    mesh = quad_grid(nx, ny, float(x0), float(x1), float(y0), float(y1), jitter=float(g["jitter"]),
                     seed=int(g["seed"]), bc=str(g["bc"]))
    assert np.array_equal(mesh.xv, g["xv"]) and np.array_equal(mesh.yv, g["yv"])
    g["mesh"] = mesh
    g.setdefault("gam", 5.0 / 3)
    g.setdefault("mu", 0.0)
    g.setdefault("eta0", 0.0)
    return g


def golden_params(g):
    """Keyword arguments of Solver / Oracle for a fixture."""
    return dict(artiv=int(g["artiv"]), k_stage=int(g["k_stage"]), gam=float(g["gam"]),
                mu=float(g["mu"]), eta0=float(g["eta0"]))


def split_topo(t):
    """Unpack topo.i32 written by oracle/ref_harness.f90 (mode setup)."""
    nc, nv, nf, ni, npair, nfree = [int(x) for x in t[:6]]
    off = 6 + 4 * nc
    if2c = t[off:off + 4 * nf].reshape(4, nf).T
    off += 4 * nf
    ibfinter = t[off:off + ni]
    off += ni
    cyc = t[off:off + 2 * npair].reshape(2, npair).T
    off += 2 * npair
    ibfree = t[off:off + nfree]
    return dict(nface=nf, if2c=if2c, ibfinter=ibfinter, cyc2fpair=cyc, ibfree=ibfree)


@pytest.fixture(scope="session")
def hip_available():
    try:
        import torch  # noqa: F401  (device probe only)
        return torch.cuda.is_available()
    except Exception:
        return False
