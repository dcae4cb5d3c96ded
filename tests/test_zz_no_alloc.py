"""chorus_diag_noalloc (no device allocation; partials in the library's fixed
per-GPU scratch array) against chorus_diag on the same contexts: div B RMS and
max bit for bit, the residual norm within 1e-13 relative (its partial sums are
added in another order), wave speed and epsilon0 equal.  Two contexts in a
row, the second with a smaller mesh and order, so the second call reuses the
array the first one filled."""
import math
import sys

import numpy as np
import pytest

import oracle as orc
from chorus_mhd import quad_grid
from chorus_mhd.solver import Solver

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(s, label):
    a = s.diag()
    b = s.diag_noalloc()
    _say(label, "diag", a, "noalloc", b)
    for k in ("divbglob", "divbmax", "eigvmax", "epsilon0"):
        assert (a[k] == b[k]) or (math.isnan(a[k]) and math.isnan(b[k])), (label, k, a[k], b[k])
    if math.isnan(a["residglob"]):
        assert math.isnan(b["residglob"]), label
    else:
        assert b["residglob"] == pytest.approx(a["residglob"], rel=1e-13, abs=0.0), label
    return b


# (n, m, N, jitter, bc, artiv, precision, steps): the second case is smaller
CASES = [(24, 20, 5, 0.10, "CYCL", 1, "fp64", 3), (8, 6, 3, 0.08, "FREE", 0, "mixed", 2)]


def test_diag_noalloc_equals_diag():
    orc.build()
    for n, m, N, jit, bc, av, prec, steps in CASES:
        label = "%dx%d N=%d %s artiv=%d %s" % (n, m, N, bc, av, prec)
        _say("case", label)
        mesh = quad_grid(n, m, 0.0, 2 * math.pi, 0.0, 2 * math.pi, jitter=jit, seed=n + m, bc=bc)
        o = orc.Oracle(mesh, N, dt=1e-3)
        o.init_ot()
        s = Solver(mesh, N, 1e-3, artiv=av, precision=prec)
        s.set_state(o.get_q())
        d0 = _check(s, label + " before any stage")
        assert math.isnan(d0["residglob"])
        s.step(steps)
        d1 = _check(s, label + " after %d steps" % steps)
        assert math.isfinite(d1["residglob"]) and d1["divbmax"] > 0.0
        s.close()
