"""Reference file formats: QUAD.INP, cell/vertex/boundary mesh files, restart dumps.

Every reader/writer here follows the reference's own I/O routines so that the
reference Fortran program and this package read byte-identical inputs:

* ``QUAD.INP``      -- readinput.f90:6-38 (positional, list-directed, 29 lines)
* cells file        -- readGRIDinfo.f90:33 (``IC IV1 IV2 IV3 IV4 ICTYPE ICKEY``)
* vertices file     -- readGRIDinfo.f90:79-80 (``FORMAT(I9,6X,3(D20.12,1X))``)
* boundary file     -- readGRIDinfo.f90:237-238 (``IB IV1 IV2 IREGION IPATCH TYPE``)
* restart dump      -- write_data.f90:44-74 / read_data_restart.f90:35-60
  (int32 iter, f64 ctime, then per cell int32 global id + Q(1:9,1:N,1:N))
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

# positional layout of QUAD.INP (readinput.f90:8-36); None marks a skipped line
_QUAD_LINES = [
    None, None, "NAMECEL", "NAMEVRT", "NAMEVRT8", "NAMEBND", "METIS_CELL",
    "NAMERESTART", None, "N", "CURVE_WALL", "vismode", "mu", "eta0", None,
    "k_stage", "dt", "MAXITER", None, "DXCYCL", "DYCYCL", None, "ARTIV", None,
    "NRE", "nwrite", None, "restart", "restart_ord",
]


@dataclass
class QuadInp:
    """Contents of ``QUAD.INP`` (readinput.f90:6-38)."""
    NAMECEL: str = "cells.dat"
    NAMEVRT: str = "verts.dat"
    NAMEVRT8: str = "verts8.dat"
    NAMEBND: str = "bound.dat"
    METIS_CELL: str = "metis.dat"
    NAMERESTART: str = "restart.dat"
    N: int = 4
    CURVE_WALL: int = 0
    vismode: int = 1
    mu: float = 0.0
    eta0: float = 0.0
    k_stage: int = 3
    dt: float = 1e-3
    MAXITER: int = 1
    DXCYCL: float = 0.0
    DYCYCL: float = 0.0
    ARTIV: int = 0
    NRE: int = 1
    nwrite: int = 1000000
    restart: int = 0
    restart_ord: int = 4

    def write(self, path: str) -> None:
        out = []
        for i, key in enumerate(_QUAD_LINES):
            if key is None:
                out.append("# line %d" % (i + 1))
                continue
            val = getattr(self, key)
            if isinstance(val, str):
                # quoted: a '/' in an unquoted list-directed string ends the read
                out.append("'%s'" % val)
            elif isinstance(val, float):
                out.append(_f17(val))
            else:
                out.append("%d" % val)
        with open(path, "w") as f:
            f.write("\n".join(out) + "\n")

    @classmethod
    def read(cls, path: str) -> "QuadInp":
        with open(path) as f:
            lines = f.read().splitlines()
        q = cls()
        for i, key in enumerate(_QUAD_LINES):
            if key is None:
                continue
            tok = lines[i].split()[0] if lines[i].split() else ""
            cur = getattr(q, key)
            if isinstance(cur, str):
                setattr(q, key, tok.strip("'\""))
            elif isinstance(cur, float):
                setattr(q, key, float(tok.replace("d", "e").replace("D", "e")))
            else:
                setattr(q, key, int(tok))
        return q


def _f17(x: float) -> str:
    """Shortest repr that round-trips (Fortran list-directed reads it exactly)."""
    return repr(float(x))


def fmt_d20(x: float) -> str:
    """One 20-character field readable by the ``D20.12`` edit descriptor.

    The field carries an explicit decimal point, so Fortran ignores the ``.12``
    and reads every digit written; both the reference and this package then
    parse the same text to the same double.
    """
    ax = abs(x)
    ndig = len(str(int(ax))) if ax >= 1 else 1
    prec = 20 - 2 - ndig  # sign + point
    s = "%20.*f" % (prec, x)
    if len(s) != 20:
        raise ValueError("coordinate %r does not fit a 20-character field" % x)
    return s


def text_round_trip(a: np.ndarray) -> np.ndarray:
    """The doubles the reference reads back from the 20-character fields of
    ``fmt_d20`` (vectorised).  A field carries 18 - d fractional digits for a
    value with d integer digits; when |x| >= 2**-4 the double's half-ulp
    (>= 2**-57 * 10**(d-1)) exceeds half the field's decimal spacing, so the
    text parses back to x itself.  Only smaller nonzero values go through the
    formatted text."""
    a = np.asarray(a, dtype=np.float64)
    out = a.copy()
    small = np.flatnonzero((np.abs(a) < 0.0625) & (a != 0.0))
    for i in small:
        out[i] = float(fmt_d20(a[i]))
    return out


@dataclass
class Mesh:
    """Unstructured quad mesh in the reference's global numbering (1-based ids).

    ``cells[ic-1] = (iv1, iv2, iv3, iv4)``, counter-clockwise (connectivity.f90:236-283);
    ``bnd[ib-1] = (iv1, iv2)`` with ``bnd_type[ib-1]`` in CYCL/FREE/...
    """
    xv: np.ndarray                  # (NVERT,) float64
    yv: np.ndarray                  # (NVERT,) float64
    cells: np.ndarray               # (NCELL, 4) int32, 1-based vertex ids
    bnd: np.ndarray                 # (NBND, 2) int32, 1-based vertex ids
    bnd_type: list = field(default_factory=list)
    dxcycl: float = 0.0
    dycycl: float = 0.0

    @property
    def ncell(self) -> int:
        return int(self.cells.shape[0])

    @property
    def nvert(self) -> int:
        return int(self.xv.shape[0])

    def write(self, directory: str, namecel="cells.dat", namevrt="verts.dat",
              namebnd="bound.dat") -> None:
        os.makedirs(directory, exist_ok=True)
        nc = self.ncell
        ids = np.arange(1, nc + 1)
        with open(os.path.join(directory, namecel), "w") as f:
            for ic in range(nc):
                c = self.cells[ic]
                f.write("%d %d %d %d %d 1 1\n" % (ids[ic], c[0], c[1], c[2], c[3]))
        with open(os.path.join(directory, namevrt), "w") as f:
            zero = fmt_d20(0.0)
            for iv in range(self.nvert):
                f.write("%9d      %s %s %s \n" % (iv + 1, fmt_d20(self.xv[iv]),
                                                  fmt_d20(self.yv[iv]), zero))
        with open(os.path.join(directory, namebnd), "w") as f:
            for ib in range(self.bnd.shape[0]):
                f.write("%d %d %d 1 1 %s\n" % (ib + 1, self.bnd[ib, 0], self.bnd[ib, 1],
                                               self.bnd_type[ib]))

    def glob2local(self) -> "Mesh":
        """The mesh in the reference's local numbering of one rank (GLOB2LOCAL,
        readGRIDinfo.f90:124-199): vertices renumbered by first appearance in
        the cells, in cell order.  The module arrays XV, YV, IVCELL a driver
        passes to the library are in this numbering (it matters for the Tecplot
        output, whose vertex nodes keep the vertex numbers)."""
        new = np.zeros(self.nvert + 1, dtype=np.int64)
        order = []
        for c in self.cells:
            for v in c:
                if new[v] == 0:
                    order.append(v)
                    new[v] = len(order)
        idx = np.array(order) - 1
        cells = new[self.cells].astype(np.int32)
        bnd = np.where(new[self.bnd] > 0, new[self.bnd], 0).astype(np.int32)
        return Mesh(self.xv[idx].copy(), self.yv[idx].copy(), cells, bnd, list(self.bnd_type),
                    self.dxcycl, self.dycycl)

    def reparsed(self) -> "Mesh":
        """The mesh as the reference reads it back from the files (text round trip)."""
        xv = text_round_trip(self.xv)
        yv = text_round_trip(self.yv)
        return Mesh(xv, yv, self.cells.copy(), self.bnd.copy(), list(self.bnd_type),
                    self.dxcycl, self.dycycl)


def read_mesh(directory: str, q: QuadInp) -> Mesh:
    """Read the three mesh files the way readGRIDinfo.f90 does (single rank)."""
    cells = []
    with open(os.path.join(directory, q.NAMECEL)) as f:
        for line in f:
            t = line.split()
            if t:
                cells.append([int(t[1]), int(t[2]), int(t[3]), int(t[4])])
    xv, yv = [], []
    with open(os.path.join(directory, q.NAMEVRT)) as f:
        for line in f:
            if not line.strip():
                continue
            xv.append(float(line[15:35].replace("D", "E").replace("d", "e")))
            yv.append(float(line[36:56].replace("D", "E").replace("d", "e")))
    bnd, btype = [], []
    with open(os.path.join(directory, q.NAMEBND)) as f:
        for line in f:
            t = line.split()
            if t:
                bnd.append([int(t[1]), int(t[2])])
                btype.append(t[5].strip("'\""))
    return Mesh(np.array(xv), np.array(yv), np.array(cells, dtype=np.int32),
                np.array(bnd, dtype=np.int32).reshape(-1, 2), btype, q.DXCYCL, q.DYCYCL)


def write_restart(path: str, it: int, ctime: float, Q: np.ndarray) -> None:
    """Write ``restart.%08d.dat`` (write_data.f90:44-74); ``Q`` is (NCELL, N, N, 9) C-order,
    i.e. Fortran ``Q(9,N,N,NCELL)`` memory order."""
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    nc = Q.shape[0]
    rec = np.dtype([("icg", "<i4"), ("q", "<f8", (Q[0].size,))])
    body = np.empty(nc, dtype=rec)
    body["icg"] = np.arange(1, nc + 1, dtype=np.int32)
    body["q"] = Q.reshape(nc, -1)
    with open(path, "wb") as f:
        f.write(struct.pack("<id", int(it), float(ctime)))
        f.write(body.tobytes())


def read_restart(path: str, N: int) -> tuple:
    """Read a restart dump (read_data_restart.f90:35-60). Returns (iter, ctime, Q[NCELL,N,N,9])."""
    with open(path, "rb") as f:
        raw = f.read()
    it, ctime = struct.unpack_from("<id", raw, 0)
    rec = np.dtype([("icg", "<i4"), ("q", "<f8", (9 * N * N,))])
    body = np.frombuffer(raw, dtype=rec, offset=12)
    nc = body.shape[0]
    Q = np.empty((nc, N, N, 9))
    Q[body["icg"] - 1] = body["q"].reshape(nc, N, N, 9)
    return it, ctime, Q
