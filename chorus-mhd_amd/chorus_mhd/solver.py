"""Host-side mirror of the reference's hot-path subroutines over the C-ABI.

``Solver`` keeps the reference's vocabulary: ``compresid`` (compresid.f90),
``correct_b`` (correct_B_via_curl_Az, setinitialcond.f90:82-92), ``step``
(the loop body of iterations.f90:20-33 with RK_S5P3 or FiveStageRk) and
``diag`` (the norms of iterations.f90:35-67).  States are numpy arrays of shape
(NCELL, N, N, 9), i.e. the memory order of Fortran ``Q(9,N,N,NCELL)``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as _lib
from .formats import Mesh

BTYPES = {"CYCL": 1, "FREE": 2}


class ChorusError(RuntimeError):
    pass


@dataclass
class Topology:
    """IF2C(NFACE,4), IBFINTER, CYC2FPAIR(NPAIR,2), IBFREE of one rank (1-based)."""
    nface: int
    if2c: np.ndarray        # (NFACE, 4)
    ibfinter: np.ndarray
    cyc2fpair: np.ndarray   # (NPAIR, 2)
    ibfree: np.ndarray


def topology(mesh: Mesh) -> Topology:
    """CONNECTIVITY + CONNECT_BDRY + MAPCYCLOC of the reference, via the library."""
    L = _lib.load()
    nc = mesh.ncell
    iv = np.asfortranarray(mesh.cells, dtype=np.int32)
    bnd = np.asfortranarray(mesh.bnd, dtype=np.int32)
    bt = np.array([BTYPES.get(t, 0) for t in mesh.bnd_type], dtype=np.int32)
    xv = np.ascontiguousarray(mesh.xv, dtype=np.float64)
    yv = np.ascontiguousarray(mesh.yv, dtype=np.float64)
    if2c = np.zeros(16 * nc, dtype=np.int32)
    ibfinter = np.zeros(4 * nc, dtype=np.int32)
    cyc = np.zeros(8 * nc, dtype=np.int32)
    ibfree = np.zeros(4 * nc, dtype=np.int32)
    nface, ninter, npair, nfree = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    ip = _lib._ip
    rc = L.chorus_topology(nc, mesh.nvert, _lib.ptr(iv, ip), _lib.ptr(xv), _lib.ptr(yv),
                           bnd.shape[0], _lib.ptr(bnd, ip), _lib.ptr(bt, ip),
                           mesh.dxcycl, mesh.dycycl, C.byref(nface), _lib.ptr(if2c, ip),
                           C.byref(ninter), _lib.ptr(ibfinter, ip), C.byref(npair),
                           _lib.ptr(cyc, ip), C.byref(nfree), _lib.ptr(ibfree, ip))
    if rc != 0:
        raise ChorusError("chorus_topology failed (%d)" % rc)
    nf, npr = nface.value, npair.value
    return Topology(nf, if2c[:4 * nf].reshape(4, nf).T.copy(), ibfinter[:ninter.value].copy(),
                    cyc[:2 * npr].reshape(2, npr).T.copy(), ibfree[:nfree.value].copy())


BC_KINDS = ("INLE", "OUTL", "SYMP", "WALL")


def boundary_lists(mesh: Mesh, t: Topology) -> dict:
    """IBFINL, IBFOUT, IBFSYMP, IBFWAL of CONNECT_BDRY (connect_bdry.f90:186-205):
    the 1-based face numbers of the mesh's INLE/OUTL/SYMP/WALL boundary
    entries, in boundary-file order.  A boundary face has no right cell in
    IF2C; its vertices are those of the left cell's local face (face k joins
    vertices k and k+1, connectivity.f90:73-83)."""
    face_of = {}
    for f in range(t.nface):
        if t.if2c[f, 1] == 0:
            lc, lf = int(t.if2c[f, 0]) - 1, int(t.if2c[f, 2])
            face_of[frozenset((int(mesh.cells[lc, lf - 1]), int(mesh.cells[lc, lf % 4])))] = f + 1
    out = {k: [] for k in BC_KINDS}
    for ib, kind in enumerate(mesh.bnd_type):
        if kind in out:
            out[kind].append(face_of[frozenset((int(mesh.bnd[ib, 0]), int(mesh.bnd[ib, 1])))])
    return {k: np.array(v, dtype=np.int32) for k, v in out.items()}


PRECISION = {"fp64": _lib.CHORUS_FP64, "mixed": _lib.CHORUS_MIXED}


def plot_topology(mesh: Mesh, nre: int):
    """Host-only plot nodes of the Tecplot output (chorus_plot_topology) of a mesh
    in file numbering (renumbered as the reference's GLOB2LOCAL does first):
    (xy (2, nnode), c2n (ncellp, 4) 1-based)."""
    local = mesh.glob2local()
    m, arrs = _c_mesh(local, 2, topology(local))
    cap = (nre + 1) ** 2 * mesh.ncell
    xy = np.zeros(2 * cap)
    c2n = np.zeros(4 * nre * nre * mesh.ncell, dtype=np.int32)
    nn, ncp = C.c_int(), C.c_int()
    rc = _lib.load().chorus_plot_topology(C.byref(m), nre, C.byref(nn), C.byref(ncp), _lib.ptr(xy),
                                          _lib.ptr(c2n, _lib._ip))
    if rc != 0:
        raise ChorusError("chorus_plot_topology failed (%d)" % rc)
    n = nn.value
    return np.stack([xy[:n], xy[n:2 * n]]), c2n[:4 * ncp.value].reshape(ncp.value, 4)


def _c_mesh(mesh: Mesh, N: int, t: Topology):
    """ChorusMesh over arrays kept alive in the returned dict (the library copies them)."""
    a = dict(
        xv=np.ascontiguousarray(mesh.xv, dtype=np.float64),
        yv=np.ascontiguousarray(mesh.yv, dtype=np.float64),
        iv=np.asfortranarray(mesh.cells, dtype=np.int32).ravel(order="F"),
        if2c=np.asfortranarray(t.if2c, dtype=np.int32).ravel(order="F"),
        ibfinter=np.ascontiguousarray(t.ibfinter, dtype=np.int32),
        cyc=np.asfortranarray(t.cyc2fpair, dtype=np.int32).ravel(order="F"),
        ibfree=np.ascontiguousarray(t.ibfree, dtype=np.int32),
    )
    ip = _lib._ip
    m = _lib.ChorusMesh(mesh.ncell, mesh.nvert, N, _lib.ptr(a["xv"]), _lib.ptr(a["yv"]),
                        _lib.ptr(a["iv"], ip), t.nface, _lib.ptr(a["if2c"], ip),
                        len(t.ibfinter), _lib.ptr(a["ibfinter"], ip),
                        t.cyc2fpair.shape[0], _lib.ptr(a["cyc"], ip),
                        len(t.ibfree), _lib.ptr(a["ibfree"], ip))
    return m, a


class Solver:
    """One MHD instance resident on the GPU."""

    def __init__(self, mesh: Mesh, N: int, dt: float, *, gam: float = 5.0 / 3, artiv: int = 0,
                 k_stage: int = 3, mu: float = 0.0, eta0: float = 0.0, precision: str = "fp64",
                 topo: Topology = None, qin=None):
        L = _lib.load()
        self.mesh, self.N = mesh, N
        self.topo = topo if topo is not None else topology(mesh)
        t = self.topo
        # keep every array alive while the context is created (it copies them)
        self._arrs = dict(
            xv=np.ascontiguousarray(mesh.xv, dtype=np.float64),
            yv=np.ascontiguousarray(mesh.yv, dtype=np.float64),
            iv=np.asfortranarray(mesh.cells, dtype=np.int32).ravel(order="F"),
            if2c=np.asfortranarray(t.if2c, dtype=np.int32).ravel(order="F"),
            ibfinter=np.ascontiguousarray(t.ibfinter, dtype=np.int32),
            cyc=np.asfortranarray(t.cyc2fpair, dtype=np.int32).ravel(order="F"),
            ibfree=np.ascontiguousarray(t.ibfree, dtype=np.int32),
        )
        a = self._arrs
        ip = _lib._ip
        m = _lib.ChorusMesh(mesh.ncell, mesh.nvert, N, _lib.ptr(a["xv"]), _lib.ptr(a["yv"]),
                            _lib.ptr(a["iv"], ip), t.nface, _lib.ptr(a["if2c"], ip),
                            len(t.ibfinter), _lib.ptr(a["ibfinter"], ip),
                            t.cyc2fpair.shape[0], _lib.ptr(a["cyc"], ip),
                            len(t.ibfree), _lib.ptr(a["ibfree"], ip))
        prec = {"fp64": _lib.CHORUS_FP64, "mixed": _lib.CHORUS_MIXED}[precision]
        p = _lib.ChorusParams(1, artiv, k_stage, mu, eta0, gam, dt, mesh.dxcycl, mesh.dycycl,
                              mesh.ncell, prec)
        self._ctx = C.c_void_p()
        if any(k in BC_KINDS for k in mesh.bnd_type):
            # INLE / OUTL / SYMP / WALL faces (chorus_init_bc)
            bl = boundary_lists(mesh, t)
            self._arrs.update({"bc_" + k: v for k, v in bl.items()})
            ip = _lib._ip
            q = (C.c_double * 9)(*([0.0] * 9 if qin is None else [float(x) for x in qin]))
            b = _lib.ChorusBC(len(bl["INLE"]), _lib.ptr(bl["INLE"], ip), len(bl["OUTL"]), _lib.ptr(bl["OUTL"], ip),
                              len(bl["SYMP"]), _lib.ptr(bl["SYMP"], ip), len(bl["WALL"]), _lib.ptr(bl["WALL"], ip), q)
            rc = L.chorus_init_bc(C.byref(m), C.byref(p), C.byref(b), C.byref(self._ctx))
        else:
            rc = L.chorus_init(C.byref(m), C.byref(p), C.byref(self._ctx))
        if rc != 0:
            msg = L.chorus_last_error(self._ctx).decode()
            L.chorus_free(self._ctx)
            self._ctx = None
            raise ChorusError("chorus_init: %s (%d)" % (msg, rc))

    # ---- plumbing -------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            raise ChorusError("%s: %s (%d)" % (what, _lib.load().chorus_last_error(self._ctx).decode(), rc))

    def close(self):
        if getattr(self, "_ctx", None):
            _lib.load().chorus_free(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.mesh.ncell, self.N, self.N, 9)

    # ---- reference subroutines -----------------------------------------
    def set_state(self, Q: np.ndarray) -> None:
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.shape)
        self._check(_lib.load().chorus_set_state(self._ctx, _lib.ptr(Q)), "chorus_set_state")

    def get_state(self) -> np.ndarray:
        Q = np.empty(self.shape)
        self._check(_lib.load().chorus_get_state(self._ctx, _lib.ptr(Q)), "chorus_get_state")
        return Q

    def compresid(self) -> np.ndarray:
        R = np.empty(self.shape)
        self._check(_lib.load().chorus_compresid(self._ctx, _lib.ptr(R)), "chorus_compresid")
        return R

    def correct_b(self) -> None:
        self._check(_lib.load().chorus_correct_b(self._ctx), "chorus_correct_b")

    def set_nodes8(self, xm: np.ndarray, ym: np.ndarray) -> None:
        """Curved cells: the edge nodes XM(NCELL,4), YM(NCELL,4) of the 8-node
        map (MAPDER_8_NODE, mapder.f90:117-253), column k the node on edge k."""
        shape = (self.mesh.ncell, 4)
        xm = np.asfortranarray(np.asarray(xm, dtype=np.float64).reshape(shape)).ravel(order="F")
        ym = np.asfortranarray(np.asarray(ym, dtype=np.float64).reshape(shape)).ravel(order="F")
        self._check(_lib.load().chorus_set_nodes8(self._ctx, _lib.ptr(xm), _lib.ptr(ym)), "chorus_set_nodes8")

    def step(self, nsteps: int = 1) -> None:
        self._check(_lib.load().chorus_step(self._ctx, nsteps), "chorus_step")

    def diag(self) -> dict:
        v = [C.c_double() for _ in range(5)]
        self._check(_lib.load().chorus_diag(self._ctx, *[C.byref(x) for x in v]), "chorus_diag")
        return dict(zip(["residglob", "divbglob", "divbmax", "eigvmax", "epsilon0"],
                        [x.value for x in v]))

    def diag_noalloc(self) -> dict:
        """diag() through chorus_diag_noalloc: no device allocation, partials in
        the library's fixed per-GPU scratch array."""
        v = [C.c_double() for _ in range(5)]
        self._check(_lib.load().chorus_diag_noalloc(self._ctx, *[C.byref(x) for x in v]), "chorus_diag_noalloc")
        return dict(zip(["residglob", "divbglob", "divbmax", "eigvmax", "epsilon0"],
                        [x.value for x in v]))

    def plot_setup(self, nre: int) -> None:
        """Plot-node topology of the Tecplot output with NRE x NRE plot cells per
        cell (tecplotter2dsetup); call before stepping (see chorus_plot_setup)."""
        local = self.mesh.glob2local()
        m, arrs = _c_mesh(local, self.N, topology(local))
        self._plot_arrs = arrs
        self._check(_lib.load().chorus_plot_setup(self._ctx, C.byref(m), int(nre)), "chorus_plot_setup")

    def write_tecplot(self, path: str, ctime: float) -> None:
        """Tecplot TDV111 file of the current state (tecplotter2d + tecpost_parallel)."""
        self._check(_lib.load().chorus_write_tecplot(self._ctx, path.encode(), float(ctime)), "chorus_write_tecplot")

    def write_restart(self, path: str, it: int, ctime: float, icg=None) -> None:
        """Restart dump of the current state (WRITE_ALL_DATA_PARALLEL,
        write_data.f90:5-83), packed on the device; icg: global cell numbers
        (default 1..NCELL)."""
        ip = None
        if icg is not None:
            icg = np.ascontiguousarray(icg, dtype=np.int32)
            if icg.shape != (self.mesh.ncell,):
                raise ValueError("icg needs one entry per cell")
            ip = icg.ctypes.data_as(C.POINTER(C.c_int))
        self._check(_lib.load().chorus_write_restart(self._ctx, path.encode(), int(it), float(ctime), ip),
                    "chorus_write_restart")

    @staticmethod
    def restart_name(it: int) -> str:
        """The reference's restart file name (write_data.f90:44)."""
        return "restart.%08d.dat" % it

    @staticmethod
    def tecplot_name(ctime: float) -> str:
        """The reference's file name for solution time ctime (tecplotter.f90:190-191):
        'TEC' + int(ctime) as 7 digits + the fraction as '.ddd' + '.plt'."""
        frac = "%6.3f" % (ctime - int(ctime))
        return "TEC%07d%s.plt" % (int(ctime), frac[2:6])

    def autotune(self) -> int:
        """Pick the fastest tile shape of the mixed-form cell kernels on this
        mesh and state (no state change); returns its index."""
        t = C.c_int()
        self._check(_lib.load().chorus_autotune(self._ctx, C.byref(t)), "chorus_autotune")
        return t.value

    def set_tile(self, tile: int) -> None:
        """Fix the tile shape of the mixed-form cell kernels (see autotune)."""
        self._check(_lib.load().chorus_set_tile(self._ctx, int(tile)), "chorus_set_tile")

    def time_steps(self, nsteps: int) -> dict:
        tot, st, ct = C.c_double(), C.c_double(), C.c_double()
        self._check(_lib.load().chorus_time_steps(self._ctx, nsteps, C.byref(tot), C.byref(st),
                                                  C.byref(ct)), "chorus_time_steps")
        return dict(ms_total=tot.value, ms_stage_avg=st.value, ms_ct_avg=ct.value)


@dataclass
class Instance:
    """One member of a batched ensemble."""
    mesh: Mesh
    N: int
    dt: float
    gam: float = 5.0 / 3
    artiv: int = 0
    k_stage: int = 3
    mu: float = 0.0
    eta0: float = 0.0
    precision: str = "mixed"
    topo: Topology = None


class Batch:
    """Many independent instances advanced together (chorus_batch_*): instances
    with equal numerics share launches; one time step is one hipGraph launch."""

    def __init__(self, members):
        L = _lib.load()
        self.members = list(members)
        n = len(self.members)
        meshes = (_lib.ChorusMesh * n)()
        params = (_lib.ChorusParams * n)()
        self._keep = []
        for i, m in enumerate(self.members):
            t = m.topo if m.topo is not None else topology(m.mesh)
            cm, arrs = _c_mesh(m.mesh, m.N, t)
            self._keep.append(arrs)
            meshes[i] = cm
            params[i] = _lib.ChorusParams(1, m.artiv, m.k_stage, m.mu, m.eta0, m.gam, m.dt, m.mesh.dxcycl,
                                          m.mesh.dycycl, m.mesh.ncell, PRECISION[m.precision])
        self._b = C.c_void_p()
        rc = L.chorus_batch_init(n, meshes, params, C.byref(self._b))
        self._keep = None
        if rc != 0:
            msg = L.chorus_batch_last_error(self._b).decode()
            L.chorus_batch_free(self._b)
            self._b = None
            raise ChorusError("chorus_batch_init: %s (%d)" % (msg, rc))

    def _check(self, rc, what):
        if rc != 0:
            raise ChorusError("%s: %s (%d)" % (what, _lib.load().chorus_batch_last_error(self._b).decode(), rc))

    def close(self):
        if getattr(self, "_b", None):
            _lib.load().chorus_batch_free(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i):
        m = self.members[i]
        return (m.mesh.ncell, m.N, m.N, 9)

    def set_state(self, i: int, Q: np.ndarray) -> None:
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.shape(i))
        self._check(_lib.load().chorus_batch_set_state(self._b, i, _lib.ptr(Q)), "chorus_batch_set_state")

    def get_state(self, i: int) -> np.ndarray:
        Q = np.empty(self.shape(i))
        self._check(_lib.load().chorus_batch_get_state(self._b, i, _lib.ptr(Q)), "chorus_batch_get_state")
        return Q

    def plot_setup(self, i: int, nre: int) -> None:
        """Tecplot plot topology of instance i (as Solver.plot_setup)."""
        m = self.members[i]
        local = m.mesh.glob2local()
        cm, arrs = _c_mesh(local, m.N, topology(local))
        self._check(_lib.load().chorus_batch_plot_setup(self._b, int(i), C.byref(cm), int(nre)),
                    "chorus_batch_plot_setup")

    def write_tecplot(self, i: int, path: str, ctime: float) -> None:
        """Tecplot file of instance i's current state (as Solver.write_tecplot)."""
        self._check(_lib.load().chorus_batch_write_tecplot(self._b, int(i), path.encode(), float(ctime)),
                    "chorus_batch_write_tecplot")

    def write_restart(self, i: int, path: str, it: int, ctime: float, icg=None) -> None:
        """Restart dump of instance i (write_data.f90:5-83), as Solver.write_restart."""
        ip = None
        if icg is not None:
            icg = np.ascontiguousarray(icg, dtype=np.int32)
            if icg.shape != (self.shape(i)[0],):
                raise ValueError("icg needs one entry per cell")
            ip = icg.ctypes.data_as(C.POINTER(C.c_int))
        self._check(_lib.load().chorus_batch_write_restart(self._b, i, path.encode(), int(it), float(ctime), ip),
                    "chorus_batch_write_restart")

    def compresid(self, i: int) -> np.ndarray:
        R = np.empty(self.shape(i))
        self._check(_lib.load().chorus_batch_compresid(self._b, i, _lib.ptr(R)), "chorus_batch_compresid")
        return R

    def step(self, nsteps: int = 1) -> None:
        self._check(_lib.load().chorus_batch_step(self._b, nsteps), "chorus_batch_step")

    def correct_b(self) -> None:
        self._check(_lib.load().chorus_batch_correct_b(self._b), "chorus_batch_correct_b")

    def divb(self) -> np.ndarray:
        out = np.empty(len(self.members))
        self._check(_lib.load().chorus_batch_divb(self._b, _lib.ptr(out)), "chorus_batch_divb")
        return out

    def groups(self):
        L = _lib.load()
        res = []
        for g in range(L.chorus_batch_ngroups(self._b)):
            v = [C.c_int() for _ in range(4)]
            self._check(L.chorus_batch_group(self._b, g, *[C.byref(x) for x in v]), "chorus_batch_group")
            res.append(dict(zip(["N", "ninst", "ncell", "tile"], [x.value for x in v])))
        return res

    def autotune(self):
        tiles = np.zeros(max(1, _lib.load().chorus_batch_ngroups(self._b)), dtype=np.int32)
        self._check(_lib.load().chorus_batch_autotune(self._b, _lib.ptr(tiles, _lib._ip)), "chorus_batch_autotune")
        return tiles.tolist()

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        self._check(_lib.load().chorus_batch_time_steps(self._b, nsteps, C.byref(ms)), "chorus_batch_time_steps")
        return ms.value
