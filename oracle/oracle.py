"""ctypes front end of the C oracle (liboracle.so) and of the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.

Arrays cross the boundary in the reference's Fortran layouts (column-major,
1-based connectivity); numpy views use the equivalent C-order shapes, e.g.
Q(9,N,N,NCELL) <-> ndarray[NCELL, N(js), N(is), 9].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
# the same build with gam = 1.4 (the reference's compile-time PARAMETER,
# setup2d.f90:68, changed by `make -C oracle ref-g14`)
REF_HARNESS_G14 = os.path.join(HERE, "_ref", "ref_harness_g14")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Topo(C.Structure):
    _fields_ = [("nface", C.c_int), ("ninter", C.c_int), ("npair", C.c_int), ("nfree", C.c_int),
                ("if2c", _ip), ("if2v", _ip), ("ibfinter", _ip), ("cyc2fpair", _ip),
                ("ibfree", _ip), ("ninl", C.c_int), ("nout", C.c_int), ("nsymp", C.c_int),
                ("nwal", C.c_int), ("ibfinl", _ip), ("ibfout", _ip), ("ibfsymp", _ip), ("ibfwal", _ip)]


class _Params(C.Structure):
    _fields_ = [("N", C.c_int), ("vismode", C.c_int), ("artiv", C.c_int), ("k_stage", C.c_int),
                ("mu", C.c_double), ("eta0", C.c_double), ("gam", C.c_double),
                ("dt", C.c_double), ("dxcycl", C.c_double), ("dycycl", C.c_double),
                ("qin", C.c_double * 9)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_topology.restype = C.c_int
        L.orc_topology.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, _ip, _ip,
                                   C.c_double, C.c_double, C.POINTER(_Topo)]
        L.orc_topo_free.argtypes = [C.POINTER(_Topo)]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_Params), C.c_int, C.c_int, _ip, _dp, _dp,
                                 C.POINTER(_Topo)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_q.argtypes = [C.c_void_p, _dp]
        L.orc_get_q.argtypes = [C.c_void_p, _dp]
        L.orc_setinitialcond_ot.argtypes = [C.c_void_p]
        L.orc_compresid.argtypes = [C.c_void_p, _dp]
        L.orc_correct_b.argtypes = [C.c_void_p]
        L.orc_steps.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.orc_array.restype = _dp
        L.orc_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long)]
        L.orc_scalar.restype = C.c_double
        L.orc_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_set_nodes8.restype = C.c_int
        L.orc_set_nodes8.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_mapder8.argtypes = [C.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


BTYPES = {"CYCL": 1, "FREE": 2, "INLE": 3, "OUTL": 4, "SYMP": 5, "WALL": 6}


class Topology:
    """IF2C / IBFINTER / CYC2FPAIR / IBFREE of a mesh, restated connectivity."""

    def __init__(self, mesh):
        L = lib()
        iv = np.asfortranarray(mesh.cells, dtype=np.int32)
        bnd = np.asfortranarray(mesh.bnd, dtype=np.int32)
        bt = np.array([BTYPES.get(t, 0) for t in mesh.bnd_type], dtype=np.int32)
        xv = np.ascontiguousarray(mesh.xv, dtype=np.float64)
        yv = np.ascontiguousarray(mesh.yv, dtype=np.float64)
        t = _Topo()
        rc = L.orc_topology(mesh.ncell, mesh.nvert, _ptr(iv, _ip), _ptr(xv, _dp), _ptr(yv, _dp),
                            bnd.shape[0], _ptr(bnd, _ip), _ptr(bt, _ip),
                            mesh.dxcycl, mesh.dycycl, C.byref(t))
        if rc != 0:
            raise RuntimeError("oracle topology failed (%d)" % rc)
        self._t = t
        nf = t.nface
        self.nface, self.ninter, self.npair, self.nfree = t.nface, t.ninter, t.npair, t.nfree
        self.if2c = np.ctypeslib.as_array(t.if2c, (4 * nf,)).reshape(4, nf).T.copy()
        self.if2v = np.ctypeslib.as_array(t.if2v, (2 * nf,)).reshape(2, nf).T.copy()
        self.ibfinter = np.ctypeslib.as_array(t.ibfinter, (max(t.ninter, 1),))[:t.ninter].copy()
        self.cyc2fpair = (np.ctypeslib.as_array(t.cyc2fpair, (max(2 * t.npair, 1),))[:2 * t.npair]
                          .reshape(2, t.npair).T.copy())
        self.ibfree = np.ctypeslib.as_array(t.ibfree, (max(t.nfree, 1),))[:t.nfree].copy()
        # IBFINL, IBFOUT, IBFSYMP, IBFWAL (connect_bdry.f90:186-205)
        self.bc_lists = {}
        for key, n, arr in (("INLE", t.ninl, t.ibfinl), ("OUTL", t.nout, t.ibfout),
                            ("SYMP", t.nsymp, t.ibfsymp), ("WALL", t.nwal, t.ibfwal)):
            self.bc_lists[key] = np.ctypeslib.as_array(arr, (max(n, 1),))[:n].copy()

    def __del__(self):
        try:
            lib().orc_topo_free(C.byref(self._t))
        except Exception:
            pass


class Oracle:
    """One FP64 instance: INIT_SETUP, compresid, RK + CT steps (reference semantics)."""

    def __init__(self, mesh, N, dt=1e-3, artiv=0, k_stage=3, mu=0.0, eta0=0.0,
                 gam=5.0 / 3, vismode=1, topo=None, qin=None):
        L = lib()
        self.mesh = mesh
        self.N = N
        self.topo = topo or Topology(mesh)
        q = (C.c_double * 9)(*([0.0] * 9 if qin is None else [float(x) for x in qin]))
        p = _Params(N, vismode, artiv, k_stage, mu, eta0, gam, dt, mesh.dxcycl, mesh.dycycl, q)
        iv = np.asfortranarray(mesh.cells, dtype=np.int32)
        self._keep = (iv,)
        self._c = L.orc_create(C.byref(p), mesh.ncell, mesh.nvert, _ptr(iv, _ip),
                               _ptr(np.ascontiguousarray(mesh.xv), _dp),
                               _ptr(np.ascontiguousarray(mesh.yv), _dp), C.byref(self.topo._t))
        if not self._c:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        try:
            lib().orc_destroy(self._c)
        except Exception:
            pass

    @property
    def shape(self):
        return (self.mesh.ncell, self.N, self.N, 9)

    def set_q(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(self.shape)
        lib().orc_set_q(self._c, _ptr(Q, _dp))

    def get_q(self):
        Q = np.empty(self.shape)
        lib().orc_get_q(self._c, _ptr(Q, _dp))
        return Q

    def init_ot(self):
        lib().orc_setinitialcond_ot(self._c)

    def compresid(self):
        R = np.empty(self.shape)
        lib().orc_compresid(self._c, _ptr(R, _dp))
        return R

    def correct_b(self):
        lib().orc_correct_b(self._c)

    def steps(self, n):
        r, d, m = C.c_double(), C.c_double(), C.c_double()
        lib().orc_steps(self._c, n, C.byref(r), C.byref(d), C.byref(m))
        return r.value, d.value, m.value

    def array(self, name):
        n = C.c_long()
        p = lib().orc_array(self._c, name.encode(), C.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, (n.value,)).copy()

    def scalar(self, name):
        return lib().orc_scalar(self._c, name.encode())

    def set_nodes8(self, xm, ym):
        """Mid-edge nodes 5..8 of every cell, (NCELL, 4) each: 8-node cells."""
        self._x8 = np.asfortranarray(xm, dtype=np.float64).ravel(order="F")
        self._y8 = np.asfortranarray(ym, dtype=np.float64).ravel(order="F")
        if lib().orc_set_nodes8(self._c, _ptr(self._x8, _dp), _ptr(self._y8, _dp)) != 0:
            raise RuntimeError("orc_set_nodes8 failed")


def mapder8(N):
    """MAPDER_8_NODE's tables (dmdxs(8,2,N,N), dmdxf1(8,2,N+1,N), dmdxf2(8,2,N,N+1)), flat in Fortran order."""
    a, b, c = np.empty(16 * N * N), np.empty(16 * (N + 1) * N), np.empty(16 * (N + 1) * N)
    lib().orc_mapder8(N, _ptr(a, _dp), _ptr(b, _dp), _ptr(c, _dp))
    return a, b, c


def run_reference(workdir: str, mode: str, timeout: float = 600.0, gam: float = 5.0 / 3) -> None:
    """Run the compiled reference (oracle/_ref/ref_harness) in ``workdir``;
    gam = 1.4 selects the build with that ratio of specific heats."""
    if gam == 5.0 / 3:
        exe = REF_HARNESS
    elif gam == 1.4:
        exe = REF_HARNESS_G14
    else:
        raise ValueError("the reference is built for gam = 5/3 and 1.4 only, not %r" % gam)
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    r = subprocess.run([exe, mode], cwd=workdir, capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
