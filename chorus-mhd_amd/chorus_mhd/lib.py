"""ctypes binding of libchorus_mhd.so (include/chorus_mhd.h).

This is the Python-side equivalent of the iso_c_binding interface a Fortran
driver uses; it loads the in-tree library and fails loudly when it is absent
or no HIP device is present -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# CHORUS_MHD_LIB selects another in-tree build of the same library (e.g. a
# variant built with different tuning macros for a measurement)
LIB_PATH = os.environ.get("CHORUS_MHD_LIB") or os.path.join(os.path.dirname(PKG), "libchorus_mhd.so")

CHORUS_FP64 = 0
CHORUS_MIXED = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class ChorusMesh(C.Structure):
    _fields_ = [("ncell", C.c_int), ("nvert", C.c_int), ("N", C.c_int),
                ("xv", _dp), ("yv", _dp), ("ivcell", _ip),
                ("nface", C.c_int), ("if2c", _ip),
                ("ninter", C.c_int), ("ibfinter", _ip),
                ("npair", C.c_int), ("cyc2fpair", _ip),
                ("nfree", C.c_int), ("ibfree", _ip)]


class ChorusParams(C.Structure):
    _fields_ = [("vismode", C.c_int), ("artiv", C.c_int), ("k_stage", C.c_int),
                ("mu", C.c_double), ("eta0", C.c_double), ("gam", C.c_double),
                ("dt", C.c_double), ("dxcycl", C.c_double), ("dycycl", C.c_double),
                ("ncellglob", C.c_int), ("precision", C.c_int)]


class ChorusBC(C.Structure):
    _fields_ = [("ninlet", C.c_int), ("ibfinl", _ip), ("noutlet", C.c_int), ("ibfout", _ip),
                ("nsymp", C.c_int), ("ibfsymp", _ip), ("nwall", C.c_int), ("ibfwal", _ip),
                ("qin", C.c_double * 9)]


# exported symbols and their signatures (include/chorus_mhd.h)
SIGNATURES = {
    "chorus_init": (C.c_int, [C.POINTER(ChorusMesh), C.POINTER(ChorusParams), C.POINTER(C.c_void_p)]),
    "chorus_init_bc": (C.c_int, [C.POINTER(ChorusMesh), C.POINTER(ChorusParams), C.POINTER(ChorusBC),
                                 C.POINTER(C.c_void_p)]),
    "chorus_set_nodes8": (C.c_int, [C.c_void_p, _dp, _dp]),
    "chorus_free": (None, [C.c_void_p]),
    "chorus_last_error": (C.c_char_p, [C.c_void_p]),
    "chorus_set_state": (C.c_int, [C.c_void_p, _dp]),
    "chorus_get_state": (C.c_int, [C.c_void_p, _dp]),
    "chorus_compresid": (C.c_int, [C.c_void_p, _dp]),
    "chorus_correct_b": (C.c_int, [C.c_void_p]),
    "chorus_step": (C.c_int, [C.c_void_p, C.c_int]),
    "chorus_diag": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "chorus_diag_noalloc": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, _dp]),
    "chorus_topology": (C.c_int, [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, _ip, _ip,
                                  C.c_double, C.c_double, _ip, _ip, _ip, _ip, _ip, _ip,
                                  _ip, _ip]),
    "chorus_time_steps": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "chorus_autotune": (C.c_int, [C.c_void_p, _ip]),
    "chorus_set_tile": (C.c_int, [C.c_void_p, C.c_int]),
    "chorus_plot_setup": (C.c_int, [C.c_void_p, C.POINTER(ChorusMesh), C.c_int]),
    "chorus_write_tecplot": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "chorus_write_restart": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_double, _ip]),
    "chorus_batch_write_restart": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, C.c_double, _ip]),
    "chorus_batch_plot_setup": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(ChorusMesh), C.c_int]),
    "chorus_batch_write_tecplot": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_double]),
    "chorus_plot_topology": (C.c_int, [C.POINTER(ChorusMesh), C.c_int, _ip, _ip, _dp, _ip]),
    "chorus_batch_init": (C.c_int, [C.c_int, C.POINTER(ChorusMesh), C.POINTER(ChorusParams), C.POINTER(C.c_void_p)]),
    "chorus_batch_free": (None, [C.c_void_p]),
    "chorus_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "chorus_batch_set_state": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "chorus_batch_get_state": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "chorus_batch_step": (C.c_int, [C.c_void_p, C.c_int]),
    "chorus_batch_compresid": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "chorus_batch_divb": (C.c_int, [C.c_void_p, _dp]),
    "chorus_batch_correct_b": (C.c_int, [C.c_void_p]),
    "chorus_batch_ngroups": (C.c_int, [C.c_void_p]),
    "chorus_batch_group": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip]),
    "chorus_batch_autotune": (C.c_int, [C.c_void_p, _ip]),
    "chorus_batch_time_steps": (C.c_int, [C.c_void_p, C.c_int, _dp]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the in-tree library (no build, no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libchorus_mhd.so is not built (%s); run __graft_entry__.build()" % path)
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)
