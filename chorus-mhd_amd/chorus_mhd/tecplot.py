"""Reader of the Tecplot TDV111 binary files written by tecpost_parallel
(tecplotter.f90:92-418) and by chorus_write_tecplot: header fields, the
variables (FP32 blocks) and the 0-based connectivity."""
from __future__ import annotations

import struct

import numpy as np


def read_plt(path: str) -> dict:
    b = open(path, "rb").read()
    if b[:8] != b"#!TDV111":
        raise ValueError("not a TDV111 file")
    i = 8

    def ints(k):
        nonlocal i
        v = struct.unpack_from("<%di" % k, b, i)
        i += 4 * k
        return v

    def string():
        nonlocal i
        out = []
        while True:
            c = struct.unpack_from("<i", b, i)[0]
            i += 4
            if c == 0:
                return "".join(out)
            out.append(chr(c))

    byte_order, ftype = ints(2)
    title = string()
    nvar = ints(1)[0]
    names = [string() for _ in range(nvar)]
    zone_marker = struct.unpack_from("<f", b, i)[0]
    i += 4
    zone = string()
    parent, strand = ints(2)
    time = struct.unpack_from("<d", b, i)[0]
    i += 8
    z = ints(12)
    eoh, zm2 = struct.unpack_from("<2f", b, i)
    i += 8
    fmt = ints(nvar + 3)
    minmax = np.frombuffer(b, dtype="<f8", count=2 * nvar, offset=i).reshape(nvar, 2)
    i += 16 * nvar
    nn, nc = z[6], z[7]
    data = np.frombuffer(b, dtype="<f4", count=nvar * nn, offset=i).reshape(nvar, nn)
    i += 4 * nvar * nn
    conn = np.frombuffer(b, dtype="<i4", count=4 * nc, offset=i).reshape(nc, 4)
    i += 16 * nc
    if i != len(b):
        raise ValueError("trailing bytes in %s" % path)
    return dict(title=title, names=names, zone=zone, time=time, zone_ints=z, formats=fmt, minmax=minmax,
                nnode=nn, ncell=nc, data=data, conn=conn, markers=(zone_marker, eoh, zm2), header_ints=(byte_order, ftype, parent, strand))
