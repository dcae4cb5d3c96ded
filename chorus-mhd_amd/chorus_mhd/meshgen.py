"""Structured quad-mesh generator emitting the reference's unstructured mesh format.

The reference ships no meshes; every workload here is an n x m grid of
counter-clockwise quads (CHECK_MESH keeps cell order as given, connectivity.f90:264-272),
with interior vertices optionally jittered by a seeded uniform offset.  Boundary
vertices stay on the grid so periodic partner faces keep matching midpoints
(map_cyc_loc.f90:24-28, tolerance 1e-5).

Numbering: vertex (i, j) -> 1 + i + j*(n+1); cell (i, j) -> 1 + i + j*n with
vertices (i,j), (i+1,j), (i+1,j+1), (i,j+1).  Boundary faces are listed
counter-clockwise: bottom (left to right), right (bottom to top), top (right to
left), left (top to bottom).
"""
from __future__ import annotations

import numpy as np

from .formats import Mesh, text_round_trip


def quad_grid(n: int, m: int, x0: float, x1: float, y0: float, y1: float,
              jitter: float = 0.0, seed: int = 0, bc: str = "CYCL") -> Mesh:
    """n x m cells on [x0,x1] x [y0,y1].

    ``jitter``: interior vertices move by U(-jitter, jitter) * (hx, hy), drawn
    from ``numpy.random.default_rng(seed)`` (x offsets for all interior vertices
    in vertex order, then y offsets).  ``bc``: boundary type on all four sides
    (``CYCL`` periodic, ``FREE``, ``INLE``, ``OUTL``, ``SYMP``, ``WALL``), or a
    tuple of four: bottom, right, top, left (``CYCL`` on opposite sides only).
    Coordinates are rounded through the 20-character vertex-file field, so the
    returned mesh is exactly what the reference reads back.
    """
    if n < 1 or m < 1:
        raise ValueError("mesh needs at least one cell per direction")
    if not 0.0 <= jitter < 0.5:
        raise ValueError("jitter must be in [0, 0.5) of the cell size")
    hx = (x1 - x0) / n
    hy = (y1 - y0) / m
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(m + 1), indexing="xy")
    xv = x0 + ii.astype(np.float64) * hx
    yv = y0 + jj.astype(np.float64) * hy
    # exact end points
    xv[:, n] = x1
    yv[m, :] = y1
    if jitter > 0.0 and n > 1 and m > 1:
        rng = np.random.default_rng(seed)
        ni = (n - 1) * (m - 1)
        dx = rng.uniform(-jitter, jitter, ni).reshape(m - 1, n - 1)
        dy = rng.uniform(-jitter, jitter, ni).reshape(m - 1, n - 1)
        xv[1:m, 1:n] += dx * hx
        yv[1:m, 1:n] += dy * hy
    xv = xv.reshape(-1)
    yv = yv.reshape(-1)
    # text round trip: what readGRIDinfo.f90:186 reads back
    xv = text_round_trip(xv)
    yv = text_round_trip(yv)

    def vid(i, j):
        return 1 + i + j * (n + 1)

    ci, cj = np.meshgrid(np.arange(n), np.arange(m), indexing="xy")
    ci = ci.reshape(-1)
    cj = cj.reshape(-1)
    cells = np.stack([vid(ci, cj), vid(ci + 1, cj), vid(ci + 1, cj + 1), vid(ci, cj + 1)],
                     axis=1).astype(np.int32)

    bnd = []
    for i in range(n):                      # bottom, left to right
        bnd.append((vid(i, 0), vid(i + 1, 0)))
    for j in range(m):                      # right, bottom to top
        bnd.append((vid(n, j), vid(n, j + 1)))
    for i in range(n - 1, -1, -1):          # top, right to left
        bnd.append((vid(i + 1, m), vid(i, m)))
    for j in range(m - 1, -1, -1):          # left, top to bottom
        bnd.append((vid(0, j + 1), vid(0, j)))
    bnd = np.array(bnd, dtype=np.int32)
    sides = (bc,) * 4 if isinstance(bc, str) else tuple(bc)
    if len(sides) != 4 or (sides[0] == "CYCL") != (sides[2] == "CYCL") or (sides[1] == "CYCL") != (sides[3] == "CYCL"):
        raise ValueError("bc: one type or four (bottom, right, top, left), CYCL on both sides of a pair")
    btype = [sides[0]] * n + [sides[1]] * m + [sides[2]] * n + [sides[3]] * m
    # DXCYCL/DYCYCL are the periodic shifts and also the AV length scale
    # h = sqrt(DXCYCL*DYCYCL/NCELLGLOB) (AV.f90:147), so they carry the domain
    # extents for FREE boundaries too.
    return Mesh(xv, yv, cells, bnd, btype, x1 - x0, y1 - y0)


def curved_grid(n: int, m: int, x0: float, x1: float, y0: float, y1: float, amp: float = 0.1,
                bc="CYCL"):
    """n x m 8-node (curved) cells on [x0,x1] x [y0,y1]: the uniform grid's
    vertices and edge midpoints moved by the smooth map
        (x, y) + amp (hx, hy) sin(2 pi (x-x0)/Lx) sin(2 pi (y-y0)/Ly),
    which fixes the boundary, so boundary edges stay straight and periodic
    partners keep matching midpoints.  Returns (mesh, xm, ym): the mesh of the
    corner vertices (numbered as quad_grid) and the mid-edge nodes 5..8 of each
    cell, (NCELL, 4) each, in the orientation of the reference's 8-node map
    (node 5 on edge 1-2, 6 on 2-3, 7 on 3-4, 8 on 4-1; mapder.f90:106-115).
    Shared edges get the same mid node from both cells."""
    base = quad_grid(n, m, x0, x1, y0, y1, bc=bc)
    hx, hy = (x1 - x0) / n, (y1 - y0) / m
    lx, ly = x1 - x0, y1 - y0

    def warp(x, y):
        s = amp * np.sin(2 * np.pi * (x - x0) / lx) * np.sin(2 * np.pi * (y - y0) / ly)
        return x + s * hx, y + s * hy

    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(m + 1), indexing="xy")
    gx, gy = x0 + ii.reshape(-1) * hx, y0 + jj.reshape(-1) * hy
    gx[(ii.reshape(-1) == n)] = x1
    gy[(jj.reshape(-1) == m)] = y1
    xv, yv = warp(gx, gy)
    xv, yv = text_round_trip(xv), text_round_trip(yv)
    ci, cj = np.meshgrid(np.arange(n), np.arange(m), indexing="xy")
    ci, cj = ci.reshape(-1).astype(np.float64), cj.reshape(-1).astype(np.float64)
    # grid positions of the mid-edge nodes: (i+1/2, j), (i+1, j+1/2), (i+1/2, j+1), (i, j+1/2)
    mi = np.stack([ci + 0.5, ci + 1.0, ci + 0.5, ci], axis=1)
    mj = np.stack([cj, cj + 0.5, cj + 1.0, cj + 0.5], axis=1)
    xm, ym = warp(x0 + mi * hx, y0 + mj * hy)
    xm = text_round_trip(xm.reshape(-1)).reshape(-1, 4)
    ym = text_round_trip(ym.reshape(-1)).reshape(-1, 4)
    mesh = Mesh(xv, yv, base.cells, base.bnd, list(base.bnd_type), base.dxcycl, base.dycycl)
    return mesh, xm, ym
