"""chorus_mhd -- MI355X-native residual evaluation and SSP-RK time step of CHORUS-MHD.

Host-side mirror of the reference's hot-path subroutines (compresid, RK_S5P3,
FiveStageRk, correct_B_via_curl_Az) over the C-ABI library
``libchorus_mhd.so`` (HIP kernels for gfx950), plus the reference's file
formats and a structured mesh generator.
"""
from .formats import QuadInp, Mesh, read_mesh, write_restart, read_restart, fmt_d20
from .meshgen import quad_grid, curved_grid

__all__ = ["QuadInp", "Mesh", "read_mesh", "write_restart", "read_restart", "fmt_d20",
           "quad_grid", "curved_grid"]
