"""MI355X-native DG-SEM wave stage (libndg_amd): hand-written HIP kernels for gfx950 behind a C ABI
(include/ndg_amd.h), with a Python mirror of the reference host API (NDG_t, Mesh2D_t, Solver2D_t).
"""
from ._abi import LIB_PATH, NDGError, device_count, lib  # noqa: F401
from .data_writer import DataWriter_t  # noqa: F401
from .sem import (BC_INFLOW, BC_NONE, BC_OUTFLOW, BC_PERIODIC, BC_SYMMETRY, BC_WALL,  # noqa: F401
                  MeshDescription, Mesh2D_t, NDG_t, Solver2D_t, inverse_quad_map, read_su2, square_mesh)

__all__ = ["NDG_t", "Mesh2D_t", "Solver2D_t", "DataWriter_t", "MeshDescription", "square_mesh", "read_su2", "inverse_quad_map", "NDGError",
           "device_count", "lib", "LIB_PATH", "BC_NONE", "BC_WALL", "BC_SYMMETRY", "BC_INFLOW",
           "BC_OUTFLOW", "BC_PERIODIC"]
