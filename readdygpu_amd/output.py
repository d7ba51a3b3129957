"""Trajectory and RDF files of the reference's stand-alone driver
(lib/ReaDDyGPU_Cpp/outputHandler/outputHandler.cpp:17-148), written by the
library (rd_traj_*, rd_rdf_write).  Formatting runs on the CPU."""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib

PDB, XML = 1, 2


class TrajectoryWriter:
    """PDB (MODEL/ATOM/ENDMDL) or XML (<traj><tplgy_coords stepId>...) frames."""

    def __init__(self, path: str, fmt: int = PDB):
        self.lib = _lib.load()
        self.t = C.c_void_p()
        rc = self.lib.rd_traj_open(path.encode(), int(fmt), C.byref(self.t))
        if rc != 0:
            raise OSError(f"cannot open {path} for trajectory output ({rc})")

    def write_frame(self, frame: int, xyz: np.ndarray, types: np.ndarray):
        x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        if len(x) != len(t):
            raise ValueError("xyz and types differ in length")
        rc = self.lib.rd_traj_write_frame(self.t, int(frame), len(t), x.ctypes.data_as(C.c_void_p),
                                          t.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise OSError("trajectory write failed")

    def close(self):
        if self.t:
            rc = self.lib.rd_traj_close(self.t)
            self.t = None
            if rc != 0:
                raise OSError("trajectory close failed")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def write_rdf(path: str, g: np.ndarray, rdf_range: float, names: Optional[Sequence[str]] = None):
    """RDF table: g [nt, nt, nbins] (one replica); the distance column is
    k * rdf_range / nbins as in RDFOutputHandler::writeOutput."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    nt, _, nb = g.shape
    arr = None
    if names is not None:
        arr = (C.c_char_p * nt)(*[n.encode() for n in names])
    rc = _lib.load().rd_rdf_write(path.encode(), nt, C.cast(arr, C.c_void_p) if arr is not None else None, nb,
                                  float(rdf_range), g.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise OSError(f"cannot write {path}")
