"""Output and restart files of the reference, through the C-ABI.

* ``CdClLog``: the Tecplot drag / lift history ``Information/Time_cDcL.dat``
  (``Delay_IO_CD_CL`` / ``noDelay_IO_CD_CL``,
  inc/dfib/virtualForceIntergrator.hpp:44-177).
* ``write_plot3d_x`` / ``read_plot3d_q``: the Plot3D grid file and the
  solution-file reader (inc/io/Plot3D/xfileWrite.hpp:6-61,
  qfileRead.hpp:6-100); ``Simulation.write_plot3d_q`` writes the solution
  file (qfile.hpp:7-104) from the device fields.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi
from .capi import GridArrays, check, dptr, load


class CdClLog:
    """simuClass's delayIO ring buffers and the file they flush to.

    ``record(loop, time, cd, cl)`` is called once per timestep with
    simuClass::loop and the time of the step (before ``finishloop``)."""

    def __init__(self, path, delay_io: int = 100, zone: str = "ELL (BICG)", mode: int = capi.CFD_CDCL_REFERENCE):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.cfd_cdcl_open(os.fsencode(path), int(delay_io), zone.encode(), int(mode), C.byref(h)),
              "cfd_cdcl_open")
        self.h = h

    def record(self, loop: int, time: float, cd: float, cl: float) -> None:
        check(self.lib.cfd_cdcl_record(self.h, int(loop), float(time), float(cd), float(cl)), "cfd_cdcl_record")

    def close(self) -> None:
        if self.h:
            self.lib.cfd_cdcl_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def write_plot3d_x(grid: GridArrays, path) -> None:
    desc = grid.desc()
    check(load().cfd_plot3d_write_x(C.byref(desc), os.fsencode(path)), "cfd_plot3d_write_x")


def read_plot3d_q(grid: GridArrays, path):
    """(nblock, (mach, alpha, Re, time), p, u, v, w, eta): ghosted FP64 arrays."""
    desc = grid.desc()
    out = [np.empty(grid.G) for _ in range(5)]
    nb = C.c_int()
    hdr = (C.c_float * 4)()
    check(load().cfd_plot3d_read_q(os.fsencode(path), C.byref(desc), C.byref(nb), hdr, *[dptr(a) for a in out]),
          "cfd_plot3d_read_q")
    return (nb.value, tuple(hdr), *out)
