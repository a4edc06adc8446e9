"""VTK output mirroring Ikarus::Vtk::Writer (ikarus/io/vtkwriter.hh:37-146) over ik_write_vtu.

    writer = Writer(assembler)                       # or a Device
    writer.addInterpolation(u, "displacement")       # a DOF vector (FlatInterleaved) as point data
    writer.addResult(PK2, DataTag.asCellData)        # ResultFunction (io/resultfunction.hh:55-110)
    writer.write("out")                              # out.vtu

Results are evaluated on the device at the requirement the assembler is bound to when write() runs.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import IkError

PK2, KIRCHHOFF, CAUCHY = capi.PK2, capi.KIRCHHOFF, capi.CAUCHY


class DataTag:
    asPointData = capi.VTK_POINT_DATA
    asCellData = capi.VTK_CELL_DATA


class Writer:
    def __init__(self, assembler):
        self.dev = getattr(assembler, "dev", assembler)
        self.problem = self.dev.problem
        self._fields = []
        self._results = 0
        self._tag = None

    def addInterpolation(self, vals, name, dataTag=DataTag.asPointData):
        """A nodal DOF vector (point data, dim components) or per-element values (cell data)."""
        p = self.problem
        v = np.ascontiguousarray(vals, dtype=np.float64)
        n = p.X.shape[0] if dataTag == DataTag.asPointData else p.ne
        if v.size % n:
            raise IkError(capi.IK_EINVAL, f"{name}: {v.size} values do not split over {n} entities")
        self._fields.append((name, v, v.size // n, dataTag))

    def addResult(self, kind, dataTag=DataTag.asPointData):
        if self._tag is not None and self._tag != dataTag:
            raise IkError(capi.IK_EINVAL, "all results of one file share one data tag")
        self._results |= 1 << kind
        self._tag = dataTag

    def addAllResults(self, dataTag=DataTag.asPointData):
        for k in (PK2, KIRCHHOFF, CAUCHY):
            self.addResult(k, dataTag)

    def write(self, filename):
        path = filename if filename.endswith(".vtu") else filename + ".vtu"
        arr = (capi.VtkField * max(1, len(self._fields)))()
        keep = []
        for i, (name, v, comps, tag) in enumerate(self._fields):
            b = name.encode()
            keep.append(b)
            arr[i].name, arr[i].values, arr[i].components, arr[i].cell = b, v.ctypes.data_as(C.c_void_p), comps, tag
        tag = self._tag if self._tag is not None else DataTag.asPointData
        self.dev.check(capi.lib().ik_write_vtu(self.dev.h, path.encode(), arr, len(self._fields), self._results, tag))
        return path
