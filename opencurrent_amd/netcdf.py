"""NetCDFGrid3DWriter / NetCDFGrid3DReader (ocustorage/gridnetcdf.h:67-183)
over the C ABI of libocu.so: grid fields (HostGrid, any layout) as 4-D
(Time, X, Y, Z) variables of NetCDF classic files."""
import ctypes as C

import numpy as np

from . import _lib
from ._types import OCU_F32, OCU_F64, OCU_I32, HostGrid, dtype_code

NC_INT, NC_FLOAT, NC_DOUBLE = 4, 5, 6
GS_CENTER_POINT, GS_U_FACE, GS_V_FACE, GS_W_FACE, GS_CUSTOM = range(5)
_NP = {NC_INT: np.int32, NC_FLOAT: np.float32, NC_DOUBLE: np.float64}


class NetCDFGrid3DWriter:
    """open(filename, nx, ny, nz, hx, hy, hz) on construction; the file is
    laid out and complete after close()."""

    def __init__(self, filename, nx, ny, nz, hx=1.0, hy=1.0, hz=1.0):
        self.h = _lib.host().ocu_nc_writer_open(str(filename).encode(), nx, ny, nz, hx, hy, hz)
        if not self.h:
            raise _lib.OcuError(f"NetCDFGrid3DWriter::open({filename}) failed")

    def define_variable(self, name, nc_type, staggering=GS_CENTER_POINT, offsets=(.5, .5, .5)):
        return _lib.host().ocu_nc_writer_define_variable(self.h, name.encode(), nc_type, staggering, *offsets) == 0

    def add_time_level(self, time):
        level = C.c_int()
        if _lib.host().ocu_nc_writer_add_time_level(self.h, time, C.byref(level)) != 0:
            return None
        return level.value

    def add_data(self, name, grid, time_level=0):
        return _lib.host().ocu_nc_writer_add_data(self.h, name.encode(), grid.ptr(), time_level) == 0

    def close(self):
        h, self.h = self.h, None
        return h is not None and _lib.host().ocu_nc_writer_close(h) == 0

    def __del__(self):
        if getattr(self, "h", None):
            self.close()


class NetCDFGrid3DReader:
    def __init__(self, filename):
        self.h = _lib.host().ocu_nc_reader_open(str(filename).encode())
        if not self.h:
            raise _lib.OcuError(f"NetCDFGrid3DReader::open({filename}) failed")
        nx, ny, nz, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        hx, hy, hz = C.c_float(), C.c_float(), C.c_float()
        _lib.host().ocu_nc_reader_info(self.h, C.byref(nx), C.byref(ny), C.byref(nz), C.byref(hx), C.byref(hy),
                                       C.byref(hz), C.byref(nt))
        self.nx, self.ny, self.nz = nx.value, ny.value, nz.value
        self.hx, self.hy, self.hz = hx.value, hy.value, hz.value
        self.num_time_levels = nt.value

    def get_time(self, level):
        t = C.c_float()
        _lib.check(_lib.host().ocu_nc_reader_time(self.h, level, C.byref(t)), "get_time")
        return t.value

    def variables(self):
        """{name: (nc_type, (nx, ny, nz))} of the 4-D variables"""
        host = _lib.host()
        out = {}
        for q in range(host.ocu_nc_reader_variable(self.h, -1, None, 0, None, None, None, None)):
            buf = C.create_string_buffer(1024)
            t, a, b, c = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            _lib.check(host.ocu_nc_reader_variable(self.h, q, buf, 1024, C.byref(t), C.byref(a), C.byref(b),
                                                   C.byref(c)), "variable")
            out[buf.value.decode()] = (t.value, (a.value, b.value, c.value))
        return out

    def read_variable(self, name, grid, time_level=0):
        return _lib.host().ocu_nc_reader_read(self.h, name.encode(), grid.ptr(), time_level) == 0

    def read(self, name, time_level=0):
        """the variable as a new HostGrid (one ghost layer, type from the file)"""
        t, n = self.variables()[name]
        g = HostGrid(*n, dtype=_NP[t])
        if not self.read_variable(name, g, time_level):
            raise _lib.OcuError(f"read_variable({name}, {time_level}) failed")
        return g

    def close(self):
        if getattr(self, "h", None):
            _lib.host().ocu_nc_reader_close(self.h)
            self.h = None

    def __del__(self):
        self.close()
