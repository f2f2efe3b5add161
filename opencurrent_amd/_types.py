"""ctypes mirrors of include/ocu_types.h and the Grid3D padded layout.

The layout follows ocustorage/grid3ddev.cu:358-400: pnz is rounded up to a
multiple of 16 elements, the allocation carries a (16 - gz)-element
pre-padding, and cell (i, j, k) lives at element
``shift + i*pny*pnz + j*pnz + k``.
"""
import ctypes as C

import numpy as np

OCU_F32, OCU_F64, OCU_I32 = 0, 1, 2

BC_INVALID, BC_NONE, BC_PERIODIC, BC_DIRICHELET, BC_NEUMANN = 0, 1, 2, 3, 4
BC_SECOND_DERIV, BC_DIRICHELET_AND_NEUMANN = 5, 6
BC_FORCED_INFLOW_VARIABLE_SLIP, BC_SCALAR_SLIP = 7, 8

CONVERGENCE_CALC_L2 = 0x01
CONVERGENCE_CALC_LINF = 0x02
CONVERGENCE_CRITERIA_L2 = 0x04
CONVERGENCE_CRITERIA_LINF = 0x08
CONVERGENCE_CRITERIA_NONE = 0x10
CONVERGENCE_L2 = 0x05
CONVERGENCE_LINF = 0x0A
CONVERGENCE_NONE = 0x10
CONVERGENCE_NONE_CALC_L2LINF = 0x13

IT_FIRST_ORDER_UPWIND, IT_SECOND_ORDER_CENTERED, IT_SECOND_ORDER_UPWIND = 1, 2, 4

DIR_XPOS, DIR_XNEG, DIR_YPOS, DIR_YNEG, DIR_ZPOS, DIR_ZNEG = 0x05, 0x06, 0x09, 0x0A, 0x11, 0x12

RELAX_DEVICE, RELAX_HOST = 0, 1
FUSED_WRITE_GHOSTS, FUSED_ZERO_INPUT = 1, 2

TS_FORWARD_EULER, TS_ADAMS_BASHFORD2 = 1, 2


class Grid3DDesc(C.Structure):
    _fields_ = [
        ("base", C.c_void_p),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("gx", C.c_int32), ("gy", C.c_int32), ("gz", C.c_int32),
        ("pnx", C.c_int32), ("pny", C.c_int32), ("pnz", C.c_int32),
        ("dtype", C.c_int32),
        ("shift", C.c_int64),
        ("allocated", C.c_int64),
    ]


class BC(C.Structure):
    _fields_ = [("type", C.c_int32), ("value", C.c_float), ("aux_value", C.c_float)]


class BCSet(C.Structure):
    _fields_ = [("xpos", BC), ("xneg", BC), ("ypos", BC), ("yneg", BC), ("zpos", BC), ("zneg", BC)]

    FACES = ("xpos", "xneg", "ypos", "yneg", "zpos", "zneg")

    @classmethod
    def uniform(cls, typ, value=0.0, aux=0.0):
        s = cls()
        for f in cls.FACES:
            setattr(s, f, BC(typ, value, aux))
        return s

    def copy(self):
        s = BCSet()
        C.memmove(C.byref(s), C.byref(self), C.sizeof(BCSet))
        return s


class MGLevelParams(C.Structure):
    _fields_ = [("h", C.c_double), ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double),
                ("omega", C.c_double), ("mods", C.c_double * 6)]


class NSStepParams(C.Structure):
    """ocu_ns_step_params (include/ocu_types.h)"""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("advection_scheme", C.c_int32), ("vertical_direction", C.c_int32),
                ("flow_periodic", C.c_int32), ("temp_periodic", C.c_int32),
                ("keep_tendencies", C.c_int32), ("use_last", C.c_int32),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("viscosity", C.c_double), ("thermal_diffusion", C.c_double),
                ("thermal_coef", C.c_double), ("a_dt", C.c_double), ("a_ab", C.c_double)]


class MGConfig(C.Structure):
    _fields_ = [("bc", BCSet), ("nu1", C.c_int32), ("nu2", C.c_int32),
                ("convergence", C.c_int32), ("make_symmetric_operator", C.c_int32),
                ("disable_relax_on_host", C.c_int32), ("prolong_quirk", C.c_int32)]

    @classmethod
    def make(cls, bc, nu1=2, nu2=2, convergence=CONVERGENCE_L2, sym=False,
             disable_relax_on_host=False, prolong_quirk=1):
        c = cls()
        c.bc = bc
        c.nu1, c.nu2 = nu1, nu2
        c.convergence = convergence
        c.make_symmetric_operator = int(sym)
        c.disable_relax_on_host = int(disable_relax_on_host)
        c.prolong_quirk = prolong_quirk
        return c


def layout(nx, ny, nz, gx=1, gy=1, gz=1, padx=0, pady=0, padz=0, dtype=OCU_F64):
    """Grid3DDevice<T>::init layout (grid3ddev.cu:358-400) as a descriptor with base=NULL."""
    d = Grid3DDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.gx, d.gy, d.gz = gx, gy, gz
    d.pnx = nx + 2 * gx + padx
    d.pny = ny + 2 * gy + pady
    pnz = nz + 2 * gz + padz
    if pnz & 15:
        pnz = ((pnz >> 4) + 1) << 4
    d.pnz = pnz
    pre = 16 - gz
    d.allocated = d.pnz * d.pny * d.pnx + pre
    d.shift = gx * d.pnz * d.pny + gy * d.pnz + gz + pre
    d.dtype = dtype
    d.base = None
    return d


def np_dtype(code):
    return {OCU_F32: np.float32, OCU_F64: np.float64, OCU_I32: np.int32}[code]


def dtype_code(dt):
    dt = np.dtype(dt)
    if dt == np.float32:
        return OCU_F32
    if dt == np.float64:
        return OCU_F64
    if dt == np.int32:
        return OCU_I32
    raise ValueError(f"unsupported dtype {dt}")


class HostGrid:
    """A numpy-backed grid with the reference layout (a Grid3DHost<T>)."""

    def __init__(self, nx, ny, nz, gx=1, gy=1, gz=1, padx=0, pady=0, padz=0, dtype=np.float64):
        self.desc = layout(nx, ny, nz, gx, gy, gz, padx, pady, padz, dtype_code(dtype))
        self.buf = np.zeros(self.desc.allocated, dtype=dtype)
        self.desc.base = self.buf.ctypes.data

    @classmethod
    def congruent(cls, other, dtype=None):
        d = other.desc
        return cls(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz,
                   d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy), d.pnz - (d.nz + 2 * d.gz),
                   dtype=dtype or other.buf.dtype)

    @property
    def shape_padded(self):
        return (self.desc.pnx, self.desc.pny, self.desc.pnz)

    def full(self):
        """View of the padded 3-D box (pnx, pny, pnz); index [i+gx, j+gy, k+gz] is cell (i,j,k)."""
        d = self.desc
        pre = 16 - d.gz
        return self.buf[pre:].reshape(d.pnx, d.pny, d.pnz)

    def interior(self):
        d = self.desc
        return self.full()[d.gx:d.gx + d.nx, d.gy:d.gy + d.ny, d.gz:d.gz + d.nz]

    def ptr(self):
        return C.byref(self.desc)

    def copy(self):
        g = HostGrid.congruent(self)
        g.buf[:] = self.buf
        return g
