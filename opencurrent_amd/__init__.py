"""opencurrent_amd - MI355X-native projection-method hot path of OpenCurrent.

The product is native code: libocu_hip.so (hand-written HIP kernels for
gfx950 behind the C ABI in include/ocu_hip.h) and libocu.so (OpenCurrent's
C++ solver classes - Sol_MultigridPressure3DDevice, Sol_ProjectDivergence3DDevice,
Eqn_IncompressibleNS3D, ... - calling that ABI, exported through
include/ocu.h).  This package loads them with ctypes and offers a thin
Python front-end; it has no pure-Python or CPU fallback.
"""
from ._types import (BC, BCSet, Grid3DDesc, HostGrid, MGConfig, MGLevelParams, layout,
                     OCU_F32, OCU_F64, BC_NONE, BC_PERIODIC, BC_DIRICHELET, BC_NEUMANN,
                     BC_FORCED_INFLOW_VARIABLE_SLIP, CONVERGENCE_L2, CONVERGENCE_LINF,
                     CONVERGENCE_NONE, IT_FIRST_ORDER_UPWIND, IT_SECOND_ORDER_CENTERED,
                     DIR_XPOS, DIR_XNEG, DIR_YPOS, DIR_YNEG, DIR_ZPOS, DIR_ZNEG,
                     RELAX_DEVICE, RELAX_HOST, TS_FORWARD_EULER, TS_ADAMS_BASHFORD2)
from ._lib import NativeLibraryMissing, OcuError, library_paths

__all__ = [n for n in dir() if not n.startswith("_")]
