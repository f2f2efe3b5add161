// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocustorage/grid3dops.h: check_valid_mac_dimensions (grid3dops.cpp:23-45)
#ifndef __OCU_STORAGE_GRID3D_OPS_H__
#define __OCU_STORAGE_GRID3D_OPS_H__

#include "ocustorage/grid3d.h"

namespace ocu {

bool check_valid_mac_dimensions(const Grid3DUntyped &u, const Grid3DUntyped &v, const Grid3DUntyped &w, int nx,
                                int ny, int nz);

}  // namespace ocu

#endif
