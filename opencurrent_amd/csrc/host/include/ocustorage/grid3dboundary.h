// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Ghost-cell fills, ocustorage/grid3dboundary.h.  Enqueued on the compute
// stream; each returns false on invalid input (same checks as the reference).
#ifndef __OCU_STORAGE_GRID3D_BOUNDARY_H__
#define __OCU_STORAGE_GRID3D_BOUNDARY_H__

#include "ocustorage/grid3d.h"
#include "ocuutil/boundary_condition.h"

namespace ocu {

ocu_bc_set to_c_bc(const BoundaryConditionSet &bc);

template <typename T>
bool apply_3d_boundary_conditions_level1(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc, double hx,
                                         double hy, double hz);
template <typename T>
bool apply_3d_boundary_conditions_level1_nocorners(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc,
                                                   double hx, double hy, double hz);
template <typename T>
bool apply_3d_mac_boundary_conditions_level1(Grid3DDevice<T> &u, Grid3DDevice<T> &v, Grid3DDevice<T> &w,
                                             const BoundaryConditionSet &bc, double hx, double hy, double hz);
// grid3dboundary.cu:904-938: two ghost layers per face (not edges or corners)
template <typename T>
bool apply_3d_boundary_conditions_level2_nocorners(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc,
                                                   double hx, double hy, double hz);
// grid3dboundary.cu:940-1016, with nx = u.nx() - 1 (see ocu_bc_fill_mac_level2)
template <typename T>
bool apply_3d_mac_boundary_conditions_level2(Grid3DDevice<T> &u, Grid3DDevice<T> &v, Grid3DDevice<T> &w,
                                             const BoundaryConditionSet &bc, double hx, double hy, double hz);

}  // namespace ocu

#endif
