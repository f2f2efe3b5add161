// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Point sampling of grids, the drop-in for ocustorage/grid3dsample.h:
// trilinear interpolation at float positions (physical units) with the
// faces' periodic wrap or clamp / Neumann extrapolation.  Enqueued on the
// compute stream; false on invalid input.
#ifndef __OCU_STORAGE_GRID3D_SAMPLE_H__
#define __OCU_STORAGE_GRID3D_SAMPLE_H__

#include "ocustorage/grid1d.h"
#include "ocustorage/grid3d.h"
#include "ocuutil/boundary_condition.h"

namespace ocu {

// grid3dsample.cu:137-171 (the definition's signature: the reference's
// header declares it without the shifts, which no definition matches)
template <typename T>
bool sample_points_3d(Grid1DDevice<T> &phi_sampled, const Grid1DDevice<float> &position_x,
                      const Grid1DDevice<float> &position_y, const Grid1DDevice<float> &position_z,
                      const Grid3DDevice<T> &phi, const BoundaryConditionSet &bc, float period_nx, float period_ny,
                      float period_nz, float hx, float hy, float hz, float shiftx, float shifty, float shiftz);

// grid3dsample.cu:174-263
template <typename T>
bool sample_points_mac_grid_3d(Grid1DDevice<T> &vx_sampled, Grid1DDevice<T> &vy_sampled, Grid1DDevice<T> &vz_sampled,
                               const Grid1DDevice<float> &position_x, const Grid1DDevice<float> &position_y,
                               const Grid1DDevice<float> &position_z, const Grid3DDevice<T> &u,
                               const Grid3DDevice<T> &v, const Grid3DDevice<T> &w, const BoundaryConditionSet &bc,
                               float hx, float hy, float hz);

}  // namespace ocu

#endif
