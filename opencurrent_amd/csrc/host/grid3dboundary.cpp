// Ghost fills over the C ABI (ocustorage/grid3dboundary.cu:719-900).
#include <cstddef>
#include <cstring>

#include "ocu_hip.h"
#include "ocustorage/grid3dboundary.h"
#include "ocuutil/kernel_wrapper.h"

namespace ocu {

static_assert(sizeof(BoundaryCondition) == sizeof(ocu_bc), "BoundaryCondition layout");
static_assert(offsetof(BoundaryCondition, value) == offsetof(ocu_bc, value), "BoundaryCondition layout");
static_assert(sizeof(BoundaryConditionSet) == sizeof(ocu_bc_set), "BoundaryConditionSet layout");

ocu_bc_set to_c_bc(const BoundaryConditionSet &bc) {
  ocu_bc_set r;
  memcpy(&r, &bc, sizeof(r));
  return r;
}

template <typename T>
bool apply_3d_boundary_conditions_level1(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc, double hx,
                                         double hy, double hz) {
  ocu_grid3d d = grid.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_bc_fill_level1(&d, &b, hx, hy, hz, ThreadManager::get_compute_stream()),
                       "apply_3d_boundary_conditions_level1");
}

template <typename T>
bool apply_3d_boundary_conditions_level1_nocorners(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc,
                                                   double hx, double hy, double hz) {
  ocu_grid3d d = grid.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_bc_fill_nocorners(&d, &b, hx, hy, hz, ThreadManager::get_compute_stream()),
                       "apply_3d_boundary_conditions_level1_nocorners");
}

template <typename T>
bool apply_3d_mac_boundary_conditions_level1(Grid3DDevice<T> &u, Grid3DDevice<T> &v, Grid3DDevice<T> &w,
                                             const BoundaryConditionSet &bc, double, double, double) {
  ocu_grid3d du = u.desc(), dv = v.desc(), dw = w.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_bc_fill_mac_level1(&du, &dv, &dw, &b, ThreadManager::get_compute_stream()),
                       "apply_3d_mac_boundary_conditions_level1");
}

template <typename T>
bool apply_3d_boundary_conditions_level2_nocorners(Grid3DDevice<T> &grid, const BoundaryConditionSet &bc,
                                                   double hx, double hy, double hz) {
  ocu_grid3d d = grid.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_bc_fill_level2_nocorners(&d, &b, hx, hy, hz, ThreadManager::get_compute_stream()),
                       "apply_3d_boundary_conditions_level2_nocorners");
}

template <typename T>
bool apply_3d_mac_boundary_conditions_level2(Grid3DDevice<T> &u, Grid3DDevice<T> &v, Grid3DDevice<T> &w,
                                             const BoundaryConditionSet &bc, double, double, double) {
  ocu_grid3d du = u.desc(), dv = v.desc(), dw = w.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_bc_fill_mac_level2(&du, &dv, &dw, &b, ThreadManager::get_compute_stream()),
                       "apply_3d_mac_boundary_conditions_level2");
}

template bool apply_3d_boundary_conditions_level2_nocorners(Grid3DDevice<float> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_boundary_conditions_level2_nocorners(Grid3DDevice<double> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_mac_boundary_conditions_level2(Grid3DDevice<float> &, Grid3DDevice<float> &, Grid3DDevice<float> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_mac_boundary_conditions_level2(Grid3DDevice<double> &, Grid3DDevice<double> &, Grid3DDevice<double> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_boundary_conditions_level1(Grid3DDevice<float> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_boundary_conditions_level1(Grid3DDevice<double> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_boundary_conditions_level1_nocorners(Grid3DDevice<float> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_boundary_conditions_level1_nocorners(Grid3DDevice<double> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_mac_boundary_conditions_level1(Grid3DDevice<float> &, Grid3DDevice<float> &, Grid3DDevice<float> &, const BoundaryConditionSet &, double, double, double);
template bool apply_3d_mac_boundary_conditions_level1(Grid3DDevice<double> &, Grid3DDevice<double> &, Grid3DDevice<double> &, const BoundaryConditionSet &, double, double, double);

}  // namespace ocu
