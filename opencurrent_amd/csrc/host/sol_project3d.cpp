// Projection of a MAC velocity onto divergence-free fields:
// Sol_ProjectDivergence3DDevice<T> (sol_project3d.cpp:26-70,
// sol_project3ddev.cu:28-135) and its mixed-precision variant.
#include "ocuequation/sol_project3d.h"

#include <cstdio>

#include "ocustorage/grid3dboundary.h"
#include "ocustorage/grid3dops.h"
#include "ocuutil/float_routines.h"

namespace ocu {

// sol_project3d.cpp:26-43
BoundaryCondition Sol_ProjectDivergence3DBase::convert_bc_to_poisson_eqn(const BoundaryCondition &b) const {
  BoundaryCondition ret;
  if (b.type == BC_PERIODIC) {
    ret.type = BC_PERIODIC;
  } else if (b.type == BC_FORCED_INFLOW_VARIABLE_SLIP) {
    ret.type = BC_NEUMANN;
    ret.value = 0;
  } else {
    printf("[ERROR] Sol_ProjectDivergence3DBase::convert_bc_to_poisson_eqn - invalid boundary condition type %d\n",
           b.type);
  }
  return ret;
}

bool Sol_ProjectDivergence3DBase::initialize_base_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                          Grid3DUntyped *u_val, Grid3DUntyped *v_val,
                                                          Grid3DUntyped *w_val) {
  if (!check_valid_mac_dimensions(*u_val, *v_val, *w_val, nx, ny, nz)) {
    printf("[ERROR] Sol_ProjectDivergence3DBase::initialize_base_storage - u,v,w grid dimensions mismatch\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_ProjectDivergence3DBase::initialize_base_storage - garbage hx,hy,hz value\n");
    return false;
  }
  _hx = hx; _hy = hy; _hz = hz;
  _nx = nx; _ny = ny; _nz = nz;
  return true;
}

// sol_project3ddev.cu:28-51: divergence congruent with u
template <typename T>
bool Sol_ProjectDivergence3DDeviceStorage<T>::initialize_device_storage(int nx, int ny, int nz, double hx, double hy,
                                                                        double hz, Grid3DDevice<T> *u_val,
                                                                        Grid3DDevice<T> *v_val,
                                                                        Grid3DDevice<T> *w_val) {
  if (!this->initialize_base_storage(nx, ny, nz, hx, hy, hz, u_val, v_val, w_val)) return false;
  u = u_val; v = v_val; w = w_val;
  if (!divergence.init(this->_nx, this->_ny, this->_nz, 1, 1, 1, u->pnx() - (this->_nx + 2),
                       u->pny() - (this->_ny + 2), u->pnz() - (this->_nz + 2))) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_device_storage - failed to initialize divergence\n");
    return false;
  }
  if (!u_val->check_layout_match(divergence)) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_device_storage - divergence layout mismatch\n");
    return false;
  }
  divergence.clear_zero();
  return true;
}

// sol_project3ddev.cu:58-71: MAC BC, divergence, pressure solve (15
// V-cycles, Linf), gradient subtraction, MAC BC
template <typename T>
bool Sol_ProjectDivergence3DDevice<T>::solve(double tolerance) {
  this->clear_error();
  double residual = 0;
  this->check_ok(apply_3d_mac_boundary_conditions_level1(*this->u, *this->v, *this->w, this->bc, this->_hx, this->_hy,
                                                         this->_hz),
                 "Sol_ProjectDivergence3DDevice::solve - could not enforce boundary conditions");
  this->check_ok(divergence_solver.solve(), "Sol_ProjectDivergence3DDevice::solve - could not calculate divergence");
  this->check_ok(pressure_solver.solve(residual, tolerance, 15),
                 "Sol_ProjectDivergence3DDevice::solve - could not solve for pressure");
  last_residual = residual;
  this->check_ok(gradient_solver.solve(),
                 "Sol_ProjectDivergence3DDevice::solve - could not subtract gradient of pressure");
  this->check_ok(apply_3d_mac_boundary_conditions_level1(*this->u, *this->v, *this->w, this->bc, this->_hx, this->_hy,
                                                         this->_hz),
                 "Sol_ProjectDivergence3DDevice::solve - could not enforce boundary conditions");
  return !this->any_error();
}

template <typename T>
bool Sol_ProjectDivergence3DDevice<T>::solve_divergence_only() {
  this->clear_error();
  this->check_ok(apply_3d_mac_boundary_conditions_level1(*this->u, *this->v, *this->w, this->bc, this->_hx, this->_hy,
                                                         this->_hz),
                 "Sol_ProjectDivergence3DDevice::solve - could not enforce boundary conditions");
  this->check_ok(divergence_solver.solve(), "Sol_ProjectDivergence3DDevice::solve - could not calculate divergence");
  return !this->any_error();
}

// sol_project3ddev.cu:76-135
template <typename T>
bool Sol_ProjectDivergence3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                          Grid3DDevice<T> *u_val, Grid3DDevice<T> *v_val,
                                                          Grid3DDevice<T> *w_val) {
  if (!this->initialize_device_storage(nx, ny, nz, hx, hy, hz, u_val, v_val, w_val)) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_storage - failed to initialize device storage\n");
    return false;
  }
  pressure_solver.bc.xneg = this->convert_bc_to_poisson_eqn(this->bc.xneg);
  pressure_solver.bc.xpos = this->convert_bc_to_poisson_eqn(this->bc.xpos);
  pressure_solver.bc.yneg = this->convert_bc_to_poisson_eqn(this->bc.yneg);
  pressure_solver.bc.ypos = this->convert_bc_to_poisson_eqn(this->bc.ypos);
  pressure_solver.bc.zneg = this->convert_bc_to_poisson_eqn(this->bc.zneg);
  pressure_solver.bc.zpos = this->convert_bc_to_poisson_eqn(this->bc.zpos);
  if (!divergence_solver.initialize_storage(nx, ny, nz, hx, hy, hz, this->u, this->v, this->w, &this->divergence)) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_storage - failed to initialize divergence_solver\n");
    return false;
  }
  if (!pressure_solver.initialize_storage(nx, ny, nz, hx, hy, hz, &this->divergence)) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_storage - failed to initialize pressure_solver\n");
    return false;
  }
  if (!gradient_solver.initialize_storage(nx, ny, nz, hx, hy, hz, this->u, this->v, this->w,
                                          &pressure_solver.pressure())) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_storage - failed to initialize gradient_solver\n");
    return false;
  }
  gradient_solver.coefficient = -1;
  if (!u_val->check_layout_match(pressure_solver.pressure())) {
    printf("[ERROR] Sol_ProjectDivergence3DDevice::initialize_storage - pressure layout mismatch\n");
    return false;
  }
  pressure_solver.pressure().clear_zero();
  pressure_solver.convergence = CONVERGENCE_LINF;
  return true;
}

// ---------------------------------------------------------- mixed precision

bool Sol_ProjectDivergenceMixed3DDeviceD::solve(double tolerance) {
  clear_error();
  double residual = 0;
  check_ok(apply_3d_mac_boundary_conditions_level1(*u, *v, *w, bc, _hx, _hy, _hz),
           "Sol_ProjectDivergenceMixed3DDeviceD::solve - could not enforce boundary conditions");
  check_ok(divergence_solver.solve(), "Sol_ProjectDivergenceMixed3DDeviceD::solve - could not calculate divergence");
  check_ok(pressure_solver.solve(residual, tolerance, 15),
           "Sol_ProjectDivergenceMixed3DDeviceD::solve - could not solve for pressure");
  last_residual = residual;
  check_ok(gradient_solver.solve(), "Sol_ProjectDivergenceMixed3DDeviceD::solve - could not subtract gradient");
  check_ok(apply_3d_mac_boundary_conditions_level1(*u, *v, *w, bc, _hx, _hy, _hz),
           "Sol_ProjectDivergenceMixed3DDeviceD::solve - could not enforce boundary conditions");
  return !any_error();
}

bool Sol_ProjectDivergenceMixed3DDeviceD::solve_divergence_only() {
  clear_error();
  check_ok(apply_3d_mac_boundary_conditions_level1(*u, *v, *w, bc, _hx, _hy, _hz));
  check_ok(divergence_solver.solve());
  return !any_error();
}

bool Sol_ProjectDivergenceMixed3DDeviceD::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                             Grid3DDevice<double> *u_val, Grid3DDevice<double> *v_val,
                                                             Grid3DDevice<double> *w_val) {
  if (!initialize_device_storage(nx, ny, nz, hx, hy, hz, u_val, v_val, w_val)) return false;
  pressure_solver.bc.xneg = convert_bc_to_poisson_eqn(bc.xneg);
  pressure_solver.bc.xpos = convert_bc_to_poisson_eqn(bc.xpos);
  pressure_solver.bc.yneg = convert_bc_to_poisson_eqn(bc.yneg);
  pressure_solver.bc.ypos = convert_bc_to_poisson_eqn(bc.ypos);
  pressure_solver.bc.zneg = convert_bc_to_poisson_eqn(bc.zneg);
  pressure_solver.bc.zpos = convert_bc_to_poisson_eqn(bc.zpos);
  pressure_solver.convergence = CONVERGENCE_LINF;
  if (!divergence_solver.initialize_storage(nx, ny, nz, hx, hy, hz, u, v, w, &divergence)) return false;
  if (!pressure_solver.initialize_storage(nx, ny, nz, hx, hy, hz, &divergence)) return false;
  if (!gradient_solver.initialize_storage(nx, ny, nz, hx, hy, hz, u, v, w, &pressure_solver.pressure())) return false;
  gradient_solver.coefficient = -1;
  pressure_solver.pressure().clear_zero();
  return true;
}

template class Sol_ProjectDivergence3DDeviceStorage<float>;
template class Sol_ProjectDivergence3DDeviceStorage<double>;
template class Sol_ProjectDivergence3DDevice<float>;
template class Sol_ProjectDivergence3DDevice<double>;

}  // namespace ocu
