// Eqn_IncompressibleNS3D<T> (eqn_incompressns3d.cu:25-399) and
// Eqn_ScalarAdvection3D<T> (eqn_scalaradvection3d.cpp) over libocu_hip.so.
#include "ocuequation/eqn_incompressns3d.h"

#include <algorithm>
#include <cstdio>

#include "ocu_hip.h"
#include "ocuequation/eqn_scalaradvection3d.h"
#include "ocustorage/grid3dboundary.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/kernel_wrapper.h"
#include "ocuutil/thread.h"

namespace ocu {

template <typename T>
Eqn_IncompressibleNS3DBase<T>::Eqn_IncompressibleNS3DBase() {
  num_steps = 0;
  verbose = false;
  _nx = _ny = _nz = 0;
  _hx = _hy = _hz = 0;
  _lastdt = 0;
  _max_divergence = 0;
  _cfl_factor = 0;
  _bouyancy = 0;
  _gravity = 0;
  _time_step = TS_ERROR;
  _vertical_direction = DIR_ZPOS;
}

// eqn_incompressns3d.cu:68-117
template <typename T>
bool Eqn_IncompressibleNS3DBase<T>::set_base_parameters(const Eqn_IncompressibleNS3DParams<T> &params) {
  _max_divergence = params.max_divergence;
  _cfl_factor = params.cfl_factor;
  _bouyancy = params.bouyancy;
  _gravity = params.gravity;
  if (!check_float(_max_divergence) || _max_divergence < 0 || _max_divergence > 1) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - invalid max_divergence %f\n", _max_divergence);
    return false;
  }
  if (_cfl_factor < 0) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - invalid cfl_factor %f\n", _cfl_factor);
    return false;
  }
  _time_step = params.time_step;
  if (_time_step != TS_ADAMS_BASHFORD2 && _time_step != TS_FORWARD_EULER) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - unknown timestep type %d\n", _time_step);
    return false;
  }
  _nx = params.nx; _ny = params.ny; _nz = params.nz;
  _hx = params.hx; _hy = params.hy; _hz = params.hz;
  _lastdt = 0;
  DirectionType vd = params.vertical_direction;
  if (vd != DIR_XPOS && vd != DIR_XNEG && vd != DIR_YPOS && vd != DIR_YNEG && vd != DIR_ZPOS && vd != DIR_ZNEG) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - bad vertical_direction %d\n", vd);
    return false;
  }
  _vertical_direction = vd;
  if (!check_float(_hx) || !check_float(_hy) || !check_float(_hz) || _hx <= 0 || _hy <= 0 || _hz <= 0) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - bad hx, hy, hz (%f, %f, %f)\n", _hx, _hy, _hz);
    return false;
  }
  _thermalbc = params.temp_bc;
  return true;
}

template <typename T>
Eqn_IncompressibleNS3D<T>::Eqn_IncompressibleNS3D() : _mixed_projection(false), _fused(false) {}

// eqn_incompressns3d.cu:119-155: the vertical component gets
// .5 * (dirmult * g * beta) * (T + T_{-vertical})
template <typename T>
void Eqn_IncompressibleNS3D<T>::add_thermal_force() {
  T direction_mult = (this->_vertical_direction & DIR_NEGATIVE_FLAG) ? 1 : -1;
  Grid3DDevice<T> &target = (this->_vertical_direction & DIR_XAXIS_FLAG) ? _deriv_udt
                            : (this->_vertical_direction & DIR_YAXIS_FLAG) ? _deriv_vdt
                                                                           : _deriv_wdt;
  ocu_grid3d d = target.desc(), t = _temp.desc();
  T coef = direction_mult * this->_gravity * this->_bouyancy;
  if (!ocu_status_ok(ocu_thermal_force(&d, &t, this->nx(), this->ny(), this->nz(), (int)this->_vertical_direction,
                                       (double)coef, ThreadManager::get_compute_stream()),
                     "Eqn_IncompressibleNS3D_add_thermal_force"))
    this->add_error();
}

// eqn_incompressns3d.cu:164-299
template <typename T>
bool Eqn_IncompressibleNS3D<T>::set_parameters(const Eqn_IncompressibleNS3DParams<T> &params) {
  if (!this->set_base_parameters(params)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on base parameters\n");
    return false;
  }
  if (!_u.init_congruent(params.init_u) || !_deriv_udt.init_congruent(params.init_u) ||
      !_last_deriv_udt.init_congruent(params.init_u) || !_u.copy_all_data(params.init_u)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on initializing u\n");
    return false;
  }
  if (!_v.init_congruent(params.init_v) || !_deriv_vdt.init_congruent(params.init_v) ||
      !_last_deriv_vdt.init_congruent(params.init_v) || !_v.copy_all_data(params.init_v)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on initializing v\n");
    return false;
  }
  if (!_w.init_congruent(params.init_w) || !_deriv_wdt.init_congruent(params.init_w) ||
      !_last_deriv_wdt.init_congruent(params.init_w) || !_w.copy_all_data(params.init_w)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on initializing w\n");
    return false;
  }
  if (!_temp.init_congruent(params.init_temp) || !_deriv_tempdt.init_congruent(params.init_temp) ||
      !_last_deriv_tempdt.init_congruent(params.init_temp) || !_temp.copy_all_data(params.init_temp)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on initializing temp\n");
    return false;
  }
  _projection_solver.bc = params.flow_bc;
  _advection_solver.interp_type = params.advection_scheme;
  _thermal_solver.interp_type = params.advection_scheme;
  _advection_solver.periodic_axes = params.flow_bc.periodic_axes();
  _thermal_solver.periodic_axes = params.temp_bc.periodic_axes();

  if (!_thermal_solver.initialize_storage(this->_nx, this->_ny, this->_nz, this->_hx, this->_hy, this->_hz, &_u, &_v,
                                          &_w, &_temp, &_deriv_tempdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _thermal_solver initialization\n");
    return false;
  }
  if (!_advection_solver.initialize_storage(this->_nx, this->_ny, this->_nz, this->_hx, this->_hy, this->_hz, &_u,
                                            &_v, &_w, &_deriv_udt, &_deriv_vdt, &_deriv_wdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _advection_solver initialization\n");
    return false;
  }
  _mixed_projection = params.mixed_precision_projection;
  if (_mixed_projection) {
    if constexpr (std::is_same<T, double>::value) {
      _mixed_projection_solver.bc = params.flow_bc;
      if (!_mixed_projection_solver.initialize_storage(this->_nx, this->_ny, this->_nz, this->_hx, this->_hy,
                                                       this->_hz, &_u, &_v, &_w)) {
        printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _mixed_projection_solver initialization\n");
        return false;
      }
    } else {
      printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - mixed-precision projection needs T = double\n");
      return false;
    }
  } else if (!_projection_solver.initialize_storage(this->_nx, this->_ny, this->_nz, this->_hx, this->_hy, this->_hz,
                                                    &_u, &_v, &_w)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _projection_solver initialization\n");
    return false;
  }
  if (!_thermal_diffusion.initialize_storage(this->_nx, this->_ny, this->_nz, this->_hx, this->_hy, this->_hz, &_temp,
                                             &_deriv_tempdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _thermal_diffusion initialization\n");
    return false;
  }
  if (!check_float(params.thermal_diffusion) || params.thermal_diffusion < 0) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - invalid thermal diffusion %f\n",
           (double)params.thermal_diffusion);
    return false;
  }
  _thermal_diffusion.coefficient = params.thermal_diffusion;
  if (!_u_diffusion.initialize_storage(this->_nx + 1, this->_ny, this->_nz, this->_hx, this->_hy, this->_hz, &_u,
                                       &_deriv_udt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _u_diffusion initialization\n");
    return false;
  }
  if (!check_float(params.viscosity) || params.viscosity < 0) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - invalid viscosity %f\n", (double)params.viscosity);
    return false;
  }
  _u_diffusion.coefficient = params.viscosity;
  if (!_v_diffusion.initialize_storage(this->_nx, this->_ny + 1, this->_nz, this->_hx, this->_hy, this->_hz, &_v,
                                       &_deriv_vdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _v_diffusion initialization\n");
    return false;
  }
  _v_diffusion.coefficient = params.viscosity;
  if (!_w_diffusion.initialize_storage(this->_nx, this->_ny, this->_nz + 1, this->_hx, this->_hy, this->_hz, &_w,
                                       &_deriv_wdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on _w_diffusion initialization\n");
    return false;
  }
  _w_diffusion.coefficient = params.viscosity;
  if (!apply_3d_mac_boundary_conditions_level1(_u, _v, _w, params.flow_bc, this->_hx, this->_hy, this->_hz)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on enforcing flow boundary conditions\n");
    return false;
  }
  if (!apply_3d_boundary_conditions_level1(_temp, this->_thermalbc, this->_hx, this->_hy, this->_hz)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on enforcing thermal boundary conditions\n");
    return false;
  }
  _deriv_udt.clear_zero();
  _deriv_vdt.clear_zero();
  _deriv_wdt.clear_zero();
  _deriv_tempdt.clear_zero();
  _last_deriv_udt.clear_zero();
  _last_deriv_vdt.clear_zero();
  _last_deriv_wdt.clear_zero();
  _last_deriv_tempdt.clear_zero();
  if (!_u.check_layout_match(_v) || !_u.check_layout_match(_w) || !_u.check_layout_match(_deriv_udt) ||
      !_u.check_layout_match(_temp) || !_u.check_layout_match(_deriv_tempdt)) {
    printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - grid layouts do not all match\n");
    return false;
  }
  // The fused step needs one congruent layout for the four grids (what
  // init_grids gives) and flow faces the projection's MAC fill rewrites
  // (periodic or forced inflow), since it leaves the normal face n of u, v,
  // w to that fill.
  _fused = false;
  if (params.fused_step && params.flow_bc.check_type(BC_PERIODIC, BC_FORCED_INFLOW_VARIABLE_SLIP) &&
      _u.shift_amount() == _v.shift_amount() && _u.shift_amount() == _w.shift_amount() &&
      _u.shift_amount() == _temp.shift_amount() && _u.gx() == _temp.gx() && _u.gy() == _temp.gy() &&
      _u.gz() == _temp.gz() && _u.num_allocated_elements() == _temp.num_allocated_elements() &&
      _u.gx() == _v.gx() && _u.gy() == _v.gy() && _u.gz() == _v.gz() && _u.gx() == _w.gx() &&
      _u.gy() == _w.gy() && _u.gz() == _w.gz()) {
    if (!_u_next.init_congruent(_u) || !_v_next.init_congruent(_v) || !_w_next.init_congruent(_w) ||
        !_temp_next.init_congruent(_temp)) {
      printf("[ERROR] Eqn_IncompressibleNS3D::set_parameters - failed on the fused step's buffers\n");
      return false;
    }
    _fused = true;
  }
  return true;
}

// eqn_incompressns3d.cu:301-328: cfl * min(h / max|u|), capped by the
// explicit-diffusion limits h^2 / (6 nu) and h^2 / (6 kappa)
template <typename T>
double Eqn_IncompressibleNS3D<T>::get_max_stable_timestep() const {
  T max_u = 0, max_v = 0, max_w = 0;
  Grid3DDevice<T>::reduce_maxabs3(_u, _v, _w, max_u, max_v, max_w);
  double ut = this->hx() / max_u;
  double vt = this->hy() / max_v;
  double wt = this->hz() / max_w;
  if (!check_float(ut)) ut = 1e10;
  if (!check_float(vt)) vt = 1e10;
  if (!check_float(wt)) wt = 1e10;
  double step = this->_cfl_factor * min3(ut, vt, wt);
  double minh = min3(this->hx(), this->hy(), this->hz());
  if (thermal_diffusion_coefficient() > 0)
    step = std::min(step, (minh * minh) / (6 * thermal_diffusion_coefficient()));
  if (viscosity_coefficient() > 0) step = std::min(step, (minh * minh) / (6 * viscosity_coefficient()));
  if (this->verbose) printf("Eqn_IncompressibleNS3D<T>::get_max_stable_timestep - return %f\n", step);
  return step;
}

// The explicit part of the step as one ocu_ns_step_fused launch into the
// _next buffers, which are then swapped in; the projection follows as in
// the reference.  Same values as the sequence below (ocu_ns_step_fused).
template <typename T>
bool Eqn_IncompressibleNS3D<T>::advance_fused(double dt) {
  T ab_coeff = -dt * dt / (2 * this->_lastdt);
  bool ab2 = this->_time_step == TS_ADAMS_BASHFORD2 && this->_lastdt > 0;
  bool keep = this->_time_step == TS_ADAMS_BASHFORD2;
  ocu_ns_step_params sp;
  sp.nx = this->nx(); sp.ny = this->ny(); sp.nz = this->nz();
  sp.advection_scheme = (int)_advection_solver.interp_type;
  sp.vertical_direction = (int)this->_vertical_direction;
  sp.flow_periodic = _advection_solver.periodic_axes;
  sp.temp_periodic = _thermal_solver.periodic_axes;
  sp.keep_tendencies = keep;
  sp.use_last = ab2;
  sp.hx = this->_hx; sp.hy = this->_hy; sp.hz = this->_hz;
  sp.viscosity = viscosity_coefficient() > 0 ? (double)viscosity_coefficient() : 0.0;
  sp.thermal_diffusion = thermal_diffusion_coefficient() > 0 ? (double)thermal_diffusion_coefficient() : 0.0;
  T direction_mult = (this->_vertical_direction & DIR_NEGATIVE_FLAG) ? 1 : -1;
  T coef = direction_mult * this->_gravity * this->_bouyancy;
  sp.thermal_coef = (double)coef;
  sp.a_dt = ab2 ? (double)(T)(dt - ab_coeff) : (double)(T)dt;
  sp.a_ab = ab2 ? (double)ab_coeff : 0.0;
  ocu_grid3d u = _u.desc(), v = _v.desc(), w = _w.desc(), t = _temp.desc();
  ocu_grid3d un = _u_next.desc(), vn = _v_next.desc(), wn = _w_next.desc(), tn = _temp_next.desc();
  ocu_grid3d du = _deriv_udt.desc(), dv = _deriv_vdt.desc(), dw = _deriv_wdt.desc(), dt_ = _deriv_tempdt.desc();
  ocu_grid3d lu = _last_deriv_udt.desc(), lv = _last_deriv_vdt.desc(), lw = _last_deriv_wdt.desc(),
             lt = _last_deriv_tempdt.desc();
  ocu_bc_set tbc = to_c_bc(this->_thermalbc);
  if (!ocu_status_ok(ocu_ns_step_fused(&u, &v, &w, &t, &un, &vn, &wn, &tn, &du, &dv, &dw, &dt_, &lu, &lv, &lw, &lt,
                                       &sp, &tbc, ThreadManager::get_compute_stream()),
                     "Eqn_IncompressibleNS3D::advance_one_step (fused)")) {
    this->add_error();
    return false;
  }
  _u.swap_storage(_u_next);
  _v.swap_storage(_v_next);
  _w.swap_storage(_w_next);
  _temp.swap_storage(_temp_next);
  if (keep) {
    // the reference copies the tendencies into the last-tendency grids
    this->_lastdt = dt;
    _last_deriv_tempdt.swap_storage(_deriv_tempdt);
    _last_deriv_udt.swap_storage(_deriv_udt);
    _last_deriv_vdt.swap_storage(_deriv_vdt);
    _last_deriv_wdt.swap_storage(_deriv_wdt);
  }
  return true;
}

// eqn_incompressns3d.cu:331-399
template <typename T>
bool Eqn_IncompressibleNS3D<T>::advance_one_step(double dt) {
  this->clear_error();
  this->num_steps++;
  if (_fused) {
    if (!advance_fused(dt)) return false;
    if constexpr (std::is_same<T, double>::value) {
      if (_mixed_projection) {
        this->check_ok(_mixed_projection_solver.solve(this->_max_divergence));
        return !this->any_error();
      }
    }
    this->check_ok(_projection_solver.solve(this->_max_divergence));
    return !this->any_error();
  }
  this->check_ok(_advection_solver.solve());
  if (viscosity_coefficient() > 0) {
    this->check_ok(_u_diffusion.solve());
    this->check_ok(_v_diffusion.solve());
    this->check_ok(_w_diffusion.solve());
  }
  add_thermal_force();
  this->check_ok(_thermal_solver.solve());
  if (thermal_diffusion_coefficient() > 0) this->check_ok(_thermal_diffusion.solve());

  T ab_coeff = -dt * dt / (2 * this->_lastdt);
  bool ab2 = this->_time_step == TS_ADAMS_BASHFORD2 && this->_lastdt > 0;
  if (ab2) {
    this->check_ok(_temp.linear_combination((T)1.0, _temp, (T)(dt - ab_coeff), _deriv_tempdt));
    this->check_ok(_temp.linear_combination((T)1.0, _temp, (T)ab_coeff, _last_deriv_tempdt));
  } else {
    this->check_ok(_temp.linear_combination((T)1.0, _temp, (T)dt, _deriv_tempdt));
  }
  this->check_ok(apply_3d_boundary_conditions_level1_nocorners(_temp, this->_thermalbc, this->_hx, this->_hy, this->_hz));

  if (ab2) {
    this->check_ok(_u.linear_combination((T)1.0, _u, (T)(dt - ab_coeff), _deriv_udt));
    this->check_ok(_u.linear_combination((T)1.0, _u, (T)ab_coeff, _last_deriv_udt));
    this->check_ok(_v.linear_combination((T)1.0, _v, (T)(dt - ab_coeff), _deriv_vdt));
    this->check_ok(_v.linear_combination((T)1.0, _v, (T)ab_coeff, _last_deriv_vdt));
    this->check_ok(_w.linear_combination((T)1.0, _w, (T)(dt - ab_coeff), _deriv_wdt));
    this->check_ok(_w.linear_combination((T)1.0, _w, (T)ab_coeff, _last_deriv_wdt));
  } else {
    this->check_ok(_u.linear_combination((T)1.0, _u, (T)dt, _deriv_udt));
    this->check_ok(_v.linear_combination((T)1.0, _v, (T)dt, _deriv_vdt));
    this->check_ok(_w.linear_combination((T)1.0, _w, (T)dt, _deriv_wdt));
  }
  if (this->_time_step == TS_ADAMS_BASHFORD2) {
    this->_lastdt = dt;
    _last_deriv_tempdt.copy_all_data(_deriv_tempdt);
    _last_deriv_udt.copy_all_data(_deriv_udt);
    _last_deriv_vdt.copy_all_data(_deriv_vdt);
    _last_deriv_wdt.copy_all_data(_deriv_wdt);
  }
  if constexpr (std::is_same<T, double>::value) {
    if (_mixed_projection) {
      this->check_ok(_mixed_projection_solver.solve(this->_max_divergence));
      return !this->any_error();
    }
  }
  this->check_ok(_projection_solver.solve(this->_max_divergence));
  return !this->any_error();
}

template class Eqn_IncompressibleNS3DBase<float>;
template class Eqn_IncompressibleNS3DBase<double>;
template class Eqn_IncompressibleNS3D<float>;
template class Eqn_IncompressibleNS3D<double>;

// ------------------------------------------------------- scalar advection

// eqn_scalaradvection3d.cpp:26-68
template <typename T>
bool Eqn_ScalarAdvection3D<T>::set_parameters(const Eqn_ScalarAdvection3DParams<T> &params) {
  bc = params.bc;
  _advection.interp_type = params.advection_scheme;
  _advection.periodic_axes = params.bc.periodic_axes();
  if (!u.init_congruent(params.u) || !v.init_congruent(params.v) || !w.init_congruent(params.w)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to initialize u,v,w grids\n");
    return false;
  }
  if (!u.copy_all_data(params.u) || !v.copy_all_data(params.v) || !w.copy_all_data(params.w)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to copy data to u,v,w grids\n");
    return false;
  }
  if (!phi.init_congruent(params.initial_values) || !phi.copy_all_data(params.initial_values)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to initialize phi grid\n");
    return false;
  }
  if (!deriv_phidt.init_congruent(params.initial_values)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to initialize deriv_phidt grid\n");
    return false;
  }
  deriv_phidt.clear_zero();
  if (!_advection.initialize_storage(params.nx, params.ny, params.nz, params.hx, params.hy, params.hz, &u, &v, &w,
                                     &phi, &deriv_phidt)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to initialize advection solver.\n");
    return false;
  }
  if (!apply_3d_boundary_conditions_level1_nocorners(phi, bc, hx(), hy(), hz())) {
    printf("[ERROR] Eqn_ScalarAdvection3D::set_parameters - failed to enforce boundary conditions.\n");
    return false;
  }
  return true;
}

// eqn_scalaradvection3d.cpp:70-86
template <typename T>
double Eqn_ScalarAdvection3D<T>::get_max_stable_timestep() const {
  T max_u = 0, max_v = 0, max_w = 0;
  Grid3DDevice<T>::reduce_maxabs3(u, v, w, max_u, max_v, max_w);
  double ut = hx() / max_u, vt = hy() / max_v, wt = hz() / max_w;
  if (!check_float(ut)) ut = 1e10;
  if (!check_float(vt)) vt = 1e10;
  if (!check_float(wt)) wt = 1e10;
  return min3(ut, vt, wt);
}

// eqn_scalaradvection3d.cpp:88-105
template <typename T>
bool Eqn_ScalarAdvection3D<T>::advance_one_step(double dt) {
  clear_error();
  if (!check_float(dt)) {
    printf("[ERROR] Eqn_ScalarAdvection3D::advance_one_step - bad dt value\n");
    return false;
  }
  check_ok(_advection.solve());
  check_ok(phi.linear_combination((T)1, phi, (T)dt, deriv_phidt));
  check_ok(apply_3d_boundary_conditions_level1_nocorners(phi, bc, hx(), hy(), hz()));
  return !any_error();
}

template class Eqn_ScalarAdvection3D<float>;
template class Eqn_ScalarAdvection3D<double>;

}  // namespace ocu
