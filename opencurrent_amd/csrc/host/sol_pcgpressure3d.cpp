// Sol_PCGPressure3DDevice<T> (ocuequation/sol_pcgpressure3d.cu:27-662) over
// libocu_hip.so.
#include "ocuequation/sol_pcgpressure3d.h"

#include <cstdio>

#include "ocu_hip.h"
#include "ocustorage/grid3dboundary.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/thread.h"

namespace ocu {

// sol_pcgpressure3d.cu:213-226
template <typename T>
Sol_PCGPressure3DDevice<T>::Sol_PCGPressure3DDevice()
    : _ws(nullptr), _host(nullptr), _host_dev(nullptr), _iterations(0) {
  _fx = _fy = _fz = 0;
  _h = 1;
  _nx = _ny = _nz = 0;
  _rhs = _coefficient = nullptr;
  convergence = CONVERGENCE_L2;
  preconditioner = PRECOND_BFBT;
  multigrid_use_fmg = true;
  multigrid_cycles = 2;
  verbose = false;
  _mg.convergence = CONVERGENCE_NONE_CALC_L2LINF;
}

template <typename T>
Sol_PCGPressure3DDevice<T>::~Sol_PCGPressure3DDevice() {
  if (_ws) ocu_free(_ws);
  if (_host) ocu_host_free(_host);
}

// sol_pcgpressure3d.cu:183-220: the no-corner fill of src, then the operator
template <typename T>
T Sol_PCGPressure3DDevice<T>::invoke_kernel_apply_laplacian(Grid3DDevice<T> &dst, Grid3DDevice<T> &src,
                                                            bool with_dot) {
  check_ok(apply_3d_boundary_conditions_level1_nocorners(src, bc, hx(), hy(), hz()));
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d d = dst.desc(), sr = src.desc(), c = _coefficient->desc();
  if (!ok_status(ocu_pcg_apply_laplacian(&d, &sr, &c, (double)fx_div_hsq(), (double)fy_div_hsq(),
                                         (double)fz_div_hsq(), _ws, with_dot ? _host_dev : nullptr, s),
                 "Sol_PCGPressure3DDevice::invoke_kernel_apply_laplacian"))
    return 0;
  if (!with_dot) return 0;
  if (!ok_status(ocu_stream_synchronize(s), "Sol_PCGPressure3DDevice::invoke_kernel_apply_laplacian")) return 0;
  return (T)*_host;
}

// sol_pcgpressure3d.cu:227-259
template <typename T>
void Sol_PCGPressure3DDevice<T>::invoke_kernel_diag_preconditioner(Grid3DDevice<T> &dst, const Grid3DDevice<T> &src) {
  ocu_grid3d d = dst.desc(), sr = src.desc(), c = _coefficient->desc();
  ok_status(ocu_pcg_diag_precond(&d, &sr, &c, (double)fx_div_hsq(), (double)fy_div_hsq(), (double)fz_div_hsq(),
                                 ThreadManager::get_compute_stream()),
            "Sol_PCGPressure3DDevice::invoke_kernel_diag_preconditioner");
}

// sol_pcgpressure3d.cu:262-293: sum of the T products, in double, rounded to T
template <typename T>
T Sol_PCGPressure3DDevice<T>::invoke_kernel_dot_product(const Grid3DDevice<T> &a, const Grid3DDevice<T> &b) {
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d da = a.desc(), db = b.desc();
  if (!ok_status(ocu_grid_dot(&da, &db, _ws, _host_dev, s), "Sol_PCGPressure3DDevice::invoke_kernel_dot_product") ||
      !ok_status(ocu_stream_synchronize(s), "Sol_PCGPressure3DDevice::invoke_kernel_dot_product"))
    return 0;
  return (T)*_host;
}

// sol_pcgpressure3d.cu:295-336
template <typename T>
void Sol_PCGPressure3DDevice<T>::apply_preconditioner(Grid3DDevice<T> &dst, const Grid3DDevice<T> &src) {
  if (preconditioner == PRECOND_JACOBI) {
    invoke_kernel_diag_preconditioner(dst, src);
  } else if (preconditioner == PRECOND_MULTIGRID || preconditioner == PRECOND_BFBT) {
    // the multigrid's right-hand side is _d_r, which is src in every call
    _mg.pressure().clear_zero();
    if (multigrid_use_fmg) check_ok(_mg.run_fmg(multigrid_cycles));
    else check_ok(_mg.run_vcycles(multigrid_cycles));
    // the sign flips: the multigrid's Laplacian is negative definite
    check_ok(dst.linear_combination((T)-1, _mg.pressure()));
    if (preconditioner == PRECOND_BFBT) invoke_kernel_diag_preconditioner(dst, dst);
  } else {
    check_ok(dst.copy_all_data(src));
  }
}

// sol_pcgpressure3d.cu:339-447
template <typename T>
bool Sol_PCGPressure3DDevice<T>::do_pcg(double tolerance, int max_iter, double &result_l2, double &result_linf) {
  // the coefficient's ghosts follow the pressure's boundary conditions
  check_ok(apply_3d_boundary_conditions_level1_nocorners(*_coefficient, bc, hx(), hy(), hz()));
  Grid3DDevice<T> &d_y = _d_y, &d_d = _d_d;

  invoke_kernel_apply_laplacian(d_y, _pressure);                    // y = L p
  check_ok(_d_r.linear_combination((T)1, *_rhs, (T)-1, d_y));      // r = rhs - y
  apply_preconditioner(d_d, _d_r);                                  // d = M^-1 r
  T delta0 = invoke_kernel_dot_product(_d_r, d_d);
  T deltanew = delta0;
  printf("[INFO] Sol_PCGPressure3DDevice<T>::do_pcg - initial residual norm %g\n", sqrt((double)delta0));

  int iteration_count = 0;
  bool converged = (iteration_count > max_iter) || (delta0 < tolerance * tolerance);
  while (!converged && !any_error()) {
    T d_dot_y = invoke_kernel_apply_laplacian(d_y, d_d, true);     // y = L d, <d, y>
    T alpha = deltanew / d_dot_y;
    check_ok(_pressure.linear_combination((T)1, _pressure, alpha, d_d));
    if (iteration_count % 25 == 0) {
      // every 25 iterations the residual is recomputed from the pressure
      invoke_kernel_apply_laplacian(d_y, _pressure);
      check_ok(_d_r.linear_combination((T)1, *_rhs, (T)-1, d_y));
    } else {
      check_ok(_d_r.linear_combination((T)1, _d_r, -alpha, d_y));
    }
    apply_preconditioner(d_y, _d_r);  // s = M^-1 r, held in y
    T deltaold = deltanew;
    deltanew = invoke_kernel_dot_product(_d_r, d_y);
    T beta = deltanew / deltaold;
    check_ok(d_d.linear_combination(beta, d_d, (T)1, d_y));      // d = s + beta d
    if (verbose) printf("%d: %g\n", iteration_count, sqrt((double)deltanew));
    iteration_count++;
    converged = (iteration_count > max_iter) || (deltanew < tolerance * tolerance);
  }
  _iterations = iteration_count;

  bool ok;
  if (deltanew <= tolerance * tolerance) {
    printf("[INFO] Sol_PCGPressure3DDevice<T>::do_pcg - Converged to %f residual in %d iterations\n",
           sqrt((double)deltanew), iteration_count);
    ok = true;
  } else {
    printf("[WARNING] Sol_PCGPressure3DDevice<T>::do_pcg - Failed to converge within %d iterations (achieved %f "
           "residual)\n",
           max_iter, sqrt((double)deltanew));
    ok = false;
  }
  T r_linf = 0;
  _d_r.reduce_maxabs(r_linf);
  result_linf = r_linf;
  result_l2 = sqrt((double)deltanew);
  return ok;
}

// sol_pcgpressure3d.cu:450-547: plain CG (not called by solve, kept for
// completeness of the class)
template <typename T>
bool Sol_PCGPressure3DDevice<T>::do_cg(double tolerance, int max_iter, double &result_l2, double &result_linf) {
  Grid3DDevice<T> &d_y = _d_y, &d_d = _d_d;
  invoke_kernel_apply_laplacian(d_y, _pressure);
  check_ok(_d_r.linear_combination((T)1, *_rhs, (T)-1, d_y));
  check_ok(d_d.copy_all_data(_d_r));
  T delta0 = invoke_kernel_dot_product(_d_r, _d_r);
  T deltanew = delta0;
  int iteration_count = 0;
  bool converged = (iteration_count > max_iter) || (delta0 < tolerance * tolerance);
  while (!converged && !any_error()) {
    T d_dot_y = invoke_kernel_apply_laplacian(d_y, d_d, true);
    T alpha = deltanew / d_dot_y;
    check_ok(_pressure.linear_combination((T)1, _pressure, alpha, d_d));
    if (iteration_count % 25 == 0) {
      invoke_kernel_apply_laplacian(d_y, _pressure);
      check_ok(_d_r.linear_combination((T)1, *_rhs, (T)-1, d_y));
    } else {
      check_ok(_d_r.linear_combination((T)1, _d_r, -alpha, d_y));
    }
    T deltaold = deltanew;
    deltanew = invoke_kernel_dot_product(_d_r, _d_r);
    T beta = deltanew / deltaold;
    check_ok(d_d.linear_combination(beta, d_d, (T)1, _d_r));
    iteration_count++;
    converged = (iteration_count > max_iter) || (deltanew < tolerance * tolerance);
  }
  _iterations = iteration_count;
  T r_linf = 0;
  _d_r.reduce_maxabs(r_linf);
  result_linf = r_linf;
  result_l2 = sqrt((double)deltanew);
  return deltanew <= tolerance * tolerance;
}

// sol_pcgpressure3d.cu:550-597
template <typename T>
bool Sol_PCGPressure3DDevice<T>::solve(double &residual, double tolerance, int max_iter) {
  clear_error();
  if (!check_float(tolerance)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::solve - garbage tolerance value %f\n", tolerance);
    return false;
  }
  if (tolerance < 0) {
    printf("[ERROR] Sol_PCGPressure3DDevice::solve - negative tolerance value %f\n", tolerance);
    return false;
  }
  if (max_iter <= 0) {
    printf("[WARNING] Sol_PCGPressure3DDevice::solve - non-positive max_iter %d\n", max_iter);
    return false;
  }
  if (!_rhs || !_ws) {
    printf("[ERROR] Sol_PCGPressure3DDevice::solve - storage not initialized\n");
    return false;
  }
  double l2 = 0, linf = 0;
  bool ok = true;
  if (!do_pcg(tolerance, max_iter, l2, linf)) {
    printf("[WARNING] Sol_PCGPressure3DDevice::solve - do_pcg failed\n");
    ok = false;
  }
  residual = (convergence == CONVERGENCE_L2) ? l2 : linf;
  return ok && !any_error();
}

// sol_pcgpressure3d.cu:600-650.  The dimension check joins its three tests
// (the reference's commas test only nz).
template <typename T>
bool Sol_PCGPressure3DDevice<T>::initialize_storage(int nx_val, int ny_val, int nz_val, float hx_val, float hy_val,
                                                    float hz_val, Grid3DDevice<T> *rhs,
                                                    Grid3DDevice<T> *coefficient) {
  if (!check_float(hx_val) || !check_float(hy_val) || !check_float(hz_val)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - garbage hx,hy,hz value %f %f %f\n", hx_val, hy_val,
           hz_val);
    return false;
  }
  if (!bc.check_type(BC_PERIODIC, BC_NEUMANN)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - unsupported boundary type\n");
    return false;
  }
  _h = min3(hx_val, hy_val, hz_val);
  _fx = (_h * _h) / (hx_val * hx_val);
  _fy = (_h * _h) / (hy_val * hy_val);
  _fz = (_h * _h) / (hz_val * hz_val);
  _nx = nx_val;
  _ny = ny_val;
  _nz = nz_val;
  if (!rhs || !coefficient) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - null rhs or coefficient\n");
    return false;
  }
  if (rhs->gx() < 1 || rhs->gy() < 1 || rhs->gz() < 1) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - rhs has invalid ghost cells (%d,%d,%d), must be "
           ">= 1\n",
           rhs->gx(), rhs->gy(), rhs->gz());
    return false;
  }
  if (rhs->nx() != nx_val || rhs->ny() != ny_val || rhs->nz() != nz_val) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - rhs dimension mismatch (%d,%d,%d) != (%d,%d,%d)\n",
           rhs->nx(), rhs->ny(), rhs->nz(), nx_val, ny_val, nz_val);
    return false;
  }
  _rhs = rhs;
  if (!coefficient->check_layout_match(*_rhs) || !coefficient->check_interior_dimension_match(*_rhs)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - coefficient layout must match rhs\n");
    return false;
  }
  _coefficient = coefficient;
  if (!_pressure.init_congruent(*_rhs) || !_d_r.init_congruent(_pressure) || !_d_y.init_congruent(_pressure) ||
      !_d_d.init_congruent(_pressure)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - could not initialize pressure\n");
    return false;
  }
  // the reference's work grids are fresh uninitialised allocations; zero
  // ones give the same interior results and defined ghosts
  _pressure.clear_zero();
  _d_r.clear_zero();
  _d_y.clear_zero();
  _d_d.clear_zero();
  if (!_ws && ocu_malloc(&_ws, ocu_reduce_workspace_bytes()) != OCU_OK) {
    _ws = nullptr;
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - could not allocate the reduction workspace\n");
    return false;
  }
  if (!_host) {
    void *h = nullptr, *hd = nullptr;
    if (ocu_host_malloc(&h, sizeof(double) * 8) != OCU_OK || ocu_host_device_pointer(&hd, h) != OCU_OK) {
      printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - could not allocate pinned memory\n");
      return false;
    }
    _host = (double *)h;
    _host_dev = (double *)hd;
  }
  _mg.nu1 = 2;
  _mg.nu2 = 2;
  _mg.bc = bc;
  if (!_mg.initialize_storage(_nx, _ny, _nz, hx_val, hy_val, hz_val, &_d_r)) {
    printf("[ERROR] Sol_PCGPressure3DDevice::initialize_storage - could not initialize multigrid solver\n");
    return false;
  }
  return true;
}

template class Sol_PCGPressure3DDevice<float>;
template class Sol_PCGPressure3DDevice<double>;

}  // namespace ocu
