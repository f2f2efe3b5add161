// MAC-grid operator solvers over libocu_hip.so: divergence
// (sol_divergence3ddev.cu), gradient (sol_gradient3ddev.cu), self-advection
// (sol_selfadvection3ddev.cu), passive advection (sol_passiveadvection3ddev.cu)
// and the centred Laplacian (sol_laplaciancent3ddev.cu).
#include <cstdio>

#include "ocu_hip.h"
#include "ocuequation/sol_divergence3d.h"
#include "ocuequation/sol_gradient3d.h"
#include "ocuequation/sol_laplaciancent3d.h"
#include "ocuequation/sol_passiveadvection3d.h"
#include "ocuequation/sol_selfadvection3d.h"
#include "ocustorage/grid3dops.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/thread.h"

namespace ocu {

// ---------------------------------------------------------------- divergence

template <typename T>
Sol_Divergence3DDevice<T>::Sol_Divergence3DDevice() {
  _nx = _ny = _nz = 0;
  _hx = _hy = _hz = 0;
  u = v = w = 0;
  divergence = 0;
}

template <typename T>
bool Sol_Divergence3DDevice<T>::solve() {
  ocu_grid3d du = u->desc(), dv = v->desc(), dw = w->desc(), dd = divergence->desc();
  return ok_status(ocu_divergence(&du, &dv, &dw, &dd, _nx, _ny, _nz, _hx, _hy, _hz, ThreadManager::get_compute_stream()),
                   "Sol_Divergence3DDevice::solve");
}

// sol_divergence3ddev.cu:91-128
template <typename T>
bool Sol_Divergence3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                   Grid3DDevice<T> *u_val, Grid3DDevice<T> *v_val,
                                                   Grid3DDevice<T> *w_val, Grid3DDevice<T> *divergence_val) {
  u = u_val; v = v_val; w = w_val;
  divergence = divergence_val;
  if (!check_valid_mac_dimensions(*u_val, *v_val, *w_val, nx, ny, nz)) {
    printf("[ERROR] Sol_Divergence3DDevice::initialize_storage - u,v,w grid dimensions mismatch\n");
    return false;
  }
  if (divergence_val->nx() != nx || divergence_val->ny() != ny || divergence_val->nz() != nz ||
      !divergence_val->check_layout_match(*u_val)) {
    printf("[ERROR] Sol_Divergence3DDevice::initialize_storage - invalid dimensions for divergence\n");
    return false;
  }
  if (u_val->gx() < 1 || u_val->gy() < 1 || u_val->gz() < 1) {
    printf("[ERROR] Sol_Divergence3DDevice::initialize_storage - must have at least one ghost cell on all sides\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_Divergence3DDevice::initialize_storage - garbage hx,hy,hz value\n");
    return false;
  }
  _hx = hx; _hy = hy; _hz = hz;
  _nx = nx; _ny = ny; _nz = nz;
  return true;
}

// ------------------------------------------------------------------ gradient

template <typename T>
Sol_Gradient3DDevice<T>::Sol_Gradient3DDevice() {
  // the reference leaves _hx.. uninitialised (`_hx = _hy = _hz;`,
  // sol_gradient3ddev.cu:55); zero here
  _hx = _hy = _hz = 0;
  u = v = w = 0;
  phi = 0;
  coefficient = (T)-1;
}

template <typename T>
bool Sol_Gradient3DDevice<T>::solve() {
  ocu_grid3d du = u->desc(), dv = v->desc(), dw = w->desc(), dp = phi->desc();
  return ok_status(ocu_gradient_sub(&du, &dv, &dw, &dp, (double)coefficient, _hx, _hy, _hz,
                                    ThreadManager::get_compute_stream()),
                   "Sol_Gradient3DDevice::solve");
}

template <typename T>
bool Sol_Gradient3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                 Grid3DDevice<T> *u_val, Grid3DDevice<T> *v_val,
                                                 Grid3DDevice<T> *w_val, Grid3DDevice<T> *phi_val) {
  u = u_val; v = v_val; w = w_val;
  phi = phi_val;
  if (!check_valid_mac_dimensions(*u_val, *v_val, *w_val, nx, ny, nz)) {
    printf("[ERROR] Sol_Gradient3DDevice::initialize_storage - u,v,w grid dimensions mismatch\n");
    return false;
  }
  if (phi_val->nx() != nx || phi_val->ny() != ny || phi_val->nz() != nz || !phi_val->check_layout_match(*u_val)) {
    printf("[ERROR] Sol_Gradient3DDevice::initialize_storage - invalid dimensions for phi\n");
    return false;
  }
  if (u_val->gx() < 1 || u_val->gy() < 1 || u_val->gz() < 1) {
    printf("[ERROR] Sol_Gradient3DDevice::initialize_storage - must have at least one ghost cell on all sides\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_Gradient3DDevice::initialize_storage - garbage hx,hy,hz value\n");
    return false;
  }
  _hx = hx; _hy = hy; _hz = hz;
  return true;
}

// ------------------------------------------------------------ self-advection

template <typename T>
Sol_SelfAdvection3DDevice<T>::Sol_SelfAdvection3DDevice() {
  _nx = _ny = _nz = 0;
  _hx = _hy = _hz = 0;
  u = v = w = 0;
  deriv_udt = deriv_vdt = deriv_wdt = 0;
  interp_type = IT_FIRST_ORDER_UPWIND;
  periodic_axes = 0;
}

template <typename T>
bool Sol_SelfAdvection3DDevice<T>::solve() {
  if (interp_type != IT_FIRST_ORDER_UPWIND && interp_type != IT_SECOND_ORDER_CENTERED &&
      interp_type != IT_SECOND_ORDER_UPWIND) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::solve - invalid interpolation type %d\n", interp_type);
    return false;
  }
  ocu_grid3d du = u->desc(), dv = v->desc(), dw = w->desc();
  ocu_grid3d a = deriv_udt->desc(), b = deriv_vdt->desc(), c = deriv_wdt->desc();
  return ok_status(ocu_self_advect_ex(&du, &dv, &dw, &a, &b, &c, _nx, _ny, _nz, _hx, _hy, _hz, (int)interp_type,
                                      periodic_axes, ThreadManager::get_compute_stream()),
                   "Sol_SelfAdvection3DDevice::solve");
}

// sol_selfadvection3ddev.cu:593-668
template <typename T>
bool Sol_SelfAdvection3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                      Grid3DDevice<T> *u_val, Grid3DDevice<T> *v_val,
                                                      Grid3DDevice<T> *w_val, Grid3DDevice<T> *deriv_udt_val,
                                                      Grid3DDevice<T> *deriv_vdt_val, Grid3DDevice<T> *deriv_wdt_val) {
  if (u_val->nx() != nx + 1 || u_val->ny() != ny || u_val->nz() != nz) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - u dimensions mismatch\n");
    return false;
  }
  if (v_val->nx() != nx || v_val->ny() != ny + 1 || v_val->nz() != nz) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - v dimensions mismatch\n");
    return false;
  }
  if (w_val->nx() != nx || w_val->ny() != ny || w_val->nz() != nz + 1) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - w dimensions mismatch\n");
    return false;
  }
  if (!u_val->check_layout_match(*v_val) || !u_val->check_layout_match(*w_val)) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - u,v,w layout mismatch\n");
    return false;
  }
  if (u_val->gx() < 1 || u_val->gy() < 1 || u_val->gz() < 1 || v_val->gx() < 1 || v_val->gy() < 1 ||
      v_val->gz() < 1 || w_val->gx() < 1 || w_val->gy() < 1 || w_val->gz() < 1) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - u,v,w need ghost cells\n");
    return false;
  }
  u = u_val; v = v_val; w = w_val;
  deriv_udt = deriv_udt_val; deriv_vdt = deriv_vdt_val; deriv_wdt = deriv_wdt_val;
  if (!deriv_udt->check_layout_match(*u) || !deriv_vdt->check_layout_match(*u) || !deriv_wdt->check_layout_match(*u)) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - deriv layout mismatch\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_SelfAdvection3DDevice::initialize_storage - garbage hx,hy,hz value\n");
    return false;
  }
  _nx = nx; _ny = ny; _nz = nz;
  _hx = hx; _hy = hy; _hz = hz;
  return true;
}

// --------------------------------------------------------- passive advection

template <typename T>
Sol_PassiveAdvection3DDevice<T>::Sol_PassiveAdvection3DDevice() {
  _nx = _ny = _nz = 0;
  _hx = _hy = _hz = 0;
  u = v = w = 0;
  phi = 0;
  deriv_phidt = 0;
  interp_type = IT_FIRST_ORDER_UPWIND;
  periodic_axes = 0;
}

template <typename T>
bool Sol_PassiveAdvection3DDevice<T>::solve() {
  if (interp_type != IT_FIRST_ORDER_UPWIND && interp_type != IT_SECOND_ORDER_CENTERED &&
      interp_type != IT_SECOND_ORDER_UPWIND) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::solve() - invalid interpolation type %d\n", interp_type);
    return false;
  }
  ocu_grid3d du = u->desc(), dv = v->desc(), dw = w->desc(), dp = phi->desc(), dd = deriv_phidt->desc();
  return ok_status(ocu_passive_advect_ex(&du, &dv, &dw, &dp, &dd, _nx, _ny, _nz, _hx, _hy, _hz, (int)interp_type,
                                         periodic_axes, ThreadManager::get_compute_stream()),
                   "Sol_PassiveAdvection3DDevice::solve");
}

// sol_passiveadvection3ddev.cu:139-200
template <typename T>
bool Sol_PassiveAdvection3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                         Grid3DDevice<T> *u_val, Grid3DDevice<T> *v_val,
                                                         Grid3DDevice<T> *w_val, Grid3DDevice<T> *phi_val,
                                                         Grid3DDevice<T> *deriv_phidt_val) {
  if (!check_valid_mac_dimensions(*u_val, *v_val, *w_val, nx, ny, nz)) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::initialize_storage - u,v,w dimensions mismatch\n");
    return false;
  }
  if (phi_val->nx() != nx || phi_val->ny() != ny || phi_val->nz() != nz) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::initialize_storage - phi dimensions mismatch\n");
    return false;
  }
  if (!u_val->check_layout_match(*phi_val) || !u_val->check_layout_match(*deriv_phidt_val)) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::initialize_storage - layout mismatch\n");
    return false;
  }
  if (phi_val->gx() < 1 || phi_val->gy() < 1 || phi_val->gz() < 1) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::initialize_storage - phi must have ghost cells\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_PassiveAdvection3DDevice::initialize_storage - garbage hx,hy,hz value\n");
    return false;
  }
  u = u_val; v = v_val; w = w_val;
  phi = phi_val;
  deriv_phidt = deriv_phidt_val;
  _nx = nx; _ny = ny; _nz = nz;
  _hx = hx; _hy = hy; _hz = hz;
  return true;
}

// ----------------------------------------------------------------- Laplacian

template <typename T>
Sol_LaplacianCentered3DDevice<T>::Sol_LaplacianCentered3DDevice() {
  _hx = _hy = _hz = 0;
  _nx = _ny = _nz = 0;
  phi = 0;
  deriv_phidt = 0;
  coefficient = 1;
}

template <typename T>
bool Sol_LaplacianCentered3DDevice<T>::solve() {
  ocu_grid3d dp = phi->desc(), dd = deriv_phidt->desc();
  return ok_status(ocu_laplacian_acc(&dp, &dd, _nx, _ny, _nz, _hx, _hy, _hz, (double)coefficient,
                                     ThreadManager::get_compute_stream()),
                   "Sol_LaplacianCentered3DDevice::solve");
}

// sol_laplaciancent3ddev.cu:112-142 (the reference checks phi->nz() < 1 where
// the ghost width is meant, :121; the ghost width is checked here)
template <typename T>
bool Sol_LaplacianCentered3DDevice<T>::initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz,
                                                          Grid3DDevice<T> *phi_val, Grid3DDevice<T> *deriv_phidt_val) {
  if (phi_val->nx() != nx || phi_val->ny() != ny || phi_val->nz() != nz) {
    printf("[ERROR] Sol_LaplacianCentered3DDevice::initialize_storage - phi dimensions mismatch\n");
    return false;
  }
  if (phi_val->gx() < 1 || phi_val->gy() < 1 || phi_val->gz() < 1) {
    printf("[ERROR] Sol_LaplacianCentered3DDevice::initialize_storage - must have >= 1 ghost cell\n");
    return false;
  }
  if (!phi_val->check_layout_match(*deriv_phidt_val)) {
    printf("[ERROR] Sol_LaplacianCentered3DDevice::initialize_storage - phi/deriv_phidt layout mismatch\n");
    return false;
  }
  if (!check_float(hx) || !check_float(hy) || !check_float(hz)) {
    printf("[ERROR] Sol_LaplacianCentered3DDevice::initialize_storage - garbage hx,hy,hz value\n");
    return false;
  }
  phi = phi_val;
  deriv_phidt = deriv_phidt_val;
  _nx = nx; _ny = ny; _nz = nz;
  _hx = hx; _hy = hy; _hz = hz;
  return true;
}

template class Sol_Divergence3DDevice<float>;
template class Sol_Divergence3DDevice<double>;
template class Sol_Gradient3DDevice<float>;
template class Sol_Gradient3DDevice<double>;
template class Sol_SelfAdvection3DDevice<float>;
template class Sol_SelfAdvection3DDevice<double>;
template class Sol_PassiveAdvection3DDevice<float>;
template class Sol_PassiveAdvection3DDevice<double>;
template class Sol_LaplacianCentered3DDevice<float>;
template class Sol_LaplacianCentered3DDevice<double>;

}  // namespace ocu
