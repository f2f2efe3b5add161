// Grid1DHost / Grid1DDevice (ocustorage/grid1dhost.cpp, grid1ddev.cu) and
// point sampling (ocustorage/grid3dsample.cu) over libocu_hip.so.
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "ocu_hip.h"
#include "ocustorage/grid1d.h"
#include "ocustorage/grid3dboundary.h"
#include "ocustorage/grid3dsample.h"
#include "ocuutil/kernel_wrapper.h"
#include "ocuutil/thread.h"

namespace ocu {

template <typename T>
Grid1DHost<T>::~Grid1DHost() {
  if (!this->_buffer) return;
  if (_pinned) ocu_host_free(this->_buffer);
  else delete[] this->_buffer;
}

template <typename T>
bool Grid1DHost<T>::init(int nx_val, int gx_val, bool pinned) {
  if (nx_val < 0 || gx_val < 0) {
    printf("[ERROR] Grid1DHost::init - invalid size %d ghosts %d\n", nx_val, gx_val);
    return false;
  }
  if (this->_buffer) {
    if (_pinned) ocu_host_free(this->_buffer);
    else delete[] this->_buffer;
    this->_buffer = this->_shifted_buffer = nullptr;
  }
  this->compute_layout(nx_val, gx_val, sizeof(T));
  _pinned = pinned;
  if (pinned) {
    void *p = nullptr;
    if (ocu_host_malloc(&p, sizeof(T) * (size_t)this->_pnx) != OCU_OK) {
      printf("[ERROR] Grid1DHost::init - pinned allocation failed\n");
      return false;
    }
    this->_buffer = (T *)p;
  } else {
    this->_buffer = new T[this->_pnx];
  }
  std::memset(this->_buffer, 0, sizeof(T) * (size_t)this->_pnx);
  this->_shifted_buffer = this->_buffer + this->_gx;
  return true;
}

template <typename T>
void Grid1DHost<T>::clear_zero() {
  if (this->_buffer) std::memset(this->_buffer, 0, sizeof(T) * (size_t)this->_pnx);
}

template <typename T>
bool Grid1DHost<T>::copy_all_data(const Grid1DHost<T> &from) {
  if (from.pnx() != this->pnx()) {
    printf("[ERROR] Grid1DHost::copy_all_data - mismatch: %d != %d\n", from.pnx(), this->pnx());
    return false;
  }
  std::memcpy(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_pnx);
  return true;
}

template <typename T>
bool Grid1DHost<T>::copy_all_data(const Grid1DDevice<T> &from) {
  if (from.pnx() != this->pnx()) {
    printf("[ERROR] Grid1DHost::copy_all_data - mismatch: %d != %d\n", from.pnx(), this->pnx());
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  return ocu_status_ok(ocu_memcpy_async(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_pnx, 1, s),
                       "Grid1DHost::copy_all_data") &&
         ocu_status_ok(ocu_stream_synchronize(s), "Grid1DHost::copy_all_data");
}

template <typename T>
Grid1DDevice<T>::~Grid1DDevice() {
  if (this->_buffer) ocu_free(this->_buffer);
}

template <typename T>
bool Grid1DDevice<T>::init(int nx_val, int gx_val) {
  if (nx_val < 0 || gx_val < 0) {
    printf("[ERROR] Grid1DDevice::init - invalid size %d ghosts %d\n", nx_val, gx_val);
    return false;
  }
  if (this->_buffer) ocu_free(this->_buffer);
  this->_buffer = this->_shifted_buffer = nullptr;
  this->compute_layout(nx_val, gx_val, sizeof(T));
  void *p = nullptr;
  if (ocu_malloc(&p, sizeof(T) * (size_t)(this->_pnx > 0 ? this->_pnx : 1)) != OCU_OK || !p) {
    printf("[ERROR] Grid1DDevice::init - allocation failed\n");
    return false;
  }
  this->_buffer = (T *)p;
  this->_shifted_buffer = this->_buffer + this->_gx;
  return true;
}

template <typename T>
bool Grid1DDevice<T>::clear_zero() {
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d d = ocu_grid3d();  // the allocation as a flat 1 x 1 x pnx grid
  d.base = this->_buffer;
  d.nx = d.ny = 1;
  d.nz = this->_pnx;
  d.pnx = d.pny = 1;
  d.pnz = this->_pnx;
  d.allocated = this->_pnx;
  d.dtype = sizeof(T) == 8 ? OCU_F64 : std::is_same<T, int>::value ? OCU_I32 : OCU_F32;
  return ocu_status_ok(ocu_grid_clear(&d, 0.0, s), "Grid1DDevice::clear_zero");
}

template <typename T>
bool Grid1DDevice<T>::copy_all_data(const Grid1DDevice<T> &from) {
  if (from.pnx() != this->pnx()) {
    printf("[ERROR] Grid1DDevice::copy_all_data - mismatch: %d != %d\n", from.pnx(), this->pnx());
    return false;
  }
  return ocu_status_ok(ocu_memcpy_async(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_pnx, 2,
                                        ThreadManager::get_compute_stream()),
                       "Grid1DDevice::copy_all_data");
}

template <typename T>
bool Grid1DDevice<T>::copy_all_data(const Grid1DHost<T> &from) {
  if (from.pnx() != this->pnx()) {
    printf("[ERROR] Grid1DDevice::copy_all_data - mismatch: %d != %d\n", from.pnx(), this->pnx());
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  return ocu_status_ok(ocu_memcpy_async(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_pnx, 0, s),
                       "Grid1DDevice::copy_all_data") &&
         ocu_status_ok(ocu_stream_synchronize(s), "Grid1DDevice::copy_all_data");
}

template class Grid1DHost<float>;
template class Grid1DHost<double>;
template class Grid1DHost<int>;
template class Grid1DDevice<float>;
template class Grid1DDevice<double>;
template class Grid1DDevice<int>;

// grid3dsample.cu:137-171.  All three position arrays must match the output
// size (the reference compares position_y twice and never position_z).
template <typename T>
bool sample_points_3d(Grid1DDevice<T> &phi_sampled, const Grid1DDevice<float> &position_x,
                      const Grid1DDevice<float> &position_y, const Grid1DDevice<float> &position_z,
                      const Grid3DDevice<T> &phi, const BoundaryConditionSet &bc, float period_nx, float period_ny,
                      float period_nz, float hx, float hy, float hz, float shiftx, float shifty, float shiftz) {
  const int n = phi_sampled.nx();
  if (position_x.nx() != n || position_y.nx() != n || position_z.nx() != n) {
    printf("[ERROR] sample_points_3d - position arrays do not have matching sizes\n");
    return false;
  }
  ocu_grid3d d = phi.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_sample_points(&phi_sampled.at(0), &position_x.at(0), &position_y.at(0), &position_z.at(0),
                                         n, &d, &b, period_nx, period_ny, period_nz, hx, hy, hz, shiftx, shifty,
                                         shiftz, ThreadManager::get_compute_stream()),
                       "sample_points_3d");
}

// grid3dsample.cu:174-263
template <typename T>
bool sample_points_mac_grid_3d(Grid1DDevice<T> &vx_sampled, Grid1DDevice<T> &vy_sampled, Grid1DDevice<T> &vz_sampled,
                               const Grid1DDevice<float> &position_x, const Grid1DDevice<float> &position_y,
                               const Grid1DDevice<float> &position_z, const Grid3DDevice<T> &u,
                               const Grid3DDevice<T> &v, const Grid3DDevice<T> &w, const BoundaryConditionSet &bc,
                               float hx, float hy, float hz) {
  const int n = vx_sampled.nx();
  if (vy_sampled.nx() != n || vz_sampled.nx() != n || position_x.nx() != n || position_y.nx() != n ||
      position_z.nx() != n) {
    printf("[ERROR] sample_points_mac_grid_3d - array sizes do not match\n");
    return false;
  }
  ocu_grid3d du = u.desc(), dv = v.desc(), dw = w.desc();
  ocu_bc_set b = to_c_bc(bc);
  return ocu_status_ok(ocu_sample_points_mac(&vx_sampled.at(0), &vy_sampled.at(0), &vz_sampled.at(0),
                                             &position_x.at(0), &position_y.at(0), &position_z.at(0), n, &du, &dv,
                                             &dw, &b, hx, hy, hz, ThreadManager::get_compute_stream()),
                       "sample_points_mac_grid_3d");
}

template bool sample_points_3d(Grid1DDevice<float> &, const Grid1DDevice<float> &, const Grid1DDevice<float> &,
                               const Grid1DDevice<float> &, const Grid3DDevice<float> &, const BoundaryConditionSet &,
                               float, float, float, float, float, float, float, float, float);
template bool sample_points_3d(Grid1DDevice<double> &, const Grid1DDevice<float> &, const Grid1DDevice<float> &,
                               const Grid1DDevice<float> &, const Grid3DDevice<double> &,
                               const BoundaryConditionSet &, float, float, float, float, float, float, float, float,
                               float);
template bool sample_points_mac_grid_3d(Grid1DDevice<float> &, Grid1DDevice<float> &, Grid1DDevice<float> &,
                                        const Grid1DDevice<float> &, const Grid1DDevice<float> &,
                                        const Grid1DDevice<float> &, const Grid3DDevice<float> &,
                                        const Grid3DDevice<float> &, const Grid3DDevice<float> &,
                                        const BoundaryConditionSet &, float, float, float);
template bool sample_points_mac_grid_3d(Grid1DDevice<double> &, Grid1DDevice<double> &, Grid1DDevice<double> &,
                                        const Grid1DDevice<float> &, const Grid1DDevice<float> &,
                                        const Grid1DDevice<float> &, const Grid3DDevice<double> &,
                                        const Grid3DDevice<double> &, const Grid3DDevice<double> &,
                                        const BoundaryConditionSet &, float, float, float);

}  // namespace ocu
