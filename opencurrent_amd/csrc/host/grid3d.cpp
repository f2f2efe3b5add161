// Grid3DHost<T> / Grid3DDevice<T> (ocustorage/grid3dhost.cu, grid3ddev.cu,
// grid3d.cpp) over libocu_hip.so.  Device operations are enqueued on the
// compute stream; operations that hand data back to the host (reductions,
// device->host copies) wait for it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>

#include "ocu_hip.h"
#include "ocustorage/grid3d.h"
#include "ocustorage/grid3dops.h"
#include "ocuutil/kernel_wrapper.h"

namespace ocu {

void Grid3DUntyped::compute_layout(int nx, int ny, int nz, int gx, int gy, int gz, int padx, int pady, int padz) {
  _nx = nx; _ny = ny; _nz = nz;
  _gx = gx; _gy = gy; _gz = gz;
  _pnx = nx + 2 * gx + padx;
  _pny = ny + 2 * gy + pady;
  _pnz = nz + 2 * gz + padz;
  if (_pnz & 15) _pnz = ((_pnz >> 4) + 1) << 4;
  int pre_padding = 16 - gz;
  _pnzpny = _pnz * _pny;
  _allocated_elements = _pnzpny * _pnx + pre_padding;
  _shift_amount = gx * _pnzpny + gy * _pnz + gz + pre_padding;
}

void Grid3DDimension::init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val) {
  _nx = nx_val; _ny = ny_val; _nz = nz_val;
  _gx = gx_val; _gy = gy_val; _gz = gz_val;
  _pnx = _nx + 2 * gx_val;
  _pny = _ny + 2 * gy_val;
  _pnz = _nz + 2 * gz_val;
  if (_pnz & 15) _pnz = ((_pnz >> 4) + 1) << 4;
  _pnzpny = _pnz * _pny;
  _allocated_elements = _pnzpny * _pnx;
  _shift_amount = _gx * _pnzpny + _gy * _pnz + _gz;
}

void Grid3DDimension::pad_for_congruence(std::vector<Grid3DDimension> &grids) {
  int max_nx = 0, max_ny = 0, max_nz = 0, max_gx = 0, max_gy = 0, max_gz = 0;
  for (auto &g : grids) {
    max_nx = std::max(max_nx, g.nx()); max_ny = std::max(max_ny, g.ny()); max_nz = std::max(max_nz, g.nz());
    max_gx = std::max(max_gx, g.gx()); max_gy = std::max(max_gy, g.gy()); max_gz = std::max(max_gz, g.gz());
  }
  for (auto &g : grids) {
    g._pnx = max_nx + 2 * max_gx;
    g._pny = max_ny + 2 * max_gy;
    g._pnz = max_nz + 2 * max_gz;
  }
}

bool check_valid_mac_dimensions(const Grid3DUntyped &u, const Grid3DUntyped &v, const Grid3DUntyped &w, int nx,
                                int ny, int nz) {
  if (u.nx() != nx + 1 || u.ny() != ny || u.nz() != nz) return false;
  if (v.nx() != nx || v.ny() != ny + 1 || v.nz() != nz) return false;
  if (w.nx() != nx || w.ny() != ny || w.nz() != nz + 1) return false;
  return u.check_layout_match(v) && u.check_layout_match(w);
}

// ---------------------------------------------------------------- host grid

template <typename T>
Grid3DHost<T>::~Grid3DHost() {
  if (this->_owns && this->_buffer) {
    if (_pinned) ocu_host_free(this->_buffer);
    else free(this->_buffer);
  }
}

template <typename T>
bool Grid3DHost<T>::init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val, bool pinned,
                         int padx, int pady, int padz) {
  if (this->_owns && this->_buffer) {
    if (_pinned) ocu_host_free(this->_buffer);
    else free(this->_buffer);
  }
  this->compute_layout(nx_val, ny_val, nz_val, gx_val, gy_val, gz_val, padx, pady, padz);
  size_t bytes = sizeof(T) * (size_t)this->_allocated_elements;
  void *p = nullptr;
  if (pinned) {
    if (ocu_host_malloc(&p, bytes) != OCU_OK) p = nullptr;
  } else {
    p = malloc(bytes);
  }
  if (!p) {
    printf("[ERROR] Grid3DHost::init - memory allocation failed\n");
    this->_buffer = this->_shifted_buffer = nullptr;
    this->_owns = false;
    return false;
  }
  memset(p, 0, bytes);
  _pinned = pinned;
  this->_owns = true;
  this->_buffer = p;
  this->_shifted_buffer = (T *)p + this->_shift_amount;
  return true;
}

template <typename T>
void Grid3DHost<T>::clear_zero() {
  memset(this->_buffer, 0, sizeof(T) * (size_t)this->_allocated_elements);
}

template <typename T>
void Grid3DHost<T>::clear(T t) {
  std::fill(this->buffer(), this->buffer() + this->_allocated_elements, t);
}

template <typename T>
template <typename S>
bool Grid3DHost<T>::copy_all_data(const Grid3DHost<S> &from) {
  if (!this->check_layout_match(from)) {
    printf("[ERROR] Grid3DHost::copy_all_data - mismatch\n");
    return false;
  }
  for (int n = 0; n < this->_allocated_elements; n++) this->buffer()[n] = (T)from.buffer()[n];
  return true;
}

template <typename T>
bool Grid3DHost<T>::copy_all_data(const Grid3DDevice<T> &from) {
  if (!this->check_layout_match(from)) {
    printf("[ERROR] Grid3DHost::copy_all_data - mismatch: (%d, %d, %d) != (%d, %d, %d)\n", this->pnx(), this->pny(),
           this->pnz(), from.pnx(), from.pny(), from.pnz());
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  if (!ocu_status_ok(ocu_memcpy_async(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_allocated_elements,
                                      1, s),
                     "Grid3DHost::copy_all_data"))
    return false;
  return ocu_status_ok(ocu_stream_synchronize(s), "Grid3DHost::copy_all_data");
}

template <typename T>
bool Grid3DHost<T>::linear_combination(T alpha1, const Grid3DHost<T> &g1) {
  if (!this->check_layout_match(g1)) return false;
  for (int n = 0; n < this->_allocated_elements; n++) this->buffer()[n] = g1.buffer()[n] * alpha1;
  return true;
}

template <typename T>
bool Grid3DHost<T>::linear_combination(T alpha1, const Grid3DHost<T> &g1, T alpha2, const Grid3DHost<T> &g2) {
  if (!this->check_layout_match(g1) || !this->check_layout_match(g2)) return false;
  for (int i = 0; i < this->nx(); i++)
    for (int j = 0; j < this->ny(); j++)
      for (int k = 0; k < this->nz(); k++)
        this->at(i, j, k) = g1.at(i, j, k) * alpha1 + g2.at(i, j, k) * alpha2;
  return true;
}

template <typename T>
static double host_reduce(const Grid3DHost<T> &g, int op) {
  double acc = 0;
  bool first = true;
  for (int i = 0; i < g.nx(); i++)
    for (int j = 0; j < g.ny(); j++)
      for (int k = 0; k < g.nz(); k++) {
        double x = (double)g.at(i, j, k);
        switch (op) {
          case 0: acc = first ? fabs(x) : std::max(acc, fabs(x)); break;
          case 1: acc = first ? x : std::max(acc, x); break;
          case 2: acc = first ? x : std::min(acc, x); break;
          case 3: acc += x; break;
          case 4: acc += x * x; break;
          case 5: if (!std::isfinite(x)) acc = x; break;
        }
        first = false;
      }
  return acc;
}

template <typename T> bool Grid3DHost<T>::reduce_maxabs(T &r) const { r = (T)host_reduce(*this, 0); return true; }
template <typename T> bool Grid3DHost<T>::reduce_max(T &r) const { r = (T)host_reduce(*this, 1); return true; }
template <typename T> bool Grid3DHost<T>::reduce_min(T &r) const { r = (T)host_reduce(*this, 2); return true; }
template <typename T> bool Grid3DHost<T>::reduce_sum(T &r) const { r = (T)host_reduce(*this, 3); return true; }
template <typename T> bool Grid3DHost<T>::reduce_sqrsum(T &r) const { r = (T)host_reduce(*this, 4); return true; }
template <typename T> bool Grid3DHost<T>::reduce_checknan(T &r) const { r = (T)host_reduce(*this, 5); return true; }

// -------------------------------------------------------------- device grid

template <typename T>
Grid3DDevice<T>::~Grid3DDevice() {
  if (this->_owns) ocu_free(this->_buffer);
}

template <typename T>
bool Grid3DDevice<T>::init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val, int padx,
                           int pady, int padz) {
  if (this->_owns && this->_buffer) ocu_free(this->_buffer);
  this->_buffer = this->_shifted_buffer = nullptr;
  this->_owns = false;
  this->compute_layout(nx_val, ny_val, nz_val, gx_val, gy_val, gz_val, padx, pady, padz);
  void *p = nullptr;
  if (ocu_malloc(&p, sizeof(T) * (size_t)this->_allocated_elements) != OCU_OK || !p) {
    printf("[ERROR] Grid3DDevice::init - hipMalloc failed: %s\n", ocu_last_error());
    return false;
  }
  this->_buffer = p;
  this->_owns = true;
  this->_shifted_buffer = (T *)p + this->_shift_amount;
  return true;
}

template <typename T>
bool Grid3DDevice<T>::init_external(const ocu_grid3d &d) {
  if (d.dtype != OcuDType<T>::value || !d.base) {
    printf("[ERROR] Grid3DDevice::init_external - descriptor type/storage mismatch\n");
    return false;
  }
  if (this->_owns && this->_buffer) ocu_free(this->_buffer);
  // nothing is held from here on, so a failed check below leaves an empty grid
  this->_buffer = nullptr;
  this->_shifted_buffer = nullptr;
  this->_owns = false;
  this->compute_layout(d.nx, d.ny, d.nz, d.gx, d.gy, d.gz, d.pnx - (d.nx + 2 * d.gx), d.pny - (d.ny + 2 * d.gy),
                       d.pnz - (d.nz + 2 * d.gz));
  if (this->_shift_amount != d.shift || this->_allocated_elements > d.allocated) {
    printf("[ERROR] Grid3DDevice::init_external - descriptor does not follow the Grid3D layout\n");
    return false;
  }
  this->_buffer = d.base;
  this->_owns = false;
  this->_shifted_buffer = (T *)d.base + this->_shift_amount;
  return true;
}

template <typename T>
bool Grid3DDevice<T>::swap_storage(Grid3DDevice<T> &other) {
  if (!this->check_layout_match(other) || this->_shift_amount != other._shift_amount ||
      this->_allocated_elements != other._allocated_elements || !this->check_interior_dimension_match(other)) {
    printf("[ERROR] Grid3DDevice::swap_storage - layout mismatch\n");
    return false;
  }
  std::swap(this->_buffer, other._buffer);
  std::swap(this->_shifted_buffer, other._shifted_buffer);
  std::swap(this->_owns, other._owns);
  return true;
}

template <typename T>
bool Grid3DDevice<T>::clear_zero() {
  ocu_grid3d d = this->desc();
  return ocu_status_ok(ocu_grid_clear(&d, 0.0, ThreadManager::get_compute_stream()), "Grid3DDevice::clear_zero");
}

template <typename T>
bool Grid3DDevice<T>::clear(T t) {
  ocu_grid3d d = this->desc();
  return ocu_status_ok(ocu_grid_clear(&d, (double)t, ThreadManager::get_compute_stream()), "Grid3DDevice::clear");
}

template <typename T>
template <typename S>
bool Grid3DDevice<T>::copy_all_data(const Grid3DDevice<S> &from) {
  if (!this->check_layout_match(from)) {
    printf("[ERROR] Grid3DDevice::copy_all_data - mismatch: (%d, %d, %d) != (%d, %d, %d)\n", this->pnx(),
           this->pny(), this->pnz(), from.pnx(), from.pny(), from.pnz());
    return false;
  }
  ocu_grid3d d = this->desc(), s = from.desc();
  return ocu_status_ok(ocu_grid_copy(&d, &s, ThreadManager::get_compute_stream()), "Grid3DDevice::copy_all_data");
}

template <typename T>
bool Grid3DDevice<T>::copy_all_data(const Grid3DHost<T> &from) {
  if (!this->check_layout_match(from)) {
    printf("[ERROR] Grid3DDevice::copy_all_data - mismatch: (%d, %d, %d) != (%d, %d, %d)\n", this->pnx(),
           this->pny(), this->pnz(), from.pnx(), from.pny(), from.pnz());
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  if (!ocu_status_ok(ocu_memcpy_async(this->_buffer, from.buffer(), sizeof(T) * (size_t)this->_allocated_elements,
                                      0, s),
                     "Grid3DDevice::copy_all_data"))
    return false;
  // the host buffer may be reused as soon as we return, like cudaMemcpy
  return ocu_status_ok(ocu_stream_synchronize(s), "Grid3DDevice::copy_all_data");
}

template <typename T>
bool Grid3DDevice<T>::linear_combination(T alpha1, const Grid3DDevice<T> &g1) {
  if (!this->check_layout_match(g1)) {
    printf("[ERROR] Grid3DDevice::linear_combination - mismatch\n");
    return false;
  }
  ocu_grid3d r = this->desc(), a = g1.desc();
  return ocu_status_ok(ocu_grid_lincomb(&r, (double)alpha1, &a, 0.0, nullptr, ThreadManager::get_compute_stream()),
                       "Grid3DDevice::linear_combination");
}

template <typename T>
bool Grid3DDevice<T>::linear_combination(T alpha1, const Grid3DDevice<T> &g1, T alpha2, const Grid3DDevice<T> &g2) {
  if (!this->check_layout_match(g1) || !this->check_layout_match(g2)) {
    printf("[ERROR] Grid3DDevice::linear_combination - mismatch\n");
    return false;
  }
  ocu_grid3d r = this->desc(), a = g1.desc(), b = g2.desc();
  return ocu_status_ok(
      ocu_grid_lincomb(&r, (double)alpha1, &a, (double)alpha2, &b, ThreadManager::get_compute_stream()),
      "Grid3DDevice::linear_combination");
}

namespace {
// per-thread scratch for blocking reductions: device workspace + result,
// pinned host result
// (kept until the process ends: freeing at thread exit could run after the
// HIP runtime has been torn down)
struct ReduceScratch {
  void *ws = nullptr;
  double *host = nullptr;
  double *host_dev = nullptr;  // host as the device sees it
  bool ready() {
    if (ws && host_dev) return true;
    if (!ws && ocu_malloc(&ws, 3 * ocu_reduce_workspace_bytes() + 64) != OCU_OK) return false;
    void *h = nullptr, *hd = nullptr;
    if (ocu_host_malloc(&h, 64) != OCU_OK || ocu_host_device_pointer(&hd, h) != OCU_OK) return false;
    host = (double *)h;
    host_dev = (double *)hd;
    return true;
  }
};
thread_local ReduceScratch g_red;
}  // namespace

template <typename T>
bool Grid3DDevice<T>::reduce_double(int op, double &result) const {
  if (!g_red.ready()) {
    printf("[ERROR] Grid3DDevice::reduce - scratch allocation failed\n");
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d d = this->desc();
  // the result goes straight into pinned host memory, with a completion tag
  // the host polls for
  const unsigned long long tag = next_completion_tag();
  if (!ocu_status_ok(ocu_grid_reduce_tagged(&d, op, g_red.ws, g_red.host_dev, tag, s), "Grid3DDevice::reduce"))
    return false;
  if (!wait_completion_tag(g_red.host + 1, tag, s)) {
    printf("[ERROR] Grid3DDevice::reduce - the result did not arrive\n");
    return false;
  }
  result = *g_red.host;
  return true;
}

// the three maxima from one pass over the grids (one partials launch and
// one fold), polled from the pinned slots 0..2 once the tag is in slot 3
template <typename T>
bool Grid3DDevice<T>::reduce_maxabs3(const Grid3DDevice<T> &a, const Grid3DDevice<T> &b, const Grid3DDevice<T> &c,
                                     T &ra, T &rb, T &rc) {
  if (!g_red.ready()) {
    printf("[ERROR] Grid3DDevice::reduce - scratch allocation failed\n");
    return false;
  }
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d da = a.desc(), db = b.desc(), dc = c.desc();
  const unsigned long long tag = next_completion_tag();
  if (!ocu_status_ok(ocu_grid_maxabs3_tagged(&da, &db, &dc, g_red.ws, g_red.host_dev, tag, s),
                     "Grid3DDevice::reduce"))
    return false;
  if (!wait_completion_tag(g_red.host + 3, tag, s)) {
    printf("[ERROR] Grid3DDevice::reduce - the results did not arrive\n");
    return false;
  }
  ra = (T)g_red.host[0];
  rb = (T)g_red.host[1];
  rc = (T)g_red.host[2];
  return true;
}

template <typename T> bool Grid3DDevice<T>::reduce_maxabs(T &r) const { double x; bool ok = reduce_double(0, x); r = (T)x; return ok; }
template <typename T> bool Grid3DDevice<T>::reduce_max(T &r) const { double x; bool ok = reduce_double(1, x); r = (T)x; return ok; }
template <typename T> bool Grid3DDevice<T>::reduce_min(T &r) const { double x; bool ok = reduce_double(2, x); r = (T)x; return ok; }
template <typename T> bool Grid3DDevice<T>::reduce_sum(T &r) const { double x; bool ok = reduce_double(3, x); r = (T)x; return ok; }
template <typename T> bool Grid3DDevice<T>::reduce_sqrsum(T &r) const { double x; bool ok = reduce_double(4, x); r = (T)x; return ok; }
template <typename T> bool Grid3DDevice<T>::reduce_checknan(T &r) const { double x; bool ok = reduce_double(5, x); r = (T)x; return ok; }

template class Grid3DHost<float>;
template class Grid3DHost<double>;
template class Grid3DHost<int>;
template class Grid3DDevice<float>;
template class Grid3DDevice<double>;

template bool Grid3DHost<float>::copy_all_data(const Grid3DHost<double> &);
template bool Grid3DHost<double>::copy_all_data(const Grid3DHost<float> &);
template bool Grid3DHost<float>::copy_all_data(const Grid3DHost<float> &);
template bool Grid3DHost<double>::copy_all_data(const Grid3DHost<double> &);
template bool Grid3DDevice<float>::copy_all_data(const Grid3DDevice<double> &);
template bool Grid3DDevice<double>::copy_all_data(const Grid3DDevice<float> &);
template bool Grid3DDevice<float>::copy_all_data(const Grid3DDevice<float> &);
template bool Grid3DDevice<double>::copy_all_data(const Grid3DDevice<double> &);

}  // namespace ocu
