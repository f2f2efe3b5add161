// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// One-dimensional grids, the drop-in for ocustorage/grid1d.h: nx cells with
// gx ghost cells on each side, padded to a multiple of 16 (4-byte types) or
// 8 (8-byte types) elements (grid1ddev.cu:306-330), element i at
// buffer + gx + i.  Device memory comes from libocu_hip.so.
#ifndef __OCU_STORAGE_GRID1D_H__
#define __OCU_STORAGE_GRID1D_H__

#include <cstddef>

#include "ocuutil/defines.h"

namespace ocu {

class Grid1DUntyped {
 protected:
  int _nx, _gx, _pnx;
  Grid1DUntyped() : _nx(0), _gx(0), _pnx(0) {}
  ~Grid1DUntyped() {}
  void compute_layout(int nx, int gx, size_t atom) {
    _nx = nx;
    _gx = gx;
    _pnx = nx + 2 * gx;
    const int shift = atom == 4 ? 4 : 3, mask = (1 << shift) - 1;
    if (_pnx & mask) _pnx = ((_pnx >> shift) + 1) << shift;
  }

 public:
  int nx() const { return _nx; }
  int pnx() const { return _pnx; }
  int gx() const { return _gx; }
};

template <typename T>
class Grid1DTypedBase : public Grid1DUntyped {
  Grid1DTypedBase(const Grid1DTypedBase &) = delete;
  Grid1DTypedBase &operator=(const Grid1DTypedBase &) = delete;

 protected:
  T *_buffer;
  T *_shifted_buffer;
  Grid1DTypedBase() : _buffer(nullptr), _shifted_buffer(nullptr) {}
  ~Grid1DTypedBase() {}

 public:
  // element reference; on a Grid1DDevice only its address is meaningful
  T &at(int i) { return _shifted_buffer[i]; }
  const T &at(int i) const { return _shifted_buffer[i]; }
  T *buffer() { return _buffer; }
  const T *buffer() const { return _buffer; }
};

template <typename T> class Grid1DDevice;

template <typename T>
class Grid1DHost : public Grid1DTypedBase<T> {
  bool _pinned;

 public:
  Grid1DHost() : _pinned(false) {}
  ~Grid1DHost();
  bool init(int nx_val, int gx_val, bool pinned = true);
  void clear_zero();
  bool copy_all_data(const Grid1DHost<T> &from);
  bool copy_all_data(const Grid1DDevice<T> &from);  // waits for the copy
};

template <typename T>
class Grid1DDevice : public Grid1DTypedBase<T> {
 public:
  Grid1DDevice() {}
  ~Grid1DDevice();
  bool init(int nx_val, int gx_val);
  bool clear_zero();
  bool copy_all_data(const Grid1DDevice<T> &from);
  bool copy_all_data(const Grid1DHost<T> &from);  // waits for the copy
};

typedef Grid1DHost<float> Grid1DHostF;
typedef Grid1DHost<double> Grid1DHostD;
typedef Grid1DHost<int> Grid1DHostI;
typedef Grid1DDevice<float> Grid1DDeviceF;
typedef Grid1DDevice<double> Grid1DDeviceD;
typedef Grid1DDevice<int> Grid1DDeviceI;

}  // namespace ocu

#endif
