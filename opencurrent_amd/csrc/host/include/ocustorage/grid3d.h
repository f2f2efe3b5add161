// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Grid3D storage with OpenCurrent's padded ghost-cell layout
// (ocustorage/grid3d.h:29-264, grid3ddev.cu:358-400): pnz rounded up to a
// multiple of 16, a (16 - gz)-element pre-padding, cell (i,j,k) at
// shift + i*pny*pnz + j*pnz + k.  Device memory comes from libocu_hip.so;
// every grid can hand out its ocu_grid3d descriptor for the C ABI.
#ifndef __OCU_STORAGE_GRID3D_H__
#define __OCU_STORAGE_GRID3D_H__

#include <cstddef>
#include <vector>

#include "ocu_types.h"
#include "ocuutil/defines.h"
#include "ocuutil/direction.h"
#include "ocuutil/thread.h"

namespace ocu {

class Grid3DUntyped {
 protected:
  int _nx, _ny, _nz;
  int _gx, _gy, _gz;
  int _pnx, _pny, _pnz;
  int _pnzpny;
  int _shift_amount;
  int _allocated_elements;

  ~Grid3DUntyped() {}
  Grid3DUntyped()
      : _nx(0), _ny(0), _nz(0), _gx(0), _gy(0), _gz(0), _pnx(0), _pny(0), _pnz(0), _pnzpny(0),
        _shift_amount(0), _allocated_elements(0) {}

  // fill the layout numbers exactly as Grid3DDevice<T>::init
  void compute_layout(int nx, int ny, int nz, int gx, int gy, int gz, int padx, int pady, int padz);

 public:
  int nx() const { return _nx; }
  int ny() const { return _ny; }
  int nz() const { return _nz; }
  int pnx() const { return _pnx; }
  int pny() const { return _pny; }
  int pnz() const { return _pnz; }
  int gx() const { return _gx; }
  int gy() const { return _gy; }
  int gz() const { return _gz; }
  int paddingx() const { return _pnx - (_nx + 2 * _gx); }
  int paddingy() const { return _pny - (_ny + 2 * _gy); }
  int paddingz() const { return _pnz - (_nz + 2 * _gz); }
  int num_allocated_elements() const { return _allocated_elements; }
  int xstride() const { return _pnzpny; }
  int ystride() const { return _pnz; }
  int zstride() const { return 1; }
  int stride(DirectionType dir) const {
    return (dir & DIR_XAXIS_FLAG) ? xstride() : (dir & DIR_YAXIS_FLAG) ? ystride() : zstride();
  }
  int shift_amount() const { return _shift_amount; }

  bool check_interior_dimension_match(const Grid3DUntyped &o) const {
    return nx() == o.nx() && ny() == o.ny() && nz() == o.nz();
  }
  bool check_layout_match(const Grid3DUntyped &o) const {
    return pnx() == o.pnx() && pny() == o.pny() && pnz() == o.pnz();
  }
};

class Grid3DDimension : public Grid3DUntyped {
 public:
  // grid3d.cpp:20-49: sizes only (no pre-padding, no allocation)
  void init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val);
  // grid3d.cpp:52-73.  The reference computes max_ny/max_nz/max_gy/max_gz
  // from max_nx/max_gx (a typo); this version takes the maximum of each
  // dimension separately, which is what the comment there describes.
  static void pad_for_congruence(std::vector<Grid3DDimension> &grids);
};

class Grid3DBase : public Grid3DUntyped {
  Grid3DBase &operator=(const Grid3DBase &) = delete;
  Grid3DBase(const Grid3DBase &) = delete;

 protected:
  void *_buffer;
  void *_shifted_buffer;
  size_t _atom_size;
  int _image_id;
  bool _owns;

  ~Grid3DBase() {}
  Grid3DBase(size_t atom) : _buffer(0), _shifted_buffer(0), _atom_size(atom), _image_id(0), _owns(false) {}

 public:
  size_t atom_size() const { return _atom_size; }
  void *ptr_untyped(int i, int j, int k) {
    return ((char *)_shifted_buffer) + _atom_size * ((size_t)i * _pnzpny + (size_t)j * _pnz + k);
  }
  const void *ptr_untyped(int i, int j, int k) const {
    return ((const char *)_shifted_buffer) + _atom_size * ((size_t)i * _pnzpny + (size_t)j * _pnz + k);
  }
};

template <typename T>
struct OcuDType;
template <> struct OcuDType<float> { enum { value = OCU_F32 }; };
template <> struct OcuDType<double> { enum { value = OCU_F64 }; };
template <> struct OcuDType<int> { enum { value = OCU_I32 }; };

template <typename T>
class Grid3DTypedBase : public Grid3DBase {
 protected:
  ~Grid3DTypedBase() {}
  Grid3DTypedBase() : Grid3DBase(sizeof(T)) {}

 public:
  // element reference; on a Grid3DDevice only its address is meaningful
  T &at(int i, int j, int k) { return *(((T *)_shifted_buffer) + (ptrdiff_t)i * _pnzpny + (ptrdiff_t)j * _pnz + k); }
  const T &at(int i, int j, int k) const {
    return *(((const T *)_shifted_buffer) + (ptrdiff_t)i * _pnzpny + (ptrdiff_t)j * _pnz + k);
  }
  const T *buffer() const { return (const T *)_buffer; }
  T *buffer() { return (T *)_buffer; }

  // descriptor for the C ABI
  ocu_grid3d desc() const {
    ocu_grid3d d;
    d.base = _buffer;
    d.nx = _nx; d.ny = _ny; d.nz = _nz;
    d.gx = _gx; d.gy = _gy; d.gz = _gz;
    d.pnx = _pnx; d.pny = _pny; d.pnz = _pnz;
    d.dtype = OcuDType<T>::value;
    d.shift = _shift_amount;
    d.allocated = _allocated_elements;
    return d;
  }
};

template <typename T> class Grid3DHost;
template <typename T> class Grid3DDevice;

template <typename T>
class Grid3DHost : public Grid3DTypedBase<T> {
  bool _pinned;

 public:
  Grid3DHost() : _pinned(false) {}
  ~Grid3DHost();

  bool init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val, bool pinned = true,
            int padx = 0, int pady = 0, int padz = 0);
  bool init_congruent(const Grid3DUntyped &other, bool pinned = true) {
    return init(other.nx(), other.ny(), other.nz(), other.gx(), other.gy(), other.gz(), pinned,
                other.paddingx(), other.paddingy(), other.paddingz());
  }

  void clear_zero();
  void clear(T t);
  template <typename S>
  bool copy_all_data(const Grid3DHost<S> &from);
  bool copy_all_data(const Grid3DDevice<T> &from);
  bool linear_combination(T alpha1, const Grid3DHost<T> &g1);
  bool linear_combination(T alpha1, const Grid3DHost<T> &g1, T alpha2, const Grid3DHost<T> &g2);

  bool reduce_maxabs(T &result) const;
  bool reduce_max(T &result) const;
  bool reduce_min(T &result) const;
  bool reduce_sum(T &result) const;
  bool reduce_sqrsum(T &result) const;
  bool reduce_checknan(T &result) const;
};

template <typename T>
class Grid3DDevice : public Grid3DTypedBase<T> {
 public:
  Grid3DDevice() {}
  ~Grid3DDevice();

  bool init(int nx_val, int ny_val, int nz_val, int gx_val, int gy_val, int gz_val, int padx = 0, int pady = 0,
            int padz = 0);
  bool init_congruent(const Grid3DUntyped &other) {
    return init(other.nx(), other.ny(), other.nz(), other.gx(), other.gy(), other.gz(), other.paddingx(),
                other.paddingy(), other.paddingz());
  }
  // Wrap device memory owned by someone else (for example a torch tensor);
  // the descriptor must describe a valid layout of type T.
  bool init_external(const ocu_grid3d &d);

  bool clear_zero();
  bool clear(T t);
  template <typename S>
  bool copy_all_data(const Grid3DDevice<S> &from);
  bool copy_all_data(const Grid3DHost<T> &from);
  bool linear_combination(T alpha1, const Grid3DDevice<T> &g1);
  bool linear_combination(T alpha1, const Grid3DDevice<T> &g1, T alpha2, const Grid3DDevice<T> &g2);

  bool reduce_maxabs(T &result) const;
  bool reduce_sum(T &result) const;
  bool reduce_sqrsum(T &result) const;
  bool reduce_max(T &result) const;
  bool reduce_min(T &result) const;
  bool reduce_checknan(T &result) const;

  // reduction in double precision (op as ocu_grid_reduce), blocking
  bool reduce_double(int op, double &result) const;
  // max |.| of three grids (a MAC velocity) in one blocking call: the three
  // reductions are launched back to back and the host waits once
  static bool reduce_maxabs3(const Grid3DDevice<T> &a, const Grid3DDevice<T> &b, const Grid3DDevice<T> &c, T &ra,
                             T &rb, T &rc);

  // exchange the storage of two grids of one layout (no reference
  // counterpart): a kernel that writes a new version of a grid out of
  // place hands it over without a copy
  bool swap_storage(Grid3DDevice<T> &other);
};

typedef Grid3DHost<float> Grid3DHostF;
typedef Grid3DHost<double> Grid3DHostD;
typedef Grid3DHost<int> Grid3DHostI;
typedef Grid3DDevice<float> Grid3DDeviceF;
typedef Grid3DDevice<double> Grid3DDeviceD;
typedef Grid3DDevice<int> Grid3DDeviceI;

}  // namespace ocu

#endif
