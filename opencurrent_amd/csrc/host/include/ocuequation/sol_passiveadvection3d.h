// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/sol_passiveadvection3d.h: flux-form advection of a cell
// scalar by a MAC velocity (sol_passiveadvection3ddev.cu)
#ifndef __OCU_EQUATION_PASSIVEADVECTION_3D_H__
#define __OCU_EQUATION_PASSIVEADVECTION_3D_H__

#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"
#include "ocuutil/interpolation.h"

namespace ocu {

template <typename T>
class Sol_PassiveAdvection3DDevice : public Solver {
  int _nx, _ny, _nz;
  double _hx, _hy, _hz;

 public:
  Sol_PassiveAdvection3DDevice();

  Grid3DDevice<T> *u, *v, *w;
  Grid3DDevice<T> *phi;
  Grid3DDevice<T> *deriv_phidt;
  InterpolationType interp_type;
  // no reference counterpart: axes (x=1, y=2, z=4) across which the
  // IT_SECOND_ORDER_UPWIND stencil wraps; the other schemes ignore it
  int periodic_axes;

  bool solve();
  bool initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz, Grid3DDevice<T> *u_val,
                          Grid3DDevice<T> *v_val, Grid3DDevice<T> *w_val, Grid3DDevice<T> *phi_val,
                          Grid3DDevice<T> *deriv_phidt_val);

  int nx() const { return _nx; }
  int ny() const { return _ny; }
  int nz() const { return _nz; }
  double hx() const { return _hx; }
  double hy() const { return _hy; }
  double hz() const { return _hz; }
};

}  // namespace ocu

#endif
