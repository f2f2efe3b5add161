// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/sol_gradient3d.h: u += c * grad(phi) (sol_gradient3ddev.cu)
#ifndef __OCU_EQUATION_GRADIENT_3D_H__
#define __OCU_EQUATION_GRADIENT_3D_H__

#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"

namespace ocu {

template <typename T>
class Sol_Gradient3DDevice : public Solver {
  double _hx, _hy, _hz;

 public:
  Sol_Gradient3DDevice();

  Grid3DDevice<T> *u, *v, *w;
  Grid3DDevice<T> *phi;
  T coefficient;

  bool solve();
  bool initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz, Grid3DDevice<T> *u_val,
                          Grid3DDevice<T> *v_val, Grid3DDevice<T> *w_val, Grid3DDevice<T> *phi_val);
};

}  // namespace ocu

#endif
