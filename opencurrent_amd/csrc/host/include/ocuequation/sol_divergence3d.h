// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/sol_divergence3d.h: MAC divergence (sol_divergence3ddev.cu)
#ifndef __OCU_EQUATION_DIVERGENCE_3D_H__
#define __OCU_EQUATION_DIVERGENCE_3D_H__

#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"

namespace ocu {

template <typename T>
class Sol_Divergence3DDevice : public Solver {
  int _nx, _ny, _nz;
  double _hx, _hy, _hz;

 public:
  Sol_Divergence3DDevice();

  Grid3DDevice<T> *u, *v, *w;
  Grid3DDevice<T> *divergence;

  bool solve();
  bool initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz, Grid3DDevice<T> *u_val,
                          Grid3DDevice<T> *v_val, Grid3DDevice<T> *w_val, Grid3DDevice<T> *divergence_val);

  int nx() const { return _nx; }
  int ny() const { return _ny; }
  int nz() const { return _nz; }
  double hx() const { return _hx; }
  double hy() const { return _hy; }
  double hz() const { return _hz; }
};

}  // namespace ocu

#endif
