// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/sol_laplaciancent3d.h: dphidt += c * Laplacian(phi)
#ifndef __OCU_EQUATION_LAPLACIANCENT_3D_H__
#define __OCU_EQUATION_LAPLACIANCENT_3D_H__

#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"

namespace ocu {

template <typename T>
class Sol_LaplacianCentered3DDevice : public Solver {
  int _nx, _ny, _nz;
  double _hx, _hy, _hz;

 public:
  Sol_LaplacianCentered3DDevice();

  Grid3DDevice<T> *phi;
  Grid3DDevice<T> *deriv_phidt;
  T coefficient;

  bool solve();
  bool initialize_storage(int nx, int ny, int nz, double hx, double hy, double hz, Grid3DDevice<T> *phi_val,
                          Grid3DDevice<T> *deriv_phidt_val);
};

}  // namespace ocu

#endif
