// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/sol_project3d.h: Helmholtz projection of a MAC velocity
// (sol_project3d.cpp, sol_project3ddev.cu).  Sol_ProjectDivergence3DDevice<T>
// uses Sol_MultigridPressure3DDevice<T>; Sol_ProjectDivergenceMixed3DDeviceD
// projects an FP64 velocity with the mixed-precision pressure solver.
#ifndef __OCU_EQUATION_PROJECT_3D_DEV_H__
#define __OCU_EQUATION_PROJECT_3D_DEV_H__

#include "ocuequation/sol_divergence3d.h"
#include "ocuequation/sol_gradient3d.h"
#include "ocuequation/sol_mgpressure3d.h"
#include "ocuequation/sol_mgpressuremixed3d.h"
#include "ocuequation/solver.h"
#include "ocuutil/boundary_condition.h"

namespace ocu {

class Sol_ProjectDivergence3DBase : public Solver {
 protected:
  int _nx, _ny, _nz;
  double _hx, _hy, _hz;

  BoundaryCondition convert_bc_to_poisson_eqn(const BoundaryCondition &bc) const;
  bool initialize_base_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                               Grid3DUntyped *u, Grid3DUntyped *v, Grid3DUntyped *w);

 public:
  BoundaryConditionSet bc;

  Sol_ProjectDivergence3DBase() : _nx(0), _ny(0), _nz(0), _hx(0), _hy(0), _hz(0) {}
};

template <typename T>
class Sol_ProjectDivergence3DDeviceStorage : public Sol_ProjectDivergence3DBase {
 protected:
  bool initialize_device_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                                 Grid3DDevice<T> *u, Grid3DDevice<T> *v, Grid3DDevice<T> *w);

 public:
  Grid3DDevice<T> *u, *v, *w;
  Grid3DDevice<T> divergence;

  Sol_ProjectDivergence3DDeviceStorage() : u(0), v(0), w(0) {}
};

template <typename T>
class Sol_ProjectDivergence3DDevice : public Sol_ProjectDivergence3DDeviceStorage<T> {
 public:
  Sol_MultigridPressure3DDevice<T> pressure_solver;
  Sol_Divergence3DDevice<T> divergence_solver;
  Sol_Gradient3DDevice<T> gradient_solver;

  // residual reported by the last pressure solve
  double last_residual;

  Sol_ProjectDivergence3DDevice() : last_residual(0) {}

  bool solve(double tolerance = 1e-5);
  bool solve_divergence_only();
  bool initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                          Grid3DDevice<T> *u, Grid3DDevice<T> *v, Grid3DDevice<T> *w);
};

class Sol_ProjectDivergenceMixed3DDeviceD : public Sol_ProjectDivergence3DDeviceStorage<double> {
 public:
  Sol_MultigridPressureMixed3DDeviceD pressure_solver;
  Sol_Divergence3DDevice<double> divergence_solver;
  Sol_Gradient3DDevice<double> gradient_solver;
  double last_residual;

  Sol_ProjectDivergenceMixed3DDeviceD() : last_residual(0) {}

  bool solve(double tolerance = 1e-5);
  bool solve_divergence_only();
  bool initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                          Grid3DDevice<double> *u, Grid3DDevice<double> *v, Grid3DDevice<double> *w);
};

}  // namespace ocu

#endif
