// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Preconditioned conjugate-gradient pressure solver, the drop-in for
// ocuequation/sol_pcgpressure3d.h: solves -div(c grad p) = rhs for a
// variable coefficient c with periodic or Neumann faces, preconditioned by
// nothing, the operator's diagonal, a few multigrid cycles of the
// constant-coefficient Laplacian, or both (BFBT).
//
// On the GPU: the operator launch also yields <d, Ld> (no product grid, no
// separate reduction), dot products are single deterministic launches whose
// FP64 result lands in pinned host memory, and the two work grids live as
// long as the solver (the reference allocates them on every solve).
#ifndef __OCU_EQUATION_PCG_PRESSURE_3D_DEV_H__
#define __OCU_EQUATION_PCG_PRESSURE_3D_DEV_H__

#include <cmath>

#include "ocuequation/sol_mgpressure3d.h"
#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"
#include "ocuutil/boundary_condition.h"
#include "ocuutil/convergence.h"

namespace ocu {

// sol_pcgpressure3d.h:34-39
enum PreconditionerType {
  PRECOND_NONE,
  PRECOND_JACOBI,
  PRECOND_MULTIGRID,
  PRECOND_BFBT,
};

template <typename T>
class Sol_PCGPressure3DDevice : public Solver {
 protected:
  double _fx, _fy, _fz;
  double _h;
  int _nx, _ny, _nz;
  Grid3DDevice<T> _pressure;
  Grid3DDevice<T> _d_r;  // CG residual, also the multigrid's right-hand side
  Grid3DDevice<T> *_rhs;
  Grid3DDevice<T> *_coefficient;
  Sol_MultigridPressure3DDevice<T> _mg;
  Grid3DDevice<T> _d_y, _d_d;  // work grids of do_pcg / do_cg
  void *_ws;                   // reduction partials (device)
  double *_host, *_host_dev;   // pinned result slot and its device address
  int _iterations;

  double hx() const { return _h / sqrt(_fx); }
  double hy() const { return _h / sqrt(_fy); }
  double hz() const { return _h / sqrt(_fz); }
  T fx_div_hsq() const { return (T)(_fx / (_h * _h)); }
  T fy_div_hsq() const { return (T)(_fy / (_h * _h)); }
  T fz_div_hsq() const { return (T)(_fz / (_h * _h)); }

  virtual bool do_pcg(double tolerance, int max_iter, double &result_l2, double &result_linf);
  virtual bool do_cg(double tolerance, int max_iter, double &result_l2, double &result_linf);

  void invoke_kernel_diag_preconditioner(Grid3DDevice<T> &dst, const Grid3DDevice<T> &src);
  // dst = L src (src's ghosts filled first); with_dot also returns <src, dst>
  T invoke_kernel_apply_laplacian(Grid3DDevice<T> &dst, Grid3DDevice<T> &src, bool with_dot = false);
  T invoke_kernel_dot_product(const Grid3DDevice<T> &a, const Grid3DDevice<T> &b);
  void apply_preconditioner(Grid3DDevice<T> &dst, const Grid3DDevice<T> &src);

 public:
  BoundaryConditionSet bc;
  ConvergenceType convergence;
  PreconditionerType preconditioner;
  int multigrid_cycles;
  bool multigrid_use_fmg;  // false -> use vcycles only
  // not in the reference: print the residual of every iteration (the
  // reference always does)
  bool verbose;

  Sol_PCGPressure3DDevice();
  ~Sol_PCGPressure3DDevice();

  bool solve(double &residual, double tolerance = 1e-6, int max_iter = 1000);
  bool initialize_storage(int nx_val, int ny_val, int nz_val, float hx_val, float hy_val, float hz_val,
                          Grid3DDevice<T> *rhs, Grid3DDevice<T> *coefficient);

  Grid3DDevice<T> &pressure() { return _pressure; }
  const Grid3DDevice<T> &pressure() const { return _pressure; }
  Grid3DDevice<T> &rhs() { return *_rhs; }
  const Grid3DDevice<T> &rhs() const { return *_rhs; }
  Grid3DDevice<T> &coefficient() { return *_coefficient; }
  const Grid3DDevice<T> &coefficient() const { return *_coefficient; }
  // not in the reference: CG iterations of the last solve
  int iterations() const { return _iterations; }
};

typedef Sol_PCGPressure3DDevice<float> Sol_PCGPressure3DDeviceF;
typedef Sol_PCGPressure3DDevice<double> Sol_PCGPressure3DDeviceD;

}  // namespace ocu

#endif
