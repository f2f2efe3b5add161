// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Mixed-precision pressure solver with the public API of
// ocuequation/sol_mgpressuremixed3d.h (a drop-in for the FP64 solver).
//
// The reference runs an FP32 FMG and then a full FP64 FMG
// (sol_mgpressuremixed3ddevd.cpp:79-100).  This class instead runs
// iterative refinement: FP64 outer residual and correction, FP32 V-cycles on
// the correction equation (homogeneous boundary conditions on every level),
// the coarsest FP32 level relaxed in FP64, and FP64 mean removal for
// singular (no Dirichlet face) problems.  solve(residual, tol, max_iter)
// keeps the reference meaning: tol is absolute in the `convergence` norm
// (RMS for CONVERGENCE_L2, max for CONVERGENCE_LINF) of the FP64 residual;
// max_iter bounds the refinement iterations.
#ifndef __OCU_EQUATION_MULTIGRID_PRESSURE_MIXED_3D_H__
#define __OCU_EQUATION_MULTIGRID_PRESSURE_MIXED_3D_H__

#include "ocuequation/sol_mgpressure3d.h"

namespace ocu {

class Sol_MultigridPressureMixed3DDeviceD : public Sol_MultigridPressure3DBase {
  Sol_MultigridPressure3DDevice<float> _mg_f;  // FP32 hierarchy: pressure() = correction, b(0) = scaled residual
  Grid3DDevice<double> _x;                     // FP64 solution
  Grid3DDevice<double> *_rhs;                  // caller's FP64 right-hand side
  Grid3DDevice<float> _r32;                    // scaled FP32 residual (level-0 rhs of _mg_f)
  double *_ws;                                 // device: NORM_PARTS quadruples, NORM_PARTS sums, the shift
  double *_host;                               // pinned: results
  double *_host_dev;                           // the same memory as the device sees it
  bool _singular;

  // never called: the refinement loop drives _mg_f directly
  virtual void apply_boundary_conditions(int) {}
  virtual void relax(int, int, RelaxOrder) {}
  virtual void restrict_residuals(int, int, double *, double *) {}
  virtual void prolong(int, int) {}
  virtual void clear_zero(int level);
  virtual bool do_fmg(double tolerance, int max_iter, double &result_l2, double &result_linf);

  bool fp64_residual(double bshift, double scale, double &sumsq, double &maxabs, double &sum, double &xsum);
  bool copy_faces() const;
  bool interior_mean(const Grid3DDevice<double> &g, double &mean);
  void vcycle_f32(int fine_level);
  void fmg_f32(int cycles0);

 public:
  // V-cycles per refinement iteration; the first iteration may start with an FMG
  int ir_cycles;
  int ir_cycles_first;
  bool ir_fmg_first;
  // after the first iteration, run as many V-cycles (at most ir_cycles) as the
  // observed contraction per cycle needs to reach half the tolerance
  bool ir_adaptive;
  // options of the FP32 hierarchy (Sol_MultigridPressure3DDevice fields)
  bool use_fused_relax;
  bool use_bottom;
  bool use_fused_prolong;  // passed to the FP32 solver (see Sol_MultigridPressure3DDevice)
  // statistics of the last solve
  int ir_iterations;
  int total_vcycles;
  double rhs_l2;  // RMS of the (mean-removed) right-hand side

  Sol_MultigridPressureMixed3DDeviceD();
  ~Sol_MultigridPressureMixed3DDeviceD();

  bool initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                          Grid3DDevice<double> *rhs);

  Grid3DDevice<double> &pressure() { return _x; }
  const Grid3DDevice<double> &pressure() const { return _x; }
  // level-0 fused sweeps of the FP32 hierarchy (Sol_MultigridPressure3DDevice::level0_sweeps)
  const long long *level0_sweeps() const { return _mg_f.level0_sweeps; }
  bool get_level_params(int level, ocu_mg_level_params *out) const { return _mg_f.get_level_params(level, out); }
};

}  // namespace ocu

#endif
