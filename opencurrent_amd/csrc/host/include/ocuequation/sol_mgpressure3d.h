// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Multigrid pressure solver, the drop-in for ocuequation/sol_mgpressure3d.h.
//
// Sol_MultigridPressure3DBase keeps the reference's FMG / V-cycle driver
// (sol_mgpressure3d.cpp:113-398) and public fields; Sol_MultigridPressure3DDevice<T>
// implements the per-level operations with the kernels of libocu_hip.so:
//  - relaxation: fused red+black sweeps (one launch per iteration, out of
//    place into a second buffer) when no face is periodic, otherwise one
//    launch per colour with the ghost refresh between colours;
//  - restrict_residuals: residual and restriction fused in one launch, the
//    level-0 norms reduced on the device (deterministic, in double) and read
//    back through pinned memory - the only synchronisation per V-cycle;
//  - the coarsest level: relaxed by one LDS-resident workgroup on the GPU
//    with the same arithmetic as the reference's host relaxation
//    (sol_mgpressure3ddev.cu:82-188), without the two PCIe copies.
#ifndef __OCU_EQUATION_MULTIGRID_PRESSURE_3D_DEV_H__
#define __OCU_EQUATION_MULTIGRID_PRESSURE_3D_DEV_H__

#include <cmath>
#include <functional>
#include <string>
#include <vector>

#include "ocuequation/solver.h"
#include "ocustorage/grid3d.h"
#include "ocuutil/boundary_condition.h"
#include "ocuutil/convergence.h"

namespace ocu {

struct int3_ {
  int x, y, z;
};

class Sol_MultigridPressure3DBase : public Solver {
 protected:
  enum RelaxOrder { RO_RED_BLACK, RO_BLACK_RED, RO_SYMMETRIC };

  int _num_levels;
  double *_h;
  double _omega;
  double _fx, _fy, _fz;
  int3_ *_dim;

  int nx(int level) const { return _dim[level].x; }
  int ny(int level) const { return _dim[level].y; }
  int nz(int level) const { return _dim[level].z; }
  double get_h(int level) const { return _h[level]; }
  double hx(int level) const { return get_h(level) / sqrt(_fx); }
  double hy(int level) const { return get_h(level) / sqrt(_fy); }
  double hz(int level) const { return get_h(level) / sqrt(_fz); }

  int num_levels(int nx_val, int ny_val, int nz_val);
  double optimal_omega(double fx, double fy, double fz);
  double bc_diag_mod(const BoundaryCondition &bc, double factor) const;

  virtual bool do_fmg(double tolerance, int max_iter, double &result_l2, double &result_linf);
  virtual bool do_vcycle(double tolerance, int max_iter, double &result_l2, double &result_linf);
  bool initialize_base_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val);

  virtual void clear_zero(int level) = 0;
  virtual void apply_boundary_conditions(int level) = 0;
  virtual void relax(int level, int iterations, RelaxOrder order) = 0;
  virtual void restrict_residuals(int fine_level, int coarse_level, double *l2, double *linf) = 0;
  virtual void prolong(int coarse_level, int fine_level) = 0;
  // bring every level's storage to the state the reference would hold
  // (implementations may defer work between the steps above)
  virtual void settle() {}
  // first level of a one-launch bottom of the V-cycle (levels bottom..coarse
  // in one workgroup), or -1; bottom_vcycle(top) runs that part from `top`
  virtual int bottom_start() const { return -1; }
  virtual void bottom_vcycle(int /*top*/) {}
  // restrict_residuals with norms in two halves: the launches, then the
  // wait for the norms and their values (the default runs it whole)
  virtual void restrict_residuals_launch(int fine_level, int coarse_level, bool norms) {
    restrict_residuals(fine_level, coarse_level, norms ? &_pend_l2 : 0, norms ? &_pend_linf : 0);
  }
  virtual void norms_read(int /*fine_level*/, double *l2, double *linf) {
    if (l2) *l2 = _pend_l2;
    if (linf) *linf = _pend_linf;
  }
  // the launches between two norm read-backs of do_fmg's finest cycles (the
  // rest of one V-cycle and the next one's level-0 smoothing and residual);
  // an implementation may record them once and replay them
  virtual void run_segment(const std::function<void()> &body) { body(); }
  double _pend_l2 = 0, _pend_linf = 0;

 public:
  BoundaryConditionSet bc;
  int nu1;
  int nu2;
  ConvergenceType convergence;
  bool make_symmetric_operator;
  // print the per-cycle residual like the reference (sol_mgpressure3d.cpp:201-202)
  bool verbose;

  // statistics of the last solve / run
  int num_vcycles;
  std::vector<double> residual_history;

  bool solve(double &residual, double tolerance = 1e-6, int max_iter = 4);
  bool run_vcycles(int count);
  bool run_fmg(int count);

  int num_levels() const { return _num_levels; }
  double omega() const { return _omega; }

  Sol_MultigridPressure3DBase();
  virtual ~Sol_MultigridPressure3DBase();
};

class Sol_MultigridPressureMixed3DDeviceD;

template <typename T>
class Sol_MultigridPressure3DDevice : public Sol_MultigridPressure3DBase {
 protected:
  friend class Sol_MultigridPressureMixed3DDeviceD;

  Grid3DDevice<T> **_r_grid;
  Grid3DDevice<T> **_u_grid;
  Grid3DDevice<T> **_b_grid;
  Grid3DDevice<T> **_u_alt;  // second buffer per level for out-of-place sweeps
  Grid3DDevice<T> _pressure;
  bool _update_bc_between_colors;
  bool _fused_ok;  // every face Neumann or Dirichlet
  // the LDS bottom can run: faces Neumann, Dirichlet or periodic in pairs,
  // with an even number of cells along each periodic axis on its levels
  bool _bottom_ok;
  // 0: the coarsest level is relaxed with the reference's host arithmetic in T;
  // 1: in double on float grids (the coarsest solve of the mixed solver)
  int _coarse_mode;
  double *_norm_ws;    // device: NORM_PARTS (sum, max) pairs + 2 results
  double *_norm_host;  // pinned host: 2 results
  double *_norm_dev;   // the same memory as the device sees it
  // _u_zero[l]: u at level l is logically the zero grid with its level-1
  // ghosts (clear_zero + apply_boundary_conditions) but memory was not
  // written; the next fused relaxation starts from zero without reading u
  static const int kMaxLevels = 40;
  bool _u_zero[kMaxLevels];
  // _u_pro[l]: the prolongation of level l+1 into level l (with its level-1
  // fill) is pending: the next fused relaxation of level l adds it while it
  // reads u, and anything else that reads level l, or writes level l+1,
  // runs it first
  bool _u_pro[kMaxLevels];
  // the fused sweep runs at level l (every face Neumann/Dirichlet, or the
  // periodic pairs it takes)
  bool _fused_lvl[kMaxLevels];

  Grid3DDevice<T> &get_u(int level) { return *_u_grid[level]; }
  Grid3DDevice<T> &get_r(int level) { return *_r_grid[level]; }
  Grid3DDevice<T> &get_b(int level) { return *_b_grid[level]; }

  virtual BoundaryConditionSet get_bc_at_level(int level) const;
  void level_params(int level, void *params_out) const;
  void free_storage();

  virtual void apply_boundary_conditions(int level);
  virtual void relax(int level, int iterations, RelaxOrder order);
  virtual void restrict_residuals(int fine_level, int coarse_level, double *l2, double *linf);
  virtual void prolong(int coarse_level, int fine_level);
  virtual void clear_zero(int level);

  void relax_colour(int level, int color);
  // one colour that also writes the level-1 ghosts it changes (no fill
  // launch); relax_bc_ok: the level's faces allow it
  void relax_colour_bc(int level, int color);
  bool relax_bc_ok(int level) const;
  void relax_coarsest(int level, int iterations, RelaxOrder order);
  bool defer_ok(int level) const;
  void materialize(int level);
  void run_prolong(int coarse_level, int fine_level);
  // level l is about to change: a pending prolongation that reads it runs first
  void flush_reader(int level) {
    if (level >= 1 && _u_pro[level - 1]) materialize(level - 1);
  }
  virtual void settle();
  int _bottom_level;  // first level of the LDS bottom, -1 if none fits
  virtual int bottom_start() const;
  virtual void bottom_vcycle(int top);

  virtual void restrict_residuals_launch(int fine_level, int coarse_level, bool norms);
  virtual void norms_read(int fine_level, double *l2, double *linf);
  virtual void run_segment(const std::function<void()> &body);
  bool _norms_pending;  // norm kernels launched, values not read yet
  unsigned long long _norm_tag;  // their completion tag (0: read after a stream synchronisation)
  bool _capturing;  // launches are being recorded into a graph (no tags: a replay would repeat them)
  // recorded segments: the solver state they start from (options, stream,
  // deferred clears and prolongations per level), the flags they leave, the
  // level-0 sweep counts they add, and the executable graph
  struct Segment {
    std::vector<long long> key;
    std::vector<char> out;
    long long sweeps[3];
    void *exec;
  };
  std::vector<Segment> _segments;
  bool _graphs_failed;  // a capture failed: run segments directly from then on
  std::vector<long long> segment_key(void *stream) const;
  void drop_segments();

 public:
  bool disable_relax_on_host;
  // one launch per red+black iteration when no face is periodic
  bool use_fused_relax;
  // also store the residual grids in restrict_residuals (read_r())
  bool store_residuals;
  // run the small coarse levels of each V-cycle as one LDS-resident launch
  bool use_bottom;
  // with use_fused_relax: a prolongation is added by the fused relaxation
  // that follows it instead of a launch of its own
  bool use_fused_prolong;
  // replay the launches between two norm read-backs of the finest V-cycles
  // as a HIP graph once they repeat (the same kernels with the same
  // arguments, so bit-identical to launching them one by one; off by
  // default: no faster than the launches themselves on MI355X)
  bool use_graphs;
  long long graph_records, graph_replays;  // segments recorded / replayed since construction
  // fused red+black sweeps launched at level 0 since construction: [0]
  // plain, [1] starting from the deferred zero grid, [2] adding the pending
  // prolongation (for profiling: the three run at different speeds)
  long long level0_sweeps[3];

  Sol_MultigridPressure3DDevice();
  ~Sol_MultigridPressure3DDevice();

  virtual bool initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val, double hy_val, double hz_val,
                                  Grid3DDevice<T> *rhs);

  Grid3DDevice<T> &pressure() { return _pressure; }
  const Grid3DDevice<T> &pressure() const { return _pressure; }

  // the relaxation parameters of a level (h, f, omega, diagonal mods) as
  // the kernels receive them; false for a level that does not exist
  bool get_level_params(int level, ocu_mg_level_params *out) const {
    if (level < 0 || level >= _num_levels || !out) return false;
    level_params(level, out);
    return true;
  }

  const Grid3DDevice<T> &read_u(int level) const { return *_u_grid[level]; }
  const Grid3DDevice<T> &read_r(int level) const { return *_r_grid[level]; }
  const Grid3DDevice<T> &read_b(int level) const { return *_b_grid[level]; }
};

typedef Sol_MultigridPressure3DDevice<float> Sol_MultigridPressure3DDeviceF;
typedef Sol_MultigridPressure3DDevice<double> Sol_MultigridPressure3DDeviceD;

}  // namespace ocu

#endif
