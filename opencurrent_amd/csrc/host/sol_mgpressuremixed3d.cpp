// Mixed-precision iterative refinement behind the
// Sol_MultigridPressureMixed3DDeviceD API (see the header for the algorithm).
#include "ocuequation/sol_mgpressuremixed3d.h"

#include <cmath>
#include <cstdio>

#include "ocu_hip.h"
#include "ocustorage/grid3dboundary.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/kernel_wrapper.h"
#include "ocuutil/thread.h"

namespace ocu {

Sol_MultigridPressureMixed3DDeviceD::Sol_MultigridPressureMixed3DDeviceD() {
  _rhs = 0;
  _ws = 0;
  _host = 0;
  _host_dev = 0;
  _singular = false;
  ir_cycles = 4;
  ir_cycles_first = 4;
  ir_fmg_first = false;
  ir_adaptive = true;
  use_fused_relax = true;
  use_bottom = true;
  use_fused_prolong = true;
  ir_iterations = 0;
  total_vcycles = 0;
  rhs_l2 = 0;
}

Sol_MultigridPressureMixed3DDeviceD::~Sol_MultigridPressureMixed3DDeviceD() {
  ocu_free(_ws);
  ocu_host_free(_host);
}

void Sol_MultigridPressureMixed3DDeviceD::clear_zero(int level) {
  if (level == 0) check_ok(_x.clear_zero());
}

bool Sol_MultigridPressureMixed3DDeviceD::initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val,
                                                             double hy_val, double hz_val, Grid3DDevice<double> *rhs) {
  if (!rhs || rhs->nx() != nx_val || rhs->ny() != ny_val || rhs->nz() != nz_val) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - rhs dimension mismatch\n");
    return false;
  }
  if (rhs->gx() < 1 || rhs->gy() < 1 || rhs->gz() < 1) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - rhs needs a ghost layer\n");
    return false;
  }
  if (!initialize_base_storage(nx_val, ny_val, nz_val, hx_val, hy_val, hz_val)) return false;
  _rhs = rhs;
  if (!_x.init_congruent(*rhs) || !_r32.init_congruent(*rhs)) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - allocation failed\n");
    return false;
  }
  _x.clear_zero();
  _r32.clear_zero();
  _mg_f.bc = bc;
  _mg_f.bc.make_homogeneous();
  _mg_f.nu1 = nu1;
  _mg_f.nu2 = nu2;
  _mg_f.convergence = CONVERGENCE_NONE;
  _mg_f._coarse_mode = 1;
  _mg_f.use_fused_relax = use_fused_relax;
  _mg_f.use_bottom = use_bottom;
  _mg_f.use_fused_prolong = use_fused_prolong;
  if (!_mg_f.initialize_storage(nx_val, ny_val, nz_val, hx_val, hy_val, hz_val, &_r32)) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - failed to initialize float solver\n");
    return false;
  }
  _singular = !(bc.xpos.check_type(BC_DIRICHELET) || bc.xneg.check_type(BC_DIRICHELET) ||
                bc.ypos.check_type(BC_DIRICHELET) || bc.yneg.check_type(BC_DIRICHELET) ||
                bc.zpos.check_type(BC_DIRICHELET) || bc.zneg.check_type(BC_DIRICHELET));
  void *w = 0, *h = 0;
  if (ocu_malloc(&w, sizeof(double) * (5 * ocu_norm_parts() + 8)) != OCU_OK ||
      ocu_host_malloc(&h, sizeof(double) * 8) != OCU_OK) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - workspace allocation failed\n");
    return false;
  }
  _ws = (double *)w;
  _host = (double *)h;
  void *hd = nullptr;
  if (ocu_host_device_pointer(&hd, h) != OCU_OK) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::initialize_storage - pinned results not device-visible\n");
    return false;
  }
  _host_dev = (double *)hd;
  return true;
}

// r = (b - bshift) - A x in FP64, written to _r32 scaled by `scale`; returns
// the FP64 norms of r and the interior sum of x (blocking: the refinement
// loop needs them)
bool Sol_MultigridPressureMixed3DDeviceD::fp64_residual(double bshift, double scale, double &sumsq, double &maxabs,
                                                        double &sum, double &xsum) {
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d x = _x.desc(), b = _rhs->desc(), r = _r32.desc();
  bool ok = ok_status(ocu_ir_residual(&x, &b, bshift, &r, scale, get_h(0), _fx, _fy, _fz, _ws, s), "ir_residual");
  // into pinned host memory, with a completion tag the host polls for
  const unsigned long long tag = next_completion_tag();
  ok = ok && ok_status(ocu_norm4_finalize_tagged(_ws, _host_dev, tag, s), "ir_residual");
  if (ok && !wait_completion_tag(_host + 4, tag, s)) {
    printf("[ERROR] Sol_MultigridPressureMixed3DDeviceD::fp64_residual - norms did not arrive\n");
    ok = false;
  }
  sumsq = _host[0];
  maxabs = _host[1];
  sum = _host[2];
  xsum = _host[3];
  return ok;
}

// every face of x (and so of the homogeneous correction) is a copy of an
// interior cell: the correction can update the face ghosts itself
bool Sol_MultigridPressureMixed3DDeviceD::copy_faces() const {
  const BoundaryCondition *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (!(b->type == BC_PERIODIC || (b->type == BC_NEUMANN && b->value == 0))) return false;
  return nu1 >= 1;
}

bool Sol_MultigridPressureMixed3DDeviceD::interior_mean(const Grid3DDevice<double> &g, double &mean) {
  double sum = 0;
  bool ok = g.reduce_double(3, sum);
  mean = sum / ((double)g.nx() * g.ny() * g.nz());
  return ok;
}

// V-cycles for refinement iteration it >= 1 when the previous iteration ran
// prev_cycles V-cycles and took the residual from prev_res to res: the
// observed contraction per cycle rho = (res/prev_res)^(1/prev_cycles) is
// extrapolated to reach half the tolerance, within [1, cap]
static int adaptive_cycles(double res, double prev_res, int prev_cycles, double tol, int cap) {
  if (!(prev_res > 0) || !(res > 0) || !(res < prev_res) || prev_cycles <= 0) return cap;
  const double rho = pow(res / prev_res, 1.0 / prev_cycles);
  const double need = log(0.5 * tol / res) / log(rho);
  if (!(need > 1)) return 1;
  const double n = ceil(need);
  return n > cap ? cap : (int)n;
}

// one FP32 V-cycle from fine_level down and back (the structure of
// do_fmg's inner loop, sol_mgpressure3d.cpp:184-239, without a convergence
// test; the coarsest level is relaxed in FP64)
void Sol_MultigridPressureMixed3DDeviceD::vcycle_f32(int fine_level) {
  int coarse = _num_levels - 1;
  // levels top..coarse in one launch when the FP32 solver has a bottom
  const int bottom = _mg_f.bottom_start();
  const int top = (bottom >= 1 && fine_level < coarse) ? max2(bottom, fine_level) : -1;
  const int stop = top >= 0 ? top : coarse;
  for (int level = fine_level; level < stop; level++) {
    _mg_f.relax(level, nu1, RO_RED_BLACK);
    _mg_f.clear_zero(level + 1);
    _mg_f.apply_boundary_conditions(level + 1);
    _mg_f.restrict_residuals(level, level + 1, 0, 0);
  }
  if (top >= 0) {
    _mg_f.bottom_vcycle(top);
  } else {
    int coarse_iters = max3(nx(coarse) * ny(coarse), ny(coarse) * nz(coarse), nx(coarse) * nz(coarse)) / 2;
    _mg_f.relax(coarse, coarse_iters, RO_RED_BLACK);
  }
  for (int level = stop - 1; level >= fine_level; level--) {
    _mg_f.prolong(level + 1, level);
    if (level > 0) _mg_f.relax(level, nu2, RO_RED_BLACK);
  }
  total_vcycles++;
}

void Sol_MultigridPressureMixed3DDeviceD::fmg_f32(int cycles0) {
  for (int level = 0; level < _num_levels - 1; level++) {
    _mg_f.clear_zero(level + 1);
    _mg_f.apply_boundary_conditions(level + 1);
    _mg_f.restrict_residuals(level, level + 1, 0, 0);
  }
  for (int fine_level = _num_levels - 1; fine_level >= 0; fine_level--) {
    int nc = fine_level == 0 ? cycles0 : 1;
    for (int c = 0; c < nc; c++) vcycle_f32(fine_level);
    if (fine_level > 0) _mg_f.prolong(fine_level, fine_level - 1);
  }
}

bool Sol_MultigridPressureMixed3DDeviceD::do_fmg(double tolerance, int max_iter, double &result_l2,
                                                 double &result_linf) {
  clear_error();
  _mg_f.clear_error();
  ir_iterations = 0;
  total_vcycles = 0;
  residual_history.clear();
  const double n = (double)nx(0) * ny(0) * nz(0);
  void *s = ThreadManager::get_compute_stream();

  // sum, sum of squares and max|b| in one pass
  ocu_grid3d bd = _rhs->desc();
  const unsigned long long tag = next_completion_tag();
  check_ok(ok_status(ocu_grid_stats3(&bd, _ws, s), "rhs statistics") &&
           ok_status(ocu_norm3_finalize_tagged(_ws, _host_dev, tag, s), "rhs statistics") &&
           wait_completion_tag(_host + 3, tag, s));
  const double bsq = _host[0], bmax = _host[1], bsum = _host[2];
  double bm = _singular ? bsum / n : 0.0;
  double var = bsq / n - bm * bm;
  rhs_l2 = sqrt(var > 0 ? var : 0);

  BoundaryConditionSet bc0 = bc;
  check_ok(apply_3d_boundary_conditions_level1(_x, bc0, hx(0), hy(0), hz(0)));

  double prev_max = bmax, prev_res = 0, xsum = 0;
  int prev_cycles = 0;
  bool converged = false;
  const bool faces = copy_faces();
  double *shift = _singular ? _ws + 5 * ocu_norm_parts() : nullptr;
  for (int it = 0;; it++) {
    int ex = 0;
    frexp(prev_max > 0 ? prev_max : 1.0, &ex);
    double scale = ldexp(1.0, -ex), inv_scale = ldexp(1.0, ex);
    double s2 = 0, mx = 0, sm = 0;
    if (!fp64_residual(bm, scale, s2, mx, sm, xsum)) {
      add_error();
      break;
    }
    result_l2 = sqrt(s2 / n);
    result_linf = mx;
    double res = (convergence & CONVERGENCE_CRITERIA_LINF) ? result_linf : result_l2;
    residual_history.push_back(res);
    if (verbose) printf("[INFO] mixed IR %d: residual L2 = %.6e, Linf = %.6e\n", it, result_l2, result_linf);
    if (res <= tolerance) {
      converged = true;
      break;
    }
    if (it >= max_iter) break;
    prev_max = mx;
    ir_iterations++;
    _mg_f.clear_zero(0);
    if (it == 0 && ir_fmg_first) {
      fmg_f32(ir_cycles_first);
      prev_cycles = 0;
    } else {
      _mg_f.apply_boundary_conditions(0);
      int nc = it == 0 ? ir_cycles_first : ir_cycles;
      if (ir_adaptive && it > 0) nc = adaptive_cycles(res, prev_res, prev_cycles, tolerance, nc);
      for (int c = 0; c < nc; c++) vcycle_f32(0);
      prev_cycles = nc;
    }
    prev_res = res;
    _mg_f.relax(0, nu1, RO_RED_BLACK);
    ocu_grid3d x = _x.desc(), e = _mg_f.pressure().desc();
    // singular problems: the correction also removes the interior mean of
    // x + e/s (from the residual pass's sum of x and one pass over e), so x
    // stays mean-free and needs no final pass
    if (shift) ok_status(ocu_ir_shift(&e, inv_scale, xsum, _ws + 4 * ocu_norm_parts(), shift, s), "ir_shift");
    ok_status(ocu_ir_correct_ex(&x, &e, inv_scale, shift, faces ? 1 : 0, s), "ir_correct");
    if (!faces) check_ok(apply_3d_boundary_conditions_level1(_x, bc0, hx(0), hy(0), hz(0)));
    if (_mg_f.any_error()) add_error();
  }
  _mg_f.settle();
  if (_singular && ir_iterations == 0) {
    double xm = 0;
    check_ok(interior_mean(_x, xm));
    ocu_grid3d x = _x.desc();
    ok_status(ocu_grid_add_scalar(&x, -xm, s), "mean removal");
    check_ok(apply_3d_boundary_conditions_level1(_x, bc0, hx(0), hy(0), hz(0)));
  }
  num_vcycles = total_vcycles;
  return converged && !any_error();
}

}  // namespace ocu
