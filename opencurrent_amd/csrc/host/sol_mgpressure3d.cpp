// Multigrid pressure solver: the FMG / V-cycle driver of
// Sol_MultigridPressure3DBase (sol_mgpressure3d.cpp:26-437) and the device
// operations of Sol_MultigridPressure3DDevice<T> (sol_mgpressure3ddev.cu:52-482)
// over libocu_hip.so.
#include "ocuequation/sol_mgpressure3d.h"

#include <cstdio>
#include <cstring>

#include "ocu_hip.h"
#include "ocustorage/grid3dboundary.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/kernel_wrapper.h"
#include "ocuutil/thread.h"

// whether a solver records and replays its V-cycle segments unless told
// otherwise: off, since replaying the graph measured no faster than the
// launches themselves on MI355X (the macro exists for tuning builds)
#ifndef OCU_MG_GRAPHS_DEFAULT
#define OCU_MG_GRAPHS_DEFAULT 0
#endif

namespace ocu {

// ------------------------------------------------------------------------
// Base: the driver (sol_mgpressure3d.cpp)
// ------------------------------------------------------------------------

Sol_MultigridPressure3DBase::Sol_MultigridPressure3DBase() {
  _num_levels = 0;
  _h = 0;
  _dim = 0;
  _omega = 1.0f;
  _fx = _fy = _fz = 1;
  nu1 = 2;
  nu2 = 2;
  convergence = CONVERGENCE_L2;
  make_symmetric_operator = false;
  verbose = false;
  num_vcycles = 0;
}

Sol_MultigridPressure3DBase::~Sol_MultigridPressure3DBase() {
  delete[] _h;
  delete[] _dim;
}

// sol_mgpressure3d.cpp:49-64
double Sol_MultigridPressure3DBase::bc_diag_mod(const BoundaryCondition &local_bc, double factor) const {
  if (local_bc.type == BC_PERIODIC) return 0;
  if (local_bc.type == BC_NONE) return 0;
  if (local_bc.type == BC_DIRICHELET) return factor;
  if (local_bc.type == BC_NEUMANN) return -factor;
  printf("[WARNING] Sol_MultigridPressure3DBase::bc_diag_mod - unknown boundary condition type %d\n", local_bc.type);
  return 0;
}

// sol_mgpressure3d.cpp:67-92: halve while every dimension is even, stopping
// before a dimension drops below 2 or reaches 5 or 7
int Sol_MultigridPressure3DBase::num_levels(int nx_val, int ny_val, int nz_val) {
  int cnx = nx_val / 2, cny = ny_val / 2, cnz = nz_val / 2;
  int levels = 0;
  while (2 * cnx == nx_val && 2 * cny == ny_val && 2 * cnz == nz_val) {
    levels++;
    nx_val = cnx; ny_val = cny; nz_val = cnz;
    cnx /= 2; cny /= 2; cnz /= 2;
    if (cnx < 2 || nx_val == 5 || nx_val == 7) break;
    if (cny < 2 || ny_val == 5 || ny_val == 7) break;
    if (cnz < 2 || nz_val == 5 || nz_val == 7) break;
  }
  return levels;
}

// Yavneh (1996) red-black SOR factor; sol_mgpressure3d.cpp:94-109 (including
// its float literal in 1.0f - c_max^2)
double Sol_MultigridPressure3DBase::optimal_omega(double fx, double fy, double fz) {
  double f_min = min3(fx, fy, fz);
  double c_max = (1.0 - f_min / (fx + fy + fz));
  double c_max_sqr = c_max * c_max;
  return 2.0 / (1.0 + sqrt(1.0f - c_max_sqr));
}

bool Sol_MultigridPressure3DBase::initialize_base_storage(int nx_val, int ny_val, int nz_val, double hx_val,
                                                          double hy_val, double hz_val) {
  if (!check_float(hx_val) || !check_float(hy_val) || !check_float(hz_val)) {
    printf("[ERROR] Sol_MultigridPressure3DBase::initialize_storage - garbage hx,hy,hz value %f %f %f\n", hx_val,
           hy_val, hz_val);
    return false;
  }
  _num_levels = num_levels(nx_val, ny_val, nz_val);
  delete[] _h;
  delete[] _dim;
  _h = new double[_num_levels > 0 ? _num_levels : 1];
  _dim = new int3_[_num_levels > 0 ? _num_levels : 1];
  _h[0] = min3(hx_val, hy_val, hz_val);
  _fx = (_h[0] * _h[0]) / (hx_val * hx_val);
  _fy = (_h[0] * _h[0]) / (hy_val * hy_val);
  _fz = (_h[0] * _h[0]) / (hz_val * hz_val);
  _omega = optimal_omega(_fx, _fy, _fz);
  _dim[0].x = nx_val;
  _dim[0].y = ny_val;
  _dim[0].z = nz_val;
  for (int level = 1; level < _num_levels; level++) {
    _h[level] = get_h(level - 1) * 2;
    _dim[level].x = nx(level - 1) / 2;
    _dim[level].y = ny(level - 1) / 2;
    _dim[level].z = nz(level - 1) / 2;
  }
  return true;
}

static inline double pick(int convergence, double l2, double linf) {
  return (convergence & CONVERGENCE_CRITERIA_L2) ? l2 : (convergence & CONVERGENCE_CRITERIA_LINF) ? linf : 1e20;
}

// sol_mgpressure3d.cpp:113-251
bool Sol_MultigridPressure3DBase::do_fmg(double tolerance, int max_iter, double &result_l2, double &result_linf) {
  clear_error();
  result_l2 = result_linf = 0;
  apply_boundary_conditions(0);
  double orig_l2 = 0, orig_linf = 0;
  restrict_residuals(0, 0, (convergence & CONVERGENCE_CALC_L2) ? &orig_l2 : 0,
                     (convergence & CONVERGENCE_CALC_LINF) ? &orig_linf : 0);
  double orig_error = pick(convergence, orig_l2, orig_linf);
  if (orig_error < tolerance) {
    if (convergence & CONVERGENCE_CALC_L2) result_l2 = orig_l2;
    if (convergence & CONVERGENCE_CALC_LINF) result_linf = orig_linf;
    return true;
  }

  int coarse_level = _num_levels - 1;
  for (int level = 0; level < _num_levels - 1; level++) {
    clear_zero(level + 1);
    apply_boundary_conditions(level + 1);
    restrict_residuals(level, level + 1, 0, 0);
  }

  const int bottom = bottom_start();
  const bool want_l2 = (convergence & CONVERGENCE_CALC_L2) != 0, want_linf = (convergence & CONVERGENCE_CALC_LINF) != 0;
  for (int fine_level = _num_levels - 1; fine_level >= 0; fine_level--) {
    int level_ncyc = (fine_level == 0) ? max_iter + 1 : 1;
    // levels top..coarse_level run as one launch when a bottom exists
    const int top = (bottom >= 1 && fine_level < coarse_level) ? max2(bottom, fine_level) : -1;
    const int stop = top >= 0 ? top : coarse_level;
    // pre-smoothing and restriction from level `from` down, the bottom (or
    // the coarsest relaxation), and the way back up
    auto down_from = [&](int from) {
      for (int level = from; level < stop; level++) {
        relax(level, nu1, RO_RED_BLACK);
        clear_zero(level + 1);
        apply_boundary_conditions(level + 1);
        restrict_residuals(level, level + 1, 0, 0);
      }
    };
    auto bottom_and_up = [&]() {
      if (top >= 0) {
        bottom_vcycle(top);
      } else {
        int cl = coarse_level;
        int coarse_iters = max3(nx(cl) * ny(cl), ny(cl) * nz(cl), nx(cl) * nz(cl)) / 2;
        relax(coarse_level, coarse_iters, make_symmetric_operator ? RO_SYMMETRIC : RO_RED_BLACK);
      }
      for (int level = stop - 1; level >= fine_level; level--) {
        prolong(level + 1, level);
        if (level > 0) relax(level, nu2, make_symmetric_operator ? RO_BLACK_RED : RO_RED_BLACK);
      }
    };
    if (fine_level > 0 || stop < 1) {
      for (int i_cyc = 0; i_cyc < level_ncyc; i_cyc++) {
        if (fine_level == 0) num_vcycles++;
        down_from(fine_level);
        bottom_and_up();
      }
    } else {
      // the finest cycles: level 0's pre-smoothing ends in the residual with
      // its norms, which are read before the rest of the cycle runs; what is
      // launched between two read-backs (the rest of one cycle and the start
      // of the next) is one segment
      auto head = [&]() {
        relax(0, nu1, RO_RED_BLACK);
        clear_zero(1);
        apply_boundary_conditions(1);
        restrict_residuals_launch(0, 1, want_l2 || want_linf);
      };
      num_vcycles++;
      head();
      for (int i_cyc = 0;; i_cyc++) {
        norms_read(0, want_l2 ? &result_l2 : 0, want_linf ? &result_linf : 0);
        double residual = pick(convergence, result_l2, result_linf);
        residual_history.push_back(residual);
        if (verbose) printf("%d: residual = %.12f,%.12f\n", i_cyc, result_linf, result_l2);
        if (residual < tolerance) {
          // ghosts of u[0] are stale after the pre-smoothing
          apply_boundary_conditions(0);
          return !any_error();
        }
        if (i_cyc + 1 >= level_ncyc) {
          down_from(1);
          bottom_and_up();
          break;
        }
        run_segment([&]() {
          down_from(1);
          bottom_and_up();
          head();
        });
        num_vcycles++;
      }
    }
    if (fine_level > 0) prolong(fine_level, fine_level - 1);
  }
  if (!(convergence & CONVERGENCE_CRITERIA_NONE) && verbose)
    printf("[WARNING] Sol_MultigridPressure3DBase::do_fmg - Failed to converge, error after: L2 = %.12f, Linf = %.12f\n",
           result_l2, result_linf);
  return false;
}

// sol_mgpressure3d.cpp:253-330
bool Sol_MultigridPressure3DBase::do_vcycle(double tolerance, int max_iter, double &result_l2, double &result_linf) {
  clear_error();
  int coarse_level = _num_levels - 1;
  apply_boundary_conditions(0);
  double orig_l2 = 0, orig_linf = 0;
  restrict_residuals(0, 0, (convergence & CONVERGENCE_CALC_L2) ? &orig_l2 : 0,
                     (convergence & CONVERGENCE_CALC_LINF) ? &orig_linf : 0);
  if (pick(convergence, orig_l2, orig_linf) < tolerance) {
    if (convergence & CONVERGENCE_CALC_L2) result_l2 = orig_l2;
    if (convergence & CONVERGENCE_CALC_LINF) result_linf = orig_linf;
    return true;
  }
  const int bottom = bottom_start();
  const int stop = bottom >= 1 ? bottom : coarse_level;
  for (int i_cyc = 0; i_cyc < max_iter; i_cyc++) {
    num_vcycles++;
    for (int level = 0; level < stop; level++) {
      relax(level, nu1, RO_RED_BLACK);
      clear_zero(level + 1);
      apply_boundary_conditions(level + 1);
      if (level == 0) {
        restrict_residuals(0, 1, (convergence & CONVERGENCE_CALC_L2) ? &result_l2 : 0,
                           (convergence & CONVERGENCE_CALC_LINF) ? &result_linf : 0);
        double residual = pick(convergence, result_l2, result_linf);
        residual_history.push_back(residual);
        if (residual < tolerance) {
          apply_boundary_conditions(0);
          return !any_error();
        }
      } else {
        restrict_residuals(level, level + 1, 0, 0);
      }
    }
    if (bottom >= 1) {
      bottom_vcycle(bottom);
    } else {
      int cl = coarse_level;
      int coarse_iters = max3(nx(cl) * ny(cl), ny(cl) * nz(cl), nx(cl) * nz(cl)) / 2;
      relax(coarse_level, coarse_iters, make_symmetric_operator ? RO_SYMMETRIC : RO_RED_BLACK);
    }
    for (int level = stop - 1; level >= 0; level--) {
      prolong(level + 1, level);
      if (level > 0) relax(level, nu2, make_symmetric_operator ? RO_BLACK_RED : RO_RED_BLACK);
    }
  }
  return false;
}

// sol_mgpressure3d.cpp:332-370: on failure retry once from a zero guess
bool Sol_MultigridPressure3DBase::solve(double &residual, double tolerance, int max_iter) {
  if (!check_float(tolerance)) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::solve - garbage tolerance value %f\n", tolerance);
    return false;
  }
  if (tolerance < 0) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::solve - negative tolerance value %f\n", tolerance);
    return false;
  }
  if (max_iter <= 0) {
    printf("[WARNING] Sol_MultigridPressure3DDevice::solve - non-positive max_iter %d\n", max_iter);
    return false;
  }
  if (_num_levels < 1) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::solve - solver not initialized\n");
    return false;
  }
  num_vcycles = 0;
  residual_history.clear();
  double l2 = 0, linf = 0;
  if (!do_fmg(tolerance, max_iter, l2, linf)) {
    printf("[WARNING] Sol_MultigridPressure3DBase::solve - do_fmg did not converge, retrying with zeroed initial "
           "search vector\n");
    clear_zero(0);
    if (!do_fmg(tolerance, max_iter, l2, linf)) {
      printf("[WARNING] Sol_MultigridPressure3DBase::solve - do_fmg failed\n");
      residual = (convergence == CONVERGENCE_L2) ? l2 : linf;
      settle();
      return false;
    }
  }
  residual = (convergence == CONVERGENCE_L2) ? l2 : linf;
  settle();
  return true;
}

bool Sol_MultigridPressure3DBase::run_vcycles(int count) {
  if (count <= 0) {
    printf("[WARNING] Sol_MultigridPressure3DDevice::run_vcycles - non-positive count %d\n", count);
    return false;
  }
  num_vcycles = 0;
  residual_history.clear();
  double l2, linf;
  do_vcycle(0, count, l2, linf);
  settle();
  return !any_error();
}

bool Sol_MultigridPressure3DBase::run_fmg(int count) {
  if (count < 0) {
    printf("[WARNING] Sol_MultigridPressure3DDevice::run_fmg - negative count %d\n", count);
    return false;
  }
  num_vcycles = 0;
  residual_history.clear();
  double l2, linf;
  do_fmg(0, count, l2, linf);
  settle();
  return !any_error();
}

// ------------------------------------------------------------------------
// Device operations
// ------------------------------------------------------------------------

template <typename T>
Sol_MultigridPressure3DDevice<T>::Sol_MultigridPressure3DDevice() {
  _r_grid = _u_grid = _b_grid = _u_alt = 0;
  _update_bc_between_colors = false;
  _fused_ok = false;
  _bottom_ok = false;
  _coarse_mode = 0;
  _norm_ws = 0;
  _norm_host = 0;
  _norm_dev = 0;
  disable_relax_on_host = false;
  use_fused_relax = true;
  store_residuals = false;
  use_bottom = true;
  _bottom_level = -1;
  use_fused_prolong = true;
  use_graphs = OCU_MG_GRAPHS_DEFAULT;
  graph_records = graph_replays = 0;
  _graphs_failed = false;
  _norms_pending = false;
  _norm_tag = 0;
  _capturing = false;
  level0_sweeps[0] = level0_sweeps[1] = level0_sweeps[2] = 0;
  for (int l = 0; l < kMaxLevels; l++) _u_zero[l] = _u_pro[l] = _fused_lvl[l] = false;
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::free_storage() {
  drop_segments();
  if (_u_grid) {
    for (int l = 0; l < _num_levels; l++) {
      if (l > 0) {
        delete _u_grid[l];
        delete _b_grid[l];
      }
      delete _r_grid[l];
      delete _u_alt[l];
    }
  }
  delete[] _u_grid;
  delete[] _b_grid;
  delete[] _r_grid;
  delete[] _u_alt;
  _u_grid = _b_grid = _r_grid = _u_alt = 0;
  ocu_free(_norm_ws);
  ocu_host_free(_norm_host);
  _norm_dev = 0;
  _norm_ws = 0;
  _norm_host = 0;
}

template <typename T>
Sol_MultigridPressure3DDevice<T>::~Sol_MultigridPressure3DDevice() {
  free_storage();
}

// sol_mgpressure3ddev.cu:52-59
template <typename T>
BoundaryConditionSet Sol_MultigridPressure3DDevice<T>::get_bc_at_level(int level) const {
  BoundaryConditionSet bc_zero = bc;
  if (level != 0) bc_zero.make_homogeneous();
  return bc_zero;
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::level_params(int level, void *out) const {
  ocu_mg_level_params &p = *(ocu_mg_level_params *)out;
  BoundaryConditionSet lbc = get_bc_at_level(level);
  p.h = get_h(level);
  p.fx = _fx; p.fy = _fy; p.fz = _fz;
  p.omega = _omega;
  p.mods[0] = bc_diag_mod(lbc.xpos, _fx);
  p.mods[1] = bc_diag_mod(lbc.xneg, _fx);
  p.mods[2] = bc_diag_mod(lbc.ypos, _fy);
  p.mods[3] = bc_diag_mod(lbc.yneg, _fy);
  p.mods[4] = bc_diag_mod(lbc.zpos, _fz);
  p.mods[5] = bc_diag_mod(lbc.zneg, _fz);
}

// A cleared level may stay unwritten when its next use is the fused
// relaxation, which can start from the zero grid without reading it.
template <typename T>
bool Sol_MultigridPressure3DDevice<T>::defer_ok(int level) const {
  if (!(use_fused_relax && _fused_lvl[level] && level < _num_levels - 1 && nu1 >= 2)) return false;
  // the zero grid's ghosts are all zero only for homogeneous faces
  BoundaryConditionSet lbc = get_bc_at_level(level);
  return lbc.xpos.value == 0 && lbc.xneg.value == 0 && lbc.ypos.value == 0 && lbc.yneg.value == 0 &&
         lbc.zpos.value == 0 && lbc.zneg.value == 0;
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::materialize(int level) {
  if (_u_pro[level]) {
    _u_pro[level] = false;
    run_prolong(level + 1, level);
  }
  if (!_u_zero[level]) return;
  _u_zero[level] = false;
  check_ok(get_u(level).clear_zero());
  apply_boundary_conditions(level);
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::settle() {
  for (int l = 0; l < _num_levels; l++) materialize(l);
}

template <typename T>
int Sol_MultigridPressure3DDevice<T>::bottom_start() const {
  if (!use_bottom || _bottom_level < 1 || !use_fused_relax || !_bottom_ok || store_residuals ||
      disable_relax_on_host || make_symmetric_operator)
    return -1;
  return _bottom_level;
}

// sol_mgpressure3d.cpp:184-239 from level `top` down to the coarsest and back,
// in one launch
template <typename T>
void Sol_MultigridPressure3DDevice<T>::bottom_vcycle(int top) {
  const int C = _num_levels - 1, nlev = C - top + 1;
  ocu_grid3d u[kMaxLevels], b[kMaxLevels];
  ocu_mg_level_params p[kMaxLevels];
  double h[3 * kMaxLevels];
  for (int l = 0; l < nlev; l++) {
    u[l] = get_u(top + l).desc();
    b[l] = get_b(top + l).desc();
    level_params(top + l, &p[l]);
    h[3 * l] = hx(top + l);
    h[3 * l + 1] = hy(top + l);
    h[3 * l + 2] = hz(top + l);
  }
  for (int l = top; l <= C; l++) {
    if (_u_pro[l]) {  // (a deferred zero stays deferred: the bottom takes it as a flag)
      _u_pro[l] = false;
      run_prolong(l + 1, l);
    }
    flush_reader(l);
  }
  ocu_bc_set cbc = to_c_bc(get_bc_at_level(top));
  int coarse_iters = max3(nx(C) * ny(C), ny(C) * nz(C), nx(C) * nz(C)) / 2;
  ok_status(ocu_mg_bottom_vcycle(u, b, p, h, nlev, &cbc, nu1, nu2, coarse_iters, _coarse_mode, _u_zero[top] ? 1 : 0,
                                 ThreadManager::get_compute_stream()),
            "Sol_MultigridPressure3DDevice::bottom_vcycle");
  for (int l = top; l <= C; l++) _u_zero[l] = false;
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::clear_zero(int level) {
  flush_reader(level);
  _u_pro[level] = false;  // the grid becomes zero: the pending correction is moot
  if (defer_ok(level)) {
    _u_zero[level] = true;
    return;
  }
  _u_zero[level] = false;
  check_ok(get_u(level).clear_zero());
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::apply_boundary_conditions(int level) {
  if (_u_zero[level]) return;  // the deferred zero grid carries its ghosts
  if (_u_pro[level]) return;   // the pending prolongation writes them
  flush_reader(level);
  if (!apply_3d_boundary_conditions_level1(get_u(level), get_bc_at_level(level), hx(level), hy(level), hz(level))) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::apply_boundary_conditions - failed at level %d \n", level);
    add_error();
  }
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::relax_colour(int level, int color) {
  materialize(level);
  flush_reader(level);
  ocu_mg_level_params p;
  level_params(level, &p);
  ocu_grid3d u = get_u(level).desc(), b = get_b(level).desc();
  ok_status(ocu_mg_relax(&u, &b, color, &p, OCU_RELAX_DEVICE, ThreadManager::get_compute_stream()),
            "Sol_MultigridPressure3DDevice::relax");
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::relax_colour_bc(int level, int color) {
  materialize(level);
  flush_reader(level);
  ocu_mg_level_params p;
  level_params(level, &p);
  ocu_bc_set cbc = to_c_bc(get_bc_at_level(level));
  ocu_grid3d u = get_u(level).desc(), b = get_b(level).desc();
  ok_status(ocu_mg_relax_bc(&u, &b, color, &p, &cbc, hx(level), hy(level), hz(level),
                            ThreadManager::get_compute_stream()),
            "Sol_MultigridPressure3DDevice::relax");
}

// the colour kernel can write the fill's ghosts itself when every face is
// Neumann or Dirichlet, or periodic with its opposite face across an even
// number of cells (the conditions of ocu_mg_relax_bc)
template <typename T>
bool Sol_MultigridPressure3DDevice<T>::relax_bc_ok(int level) const {
  const BoundaryConditionSet b = get_bc_at_level(level);
  const BoundaryCondition *lo[3] = {&b.xneg, &b.yneg, &b.zneg}, *hi[3] = {&b.xpos, &b.ypos, &b.zpos};
  const int n[3] = {nx(level), ny(level), nz(level)};
  for (int a = 0; a < 3; a++) {
    const bool pl = lo[a]->type == BC_PERIODIC, ph = hi[a]->type == BC_PERIODIC;
    if (pl != ph) return false;
    if (pl) {
      if (n[a] % 2) return false;
      continue;
    }
    for (const BoundaryCondition *c : {lo[a], hi[a]})
      if (c->type != BC_NEUMANN && c->type != BC_DIRICHELET) return false;
  }
  return true;
}

// relax_on_host (sol_mgpressure3ddev.cu:134-188) on the GPU: one LDS
// workgroup, host arithmetic, no-corner fill after each colour, level-1 fill
// at the end
template <typename T>
void Sol_MultigridPressure3DDevice<T>::relax_coarsest(int level, int iterations, RelaxOrder order) {
  materialize(level);
  flush_reader(level);
  ocu_mg_level_params p;
  level_params(level, &p);
  ocu_grid3d u = get_u(level).desc(), b = get_b(level).desc();
  ocu_bc_set cbc = to_c_bc(get_bc_at_level(level));
  int ord = order == RO_RED_BLACK ? 0 : order == RO_BLACK_RED ? 1 : 2;
  ok_status(ocu_mg_coarse_relax(&u, &b, iterations, ord, &p, &cbc, hx(level), hy(level), hz(level), _coarse_mode,
                                ThreadManager::get_compute_stream()),
            "Sol_MultigridPressure3DDevice::relax_on_host");
}

// relax (sol_mgpressure3ddev.cu:190-248)
template <typename T>
void Sol_MultigridPressure3DDevice<T>::relax(int level, int iterations, RelaxOrder order) {
  if (!disable_relax_on_host && level == _num_levels - 1) {
    relax_coarsest(level, iterations, order);
    return;
  }
  if (order == RO_SYMMETRIC) {
    if (iterations % 2 == 1) iterations++;
    if (iterations < 4) iterations = 4;
  }
  auto first_colour = [&](int it) {
    return (order == RO_BLACK_RED || (order == RO_SYMMETRIC && it >= iterations / 2)) ? 1 : 0;
  };

  flush_reader(level);
  if (iterations <= 0) {
    materialize(level);
    return;
  }
  if (use_fused_relax && _fused_lvl[level]) {
    // pairs of out-of-place sweeps u -> alt -> u, so the result always ends
    // in u; each sweep also writes the level-1 ghosts of its output (the
    // fill the reference runs after every iteration), and the first one
    // starts from the zero grid when u is a deferred clear, or adds a
    // pending prolongation while it reads u.  An odd last iteration runs in
    // place, one launch per colour.
    ocu_mg_level_params p;
    level_params(level, &p);
    ocu_bc_set cbc = to_c_bc(get_bc_at_level(level));
    void *s = ThreadManager::get_compute_stream();
    Grid3DDevice<T> &u = get_u(level);
    Grid3DDevice<T> &alt = *_u_alt[level];
    ocu_grid3d du = u.desc(), da = alt.desc(), db = get_b(level).desc();
    const double lhx = hx(level), lhy = hy(level), lhz = hz(level);
    int it = 0;
    for (; it + 1 < iterations; it += 2) {
      if (level == 0) level0_sweeps[_u_pro[0] ? 2 : _u_zero[0] ? 1 : 0]++;
      if (level == 0) level0_sweeps[0]++;  // the second sweep of the pair
      if (_u_pro[level]) {
        _u_pro[level] = false;
        ocu_grid3d dc = get_u(level + 1).desc();
        ok_status(ocu_mg_relax_fused_prolong_bc(&du, &da, &db, &dc, 1, &p, &cbc, lhx, lhy, lhz, first_colour(it),
                                                OCU_FUSED_WRITE_GHOSTS, s),
                  "Sol_MultigridPressure3DDevice::relax");
      } else {
        const int zero = _u_zero[level] ? OCU_FUSED_ZERO_INPUT : 0;
        _u_zero[level] = false;
        ok_status(ocu_mg_relax_fused_bc(&du, &da, &db, &p, &cbc, lhx, lhy, lhz, first_colour(it),
                                        OCU_FUSED_WRITE_GHOSTS | zero, s),
                  "Sol_MultigridPressure3DDevice::relax");
      }
      ok_status(ocu_mg_relax_fused_bc(&da, &du, &db, &p, &cbc, lhx, lhy, lhz, first_colour(it + 1),
                                      OCU_FUSED_WRITE_GHOSTS, s),
                "Sol_MultigridPressure3DDevice::relax");
    }
    for (; it < iterations; it++) {
      int c0 = first_colour(it);
      relax_colour_bc(level, c0);
      relax_colour_bc(level, !c0);
    }
  } else if (relax_bc_ok(level)) {
    // each colour launch writes the ghosts of the cells it updates, which is
    // what the reference's fill between colours (periodic faces) or after
    // each iteration computes; the ghosts were filled before the call
    for (int it = 0; it < iterations; it++)
      for (int rb = 0; rb < 2; rb++) relax_colour_bc(level, (first_colour(it) == 1) ? !rb : rb);
  } else {
    materialize(level);
    for (int it = 0; it < iterations; it++) {
      for (int rb = 0; rb < 2; rb++) {
        int color = (first_colour(it) == 1) ? !rb : rb;
        relax_colour(level, color);
        if (_update_bc_between_colors) apply_boundary_conditions(level);
      }
      if (!_update_bc_between_colors) apply_boundary_conditions(level);
    }
  }
  if (any_error()) printf("[ERROR] Sol_MultigridPressure3DDevice::relax - failed at level %d\n", level);
}

// restrict_residuals (sol_mgpressure3ddev.cu:250-279): residual and
// restriction in one launch; level-0 norms reduced on the device
template <typename T>
void Sol_MultigridPressure3DDevice<T>::restrict_residuals(int fine_level, int coarse_level, double *l2,
                                                          double *linf) {
  restrict_residuals_launch(fine_level, coarse_level, l2 || linf);
  if (l2 || linf) norms_read(fine_level, l2, linf);
  if (any_error())
    printf("[ERROR] Sol_MultigridPressure3DDevice::restrict_residuals - failed at level %d -> %d\n", fine_level,
           coarse_level);
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::restrict_residuals_launch(int fine_level, int coarse_level, bool norms) {
  materialize(fine_level);
  void *s = ThreadManager::get_compute_stream();
  ocu_grid3d u = get_u(fine_level).desc(), b = get_b(fine_level).desc(), r = get_r(fine_level).desc();
  double h = get_h(fine_level);
  if (coarse_level != fine_level) {
    ocu_grid3d bc = get_b(coarse_level).desc();
    ok_status(ocu_mg_residual_restrict(&u, &b, store_residuals ? &r : nullptr, &bc, h, _fx, _fy, _fz,
                                       norms ? _norm_ws : nullptr, s),
              "Sol_MultigridPressure3DDevice::restrict_residuals");
  } else {
    ok_status(ocu_mg_residual_norm(&u, &b, &r, h, _fx, _fy, _fz, norms ? _norm_ws : nullptr, s),
              "Sol_MultigridPressure3DDevice::restrict_residuals");
  }
  if (norms) {
    // the finalize kernel writes the two results straight into pinned host
    // memory and a completion tag after them: no copy launch, and the host
    // polls for the tag rather than synchronising the stream
    _norm_tag = _capturing ? 0 : next_completion_tag();
    if (_norm_tag)
      ok_status(ocu_norm_finalize_tagged(_norm_ws, _norm_dev, _norm_tag, s),
                "Sol_MultigridPressure3DDevice::restrict_residuals");
    else
      ok_status(ocu_norm_finalize(_norm_ws, _norm_dev, s), "Sol_MultigridPressure3DDevice::restrict_residuals");
    _norms_pending = true;
  }
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::norms_read(int fine_level, double *l2, double *linf) {
  if (!_norms_pending) return;
  _norms_pending = false;
  void *s = ThreadManager::get_compute_stream();
  if (_norm_tag) {
    if (!wait_completion_tag(_norm_host + 2, _norm_tag, s)) {
      printf("[ERROR] Sol_MultigridPressure3DDevice::restrict_residuals - norms did not arrive\n");
      add_error();
    }
  } else {
    ok_status(ocu_stream_synchronize(s), "Sol_MultigridPressure3DDevice::restrict_residuals");
  }
  double n = (double)nx(fine_level) * ny(fine_level) * nz(fine_level);
  if (l2) *l2 = sqrt(_norm_host[0] / n);
  if (linf) *linf = _norm_host[1];
}

// what a segment's launches depend on besides the grids (fixed from
// initialize_storage on): options, faces, the stream, and the deferred
// clears and prolongations of every level
template <typename T>
std::vector<long long> Sol_MultigridPressure3DDevice<T>::segment_key(void *stream) const {
  std::vector<long long> k = {nu1, nu2, use_fused_relax, use_bottom, use_fused_prolong, store_residuals,
                              disable_relax_on_host, make_symmetric_operator, (long long)convergence, _coarse_mode,
                              (long long)(intptr_t)stream, _num_levels};
  const BoundaryCondition *f[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (const BoundaryCondition *c : f) {
    long long v, a;
    memcpy(&v, &c->value, sizeof(v));
    memcpy(&a, &c->aux_value, sizeof(a));
    k.push_back(c->type);
    k.push_back(v);
    k.push_back(a);
  }
  for (int l = 0; l < _num_levels; l++) k.push_back(_u_zero[l] | _u_pro[l] << 1);
  return k;
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::drop_segments() {
  for (Segment &g : _segments) ocu_graph_destroy(g.exec);
  _segments.clear();
}

// The segment's launches go to a HIP graph the first time the solver is in
// a given state at the start of a segment; later segments from that state
// launch the graph and take the flags the recording left.  A failed
// recording restores the state and runs the launches directly.
template <typename T>
void Sol_MultigridPressure3DDevice<T>::run_segment(const std::function<void()> &body) {
  void *s = ThreadManager::get_compute_stream();
  if (!use_graphs || _graphs_failed || !s) {
    body();
    return;
  }
  const std::vector<long long> key = segment_key(s);
  for (Segment &g : _segments) {
    if (g.key != key) continue;
    if (!ok_status(ocu_graph_launch(g.exec, s), "Sol_MultigridPressure3DDevice::run_segment")) return;
    for (int l = 0; l < _num_levels; l++) {
      _u_zero[l] = g.out[l] & 1;
      _u_pro[l] = (g.out[l] >> 1) & 1;
    }
    for (int q = 0; q < 3; q++) level0_sweeps[q] += g.sweeps[q];
    _norms_pending = true;
    _norm_tag = 0;
    graph_replays++;
    return;
  }
  // the state to return to if the recording fails
  std::vector<char> z(_num_levels), p(_num_levels);
  for (int l = 0; l < _num_levels; l++) z[l] = _u_zero[l], p[l] = _u_pro[l];
  long long sw[3] = {level0_sweeps[0], level0_sweeps[1], level0_sweeps[2]};
  const bool pending = _norms_pending, err = any_error();
  const unsigned long long tag = _norm_tag;
  void *exec = nullptr;
  bool ok = ocu_stream_begin_capture(s) == OCU_OK;
  if (ok) {
    _capturing = true;
    body();
    _capturing = false;
    ok = ocu_stream_end_capture(s, &exec) == OCU_OK && exec && (err || !any_error());
  }
  if (!ok) {
    if (exec) ocu_graph_destroy(exec);
    for (int l = 0; l < _num_levels; l++) _u_zero[l] = z[l], _u_pro[l] = p[l];
    for (int q = 0; q < 3; q++) level0_sweeps[q] = sw[q];
    _norms_pending = pending;
    _norm_tag = tag;
    if (!err) clear_error();
    _graphs_failed = true;
    body();
    return;
  }
  Segment g;
  g.key = key;
  g.out.resize(_num_levels);
  for (int l = 0; l < _num_levels; l++) g.out[l] = (char)(_u_zero[l] | _u_pro[l] << 1);
  for (int q = 0; q < 3; q++) g.sweeps[q] = level0_sweeps[q] - sw[q];
  g.exec = exec;
  if (_segments.size() >= 4) {
    ocu_graph_destroy(_segments.front().exec);
    _segments.erase(_segments.begin());
  }
  _segments.push_back(g);
  graph_records++;
  ok_status(ocu_graph_launch(exec, s), "Sol_MultigridPressure3DDevice::run_segment");
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::prolong(int coarse_level, int fine_level) {
  materialize(coarse_level);
  materialize(fine_level);
  flush_reader(fine_level);
  ocu_grid3d uc = get_u(coarse_level).desc(), uf = get_u(fine_level).desc();
  if (use_fused_prolong && use_fused_relax && _fused_ok && coarse_level == fine_level + 1) {
    ocu_bc_set cbc = to_c_bc(get_bc_at_level(fine_level));
    if (ocu_mg_relax_fused_prolong_ok(&uf, &uc, &cbc)) {
      _u_pro[fine_level] = true;  // the next fused relaxation of the fine level adds it
      return;
    }
  }
  run_prolong(coarse_level, fine_level);
}

template <typename T>
void Sol_MultigridPressure3DDevice<T>::run_prolong(int coarse_level, int fine_level) {
  ocu_grid3d uc = get_u(coarse_level).desc(), uf = get_u(fine_level).desc();
  void *s = ThreadManager::get_compute_stream();
  ocu_bc_set fbc = to_c_bc(get_bc_at_level(fine_level));
  if (use_fused_relax && ocu_mg_prolong_bc_ok(&uc, &uf, &fbc)) {
    // prolongation and the level-1 fill of the fine grid in one launch
    ok_status(ocu_mg_prolong_bc(&uc, &uf, 1, &fbc, hx(fine_level), hy(fine_level), hz(fine_level), s),
              "Sol_MultigridPressure3DDevice::prolong");
  } else {
    ok_status(ocu_mg_prolong(&uc, &uf, 1, s), "Sol_MultigridPressure3DDevice::prolong");
    apply_boundary_conditions(fine_level);
  }
  if (any_error())
    printf("[ERROR] Sol_MultigridPressure3DDevice::prolong - failed at level %d -> %d\n", coarse_level, fine_level);
}

// sol_mgpressure3ddev.cu:376-482
template <typename T>
bool Sol_MultigridPressure3DDevice<T>::initialize_storage(int nx_val, int ny_val, int nz_val, double hx_val,
                                                          double hy_val, double hz_val, Grid3DDevice<T> *rhs) {
  free_storage();
  if (!initialize_base_storage(nx_val, ny_val, nz_val, hx_val, hy_val, hz_val)) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - error in grid calculation\n");
    return false;
  }
  if (_num_levels < 1 || _num_levels > kMaxLevels) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - (%d,%d,%d) admits no multigrid level\n",
           nx_val, ny_val, nz_val);
    return false;
  }
  for (int l = 0; l < kMaxLevels; l++) _u_zero[l] = _u_pro[l] = false;
  _update_bc_between_colors = bc.xpos.check_type(BC_PERIODIC) || bc.xneg.check_type(BC_PERIODIC) ||
                              bc.ypos.check_type(BC_PERIODIC) || bc.yneg.check_type(BC_PERIODIC) ||
                              bc.zpos.check_type(BC_PERIODIC) || bc.zneg.check_type(BC_PERIODIC);
  // out-of-place sweeps need every ghost to be a function of the interior
  // (Neumann/Dirichlet): periodic ghosts change between colours and
  // BC_NONE ghosts are never written, so those keep the in-place path
  _fused_ok = bc.check_type(BC_NEUMANN, BC_DIRICHELET);
  if (!rhs) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - null rhs\n");
    return false;
  }
  int b_gx = rhs->gx(), b_gy = rhs->gy(), b_gz = rhs->gz();
  if (b_gx < 1 || b_gy < 1 || b_gz < 1) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - rhs has invalid ghost cells (%d,%d,%d), must "
           "be >= 1\n",
           b_gx, b_gy, b_gz);
    return false;
  }
  // the reference's check (:398) joins the three tests with commas and so
  // only tests nz; all three are tested here
  if (rhs->nx() != nx_val || rhs->ny() != ny_val || rhs->nz() != nz_val) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - rhs dimension mismatch (%d,%d,%d) != "
           "(%d,%d,%d)\n",
           rhs->nx(), rhs->ny(), rhs->nz(), nx_val, ny_val, nz_val);
    return false;
  }
  _r_grid = new Grid3DDevice<T> *[_num_levels];
  _u_grid = new Grid3DDevice<T> *[_num_levels];
  _b_grid = new Grid3DDevice<T> *[_num_levels];
  _u_alt = new Grid3DDevice<T> *[_num_levels];
  memset(_r_grid, 0, sizeof(void *) * _num_levels);
  memset(_u_grid, 0, sizeof(void *) * _num_levels);
  memset(_b_grid, 0, sizeof(void *) * _num_levels);
  memset(_u_alt, 0, sizeof(void *) * _num_levels);

  int px = rhs->paddingx(), py = rhs->paddingy(), pz = rhs->paddingz();
  _b_grid[0] = rhs;
  _u_grid[0] = &_pressure;
  _r_grid[0] = new Grid3DDevice<T>();
  _u_alt[0] = new Grid3DDevice<T>();
  if (!_r_grid[0]->init(nx_val, ny_val, nz_val, b_gx, b_gy, b_gz, px, py, pz) ||
      !_u_grid[0]->init(nx_val, ny_val, nz_val, b_gx, b_gy, b_gz, px, py, pz) ||
      !_u_alt[0]->init(nx_val, ny_val, nz_val, b_gx, b_gy, b_gz, px, py, pz)) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - could not initialize level 0\n");
    return false;
  }
  _u_grid[0]->clear_zero();
  _u_alt[0]->clear_zero();
  _r_grid[0]->clear_zero();

  for (int level = 1; level < _num_levels; level++) {
    _u_grid[level] = new Grid3DDevice<T>();
    _b_grid[level] = new Grid3DDevice<T>();
    _r_grid[level] = new Grid3DDevice<T>();
    _u_alt[level] = new Grid3DDevice<T>();
    if (!_u_grid[level]->init(nx(level), ny(level), nz(level), 1, 1, 1) ||
        !_b_grid[level]->init(nx(level), ny(level), nz(level), 1, 1, 1) ||
        !_r_grid[level]->init(nx(level), ny(level), nz(level), 1, 1, 1) ||
        !_u_alt[level]->init(nx(level), ny(level), nz(level), 1, 1, 1)) {
      printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - could not initialize level %d\n", level);
      return false;
    }
    _u_grid[level]->clear_zero();
    _b_grid[level]->clear_zero();
    _r_grid[level]->clear_zero();
    _u_alt[level]->clear_zero();
  }

  void *ws = nullptr, *hh = nullptr;
  if (ocu_malloc(&ws, sizeof(double) * (2 * ocu_norm_parts() + 8)) != OCU_OK ||
      ocu_host_malloc(&hh, sizeof(double) * 8) != OCU_OK) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - workspace allocation failed\n");
    return false;
  }
  _norm_ws = (double *)ws;
  _norm_host = (double *)hh;
  void *hd = nullptr;
  if (ocu_host_device_pointer(&hd, hh) != OCU_OK) {
    printf("[ERROR] Sol_MultigridPressure3DDevice::initialize_storage - pinned results not device-visible\n");
    return false;
  }
  _norm_dev = (double *)hd;

  // levels whose relaxation runs as fused sweeps (periodic pairs included)
  for (int l = 0; l < _num_levels; l++) {
    ocu_bc_set cbc = to_c_bc(get_bc_at_level(l));
    ocu_grid3d du = get_u(l).desc(), da = _u_alt[l]->desc(), db = get_b(l).desc();
    _fused_lvl[l] = _fused_ok || ocu_mg_relax_fused_ok(&du, &da, &db, &cbc) == 1;
  }

  // the smallest level (>= 1) from which every coarser level fits one
  // workgroup's LDS together
  _bottom_level = -1;
  const int C = _num_levels - 1;
  for (int top = C - 1; top >= 1; top--) {
    ocu_grid3d d[kMaxLevels];
    for (int l = top; l <= C; l++) d[l - top] = get_u(l).desc();
    size_t bytes = ocu_mg_bottom_lds_bytes(d, C - top + 1, _coarse_mode);
    if (bytes == 0 || bytes > ocu_mg_bottom_lds_max()) break;
    _bottom_level = top;
  }
  _bottom_ok = _fused_ok;
  if (!_fused_ok && _bottom_level >= 1 && relax_bc_ok(0)) {
    // periodic pairs: every bottom level needs an even count along them
    _bottom_ok = true;
    const BoundaryCondition *hi[3] = {&bc.xpos, &bc.ypos, &bc.zpos};
    for (int l = _bottom_level; l <= C; l++) {
      const int n[3] = {nx(l), ny(l), nz(l)};
      for (int a = 0; a < 3; a++)
        if (hi[a]->type == BC_PERIODIC && n[a] % 2) _bottom_ok = false;
    }
  }
  return !any_error();
}

template class Sol_MultigridPressure3DDevice<float>;
template class Sol_MultigridPressure3DDevice<double>;

}  // namespace ocu
