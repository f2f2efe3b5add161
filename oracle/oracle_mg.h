/*
 * ORACLE - test infrastructure only (see oracle_grid.h).
 *
 * CPU restatement of Sol_MultigridPressure3D: level hierarchy, red-black
 * Gauss-Seidel relaxation in its three reference operation orders, residual,
 * restriction, prolongation (with the reference's corner quirk), the
 * coarsest-level host relaxation and the FMG / V-cycle driver; plus the
 * mixed-precision iterative-refinement solver, which has no reference
 * counterpart and is restated here from its own definition.
 *
 * Loops over i are OpenMP-parallel where the result does not depend on the
 * schedule (one colour of red-black GS, residual, restrict, prolong); sums
 * are accumulated per i-plane and then serially, so results are identical
 * for any thread count.
 */
#ifndef OCU_ORACLE_MG_H
#define OCU_ORACLE_MG_H

#include <array>
#include <cstdio>
#include <memory>

#include "oracle_grid.h"

namespace orc {

/* ------------------------------------------------------------------ */
/* Level set-up (sol_mgpressure3d.cpp)                                   */
/* ------------------------------------------------------------------ */

/* num_levels, sol_mgpressure3d.cpp:67-92 */
inline int num_levels(int nx, int ny, int nz) {
  int cnx = nx / 2, cny = ny / 2, cnz = nz / 2;
  int levels = 0;
  while (2 * cnx == nx && 2 * cny == ny && 2 * cnz == nz) {
    levels++;
    nx = cnx; ny = cny; nz = cnz;
    cnx /= 2; cny /= 2; cnz /= 2;
    if (cnx < 2 || nx == 5 || nx == 7) break;
    if (cny < 2 || ny == 5 || ny == 7) break;
    if (cnz < 2 || nz == 5 || nz == 7) break;
  }
  return levels;
}

inline double min3(double a, double b, double c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }

/* optimal_omega (Yavneh 1996), sol_mgpressure3d.cpp:94-109.  Note the
 * reference's `sqrt(1.0f - c_max_sqr)`: the literal is float but the
 * expression is double. */
inline double optimal_omega(double fx, double fy, double fz) {
  double f_min = min3(fx, fy, fz);
  double c_max = (1.0 - f_min / (fx + fy + fz));
  double c_max_sqr = c_max * c_max;
  return 2.0 / (1.0 + std::sqrt(1.0f - c_max_sqr));
}

/* bc_diag_mod, sol_mgpressure3d.cpp:49-64 */
inline double bc_diag_mod(const ocu_bc &b, double factor) {
  if (b.type == OCU_BC_PERIODIC) return 0;
  if (b.type == OCU_BC_NONE) return 0;
  if (b.type == OCU_BC_DIRICHELET) return factor;
  if (b.type == OCU_BC_NEUMANN) return -factor;
  return 0;
}

/* get_bc_at_level, sol_mgpressure3ddev.cu:52-59: coarse levels solve the
 * homogeneous problem. */
inline ocu_bc_set bc_at_level(const ocu_bc_set &bc, int level) {
  ocu_bc_set r = bc;
  if (level != 0) {
    ocu_bc *all[6] = {&r.xpos, &r.xneg, &r.ypos, &r.yneg, &r.zpos, &r.zneg};
    for (auto b : all) { b->value = 0; b->aux_value = 0; }
  }
  return r;
}

inline ocu_mg_level_params level_params(double h, double fx, double fy, double fz,
                                        double omega, const ocu_bc_set &lbc) {
  ocu_mg_level_params p;
  p.h = h; p.fx = fx; p.fy = fy; p.fz = fz; p.omega = omega;
  p.mods[0] = bc_diag_mod(lbc.xpos, fx);
  p.mods[1] = bc_diag_mod(lbc.xneg, fx);
  p.mods[2] = bc_diag_mod(lbc.ypos, fy);
  p.mods[3] = bc_diag_mod(lbc.yneg, fy);
  p.mods[4] = bc_diag_mod(lbc.zpos, fz);
  p.mods[5] = bc_diag_mod(lbc.zneg, fz);
  return p;
}

/* ------------------------------------------------------------------ */
/* One colour of red-black Gauss-Seidel                                 */
/* ------------------------------------------------------------------ */

/* Device FP64, Sol_MultigridPressure3DDeviceD_relax_DIAG
 * (sol_mgpressure3ddevd.cu:542-599) with the per-(i,j) diagonal table of
 * Sol_MultigridPressure3DDevice_initialize_diag_grids
 * (sol_mgpressure3ddev.cu:27-43, filled at :346-373), launched over nz/2
 * threads in k (:1246-1268). */
inline void relax_color_device(const ocu_grid3d &ug, const ocu_grid3d &bg, int color,
                               const ocu_mg_level_params &P, double * /*tag*/) {
  View<double> u(ug), b(bg);
  const int nx = ug.nx, ny = ug.ny, nz = ug.nz;
  const double ld = (double)(2 * P.fx + 2 * P.fy + 2 * P.fz);
  const double neg_hsq = (double)-P.h * P.h;
  const double fx = P.fx, fy = P.fy, fz = P.fz, omega = P.omega;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      double d2 = ld;
      if (i == nx - 1) d2 += P.mods[0];
      if (i == 0) d2 += P.mods[1];
      if (j == ny - 1) d2 += P.mods[2];
      if (j == 0) d2 += P.mods[3];
      for (int kk = 0; kk < nz / 2; kk++) {
        int k = 2 * kk + ((i + j + color) % 2);
        if (k >= nz) continue;
        double ucent = u.at(i, j, k);
        double residual = -ld * ucent;
        residual += fz * (u.at(i, j, k - 1) + u.at(i, j, k + 1));
        residual += fy * (u.at(i, j - 1, k) + u.at(i, j + 1, k));
        residual += fx * (u.at(i - 1, j, k) + u.at(i + 1, j, k));
        residual += neg_hsq * b.at(i, j, k);
        double diag = d2;
        if (k == nz - 1) diag += P.mods[4];
        if (k == 0) diag += P.mods[5];
        u.at(i, j, k) = ucent + omega * residual / diag;
      }
    }
}

/* Device FP32, Sol_MultigridPressure3DDeviceF_relax
 * (sol_mgpressure3ddevf.cu:200-252, launched over nz/2 threads at :367-395);
 * every constant is rounded to float by the launcher. */
inline void relax_color_device(const ocu_grid3d &ug, const ocu_grid3d &bg, int color,
                               const ocu_mg_level_params &P, float * /*tag*/) {
  View<float> u(ug), b(bg);
  const int nx = ug.nx, ny = ug.ny, nz = ug.nz;
  const float omega = (float)P.omega, hsq = (float)(P.h * P.h);
  const float fx = (float)P.fx, fy = (float)P.fy, fz = (float)P.fz;
  const float ld = (float)(2 * P.fx + 2 * P.fy + 2 * P.fz);
  float m[6];
  for (int q = 0; q < 6; q++) m[q] = (float)P.mods[q];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int kk = 0; kk < nz / 2; kk++) {
        int k = 2 * kk + ((i + j + color) % 2);
        if (k >= nz) continue;
        float bval = b.at(i, j, k);
        float ucent = u.at(i, j, k);
        float residual = (-hsq * bval - ld * ucent);
        residual += fz * (u.at(i, j, k - 1) + u.at(i, j, k + 1));
        residual += fy * (u.at(i, j - 1, k) + u.at(i, j + 1, k));
        residual += fx * (u.at(i - 1, j, k) + u.at(i + 1, j, k));
        float diag = ld;
        if (i == nx - 1) diag += m[0];
        if (i == 0) diag += m[1];
        if (j == ny - 1) diag += m[2];
        if (j == 0) diag += m[3];
        if (k == nz - 1) diag += m[4];
        if (k == 0) diag += m[5];
        u.at(i, j, k) = ucent + omega * residual / diag;
      }
}

/* Host relaxation, Sol_MultigridPressure3DDevice<T>::do_cpu_solve
 * (sol_mgpressure3ddev.cu:82-125).  laplace_diag, h and the mods are T;
 * _fx/_fy/_fz and the three neighbour products are double, so for T=float
 * each `residual += _f * (a+b)` is a double add rounded back to float. */
template <typename T>
void relax_color_host(const ocu_grid3d &ug, const ocu_grid3d &bg, int color,
                      const ocu_mg_level_params &P) {
  View<T> u(ug), b(bg);
  const int nx = ug.nx, ny = ug.ny, nz = ug.nz;
  const T laplace_diag = 2 * P.fx + 2 * P.fy + 2 * P.fz;
  const T h = (T)P.h;
  const T xpos = (T)P.mods[0], xneg = (T)P.mods[1], ypos = (T)P.mods[2],
          yneg = (T)P.mods[3], zpos = (T)P.mods[4], zneg = (T)P.mods[5];
  const double fx = P.fx, fy = P.fy, fz = P.fz;
  const T omega = (T)P.omega;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++) {
    T x_mod = (i == 0) ? xneg : (i == nx - 1) ? xpos : 0;
    for (int j = 0; j < ny; j++) {
      T y_mod = (j == 0) ? yneg : (j == ny - 1) ? ypos : 0;
      int k_start = (i + j + color) % 2;
      for (int k = k_start; k < nz; k += 2) {
        T z_mod = (k == 0) ? zneg : (k == nz - 1) ? zpos : 0;
        T residual = (-h * h * b.at(i, j, k) - laplace_diag * u.at(i, j, k));
        residual += fz * (u.at(i, j, k - 1) + u.at(i, j, k + 1));
        residual += fy * (u.at(i, j + 1, k) + u.at(i, j - 1, k));
        residual += fx * (u.at(i + 1, j, k) + u.at(i - 1, j, k));
        u.at(i, j, k) += omega * residual / (laplace_diag + x_mod + y_mod + z_mod);
      }
    }
  }
}

template <typename T>
void relax_color(const ocu_grid3d &u, const ocu_grid3d &b, int color,
                 const ocu_mg_level_params &P, int variant) {
  if (variant == OCU_RELAX_HOST) relax_color_host<T>(u, b, color, P);
  else relax_color_device(u, b, color, P, (T *)nullptr);
}

/* ------------------------------------------------------------------ */
/* residual / restrict / prolong                                         */
/* ------------------------------------------------------------------ */

/* calculate_residual: kernels sol_mgpressure3ddevd.cu:94-114 and
 * sol_mgpressure3ddevf.cu:68-90, constants from the launcher
 * (devd.cu:1396-1399, devf.cu:416-419; rounded to float for T=float). */
template <typename T>
void residual(const ocu_grid3d &ug, const ocu_grid3d &bg, const ocu_grid3d &rg,
              double h, double fxd, double fyd, double fzd) {
  View<T> u(ug), b(bg), r(rg);
  const T fx = (T)(fxd / (h * h)), fy = (T)(fyd / (h * h)), fz = (T)(fzd / (h * h));
  const T diag = (T)(2 * (fxd + fyd + fzd) / (h * h));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rg.nx; i++)
    for (int j = 0; j < rg.ny; j++)
      for (int k = 0; k < rg.nz; k++)
        r.at(i, j, k) = b.at(i, j, k) -
                        (fz * (u.at(i, j, k - 1) + u.at(i, j, k + 1)) +
                         fy * (u.at(i, j - 1, k) + u.at(i, j + 1, k)) +
                         fx * (u.at(i - 1, j, k) + u.at(i + 1, j, k)) - diag * u.at(i, j, k));
}

/* restrict: sol_mgpressure3ddevd.cu:65-91, devf.cu:38-65.  `.125 * (sum)`
 * with a double literal: the sum is in T, the scaling in double. */
template <typename T>
void restrict_(const ocu_grid3d &rg, const ocu_grid3d &bcg) {
  View<T> r(rg), bc(bcg);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < bcg.nx; i++)
    for (int j = 0; j < bcg.ny; j++)
      for (int k = 0; k < bcg.nz; k++) {
        int I = 2 * i, J = 2 * j, K = 2 * k;
        T s = r.at(I, J, K) + r.at(I, J, K + 1) + r.at(I, J + 1, K) + r.at(I, J + 1, K + 1) +
              r.at(I + 1, J, K) + r.at(I + 1, J, K + 1) + r.at(I + 1, J + 1, K) +
              r.at(I + 1, J + 1, K + 1);
        bc.at(i, j, k) = (T)(.125 * s);
      }
}

template <typename T> struct ProlongW;
template <> struct ProlongW<double> {
  static constexpr double c0 = 0.421875, c1 = 0.140625, c2 = 0.046875, c3 = 0.015625;
};
template <> struct ProlongW<float> {
  static constexpr float c0 = 0.421875f, c1 = 0.140625f, c2 = 0.046875f, c3 = 0.015625f;
};

/* prolong + correct: sol_mgpressure3ddevd.cu:118-223, devf.cu:94-197.
 * Child (2iC+dx, 2jC+dy, 2kC+dz) takes its neighbours toward s = 2d-1 on
 * each axis.  With quirk=1 the child (dx,dy,dz)=(0,1,1) uses the +x corner
 * (devd.cu:181-185, devf.cu:155-159) like the reference. */
template <typename T>
void prolong(const ocu_grid3d &ucg, const ocu_grid3d &ufg, int quirk) {
  View<T> uc(ucg), uf(ufg);
  typedef ProlongW<T> W;
#pragma omp parallel for schedule(static)
  for (int ic = 0; ic < ucg.nx; ic++)
    for (int jc = 0; jc < ucg.ny; jc++)
      for (int kc = 0; kc < ucg.nz; kc++) {
        const T C = uc.at(ic, jc, kc);
        for (int dx = 0; dx < 2; dx++)
          for (int dy = 0; dy < 2; dy++)
            for (int dz = 0; dz < 2; dz++) {
              int sx = 2 * dx - 1, sy = 2 * dy - 1, sz = 2 * dz - 1;
              int cx = (quirk && dx == 0 && dy == 1 && dz == 1) ? 1 : sx;
              T X = uc.at(ic + sx, jc, kc), Y = uc.at(ic, jc + sy, kc), Z = uc.at(ic, jc, kc + sz);
              T XY = uc.at(ic + sx, jc + sy, kc), XZ = uc.at(ic + sx, jc, kc + sz),
                YZ = uc.at(ic, jc + sy, kc + sz);
              T XYZ = uc.at(ic + cx, jc + sy, kc + sz);
              T &f = uf.at(2 * ic + dx, 2 * jc + dy, 2 * kc + dz);
              f = f + W::c0 * C + W::c1 * (X + Y + Z) + W::c2 * (XY + XZ + YZ) + W::c3 * XYZ;
            }
      }
}

/* interior norms of r: returns sum of squares and max |r| (accumulated in
 * double; deterministic order). */
template <typename T>
void norms(const ocu_grid3d &rg, double &sumsq, double &maxabs) {
  View<T> r(rg);
  std::vector<double> ps(rg.nx), pm(rg.nx);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rg.nx; i++) {
    double s = 0, m = 0;
    for (int j = 0; j < rg.ny; j++)
      for (int k = 0; k < rg.nz; k++) {
        double x = (double)r.at(i, j, k);
        s += x * x;
        m = std::max(m, std::fabs(x));
      }
    ps[i] = s; pm[i] = m;
  }
  sumsq = 0; maxabs = 0;
  for (int i = 0; i < rg.nx; i++) { sumsq += ps[i]; maxabs = std::max(maxabs, pm[i]); }
}

/* ------------------------------------------------------------------ */
/* The FMG / V-cycle driver                                             */
/* ------------------------------------------------------------------ */

/* Sol_MultigridPressure3DBase (sol_mgpressure3d.cpp:26-437) together with
 * the device storage and overrides of Sol_MultigridPressure3DDevice<T>
 * (sol_mgpressure3ddev.cu:52-482). */
template <typename T>
struct MGSolver {
  enum RelaxOrder { RO_RED_BLACK, RO_BLACK_RED, RO_SYMMETRIC };

  // public state, defaults from sol_mgpressure3d.cpp:26-40 / ...dev.cu:310-318
  ocu_bc_set bc;
  int nu1 = 2, nu2 = 2;
  int convergence = OCU_CONVERGENCE_L2;
  bool make_symmetric_operator = false;
  bool disable_relax_on_host = false;
  int prolong_quirk = 1;

  // hierarchy
  int L = 0;
  std::vector<double> h;
  double fx = 1, fy = 1, fz = 1, omega = 1;
  std::vector<std::array<int, 3>> dim;
  bool update_bc_between_colors = false;
  ocu_grid3d u0{}, b0{};  // external level-0 grids
  std::vector<HostGrid<T>> U, B, R;  // index 0 of U and B unused
  bool any_error = false;

  // statistics
  int num_vcycles = 0;
  std::vector<double> history;  // residual after each level-0 pre-smoothing

  MGSolver() { std::memset(&bc, 0, sizeof(bc)); }

  const ocu_grid3d &u(int l) const { return l == 0 ? u0 : U[l].desc; }
  const ocu_grid3d &b(int l) const { return l == 0 ? b0 : B[l].desc; }
  const ocu_grid3d &r(int l) const { return R[l].desc; }
  double get_h(int l) const { return h[l]; }
  double hx(int l) const { return get_h(l) / std::sqrt(fx); }
  double hy(int l) const { return get_h(l) / std::sqrt(fy); }
  double hz(int l) const { return get_h(l) / std::sqrt(fz); }

  bool initialize_base_storage(int nx, int ny, int nz, double hxv, double hyv, double hzv) {
    if (!std::isfinite(hxv) || !std::isfinite(hyv) || !std::isfinite(hzv)) return false;
    L = num_levels(nx, ny, nz);
    h.assign(L, 0);
    dim.assign(L, {0, 0, 0});
    h[0] = min3(hxv, hyv, hzv);
    fx = (h[0] * h[0]) / (hxv * hxv);
    fy = (h[0] * h[0]) / (hyv * hyv);
    fz = (h[0] * h[0]) / (hzv * hzv);
    omega = optimal_omega(fx, fy, fz);
    dim[0] = {nx, ny, nz};
    for (int l = 1; l < L; l++) {
      dim[l] = {dim[l - 1][0] / 2, dim[l - 1][1] / 2, dim[l - 1][2] / 2};
      h[l] = h[l - 1] * 2;
    }
    return true;
  }

  /* sol_mgpressure3ddev.cu:376-482; pressure is the caller's level-0 u,
   * congruent with rhs. */
  bool initialize_storage(int nx, int ny, int nz, double hxv, double hyv, double hzv,
                          const ocu_grid3d &rhs, const ocu_grid3d &pressure) {
    if (!initialize_base_storage(nx, ny, nz, hxv, hyv, hzv)) return false;
    if (L < 1) return false;
    update_bc_between_colors = bc.xpos.type == OCU_BC_PERIODIC || bc.xneg.type == OCU_BC_PERIODIC ||
                               bc.ypos.type == OCU_BC_PERIODIC || bc.yneg.type == OCU_BC_PERIODIC ||
                               bc.zpos.type == OCU_BC_PERIODIC || bc.zneg.type == OCU_BC_PERIODIC;
    if (rhs.gx < 1 || rhs.gy < 1 || rhs.gz < 1) return false;
    if (rhs.nx != nx || rhs.ny != ny || rhs.nz != nz) return false;
    if (!layout_match(rhs, pressure)) return false;
    b0 = rhs;
    u0 = pressure;
    U.resize(L); B.resize(L); R.resize(L);
    R[0].init_congruent(rhs);
    for (int l = 1; l < L; l++) {
      U[l].init(dim[l][0], dim[l][1], dim[l][2], 1, 1, 1);
      B[l].init(dim[l][0], dim[l][1], dim[l][2], 1, 1, 1);
      R[l].init(dim[l][0], dim[l][1], dim[l][2], 1, 1, 1);
    }
    return true;
  }

  ocu_mg_level_params params(int level) const {
    return level_params(get_h(level), fx, fy, fz, omega, bc_at_level(bc, level));
  }

  void clear_zero(int level) {
    const ocu_grid3d &g = u(level);
    std::memset(g.base, 0, sizeof(T) * (size_t)g.allocated);
  }

  void apply_boundary_conditions(int level) {
    if (!bc_fill_level1<T>(u(level), bc_at_level(bc, level), hx(level), hy(level), hz(level)))
      any_error = true;
  }

  /* relax_on_host (sol_mgpressure3ddev.cu:134-188): host BC (nocorners)
   * after every colour, then device BC on write-back. */
  void relax_on_host(int level, int iterations, RelaxOrder order) {
    if (order == RO_SYMMETRIC) {
      if (iterations % 2 == 1) iterations++;
      if (iterations < 4) iterations = 4;
    }
    ocu_mg_level_params P = params(level);
    ocu_bc_set hbc = bc_at_level(bc, level);
    for (int it = 0; it < iterations; it++)
      for (int rb = 0; rb < 2; rb++) {
        int color = rb;
        if (order == RO_BLACK_RED || (order == RO_SYMMETRIC && it >= iterations / 2)) color = !rb;
        relax_color_host<T>(u(level), b(level), color, P);
        bc_fill_nocorners_host<T>(u(level), hbc, hx(level), hy(level), hz(level));
      }
    apply_boundary_conditions(level);
  }

  /* relax (sol_mgpressure3ddev.cu:190-248) */
  void relax(int level, int iterations, RelaxOrder order) {
    if (!disable_relax_on_host && level == L - 1) {
      relax_on_host(level, iterations, order);
      return;
    }
    if (order == RO_SYMMETRIC) {
      if (iterations % 2 == 1) iterations++;
      if (iterations < 4) iterations = 4;
    }
    ocu_mg_level_params P = params(level);
    for (int it = 0; it < iterations; it++) {
      for (int rb = 0; rb < 2; rb++) {
        int color = rb;
        if (order == RO_BLACK_RED || (order == RO_SYMMETRIC && it >= iterations / 2)) color = !rb;
        relax_color<T>(u(level), b(level), color, P, OCU_RELAX_DEVICE);
        if (update_bc_between_colors) apply_boundary_conditions(level);
      }
      if (!update_bc_between_colors) apply_boundary_conditions(level);
    }
  }

  /* restrict_residuals (sol_mgpressure3ddev.cu:250-279).  Norms: RMS and
   * max over the fine interior (sums in double here). */
  void restrict_residuals(int fine, int coarse, double *l2, double *linf) {
    residual<T>(u(fine), b(fine), r(fine), get_h(fine), fx, fy, fz);
    if (coarse != fine) restrict_<T>(r(fine), b(coarse));
    if (l2 || linf) {
      double s, m;
      norms<T>(r(fine), s, m);
      if (l2) *l2 = std::sqrt(s / ((double)dim[fine][0] * dim[fine][1] * dim[fine][2]));
      if (linf) *linf = m;
    }
  }

  void prolong_(int coarse, int fine) {
    prolong<T>(u(coarse), u(fine), prolong_quirk);
    apply_boundary_conditions(fine);
  }

  double crit(double l2, double linf) const {
    return (convergence & OCU_CONVERGENCE_CRITERIA_L2) ? l2
         : (convergence & OCU_CONVERGENCE_CRITERIA_LINF) ? linf : 1e20;
  }

  int coarse_iters() const {
    int c = L - 1;
    int a = dim[c][0] * dim[c][1], bb = dim[c][1] * dim[c][2], cc = dim[c][0] * dim[c][2];
    int m = a > bb ? (a > cc ? a : cc) : (bb > cc ? bb : cc);
    return m / 2;
  }

  /* do_fmg, sol_mgpressure3d.cpp:113-251 */
  bool do_fmg(double tolerance, int max_iter, double &result_l2, double &result_linf) {
    any_error = false;
    result_l2 = result_linf = 0;
    apply_boundary_conditions(0);
    double orig_l2 = 0, orig_linf = 0;
    restrict_residuals(0, 0, (convergence & OCU_CONVERGENCE_CALC_L2) ? &orig_l2 : nullptr,
                       (convergence & OCU_CONVERGENCE_CALC_LINF) ? &orig_linf : nullptr);
    double orig_error = crit(orig_l2, orig_linf);
    if (orig_error < tolerance) {
      if (convergence & OCU_CONVERGENCE_CALC_L2) result_l2 = orig_l2;
      if (convergence & OCU_CONVERGENCE_CALC_LINF) result_linf = orig_linf;
      return true;
    }
    int coarse_level = L - 1;
    for (int level = 0; level < L - 1; level++) {
      clear_zero(level + 1);
      apply_boundary_conditions(level + 1);
      restrict_residuals(level, level + 1, nullptr, nullptr);
    }
    for (int fine_level = L - 1; fine_level >= 0; fine_level--) {
      int level_ncyc = (fine_level == 0) ? max_iter + 1 : 1;
      for (int i_cyc = 0; i_cyc < level_ncyc; i_cyc++) {
        if (fine_level == 0) num_vcycles++;
        for (int level = fine_level; level < coarse_level; level++) {
          relax(level, nu1, RO_RED_BLACK);
          clear_zero(level + 1);
          apply_boundary_conditions(level + 1);
          if (level == 0) {
            restrict_residuals(0, 1, (convergence & OCU_CONVERGENCE_CALC_L2) ? &result_l2 : nullptr,
                               (convergence & OCU_CONVERGENCE_CALC_LINF) ? &result_linf : nullptr);
            double res = crit(result_l2, result_linf);
            history.push_back(res);
            if (res < tolerance) {
              apply_boundary_conditions(0);
              return !any_error;
            }
          } else {
            restrict_residuals(level, level + 1, nullptr, nullptr);
          }
        }
        relax(coarse_level, coarse_iters(), make_symmetric_operator ? RO_SYMMETRIC : RO_RED_BLACK);
        for (int level = coarse_level - 1; level >= fine_level; level--) {
          prolong_(level + 1, level);
          if (level > 0) relax(level, nu2, make_symmetric_operator ? RO_BLACK_RED : RO_RED_BLACK);
        }
      }
      if (fine_level > 0) prolong_(fine_level, fine_level - 1);
    }
    return false;
  }

  /* do_vcycle, sol_mgpressure3d.cpp:253-330 */
  bool do_vcycle(double tolerance, int max_iter, double &result_l2, double &result_linf) {
    any_error = false;
    int coarse_level = L - 1;
    apply_boundary_conditions(0);
    double orig_l2 = 0, orig_linf = 0;
    restrict_residuals(0, 0, (convergence & OCU_CONVERGENCE_CALC_L2) ? &orig_l2 : nullptr,
                       (convergence & OCU_CONVERGENCE_CALC_LINF) ? &orig_linf : nullptr);
    if (crit(orig_l2, orig_linf) < tolerance) {
      if (convergence & OCU_CONVERGENCE_CALC_L2) result_l2 = orig_l2;
      if (convergence & OCU_CONVERGENCE_CALC_LINF) result_linf = orig_linf;
      return true;
    }
    for (int i_cyc = 0; i_cyc < max_iter; i_cyc++) {
      num_vcycles++;
      for (int level = 0; level < coarse_level; level++) {
        relax(level, nu1, RO_RED_BLACK);
        clear_zero(level + 1);
        apply_boundary_conditions(level + 1);
        if (level == 0) {
          restrict_residuals(0, 1, (convergence & OCU_CONVERGENCE_CALC_L2) ? &result_l2 : nullptr,
                             (convergence & OCU_CONVERGENCE_CALC_LINF) ? &result_linf : nullptr);
          double res = crit(result_l2, result_linf);
          history.push_back(res);
          if (res < tolerance) {
            apply_boundary_conditions(0);
            return !any_error;
          }
        } else {
          restrict_residuals(level, level + 1, nullptr, nullptr);
        }
      }
      relax(coarse_level, coarse_iters(), make_symmetric_operator ? RO_SYMMETRIC : RO_RED_BLACK);
      for (int level = coarse_level - 1; level >= 0; level--) {
        prolong_(level + 1, level);
        if (level > 0) relax(level, nu2, make_symmetric_operator ? RO_BLACK_RED : RO_RED_BLACK);
      }
    }
    return false;
  }

  /* solve, sol_mgpressure3d.cpp:332-370: retry once from a zero guess. */
  bool solve(double &residual_out, double tolerance, int max_iter) {
    if (!std::isfinite(tolerance) || tolerance < 0 || max_iter <= 0) return false;
    double l2, linf;
    if (!do_fmg(tolerance, max_iter, l2, linf)) {
      clear_zero(0);
      if (!do_fmg(tolerance, max_iter, l2, linf)) {
        residual_out = (convergence == OCU_CONVERGENCE_L2) ? l2 : linf;
        return false;
      }
    }
    residual_out = (convergence == OCU_CONVERGENCE_L2) ? l2 : linf;
    return true;
  }

  bool run_vcycles(int count) {
    if (count <= 0) return false;
    double l2, linf;
    do_vcycle(0, count, l2, linf);
    return !any_error;
  }

  bool run_fmg(int count) {
    if (count < 0) return false;
    double l2, linf;
    do_fmg(0, count, l2, linf);
    return !any_error;
  }
};

/* ------------------------------------------------------------------ */
/* Mixed-precision iterative refinement (no reference counterpart)      */
/* ------------------------------------------------------------------ */

/* FP64 residual of the level-0 problem with the right-hand side shifted by
 * a constant: r = (b - bshift) - A x, evaluated with the same operator and
 * operand order as the reference FP64 residual kernel. */
inline void residual_shifted_f64(const ocu_grid3d &xg, const ocu_grid3d &bg, double bshift,
                                 std::vector<double> &rout, const ocu_grid3d &shape,
                                 double h, double fxd, double fyd, double fzd) {
  View<double> x(xg), b(bg);
  const double fx = fxd / (h * h), fy = fyd / (h * h), fz = fzd / (h * h);
  const double diag = 2 * (fxd + fyd + fzd) / (h * h);
  const int nx = shape.nx, ny = shape.ny, nz = shape.nz;
  rout.resize((size_t)nx * ny * nz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        rout[((size_t)i * ny + j) * nz + k] =
            (b.at(i, j, k) - bshift) -
            (fz * (x.at(i, j, k - 1) + x.at(i, j, k + 1)) + fy * (x.at(i, j - 1, k) + x.at(i, j + 1, k)) +
             fx * (x.at(i - 1, j, k) + x.at(i + 1, j, k)) - diag * x.at(i, j, k));
}

/* The singular case: no Dirichlet face means constants are in the null
 * space (a pure Neumann/periodic operator). */
inline bool is_singular(const ocu_bc_set &bc) {
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (b->type == OCU_BC_DIRICHELET) return false;
  return true;
}

/* Iterative refinement: FP64 outer residual and correction, FP32 V-cycles
 * on the correction equation with homogeneous boundary conditions at every
 * level, FP64 coarsest-level relaxation.  Exactly the algorithm of the HIP
 * implementation (MixedIR in the host library):
 *
 *   if singular: bm = mean(b) else bm = 0
 *   loop it = 0..max_ir:
 *     r = (b - bm) - A x                         (FP64)
 *     res = RMS(r) or max|r|; stop if res <= target
 *     s = 2^-e, e = exponent of max|r|          (power-of-two scaling)
 *     e32 = 0; b32 = (float)(r * s)
 *     it == 0 and fmg_first: FP32 FMG with cycles_first V-cycles at level 0
 *     else:                  cycles V-cycles
 *     (each V-cycle: relax(l, nu1) + restrict going down, coarsest relaxed in
 *      FP64, prolong + relax(l, nu2) going up for l > 0), then relax(0, nu1)
 *     singular: shift = (sum x + sum e32 / s) / N, the mean of x + e32/s
 *     x = (x + (double)e32 / s) - shift (singular) or x + (double)e32 / s ;
 *     BC(x)
 *   if singular and no correction was made: x -= mean(x); BC(x)
 */
/* V-cycles for refinement iteration it >= 1 when the previous iteration
 * ran prev_cycles V-cycles and took the residual from prev_res to res: the
 * observed contraction per cycle rho = (res/prev_res)^(1/prev_cycles) is
 * extrapolated to reach half the tolerance, within [1, cap]. */
inline int adaptive_cycles(double res, double prev_res, int prev_cycles, double tol, int cap) {
  if (!(prev_res > 0) || !(res > 0) || !(res < prev_res) || prev_cycles <= 0) return cap;
  const double rho = std::pow(res / prev_res, 1.0 / prev_cycles);
  const double need = std::log(0.5 * tol / res) / std::log(rho);
  if (!(need > 1)) return 1;
  const double n = std::ceil(need);
  return n > cap ? cap : (int)n;
}

struct MixedIR {
  MGSolver<float> mg;  // FP32 hierarchy; level-0 u = e32, b = r32
  HostGrid<float> e32, b32;
  ocu_grid3d x{}, b{};
  ocu_bc_set bc;
  int nu1 = 2, nu2 = 2;
  int cycles = 4, cycles_first = 4;
  bool fmg_first = false;
  bool adaptive = false;  // later iterations: adaptive_cycles(...) capped by cycles
  int convergence = OCU_CONVERGENCE_L2;
  int prolong_quirk = 1;
  bool singular = false;
  // statistics
  int ir_iterations = 0, total_vcycles = 0;
  std::vector<double> history;

  MixedIR() { std::memset(&bc, 0, sizeof(bc)); }

  bool initialize_storage(int nx, int ny, int nz, double hxv, double hyv, double hzv,
                          const ocu_grid3d &rhs64, const ocu_grid3d &x64) {
    b = rhs64; x = x64;
    e32.init_congruent(rhs64);
    b32.init_congruent(rhs64);
    mg.bc = bc_at_level(bc, 1);  // homogeneous on every level
    mg.nu1 = nu1; mg.nu2 = nu2;
    mg.prolong_quirk = prolong_quirk;
    mg.disable_relax_on_host = true;  // coarsest handled by coarse_solve_f64
    singular = is_singular(bc);
    return mg.initialize_storage(nx, ny, nz, hxv, hyv, hzv, b32.desc, e32.desc);
  }

  /* Coarsest level: relaxation carried out in double on a copy of the float
   * grids (host operation order), result rounded back to float. */
  void coarse_solve_f64(int iters) {
    int l = mg.L - 1;
    HostGrid<double> ud, bd;
    ud.init_congruent(mg.u(l));
    bd.init_congruent(mg.b(l));
    const float *us = (const float *)mg.u(l).base, *bs = (const float *)mg.b(l).base;
    for (int64_t n = 0; n < ud.desc.allocated; n++) { ud.buf[n] = us[n]; bd.buf[n] = bs[n]; }
    ocu_bc_set hbc = bc_at_level(bc, l);
    ocu_mg_level_params P = level_params(mg.get_h(l), mg.fx, mg.fy, mg.fz, mg.omega, hbc);
    for (int it = 0; it < iters; it++)
      for (int color = 0; color < 2; color++) {
        relax_color_host<double>(ud.desc, bd.desc, color, P);
        bc_fill_level1<double>(ud.desc, hbc, mg.hx(l), mg.hy(l), mg.hz(l));
      }
    float *ud_out = (float *)mg.u(l).base;
    for (int64_t n = 0; n < ud.desc.allocated; n++) ud_out[n] = (float)ud.buf[n];
  }

  void vcycle_f32(int fine_level) {
    int coarse = mg.L - 1;
    for (int level = fine_level; level < coarse; level++) {
      mg.relax(level, nu1, MGSolver<float>::RO_RED_BLACK);
      mg.clear_zero(level + 1);
      mg.apply_boundary_conditions(level + 1);
      mg.restrict_residuals(level, level + 1, nullptr, nullptr);
    }
    coarse_solve_f64(mg.coarse_iters());
    for (int level = coarse - 1; level >= fine_level; level--) {
      mg.prolong_(level + 1, level);
      if (level > 0) mg.relax(level, nu2, MGSolver<float>::RO_RED_BLACK);
    }
  }

  void fmg_f32(int ncyc0) {
    // residual of e32 = 0 is b32: restrict it down
    for (int level = 0; level < mg.L - 1; level++) {
      mg.clear_zero(level + 1);
      mg.apply_boundary_conditions(level + 1);
      mg.restrict_residuals(level, level + 1, nullptr, nullptr);
    }
    for (int fine_level = mg.L - 1; fine_level >= 0; fine_level--) {
      int nc = (fine_level == 0) ? ncyc0 : 1;
      for (int c = 0; c < nc; c++) {
        vcycle_f32(fine_level);
        total_vcycles++;
      }
      if (fine_level > 0) mg.prolong_(fine_level, fine_level - 1);
    }
  }

  double mean_interior(const ocu_grid3d &g) {
    View<double> v(g);
    std::vector<double> ps(g.nx);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g.nx; i++) {
      double s = 0;
      for (int j = 0; j < g.ny; j++)
        for (int k = 0; k < g.nz; k++) s += v.at(i, j, k);
      ps[i] = s;
    }
    double s = 0;
    for (int i = 0; i < g.nx; i++) s += ps[i];
    return s / ((double)g.nx * g.ny * g.nz);
  }

  /* tol is absolute in the chosen norm (RMS for L2, max for LINF) */
  bool solve(double &res_out, double tol, int max_ir) {
    const int nx = x.nx, ny = x.ny, nz = x.nz;
    double bm = singular ? mean_interior(b) : 0.0;
    ocu_bc_set bc0 = bc;
    bc_fill_level1<double>(x, bc0, mg.hx(0), mg.hy(0), mg.hz(0));
    std::vector<double> r;
    View<float> bv(b32.desc), ev(e32.desc);
    View<double> xv(x);
    double res = 0;
    bool ok = false;
    // the cast scale of iteration it comes from max|r| of iteration it-1
    // (max|b| for the first), so the scaled cast can share the residual pass
    double prev_max = 0, prev_res = 0;
    int prev_cycles = 0;
    {
      View<double> bb(b);
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) prev_max = std::max(prev_max, std::fabs(bb.at(i, j, k)));
    }
    for (int it = 0;; it++) {
      int ex = 0;
      std::frexp(prev_max > 0 ? prev_max : 1.0, &ex);
      double scale = std::ldexp(1.0, -ex), inv = std::ldexp(1.0, ex);
      residual_shifted_f64(x, b, bm, r, x, mg.get_h(0), mg.fx, mg.fy, mg.fz);
      const double xsum = mean_interior(x) * ((double)nx * ny * nz);
      double s2 = 0, mx = 0;
      for (double q : r) { s2 += q * q; mx = std::max(mx, std::fabs(q)); }
      res = (convergence & OCU_CONVERGENCE_CRITERIA_LINF) ? mx
            : std::sqrt(s2 / ((double)nx * ny * nz));
      history.push_back(res);
      if (res <= tol) { ok = true; break; }
      if (it >= max_ir) break;
      prev_max = mx;
      ir_iterations++;
      e32.clear_zero();
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++)
            bv.at(i, j, k) = (float)(r[((size_t)i * ny + j) * nz + k] * scale);
      if (it == 0 && fmg_first) {
        fmg_f32(cycles_first);
        prev_cycles = 0;
      } else {
        int nc = it == 0 ? cycles_first : cycles;
        if (adaptive && it > 0) nc = adaptive_cycles(res, prev_res, prev_cycles, tol, nc);
        mg.apply_boundary_conditions(0);
        for (int c = 0; c < nc; c++) { vcycle_f32(0); total_vcycles++; }
        prev_cycles = nc;
      }
      prev_res = res;
      mg.relax(0, nu1, MGSolver<float>::RO_RED_BLACK);
      double shift = 0;
      if (singular) {
        double esum = 0;
        for (int i = 0; i < nx; i++)
          for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) esum += (double)ev.at(i, j, k);
        shift = (xsum + inv * esum) / ((double)nx * ny * nz);
      }
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) {
            const double v = xv.at(i, j, k) + (double)ev.at(i, j, k) * inv;
            xv.at(i, j, k) = singular ? v - shift : v;
          }
      bc_fill_level1<double>(x, bc0, mg.hx(0), mg.hy(0), mg.hz(0));
    }
    if (singular && ir_iterations == 0) {
      double xm = mean_interior(x);
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) xv.at(i, j, k) = xv.at(i, j, k) - xm;
      bc_fill_level1<double>(x, bc0, mg.hx(0), mg.hy(0), mg.hz(0));
    }
    res_out = res;
    return ok;
  }
};

}  // namespace orc

#endif
