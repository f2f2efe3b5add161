/*
 * ORACLE - test infrastructure only.
 *
 * CPU restatement of Sol_PCGPressure3DDevice<T>
 * (ocuequation/sol_pcgpressure3d.cu:27-662): preconditioned conjugate
 * gradients for the variable-coefficient pressure equation
 *     -div(c grad p) = rhs
 * with face coefficients (c_i + c_nb)/2, preconditioned by nothing, the
 * coefficient diagonal (Jacobi), a few multigrid cycles of the constant
 * coefficient Laplacian, or both (BFBT).  Floating-point expressions keep the
 * reference's operand order and precisions: the stencil coefficients are T
 * values carried in double kernel parameters, so the stencil and diagonal
 * sums are evaluated in double and rounded to T once.  Dot products: the
 * reference multiplies in T and sums the products in T with its reduction
 * tree; here (and in the GPU kernels) the T products are summed in double
 * and the sum is rounded to T, a deliberate difference that makes the sum
 * order-insensitive to within double rounding.
 */
#ifndef OCU_ORACLE_PCG_H
#define OCU_ORACLE_PCG_H

#include <cstdio>

#include "oracle_grid.h"
#include "oracle_mg.h"

namespace orc {

enum { PRECOND_NONE = 0, PRECOND_JACOBI = 1, PRECOND_MULTIGRID = 2, PRECOND_BFBT = 3 };

/* Sol_PCGPressure3DDevice_apply_laplacian (sol_pcgpressure3d.cu:88-126):
 * dst = -(sum over the six neighbours of f/h^2 (src_nb - src) (.5)(c_nb + c))
 * on the interior, with f/h^2 T values promoted to double. */
template <typename T>
void pcg_apply_laplacian(const ocu_grid3d &dg, const ocu_grid3d &sg, const ocu_grid3d &cg, T fxh, T fyh, T fzh) {
  View<T> D(dg), S(sg), Cf(cg);
  const double fx = fxh, fy = fyh, fz = fzh;
  for (int i = 0; i < sg.nx; i++)
    for (int j = 0; j < sg.ny; j++)
      for (int k = 0; k < sg.nz; k++) {
        const T c = S.at(i, j, k), cc = Cf.at(i, j, k);
        const double r = fz * (S.at(i, j, k - 1) - c) * (T)0.5 * (Cf.at(i, j, k - 1) + cc) +
                         fz * (S.at(i, j, k + 1) - c) * (T)0.5 * (Cf.at(i, j, k + 1) + cc) +
                         fy * (S.at(i, j - 1, k) - c) * (T)0.5 * (Cf.at(i, j - 1, k) + cc) +
                         fy * (S.at(i, j + 1, k) - c) * (T)0.5 * (Cf.at(i, j + 1, k) + cc) +
                         fx * (S.at(i - 1, j, k) - c) * (T)0.5 * (Cf.at(i - 1, j, k) + cc) +
                         fx * (S.at(i + 1, j, k) - c) * (T)0.5 * (Cf.at(i + 1, j, k) + cc);
        D.at(i, j, k) = (T)(-r);
      }
}

/* Sol_PCGPressure3DDevice_diag_preconditioner (sol_pcgpressure3d.cu:45-83):
 * dst = src / (the diagonal of the operator above), in double */
template <typename T>
void pcg_diag_precond(const ocu_grid3d &dg, const ocu_grid3d &sg, const ocu_grid3d &cg, T fxh, T fyh, T fzh) {
  View<T> D(dg), S(sg), Cf(cg);
  const double fx = fxh, fy = fyh, fz = fzh;
  for (int i = 0; i < sg.nx; i++)
    for (int j = 0; j < sg.ny; j++)
      for (int k = 0; k < sg.nz; k++) {
        const T cc = Cf.at(i, j, k);
        const double diag = fz * (T)0.5 * (Cf.at(i, j, k - 1) + cc) + fz * (T)0.5 * (Cf.at(i, j, k + 1) + cc) +
                            fy * (T)0.5 * (Cf.at(i, j - 1, k) + cc) + fy * (T)0.5 * (Cf.at(i, j + 1, k) + cc) +
                            fx * (T)0.5 * (Cf.at(i - 1, j, k) + cc) + fx * (T)0.5 * (Cf.at(i + 1, j, k) + cc);
        D.at(i, j, k) = (T)(S.at(i, j, k) / diag);
      }
}

/* invoke_kernel_dot_product (sol_pcgpressure3d.cu:262-293): T products of
 * the interiors, summed (in double here, in i,j,k order) and rounded to T */
template <typename T>
T pcg_dot(const ocu_grid3d &ag, const ocu_grid3d &bg) {
  View<T> A(ag), B(bg);
  double acc = 0;
  for (int i = 0; i < ag.nx; i++)
    for (int j = 0; j < ag.ny; j++)
      for (int k = 0; k < ag.nz; k++) acc += (double)(T)(A.at(i, j, k) * B.at(i, j, k));
  return (T)acc;
}

template <typename T>
struct PCGSolver {
  // public state, sol_pcgpressure3d.cu:213-226 / :600-660
  ocu_bc_set bc;
  int convergence = OCU_CONVERGENCE_L2;
  int preconditioner = PRECOND_BFBT;
  int multigrid_cycles = 2;
  bool multigrid_use_fmg = true;

  double fx = 0, fy = 0, fz = 0, h = 1;
  int nx = 0, ny = 0, nz = 0;
  ocu_grid3d rhs{}, coeff{};
  HostGrid<T> pressure, d_r, mg_u;  // mg_u: the multigrid's own solution grid
  MGSolver<T> mg;
  bool any_error = false;
  int iterations = 0;

  PCGSolver() { std::memset(&bc, 0, sizeof(bc)); }

  double hx() const { return h / std::sqrt(fx); }
  double hy() const { return h / std::sqrt(fy); }
  double hz() const { return h / std::sqrt(fz); }
  T fxh() const { return (T)(fx / (h * h)); }
  T fyh() const { return (T)(fy / (h * h)); }
  T fzh() const { return (T)(fz / (h * h)); }

  /* initialize_storage, sol_pcgpressure3d.cu:600-650 (float spacings) */
  bool initialize_storage(int nxv, int nyv, int nzv, float hxv, float hyv, float hzv, const ocu_grid3d &rhs_,
                          const ocu_grid3d &coeff_) {
    if (!std::isfinite(hxv) || !std::isfinite(hyv) || !std::isfinite(hzv)) return false;
    const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
    for (auto b : all)
      if (b->type != OCU_BC_PERIODIC && b->type != OCU_BC_NEUMANN) return false;
    // min3 of the float spacings; hx*hx is a float product
    h = min3(hxv, hyv, hzv);
    fx = (h * h) / (hxv * hxv);
    fy = (h * h) / (hyv * hyv);
    fz = (h * h) / (hzv * hzv);
    nx = nxv, ny = nyv, nz = nzv;
    if (rhs_.gx < 1 || rhs_.gy < 1 || rhs_.gz < 1) return false;
    if (rhs_.nx != nxv || rhs_.ny != nyv || rhs_.nz != nzv) return false;
    if (!layout_match(coeff_, rhs_)) return false;
    rhs = rhs_, coeff = coeff_;
    pressure.init_congruent(rhs);
    d_r.init_congruent(rhs);
    mg.nu1 = 2;
    mg.nu2 = 2;
    mg.bc = bc;
    mg.convergence = OCU_CONVERGENCE_NONE_CALC_L2LINF;
    mg_u.init_congruent(rhs);
    return mg.initialize_storage(nx, ny, nz, hxv, hyv, hzv, d_r.desc, mg_u.desc);
  }

  /* invoke_kernel_apply_laplacian (:183-220): no-corner fill of src, then A src */
  void apply_laplacian(const ocu_grid3d &dst, const ocu_grid3d &src) {
    if (!bc_fill_nocorners_device<T>(src, bc, hx(), hy(), hz())) any_error = true;
    pcg_apply_laplacian<T>(dst, src, coeff, fxh(), fyh(), fzh());
  }

  /* apply_preconditioner (:295-336) */
  void apply_preconditioner(const ocu_grid3d &dst, const ocu_grid3d &src) {
    if (preconditioner == PRECOND_JACOBI) {
      pcg_diag_precond<T>(dst, src, coeff, fxh(), fyh(), fzh());
    } else if (preconditioner == PRECOND_MULTIGRID || preconditioner == PRECOND_BFBT) {
      mg.clear_zero(0);
      if (multigrid_use_fmg) mg.run_fmg(multigrid_cycles);
      else mg.run_vcycles(multigrid_cycles);
      lincomb1<T>(dst, (T)-1, mg.u(0));
      if (preconditioner == PRECOND_BFBT) pcg_diag_precond<T>(dst, dst, coeff, fxh(), fyh(), fzh());
    } else {
      std::memcpy(dst.base, src.base, sizeof(T) * (size_t)dst.allocated);
    }
  }

  /* do_pcg (:339-447) */
  bool do_pcg(double tolerance, int max_iter, double &result_l2, double &result_linf) {
    if (!bc_fill_nocorners_device<T>(coeff, bc, hx(), hy(), hz())) any_error = true;
    HostGrid<T> y, d;
    y.init_congruent(pressure.desc);
    d.init_congruent(pressure.desc);
    apply_laplacian(y.desc, pressure.desc);
    lincomb2<T>(d_r.desc, (T)1, rhs, (T)-1, y.desc);
    apply_preconditioner(d.desc, d_r.desc);
    T delta0 = pcg_dot<T>(d_r.desc, d.desc);
    T deltanew = delta0;
    int it = 0;
    bool converged = (it > max_iter) || (delta0 < tolerance * tolerance);
    while (!converged) {
      apply_laplacian(y.desc, d.desc);
      T d_dot_y = pcg_dot<T>(d.desc, y.desc);
      T alpha = deltanew / d_dot_y;
      lincomb2<T>(pressure.desc, (T)1, pressure.desc, alpha, d.desc);
      if (it % 25 == 0) {
        apply_laplacian(y.desc, pressure.desc);
        lincomb2<T>(d_r.desc, (T)1, rhs, (T)-1, y.desc);
      } else {
        lincomb2<T>(d_r.desc, (T)1, d_r.desc, -alpha, y.desc);
      }
      apply_preconditioner(y.desc, d_r.desc);
      T deltaold = deltanew;
      deltanew = pcg_dot<T>(d_r.desc, y.desc);
      T beta = deltanew / deltaold;
      lincomb2<T>(d.desc, beta, d.desc, (T)1, y.desc);
      it++;
      converged = (it > max_iter) || (deltanew < tolerance * tolerance);
    }
    iterations = it;
    bool ok = deltanew <= tolerance * tolerance;
    result_linf = reduce<T>(d_r.desc, 0);
    result_l2 = std::sqrt((double)deltanew);
    return ok;
  }

  /* solve (:550-597) */
  bool solve(double &residual, double tolerance, int max_iter) {
    any_error = false;
    if (!std::isfinite(tolerance) || tolerance < 0 || max_iter <= 0) return false;
    double l2 = 0, linf = 0;
    bool ok = do_pcg(tolerance, max_iter, l2, linf);
    residual = (convergence == OCU_CONVERGENCE_L2) ? l2 : linf;
    return ok && !any_error;
  }
};

}  // namespace orc

#endif
