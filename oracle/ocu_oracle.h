/*
 * ORACLE - test infrastructure only.  C entry points of the CPU restatement
 * of OpenCurrent's projection-method hot path.  Grids are host buffers
 * described by ocu_grid3d (base = host pointer).  Return value: 1 = the
 * reference call would return true, 0 = false, negative = bad arguments.
 */
#ifndef OCU_ORACLE_H
#define OCU_ORACLE_H

#include "../include/ocu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int orc_layout(int nx, int ny, int nz, int gx, int gy, int gz, int padx, int pady, int padz,
               int dtype, ocu_grid3d *out);

/* kind: 0 level1 (x,y,z passes), 1 device no-corners, 2 host no-corners */
int orc_bc_fill(int kind, const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz);
int orc_bc_fill_mac(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc);

int orc_set_threads(int n);
int orc_lincomb1(const ocu_grid3d *r, double a1, const ocu_grid3d *g1);
int orc_lincomb2(const ocu_grid3d *r, double a1, const ocu_grid3d *g1, double a2, const ocu_grid3d *g2);
int orc_copy_convert(const ocu_grid3d *dst, const ocu_grid3d *src);
/* op: 0 maxabs, 1 max, 2 min, 3 sum, 4 sqrsum, 5 checknan */
double orc_reduce(const ocu_grid3d *g, int op);

int orc_mg_num_levels(int nx, int ny, int nz);
double orc_mg_optimal_omega(double fx, double fy, double fz);
int orc_mg_level_params(const ocu_bc_set *bc, int level, double hx, double hy, double hz,
                        ocu_mg_level_params *out);
int orc_mg_relax(const ocu_grid3d *u, const ocu_grid3d *b, int color,
                 const ocu_mg_level_params *p, int variant);
int orc_mg_residual(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r,
                    double h, double fx, double fy, double fz);
int orc_mg_restrict(const ocu_grid3d *r, const ocu_grid3d *bc);
int orc_mg_prolong(const ocu_grid3d *uc, const ocu_grid3d *uf, int quirk);

/* Full solver.  mode 0: solve(residual, tol, max_iter); 1: run_vcycles(max_iter);
 * 2: run_fmg(max_iter).  u0 is the pressure (initial guess in, solution out),
 * congruent with b0.  history receives the level-0 residual after each
 * pre-smoothing (up to history_cap values). */
int orc_mg_solve(const ocu_grid3d *u0, const ocu_grid3d *b0, double hx, double hy, double hz,
                 const ocu_mg_config *cfg, double tol, int max_iter, int mode,
                 double *residual, int *vcycles, double *history, int history_cap, int *history_len);

/* Mixed-precision iterative refinement (FP32 V-cycles in FP64 IR).  x, b are
 * FP64.  tol is absolute in the cfg->convergence norm. */
int orc_ir_solve(const ocu_grid3d *x, const ocu_grid3d *b, double hx, double hy, double hz,
                 const ocu_mg_config *cfg, double tol, int max_ir, int cycles, int cycles_first,
                 int fmg_first, int adaptive, double *residual, int *ir_iters, int *vcycles);

int orc_divergence(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *div,
                   int nx, int ny, int nz, double hx, double hy, double hz);
int orc_gradient_sub(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *p,
                     double coefficient, double hx, double hy, double hz);
int orc_self_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                    const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz,
                    double hx, double hy, double hz, int scheme);
int orc_passive_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                       const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                       double hx, double hy, double hz, int scheme);
/* The two above with OCU_IT_SECOND_ORDER_UPWIND allowed; periodic_axes bit a
 * (x=1, y=2, z=4) lets that scheme's stencil wrap along axis a. */
int orc_self_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                       const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz,
                       double hx, double hy, double hz, int scheme, int periodic_axes);
int orc_passive_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                          const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                          double hx, double hy, double hz, int scheme, int periodic_axes);
int orc_laplacian_acc(const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                      double hx, double hy, double hz, double coefficient);
int orc_thermal_force(const ocu_grid3d *d, const ocu_grid3d *t, int nx, int ny, int nz,
                      int vertical_direction, double coefficient);

/* Adjoints (no reference counterpart): see oracle_ns.h.
 * lam_out = lam + A_phi^T(scale*lam); adj_u/v/w += A_u^T(scale*lam). */
int orc_passive_advect_adjoint(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                               const ocu_grid3d *phi, const ocu_grid3d *lam, const ocu_grid3d *lam_out,
                               const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                               double scale, int nx, int ny, int nz, double hx, double hy, double hz,
                               int scheme);
int orc_bc_adjoint_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc);
int orc_mac_adjoint_normal(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc);
int orc_divergence_adjoint(const ocu_grid3d *lam, const ocu_grid3d *adj_u, const ocu_grid3d *adj_v,
                           const ocu_grid3d *adj_w, int nx, int ny, int nz, double hx, double hy, double hz);
int orc_gradient_adjoint(const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                         const ocu_grid3d *adj_p, double coefficient, double hx, double hy, double hz);

/* One projection, Sol_ProjectDivergence3DDevice<T>::solve(tol) on a fresh
 * solver (zero pressure guess).  pressure_out (congruent with u) receives p. */
int orc_project(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                int nx, int ny, int nz, double hx, double hy, double hz, double tol,
                const ocu_grid3d *pressure_out, double *residual);

typedef struct orc_ns_params {
  int32_t nx, ny, nz;
  double hx, hy, hz;
  double max_divergence;
  ocu_bc_set temp_bc, flow_bc;
  int32_t vertical_direction;
  double gravity, bouyancy, thermal_diffusion, viscosity;
  int32_t advection_scheme, time_step;
  double cfl_factor;
} orc_ns_params;

void *orc_ns_create(int dtype, const orc_ns_params *p, const ocu_grid3d *u, const ocu_grid3d *v,
                    const ocu_grid3d *w, const ocu_grid3d *t);
int orc_ns_advance_one_step(void *h, double dt);
int orc_ns_advance(void *h, double dt);
double orc_ns_max_stable_timestep(void *h);
/* which: 0 u, 1 v, 2 w, 3 temperature, 4 pressure; out must be congruent */
int orc_ns_get(void *h, int which, const ocu_grid3d *out);
void orc_ns_destroy(void *h);

int orc_scalar_advect_step(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_grid3d *phi, const ocu_grid3d *dphi, const ocu_bc_set *bc,
                           int nx, int ny, int nz, double hx, double hy, double hz, int scheme,
                           double dt);

/* Sol_PCGPressure3DDevice<T> (sol_pcgpressure3d.cu): fxh/fyh/fzh are the
 * T values f/h^2 the kernels receive */
int orc_pcg_apply_laplacian(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fxh,
                            double fyh, double fzh);
int orc_pcg_diag_precond(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fxh,
                         double fyh, double fzh);
double orc_pcg_dot(const ocu_grid3d *a, const ocu_grid3d *b);
/* solve(residual, tol, max_iter) from the guess in pressure (overwritten by
 * the result); the coefficient's ghosts are filled in place as the reference
 * does.  preconditioner: 0 none, 1 Jacobi, 2 multigrid, 3 BFBT. */
int orc_pcg_solve(const ocu_grid3d *rhs, const ocu_grid3d *coeff, const ocu_grid3d *pressure, float hx, float hy,
                  float hz, const ocu_bc_set *bc, int preconditioner, int multigrid_cycles, int use_fmg,
                  int convergence, double tol, int max_iter, double *residual, int *iterations);

/* two-layer ghost fills (grid3dboundary.cu:904-1016) and point sampling
 * (grid3dsample.cu); positions are float host arrays, out has phi's type */
int orc_bc_fill_level2_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz);
int orc_bc_fill_mac_level2(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc);
int orc_sample_points(void *out, const float *px, const float *py, const float *pz, int n, const ocu_grid3d *phi,
                      const ocu_bc_set *bc, float period_nx, float period_ny, float period_nz, float hx, float hy,
                      float hz, float shiftx, float shifty, float shiftz);
int orc_sample_points_mac(void *vx, void *vy, void *vz, const float *px, const float *py, const float *pz, int n,
                          const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                          float hx, float hy, float hz);

#ifdef __cplusplus
}
#endif

#endif
