/*
 * ORACLE - test infrastructure only.  C entry points over the templates in
 * oracle_grid.h / oracle_mg.h / oracle_ns.h (dispatch on dtype).
 */
#include "ocu_oracle.h"

#include <omp.h>

#include "oracle_ns.h"
#include "oracle_pcg.h"
#include "oracle_bc2.h"

using namespace orc;

#define DISPATCH(dt, CALL_F, CALL_D) \
  ((dt) == OCU_F64 ? (CALL_D) : (dt) == OCU_F32 ? (CALL_F) : -1)

/* C linkage comes from the declarations in ocu_oracle.h */

int orc_layout(int nx, int ny, int nz, int gx, int gy, int gz, int padx, int pady, int padz,
               int dtype, ocu_grid3d *out) {
  if (!out) return -1;
  *out = layout(nx, ny, nz, gx, gy, gz, padx, pady, padz, dtype);
  return 1;
}

int orc_bc_fill(int kind, const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz) {
  if (kind == 0) return DISPATCH(g->dtype, bc_fill_level1<float>(*g, *bc, hx, hy, hz), bc_fill_level1<double>(*g, *bc, hx, hy, hz));
  if (kind == 1) return DISPATCH(g->dtype, bc_fill_nocorners_device<float>(*g, *bc, hx, hy, hz), bc_fill_nocorners_device<double>(*g, *bc, hx, hy, hz));
  if (kind == 2) return DISPATCH(g->dtype, bc_fill_nocorners_host<float>(*g, *bc, hx, hy, hz), bc_fill_nocorners_host<double>(*g, *bc, hx, hy, hz));
  return -1;
}

int orc_bc_fill_mac(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc) {
  return DISPATCH(u->dtype, bc_fill_mac_level1<float>(*u, *v, *w, *bc), bc_fill_mac_level1<double>(*u, *v, *w, *bc));
}

int orc_lincomb1(const ocu_grid3d *r, double a1, const ocu_grid3d *g1) {
  return DISPATCH(r->dtype, lincomb1<float>(*r, (float)a1, *g1), lincomb1<double>(*r, a1, *g1));
}

int orc_lincomb2(const ocu_grid3d *r, double a1, const ocu_grid3d *g1, double a2, const ocu_grid3d *g2) {
  return DISPATCH(r->dtype, lincomb2<float>(*r, (float)a1, *g1, (float)a2, *g2),
                  lincomb2<double>(*r, a1, *g1, a2, *g2));
}

/* copy_all_data<T,S>, grid3ddev.cu:25-33 / :216-233: element-wise (T)from
 * over the whole allocation. */
int orc_copy_convert(const ocu_grid3d *dst, const ocu_grid3d *src) {
  if (!layout_match(*dst, *src)) return 0;
  for (int64_t n = 0; n < dst->allocated; n++) {
    double x = src->dtype == OCU_F64 ? ((const double *)src->base)[n] : (double)((const float *)src->base)[n];
    if (dst->dtype == OCU_F64) ((double *)dst->base)[n] = x;
    else if (src->dtype == OCU_F32) ((float *)dst->base)[n] = ((const float *)src->base)[n];
    else ((float *)dst->base)[n] = (float)x;
  }
  return 1;
}

double orc_reduce(const ocu_grid3d *g, int op) {
  return g->dtype == OCU_F64 ? reduce<double>(*g, op) : reduce<float>(*g, op);
}

/* OpenMP threads of the oracle's parallel loops (the CPU baseline's thread
 * count); returns the count now in effect */
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_mg_num_levels(int nx, int ny, int nz) { return num_levels(nx, ny, nz); }
double orc_mg_optimal_omega(double fx, double fy, double fz) { return optimal_omega(fx, fy, fz); }

int orc_mg_level_params(const ocu_bc_set *bc, int level, double hx, double hy, double hz,
                        ocu_mg_level_params *out) {
  double h0 = min3(hx, hy, hz);
  double fx = (h0 * h0) / (hx * hx), fy = (h0 * h0) / (hy * hy), fz = (h0 * h0) / (hz * hz);
  double h = h0;
  for (int l = 0; l < level; l++) h *= 2;
  *out = level_params(h, fx, fy, fz, optimal_omega(fx, fy, fz), bc_at_level(*bc, level));
  return 1;
}

int orc_mg_relax(const ocu_grid3d *u, const ocu_grid3d *b, int color, const ocu_mg_level_params *p, int variant) {
  if (u->dtype == OCU_F64) relax_color<double>(*u, *b, color, *p, variant);
  else relax_color<float>(*u, *b, color, *p, variant);
  return 1;
}

int orc_mg_residual(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r, double h,
                    double fx, double fy, double fz) {
  if (u->dtype == OCU_F64) residual<double>(*u, *b, *r, h, fx, fy, fz);
  else residual<float>(*u, *b, *r, h, fx, fy, fz);
  return 1;
}

int orc_mg_restrict(const ocu_grid3d *r, const ocu_grid3d *bc) {
  if (r->dtype == OCU_F64) restrict_<double>(*r, *bc);
  else restrict_<float>(*r, *bc);
  return 1;
}

int orc_mg_prolong(const ocu_grid3d *uc, const ocu_grid3d *uf, int quirk) {
  if (uc->dtype == OCU_F64) prolong<double>(*uc, *uf, quirk);
  else prolong<float>(*uc, *uf, quirk);
  return 1;
}

template <typename T>
static int mg_solve_t(const ocu_grid3d *u0, const ocu_grid3d *b0, double hx, double hy, double hz,
                      const ocu_mg_config *cfg, double tol, int max_iter, int mode, double *residual,
                      int *vcycles, double *history, int history_cap, int *history_len) {
  MGSolver<T> s;
  s.bc = cfg->bc;
  s.nu1 = cfg->nu1; s.nu2 = cfg->nu2;
  s.convergence = cfg->convergence;
  s.make_symmetric_operator = cfg->make_symmetric_operator != 0;
  s.disable_relax_on_host = cfg->disable_relax_on_host != 0;
  s.prolong_quirk = cfg->prolong_quirk;
  if (!s.initialize_storage(b0->nx, b0->ny, b0->nz, hx, hy, hz, *b0, *u0)) return -1;
  bool ok;
  double res = 0;
  if (mode == 0) ok = s.solve(res, tol, max_iter);
  else if (mode == 1) ok = s.run_vcycles(max_iter);
  else ok = s.run_fmg(max_iter);
  if (residual) *residual = res;
  if (vcycles) *vcycles = s.num_vcycles;
  if (history_len) *history_len = (int)s.history.size();
  if (history)
    for (int n = 0; n < (int)s.history.size() && n < history_cap; n++) history[n] = s.history[n];
  return ok ? 1 : 0;
}

int orc_mg_solve(const ocu_grid3d *u0, const ocu_grid3d *b0, double hx, double hy, double hz,
                 const ocu_mg_config *cfg, double tol, int max_iter, int mode, double *residual,
                 int *vcycles, double *history, int history_cap, int *history_len) {
  if (u0->dtype == OCU_F64)
    return mg_solve_t<double>(u0, b0, hx, hy, hz, cfg, tol, max_iter, mode, residual, vcycles, history, history_cap, history_len);
  return mg_solve_t<float>(u0, b0, hx, hy, hz, cfg, tol, max_iter, mode, residual, vcycles, history, history_cap, history_len);
}

int orc_ir_solve(const ocu_grid3d *x, const ocu_grid3d *b, double hx, double hy, double hz,
                 const ocu_mg_config *cfg, double tol, int max_ir, int cycles, int cycles_first,
                 int fmg_first, int adaptive, double *residual, int *ir_iters, int *vcycles) {
  if (x->dtype != OCU_F64 || b->dtype != OCU_F64) return -1;
  MixedIR s;
  s.bc = cfg->bc;
  s.nu1 = cfg->nu1; s.nu2 = cfg->nu2;
  s.convergence = cfg->convergence;
  s.prolong_quirk = cfg->prolong_quirk;
  s.cycles = cycles; s.cycles_first = cycles_first; s.fmg_first = fmg_first != 0;
  s.adaptive = adaptive != 0;
  if (!s.initialize_storage(b->nx, b->ny, b->nz, hx, hy, hz, *b, *x)) return -1;
  double res = 0;
  bool ok = s.solve(res, tol, max_ir);
  if (residual) *residual = res;
  if (ir_iters) *ir_iters = s.ir_iterations;
  if (vcycles) *vcycles = s.total_vcycles;
  return ok ? 1 : 0;
}

int orc_divergence(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *div,
                   int nx, int ny, int nz, double hx, double hy, double hz) {
  if (u->dtype == OCU_F64) divergence<double>(*u, *v, *w, *div, nx, ny, nz, hx, hy, hz);
  else divergence<float>(*u, *v, *w, *div, nx, ny, nz, hx, hy, hz);
  return 1;
}

int orc_gradient_sub(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *p,
                     double coefficient, double hx, double hy, double hz) {
  if (u->dtype == OCU_F64) gradient_sub<double>(*u, *v, *w, *p, coefficient, hx, hy, hz);
  else gradient_sub<float>(*u, *v, *w, *p, (float)coefficient, hx, hy, hz);
  return 1;
}

int orc_self_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                    const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz,
                    double hx, double hy, double hz, int scheme) {
  if (u->dtype == OCU_F64) self_advect<double>(*u, *v, *w, *du, *dv, *dw, nx, ny, nz, hx, hy, hz, scheme);
  else self_advect<float>(*u, *v, *w, *du, *dv, *dw, nx, ny, nz, hx, hy, hz, scheme);
  return 1;
}

int orc_passive_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                       const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                       double hx, double hy, double hz, int scheme) {
  if (u->dtype == OCU_F64) passive_advect<double>(*u, *v, *w, *phi, *dphi, nx, ny, nz, hx, hy, hz, scheme);
  else passive_advect<float>(*u, *v, *w, *phi, *dphi, nx, ny, nz, hx, hy, hz, scheme);
  return 1;
}

int orc_self_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                       const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz,
                       double hx, double hy, double hz, int scheme, int periodic_axes) {
  if (u->dtype == OCU_F64)
    self_advect<double>(*u, *v, *w, *du, *dv, *dw, nx, ny, nz, hx, hy, hz, scheme, periodic_axes);
  else self_advect<float>(*u, *v, *w, *du, *dv, *dw, nx, ny, nz, hx, hy, hz, scheme, periodic_axes);
  return 1;
}

int orc_passive_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                          const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                          double hx, double hy, double hz, int scheme, int periodic_axes) {
  if (u->dtype == OCU_F64)
    passive_advect<double>(*u, *v, *w, *phi, *dphi, nx, ny, nz, hx, hy, hz, scheme, periodic_axes);
  else passive_advect<float>(*u, *v, *w, *phi, *dphi, nx, ny, nz, hx, hy, hz, scheme, periodic_axes);
  return 1;
}

int orc_laplacian_acc(const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                      double hx, double hy, double hz, double coefficient) {
  if (phi->dtype == OCU_F64) laplacian_acc<double>(*phi, *dphi, nx, ny, nz, hx, hy, hz, coefficient);
  else laplacian_acc<float>(*phi, *dphi, nx, ny, nz, hx, hy, hz, (float)coefficient);
  return 1;
}

int orc_thermal_force(const ocu_grid3d *d, const ocu_grid3d *t, int nx, int ny, int nz,
                      int vd, double coefficient) {
  int64_t stride = (vd & OCU_DIR_XAXIS_FLAG) ? (int64_t)t->pny * t->pnz : (vd & OCU_DIR_YAXIS_FLAG) ? t->pnz : 1;
  if (d->dtype == OCU_F64) thermal_force<double>(*d, *t, nx, ny, nz, stride, coefficient);
  else thermal_force<float>(*d, *t, nx, ny, nz, stride, (float)coefficient);
  return 1;
}

template <typename T>
static int project_t(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                     int nx, int ny, int nz, double hx, double hy, double hz, double tol,
                     const ocu_grid3d *pout, double *residual) {
  Projection<T> p;
  p.bc = *bc;
  if (!p.initialize_storage(nx, ny, nz, hx, hy, hz, *u, *v, *w)) return -1;
  bool ok = p.solve(tol);
  if (residual) *residual = p.last_residual;
  if (pout && pout->base) std::memcpy(pout->base, p.pressure.buf.data(), sizeof(T) * (size_t)pout->allocated);
  return ok ? 1 : 0;
}

int orc_project(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                int nx, int ny, int nz, double hx, double hy, double hz, double tol,
                const ocu_grid3d *pressure_out, double *residual) {
  if (u->dtype == OCU_F64) return project_t<double>(u, v, w, bc, nx, ny, nz, hx, hy, hz, tol, pressure_out, residual);
  return project_t<float>(u, v, w, bc, nx, ny, nz, hx, hy, hz, tol, pressure_out, residual);
}

struct NSHandle {
  int dtype;
  NS<float> *f = nullptr;
  NS<double> *d = nullptr;
};

static NSParams to_params(const orc_ns_params *p) {
  NSParams q;
  q.nx = p->nx; q.ny = p->ny; q.nz = p->nz;
  q.hx = p->hx; q.hy = p->hy; q.hz = p->hz;
  q.max_divergence = p->max_divergence;
  q.temp_bc = p->temp_bc; q.flow_bc = p->flow_bc;
  q.vertical_direction = p->vertical_direction;
  q.gravity = p->gravity; q.bouyancy = p->bouyancy;
  q.thermal_diffusion = p->thermal_diffusion; q.viscosity = p->viscosity;
  q.advection_scheme = p->advection_scheme; q.time_step = p->time_step;
  q.cfl_factor = p->cfl_factor;
  return q;
}

void *orc_ns_create(int dtype, const orc_ns_params *p, const ocu_grid3d *u, const ocu_grid3d *v,
                    const ocu_grid3d *w, const ocu_grid3d *t) {
  NSHandle *h = new NSHandle();
  h->dtype = dtype;
  bool ok;
  if (dtype == OCU_F64) { h->d = new NS<double>(); ok = h->d->set_parameters(to_params(p), *u, *v, *w, *t); }
  else { h->f = new NS<float>(); ok = h->f->set_parameters(to_params(p), *u, *v, *w, *t); }
  if (!ok) { delete h->d; delete h->f; delete h; return nullptr; }
  return h;
}

int orc_ns_advance_one_step(void *hh, double dt) {
  NSHandle *h = (NSHandle *)hh;
  return (h->d ? h->d->advance_one_step(dt) : h->f->advance_one_step(dt)) ? 1 : 0;
}

int orc_ns_advance(void *hh, double dt) {
  NSHandle *h = (NSHandle *)hh;
  return (h->d ? h->d->advance(dt) : h->f->advance(dt)) ? 1 : 0;
}

double orc_ns_max_stable_timestep(void *hh) {
  NSHandle *h = (NSHandle *)hh;
  return h->d ? h->d->get_max_stable_timestep() : h->f->get_max_stable_timestep();
}

template <typename T>
static int ns_get_t(NS<T> *s, int which, const ocu_grid3d *out) {
  const HostGrid<T> *g = which == 0 ? &s->u : which == 1 ? &s->v : which == 2 ? &s->w
                       : which == 3 ? &s->temp : &s->proj.pressure;
  if (!layout_match(g->desc, *out)) return -1;
  std::memcpy(out->base, g->buf.data(), sizeof(T) * (size_t)out->allocated);
  return 1;
}

int orc_ns_get(void *hh, int which, const ocu_grid3d *out) {
  NSHandle *h = (NSHandle *)hh;
  return h->d ? ns_get_t(h->d, which, out) : ns_get_t(h->f, which, out);
}

void orc_ns_destroy(void *hh) {
  NSHandle *h = (NSHandle *)hh;
  if (!h) return;
  delete h->d;
  delete h->f;
  delete h;
}

int orc_scalar_advect_step(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_grid3d *phi, const ocu_grid3d *dphi, const ocu_bc_set *bc,
                           int nx, int ny, int nz, double hx, double hy, double hz, int scheme,
                           double dt) {
  bool ok = phi->dtype == OCU_F64
                ? scalar_advect_step<double>(*u, *v, *w, *phi, *dphi, *bc, nx, ny, nz, hx, hy, hz, scheme, dt)
                : scalar_advect_step<float>(*u, *v, *w, *phi, *dphi, *bc, nx, ny, nz, hx, hy, hz, scheme, dt);
  return ok ? 1 : 0;
}


int orc_passive_advect_adjoint(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                               const ocu_grid3d *phi, const ocu_grid3d *lam, const ocu_grid3d *lam_out,
                               const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                               double scale, int nx, int ny, int nz, double hx, double hy, double hz,
                               int scheme) {
  if (!u || !v || !w || !phi || !lam || !lam_out || !adj_u || !adj_v || !adj_w) return -1;
  return DISPATCH(u->dtype,
                  (passive_advect_adjoint<float>(*u, *v, *w, *phi, *lam, *lam_out, *adj_u, *adj_v, *adj_w, scale, nx,
                                                 ny, nz, hx, hy, hz, scheme), 1),
                  (passive_advect_adjoint<double>(*u, *v, *w, *phi, *lam, *lam_out, *adj_u, *adj_v, *adj_w, scale,
                                                  nx, ny, nz, hx, hy, hz, scheme), 1));
}

int orc_bc_adjoint_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc) {
  if (!g || !bc) return -1;
  return DISPATCH(g->dtype, (bc_adjoint_nocorners<float>(*g, *bc), 1), (bc_adjoint_nocorners<double>(*g, *bc), 1));
}

int orc_mac_adjoint_normal(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc) {
  if (!u || !v || !w || !bc) return -1;
  return DISPATCH(u->dtype, (mac_adjoint_normal<float>(*u, *v, *w, *bc), 1),
                  (mac_adjoint_normal<double>(*u, *v, *w, *bc), 1));
}

int orc_divergence_adjoint(const ocu_grid3d *lam, const ocu_grid3d *adj_u, const ocu_grid3d *adj_v,
                           const ocu_grid3d *adj_w, int nx, int ny, int nz, double hx, double hy, double hz) {
  if (!lam || !adj_u || !adj_v || !adj_w) return -1;
  return DISPATCH(lam->dtype, (divergence_adjoint<float>(*lam, *adj_u, *adj_v, *adj_w, nx, ny, nz, hx, hy, hz), 1),
                  (divergence_adjoint<double>(*lam, *adj_u, *adj_v, *adj_w, nx, ny, nz, hx, hy, hz), 1));
}

int orc_gradient_adjoint(const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                         const ocu_grid3d *adj_p, double coefficient, double hx, double hy, double hz) {
  if (!adj_u || !adj_v || !adj_w || !adj_p) return -1;
  return DISPATCH(adj_p->dtype,
                  (gradient_adjoint<float>(*adj_u, *adj_v, *adj_w, *adj_p, (float)coefficient, hx, hy, hz), 1),
                  (gradient_adjoint<double>(*adj_u, *adj_v, *adj_w, *adj_p, (double)coefficient, hx, hy, hz), 1));
}

int orc_pcg_apply_laplacian(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fxh,
                            double fyh, double fzh) {
  if (!dst || !src || !coeff) return -1;
  return DISPATCH(src->dtype,
                  (pcg_apply_laplacian<float>(*dst, *src, *coeff, (float)fxh, (float)fyh, (float)fzh), 1),
                  (pcg_apply_laplacian<double>(*dst, *src, *coeff, fxh, fyh, fzh), 1));
}

int orc_pcg_diag_precond(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fxh,
                         double fyh, double fzh) {
  if (!dst || !src || !coeff) return -1;
  return DISPATCH(src->dtype, (pcg_diag_precond<float>(*dst, *src, *coeff, (float)fxh, (float)fyh, (float)fzh), 1),
                  (pcg_diag_precond<double>(*dst, *src, *coeff, fxh, fyh, fzh), 1));
}

double orc_pcg_dot(const ocu_grid3d *a, const ocu_grid3d *b) {
  return a->dtype == OCU_F64 ? (double)pcg_dot<double>(*a, *b) : (double)pcg_dot<float>(*a, *b);
}

template <typename T>
static int pcg_solve_t(const ocu_grid3d *rhs, const ocu_grid3d *coeff, const ocu_grid3d *pressure, float hx,
                       float hy, float hz, const ocu_bc_set *bc, int preconditioner, int cycles, int use_fmg,
                       int convergence, double tol, int max_iter, double *residual, int *iterations) {
  PCGSolver<T> s;
  s.bc = *bc;
  s.preconditioner = preconditioner;
  s.multigrid_cycles = cycles;
  s.multigrid_use_fmg = use_fmg != 0;
  s.convergence = convergence;
  if (!s.initialize_storage(rhs->nx, rhs->ny, rhs->nz, hx, hy, hz, *rhs, *coeff)) return -1;
  if (!layout_match(*pressure, s.pressure.desc)) return -1;
  std::memcpy(s.pressure.buf.data(), pressure->base, sizeof(T) * (size_t)pressure->allocated);
  double res = 0;
  bool ok = s.solve(res, tol, max_iter);
  std::memcpy(pressure->base, s.pressure.buf.data(), sizeof(T) * (size_t)pressure->allocated);
  if (residual) *residual = res;
  if (iterations) *iterations = s.iterations;
  return ok ? 1 : 0;
}

int orc_pcg_solve(const ocu_grid3d *rhs, const ocu_grid3d *coeff, const ocu_grid3d *pressure, float hx, float hy,
                  float hz, const ocu_bc_set *bc, int preconditioner, int multigrid_cycles, int use_fmg,
                  int convergence, double tol, int max_iter, double *residual, int *iterations) {
  if (!rhs || !coeff || !pressure || !bc) return -1;
  if (rhs->dtype != coeff->dtype || rhs->dtype != pressure->dtype) return -1;
  return DISPATCH(rhs->dtype,
                  pcg_solve_t<float>(rhs, coeff, pressure, hx, hy, hz, bc, preconditioner, multigrid_cycles, use_fmg,
                                     convergence, tol, max_iter, residual, iterations),
                  pcg_solve_t<double>(rhs, coeff, pressure, hx, hy, hz, bc, preconditioner, multigrid_cycles,
                                      use_fmg, convergence, tol, max_iter, residual, iterations));
}

int orc_bc_fill_level2_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz) {
  if (!g || !bc) return -1;
  return DISPATCH(g->dtype, bc_fill_level2_nocorners<float>(*g, *bc, (float)hx, (float)hy, (float)hz) ? 1 : 0,
                  bc_fill_level2_nocorners<double>(*g, *bc, hx, hy, hz) ? 1 : 0);
}

int orc_bc_fill_mac_level2(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc) {
  if (!u || !v || !w || !bc) return -1;
  return DISPATCH(u->dtype, bc_fill_mac_level2<float>(*u, *v, *w, *bc) ? 1 : 0,
                  bc_fill_mac_level2<double>(*u, *v, *w, *bc) ? 1 : 0);
}

int orc_sample_points(void *out, const float *px, const float *py, const float *pz, int n, const ocu_grid3d *phi,
                      const ocu_bc_set *bc, float period_nx, float period_ny, float period_nz, float hx, float hy,
                      float hz, float shiftx, float shifty, float shiftz) {
  if (!out || !phi || !bc) return -1;
  return DISPATCH(phi->dtype,
                  sample_points<float>((float *)out, px, py, pz, n, *phi, *bc, period_nx, period_ny, period_nz, hx,
                                       hy, hz, shiftx, shifty, shiftz) ? 1 : 0,
                  sample_points<double>((double *)out, px, py, pz, n, *phi, *bc, period_nx, period_ny, period_nz, hx,
                                        hy, hz, shiftx, shifty, shiftz) ? 1 : 0);
}

int orc_sample_points_mac(void *vx, void *vy, void *vz, const float *px, const float *py, const float *pz, int n,
                          const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                          float hx, float hy, float hz) {
  if (!vx || !vy || !vz || !u || !v || !w || !bc) return -1;
  return DISPATCH(u->dtype,
                  sample_points_mac<float>((float *)vx, (float *)vy, (float *)vz, px, py, pz, n, *u, *v, *w, *bc, hx,
                                           hy, hz) ? 1 : 0,
                  sample_points_mac<double>((double *)vx, (double *)vy, (double *)vz, px, py, pz, n, *u, *v, *w, *bc,
                                            hx, hy, hz) ? 1 : 0);
}
