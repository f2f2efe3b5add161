// Handle-based C ABI (include/ocu.h) over the drop-in classes.
#include "ocu.h"

#include <cstdio>
#include <cstring>
#include <new>

#include "ocu_hip.h"
#include "ocuequation/eqn_incompressns3d.h"
#include "ocuequation/eqn_scalaradvection3d.h"
#include "ocuequation/sol_mgpressure3d.h"
#include "ocuequation/sol_mgpressuremixed3d.h"
#include "ocuequation/sol_pcgpressure3d.h"
#include "ocustorage/gridnetcdf.h"
#include "ocuequation/sol_project3d.h"
#include "ocuutil/thread.h"

using namespace ocu;

namespace {

BoundaryCondition from_c1(const ocu_bc &b) {
  BoundaryCondition r;
  r.type = (BoundaryConditionType)b.type;
  r.value = b.value;
  r.aux_value = b.aux_value;
  return r;
}

BoundaryConditionSet from_c(const ocu_bc_set &b) {
  return BoundaryConditionSet(from_c1(b.xpos), from_c1(b.xneg), from_c1(b.ypos), from_c1(b.yneg), from_c1(b.zpos),
                              from_c1(b.zneg));
}

void apply_cfg(Sol_MultigridPressure3DBase &s, const ocu_mg_config *cfg) {
  s.bc = from_c(cfg->bc);
  s.nu1 = cfg->nu1;
  s.nu2 = cfg->nu2;
  s.convergence = (ConvergenceType)cfg->convergence;
  s.make_symmetric_operator = cfg->make_symmetric_operator != 0;
}

void copy_stats(const std::vector<double> &hist, double *history, int cap, int *len) {
  if (len) *len = (int)hist.size();
  if (history)
    for (int n = 0; n < (int)hist.size() && n < cap; n++) history[n] = hist[n];
}

struct MGHandle {
  int dtype;
  Grid3DDevice<float> rhs_f;
  Grid3DDevice<double> rhs_d;
  Sol_MultigridPressure3DDevice<float> *f = nullptr;
  Sol_MultigridPressure3DDevice<double> *d = nullptr;
  ~MGHandle() {
    delete f;
    delete d;
  }
  Sol_MultigridPressure3DBase &base() { return d ? (Sol_MultigridPressure3DBase &)*d : *f; }
};

struct MixedHandle {
  Grid3DDevice<double> rhs;
  Sol_MultigridPressureMixed3DDeviceD s;
};

struct ProjHandle {
  int dtype, mixed;
  Grid3DDevice<float> uf, vf, wf;
  Grid3DDevice<double> ud, vd, wd;
  Sol_ProjectDivergence3DDevice<float> *pf = nullptr;
  Sol_ProjectDivergence3DDevice<double> *pd = nullptr;
  Sol_ProjectDivergenceMixed3DDeviceD *pm = nullptr;
  ~ProjHandle() {
    delete pf;
    delete pd;
    delete pm;
  }
};

struct NSHandle {
  int dtype;
  Eqn_IncompressibleNS3D<float> *f = nullptr;
  Eqn_IncompressibleNS3D<double> *d = nullptr;
  ~NSHandle() {
    delete f;
    delete d;
  }
};

// A host descriptor handed in for one of the init grids must be exactly the
// layout and type the params' init_grids gave that grid: its buffer is
// copied over the whole allocation.
template <typename T>
bool host_grid_from(Grid3DHost<T> &dst, const ocu_grid3d *src, const char *what) {
  ocu_grid3d want = dst.desc();
  if (!src || !src->base || src->dtype != want.dtype || src->nx != want.nx || src->ny != want.ny ||
      src->nz != want.nz || src->gx != want.gx || src->gy != want.gy || src->gz != want.gz ||
      src->pnx != want.pnx || src->pny != want.pny || src->pnz != want.pnz || src->shift != want.shift ||
      src->allocated != want.allocated) {
    printf("[ERROR] %s - descriptor does not match init_grids' layout (%d,%d,%d) padded (%d,%d,%d) dtype %d\n",
           what, want.nx, want.ny, want.nz, want.pnx, want.pny, want.pnz, want.dtype);
    return false;
  }
  memcpy(dst.buffer(), src->base, (size_t)want.allocated * sizeof(T));
  return true;
}

template <typename T>
bool fill_params(Eqn_IncompressibleNS3DParams<T> &P, const ocu_ns_params *p, const ocu_grid3d *iu,
                 const ocu_grid3d *iv, const ocu_grid3d *iw, const ocu_grid3d *it) {
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return false;
  P.init_grids(p->nx, p->ny, p->nz, true);
  P.hx = p->hx; P.hy = p->hy; P.hz = p->hz;
  P.max_divergence = p->max_divergence;
  P.temp_bc = from_c(p->temp_bc);
  P.flow_bc = from_c(p->flow_bc);
  P.vertical_direction = (DirectionType)p->vertical_direction;
  P.gravity = (T)p->gravity;
  P.bouyancy = (T)p->bouyancy;
  P.thermal_diffusion = (T)p->thermal_diffusion;
  P.viscosity = (T)p->viscosity;
  P.advection_scheme = (InterpolationType)p->advection_scheme;
  P.time_step = (TimeStepType)p->time_step;
  P.cfl_factor = p->cfl_factor;
  P.mixed_precision_projection = p->mixed_projection != 0;
  P.fused_step = p->unfused_step == 0;
  return host_grid_from(P.init_u, iu, "ocu_ns_create init_u") && host_grid_from(P.init_v, iv, "ocu_ns_create init_v") &&
         host_grid_from(P.init_w, iw, "ocu_ns_create init_w") &&
         host_grid_from(P.init_temp, it, "ocu_ns_create init_temp");
}

template <typename T>
bool fill_scalar_params(Eqn_ScalarAdvection3DParams<T> &P, const ocu_scalar_params *p, const ocu_grid3d *iu,
                        const ocu_grid3d *iv, const ocu_grid3d *iw, const ocu_grid3d *iphi) {
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0) return false;
  P.init_grids(p->nx, p->ny, p->nz, true);
  P.hx = p->hx; P.hy = p->hy; P.hz = p->hz;
  P.bc = from_c(p->bc);
  P.advection_scheme = (InterpolationType)p->advection_scheme;
  return host_grid_from(P.u, iu, "ocu_scalar_create init_u") && host_grid_from(P.v, iv, "ocu_scalar_create init_v") &&
         host_grid_from(P.w, iw, "ocu_scalar_create init_w") &&
         host_grid_from(P.initial_values, iphi, "ocu_scalar_create init_phi");
}

struct ScalarHandle {
  Eqn_ScalarAdvection3D<float> *f = nullptr;
  Eqn_ScalarAdvection3D<double> *d = nullptr;
  ~ScalarHandle() {
    delete f;
    delete d;
  }
};

}  // namespace

extern "C" {

int ocu_set_compute_stream(void *stream) {
  ThreadManager::set_compute_stream(stream);
  return OCU_OK;
}

void *ocu_get_compute_stream(void) { return ThreadManager::get_compute_stream(); }

// ---------------------------------------------------------------- multigrid

void *ocu_mg_create(const ocu_grid3d *rhs, double hx, double hy, double hz, const ocu_mg_config *cfg,
                    int options) {
  if (!rhs || !cfg || (rhs->dtype != OCU_F32 && rhs->dtype != OCU_F64)) return nullptr;
  MGHandle *h = new (std::nothrow) MGHandle();
  if (!h) return nullptr;
  h->dtype = rhs->dtype;
  bool ok;
  if (rhs->dtype == OCU_F64) {
    h->d = new Sol_MultigridPressure3DDevice<double>();
    apply_cfg(*h->d, cfg);
    h->d->disable_relax_on_host = cfg->disable_relax_on_host != 0;
    h->d->use_fused_relax = (options & OCU_MG_FUSED_RELAX) != 0;
    h->d->use_bottom = (options & OCU_MG_LDS_BOTTOM) != 0;
    h->d->use_fused_prolong = (options & OCU_MG_FUSED_PROLONG) != 0;
    h->d->use_graphs = (options & OCU_MG_GRAPHS) != 0;
    ok = h->rhs_d.init_external(*rhs) &&
         h->d->initialize_storage(rhs->nx, rhs->ny, rhs->nz, hx, hy, hz, &h->rhs_d);
  } else {
    h->f = new Sol_MultigridPressure3DDevice<float>();
    apply_cfg(*h->f, cfg);
    h->f->disable_relax_on_host = cfg->disable_relax_on_host != 0;
    h->f->use_fused_relax = (options & OCU_MG_FUSED_RELAX) != 0;
    h->f->use_bottom = (options & OCU_MG_LDS_BOTTOM) != 0;
    h->f->use_fused_prolong = (options & OCU_MG_FUSED_PROLONG) != 0;
    h->f->use_graphs = (options & OCU_MG_GRAPHS) != 0;
    ok = h->rhs_f.init_external(*rhs) &&
         h->f->initialize_storage(rhs->nx, rhs->ny, rhs->nz, hx, hy, hz, &h->rhs_f);
  }
  if (!ok) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_mg_pressure(void *hh, ocu_grid3d *out) {
  MGHandle *h = (MGHandle *)hh;
  if (!h || !out) return OCU_ERR_INVALID;
  *out = h->d ? h->d->pressure().desc() : h->f->pressure().desc();
  return OCU_OK;
}

int ocu_mg_solve(void *hh, double tol, int max_iter, double *residual) {
  MGHandle *h = (MGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  double r = 0;
  bool ok = h->base().solve(r, tol, max_iter);
  if (residual) *residual = r;
  return ok ? 1 : 0;
}

int ocu_mg_run_vcycles(void *hh, int count) {
  MGHandle *h = (MGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return h->base().run_vcycles(count) ? 1 : 0;
}

int ocu_mg_run_fmg(void *hh, int count) {
  MGHandle *h = (MGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return h->base().run_fmg(count) ? 1 : 0;
}

int ocu_mg_stats(void *hh, int *num_levels, int *vcycles, double *history, int cap, int *len) {
  MGHandle *h = (MGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  Sol_MultigridPressure3DBase &b = h->base();
  if (num_levels) *num_levels = b.num_levels();
  if (vcycles) *vcycles = b.num_vcycles;
  copy_stats(b.residual_history, history, cap, len);
  return OCU_OK;
}

void ocu_mg_destroy(void *hh) { delete (MGHandle *)hh; }

int ocu_mg_get_level_params(void *hh, int level, ocu_mg_level_params *out) {
  MGHandle *h = (MGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return (h->d ? h->d->get_level_params(level, out) : h->f->get_level_params(level, out)) ? OCU_OK
                                                                                          : OCU_ERR_INVALID;
}

int ocu_mg_sweep_counts(void *hh, long long *counts) {
  MGHandle *h = (MGHandle *)hh;
  if (!h || !counts) return OCU_ERR_INVALID;
  const long long *c = h->d ? h->d->level0_sweeps : h->f->level0_sweeps;
  for (int q = 0; q < 3; q++) counts[q] = c[q];
  return OCU_OK;
}

int ocu_mg_graph_counts(void *hh, long long *counts) {
  MGHandle *h = (MGHandle *)hh;
  if (!h || !counts) return OCU_ERR_INVALID;
  counts[0] = h->d ? h->d->graph_records : h->f->graph_records;
  counts[1] = h->d ? h->d->graph_replays : h->f->graph_replays;
  return OCU_OK;
}

// ------------------------------------------------------------------ mixed IR

void *ocu_mgmixed_create(const ocu_grid3d *rhs, double hx, double hy, double hz, const ocu_mg_config *cfg,
                         int cycles, int cycles_first, int fmg_first, int options) {
  if (!rhs || !cfg || rhs->dtype != OCU_F64) return nullptr;
  MixedHandle *h = new (std::nothrow) MixedHandle();
  if (!h) return nullptr;
  apply_cfg(h->s, cfg);
  h->s.ir_cycles = cycles;
  h->s.ir_cycles_first = cycles_first;
  h->s.ir_fmg_first = fmg_first != 0;
  h->s.use_fused_relax = (options & OCU_MG_FUSED_RELAX) != 0;
  h->s.use_bottom = (options & OCU_MG_LDS_BOTTOM) != 0;
  h->s.use_fused_prolong = (options & OCU_MG_FUSED_PROLONG) != 0;
  h->s.ir_adaptive = (options & OCU_MG_ADAPTIVE_IR) != 0;
  if (!h->rhs.init_external(*rhs) || !h->s.initialize_storage(rhs->nx, rhs->ny, rhs->nz, hx, hy, hz, &h->rhs)) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_mgmixed_pressure(void *hh, ocu_grid3d *out) {
  MixedHandle *h = (MixedHandle *)hh;
  if (!h || !out) return OCU_ERR_INVALID;
  *out = h->s.pressure().desc();
  return OCU_OK;
}

int ocu_mgmixed_solve(void *hh, double tol, int max_iter, double *residual) {
  MixedHandle *h = (MixedHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  double r = 0;
  bool ok = h->s.solve(r, tol, max_iter);
  if (residual) *residual = r;
  return ok ? 1 : 0;
}

int ocu_mgmixed_stats(void *hh, int *ir_iterations, int *vcycles, double *rhs_rms, double *history, int cap,
                      int *len) {
  MixedHandle *h = (MixedHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  if (ir_iterations) *ir_iterations = h->s.ir_iterations;
  if (vcycles) *vcycles = h->s.total_vcycles;
  if (rhs_rms) *rhs_rms = h->s.rhs_l2;
  copy_stats(h->s.residual_history, history, cap, len);
  return OCU_OK;
}

void ocu_mgmixed_destroy(void *hh) { delete (MixedHandle *)hh; }

int ocu_mgmixed_get_level_params(void *hh, int level, ocu_mg_level_params *out) {
  MixedHandle *h = (MixedHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return h->s.get_level_params(level, out) ? OCU_OK : OCU_ERR_INVALID;
}

int ocu_mgmixed_sweep_counts(void *hh, long long *counts) {
  MixedHandle *h = (MixedHandle *)hh;
  if (!h || !counts) return OCU_ERR_INVALID;
  const long long *c = h->s.level0_sweeps();
  for (int q = 0; q < 3; q++) counts[q] = c[q];
  return OCU_OK;
}

// ---------------------------------------------------------------- projection

void *ocu_project_create(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                         int nx, int ny, int nz, double hx, double hy, double hz, int mixed) {
  if (!u || !v || !w || !bc) return nullptr;
  ProjHandle *h = new (std::nothrow) ProjHandle();
  if (!h) return nullptr;
  h->dtype = u->dtype;
  h->mixed = mixed;
  bool ok = false;
  if (u->dtype == OCU_F64) {
    ok = h->ud.init_external(*u) && h->vd.init_external(*v) && h->wd.init_external(*w);
    if (ok && mixed) {
      h->pm = new Sol_ProjectDivergenceMixed3DDeviceD();
      h->pm->bc = from_c(*bc);
      ok = h->pm->initialize_storage(nx, ny, nz, hx, hy, hz, &h->ud, &h->vd, &h->wd);
    } else if (ok) {
      h->pd = new Sol_ProjectDivergence3DDevice<double>();
      h->pd->bc = from_c(*bc);
      ok = h->pd->initialize_storage(nx, ny, nz, hx, hy, hz, &h->ud, &h->vd, &h->wd);
    }
  } else if (u->dtype == OCU_F32 && !mixed) {
    ok = h->uf.init_external(*u) && h->vf.init_external(*v) && h->wf.init_external(*w);
    if (ok) {
      h->pf = new Sol_ProjectDivergence3DDevice<float>();
      h->pf->bc = from_c(*bc);
      ok = h->pf->initialize_storage(nx, ny, nz, hx, hy, hz, &h->uf, &h->vf, &h->wf);
    }
  }
  if (!ok) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_project_solve(void *hh, double tol, double *residual) {
  ProjHandle *h = (ProjHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  bool ok;
  double r;
  if (h->pm) { ok = h->pm->solve(tol); r = h->pm->last_residual; }
  else if (h->pd) { ok = h->pd->solve(tol); r = h->pd->last_residual; }
  else { ok = h->pf->solve(tol); r = h->pf->last_residual; }
  if (residual) *residual = r;
  return ok ? 1 : 0;
}

int ocu_project_solve_divergence_only(void *hh) {
  ProjHandle *h = (ProjHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  bool ok = h->pm ? h->pm->solve_divergence_only() : h->pd ? h->pd->solve_divergence_only()
                                                           : h->pf->solve_divergence_only();
  return ok ? 1 : 0;
}

int ocu_project_divergence(void *hh, ocu_grid3d *out) {
  ProjHandle *h = (ProjHandle *)hh;
  if (!h || !out) return OCU_ERR_INVALID;
  *out = h->pm ? h->pm->divergence.desc() : h->pd ? h->pd->divergence.desc() : h->pf->divergence.desc();
  return OCU_OK;
}

int ocu_project_pressure(void *hh, ocu_grid3d *out) {
  ProjHandle *h = (ProjHandle *)hh;
  if (!h || !out) return OCU_ERR_INVALID;
  *out = h->pm ? h->pm->pressure_solver.pressure().desc()
         : h->pd ? h->pd->pressure_solver.pressure().desc()
                 : h->pf->pressure_solver.pressure().desc();
  return OCU_OK;
}

void ocu_project_destroy(void *hh) { delete (ProjHandle *)hh; }

// ------------------------------------------------------------------- NS eqn

void *ocu_ns_create(int dtype, const ocu_ns_params *p, const ocu_grid3d *iu, const ocu_grid3d *iv,
                    const ocu_grid3d *iw, const ocu_grid3d *it) {
  if (!p || !iu || !iv || !iw || !it) return nullptr;
  NSHandle *h = new (std::nothrow) NSHandle();
  if (!h) return nullptr;
  h->dtype = dtype;
  bool ok = false;
  if (dtype == OCU_F64) {
    Eqn_IncompressibleNS3DParams<double> P;
    if (fill_params(P, p, iu, iv, iw, it)) {
      h->d = new Eqn_IncompressibleNS3D<double>();
      ok = h->d->set_parameters(P);
    }
  } else if (dtype == OCU_F32) {
    Eqn_IncompressibleNS3DParams<float> P;
    if (fill_params(P, p, iu, iv, iw, it)) {
      h->f = new Eqn_IncompressibleNS3D<float>();
      ok = h->f->set_parameters(P);
    }
  }
  if (!ok) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_ns_advance_one_step(void *hh, double dt) {
  NSHandle *h = (NSHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return (h->d ? h->d->advance_one_step(dt) : h->f->advance_one_step(dt)) ? 1 : 0;
}

int ocu_ns_advance(void *hh, double dt) {
  NSHandle *h = (NSHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return (h->d ? h->d->advance(dt) : h->f->advance(dt)) ? 1 : 0;
}

int ocu_ns_max_stable_timestep(void *hh, double *dt) {
  NSHandle *h = (NSHandle *)hh;
  if (!h || !dt) return OCU_ERR_INVALID;
  *dt = h->d ? h->d->get_max_stable_timestep() : h->f->get_max_stable_timestep();
  return OCU_OK;
}

int ocu_ns_field(void *hh, int which, ocu_grid3d *out) {
  NSHandle *h = (NSHandle *)hh;
  if (!h || !out || which < 0 || which > 4) return OCU_ERR_INVALID;
  if (h->d) {
    *out = which == 0 ? h->d->get_u().desc() : which == 1 ? h->d->get_v().desc() : which == 2 ? h->d->get_w().desc()
           : which == 3 ? h->d->get_temperature().desc() : h->d->get_pressure().desc();
  } else {
    *out = which == 0 ? h->f->get_u().desc() : which == 1 ? h->f->get_v().desc() : which == 2 ? h->f->get_w().desc()
           : which == 3 ? h->f->get_temperature().desc() : h->f->get_pressure().desc();
  }
  return OCU_OK;
}

int ocu_ns_num_steps(void *hh) {
  NSHandle *h = (NSHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return h->d ? h->d->num_steps : h->f->num_steps;
}

void ocu_ns_destroy(void *hh) { delete (NSHandle *)hh; }

// ------------------------------------------------------- scalar advection

void *ocu_scalar_create(int dtype, const ocu_scalar_params *p, const ocu_grid3d *iu, const ocu_grid3d *iv,
                        const ocu_grid3d *iw, const ocu_grid3d *iphi) {
  if (!p) return nullptr;
  ScalarHandle *h = new (std::nothrow) ScalarHandle();
  if (!h) return nullptr;
  bool ok = false;
  if (dtype == OCU_F64) {
    Eqn_ScalarAdvection3DParams<double> P;
    if (fill_scalar_params(P, p, iu, iv, iw, iphi)) {
      h->d = new Eqn_ScalarAdvection3D<double>();
      ok = h->d->set_parameters(P);
    }
  } else if (dtype == OCU_F32) {
    Eqn_ScalarAdvection3DParams<float> P;
    if (fill_scalar_params(P, p, iu, iv, iw, iphi)) {
      h->f = new Eqn_ScalarAdvection3D<float>();
      ok = h->f->set_parameters(P);
    }
  }
  if (!ok) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_scalar_advance_one_step(void *hh, double dt) {
  ScalarHandle *h = (ScalarHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return (h->d ? h->d->advance_one_step(dt) : h->f->advance_one_step(dt)) ? 1 : 0;
}

int ocu_scalar_advance(void *hh, double dt) {
  ScalarHandle *h = (ScalarHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return (h->d ? h->d->advance(dt) : h->f->advance(dt)) ? 1 : 0;
}

int ocu_scalar_max_stable_timestep(void *hh, double *dt) {
  ScalarHandle *h = (ScalarHandle *)hh;
  if (!h || !dt) return OCU_ERR_INVALID;
  *dt = h->d ? h->d->get_max_stable_timestep() : h->f->get_max_stable_timestep();
  return OCU_OK;
}

int ocu_scalar_field(void *hh, int which, ocu_grid3d *out) {
  ScalarHandle *h = (ScalarHandle *)hh;
  if (!h || !out || which < 0 || which > 4) return OCU_ERR_INVALID;
  if (h->d) {
    *out = which == 0 ? h->d->u.desc() : which == 1 ? h->d->v.desc() : which == 2 ? h->d->w.desc()
           : which == 3 ? h->d->phi.desc() : h->d->deriv_phidt.desc();
  } else {
    *out = which == 0 ? h->f->u.desc() : which == 1 ? h->f->v.desc() : which == 2 ? h->f->w.desc()
           : which == 3 ? h->f->phi.desc() : h->f->deriv_phidt.desc();
  }
  return OCU_OK;
}

void ocu_scalar_destroy(void *hh) { delete (ScalarHandle *)hh; }

}  // extern "C"

// ---------------------------------------------------------------- PCG

namespace {
struct PCGHandle {
  Sol_PCGPressure3DDevice<double> *d = nullptr;
  Sol_PCGPressure3DDevice<float> *f = nullptr;
  Grid3DDevice<double> rhs_d, coeff_d;
  Grid3DDevice<float> rhs_f, coeff_f;
  ~PCGHandle() {
    delete d;
    delete f;
  }
};

template <typename T>
bool pcg_init(Sol_PCGPressure3DDevice<T> &s, Grid3DDevice<T> &rhs, Grid3DDevice<T> &coeff, const ocu_grid3d *r,
              const ocu_grid3d *c, float hx, float hy, float hz, const ocu_bc_set *bc, int convergence,
              int preconditioner, int cycles, int use_fmg) {
  s.bc = from_c(*bc);
  s.convergence = (ConvergenceType)convergence;
  s.preconditioner = (PreconditionerType)preconditioner;
  s.multigrid_cycles = cycles;
  s.multigrid_use_fmg = use_fmg != 0;
  return rhs.init_external(*r) && coeff.init_external(*c) &&
         s.initialize_storage(r->nx, r->ny, r->nz, hx, hy, hz, &rhs, &coeff);
}
}  // namespace

extern "C" {

void *ocu_pcg_create(const ocu_grid3d *rhs, const ocu_grid3d *coeff, float hx, float hy, float hz,
                     const ocu_bc_set *bc, int convergence, int preconditioner, int multigrid_cycles,
                     int multigrid_use_fmg) {
  if (!rhs || !coeff || !bc || (rhs->dtype != OCU_F32 && rhs->dtype != OCU_F64) || coeff->dtype != rhs->dtype)
    return nullptr;
  if (preconditioner < 0 || preconditioner > 3) return nullptr;
  PCGHandle *h = new (std::nothrow) PCGHandle();
  if (!h) return nullptr;
  bool ok;
  if (rhs->dtype == OCU_F64) {
    h->d = new Sol_PCGPressure3DDevice<double>();
    ok = pcg_init(*h->d, h->rhs_d, h->coeff_d, rhs, coeff, hx, hy, hz, bc, convergence, preconditioner,
                  multigrid_cycles, multigrid_use_fmg);
  } else {
    h->f = new Sol_PCGPressure3DDevice<float>();
    ok = pcg_init(*h->f, h->rhs_f, h->coeff_f, rhs, coeff, hx, hy, hz, bc, convergence, preconditioner,
                  multigrid_cycles, multigrid_use_fmg);
  }
  if (!ok) {
    delete h;
    return nullptr;
  }
  return h;
}

int ocu_pcg_solve(void *hh, double tol, int max_iter, double *residual) {
  PCGHandle *h = (PCGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  double r = 0;
  bool ok = h->d ? h->d->solve(r, tol, max_iter) : h->f->solve(r, tol, max_iter);
  if (residual) *residual = r;
  return ok ? 1 : 0;
}

int ocu_pcg_iterations(void *hh) {
  PCGHandle *h = (PCGHandle *)hh;
  if (!h) return OCU_ERR_INVALID;
  return h->d ? h->d->iterations() : h->f->iterations();
}

int ocu_pcg_pressure(void *hh, ocu_grid3d *out) {
  PCGHandle *h = (PCGHandle *)hh;
  if (!h || !out) return OCU_ERR_INVALID;
  *out = h->d ? h->d->pressure().desc() : h->f->pressure().desc();
  return OCU_OK;
}

void ocu_pcg_destroy(void *hh) { delete (PCGHandle *)hh; }

}  // extern "C"

// ---------------------------------------------------------------- NetCDF

namespace {
// the layout of a descriptor, for the writer's and reader's strided access
struct DescLayout : public Grid3DUntyped {
  explicit DescLayout(const ocu_grid3d &d) {
    _nx = d.nx, _ny = d.ny, _nz = d.nz;
    _gx = d.gx, _gy = d.gy, _gz = d.gz;
    _pnx = d.pnx, _pny = d.pny, _pnz = d.pnz;
    _pnzpny = d.pnz * d.pny;
    _shift_amount = (int)d.shift;
    _allocated_elements = (int)d.allocated;
  }
};
nc_type nc_of(int dtype) { return dtype == OCU_F64 ? NC_DOUBLE : dtype == OCU_F32 ? NC_FLOAT : NC_INT; }
size_t atom_of(int dtype) { return dtype == OCU_F64 ? 8 : 4; }
}  // namespace

extern "C" {

void *ocu_nc_writer_open(const char *filename, int nx, int ny, int nz, float hx, float hy, float hz) {
  NetCDFGrid3DWriter *w = new (std::nothrow) NetCDFGrid3DWriter();
  if (!w) return nullptr;
  if (!w->open(filename, nx, ny, nz, hx, hy, hz)) {
    delete w;
    return nullptr;
  }
  return w;
}

int ocu_nc_writer_define_variable(void *h, const char *name, int nc_type_, int staggering, float offsetx,
                                  float offsety, float offsetz) {
  if (!h || staggering < 0 || staggering > 4) return OCU_ERR_INVALID;
  return ((NetCDFGrid3DWriter *)h)->define_variable(name, nc_type_, (GridStaggering)staggering, offsetx, offsety,
                                                    offsetz)
             ? OCU_OK
             : OCU_ERR_INVALID;
}

int ocu_nc_writer_add_time_level(void *h, float time, int *level) {
  if (!h) return OCU_ERR_INVALID;
  size_t l = 0;
  if (!((NetCDFGrid3DWriter *)h)->add_time_level(time, l)) return OCU_ERR_INVALID;
  if (level) *level = (int)l;
  return OCU_OK;
}

int ocu_nc_writer_add_data(void *h, const char *name, const ocu_grid3d *host_grid, int time_level) {
  if (!h || !host_grid || !host_grid->base) return OCU_ERR_INVALID;
  DescLayout L(*host_grid);
  const void *interior = (const char *)host_grid->base + atom_of(host_grid->dtype) * (size_t)host_grid->shift;
  return ((NetCDFGrid3DWriter *)h)->add_data_layout(name, L, nc_of(host_grid->dtype), interior, time_level)
             ? OCU_OK
             : OCU_ERR_INVALID;
}

int ocu_nc_writer_close(void *h) {
  if (!h) return OCU_ERR_INVALID;
  NetCDFGrid3DWriter *w = (NetCDFGrid3DWriter *)h;
  const bool ok = w->close();
  delete w;
  return ok ? OCU_OK : OCU_ERR_RUNTIME;
}

void *ocu_nc_reader_open(const char *filename) {
  NetCDFGrid3DReader *r = new (std::nothrow) NetCDFGrid3DReader();
  if (!r) return nullptr;
  if (!r->open(filename)) {
    delete r;
    return nullptr;
  }
  return r;
}

int ocu_nc_reader_info(void *h, int *nx, int *ny, int *nz, float *hx, float *hy, float *hz, int *num_time_levels) {
  if (!h) return OCU_ERR_INVALID;
  NetCDFGrid3DReader *r = (NetCDFGrid3DReader *)h;
  if (nx) *nx = r->nx();
  if (ny) *ny = r->ny();
  if (nz) *nz = r->nz();
  if (hx) *hx = r->hx();
  if (hy) *hy = r->hy();
  if (hz) *hz = r->hz();
  if (num_time_levels) *num_time_levels = r->num_time_levels();
  return OCU_OK;
}

int ocu_nc_reader_time(void *h, int time_level, float *time) {
  NetCDFGrid3DReader *r = (NetCDFGrid3DReader *)h;
  if (!r || !time || time_level < 0 || time_level >= r->num_time_levels()) return OCU_ERR_INVALID;
  *time = r->get_time(time_level);
  return OCU_OK;
}

int ocu_nc_reader_variable(void *h, int index, char *name, int name_cap, int *nc_type_, int *nx, int *ny, int *nz) {
  NetCDFGrid3DReader *r = (NetCDFGrid3DReader *)h;
  if (!r) return OCU_ERR_INVALID;
  const std::vector<std::string> names = r->list_variables();
  if (index < 0) return (int)names.size();
  if (index >= (int)names.size() || !name || name_cap <= (int)names[index].size()) return OCU_ERR_INVALID;
  std::memcpy(name, names[index].c_str(), names[index].size() + 1);
  nc_type t = 0;
  int a = 0, b = 0, c = 0;
  r->variable_type(names[index].c_str(), t);
  r->variable_size(names[index].c_str(), a, b, c);
  if (nc_type_) *nc_type_ = t;
  if (nx) *nx = a;
  if (ny) *ny = b;
  if (nz) *nz = c;
  return OCU_OK;
}

int ocu_nc_reader_read(void *h, const char *name, const ocu_grid3d *host_grid, int time_level) {
  if (!h || !host_grid || !host_grid->base) return OCU_ERR_INVALID;
  DescLayout L(*host_grid);
  void *interior = (char *)host_grid->base + atom_of(host_grid->dtype) * (size_t)host_grid->shift;
  return ((NetCDFGrid3DReader *)h)->read_variable_layout(name, L, nc_of(host_grid->dtype), interior, time_level)
             ? OCU_OK
             : OCU_ERR_INVALID;
}

void ocu_nc_reader_close(void *h) {
  NetCDFGrid3DReader *r = (NetCDFGrid3DReader *)h;
  if (!r) return;
  if (r->is_open()) r->close();
  delete r;
}

}  // extern "C"
