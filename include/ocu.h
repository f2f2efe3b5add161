/*
 * ocu.h - C ABI of libocu.so, the host drop-in classes (namespace ocu, the
 * reference's class names and signatures) behind opaque handles, for callers
 * that are not C++ (the Python package binds these with ctypes).
 *
 * Grids passed in are device grids owned by the caller (ocu_grid3d with a
 * device base pointer and the Grid3D layout); they must outlive the handle,
 * as the reference solvers keep raw pointers to their inputs
 * (sol_mgpressure3ddev.cu:418, sol_divergence3ddev.cu:95-98).  Grids handed
 * out (pressure, divergence, fields) are owned by the handle.
 *
 * Return values: handles are NULL on failure; ints are OCU_OK (0) or a
 * negative ocu_status, except *_solve which return 1 (converged / true),
 * 0 (not converged / false) or a negative status.  ocu_last_error() (from
 * ocu_hip.h) holds the message of a negative status.
 */
#ifndef OCU_H
#define OCU_H

#include "ocu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* stream every solver enqueues on (ThreadManager::get_compute_stream) */
int ocu_set_compute_stream(void *stream);
void *ocu_get_compute_stream(void);

/* options of the multigrid solvers (OR-ed; OCU_MG_DEFAULT = all but
 * OCU_MG_GRAPHS):
 *   OCU_MG_FUSED_RELAX: one launch per red+black iteration, ghosts written by
 *     the sweep, deferred zero clears (Neumann/Dirichlet faces);
 *   OCU_MG_LDS_BOTTOM: the coarse levels of each V-cycle as one launch of a
 *     single LDS-resident workgroup;
 *   OCU_MG_FUSED_PROLONG (with OCU_MG_FUSED_RELAX): a prolongation into a
 *     level with rows of 129..256 cells is added by the fused sweep that
 *     follows it instead of a launch of its own;
 *   OCU_MG_GRAPHS (FP64/FP32 solver): the launches between two residual
 *     read-backs of the finest V-cycles are recorded once as a HIP graph
 *     and replayed while they repeat (the compute stream must not be the
 *     default stream; opt-in: the replay measured no faster than the
 *     launches on MI355X);
 *   (these four give results bit-identical to the plain per-level path)
 *   OCU_MG_ADAPTIVE_IR (mixed solver): after the first refinement iteration,
 *     as many V-cycles (at most `cycles`) as the observed contraction per
 *     cycle needs to reach half the tolerance. */
enum ocu_mg_options {
  OCU_MG_FUSED_RELAX = 1,
  OCU_MG_LDS_BOTTOM = 2,
  OCU_MG_ADAPTIVE_IR = 4,
  OCU_MG_FUSED_PROLONG = 8,
  OCU_MG_GRAPHS = 16,
  OCU_MG_DEFAULT = 15
};

/* ---- Sol_MultigridPressure3DDevice<T> (ocuequation/sol_mgpressure3d.h:36-181) ----
 * T follows rhs->dtype.  initialize_storage(nx,ny,nz,hx,hy,hz,&rhs) with the
 * public fields taken from cfg. */
void *ocu_mg_create(const ocu_grid3d *rhs, double hx, double hy, double hz, const ocu_mg_config *cfg,
                    int options);
int ocu_mg_pressure(void *h, ocu_grid3d *out);
/* solve(residual, tol, max_iter) (sol_mgpressure3d.cpp:332-370) */
int ocu_mg_solve(void *h, double tol, int max_iter, double *residual);
int ocu_mg_run_vcycles(void *h, int count);
int ocu_mg_run_fmg(void *h, int count);
/* level count, V-cycles of the last call and the level-0 residual history */
int ocu_mg_stats(void *h, int *num_levels, int *vcycles, double *history, int cap, int *len);
void ocu_mg_destroy(void *h);
/* fused red+black sweeps launched at level 0 since the handle was created,
 * by kind: counts[0] plain, [1] starting from the deferred zero grid, [2]
 * adding the pending prolongation (profiling: the kinds run at different
 * speeds); ocu_mgmixed_sweep_counts: those of the FP32 hierarchy */
int ocu_mg_sweep_counts(void *h, long long *counts);
/* HIP graphs of V-cycle segments since the handle was created: counts[0]
 * recorded, counts[1] replayed */
int ocu_mg_graph_counts(void *h, long long *counts);
/* relaxation parameters of a level as the solver hands them to the kernels
 * (ocu_mgmixed_get_level_params: the FP32 hierarchy's) */
int ocu_mg_get_level_params(void *h, int level, ocu_mg_level_params *out);

/* ---- Sol_MultigridPressureMixed3DDeviceD: FP32 V-cycles in FP64 iterative
 * refinement behind the reference's mixed-solver API
 * (ocuequation/sol_mgpressuremixed3d.h).  rhs is FP64.  cycles /
 * cycles_first are V-cycles per refinement iteration; fmg_first starts the
 * first iteration with an FP32 FMG. */
void *ocu_mgmixed_create(const ocu_grid3d *rhs, double hx, double hy, double hz, const ocu_mg_config *cfg,
                         int cycles, int cycles_first, int fmg_first, int options);
int ocu_mgmixed_pressure(void *h, ocu_grid3d *out);
/* tol is absolute in cfg->convergence's norm; max_iter bounds the
 * refinement iterations */
int ocu_mgmixed_solve(void *h, double tol, int max_iter, double *residual);
int ocu_mgmixed_stats(void *h, int *ir_iterations, int *vcycles, double *rhs_rms, double *history, int cap,
                      int *len);
void ocu_mgmixed_destroy(void *h);
int ocu_mgmixed_sweep_counts(void *h, long long *counts);
int ocu_mgmixed_get_level_params(void *h, int level, ocu_mg_level_params *out);

/* ---- Sol_PCGPressure3DDevice<T> (ocuequation/sol_pcgpressure3d.h:41-101) ----
 * T follows rhs->dtype; rhs and coeff are the caller's congruent device grids
 * (the coefficient's ghosts are filled in place by each solve, as the
 * reference does).  initialize_storage(nx,ny,nz,hx,hy,hz,&rhs,&coeff) with
 * the public fields bc (periodic or Neumann faces), convergence,
 * preconditioner (0 none, 1 Jacobi, 2 multigrid, 3 BFBT), multigrid_cycles
 * and multigrid_use_fmg.  Each solve starts from the current pressure. */
void *ocu_pcg_create(const ocu_grid3d *rhs, const ocu_grid3d *coeff, float hx, float hy, float hz,
                     const ocu_bc_set *bc, int convergence, int preconditioner, int multigrid_cycles,
                     int multigrid_use_fmg);
/* solve(residual, tol, max_iter) (sol_pcgpressure3d.cu:550-597) */
int ocu_pcg_solve(void *h, double tol, int max_iter, double *residual);
int ocu_pcg_iterations(void *h);
int ocu_pcg_pressure(void *h, ocu_grid3d *out);
void ocu_pcg_destroy(void *h);

/* ---- NetCDFGrid3DWriter / NetCDFGrid3DReader (ocustorage/gridnetcdf.h:67-183) ----
 * Grid fields in NetCDF classic files (64-bit offset variant written; both
 * variants read).  Grids are host descriptors (base = host pointer); the
 * element type picks the variable type (OCU_F32 -> NC_FLOAT 5, OCU_F64 ->
 * NC_DOUBLE 6, OCU_I32 -> NC_INT 4).  staggering: 0 cell centre, 1 u face,
 * 2 v face, 3 w face, 4 custom (the offsets given).  The writer's file is
 * complete after ocu_nc_writer_close, which also frees the handle. */
void *ocu_nc_writer_open(const char *filename, int nx, int ny, int nz, float hx, float hy, float hz);
int ocu_nc_writer_define_variable(void *h, const char *name, int nc_type, int staggering, float offsetx,
                                  float offsety, float offsetz);
int ocu_nc_writer_add_time_level(void *h, float time, int *level);
int ocu_nc_writer_add_data(void *h, const char *name, const ocu_grid3d *host_grid, int time_level);
int ocu_nc_writer_close(void *h);
void *ocu_nc_reader_open(const char *filename);
int ocu_nc_reader_info(void *h, int *nx, int *ny, int *nz, float *hx, float *hy, float *hz, int *num_time_levels);
int ocu_nc_reader_time(void *h, int time_level, float *time);
/* index < 0 returns the number of 4-D variables; else its name (NUL
 * terminated, name_cap bytes available), type and sizes */
int ocu_nc_reader_variable(void *h, int index, char *name, int name_cap, int *nc_type, int *nx, int *ny, int *nz);
int ocu_nc_reader_read(void *h, const char *name, const ocu_grid3d *host_grid, int time_level);
void ocu_nc_reader_close(void *h);

/* ---- Sol_ProjectDivergence3DDevice<T> (ocuequation/sol_project3d.h:74-93) ----
 * u, v, w: congruent MAC grids (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1).
 * mixed = 1 (FP64 u) uses the mixed-precision pressure solver. */
void *ocu_project_create(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                         int nx, int ny, int nz, double hx, double hy, double hz, int mixed);
int ocu_project_solve(void *h, double tol, double *residual);
int ocu_project_solve_divergence_only(void *h);
int ocu_project_divergence(void *h, ocu_grid3d *out);
int ocu_project_pressure(void *h, ocu_grid3d *out);
void ocu_project_destroy(void *h);

/* ---- Eqn_IncompressibleNS3D<T> (ocuequation/eqn_incompressns3d.h:33-175) ---- */
typedef struct ocu_ns_params {
  int32_t nx, ny, nz;
  double hx, hy, hz;
  double max_divergence;
  ocu_bc_set temp_bc, flow_bc;
  int32_t vertical_direction;
  double gravity, bouyancy, thermal_diffusion, viscosity;
  int32_t advection_scheme, time_step;
  double cfl_factor;
  /* not in the reference: FP64 equations projected with FP32 V-cycles in
   * FP64 iterative refinement (dtype OCU_F64 only) */
  int32_t mixed_projection;
  /* not in the reference: 1 runs the reference's operator-by-operator step;
   * 0 (default) runs the explicit part as one fused launch where it applies
   * (same values; ocu_ns_step_fused in ocu_hip.h) */
  int32_t unfused_step;
} ocu_ns_params;

/* init_* describe host buffers (base = host pointer) laid out by
 * Eqn_IncompressibleNS3DParams::init_grids (eqn_incompressns3d.h:78-86) for
 * p->nx, ny, nz, element type dtype; a descriptor that differs from that
 * layout or type is rejected (NULL, ocu_last_error says which). */
void *ocu_ns_create(int dtype, const ocu_ns_params *p, const ocu_grid3d *init_u, const ocu_grid3d *init_v,
                    const ocu_grid3d *init_w, const ocu_grid3d *init_temp);
int ocu_ns_advance_one_step(void *h, double dt);
int ocu_ns_advance(void *h, double dt);
int ocu_ns_max_stable_timestep(void *h, double *dt);
/* which: 0 u, 1 v, 2 w, 3 temperature, 4 pressure (device descriptors).
 * A descriptor stays valid until the next step: the fused step writes the
 * new u, v, w, temperature into a second set of buffers and swaps them in. */
int ocu_ns_field(void *h, int which, ocu_grid3d *out);
int ocu_ns_num_steps(void *h);
void ocu_ns_destroy(void *h);

/* ---- Eqn_ScalarAdvection3D<T> (ocuequation/eqn_scalaradvection3d.h,
 * eqn_scalaradvection3d.cpp): forward-Euler passive advection of a scalar
 * by a fixed MAC velocity, with the no-corner ghost fill after each step. */
typedef struct ocu_scalar_params {
  int32_t nx, ny, nz;
  double hx, hy, hz;
  ocu_bc_set bc;
  int32_t advection_scheme;
} ocu_scalar_params;

/* init_* describe host buffers laid out by
 * Eqn_ScalarAdvection3DParams::init_grids (u, v, w: MAC paddings (0,1,1),
 * (1,0,1), (1,1,0); initial values (1,1,1)), checked like ocu_ns_create's. */
void *ocu_scalar_create(int dtype, const ocu_scalar_params *p, const ocu_grid3d *init_u, const ocu_grid3d *init_v,
                        const ocu_grid3d *init_w, const ocu_grid3d *init_phi);
int ocu_scalar_advance_one_step(void *h, double dt);
int ocu_scalar_advance(void *h, double dt);
int ocu_scalar_max_stable_timestep(void *h, double *dt);
/* which: 0 u, 1 v, 2 w, 3 phi, 4 deriv_phidt (device descriptors) */
int ocu_scalar_field(void *h, int which, ocu_grid3d *out);
void ocu_scalar_destroy(void *h);

#ifdef __cplusplus
}
#endif

#endif /* OCU_H */
