/*
 * ocu_hip.h - C ABI of libocu_hip.so, the hand-written HIP kernels (gfx950)
 * behind OpenCurrent's projection-method hot path.
 *
 * Conventions
 *  - Every entry point returns an int status (ocu_status, 0 = OK).  On error
 *    ocu_last_error() returns a thread-local message.
 *  - Grids are ocu_grid3d descriptors whose base points to device memory.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Kernels are enqueued asynchronously; nothing synchronises unless the
 *    entry point says so.  Entry points never allocate device memory: any
 *    scratch space is an explicit argument.
 *  - Each kernel entry point names the reference routine it replaces; the
 *    reference binds these as CUDA launches inside the named C++ methods.
 */
#ifndef OCU_HIP_H
#define OCU_HIP_H

#include "ocu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- runtime: errors, memory, streams, events, graphs ------- */

/* thread-local text of the last error (empty string if none) */
const char *ocu_last_error(void);
int ocu_get_device_count(int *count);
int ocu_set_device(int device);
int ocu_device_synchronize(void);

/* device memory (hipMalloc / hipFree): replaces ocuutil/memory.cu:26-91 and
 * the cudaMalloc in Grid3DDevice<T>::init (ocustorage/grid3ddev.cu:391). */
int ocu_malloc(void **ptr, size_t bytes);
int ocu_free(void *ptr);
/* pinned host memory (hipHostMalloc), host_malloc(pinned=true) in memory.cu */
int ocu_host_malloc(void **ptr, size_t bytes);
int ocu_host_free(void *ptr);
/* device address of ocu_host_malloc memory (kernels write results there
 * directly; read them after synchronising the stream) */
int ocu_host_device_pointer(void **dev, void *host);
/* kind: 0 host->device, 1 device->host, 2 device->device, 3 host->host.
 * Asynchronous on `stream` (cudaMemcpy in grid3ddev.cu:205, :245-265). */
int ocu_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);

int ocu_stream_create(void **stream);
int ocu_stream_destroy(void *stream);
int ocu_stream_synchronize(void *stream);

int ocu_event_create(void **event);
int ocu_event_destroy(void *event);
int ocu_event_record(void *event, void *stream);
int ocu_event_synchronize(void *event);
int ocu_event_elapsed_ms(float *ms, void *start, void *stop);

/* HIP-graph capture of everything enqueued on `stream` between begin and
 * end; end instantiates the graph and returns the executable. */
int ocu_stream_begin_capture(void *stream);
int ocu_stream_end_capture(void *stream, void **graph_exec);
int ocu_graph_launch(void *graph_exec, void *stream);
int ocu_graph_destroy(void *graph_exec);

/* ---------------- whole-grid operations (ocustorage/grid3ddev.cu) -------- */

/* every allocated element = value; Grid3DDevice<T>::clear / clear_zero
 * (grid3ddev.cu:177-195).  A fill kernel, so it is safe inside a captured
 * graph. */
int ocu_grid_clear(const ocu_grid3d *g, double value, void *stream);
/* dst[n] = (T_dst) src[n] over the allocation; copy_all_data<T,S>
 * (grid3ddev.cu:25-33, 216-281).  Layouts must match. */
int ocu_grid_copy(const ocu_grid3d *dst, const ocu_grid3d *src, void *stream);
/* zero the first ghost layer (faces, edges, corners); no reference
 * counterpart (adjoint fields carry no ghost values) */
int ocu_grid_clear_ghosts(const ocu_grid3d *g, void *stream);
/* r = a1*g1 (+ a2*g2); Grid3DDevice<T>::linear_combination
 * (grid3ddev.cu:283-349) including its 3%-waste rule for choosing between
 * an interior-only and a whole-allocation update.  g2 may be NULL. */
int ocu_grid_lincomb(const ocu_grid3d *r, double a1, const ocu_grid3d *g1, double a2,
                     const ocu_grid3d *g2, void *stream);
/* interior += value (mean removal) */
int ocu_grid_add_scalar(const ocu_grid3d *g, double value, void *stream);

/* Reductions over the interior; Grid3DDevice<T>::reduce_* (grid3ddev.cu:93-174,
 * grid3dops.cu:26-130).  op: 0 maxabs, 1 max, 2 min, 3 sum, 4 sqrsum,
 * 5 checknan.  Deterministic two-stage tree, accumulated in double.  The
 * result (one double) is written to device memory `out`; `workspace` must
 * hold ocu_reduce_workspace_bytes() bytes of device memory. */
size_t ocu_reduce_workspace_bytes(void);
int ocu_grid_reduce(const ocu_grid3d *g, int op, void *workspace, double *out, void *stream);
/* ocu_grid_reduce that also signals its completion (no counterpart in the
 * reference; for hosts that poll instead of synchronising the stream): out
 * is pinned host memory as the device sees it with room for two doubles;
 * after the result out[0], with a system-scope fence between, the 8 bytes
 * of out[1] receive `tag` (nonzero).  A host that reads tag from out[1]
 * (an acquire read) then finds the result in out[0]. */
int ocu_grid_reduce_tagged(const ocu_grid3d *g, int op, void *workspace, double *out, unsigned long long tag,
                           void *stream);
/* max |a|, max |b| and max |c| over the interiors of three grids of one
 * type in one pass, with the completion signal of ocu_grid_reduce_tagged
 * (no counterpart in the reference; it takes the place of the three
 * reduce_maxabs calls of get_max_stable_timestep, eqn_incompressns3d.cu:
 * 302-307): out is pinned host memory as the device sees it with room for
 * four doubles; after the maxima out[0..2], with a system-scope fence
 * between, out[3] receives `tag` (nonzero).  `workspace` must hold
 * 3 * ocu_reduce_workspace_bytes() bytes of device memory.  The maxima are
 * those of three ocu_grid_reduce calls with op 0, bit for bit. */
int ocu_grid_maxabs3_tagged(const ocu_grid3d *a, const ocu_grid3d *b, const ocu_grid3d *c, void *workspace,
                            double *out, unsigned long long tag, void *stream);
/* ocu_grid_reduce without the workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP-graph capture, tight serving loops).  The partials
 * live in a library-owned array of NORM_PARTS = 2048 doubles (16 KiB) per
 * GPU, held until the process ends; every grid ocu_grid_reduce accepts fits
 * (the partial count does not depend on the grid size; grids with nx*ny >=
 * INT_MAX rows are rejected as there).  Calls that overlap on one GPU (two
 * streams or two host threads) race on that array. */
int ocu_grid_reduce_noalloc(const ocu_grid3d *g, int op, double *out, void *stream);

/* ---------------- ghost-cell fills (ocustorage/grid3dboundary.cu) -------- */

/* apply_3d_boundary_conditions_level1 (grid3dboundary.cu:758-820): same
 * values as the reference's x, y, z passes, computed in one launch. */
int ocu_bc_fill_level1(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy,
                       double hz, void *stream);
/* apply_3d_boundary_conditions_level1_nocorners, device version
 * (grid3dboundary.cu:28-119, 719-753) */
int ocu_bc_fill_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy,
                          double hz, void *stream);
/* apply_3d_mac_boundary_conditions_level1 (grid3dboundary.cu:825-900) */
int ocu_bc_fill_mac_level1(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_bc_set *bc, void *stream);

/* apply_3d_boundary_conditions_level2_nocorners (grid3dboundary.cu:122-232,
 * 904-938): the two ghost layers of each face (not edges or corners), from
 * the interior: periodic copies, Dirichlet -phi0 + 2*value in both layers,
 * Neumann phi0 -/+ h*value and phi0 -/+ 2*h*value.  Needs 2 ghost layers. */
int ocu_bc_fill_level2_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy,
                                 double hz, void *stream);
/* apply_3d_mac_boundary_conditions_level2 (grid3dboundary.cu:350-549,
 * 940-1016): x, y, z passes over two ghost layers (periodic or forced
 * inflow / variable slip).  nx is u->nx - 1: the reference computes
 * u.nx() - u.gz(), which with the two ghost layers it requires rejects
 * every congruent MAC triple. */
int ocu_bc_fill_mac_level2(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_bc_set *bc, void *stream);

/* ---------------- point sampling (ocustorage/grid3dsample.cu) ------------- */

/* sample_points_3d (grid3dsample.cu:91-171): trilinear interpolation of phi
 * (ghosts filled) at n float positions (device arrays, physical units,
 * shifted by shift* cells), with periodic wrap (period_n* cells) or the
 * clamp / Neumann extrapolation of the faces (periodic, Neumann, Dirichlet,
 * scalar slip).  phi_sampled: n values of phi's type (device), each the
 * float result of the lookup.  The cell index is kept in the first ghost
 * layer (the reference reads one past it for a position exactly at the upper
 * periodic end). */
int ocu_sample_points(void *phi_sampled, const float *position_x, const float *position_y,
                      const float *position_z, int n, const ocu_grid3d *phi, const ocu_bc_set *bc,
                      float period_nx, float period_ny, float period_nz, float hx, float hy, float hz,
                      float shiftx, float shifty, float shiftz, void *stream);
/* sample_points_mac_grid_3d (grid3dsample.cu:174-263): u, v, w of a MAC
 * velocity at n positions, with the reference's translation of the MAC
 * faces (periodic or forced inflow) into sampling conditions. */
int ocu_sample_points_mac(void *vx, void *vy, void *vz, const float *position_x, const float *position_y,
                          const float *position_z, int n, const ocu_grid3d *u, const ocu_grid3d *v,
                          const ocu_grid3d *w, const ocu_bc_set *bc, float hx, float hy, float hz,
                          void *stream);

/* ---------------- multigrid (ocuequation/sol_mgpressure3ddev*.cu) -------- */

/* One colour (cells with (i+j+k)%2 == color) of red-black Gauss-Seidel,
 * in place.  variant OCU_RELAX_DEVICE = Sol_MultigridPressure3DDeviceD_relax_DIAG
 * (devd.cu:542-599) for FP64 grids / Sol_MultigridPressure3DDeviceF_relax
 * (devf.cu:200-252) for FP32 grids; OCU_RELAX_HOST = do_cpu_solve
 * (sol_mgpressure3ddev.cu:82-125) arithmetic. */
int ocu_mg_relax(const ocu_grid3d *u, const ocu_grid3d *b, int color,
                 const ocu_mg_level_params *p, int variant, void *stream);
/* One full iteration - colour first_color then the other - with the
 * OCU_RELAX_DEVICE arithmetic, reading u_in and writing the interior of u_out
 * (congruent with u_in, must not alias it; the ghosts of u_out are not
 * written).  Bit-identical to two ocu_mg_relax calls when no face is
 * periodic (the reference refreshes ghosts between colours only for periodic
 * faces, sol_mgpressure3ddev.cu:227-241); returns OCU_ERR_UNSUPPORTED if a
 * face of `bc` is periodic. */
/* ocu_mg_relax (OCU_RELAX_DEVICE) that also writes the level-1 ghosts of
 * the cells it updates, so that relax-colour + apply_3d_boundary_conditions_
 * level1(u, bc, hx, hy, hz) (the reference's fill between colours for
 * periodic faces, sol_mgpressure3ddev.cu:227-241) is one launch.  Equal to
 * the pair bit for bit when u's ghosts were filled before the call.  Faces
 * Neumann, Dirichlet or periodic in pairs with an even number of cells along
 * a periodic axis (else OCU_ERR_UNSUPPORTED).  No reference counterpart as
 * one call. */
int ocu_mg_relax_bc(const ocu_grid3d *u, const ocu_grid3d *b, int color, const ocu_mg_level_params *p,
                    const ocu_bc_set *bc, double hx, double hy, double hz, void *stream);
int ocu_mg_relax_fused(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                       const ocu_mg_level_params *p, const ocu_bc_set *bc, int first_color,
                       void *stream);
/* 1 when ocu_mg_relax_fused_bc with ghost writes runs for these grids and
 * faces: every face Neumann or Dirichlet, or periodic x and/or y pairs (even
 * counts along them, z not periodic, rows of at most 256 cells split evenly
 * over the lanes, aligned); else 0 */
int ocu_mg_relax_fused_ok(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                          const ocu_bc_set *bc);
/* ocu_mg_relax_fused with two options (flags, OR-ed):
 *   OCU_FUSED_WRITE_GHOSTS: also write the first ghost layer of u_out exactly
 *     as apply_3d_boundary_conditions_level1(u_out, bc, hx, hy, hz) would
 *     (grid3dboundary.cu:237-347), so no separate fill is needed;
 *   OCU_FUSED_ZERO_INPUT: u_in is taken to be the zero grid with its ghosts
 *     filled by that same fill (the state after clear_zero + level-1 fill,
 *     sol_mgpressure3d.cpp:168-170) and is not read; u_in may equal u_out.
 * Either flag needs every face Neumann or Dirichlet, except that ghost
 * writes also take periodic x/y pairs on whole rows (ocu_mg_relax_fused_ok;
 * else OCU_ERR_UNSUPPORTED).  With periodic faces the sweep equals the
 * reference's two colours with a fill between and after them. */
int ocu_mg_relax_fused_bc(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                          const ocu_mg_level_params *p, const ocu_bc_set *bc, double hx,
                          double hy, double hz, int first_color, int flags, void *stream);
/* ocu_mg_prolong_bc(u_coarse, u_in, quirk, bc, hx, hy, hz) followed by
 * ocu_mg_relax_fused_bc(u_in, u_out, b, ..., flags) in one launch, bit for
 * bit, except that u_in is left as it was (the sweep reads u_in plus the
 * prolonged coarse correction and its ghost transforms).  flags: only
 * OCU_FUSED_WRITE_GHOSTS.  Needs fine = 2 x coarse, 129..256 cells per row
 * with aligned rows (ocu_mg_relax_fused_prolong_ok), Neumann or Dirichlet
 * faces, the coarse ghost layer filled; else OCU_ERR_UNSUPPORTED.  No
 * reference counterpart as one call. */
int ocu_mg_relax_fused_prolong_bc(const ocu_grid3d *u_in, const ocu_grid3d *u_out,
                                  const ocu_grid3d *b, const ocu_grid3d *u_coarse, int quirk,
                                  const ocu_mg_level_params *p, const ocu_bc_set *bc, double hx,
                                  double hy, double hz, int first_color, int flags, void *stream);
/* 1 when ocu_mg_relax_fused_prolong_bc can run for this fine/coarse pair
 * and these faces (u stands for u_in, u_out and b), else 0 */
int ocu_mg_relax_fused_prolong_ok(const ocu_grid3d *u, const ocu_grid3d *u_coarse,
                                  const ocu_bc_set *bc);
/* r = b - A u over the interior; calculate_residual (devd.cu:94-114,
 * devf.cu:68-90; constants as in the launchers devd.cu:1372-1405). */
int ocu_mg_residual(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r, double h,
                    double fx, double fy, double fz, void *stream);
/* b_coarse = .125 * sum of the 8 children of r; restrict (devd.cu:65-91) */
int ocu_mg_restrict(const ocu_grid3d *r, const ocu_grid3d *b_coarse, void *stream);
/* Fused residual + restriction: the fine residual of each child is computed
 * in registers, summed in the restrict kernel's order and written to
 * b_coarse; r is written too when r != NULL.  If norm_ws != NULL, per-block
 * partial (sum r^2, max|r|) in double go to norm_ws (OCU_NORM_PARTS pairs)
 * for ocu_norm_finalize.  Same values as residual followed by restrict. */
int ocu_mg_residual_restrict(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r,
                             const ocu_grid3d *b_coarse, double h, double fx, double fy,
                             double fz, double *norm_ws, void *stream);
/* r = b - A u with optional norm partials (the level-0 convergence check of
 * restrict_residuals(0,0,...), sol_mgpressure3ddev.cu:250-279) */
int ocu_mg_residual_norm(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r,
                         double h, double fx, double fy, double fz, double *norm_ws,
                         void *stream);
/* out[0] = sum r^2, out[1] = max |r| (device doubles) from the partials */
int ocu_norm_finalize(const double *norm_ws, double *out, void *stream);
/* the same with a completion signal (no counterpart in the reference): out
 * has room for three doubles and receives `tag` (nonzero) in the 8 bytes of
 * out[2] after the two results, with a system-scope fence between */
int ocu_norm_finalize_tagged(const double *norm_ws, double *out, unsigned long long tag, void *stream);
/* u_fine += trilinear prolongation of u_coarse; prolong (devd.cu:118-223).
 * quirk = 1 reproduces the reference corner choice (devd.cu:181-185). */
int ocu_mg_prolong(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, int quirk, void *stream);
/* ocu_mg_prolong followed by apply_3d_boundary_conditions_level1(u_fine,
 * bc, hx, hy, hz) in the same launch (the pair in prolong,
 * sol_mgpressure3ddev.cu:281-300).  Needs u_fine exactly twice u_coarse
 * and every face Neumann or Dirichlet. */
int ocu_mg_prolong_bc(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, int quirk,
                      const ocu_bc_set *bc, double hx, double hy, double hz, void *stream);
/* 1 when ocu_mg_prolong_bc runs for this pair and these faces: fine exactly
 * twice the coarse grid and either every face Neumann or Dirichlet, or
 * periodic x/y pairs with a non-periodic z on whole rows; else 0 */
int ocu_mg_prolong_bc_ok(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine,
                         const ocu_bc_set *bc);
/* The bottom of a V-cycle in one workgroup: levels top..top+nlev-1 (the
 * last the coarsest) of Sol_MultigridPressure3DDevice, held in LDS.  Does
 * what the per-level calls do on them (sol_mgpressure3d.cpp:184-239):
 * relax(nu1) with the OCU_RELAX_DEVICE arithmetic and a level-1 fill after
 * every iteration, a zero start on the next level, residual + restriction,
 * the coarsest relaxation (ocu_mg_coarse_relax semantics, coarse_mode 0/1,
 * coarse_iters red-black iterations), prolongation with the level-1 fill and
 * relax(nu2); every level's u (with ghosts) and b (below the top) is written
 * back.  Arrays u[l], b[l], p[l] and h[3l..3l+2] (hx,hy,hz) describe level
 * top+l; zero_top: the top level's u is the zero grid (not read; needs
 * homogeneous faces).  Faces must be Neumann or Dirichlet; the boxes must
 * fit ocu_mg_bottom_lds_bytes() <= ocu_mg_bottom_lds_max(). */
int ocu_mg_bottom_vcycle(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_mg_level_params *p,
                         const double *h, int nlev, const ocu_bc_set *bc, int nu1, int nu2,
                         int coarse_iters, int coarse_mode, int zero_top, void *stream);
/* LDS bytes a bottom launch over these levels needs (0: invalid request) */
size_t ocu_mg_bottom_lds_bytes(const ocu_grid3d *u, int nlev, int coarse_mode);
/* LDS bytes one bottom (or coarse-relax) workgroup may use: 160 KiB */
size_t ocu_mg_bottom_lds_max(void);
/* Coarsest-level relaxation in one workgroup (replaces relax_on_host,
 * sol_mgpressure3ddev.cu:134-188, and its two PCIe round trips).
 * mode 0: reference semantics - `iterations` RB-GS iterations with the
 *   do_cpu_solve arithmetic in the grid's type, the host no-corner ghost
 *   fill after every colour, then apply_3d_boundary_conditions_level1.
 * mode 1: FP32 grids relaxed in double (copy in, `iterations` host-order
 *   RB-GS iterations with a level-1 ghost fill after every colour, round
 *   back to float) - the coarsest solve of the mixed-precision solver.
 * order: 0 red-black, 1 black-red, 2 symmetric (as relax()).  The grid must
 * fit in 64 KiB of LDS (u and b in the compute type, ghosts included). */
int ocu_mg_coarse_relax(const ocu_grid3d *u, const ocu_grid3d *b, int iterations, int order,
                        const ocu_mg_level_params *p, const ocu_bc_set *bc, double hx, double hy,
                        double hz, int mode, void *stream);

/* ---------------- mixed-precision iterative refinement -------------------- */

/* r = (b - bshift) - A x in FP64 (x, b FP64), r32 = (float)(r*scale) (FP32,
 * congruent); partial (sum r^2, max|r|, sum r, sum x) per block to ws
 * (OCU_NORM_PARTS quadruples) for ocu_norm4_finalize. */
int ocu_ir_residual(const ocu_grid3d *x, const ocu_grid3d *b, double bshift,
                    const ocu_grid3d *r32, double scale, double h, double fx, double fy,
                    double fz, double *ws, void *stream);
/* out[0..2] / out[0..3]: the triples / quadruples of ws reduced (slot 1 by
 * max, the others by sum) */
int ocu_norm3_finalize(const double *ws, double *out, void *stream);
int ocu_norm4_finalize(const double *ws, double *out, void *stream);
/* the same with a completion signal (no counterpart in the reference): after
 * the 3 (4) results, with a system-scope fence between, `tag` (nonzero) goes
 * into the 8 bytes of out[3] (out[4]) */
int ocu_norm3_finalize_tagged(const double *ws, double *out, unsigned long long tag, void *stream);
int ocu_norm4_finalize_tagged(const double *ws, double *out, unsigned long long tag, void *stream);
/* shift[0] = (xsum + inv_scale * sum of e over the interior) / (nx*ny*nz),
 * the interior mean of x + e*inv_scale when xsum is the interior sum of x
 * (e FP32; ws: OCU_NORM_PARTS doubles; shift: device memory). */
int ocu_ir_shift(const ocu_grid3d *e, double inv_scale, double xsum, double *ws, double *shift,
                 void *stream);
/* the same triples for an FP64 grid itself (sum g^2, max|g|, sum g over
 * the interior): the three statistics of the right-hand side in one pass */
int ocu_grid_stats3(const ocu_grid3d *g, double *ws, void *stream);
/* x += (double)e * inv_scale over the interior (x FP64, e FP32) */
int ocu_ir_correct(const ocu_grid3d *x, const ocu_grid3d *e, double inv_scale, void *stream);
/* x = (x + (double)e * inv_scale) - shift[0] (shift: device memory, or null
 * for none) over the interior, and with faces != 0 also over the six
 * face-ghost layers (not edges or corners).  Face mode is the ghost fill
 * itself when every face of x and e is a copy (homogeneous Neumann or
 * periodic) and both ghost layers are filled on entry. */
int ocu_ir_correct_ex(const ocu_grid3d *x, const ocu_grid3d *e, double inv_scale,
                      const double *shift, int faces, void *stream);
/* out[0] = interior sum of g (double) via ws (OCU_NORM_PARTS doubles) */
int ocu_grid_sum(const ocu_grid3d *g, double *ws, double *out, void *stream);

/* number of per-block partial slots used by the norm entry points */
int ocu_norm_parts(void);

/* ---------------- MAC-grid operators (ocuequation/sol_*3ddev.cu) ----------- */

/* Sol_Divergence3DDevice_calculate_divergence (sol_divergence3ddev.cu:23-45) */
int ocu_divergence(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                   const ocu_grid3d *div, int nx, int ny, int nz, double hx, double hy,
                   double hz, void *stream);
/* Sol_Gradient3DDevice_subtract_grad (sol_gradient3ddev.cu:24-47) over the
 * cells of p */
int ocu_gradient_sub(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                     const ocu_grid3d *p, double coefficient, double hx, double hy, double hz,
                     void *stream);
/* Advection3DF/D_apply_upwind_TEX (sol_selfadvection3ddev.cu:136-337) */
int ocu_self_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                    const ocu_grid3d *du, const ocu_grid3d *dv, const ocu_grid3d *dw, int nx,
                    int ny, int nz, double hx, double hy, double hz, int scheme, void *stream);
/* Sol_PassiveAdvection3D_apply_interp (sol_passiveadvection3ddev.cu:31-78) */
int ocu_passive_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                       const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                       double hx, double hy, double hz, int scheme, void *stream);
/* The two entry points above (same references, same argument meaning) with
 * one more scheme, OCU_IT_SECOND_ORDER_UPWIND, which has no reference
 * counterpart.  Bit a of periodic_axes (x=1, y=2, z=4) marks axis a
 * periodic: that scheme's two-cell upstream stencil then wraps across it
 * instead of falling back to first order at the boundary faces.  The
 * periodic ghost layer must be filled on entry, as for the other schemes. */
int ocu_self_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                       const ocu_grid3d *du, const ocu_grid3d *dv, const ocu_grid3d *dw, int nx,
                       int ny, int nz, double hx, double hy, double hz, int scheme,
                       int periodic_axes, void *stream);
int ocu_passive_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                          const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                          double hx, double hy, double hz, int scheme, int periodic_axes,
                          void *stream);
/* ---- adjoints (no reference counterpart; for gradients of losses through
 * the advection-projection path) ----
 * Adjoint of ocu_passive_advect with the velocity's upwind selection fixed:
 * lam is the adjoint of dphi on the interior (its ghost layer must be zero);
 * writes lam_out = lam + A_phi^T(scale*lam) over the ghost box (with scale =
 * dt this is the adjoint of phi + dt*A(u,phi)) and accumulates
 * adj_u/v/w += A_u^T(scale*lam) on the faces [0,n] of interior rows.  All
 * grids congruent with u; lam_out must not alias lam. */
int ocu_passive_advect_adjoint(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                               const ocu_grid3d *phi, const ocu_grid3d *lam, const ocu_grid3d *lam_out,
                               const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                               double scale, int nx, int ny, int nz, double hx, double hy, double hz,
                               int scheme, void *stream);
/* Adjoint of ocu_bc_fill_nocorners: folds every face ghost's adjoint into
 * the interior cell it derives from (periodic/Neumann +1, Dirichlet -1,
 * BC_NONE dropped) and zeroes the face ghosts. */
int ocu_bc_adjoint_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, void *stream);
/* Adjoint of the normal-face part of ocu_bc_fill_mac_level1 on the faces an
 * advection reads: periodic high face n duplicates face 0 (adj(0) +=
 * adj(n)); forced-inflow low face 0 is fixed (adj(0) = 0); adj(n) = 0. */
int ocu_mac_adjoint_normal(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_bc_set *bc, void *stream);
/* Adjoint of ocu_divergence (sol_divergence3ddev.cu:23-45): adj_u/v/w +=
 * D^T lam on the normal faces [0,n] of interior rows, i.e. adj_f +=
 * (1/h)*(lam_{f-1} - lam_f); lam is read on its interior only (values off
 * the interior count as zero).  lam congruent with adj_u. */
int ocu_divergence_adjoint(const ocu_grid3d *lam, const ocu_grid3d *adj_u, const ocu_grid3d *adj_v,
                           const ocu_grid3d *adj_w, int nx, int ny, int nz, double hx, double hy,
                           double hz, void *stream);
/* Adjoint in p of ocu_gradient_sub (sol_gradient3ddev.cu:24-47): adj_p +=
 * G^T(adj_u, adj_v, adj_w) over p's interior and its low face ghosts (the
 * cells the gradient reads), with coefficient and 1/h as in the forward
 * call; adj_u/v/w are read on the faces [0,n-1] of interior rows.  The
 * adjoint in u is the identity. */
int ocu_gradient_adjoint(const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                         const ocu_grid3d *adj_p, double coefficient, double hx, double hy, double hz,
                         void *stream);
/* Sol_LaplacianCentered3DDevice_stencil (sol_laplaciancent3ddev.cu:27-62) */
int ocu_laplacian_acc(const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz,
                      double hx, double hy, double hz, double coefficient, void *stream);
/* Eqn_IncompressibleNS3D_add_thermal_force (eqn_incompressns3d.cu:25-44) */
int ocu_thermal_force(const ocu_grid3d *d, const ocu_grid3d *t, int nx, int ny, int nz,
                      int vertical_direction, double coefficient, void *stream);

/* ---------------- PCG pressure solver (ocuequation/sol_pcgpressure3d.cu) -- */

/* Sol_PCGPressure3DDevice_apply_laplacian (sol_pcgpressure3d.cu:88-126):
 * dst = -div(c grad src) on the interior with face coefficients
 * (c_i + c_nb)/2, from src's and coeff's ghosts as they are (the caller
 * fills them).  f*_div_hsq are the T values f/h^2 (passed as double, as the
 * reference's kernel parameters).  When dot_out is given, also
 * *dot_out = sum over the interior of (T)(src * dst) in double (device
 * memory; workspace holds ocu_reduce_workspace_bytes()), the quantity of
 * invoke_kernel_dot_product(src, dst) (:262-293).  dst must not alias src. */
int ocu_pcg_apply_laplacian(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff,
                            double fx_div_hsq, double fy_div_hsq, double fz_div_hsq, void *workspace,
                            double *dot_out, void *stream);
/* Sol_PCGPressure3DDevice_diag_preconditioner (sol_pcgpressure3d.cu:45-83):
 * dst = src / (sum over the six faces of f/h^2 (c_i + c_nb)/2); dst may be
 * src. */
int ocu_pcg_diag_precond(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff,
                         double fx_div_hsq, double fy_div_hsq, double fz_div_hsq, void *stream);
/* *out = sum over the interior of (T)(a * b), accumulated in double;
 * invoke_kernel_dot_product (sol_pcgpressure3d.cu:262-293) without its
 * temporary grid.  workspace: ocu_reduce_workspace_bytes(). */
int ocu_grid_dot(const ocu_grid3d *a, const ocu_grid3d *b, void *workspace, double *out, void *stream);

/* The explicit part of Eqn_IncompressibleNS3D<T>::advance_one_step
 * (eqn_incompressns3d.cu:331-393) in one launch: self-advection, viscous
 * diffusion and thermal force of u, v, w, passive advection and diffusion
 * of the temperature, the forward-Euler (or Adams-Bashforth) combination
 * linear_combination(1, x, a_dt, dx) [+ linear_combination(1, x, a_ab,
 * last_dx)] with the reference's whole-allocation / interior choice
 * (grid3ddev.cu:316), and the temperature's no-corner ghost fill
 * (grid3dboundary.cu:28-119).  Reads u, v, w, temp and writes the new
 * values into u_new .. temp_new (congruent, not aliasing): afterwards the
 * new buffers hold, element for element, what the reference's in-place
 * sequence leaves in the old ones, except the normal face n of u, v, w
 * (overwritten by the MAC boundary fill of the projection that follows;
 * flow faces periodic or forced-inflow).  keep_tendencies stores the
 * tendencies in du .. dtemp (the copies into the "last" grids become a
 * swap); use_last reads last_du .. last_dtemp.  Unused tendency grids may
 * be NULL.  No reference counterpart as one call. */
int ocu_ns_step_fused(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *temp,
                      const ocu_grid3d *u_new, const ocu_grid3d *v_new, const ocu_grid3d *w_new,
                      const ocu_grid3d *temp_new, const ocu_grid3d *du, const ocu_grid3d *dv,
                      const ocu_grid3d *dw, const ocu_grid3d *dtemp, const ocu_grid3d *last_du,
                      const ocu_grid3d *last_dv, const ocu_grid3d *last_dw, const ocu_grid3d *last_dtemp,
                      const ocu_ns_step_params *p, const ocu_bc_set *temp_bc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OCU_HIP_H */
