// The explicit part of one Eqn_IncompressibleNS3D step in one launch.
//
// The reference's step (eqn_incompressns3d.cu:331-393) runs self-advection,
// three momentum Laplacians, the thermal force, passive advection of T and
// its Laplacian, each writing or updating a tendency grid, then the
// forward-Euler / Adams-Bashforth combinations (linear_combination,
// grid3ddev.cu:283-349) and the no-corner ghost fill of T, then (AB2) copies
// the tendencies to the "last" grids: about 38 grid passes of memory traffic
// per cell.  Here one thread per element of the padded allocation computes
// every tendency of its cell in registers (the same device functions, in the
// same order, as the per-operator kernels), combines it with the old value
// and writes the new u, v, w and T into a second set of buffers (the caller
// swaps them in afterwards): 4 grids read, 4 written (plus, for AB2, the
// last tendencies read and this step's written).  Every element of the new
// buffers receives exactly what the in-place sequence leaves in the old one:
//   - cells the combination covers get old*1 + d*a_dt (+ last*a_ab);
//   - with the reference's whole-allocation combination (allocation waste
//     <= 3%) the other elements get old*1 + 0*a_dt (+ 0*a_ab), the tendency
//     grids being zero outside the cells their operators write;
//   - with the interior-only combination they keep the old value;
//   - T's face ghosts take the no-corner fill of the new interior value
//     they derive from (recomputed by the ghost's own thread).
#include "advect.h"

using namespace ocu_hip;

namespace {

struct TFace {
  int type;
  float value;
};

template <typename T>
struct NSStep {
  const T *u, *v, *w, *t;       // old values (element 0 of the allocation)
  T *un, *vn, *wn, *tn;         // new values
  T *du, *dv, *dw, *dtt;        // AB2: this step's tendencies (written)
  const T *lu, *lv, *lw, *lt;   // AB2: the last step's tendencies (read)
  Inv<T> ih, ih2;
  T nu, kappa, tcoef, a_dt, a_ab, th[3];
  TFace tf[6];  // xpos, xneg, ypos, yneg, zpos, zneg
  int64_t xs, ys, shift, vnbr;
  int pre, nx, ny, nz, gx, gy, gz, pnx, pny, pnz;
  int scheme, fper, tper, vaxis, flat, keep, use_last, visc, tdiff;
};

template <typename T>
__device__ __forceinline__ T combine(const NSStep<T> &s, T old, T d, T last) {
  T r = old * (T)1 + d * s.a_dt;
  if (s.use_last) r = r * (T)1 + last * s.a_ab;
  return r;
}

// an element no operator writes: the whole-allocation combination adds
// zero tendencies, the interior one leaves it alone
template <typename T>
__device__ __forceinline__ T untouched(const NSStep<T> &s, int grid, T old) {
  return ((s.flat >> grid) & 1) ? combine(s, old, (T)0, (T)0) : old;
}

// T's tendency at interior cell c (offset from cell (0,0,0)): passive
// advection, then + kappa*lap(T)
template <typename T>
__device__ __forceinline__ T temp_tendency(const NSStep<T> &s, int64_t c, int i, int j, int k) {
  const T *U = s.u + s.shift, *V = s.v + s.shift, *W = s.w + s.shift, *Tt = s.t + s.shift;
  T d = s.scheme == OCU_IT_SECOND_ORDER_UPWIND
            ? passive_advect_up2_cell(U, V, W, Tt, c, s.xs, s.ys, s.ih, i, j, k, s.nx, s.ny, s.nz, s.tper)
            : passive_advect_cell(U, V, W, Tt, c, s.xs, s.ys, s.ih, s.scheme);
  if (s.tdiff) d = d + s.kappa * laplacian_cell(Tt, c, s.xs, s.ys, s.ih2);
  return d;
}

// no-corner ghost value (grid3dboundary.cu:28-119): Dirichlet in double
// (`-phi + 2.0 * (T)value`), Neumann phi -/+ h*value in T, periodic a copy
template <typename T>
__device__ __forceinline__ T nocorner_xform(const TFace &f, bool neg, T h, T a) {
  if (f.type == OCU_BC_DIRICHELET) return (T)(-(double)a + 2.0 * (double)(T)f.value);
  if (f.type == OCU_BC_NEUMANN) return neg ? a - h * (T)f.value : a + h * (T)f.value;
  return a;
}

__device__ __forceinline__ bool tf_writes(const TFace &f) {
  return f.type == OCU_BC_PERIODIC || f.type == OCU_BC_DIRICHELET || f.type == OCU_BC_NEUMANN;
}

template <typename T>
__global__ __launch_bounds__(256) void k_ns_step(NSStep<T> s) {
  const int kp = blockIdx.x * 64 + threadIdx.x, jp = blockIdx.y * 4 + threadIdx.y, ip = blockIdx.z;
  // the pre-padding before the padded box: untouched elements
  if (ip == 0 && jp == 0 && kp < s.pre) {
    const int64_t e = kp;
    s.un[e] = untouched(s, 0, s.u[e]);
    s.vn[e] = untouched(s, 1, s.v[e]);
    s.wn[e] = untouched(s, 2, s.w[e]);
    s.tn[e] = untouched(s, 3, s.t[e]);
  }
  if (kp >= s.pnz || jp >= s.pny) return;
  const int i = ip - s.gx, j = jp - s.gy, k = kp - s.gz;
  const int64_t e = (int64_t)s.pre + ((int64_t)ip * s.pny + jp) * s.pnz + kp;  // element of the allocation
  const int64_t c = e - s.shift;                                               // offset from cell (0,0,0)
  const T *U = s.u + s.shift, *V = s.v + s.shift, *W = s.w + s.shift, *Tt = s.t + s.shift;

  const bool ci = i >= 0 && i < s.nx, cj = j >= 0 && j < s.ny, ck = k >= 0 && k < s.nz;
  // interiors of the four grids: u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), T (nx,ny,nz)
  const bool in_u = i >= 0 && i <= s.nx && cj && ck;
  const bool in_v = ci && j >= 0 && j <= s.ny && ck;
  const bool in_w = ci && cj && k >= 0 && k <= s.nz;
  const bool in_t = ci && cj && ck;

  T du = 0, dv = 0, dw = 0;
  if (in_t) {  // advection covers the cell box
    if (s.scheme == OCU_IT_SECOND_ORDER_UPWIND)
      self_advect_up2_cell(U, V, W, c, s.xs, s.ys, s.ih, i, j, k, s.nx, s.ny, s.nz, s.fper, du, dv, dw);
    else
      self_advect_cell(U, V, W, c, s.xs, s.ys, s.ih, s.scheme, du, dv, dw);
  }
  if (s.visc) {
    if (in_u) du = du + s.nu * laplacian_cell(U, c, s.xs, s.ys, s.ih2);
    if (in_v) dv = dv + s.nu * laplacian_cell(V, c, s.xs, s.ys, s.ih2);
    if (in_w) dw = dw + s.nu * laplacian_cell(W, c, s.xs, s.ys, s.ih2);
  }
  if (in_t) {  // the thermal force covers the cell box of the vertical component
    const T f = thermal_cell(Tt, c, s.vnbr, s.tcoef);
    if (s.vaxis == 0) du = du + f;
    else if (s.vaxis == 1) dv = dv + f;
    else dw = dw + f;
  }

  if (in_u) {
    s.un[e] = combine(s, U[c], du, s.use_last ? s.lu[e] : (T)0);
    if (s.keep) s.du[e] = du;
  } else {
    s.un[e] = untouched(s, 0, U[c]);
  }
  if (in_v) {
    s.vn[e] = combine(s, V[c], dv, s.use_last ? s.lv[e] : (T)0);
    if (s.keep) s.dv[e] = dv;
  } else {
    s.vn[e] = untouched(s, 1, V[c]);
  }
  if (in_w) {
    s.wn[e] = combine(s, W[c], dw, s.use_last ? s.lw[e] : (T)0);
    if (s.keep) s.dw[e] = dw;
  } else {
    s.wn[e] = untouched(s, 2, W[c]);
  }

  if (in_t) {
    const T dt = temp_tendency(s, c, i, j, k);
    s.tn[e] = combine(s, Tt[c], dt, s.use_last ? s.lt[e] : (T)0);
    if (s.keep) s.dtt[e] = dt;
    return;
  }
  // a face ghost of T (one index at -1 or n, the others interior) that the
  // no-corner fill writes: the fill of the new value of its source cell
  const bool gx_ = (i == -1 || i == s.nx) && cj && ck, gy_ = ci && (j == -1 || j == s.ny) && ck,
             gz_ = ci && cj && (k == -1 || k == s.nz);
  int face = -1, si = i, sj = j, sk = k, ax = 0;
  bool neg = false;
  if (gx_) {
    neg = i < 0, face = neg ? 1 : 0, ax = 0;
    si = s.tf[face].type == OCU_BC_PERIODIC ? (neg ? s.nx - 1 : 0) : (neg ? 0 : s.nx - 1);
  } else if (gy_) {
    neg = j < 0, face = neg ? 3 : 2, ax = 1;
    sj = s.tf[face].type == OCU_BC_PERIODIC ? (neg ? s.ny - 1 : 0) : (neg ? 0 : s.ny - 1);
  } else if (gz_) {
    neg = k < 0, face = neg ? 5 : 4, ax = 2;
    sk = s.tf[face].type == OCU_BC_PERIODIC ? (neg ? s.nz - 1 : 0) : (neg ? 0 : s.nz - 1);
  }
  if (face >= 0 && tf_writes(s.tf[face])) {
    const int64_t cs = (int64_t)si * s.xs + (int64_t)sj * s.ys + sk;
    const T dt = temp_tendency(s, cs, si, sj, sk);
    const T tnew = combine(s, Tt[cs], dt, s.use_last ? s.lt[cs + s.shift] : (T)0);
    s.tn[e] = nocorner_xform(s.tf[face], neg, s.th[ax], tnew);
  } else {
    s.tn[e] = untouched(s, 3, Tt[c]);
  }
}

int check_congruent(const ocu_grid3d *a, const ocu_grid3d *b, const char *what) {
  if (int e = check_grid(b, what)) return e;
  OCU_REQUIRE(layout_match(a, b) && a->dtype == b->dtype && a->gx == b->gx && a->gy == b->gy && a->gz == b->gz &&
                  a->shift == b->shift && a->allocated == b->allocated,
              "%s - grids are not congruent", what);
  return OCU_OK;
}

// the reference's choice in linear_combination (grid3ddev.cu:316): more
// than 3% of the allocation outside the interior -> interior only
bool whole_allocation(const ocu_grid3d *g) {
  const int64_t interior = (int64_t)g->nx * g->ny * g->nz;
  return !((g->allocated - interior) / (float)g->allocated > .03f);
}

template <typename T>
int launch_ns_step(const ocu_grid3d *const *in, const ocu_grid3d *const *out, const ocu_grid3d *const *d,
                   const ocu_grid3d *const *last, const ocu_ns_step_params *p, const ocu_bc_set *tbc,
                   hipStream_t stream) {
  NSStep<T> s;
  const ocu_grid3d *g = in[0];
  s.u = (const T *)in[0]->base, s.v = (const T *)in[1]->base, s.w = (const T *)in[2]->base,
  s.t = (const T *)in[3]->base;
  s.un = (T *)out[0]->base, s.vn = (T *)out[1]->base, s.wn = (T *)out[2]->base, s.tn = (T *)out[3]->base;
  s.keep = p->keep_tendencies != 0, s.use_last = p->use_last != 0;
  s.du = s.keep ? (T *)d[0]->base : nullptr, s.dv = s.keep ? (T *)d[1]->base : nullptr,
  s.dw = s.keep ? (T *)d[2]->base : nullptr, s.dtt = s.keep ? (T *)d[3]->base : nullptr;
  s.lu = s.use_last ? (const T *)last[0]->base : nullptr, s.lv = s.use_last ? (const T *)last[1]->base : nullptr,
  s.lw = s.use_last ? (const T *)last[2]->base : nullptr, s.lt = s.use_last ? (const T *)last[3]->base : nullptr;
  s.ih = Inv<T>{(T)(1 / p->hx), (T)(1 / p->hy), (T)(1 / p->hz)};
  s.ih2 = Inv<T>{(T)(1 / (p->hx * p->hx)), (T)(1 / (p->hy * p->hy)), (T)(1 / (p->hz * p->hz))};
  s.nu = (T)p->viscosity, s.kappa = (T)p->thermal_diffusion, s.tcoef = (T)p->thermal_coef;
  s.a_dt = (T)p->a_dt, s.a_ab = (T)p->a_ab;
  s.th[0] = (T)p->hx, s.th[1] = (T)p->hy, s.th[2] = (T)p->hz;
  const ocu_bc *all[6] = {&tbc->xpos, &tbc->xneg, &tbc->ypos, &tbc->yneg, &tbc->zpos, &tbc->zneg};
  for (int f = 0; f < 6; f++) s.tf[f] = TFace{all[f]->type, all[f]->value};
  s.xs = (int64_t)g->pny * g->pnz, s.ys = g->pnz, s.shift = g->shift;
  const int vd = p->vertical_direction;
  s.vaxis = (vd & OCU_DIR_XAXIS_FLAG) ? 0 : (vd & OCU_DIR_YAXIS_FLAG) ? 1 : 2;
  s.vnbr = s.vaxis == 0 ? s.xs : s.vaxis == 1 ? s.ys : 1;
  s.pre = (int)(g->shift - ((int64_t)g->gx * g->pny * g->pnz + (int64_t)g->gy * g->pnz + g->gz));
  s.nx = p->nx, s.ny = p->ny, s.nz = p->nz, s.gx = g->gx, s.gy = g->gy, s.gz = g->gz;
  s.pnx = g->pnx, s.pny = g->pny, s.pnz = g->pnz;
  s.scheme = p->advection_scheme, s.fper = p->flow_periodic, s.tper = p->temp_periodic;
  s.flat = 0;
  for (int q = 0; q < 4; q++) s.flat |= whole_allocation(in[q]) << q;
  s.visc = p->viscosity > 0, s.tdiff = p->thermal_diffusion > 0;
  const dim3 grid((g->pnz + 63) / 64, (g->pny + 3) / 4, g->pnx);
  k_ns_step<T><<<grid, dim3(64, 4, 1), 0, stream>>>(s);
  return OCU_OK;
}

}  // namespace

extern "C" {

int ocu_ns_step_fused(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *temp,
                      const ocu_grid3d *u_new, const ocu_grid3d *v_new, const ocu_grid3d *w_new,
                      const ocu_grid3d *temp_new, const ocu_grid3d *du, const ocu_grid3d *dv, const ocu_grid3d *dw,
                      const ocu_grid3d *dtemp, const ocu_grid3d *last_du, const ocu_grid3d *last_dv,
                      const ocu_grid3d *last_dw, const ocu_grid3d *last_dtemp, const ocu_ns_step_params *p,
                      const ocu_bc_set *temp_bc, void *stream) {
  const char *what = "ocu_ns_step_fused";
  OCU_REQUIRE(p && temp_bc, "%s: null parameters", what);
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  OCU_REQUIRE(u && v && w && temp, "%s: null grid", what);
  if (int e = check_grid(u, what)) return e;
  OCU_REQUIRE(u->nx == nx + 1 && u->ny == ny && u->nz == nz, "%s - u dimensions mismatch", what);
  OCU_REQUIRE(v && v->nx == nx && v->ny == ny + 1 && v->nz == nz, "%s - v dimensions mismatch", what);
  OCU_REQUIRE(w && w->nx == nx && w->ny == ny && w->nz == nz + 1, "%s - w dimensions mismatch", what);
  OCU_REQUIRE(temp && temp->nx == nx && temp->ny == ny && temp->nz == nz, "%s - temperature dimensions mismatch", what);
  OCU_REQUIRE(u->dtype == OCU_F32 || u->dtype == OCU_F64, "%s - dtype", what);
  OCU_REQUIRE(u->gx >= 1 && u->gy >= 1 && u->gz >= 1, "%s - must have at least one ghost cell", what);
  const ocu_grid3d *in[4] = {u, v, w, temp}, *out[4] = {u_new, v_new, w_new, temp_new};
  const ocu_grid3d *d[4] = {du, dv, dw, dtemp}, *last[4] = {last_du, last_dv, last_dw, last_dtemp};
  // all grids share one layout (the congruent MAC layout of init_grids)
  for (int q = 0; q < 4; q++) {
    OCU_REQUIRE(in[q] && out[q], "%s: null grid", what);
    if (int e = check_congruent(u, in[q], what)) return e;
    if (int e = check_congruent(u, out[q], what)) return e;
    OCU_REQUIRE(out[q]->base != in[q]->base, "%s - the new buffers must not alias the old ones", what);
    if (p->keep_tendencies) {
      OCU_REQUIRE(d[q], "%s - keep_tendencies needs the tendency grids", what);
      if (int e = check_congruent(u, d[q], what)) return e;
    }
    if (p->use_last) {
      OCU_REQUIRE(last[q], "%s - use_last needs the last-tendency grids", what);
      if (int e = check_congruent(u, last[q], what)) return e;
    }
    if (p->keep_tendencies && p->use_last)
      OCU_REQUIRE(d[q]->base != last[q]->base, "%s - tendency and last tendency must differ", what);
  }
  OCU_REQUIRE(p->advection_scheme == OCU_IT_FIRST_ORDER_UPWIND || p->advection_scheme == OCU_IT_SECOND_ORDER_CENTERED ||
                  p->advection_scheme == OCU_IT_SECOND_ORDER_UPWIND,
              "%s - invalid interpolation type %d", what, p->advection_scheme);
  OCU_REQUIRE(p->flow_periodic >= 0 && p->flow_periodic <= 7 && p->temp_periodic >= 0 && p->temp_periodic <= 7,
              "%s - invalid periodic axes mask", what);
  const int vd = p->vertical_direction;
  OCU_REQUIRE(vd == OCU_DIR_XPOS || vd == OCU_DIR_XNEG || vd == OCU_DIR_YPOS || vd == OCU_DIR_YNEG ||
                  vd == OCU_DIR_ZPOS || vd == OCU_DIR_ZNEG,
              "%s - bad vertical_direction %d", what, vd);
  const ocu_bc *all[6] = {&temp_bc->xpos, &temp_bc->xneg, &temp_bc->ypos, &temp_bc->yneg, &temp_bc->zpos,
                          &temp_bc->zneg};
  for (auto b : all)
    OCU_REQUIRE(b->type == OCU_BC_NONE || b->type == OCU_BC_PERIODIC || b->type == OCU_BC_DIRICHELET ||
                    b->type == OCU_BC_NEUMANN,
                "%s - invalid temperature boundary condition type %d", what, b->type);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  if (u->dtype == OCU_F64) launch_ns_step<double>(in, out, d, last, p, temp_bc, S(stream));
  else launch_ns_step<float>(in, out, d, last, p, temp_bc, S(stream));
  OCU_HIP_LAUNCHED(what);
}

}  // extern "C"
