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

// One element (ip, jp, kp) of the padded box (indices from the box corner,
// ghosts included): the new value of each grid there.
template <typename T>
__device__ __forceinline__ void ns_element(const NSStep<T> &s, int ip, int jp, int kp) {
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
  // (the face and spacing picked by selects: an indexed read of the
  // parameter arrays would copy them to scratch memory)
  int si = i, sj = j, sk = k;
  bool neg = false, has = false;
  TFace f{OCU_BC_NONE, 0.f};
  T h = 0;
  if (gx_) {
    neg = i < 0, f = neg ? s.tf[1] : s.tf[0], h = s.th[0], has = true;
    si = f.type == OCU_BC_PERIODIC ? (neg ? s.nx - 1 : 0) : (neg ? 0 : s.nx - 1);
  } else if (gy_) {
    neg = j < 0, f = neg ? s.tf[3] : s.tf[2], h = s.th[1], has = true;
    sj = f.type == OCU_BC_PERIODIC ? (neg ? s.ny - 1 : 0) : (neg ? 0 : s.ny - 1);
  } else if (gz_) {
    neg = k < 0, f = neg ? s.tf[5] : s.tf[4], h = s.th[2], has = true;
    sk = f.type == OCU_BC_PERIODIC ? (neg ? s.nz - 1 : 0) : (neg ? 0 : s.nz - 1);
  }
  if (has && tf_writes(f)) {
    const int64_t cs = (int64_t)si * s.xs + (int64_t)sj * s.ys + sk;
    const T dt = temp_tendency(s, cs, si, sj, sk);
    const T tnew = combine(s, Tt[cs], dt, s.use_last ? s.lt[cs + s.shift] : (T)0);
    s.tn[e] = nocorner_xform(f, neg, h, tnew);
  } else {
    s.tn[e] = untouched(s, 3, Tt[c]);
  }
}

// the pre-padding before the padded box: untouched elements
template <typename T>
__device__ __forceinline__ void ns_prepad(const NSStep<T> &s, int64_t e) {
  s.un[e] = untouched(s, 0, s.u[e]);
  s.vn[e] = untouched(s, 1, s.v[e]);
  s.wn[e] = untouched(s, 2, s.w[e]);
  s.tn[e] = untouched(s, 3, s.t[e]);
}

// The interior core is the box [bx, nx-bx) x [by, ny-by) x [bz, nz-bz) of
// T's interior (a band of b = 2 cells is left out along a periodic axis when
// second-order upwind samples would wrap around it); k_ns_march takes the
// core.  Every other element - the bands, the u, v and w faces at i = nx,
// j = ny and k = nz, the ghost layers, the padding - goes here, one thread
// each, after the pre-padding: the planes outside the core's x range (count
// na), then the rows outside its y range of the core's planes (nb), then the
// row ends outside its z range of the core's rows (nc).
struct Core {
  int bx, by, bz;
};

template <typename T>
__global__ __launch_bounds__(256) void k_ns_shell(NSStep<T> s, Core b, int64_t na, int64_t nb, int64_t nc) {
  int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q < s.pre) {
    ns_prepad(s, q);
    return;
  }
  q -= s.pre;
  const int cx = s.nx - 2 * b.bx, cy = s.ny - 2 * b.by, cz = s.nz - 2 * b.bz;  // core extents
  const int lx = s.gx + b.bx, ly = s.gy + b.by, lz = s.gz + b.bz;             // its low corner (box indices)
  int ip, jp, kp;
  if (q < na) {
    const int64_t plane = (int64_t)s.pny * s.pnz, pl = q / plane, r = q % plane;
    ip = pl < lx ? (int)pl : (int)pl + cx;
    jp = (int)(r / s.pnz);
    kp = (int)(r % s.pnz);
  } else if ((q -= na) < nb) {
    const int64_t per = (int64_t)(s.pny - cy) * s.pnz, pl = q / per, r = q % per;
    const int rr = (int)(r / s.pnz);
    ip = (int)pl + lx;
    jp = rr < ly ? rr : rr + cy;
    kp = (int)(r % s.pnz);
  } else if ((q -= nb) < nc) {
    const int per = s.pnz - cz, t = (int)(q % per);
    const int64_t rowi = q / per;
    ip = (int)(rowi / cy) + lx;
    jp = (int)(rowi % cy) + ly;
    kp = t < lz ? t : t + cz;
  } else {
    return;
  }
  ns_element(s, ip, jp, kp);
}

// ------------------------------------------------------------------------
// the interior core, marching along x through an LDS window
// ------------------------------------------------------------------------
//
// A block owns a tile of MTJ rows x MTK cells (j, k) of T's interior (its
// cells in the core) and an x-chunk of planes.  It keeps a window of planes p-2..p+2 of u, v, w and T
// in LDS, each plane with a two-cell halo in j and k (every sample the
// cell functions read: second-order upwind reaches two cells, the
// self-advection velocities one diagonal), in six slots, so the plane
// loading for p+1 goes to the slot no thread reads in plane p and one
// barrier per plane orders the accesses.  The cell functions read the
// window through WAt, whose offsets use a virtual x stride of 2^20 elements
// (no sample of a core cell wraps around a periodic axis).  Each cell's
// tendencies and new values are those of ns_element, in the same order.
constexpr int MTK = 64, MTJ = 8, MWK = MTK + 4, MWJ = MTJ + 4, MPLANE = MWJ * MWK, MSLOTS = 6;
constexpr int MXSH = 20;
constexpr int64_t MXS = (int64_t)1 << MXSH;
constexpr int MLOADS = (MPLANE + MTK * MTJ - 1) / (MTK * MTJ);  // window elements per thread and plane

template <typename T>
struct WAt {
  const T *base;               // this field's slot 0
  int cell;                    // the cell within a window plane
  int sm2, sm1, s0, sp1, sp2;  // element offsets of the slots of planes p-2..p+2
  __device__ __forceinline__ T operator()(int64_t o) const {
    const int di = (int)((o + (MXS >> 1)) >> MXSH);
    const int r = (int)(o - ((int64_t)di << MXSH));
    // a chain of selects, not an array: constant offsets fold to one slot,
    // and a computed one keeps the slots in registers
    const int sl = di == -2 ? sm2 : di == -1 ? sm1 : di == 0 ? s0 : di == 1 ? sp1 : sp2;
    return base[sl + cell + r];
  }
};

template <typename T>
struct Tend {
  T du, dv, dw, dt;
};

// the tendencies of interior cell (i, j, k) through accessors A (window or
// global): the sequence of ns_element for a cell of T's interior (returned
// by value, so that the two callers' results stay in registers)
template <bool WRAP, typename T, typename A>
__device__ __forceinline__ Tend<T> ns_tendencies(const NSStep<T> &s, const A &U, const A &V, const A &W, const A &Tt,
                                                 int64_t xs, int64_t ys, int64_t vnbr, int i, int j, int k) {
  T du, dv, dw, dt;
  if (s.scheme == OCU_IT_SECOND_ORDER_UPWIND)
    self_advect_up2_at<WRAP, T>(U, V, W, xs, ys, s.ih, i, j, k, s.nx, s.ny, s.nz, s.fper, du, dv, dw);
  else
    self_advect_at<T>(U, V, W, xs, ys, s.ih, s.scheme, du, dv, dw);
  if (s.visc) {
    du = du + s.nu * laplacian_at<T>(U, xs, ys, s.ih2);
    dv = dv + s.nu * laplacian_at<T>(V, xs, ys, s.ih2);
    dw = dw + s.nu * laplacian_at<T>(W, xs, ys, s.ih2);
  }
  const T f = thermal_at<T>(Tt, vnbr, s.tcoef);
  if (s.vaxis == 0) du = du + f;
  else if (s.vaxis == 1) dv = dv + f;
  else dw = dw + f;
  dt = s.scheme == OCU_IT_SECOND_ORDER_UPWIND
           ? passive_advect_up2_at<WRAP, T>(U, V, W, Tt, xs, ys, s.ih, i, j, k, s.nx, s.ny, s.nz, s.tper)
           : passive_advect_at<T>(U, V, W, Tt, xs, ys, s.ih, s.scheme);
  if (s.tdiff) dt = dt + s.kappa * laplacian_at<T>(Tt, xs, ys, s.ih2);
  return Tend<T>{du, dv, dw, dt};
}

template <typename T>
__global__ __launch_bounds__(MTK *MTJ) void k_ns_march(NSStep<T> s, Core b, int tiles_k, int tiles_j, int ichunk,
                                                       int ntiles) {
  extern __shared__ unsigned char ns_smem[];
  T *win = reinterpret_cast<T *>(ns_smem);  // [4 fields][MSLOTS][MPLANE]
  // blocks are dealt round-robin to the 8 XCDs: a contiguous run of tiles
  // per XCD
  const int per = gridDim.x / 8;
  const int logical = (blockIdx.x % 8) * per + (blockIdx.x / 8);
  if (logical >= ntiles) return;  // uniform per block
  const int tk = logical % tiles_k, rem = logical / tiles_k, tj = rem % tiles_j, ti = rem / tiles_j;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * MTK + tx;
  const int k0 = tk * MTK, j0 = tj * MTJ, i0 = ti * ichunk, i1 = min(s.nx, i0 + ichunk);
  const int k = k0 + tx, j = j0 + ty;
  const bool mine = k >= b.bz && k < s.nz - b.bz && j >= b.by && j < s.ny - b.by;
  const T *src[4] = {s.u, s.v, s.w, s.t};

  // element offsets (from element 0) of this thread's window positions in
  // plane 0 (-1: outside the padded box in j or k)
  int64_t wpos[MLOADS];
#pragma unroll
  for (int m = 0; m < MLOADS; m++) {
    const int t = tid + m * MTK * MTJ;
    const int wj = t / MWK, wk = t % MWK;
    const int jp = j0 - 2 + wj + s.gy, kp = k0 - 2 + wk + s.gz;
    wpos[m] = (t < MPLANE && jp >= 0 && jp < s.pny && kp >= 0 && kp < s.pnz)
                  ? (int64_t)s.pre + (int64_t)jp * s.pnz + kp
                  : (int64_t)-1;
  }
  auto slot = [](int q) { return ((q % MSLOTS) + MSLOTS) % MSLOTS; };
  auto fetch = [&](int q, T (&v)[4][MLOADS]) {
    const int ip = q + s.gx;
    const bool in = ip >= 0 && ip < s.pnx;
#pragma unroll
    for (int m = 0; m < MLOADS; m++) {
      const bool ok = in && wpos[m] >= 0;
      const int64_t e = wpos[m] + (int64_t)ip * s.pny * s.pnz;
#pragma unroll
      for (int f = 0; f < 4; f++) v[f][m] = ok ? src[f][e] : (T)0;
    }
  };
  auto store = [&](int q, const T (&v)[4][MLOADS]) {
    T *dst = win + slot(q) * MPLANE;
#pragma unroll
    for (int m = 0; m < MLOADS; m++) {
      const int t = tid + m * MTK * MTJ;
      if (t < MPLANE) {
#pragma unroll
        for (int f = 0; f < 4; f++) dst[f * MSLOTS * MPLANE + t] = v[f][m];
      }
    }
  };
  T v[4][MLOADS];
  for (int q = i0 - 2; q <= i0 + 2; q++) {
    fetch(q, v);
    store(q, v);
  }
  __syncthreads();

  const int cell = (ty + 2) * MWK + tx + 2;
  const int64_t xs = s.xs, ys = s.ys;
  const int64_t vnbr_w = s.vaxis == 0 ? MXS : s.vaxis == 1 ? (int64_t)MWK : (int64_t)1;
  for (int p = i0; p < i1; p++) {
    fetch(p + 3, v);  // in flight while plane p is computed
    if (mine && p >= b.bx && p < s.nx - b.bx) {
      const int64_t c = (int64_t)p * xs + (int64_t)j * ys + k;  // offset from cell (0,0,0)
      const int sm2 = slot(p - 2) * MPLANE, sm1 = slot(p - 1) * MPLANE, s0 = slot(p) * MPLANE,
                sp1 = slot(p + 1) * MPLANE, sp2 = slot(p + 2) * MPLANE;
      auto acc = [&](int f) { return WAt<T>{win + f * MSLOTS * MPLANE, cell, sm2, sm1, s0, sp1, sp2}; };
      const Tend<T> d = ns_tendencies<false, T>(s, acc(0), acc(1), acc(2), acc(3), MXS, (int64_t)MWK, vnbr_w, p, j, k);
      const int64_t e = c + s.shift;
      const T *olds = win + slot(p) * MPLANE + cell;
      const T ou = olds[0], ov = olds[MSLOTS * MPLANE], ow = olds[2 * MSLOTS * MPLANE],
              ot = olds[3 * MSLOTS * MPLANE];
      s.un[e] = combine(s, ou, d.du, s.use_last ? s.lu[e] : (T)0);
      s.vn[e] = combine(s, ov, d.dv, s.use_last ? s.lv[e] : (T)0);
      s.wn[e] = combine(s, ow, d.dw, s.use_last ? s.lw[e] : (T)0);
      s.tn[e] = combine(s, ot, d.dt, s.use_last ? s.lt[e] : (T)0);
      if (s.keep) {
        s.du[e] = d.du;
        s.dv[e] = d.dv;
        s.dw[e] = d.dw;
        s.dtt[e] = d.dt;
      }
    }
    store(p + 3, v);
    __syncthreads();
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
  // the core: bands of two cells along the periodic axes whose second-order
  // upwind samples wrap (flow or temperature)
  Core b{0, 0, 0};
  if (s.scheme == OCU_IT_SECOND_ORDER_UPWIND) {
    const int per = s.fper | s.tper;
    b.bx = (per & 1) ? 2 : 0, b.by = (per & 2) ? 2 : 0, b.bz = (per & 4) ? 2 : 0;
  }
  const bool core = s.nx > 2 * b.bx && s.ny > 2 * b.by && s.nz > 2 * b.bz;
  if (!core) b = Core{(s.nx + 1) / 2, (s.ny + 1) / 2, (s.nz + 1) / 2};  // all to the shell
  const int cx = max(s.nx - 2 * b.bx, 0), cy = max(s.ny - 2 * b.by, 0), cz = max(s.nz - 2 * b.bz, 0);
  // the shell: everything outside the core
  const int64_t na = (int64_t)(s.pnx - cx) * s.pny * s.pnz, nb = (int64_t)cx * (s.pny - cy) * s.pnz,
                nc = (int64_t)cx * cy * (s.pnz - cz);
  const int64_t nshell = s.pre + na + nb + nc;
  k_ns_shell<T><<<(unsigned)((nshell + 255) / 256), 256, 0, stream>>>(s, b, na, nb, nc);
  OCU_HIP_TRY(hipGetLastError());
  if (!core) return OCU_OK;
  // the interior
  const int tiles_k = (s.nz + MTK - 1) / MTK, tiles_j = (s.ny + MTJ - 1) / MTJ;
  const size_t lds = (size_t)4 * MSLOTS * MPLANE * sizeof(T);
  OCU_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ns_march<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int dev = 0, cus = 256, per = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_march<T>, MTK * MTJ, lds) != hipSuccess || per < 1)
    per = 1;
  // planes per block: the fewest rounds of resident blocks, each block
  // loading four planes beyond its own
  const int64_t slots = (int64_t)cus * per;
  int ichunk = s.nx;
  int64_t best = INT64_MAX;
  for (int ic = 1;; ic = ic < 8 ? ic + 1 : ic * 2) {
    const int c = min(ic, s.nx);  // the last candidate is the whole extent
    const int64_t nb_ = (int64_t)tiles_k * tiles_j * ((s.nx + c - 1) / c);
    const int64_t cost = ((nb_ + slots - 1) / slots) * (c + 4);
    if (cost < best) best = cost, ichunk = c;
    if (c == s.nx) break;
  }
  const int ntiles = tiles_k * tiles_j * ((s.nx + ichunk - 1) / ichunk);
  const int nblocks = ((ntiles + 7) / 8) * 8;
  k_ns_march<T><<<nblocks, dim3(MTK, MTJ), lds, stream>>>(s, b, tiles_k, tiles_j, ichunk, ntiles);
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
  const int e = u->dtype == OCU_F64 ? launch_ns_step<double>(in, out, d, last, p, temp_bc, S(stream))
                                    : launch_ns_step<float>(in, out, d, last, p, temp_bc, S(stream));
  if (e != OCU_OK) return e;
  OCU_HIP_LAUNCHED(what);
}

}  // extern "C"
