// MAC-grid operators around the pressure solve: divergence, gradient
// subtraction, self-/passive advection, centred Laplacian, thermal force.
// u, v, w, phi and their derivatives share one padded layout (the reference
// requires congruent grids, sol_selfadvection3ddev.cu:622-626), so one
// index serves every array.
#include "ocu_hip_common.h"
#include "advect.h"
#include "rows.h"

using namespace ocu_hip;

namespace {

// sol_divergence3ddev.cu:23-45
template <typename T>
__global__ __launch_bounds__(256) void k_divergence(const T *__restrict__ u, const T *__restrict__ v,
                                                    const T *__restrict__ w, T *__restrict__ d, Inv<T> ih,
                                                    int64_t xs, int64_t ys, int nx, int ny, int nz) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  d[id] = ih.x * (u[id + xs] - u[id]) + ih.y * (v[id + ys] - v[id]) + ih.z * (w[id + 1] - w[id]);
}

// sol_gradient3ddev.cu:24-47
template <typename T>
__global__ __launch_bounds__(256) void k_gradient(T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                                                  const T *__restrict__ p, T coef, Inv<T> ih, int64_t xs, int64_t ys,
                                                  int nx, int ny, int nz) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  T pc = p[id];
  u[id] += coef * ih.x * (pc - p[id - xs]);
  v[id] += coef * ih.y * (pc - p[id - ys]);
  w[id] += coef * ih.z * (pc - p[id - 1]);
}

// Advection3DF/D_apply_upwind_TEX (sol_selfadvection3ddev.cu:136-337)
template <typename T>
__global__ __launch_bounds__(256) void k_self_advect(const T *__restrict__ U, const T *__restrict__ V,
                                                     const T *__restrict__ W, T *__restrict__ DU, T *__restrict__ DV,
                                                     T *__restrict__ DW, Inv<T> ih, int64_t xs, int64_t ys, int nx,
                                                     int ny, int nz, int scheme) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  T du, dv, dw;
  self_advect_cell(U, V, W, id, xs, ys, ih, scheme, du, dv, dw);
  DU[id] = du;
  DV[id] = dv;
  DW[id] = dw;
}

// Sol_PassiveAdvection3D_apply_interp (sol_passiveadvection3ddev.cu:31-78),
// marching along x: a thread takes the cells (i, j, k) of planes i0 .. i1-1
// and keeps phi at i-1, i, i+1 and u at the faces i, i+1 of its (j, k) in
// registers, so every plane of u and phi is read once.
template <typename T>
__global__ __launch_bounds__(256) void k_passive_advect_march(const T *__restrict__ u, const T *__restrict__ v,
                                                              const T *__restrict__ w, const T *__restrict__ phi,
                                                              T *__restrict__ d, Inv<T> ih, int64_t xs, int64_t ys,
                                                              int nx, int ny, int nz, int scheme, int ichunk) {
  __builtin_assume(xs > ys && ys > 1);
  const int k = blockIdx.x * TK + threadIdx.x;
  const int j = blockIdx.y * TJ + threadIdx.y;
  const int i0 = blockIdx.z * ichunk, i1 = min(nx, i0 + ichunk);
  if (j >= ny || k >= nz) return;
  const int64_t row = (int64_t)j * ys + k;
  // planes -1 .. nx of phi and -1 .. nx+1 of u are allocated
  T pm = phi[(int64_t)(i0 - 1) * xs + row], pc = phi[(int64_t)i0 * xs + row];
  T uc = u[(int64_t)i0 * xs + row];
  for (int i = i0; i < i1; i++) {
    const int64_t id = (int64_t)i * xs + row;
    const T pn = phi[id + xs], un = u[id + xs];
    d[id] = passive_advect_at<T>(XAt<T>{u + id, (T)0, uc, un, xs}, GAt<T>{v + id}, GAt<T>{w + id},
                                 XAt<T>{phi + id, pm, pc, pn, xs}, xs, ys, ih, scheme);
    pm = pc;
    pc = pn;
    uc = un;
  }
}

// ---- second-order (linear) upwind (no reference counterpart; advect.h) ----

template <typename T>
__global__ __launch_bounds__(256) void k_self_advect_up2(const T *__restrict__ U, const T *__restrict__ V,
                                                         const T *__restrict__ W, T *__restrict__ DU,
                                                         T *__restrict__ DV, T *__restrict__ DW, Inv<T> ih,
                                                         int64_t xs, int64_t ys, int nx, int ny, int nz, int pmask) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  T du, dv, dw;
  self_advect_up2_cell(U, V, W, id, xs, ys, ih, i, j, k, nx, ny, nz, pmask, du, dv, dw);
  DU[id] = du;
  DV[id] = dv;
  DW[id] = dw;
}

template <typename T>
__global__ __launch_bounds__(256) void k_passive_advect_up2(const T *__restrict__ u, const T *__restrict__ v,
                                                            const T *__restrict__ w, const T *__restrict__ phi,
                                                            T *__restrict__ d, Inv<T> ih, int64_t xs, int64_t ys,
                                                            int nx, int ny, int nz, int pmask) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  d[id] = passive_advect_up2_cell(u, v, w, phi, id, xs, ys, ih, i, j, k, nx, ny, nz, pmask);
}

// ---- adjoints (no reference counterpart) ----
//
// passive advection is dphi_c = -sum_axes (F_{c+1} - F_c)/h with face fluxes
// F_f = u_f * I(u_f, phi_{f-1}, phi_f), linear in phi for fixed u and in u for
// a fixed upwind selection.  For lam (adjoint of dphi, zero ghost layer) and
// g_f = scale * ((lam_f - lam_{f-1}) / h), each cell of the ghost box
// gathers  lam_out = lam + sum_faces (g_f * u_f) * dI/dphi_c  (x faces low
// then high, then y, then z) and, as the owner of its low faces, accumulates
// adj_u_f += g_f * I(u_f, ...).
template <typename T>
__device__ __forceinline__ T interp_w0(int scheme, T u) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return (T).5;
  return u < 0 ? (T)0 : (T)1;
}
template <typename T>
__device__ __forceinline__ T interp_w1(int scheme, T u) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return (T).5;
  return u < 0 ? (T)1 : (T)0;
}

// Marching along x: a thread takes the cells (i, j, k) of planes i0 ..
// i1-1 of the ghost box (planes -1 .. nx) and keeps lam at i-1, i, i+1, u
// at the faces i and i+1 and phi at i-1, i of its (j, k) in registers from
// plane to plane, so every plane of lam, u and phi is read once.
template <typename T>
__global__ __launch_bounds__(256) void k_passive_advect_adjoint_march(
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w, const T *__restrict__ phi,
    const T *__restrict__ lam, T *__restrict__ out, T *__restrict__ au, T *__restrict__ av, T *__restrict__ aw,
    Inv<T> ih, T sc, int64_t xs, int64_t ys, int nx, int ny, int nz, int scheme, int ichunk) {
  const int k = (int)(blockIdx.x * TK + threadIdx.x) - 1;
  const int j = (int)(blockIdx.y * TJ + threadIdx.y) - 1;
  const int i0 = (int)blockIdx.z * ichunk - 1, i1 = min(nx + 1, i0 + ichunk);
  if (j > ny || k > nz) return;
  const bool ji = j >= 0 && j < ny, ki = k >= 0 && k < nz;
  const int64_t row = (int64_t)j * ys + k;
  // planes -1 .. nx of lam and phi, -1 .. nx+1 of u are allocated
  T lm = i0 >= 0 ? lam[(int64_t)(i0 - 1) * xs + row] : (T)0;
  T lc = lam[(int64_t)i0 * xs + row];
  T uc = i0 >= 0 ? u[(int64_t)i0 * xs + row] : (T)0;
  T pm = i0 >= 0 ? phi[(int64_t)(i0 - 1) * xs + row] : (T)0;
  for (int i = i0; i < i1; i++) {
    const int64_t id = (int64_t)i * xs + row;
    const bool ii = i >= 0 && i < nx;
    const T lp = i < nx ? lam[id + xs] : (T)0;
    const T un = i < nx ? u[id + xs] : (T)0;
    const T pc = phi[id];
    T acc = 0;
    if (ji && ki) {
      if (i >= 0) {
        const T g = sc * (ih.x * (lc - lm));
        const T f = uc;
        acc += (g * f) * interp_w1(scheme, f);
        au[id] += g * interp(scheme, f, pm, pc);
      }
      if (i < nx) {
        const T g = sc * (ih.x * (lp - lc));
        const T f = un;
        acc += (g * f) * interp_w0(scheme, f);
      }
    }
    if (ii && ki) {
      if (j >= 0) {
        const T g = sc * (ih.y * (lc - lam[id - ys]));
        const T f = v[id];
        acc += (g * f) * interp_w1(scheme, f);
        av[id] += g * interp(scheme, f, phi[id - ys], pc);
      }
      if (j < ny) {
        const T g = sc * (ih.y * (lam[id + ys] - lc));
        const T f = v[id + ys];
        acc += (g * f) * interp_w0(scheme, f);
      }
    }
    if (ii && ji) {
      if (k >= 0) {
        const T g = sc * (ih.z * (lc - lam[id - 1]));
        const T f = w[id];
        acc += (g * f) * interp_w1(scheme, f);
        aw[id] += g * interp(scheme, f, phi[id - 1], pc);
      }
      if (k < nz) {
        const T g = sc * (ih.z * (lam[id + 1] - lc));
        const T f = w[id + 1];
        acc += (g * f) * interp_w0(scheme, f);
      }
    }
    out[id] = lc + acc;
    lm = lc;
    lc = lp;
    uc = un;
    pm = pc;
  }
}

// adjoint of the device no-corner fill: each face ghost is c * (source) +
// const with c = 1 (periodic copy, Neumann), -1 (Dirichlet), 0 (none); the
// source cell gathers c * adj(ghost) (x, y, z; low then high face) and zeroes
// the ghost it read - every face ghost has exactly one source.
struct GhostCoef {
  float c[6];   // xpos, xneg, ypos, yneg, zpos, zneg
  int per[6];   // periodic: the ghost's source is the opposite end
};

template <typename T>
__global__ __launch_bounds__(256) void k_bc_adjoint_nocorners(T *__restrict__ a, GhostCoef gc, int64_t xs,
                                                              int64_t ys, int nx, int ny, int nz) {
  const int k = blockIdx.x * TK + threadIdx.x;
  const int j = blockIdx.y * TJ + threadIdx.y;
  const int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const bool xl = i == (gc.per[1] ? nx - 1 : 0), xh = i == (gc.per[0] ? 0 : nx - 1);
  const bool yl = j == (gc.per[3] ? ny - 1 : 0), yh = j == (gc.per[2] ? 0 : ny - 1);
  const bool zl = k == (gc.per[5] ? nz - 1 : 0), zh = k == (gc.per[4] ? 0 : nz - 1);
  if (!(xl || xh || yl || yh || zl || zh)) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  T acc = a[id];
  auto take = [&](bool on, int face, int64_t gid) {
    if (!on) return;
    acc += (T)gc.c[face] * a[gid];
    a[gid] = (T)0;
  };
  take(xl, 1, (int64_t)(-1) * xs + (int64_t)j * ys + k);
  take(xh, 0, (int64_t)nx * xs + (int64_t)j * ys + k);
  take(yl, 3, (int64_t)i * xs + (int64_t)(-1) * ys + k);
  take(yh, 2, (int64_t)i * xs + (int64_t)ny * ys + k);
  take(zl, 5, (int64_t)i * xs + (int64_t)j * ys - 1);
  take(zh, 4, (int64_t)i * xs + (int64_t)j * ys + nz);
  a[id] = acc;
}

// adjoint of the MAC fill on the normal faces [0,n]: a periodic high face
// duplicates face 0 (adj(0) += adj(n)), a forced-inflow low face fixes face 0
// (adj(0) dropped); face n is never a parameter (adj(n) = 0)
template <typename T>
__global__ void k_mac_adjoint_normal(T *__restrict__ a, int64_t s_n, int64_t s_a, int64_t s_b, int n, int na,
                                     int nb, int per_hi, int inflow_lo) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= na * nb) return;
  const int ia = q / nb, ib = q % nb;
  T *p = a + (int64_t)ia * s_a + (int64_t)ib * s_b;
  if (per_hi) p[0] += p[(int64_t)n * s_n];
  if (inflow_lo) p[0] = (T)0;
  p[(int64_t)n * s_n] = (T)0;
}

// adjoint of the divergence: one thread per index (i,j,k) of the face box
// [0,nx]x[0,ny]x[0,nz] updates the u, v and w faces at that index that lie on
// interior rows: adj_f += ih * (lam_{f-1} - lam_f), lam = 0 off the interior
template <typename T>
__global__ __launch_bounds__(256) void k_divergence_adjoint(const T *__restrict__ lam, T *__restrict__ au,
                                                            T *__restrict__ av, T *__restrict__ aw, Inv<T> ih,
                                                            int64_t xs, int64_t ys, int nx, int ny, int nz) {
  const int k = blockIdx.x * TK + threadIdx.x;
  const int j = blockIdx.y * TJ + threadIdx.y;
  const int i = blockIdx.z;
  if (j > ny || k > nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  const bool ii = i < nx, ji = j < ny, ki = k < nz;
  const T c = (ii && ji && ki) ? lam[id] : (T)0;
  if (ji && ki) au[id] += ih.x * ((i > 0 ? lam[id - xs] : (T)0) - c);
  if (ii && ki) av[id] += ih.y * ((j > 0 ? lam[id - ys] : (T)0) - c);
  if (ii && ji) aw[id] += ih.z * ((k > 0 ? lam[id - 1] : (T)0) - c);
}

// adjoint of the gradient subtraction in p: one thread per cell of
// [-1,nx-1]x[-1,ny-1]x[-1,nz-1] with at most one index at -1 (the interior and
// the low face ghosts the gradient reads) gathers
// adj_p += (cx*(U_c - U_{c+x}) + cy*(V_c - V_{c+y})) + cz*(W_c - W_{c+z}),
// U/V/W = the velocity adjoints on the written faces [0,n-1], 0 elsewhere
template <typename T>
__global__ __launch_bounds__(256) void k_gradient_adjoint(const T *__restrict__ au, const T *__restrict__ av,
                                                          const T *__restrict__ aw, T *__restrict__ ap, Inv<T> cf,
                                                          int64_t xs, int64_t ys, int nx, int ny, int nz) {
  const int k = (int)(blockIdx.x * TK + threadIdx.x) - 1;
  const int j = (int)(blockIdx.y * TJ + threadIdx.y) - 1;
  const int i = (int)blockIdx.z - 1;
  if (j >= ny || k >= nz) return;
  if ((i < 0) + (j < 0) + (k < 0) > 1) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  const bool in = i >= 0 && j >= 0 && k >= 0;
  const T u0 = in ? au[id] : (T)0, v0 = in ? av[id] : (T)0, w0 = in ? aw[id] : (T)0;
  const T u1 = (j >= 0 && k >= 0 && i + 1 < nx) ? au[id + xs] : (T)0;
  const T v1 = (i >= 0 && k >= 0 && j + 1 < ny) ? av[id + ys] : (T)0;
  const T w1 = (i >= 0 && j >= 0 && k + 1 < nz) ? aw[id + 1] : (T)0;
  ap[id] += (cf.x * (u0 - u1) + cf.y * (v0 - v1)) + cf.z * (w0 - w1);
}

// Sol_LaplacianCentered3DDevice_stencil (sol_laplaciancent3ddev.cu:27-62)
template <typename T>
__global__ __launch_bounds__(256) void k_laplacian(const T *__restrict__ phi, T *__restrict__ d, Inv<T> ih2, T coef,
                                                   int64_t xs, int64_t ys, int nx, int ny, int nz) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  T dp = d[id];
  d[id] = dp + coef * laplacian_cell(phi, id, xs, ys, ih2);
}

// The same d += coef * lap(phi) by rows: each wave owns a row j (V cells
// per lane) and marches along the planes of its i-chunk with planes i-1, i,
// i+1 of its row in registers; rows j-1 and j+1 of plane i are read beside
// them (other waves' rows: L2 hits) and k-neighbours come by DPP with the
// row's z ghosts.  Same operations in the same order as k_laplacian.  Tiles
// are dealt to the XCDs in contiguous runs (gridDim.x a multiple of 8).
constexpr int LAP_NW = 4;
template <typename T, int V>
__global__ __launch_bounds__(64 * LAP_NW) void k_laplacian_rows(DGrid<T> phi, DGrid<T> d, Inv<T> ih2, T coef, int nx,
                                                                int ny, int nz, int tiles_j, int ichunk,
                                                                int ntiles) {
  typedef Vec<T, V> VT;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int per = gridDim.x / 8;
  const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (t >= ntiles) return;
  const int tj = t % tiles_j, ti = t / tiles_j;
  const int j = tj * LAP_NW + w;
  if (j >= ny) return;
  const int k0 = lane * V;
  const bool active = k0 < nz;
  const int kc = active ? k0 : 0;
  const int hi_lane = (nz == 64 * V) ? 63 : nz / V;
  const int gk = lane == 0 ? -1 : nz;
  const bool g_in_data = lane == hi_lane && nz < 64 * V;
  const int i0 = ti * ichunk, i1 = min(nx, i0 + ichunk);
  const T *pr = phi.p + (int64_t)j * phi.ys;
  T *dr = d.p + (int64_t)j * d.ys + kc;
  auto row = [&](int i, int dj) {
    return *reinterpret_cast<const VT *>(pr + (int64_t)i * phi.xs + (int64_t)dj * phi.ys + kc);
  };
  auto rowg = [&](int i, VT &v, T &g) {
    const T *p = pr + (int64_t)i * phi.xs;
    v = *reinterpret_cast<const VT *>(p + kc);
    g = p[gk];
    if (g_in_data) v.v[0] = g;
  };
  VT prev = row(i0 - 1, 0), cur, nxt;
  T gc, gn;
  rowg(i0, cur, gc);
  for (int i = i0; i < i1; ++i) {
    rowg(i + 1, nxt, gn);
    const VT ym = row(i, -1), yp = row(i, 1);
    VT dv = *reinterpret_cast<const VT *>(dr + (int64_t)i * d.xs);
    const T lo = from_lower(cur.v[V - 1], gc), hi = from_upper(cur.v[0], gc);
#pragma unroll
    for (int e = 0; e < V; e++) {
      const T p = cur.v[e], zm = e > 0 ? cur.v[e - 1] : lo, zp = e < V - 1 ? cur.v[e + 1] : hi;
      const T lap = ih2.z * (zp + zm - ((T)2) * p) + ih2.y * (yp.v[e] + ym.v[e] - ((T)2) * p) +
                    ih2.x * (nxt.v[e] + prev.v[e] - ((T)2) * p);
      dv.v[e] = dv.v[e] + coef * lap;
    }
    if (active) *reinterpret_cast<VT *>(dr + (int64_t)i * d.xs) = dv;
    prev = cur;
    cur = nxt;
    gc = gn;
  }
}

template <typename T>
int launch_laplacian_rows(const ocu_grid3d *phi, const ocu_grid3d *dphi, Inv<T> ih2, T coef, int nx, int ny, int nz,
                          hipStream_t stream) {
  const int V = nz <= 64 ? 1 : nz <= 128 ? 2 : nz <= 256 ? 4 : 0;
  if (!V || nz % V || !row_aligned(phi, V) || !row_aligned(dphi, V)) return 0;
  const int tiles_j = (ny + LAP_NW - 1) / LAP_NW;
  // about 2048 blocks: short chunks re-read the x-neighbour planes
  const int ichunk = std::max(1, (int)(((int64_t)nx * tiles_j + 2047) / 2048));
  const int ntiles = tiles_j * ((nx + ichunk - 1) / ichunk);
  const int nblocks = (ntiles + 7) / 8 * 8;
  const dim3 blk(64, LAP_NW);
#define LAPR(VV)                                                                                                   \
  k_laplacian_rows<T, VV><<<nblocks, blk, 0, stream>>>(dgrid<T>(phi), dgrid<T>(dphi), ih2, coef, nx, ny, nz, tiles_j, \
                                                       ichunk, ntiles)
  if (V == 1) LAPR(1);
  else if (V == 2) LAPR(2);
  else LAPR(4);
#undef LAPR
  return 1;
}

// Eqn_IncompressibleNS3D_add_thermal_force (eqn_incompressns3d.cu:25-44)
template <typename T>
__global__ __launch_bounds__(256) void k_thermal(T *__restrict__ d, const T *__restrict__ t, T coef, int64_t xs,
                                                 int64_t ys, int64_t nbr, int nx, int ny, int nz) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= ny || k >= nz) return;
  const int64_t id = (int64_t)i * xs + (int64_t)j * ys + k;
  d[id] += thermal_cell(t, id, nbr, coef);
}

template <typename T>
T *at0(const ocu_grid3d *g) {
  return (T *)g->base + g->shift;
}

int check_mac(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, int nx, int ny, int nz,
              const char *what) {
  if (int e = check_grid(u, what)) return e;
  if (int e = check_grid(v, what)) return e;
  if (int e = check_grid(w, what)) return e;
  // check_valid_mac_dimensions (ocustorage/grid3dops.cpp:23-45)
  OCU_REQUIRE(u->nx == nx + 1 && u->ny == ny && u->nz == nz, "%s - u dimensions mismatch", what);
  OCU_REQUIRE(v->nx == nx && v->ny == ny + 1 && v->nz == nz, "%s - v dimensions mismatch", what);
  OCU_REQUIRE(w->nx == nx && w->ny == ny && w->nz == nz + 1, "%s - w dimensions mismatch", what);
  OCU_REQUIRE(layout_match(u, v) && layout_match(u, w), "%s - u,v,w layout mismatch", what);
  OCU_REQUIRE(u->dtype == v->dtype && u->dtype == w->dtype && u->dtype != OCU_I32, "%s - dtype mismatch", what);
  OCU_REQUIRE(u->gx >= 1 && u->gy >= 1 && u->gz >= 1, "%s - must have at least one ghost cell", what);
  return OCU_OK;
}

int check_same(const ocu_grid3d *a, const ocu_grid3d *b, const char *what) {
  if (int e = check_grid(b, what)) return e;
  OCU_REQUIRE(layout_match(a, b) && a->dtype == b->dtype, "%s - layout mismatch", what);
  return OCU_OK;
}

template <typename T>
Inv<T> inv(double hx, double hy, double hz) {
  return Inv<T>{(T)(1 / hx), (T)(1 / hy), (T)(1 / hz)};
}

}  // namespace

extern "C" {

int ocu_divergence(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *div, int nx,
                   int ny, int nz, double hx, double hy, double hz, void *stream) {
  const char *what = "Sol_Divergence3DDevice::solve";
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if (int e = check_same(u, div, what)) return e;
  OCU_REQUIRE(div->nx == nx && div->ny == ny && div->nz == nz, "%s - invalid dimensions for divergence", what);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  dim3 g = stencil_grid(nx, ny, nz);
  if (u->dtype == OCU_F64)
    k_divergence<double><<<g, stencil_block(), 0, S(stream)>>>(at0<double>(u), at0<double>(v), at0<double>(w),
                                                               at0<double>(div), inv<double>(hx, hy, hz), xs, ys, nx, ny, nz);
  else
    k_divergence<float><<<g, stencil_block(), 0, S(stream)>>>(at0<float>(u), at0<float>(v), at0<float>(w),
                                                              at0<float>(div), inv<float>(hx, hy, hz), xs, ys, nx, ny, nz);
  OCU_HIP_LAUNCHED("ocu_divergence");
}

int ocu_gradient_sub(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *p,
                     double coefficient, double hx, double hy, double hz, void *stream) {
  const char *what = "Sol_Gradient3DDevice::solve";
  if (int e = check_grid(p, what)) return e;
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if (int e = check_same(u, p, what)) return e;
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  dim3 g = stencil_grid(nx, ny, nz);
  if (u->dtype == OCU_F64)
    k_gradient<double><<<g, stencil_block(), 0, S(stream)>>>(at0<double>(u), at0<double>(v), at0<double>(w),
                                                             at0<double>(p), coefficient, inv<double>(hx, hy, hz), xs,
                                                             ys, nx, ny, nz);
  else
    k_gradient<float><<<g, stencil_block(), 0, S(stream)>>>(at0<float>(u), at0<float>(v), at0<float>(w), at0<float>(p),
                                                            (float)coefficient, inv<float>(hx, hy, hz), xs, ys, nx, ny,
                                                            nz);
  OCU_HIP_LAUNCHED("ocu_gradient_sub");
}

int ocu_self_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                    const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz, double hx, double hy,
                    double hz, int scheme, void *stream) {
  OCU_REQUIRE(scheme == OCU_IT_FIRST_ORDER_UPWIND || scheme == OCU_IT_SECOND_ORDER_CENTERED,
              "Sol_SelfAdvection3DDevice::solve - invalid interpolation type %d", scheme);
  return ocu_self_advect_ex(u, v, w, du, dv, dw, nx, ny, nz, hx, hy, hz, scheme, 0, stream);
}

int ocu_self_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *du,
                       const ocu_grid3d *dv, const ocu_grid3d *dw, int nx, int ny, int nz, double hx, double hy,
                       double hz, int scheme, int periodic_axes, void *stream) {
  const char *what = "Sol_SelfAdvection3DDevice::solve";
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if (int e = check_same(u, du, what)) return e;
  if (int e = check_same(u, dv, what)) return e;
  if (int e = check_same(u, dw, what)) return e;
  OCU_REQUIRE(scheme == OCU_IT_FIRST_ORDER_UPWIND || scheme == OCU_IT_SECOND_ORDER_CENTERED ||
                  scheme == OCU_IT_SECOND_ORDER_UPWIND,
              "%s - invalid interpolation type %d", what, scheme);
  OCU_REQUIRE(periodic_axes >= 0 && periodic_axes <= 7, "%s - invalid periodic axes mask %d", what, periodic_axes);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  dim3 g = stencil_grid(nx, ny, nz);
  if (scheme == OCU_IT_SECOND_ORDER_UPWIND) {
    if (u->dtype == OCU_F64)
      k_self_advect_up2<double><<<g, stencil_block(), 0, S(stream)>>>(
          at0<double>(u), at0<double>(v), at0<double>(w), at0<double>(du), at0<double>(dv), at0<double>(dw),
          inv<double>(hx, hy, hz), xs, ys, nx, ny, nz, periodic_axes);
    else
      k_self_advect_up2<float><<<g, stencil_block(), 0, S(stream)>>>(
          at0<float>(u), at0<float>(v), at0<float>(w), at0<float>(du), at0<float>(dv), at0<float>(dw),
          inv<float>(hx, hy, hz), xs, ys, nx, ny, nz, periodic_axes);
    OCU_HIP_LAUNCHED("ocu_self_advect");
  }
  if (u->dtype == OCU_F64)
    k_self_advect<double><<<g, stencil_block(), 0, S(stream)>>>(at0<double>(u), at0<double>(v), at0<double>(w),
                                                                at0<double>(du), at0<double>(dv), at0<double>(dw),
                                                                inv<double>(hx, hy, hz), xs, ys, nx, ny, nz, scheme);
  else
    k_self_advect<float><<<g, stencil_block(), 0, S(stream)>>>(at0<float>(u), at0<float>(v), at0<float>(w),
                                                               at0<float>(du), at0<float>(dv), at0<float>(dw),
                                                               inv<float>(hx, hy, hz), xs, ys, nx, ny, nz, scheme);
  OCU_HIP_LAUNCHED("ocu_self_advect");
}

int ocu_passive_advect(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *phi,
                       const ocu_grid3d *dphi, int nx, int ny, int nz, double hx, double hy, double hz, int scheme,
                       void *stream) {
  OCU_REQUIRE(scheme == OCU_IT_FIRST_ORDER_UPWIND || scheme == OCU_IT_SECOND_ORDER_CENTERED,
              "Sol_PassiveAdvection3DDevice::solve - invalid interpolation type %d", scheme);
  return ocu_passive_advect_ex(u, v, w, phi, dphi, nx, ny, nz, hx, hy, hz, scheme, 0, stream);
}

int ocu_passive_advect_ex(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *phi,
                          const ocu_grid3d *dphi, int nx, int ny, int nz, double hx, double hy, double hz,
                          int scheme, int periodic_axes, void *stream) {
  const char *what = "Sol_PassiveAdvection3DDevice::solve";
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if (int e = check_same(u, phi, what)) return e;
  if (int e = check_same(u, dphi, what)) return e;
  OCU_REQUIRE(scheme == OCU_IT_FIRST_ORDER_UPWIND || scheme == OCU_IT_SECOND_ORDER_CENTERED ||
                  scheme == OCU_IT_SECOND_ORDER_UPWIND,
              "%s - invalid interpolation type %d", what, scheme);
  OCU_REQUIRE(periodic_axes >= 0 && periodic_axes <= 7, "%s - invalid periodic axes mask %d", what, periodic_axes);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  dim3 g = stencil_grid(nx, ny, nz);
  if (scheme == OCU_IT_SECOND_ORDER_UPWIND) {
    if (u->dtype == OCU_F64)
      k_passive_advect_up2<double><<<g, stencil_block(), 0, S(stream)>>>(
          at0<double>(u), at0<double>(v), at0<double>(w), at0<double>(phi), at0<double>(dphi),
          inv<double>(hx, hy, hz), xs, ys, nx, ny, nz, periodic_axes);
    else
      k_passive_advect_up2<float><<<g, stencil_block(), 0, S(stream)>>>(
          at0<float>(u), at0<float>(v), at0<float>(w), at0<float>(phi), at0<float>(dphi),
          inv<float>(hx, hy, hz), xs, ys, nx, ny, nz, periodic_axes);
    OCU_HIP_LAUNCHED("ocu_passive_advect");
  }
  // planes per thread: long marches, but enough blocks to fill the chip
  const int64_t tiles = (int64_t)g.x * g.y;
  int ichunk = 16;
  while (ichunk > 1 && tiles * ((nx + ichunk - 1) / ichunk) < 4096) ichunk /= 2;
  const dim3 gm(g.x, g.y, (nx + ichunk - 1) / ichunk);
  if (u->dtype == OCU_F64)
    k_passive_advect_march<double><<<gm, stencil_block(), 0, S(stream)>>>(
        at0<double>(u), at0<double>(v), at0<double>(w), at0<double>(phi), at0<double>(dphi), inv<double>(hx, hy, hz),
        xs, ys, nx, ny, nz, scheme, ichunk);
  else
    k_passive_advect_march<float><<<gm, stencil_block(), 0, S(stream)>>>(
        at0<float>(u), at0<float>(v), at0<float>(w), at0<float>(phi), at0<float>(dphi), inv<float>(hx, hy, hz), xs,
        ys, nx, ny, nz, scheme, ichunk);
  OCU_HIP_LAUNCHED("ocu_passive_advect");
}

int ocu_passive_advect_adjoint(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_grid3d *phi,
                               const ocu_grid3d *lam, const ocu_grid3d *lam_out, const ocu_grid3d *adj_u,
                               const ocu_grid3d *adj_v, const ocu_grid3d *adj_w, double scale, int nx, int ny, int nz,
                               double hx, double hy, double hz, int scheme, void *stream) {
  const char *what = "ocu_passive_advect_adjoint";
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if (int e = check_mac(adj_u, adj_v, adj_w, nx, ny, nz, what)) return e;
  if (int e = check_same(u, adj_u, what)) return e;
  if (int e = check_same(u, phi, what)) return e;
  if (int e = check_same(u, lam, what)) return e;
  if (int e = check_same(u, lam_out, what)) return e;
  OCU_REQUIRE(lam->base != lam_out->base, "%s - lam_out must not alias lam", what);
  OCU_REQUIRE(scheme == OCU_IT_FIRST_ORDER_UPWIND || scheme == OCU_IT_SECOND_ORDER_CENTERED,
              "%s - invalid interpolation type %d", what, scheme);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  const int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  // planes per thread: long marches, but enough blocks to fill the chip
  const int64_t tiles = (int64_t)((nz + 2 + TK - 1) / TK) * ((ny + 2 + TJ - 1) / TJ);
  int ichunk = 16;
  while (ichunk > 1 && tiles * ((nx + 2 + ichunk - 1) / ichunk) < 4096) ichunk /= 2;
  const dim3 g((nz + 2 + TK - 1) / TK, (ny + 2 + TJ - 1) / TJ, (nx + 2 + ichunk - 1) / ichunk);
  if (u->dtype == OCU_F64)
    k_passive_advect_adjoint_march<double><<<g, stencil_block(), 0, S(stream)>>>(
        at0<double>(u), at0<double>(v), at0<double>(w), at0<double>(phi), at0<double>(lam), at0<double>(lam_out),
        at0<double>(adj_u), at0<double>(adj_v), at0<double>(adj_w), inv<double>(hx, hy, hz), (double)scale, xs, ys,
        nx, ny, nz, scheme, ichunk);
  else
    k_passive_advect_adjoint_march<float><<<g, stencil_block(), 0, S(stream)>>>(
        at0<float>(u), at0<float>(v), at0<float>(w), at0<float>(phi), at0<float>(lam), at0<float>(lam_out),
        at0<float>(adj_u), at0<float>(adj_v), at0<float>(adj_w), inv<float>(hx, hy, hz), (float)scale, xs, ys, nx,
        ny, nz, scheme, ichunk);
  OCU_HIP_LAUNCHED(what);
}

int ocu_bc_adjoint_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, void *stream) {
  const char *what = "ocu_bc_adjoint_nocorners";
  if (int e = check_grid(g, what)) return e;
  OCU_REQUIRE(bc, "%s: null bc", what);
  OCU_REQUIRE(g->gx >= 1 && g->gy >= 1 && g->gz >= 1, "%s - must have at least 1 ghost point", what);
  OCU_REQUIRE(g->dtype != OCU_I32, "%s: int grid", what);
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  GhostCoef gc;
  for (int f = 0; f < 6; f++) {
    const int t = all[f]->type;
    OCU_REQUIRE(t == OCU_BC_NONE || t == OCU_BC_PERIODIC || t == OCU_BC_DIRICHELET || t == OCU_BC_NEUMANN,
                "%s - invalid boundary condition type %d", what, t);
    gc.c[f] = t == OCU_BC_DIRICHELET ? -1.0f : t == OCU_BC_NONE ? 0.0f : 1.0f;
    gc.per[f] = t == OCU_BC_PERIODIC;
  }
  if ((int64_t)g->nx * g->ny * g->nz == 0) return OCU_OK;
  const int64_t xs = (int64_t)g->pny * g->pnz, ys = g->pnz;
  const dim3 grid = stencil_grid(g->nx, g->ny, g->nz);
  if (g->dtype == OCU_F64)
    k_bc_adjoint_nocorners<double><<<grid, stencil_block(), 0, S(stream)>>>(at0<double>(g), gc, xs, ys, g->nx, g->ny,
                                                                             g->nz);
  else
    k_bc_adjoint_nocorners<float><<<grid, stencil_block(), 0, S(stream)>>>(at0<float>(g), gc, xs, ys, g->nx, g->ny,
                                                                            g->nz);
  OCU_HIP_LAUNCHED(what);
}

int ocu_mac_adjoint_normal(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                           void *stream) {
  const char *what = "ocu_mac_adjoint_normal";
  OCU_REQUIRE(u && v && w && bc, "%s: null argument", what);
  const int nx = u->nx - 1, ny = u->ny, nz = u->nz;
  if (int e = check_mac(u, v, w, nx, ny, nz, what)) return e;
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  const int64_t xs = (int64_t)u->pny * u->pnz, ys = u->pnz;
  auto launch = [&](const ocu_grid3d *g, int64_t s_n, int64_t s_a, int64_t s_b, int n, int na, int nb, int per_hi,
                    int inflow_lo) {
    const int cnt = na * nb, nb_ = (cnt + 255) / 256;
    if (g->dtype == OCU_F64)
      k_mac_adjoint_normal<double><<<nb_, 256, 0, S(stream)>>>(at0<double>(g), s_n, s_a, s_b, n, na, nb, per_hi,
                                                                inflow_lo);
    else
      k_mac_adjoint_normal<float><<<nb_, 256, 0, S(stream)>>>(at0<float>(g), s_n, s_a, s_b, n, na, nb, per_hi,
                                                               inflow_lo);
  };
  const int IN = OCU_BC_FORCED_INFLOW_VARIABLE_SLIP, PER = OCU_BC_PERIODIC;
  launch(u, xs, ys, 1, nx, ny, nz, bc->xpos.type == PER, bc->xneg.type == IN);
  launch(v, ys, xs, 1, ny, nx, nz, bc->ypos.type == PER, bc->yneg.type == IN);
  launch(w, 1, xs, ys, nz, nx, ny, bc->zpos.type == PER, bc->zneg.type == IN);
  OCU_HIP_LAUNCHED(what);
}

int ocu_divergence_adjoint(const ocu_grid3d *lam, const ocu_grid3d *adj_u, const ocu_grid3d *adj_v,
                           const ocu_grid3d *adj_w, int nx, int ny, int nz, double hx, double hy, double hz,
                           void *stream) {
  const char *what = "ocu_divergence_adjoint";
  if (int e = check_mac(adj_u, adj_v, adj_w, nx, ny, nz, what)) return e;
  if (int e = check_same(adj_u, lam, what)) return e;
  OCU_REQUIRE(lam->nx == nx && lam->ny == ny && lam->nz == nz, "%s - invalid dimensions for lam", what);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  const int64_t xs = (int64_t)adj_u->pny * adj_u->pnz, ys = adj_u->pnz;
  const dim3 g((nz + 1 + TK - 1) / TK, (ny + 1 + TJ - 1) / TJ, nx + 1);
  if (lam->dtype == OCU_F64)
    k_divergence_adjoint<double><<<g, stencil_block(), 0, S(stream)>>>(
        at0<double>(lam), at0<double>(adj_u), at0<double>(adj_v), at0<double>(adj_w), inv<double>(hx, hy, hz), xs, ys,
        nx, ny, nz);
  else
    k_divergence_adjoint<float><<<g, stencil_block(), 0, S(stream)>>>(
        at0<float>(lam), at0<float>(adj_u), at0<float>(adj_v), at0<float>(adj_w), inv<float>(hx, hy, hz), xs, ys, nx,
        ny, nz);
  OCU_HIP_LAUNCHED(what);
}

int ocu_gradient_adjoint(const ocu_grid3d *adj_u, const ocu_grid3d *adj_v, const ocu_grid3d *adj_w,
                         const ocu_grid3d *adj_p, double coefficient, double hx, double hy, double hz, void *stream) {
  const char *what = "ocu_gradient_adjoint";
  if (int e = check_grid(adj_p, what)) return e;
  const int nx = adj_p->nx, ny = adj_p->ny, nz = adj_p->nz;
  if (int e = check_mac(adj_u, adj_v, adj_w, nx, ny, nz, what)) return e;
  if (int e = check_same(adj_u, adj_p, what)) return e;
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  const int64_t xs = (int64_t)adj_u->pny * adj_u->pnz, ys = adj_u->pnz;
  const dim3 g((nz + 1 + TK - 1) / TK, (ny + 1 + TJ - 1) / TJ, nx + 1);
  if (adj_p->dtype == OCU_F64) {
    const Inv<double> ih = inv<double>(hx, hy, hz), c = {coefficient * ih.x, coefficient * ih.y, coefficient * ih.z};
    k_gradient_adjoint<double><<<g, stencil_block(), 0, S(stream)>>>(at0<double>(adj_u), at0<double>(adj_v),
                                                                      at0<double>(adj_w), at0<double>(adj_p), c, xs,
                                                                      ys, nx, ny, nz);
  } else {
    const Inv<float> ih = inv<float>(hx, hy, hz);
    const float cf = (float)coefficient;
    const Inv<float> c = {cf * ih.x, cf * ih.y, cf * ih.z};
    k_gradient_adjoint<float><<<g, stencil_block(), 0, S(stream)>>>(at0<float>(adj_u), at0<float>(adj_v),
                                                                     at0<float>(adj_w), at0<float>(adj_p), c, xs, ys,
                                                                     nx, ny, nz);
  }
  OCU_HIP_LAUNCHED(what);
}

int ocu_laplacian_acc(const ocu_grid3d *phi, const ocu_grid3d *dphi, int nx, int ny, int nz, double hx, double hy,
                      double hz, double coefficient, void *stream) {
  const char *what = "Sol_LaplacianCentered3DDevice::solve";
  if (int e = check_grid(phi, what)) return e;
  if (int e = check_same(phi, dphi, what)) return e;
  OCU_REQUIRE(phi->nx == nx && phi->ny == ny && phi->nz == nz, "%s - phi dimensions mismatch", what);
  OCU_REQUIRE(phi->gx >= 1 && phi->gy >= 1 && phi->gz >= 1, "%s - must have >= 1 ghost cell", what);
  OCU_REQUIRE(phi->dtype != OCU_I32, "%s - int grid", what);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)phi->pny * phi->pnz, ys = phi->pnz;
  dim3 g = stencil_grid(nx, ny, nz);
  const Inv<double> ih2d{(double)(1 / (hx * hx)), (double)(1 / (hy * hy)), (double)(1 / (hz * hz))};
  const Inv<float> ih2f{(float)(1 / (hx * hx)), (float)(1 / (hy * hy)), (float)(1 / (hz * hz))};
  if (phi->dtype == OCU_F64 ? launch_laplacian_rows<double>(phi, dphi, ih2d, coefficient, nx, ny, nz, S(stream))
                            : launch_laplacian_rows<float>(phi, dphi, ih2f, (float)coefficient, nx, ny, nz,
                                                           S(stream)))
    OCU_HIP_LAUNCHED("ocu_laplacian_acc");
  if (phi->dtype == OCU_F64)
    k_laplacian<double><<<g, stencil_block(), 0, S(stream)>>>(
        at0<double>(phi), at0<double>(dphi),
        Inv<double>{(double)(1 / (hx * hx)), (double)(1 / (hy * hy)), (double)(1 / (hz * hz))}, coefficient, xs, ys,
        nx, ny, nz);
  else
    k_laplacian<float><<<g, stencil_block(), 0, S(stream)>>>(
        at0<float>(phi), at0<float>(dphi),
        Inv<float>{(float)(1 / (hx * hx)), (float)(1 / (hy * hy)), (float)(1 / (hz * hz))}, (float)coefficient, xs,
        ys, nx, ny, nz);
  OCU_HIP_LAUNCHED("ocu_laplacian_acc");
}

int ocu_thermal_force(const ocu_grid3d *d, const ocu_grid3d *t, int nx, int ny, int nz, int vd, double coefficient,
                      void *stream) {
  const char *what = "Eqn_IncompressibleNS3D::add_thermal_force";
  if (int e = check_grid(d, what)) return e;
  if (int e = check_same(d, t, what)) return e;
  OCU_REQUIRE(vd == OCU_DIR_XPOS || vd == OCU_DIR_XNEG || vd == OCU_DIR_YPOS || vd == OCU_DIR_YNEG ||
                  vd == OCU_DIR_ZPOS || vd == OCU_DIR_ZNEG,
              "%s - bad vertical_direction %d", what, vd);
  if ((int64_t)nx * ny * nz == 0) return OCU_OK;
  int64_t xs = (int64_t)t->pny * t->pnz, ys = t->pnz;
  int64_t nbr = (vd & OCU_DIR_XAXIS_FLAG) ? xs : (vd & OCU_DIR_YAXIS_FLAG) ? ys : 1;
  dim3 g = stencil_grid(nx, ny, nz);
  if (d->dtype == OCU_F64)
    k_thermal<double><<<g, stencil_block(), 0, S(stream)>>>(at0<double>(d), at0<double>(t), coefficient, xs, ys, nbr,
                                                            nx, ny, nz);
  else
    k_thermal<float><<<g, stencil_block(), 0, S(stream)>>>(at0<float>(d), at0<float>(t), (float)coefficient, xs, ys,
                                                           nbr, nx, ny, nz);
  OCU_HIP_LAUNCHED("ocu_thermal_force");
}

}  // extern "C"
