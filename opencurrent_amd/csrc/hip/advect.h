// Cell-level bodies of the MAC-grid operators of the NS step: self- and
// passive advection (first-order upwind, centred and second-order upwind),
// the centred Laplacian and the thermal force.  The per-operator kernels
// (ns_ops.hip) and the fused NS step (ns_step.hip) both call these, so the
// two paths perform the same operations in the same order.
#pragma once

#include "ocu_hip_common.h"

namespace ocu_hip {

template <typename T>
struct Inv {
  T x, y, z;
};

// interpolation.h:35-48
template <typename T>
__device__ __forceinline__ T interp(int scheme, T u, T x0, T x1) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return ((T).5) * (x1 + x0);
  return u < 0 ? x1 : x0;
}

// Reading an array around one cell: A(o) is the element at offset o from the
// cell (o = di*xs + dj*ys + dk in the strides the caller passes).  GAt reads
// global memory; the marching NS step passes accessors over its LDS window.
template <typename T>
struct GAt {
  const T *__restrict__ p;  // the cell
  __device__ __forceinline__ T operator()(int64_t o) const { return p[o]; }
};
// GAt with the samples at x offsets -xs, 0, +xs held in registers (a
// kernel marching along x); with xs > ys > 1 known to the compiler every
// offset test folds away
template <typename T>
struct XAt {
  const T *__restrict__ p;
  T m, c, n;
  int64_t xs;
  __device__ __forceinline__ T operator()(int64_t o) const {
    return o == xs ? n : o == -xs ? m : o == 0 ? c : p[o];
  }
};

// Self-advection tendencies of the u, v and w faces at a cell (flux form,
// Advection3DF/D_apply_upwind_TEX, sol_selfadvection3ddev.cu:136-337)
template <typename T, typename A>
__device__ __forceinline__ void self_advect_at(const A &U, const A &V, const A &W, int64_t xs, int64_t ys, Inv<T> ih,
                                               int scheme, T &du, T &dv, T &dw) {
  const int64_t pi = xs, pj = ys, pk = 1, mi = -xs, mj = -ys, mk = -1;
  const T half = (T).5;
  T u0 = U(0), v0 = V(0), w0 = W(0);

  T u_iph = half * (u0 + U(pi));
  T u_imh = half * (u0 + U(mi));
  T duu = (u_iph * interp(scheme, u_iph, u0, U(pi))) - (u_imh * interp(scheme, u_imh, U(mi), u0));
  T v_jph = half * (V(pj) + V(pj - xs));
  T v_jmh = half * (v0 + V(mi));
  T dvu = (v_jph * interp(scheme, v_jph, u0, U(pj))) - (v_jmh * interp(scheme, v_jmh, U(mj), u0));
  T w_kph = half * (W(pk) + W(pk - xs));
  T w_kmh = half * (w0 + W(mi));
  T dwu = (w_kph * interp(scheme, w_kph, u0, U(pk))) - (w_kmh * interp(scheme, w_kmh, U(mk), u0));
  du = -duu * ih.x - dvu * ih.y - dwu * ih.z;

  T u_atv_iph = half * (U(pi) + U(pi - ys));
  T u_atv_imh = half * (u0 + U(mj));
  T duv = (u_atv_iph * interp(scheme, u_atv_iph, v0, V(pi))) - (u_atv_imh * interp(scheme, u_atv_imh, V(mi), v0));
  T v_iph = half * (v0 + V(pj));
  T v_imh = half * (v0 + V(mj));
  T dvv = (v_iph * interp(scheme, v_iph, v0, V(pj))) - (v_imh * interp(scheme, v_imh, V(mj), v0));
  T w_atv_kph = half * (W(pk) + W(pk - ys));
  T w_atv_kmh = half * (w0 + W(mj));
  T dwv = (w_atv_kph * interp(scheme, w_atv_kph, v0, V(pk))) - (w_atv_kmh * interp(scheme, w_atv_kmh, V(mk), v0));
  dv = -duv * ih.x - dvv * ih.y - dwv * ih.z;

  T u_atw_iph = half * (U(pi) + U(pi - 1));
  T u_atw_imh = half * (u0 + U(mk));
  T duw = (u_atw_iph * interp(scheme, u_atw_iph, w0, W(pi))) - (u_atw_imh * interp(scheme, u_atw_imh, W(mi), w0));
  T v_atw_jph = half * (V(pj) + V(pj - 1));
  T v_atw_jmh = half * (v0 + V(mk));
  T dvw = (v_atw_jph * interp(scheme, v_atw_jph, w0, W(pj))) - (v_atw_jmh * interp(scheme, v_atw_jmh, W(mj), w0));
  T w_iph = half * (w0 + W(pk));
  T w_imh = half * (w0 + W(mk));
  T dww = (w_iph * interp(scheme, w_iph, w0, W(pk))) - (w_imh * interp(scheme, w_imh, W(mk), w0));
  dw = -duw * ih.x - dvw * ih.y - dww * ih.z;
}
template <typename T>
__device__ __forceinline__ void self_advect_cell(const T *__restrict__ U, const T *__restrict__ V,
                                                 const T *__restrict__ W, int64_t id, int64_t xs, int64_t ys,
                                                 Inv<T> ih, int scheme, T &du, T &dv, T &dw) {
  self_advect_at<T>(GAt<T>{U + id}, GAt<T>{V + id}, GAt<T>{W + id}, xs, ys, ih, scheme, du, dv, dw);
}

// Passive advection tendency of phi at a cell
// (Sol_PassiveAdvection3D_apply_interp, sol_passiveadvection3ddev.cu:31-78)
template <typename T, typename AU, typename AV, typename AW, typename AP>
__device__ __forceinline__ T passive_advect_at(const AU &u, const AV &v, const AW &w, const AP &phi, int64_t xs,
                                               int64_t ys, Inv<T> ih, int scheme) {
  T p = phi(0);
  T u_pi = u(xs), u0 = u(0);
  T du = u_pi * interp(scheme, u_pi, p, phi(xs)) - u0 * interp(scheme, u0, phi(-xs), p);
  T v_pj = v(ys), v0 = v(0);
  T dv = v_pj * interp(scheme, v_pj, p, phi(ys)) - v0 * interp(scheme, v0, phi(-ys), p);
  T w_pk = w(1), w0 = w(0);
  T dw = w_pk * interp(scheme, w_pk, p, phi(1)) - w0 * interp(scheme, w0, phi(-1), p);
  return (-du * ih.x) - (dv * ih.y) - (dw * ih.z);
}
template <typename T>
__device__ __forceinline__ T passive_advect_cell(const T *__restrict__ u, const T *__restrict__ v,
                                                 const T *__restrict__ w, const T *__restrict__ phi, int64_t id,
                                                 int64_t xs, int64_t ys, Inv<T> ih, int scheme) {
  return passive_advect_at<T>(GAt<T>{u + id}, GAt<T>{v + id}, GAt<T>{w + id}, GAt<T>{phi + id}, xs, ys, ih, scheme);
}

// ---- second-order (linear) upwind, OCU_IT_SECOND_ORDER_UPWIND (no reference
// counterpart) ----
//
// The face value between xL and xR is extrapolated from the two upstream
// samples, xL + .5*(xL - xLL) for vel >= 0 and xR + .5*(xR - xRR) for vel < 0
// (vel == 0 takes the low side, as the reference's upwind does).  Only the
// second upstream sample can leave a ghost layer: along axis a the layer of
// a cell-centred or tangential array spans [-1, n_a], that of the face array
// normal to a [-1, n_a + 1].  Beyond it the sample wraps (periodic axis) or
// is missing, and the face falls back to first order.
template <typename T>
__device__ __forceinline__ T up2(T vel, T xLL, bool hLL, T xL, T xR, T xRR, bool hRR) {
  if (vel < 0) return hRR ? xR + ((T).5) * (xR - xRR) : xR;
  return hLL ? xL + ((T).5) * (xL - xLL) : xL;
}

// Offsets (in elements) of the samples two cells below and above cell c.
struct Far2 {
  int64_t lo, hi;
  bool hlo, hhi;
};

__device__ __forceinline__ Far2 far2(int c, int n, int top, int64_t s, bool per) {
  Far2 f;
  const int wlo = (((c - 2) % n) + n) % n, whi = (c + 2) % n;
  f.hlo = c >= 1 || per;
  f.hhi = c + 2 <= top || per;
  f.lo = (c >= 1 ? (int64_t)-2 : (int64_t)(wlo - c)) * s;
  f.hi = (c + 2 <= top ? (int64_t)2 : (int64_t)(whi - c)) * s;
  return f;
}
// far2 for a cell whose samples do not wrap (not within two cells of a
// periodic boundary): the same flags and offsets with the offsets constant
__device__ __forceinline__ Far2 far2_nowrap(int c, int top, int64_t s) {
  Far2 f;
  f.hlo = c >= 1;
  f.hhi = c + 2 <= top;
  f.lo = -2 * s;
  f.hi = 2 * s;
  return f;
}

// (vh * face_hi) - (vl * face_lo) for the array X along the axis of stride s
template <typename T, typename A>
__device__ __forceinline__ T dflux2(const A &X, int64_t s, const Far2 &f, T vh, T vl) {
  const T x0 = X(0), xm = X(-s), xp = X(s);
  const T xpp = f.hhi ? X(f.hi) : (T)0;
  const T xmm = f.hlo ? X(f.lo) : (T)0;
  return (vh * up2(vh, xm, true, x0, xp, xpp, f.hhi)) - (vl * up2(vl, xmm, f.hlo, xm, x0, xp, true));
}

// self-advection with second-order upwind face values: the face velocities
// and flux terms of self_advect_at.  WRAP: far samples may wrap around a
// periodic axis (far2); otherwise the caller guarantees they do not.
template <bool WRAP, typename T, typename A>
__device__ __forceinline__ void self_advect_up2_at(const A &U, const A &V, const A &W, int64_t xs, int64_t ys,
                                                   Inv<T> ih, int i, int j, int k, int nx, int ny, int nz, int pmask,
                                                   T &DU, T &DV, T &DW) {
  const int64_t pi = xs, pj = ys, pk = 1, mi = -xs, mj = -ys, mk = -1;
  const bool px = pmask & 1, py = pmask & 2, pz = pmask & 4;
  auto far = [&](int c, int n, int top, int64_t s, bool per) {
    return WRAP ? far2(c, n, top, s, per) : far2_nowrap(c, top, s);
  };
  // tangential (t) and normal (n) ghost extents per axis
  const Far2 fxt = far(i, nx, nx, xs, px), fxn = far(i, nx, nx + 1, xs, px);
  const Far2 fyt = far(j, ny, ny, ys, py), fyn = far(j, ny, ny + 1, ys, py);
  const Far2 fzt = far(k, nz, nz, 1, pz), fzn = far(k, nz, nz + 1, 1, pz);
  const T half = (T).5;
  const T u0 = U(0), v0 = V(0), w0 = W(0);

  const T duu = dflux2(U, xs, fxn, half * (u0 + U(pi)), half * (u0 + U(mi)));
  const T dvu = dflux2(U, ys, fyt, half * (V(pj) + V(pj - xs)), half * (v0 + V(mi)));
  const T dwu = dflux2(U, (int64_t)1, fzt, half * (W(pk) + W(pk - xs)), half * (w0 + W(mi)));
  DU = -duu * ih.x - dvu * ih.y - dwu * ih.z;

  const T duv = dflux2(V, xs, fxt, half * (U(pi) + U(pi - ys)), half * (u0 + U(mj)));
  const T dvv = dflux2(V, ys, fyn, half * (v0 + V(pj)), half * (v0 + V(mj)));
  const T dwv = dflux2(V, (int64_t)1, fzt, half * (W(pk) + W(pk - ys)), half * (w0 + W(mj)));
  DV = -duv * ih.x - dvv * ih.y - dwv * ih.z;

  const T duw = dflux2(W, xs, fxt, half * (U(pi) + U(pi - 1)), half * (u0 + U(mk)));
  const T dvw = dflux2(W, ys, fyt, half * (V(pj) + V(pj - 1)), half * (v0 + V(mk)));
  const T dww = dflux2(W, (int64_t)1, fzn, half * (w0 + W(pk)), half * (w0 + W(mk)));
  DW = -duw * ih.x - dvw * ih.y - dww * ih.z;
}
template <typename T>
__device__ __forceinline__ void self_advect_up2_cell(const T *__restrict__ U, const T *__restrict__ V,
                                                     const T *__restrict__ W, int64_t id, int64_t xs, int64_t ys,
                                                     Inv<T> ih, int i, int j, int k, int nx, int ny, int nz,
                                                     int pmask, T &DU, T &DV, T &DW) {
  self_advect_up2_at<true, T>(GAt<T>{U + id}, GAt<T>{V + id}, GAt<T>{W + id}, xs, ys, ih, i, j, k, nx, ny, nz,
                              pmask, DU, DV, DW);
}

// passive advection with second-order upwind face values
template <bool WRAP, typename T, typename A>
__device__ __forceinline__ T passive_advect_up2_at(const A &u, const A &v, const A &w, const A &phi, int64_t xs,
                                                   int64_t ys, Inv<T> ih, int i, int j, int k, int nx, int ny, int nz,
                                                   int pmask) {
  auto far = [&](int c, int n, int top, int64_t s, bool per) {
    return WRAP ? far2(c, n, top, s, per) : far2_nowrap(c, top, s);
  };
  const T du = dflux2(phi, xs, far(i, nx, nx, xs, pmask & 1), u(xs), u(0));
  const T dv = dflux2(phi, ys, far(j, ny, ny, ys, pmask & 2), v(ys), v(0));
  const T dw = dflux2(phi, (int64_t)1, far(k, nz, nz, 1, pmask & 4), w(1), w(0));
  return (-du * ih.x) - (dv * ih.y) - (dw * ih.z);
}
template <typename T>
__device__ __forceinline__ T passive_advect_up2_cell(const T *__restrict__ u, const T *__restrict__ v,
                                                     const T *__restrict__ w, const T *__restrict__ phi, int64_t id,
                                                     int64_t xs, int64_t ys, Inv<T> ih, int i, int j, int k, int nx,
                                                     int ny, int nz, int pmask) {
  return passive_advect_up2_at<true, T>(GAt<T>{u + id}, GAt<T>{v + id}, GAt<T>{w + id}, GAt<T>{phi + id}, xs, ys, ih,
                                        i, j, k, nx, ny, nz, pmask);
}

// centred Laplacian of phi at a cell, the sum of
// Sol_LaplacianCentered3DDevice_stencil (sol_laplaciancent3ddev.cu:27-62);
// the caller adds coef * lap
template <typename T, typename A>
__device__ __forceinline__ T laplacian_at(const A &phi, int64_t xs, int64_t ys, Inv<T> ih2) {
  const T p = phi(0);
  return ih2.z * (phi(1) + phi(-1) - ((T)2) * p) + ih2.y * (phi(ys) + phi(-ys) - ((T)2) * p) +
         ih2.x * (phi(xs) + phi(-xs) - ((T)2) * p);
}
template <typename T>
__device__ __forceinline__ T laplacian_cell(const T *__restrict__ phi, int64_t id, int64_t xs, int64_t ys,
                                            Inv<T> ih2) {
  return laplacian_at<T>(GAt<T>{phi + id}, xs, ys, ih2);
}

// Eqn_IncompressibleNS3D_add_thermal_force (eqn_incompressns3d.cu:25-44):
// the increment .5 * coef * (T + T_{-vertical}) of the vertical component
template <typename T, typename A>
__device__ __forceinline__ T thermal_at(const A &t, int64_t nbr, T coef) {
  return ((T).5) * coef * (t(0) + t(-nbr));
}
template <typename T>
__device__ __forceinline__ T thermal_cell(const T *__restrict__ t, int64_t id, int64_t nbr, T coef) {
  return thermal_at<T>(GAt<T>{t + id}, nbr, coef);
}

}  // namespace ocu_hip
