/*
 * ORACLE - test infrastructure only (see oracle_grid.h).
 *
 * CPU restatement of the MAC-grid operators around the pressure solve:
 * divergence, gradient subtraction, self- and passive advection, the
 * centred Laplacian, the Boussinesq thermal force, the projection solver and
 * the Eqn_IncompressibleNS3D / Eqn_ScalarAdvection3D time steps.
 */
#ifndef OCU_ORACLE_NS_H
#define OCU_ORACLE_NS_H

#include "oracle_mg.h"

namespace orc {

/* interpolation.h:35-48 */
template <typename T>
inline T interp(int scheme, T u, T x0, T x1) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return ((T).5) * (x1 + x0);
  return u < 0 ? x1 : x0;
}

/* Second-order (linear) upwind, OCU_IT_SECOND_ORDER_UPWIND (no reference
 * counterpart): the face value between xL and xR extrapolated from the two
 * upstream cells, xL + .5*(xL - xLL) for vel >= 0, xR + .5*(xR - xRR) for
 * vel < 0 (the reference's upwind picks xL at vel == 0); a missing second
 * upstream cell (h* false) gives the first-order value. */
template <typename T>
inline T up2(T vel, T xLL, bool hLL, T xL, T xR, T xRR, bool hRR) {
  if (vel < 0) return hRR ? xR + ((T).5) * (xR - xRR) : xR;
  return hLL ? xL + ((T).5) * (xL - xLL) : xL;
}

/* The cell c+d (d = +-2) along an axis of a grid whose ghost layer spans
 * [-1, hi]: present in the layer, or wrapped by the period P when the axis
 * is periodic, or missing.  Returns the offset to use (in cells). */
inline bool far_cell(int c, int d, int hi, int P, bool per, int &off) {
  const int cc = c + d;
  if (cc >= -1 && cc <= hi) {
    off = d;
    return true;
  }
  if (!per) return false;
  off = ((cc % P) + P) % P - c;
  return true;
}

/* Bit a set when both faces of axis a are periodic: the axes along which
 * the second-order upwind stencil wraps. */
inline int periodic_axes(const ocu_bc_set &bc) {
  auto pp = [](const ocu_bc &a, const ocu_bc &b) { return a.type == OCU_BC_PERIODIC && b.type == OCU_BC_PERIODIC; };
  return (pp(bc.xneg, bc.xpos) ? 1 : 0) | (pp(bc.yneg, bc.ypos) ? 2 : 0) | (pp(bc.zneg, bc.zpos) ? 4 : 0);
}

/* Sol_Divergence3DDevice_calculate_divergence, sol_divergence3ddev.cu:23-45;
 * 1/h rounded to T by the launcher (:80-84). */
template <typename T>
void divergence(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                const ocu_grid3d &dg, int nx, int ny, int nz, double hx, double hy, double hz) {
  View<T> u(ug), v(vg), w(wg), d(dg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        d.at(i, j, k) = ix * (u.at(i + 1, j, k) - u.at(i, j, k)) +
                        iy * (v.at(i, j + 1, k) - v.at(i, j, k)) +
                        iz * (w.at(i, j, k + 1) - w.at(i, j, k));
}

/* Sol_Gradient3DDevice_subtract_grad, sol_gradient3ddev.cu:24-47: over the
 * cells of phi, u += c*(1/hx)*(p - p_{i-1}) etc. */
template <typename T>
void gradient_sub(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                  const ocu_grid3d &pg, T coefficient, double hx, double hy, double hz) {
  View<T> u(ug), v(vg), w(wg), p(pg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < pg.nx; i++)
    for (int j = 0; j < pg.ny; j++)
      for (int k = 0; k < pg.nz; k++) {
        T pc = p.at(i, j, k);
        u.at(i, j, k) += coefficient * ix * (pc - p.at(i - 1, j, k));
        v.at(i, j, k) += coefficient * iy * (pc - p.at(i, j - 1, k));
        w.at(i, j, k) += coefficient * iz * (pc - p.at(i, j, k - 1));
      }
}

/* Self-advection of a MAC velocity in flux form,
 * Advection3DF/D_apply_upwind_TEX (sol_selfadvection3ddev.cu:136-337; the
 * naive kernel :29-114 has the same arithmetic).  Writes du,dv,dw on
 * i<nx, j<ny, k<nz only. */
template <typename T>
void self_advect_up2(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg, const ocu_grid3d &dug,
                     const ocu_grid3d &dvg, const ocu_grid3d &dwg, int nx, int ny, int nz, double hx, double hy,
                     double hz, int pmask);
template <typename T>
void self_advect(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                 const ocu_grid3d &dug, const ocu_grid3d &dvg, const ocu_grid3d &dwg,
                 int nx, int ny, int nz, double hx, double hy, double hz, int scheme, int pmask = 0) {
  if (scheme == OCU_IT_SECOND_ORDER_UPWIND) {
    self_advect_up2<T>(ug, vg, wg, dug, dvg, dwg, nx, ny, nz, hx, hy, hz, pmask);
    return;
  }
  View<T> U(ug), V(vg), W(wg), DU(dug), DV(dvg), DW(dwg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
  const T half = (T).5;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        auto u = [&](int a, int b, int c) { return U.at(i + a, j + b, k + c); };
        auto v = [&](int a, int b, int c) { return V.at(i + a, j + b, k + c); };
        auto w = [&](int a, int b, int c) { return W.at(i + a, j + b, k + c); };
        T u0 = u(0, 0, 0), v0 = v(0, 0, 0), w0 = w(0, 0, 0);
        // dudt
        T u_iph = half * (u0 + u(1, 0, 0));
        T u_imh = half * (u0 + u(-1, 0, 0));
        T duu = (u_iph * interp<T>(scheme, u_iph, u0, u(1, 0, 0))) -
                (u_imh * interp<T>(scheme, u_imh, u(-1, 0, 0), u0));
        T v_jph = half * (v(0, 1, 0) + v(-1, 1, 0));
        T v_jmh = half * (v0 + v(-1, 0, 0));
        T dvu = (v_jph * interp<T>(scheme, v_jph, u0, u(0, 1, 0))) -
                (v_jmh * interp<T>(scheme, v_jmh, u(0, -1, 0), u0));
        T w_kph = half * (w(0, 0, 1) + w(-1, 0, 1));
        T w_kmh = half * (w0 + w(-1, 0, 0));
        T dwu = (w_kph * interp<T>(scheme, w_kph, u0, u(0, 0, 1))) -
                (w_kmh * interp<T>(scheme, w_kmh, u(0, 0, -1), u0));
        DU.at(i, j, k) = -duu * ix - dvu * iy - dwu * iz;
        // dvdt
        T u_atv_iph = half * (u(1, 0, 0) + u(1, -1, 0));
        T u_atv_imh = half * (u0 + u(0, -1, 0));
        T duv = (u_atv_iph * interp<T>(scheme, u_atv_iph, v0, v(1, 0, 0))) -
                (u_atv_imh * interp<T>(scheme, u_atv_imh, v(-1, 0, 0), v0));
        T v_iph = half * (v0 + v(0, 1, 0));
        T v_imh = half * (v0 + v(0, -1, 0));
        T dvv = (v_iph * interp<T>(scheme, v_iph, v0, v(0, 1, 0))) -
                (v_imh * interp<T>(scheme, v_imh, v(0, -1, 0), v0));
        T w_atv_kph = half * (w(0, 0, 1) + w(0, -1, 1));
        T w_atv_kmh = half * (w0 + w(0, -1, 0));
        T dwv = (w_atv_kph * interp<T>(scheme, w_atv_kph, v0, v(0, 0, 1))) -
                (w_atv_kmh * interp<T>(scheme, w_atv_kmh, v(0, 0, -1), v0));
        DV.at(i, j, k) = -duv * ix - dvv * iy - dwv * iz;
        // dwdt
        T u_atw_iph = half * (u(1, 0, 0) + u(1, 0, -1));
        T u_atw_imh = half * (u0 + u(0, 0, -1));
        T duw = (u_atw_iph * interp<T>(scheme, u_atw_iph, w0, w(1, 0, 0))) -
                (u_atw_imh * interp<T>(scheme, u_atw_imh, w(-1, 0, 0), w0));
        T v_atw_jph = half * (v(0, 1, 0) + v(0, 1, -1));
        T v_atw_jmh = half * (v0 + v(0, 0, -1));
        T dvw = (v_atw_jph * interp<T>(scheme, v_atw_jph, w0, w(0, 1, 0))) -
                (v_atw_jmh * interp<T>(scheme, v_atw_jmh, w(0, -1, 0), w0));
        T w_iph = half * (w0 + w(0, 0, 1));
        T w_imh = half * (w0 + w(0, 0, -1));
        T dww = (w_iph * interp<T>(scheme, w_iph, w0, w(0, 0, 1))) -
                (w_imh * interp<T>(scheme, w_imh, w(0, 0, -1), w0));
        DW.at(i, j, k) = -duw * ix - dvw * iy - dww * iz;
      }
}

/* Self-advection with OCU_IT_SECOND_ORDER_UPWIND (no reference counterpart):
 * the flux-form terms and face velocities of self_advect above, each face
 * value taken by up2 from the two upstream samples of the advected component
 * along the flux axis.  The component's ghost layer along axis a spans
 * [-1, n_a] or, for the component normal to a, [-1, n_a + 1]; a second
 * upstream sample beyond it wraps when bit a of pmask is set and is
 * otherwise missing (first order at that face). */
template <typename T>
void self_advect_up2(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg, const ocu_grid3d &dug,
                     const ocu_grid3d &dvg, const ocu_grid3d &dwg, int nx, int ny, int nz, double hx, double hy,
                     double hz, int pmask) {
  View<T> U(ug), V(vg), W(wg), DU(dug), DV(dvg), DW(dwg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
  const T half = (T).5;
  const int n[3] = {nx, ny, nz};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const int c[3] = {i, j, k};
        auto u = [&](int a, int b, int cc) { return U.at(i + a, j + b, k + cc); };
        auto v = [&](int a, int b, int cc) { return V.at(i + a, j + b, k + cc); };
        auto w = [&](int a, int b, int cc) { return W.at(i + a, j + b, k + cc); };
        // (vh * face_hi) - (vl * face_lo) for component X (normal axis xa)
        // along axis a
        auto dflux = [&](const View<T> &X, int xa, int a, T vh, T vl) {
          auto at = [&](int dd) { return X.at(i + (a == 0) * dd, j + (a == 1) * dd, k + (a == 2) * dd); };
          const bool per = (pmask >> a) & 1;
          const int hi = n[a] + (xa == a);
          int ohi = 0, olo = 0;
          const bool hhi = far_cell(c[a], 2, hi, n[a], per, ohi);
          const bool hlo = far_cell(c[a], -2, hi, n[a], per, olo);
          const T x0 = at(0), xm = at(-1), xp = at(1);
          const T xpp = hhi ? at(ohi) : (T)0, xmm = hlo ? at(olo) : (T)0;
          return (vh * up2<T>(vh, xm, true, x0, xp, xpp, hhi)) - (vl * up2<T>(vl, xmm, hlo, xm, x0, xp, true));
        };
        const T u0 = u(0, 0, 0), v0 = v(0, 0, 0), w0 = w(0, 0, 0);
        const T duu = dflux(U, 0, 0, half * (u0 + u(1, 0, 0)), half * (u0 + u(-1, 0, 0)));
        const T dvu = dflux(U, 0, 1, half * (v(0, 1, 0) + v(-1, 1, 0)), half * (v0 + v(-1, 0, 0)));
        const T dwu = dflux(U, 0, 2, half * (w(0, 0, 1) + w(-1, 0, 1)), half * (w0 + w(-1, 0, 0)));
        DU.at(i, j, k) = -duu * ix - dvu * iy - dwu * iz;
        const T duv = dflux(V, 1, 0, half * (u(1, 0, 0) + u(1, -1, 0)), half * (u0 + u(0, -1, 0)));
        const T dvv = dflux(V, 1, 1, half * (v0 + v(0, 1, 0)), half * (v0 + v(0, -1, 0)));
        const T dwv = dflux(V, 1, 2, half * (w(0, 0, 1) + w(0, -1, 1)), half * (w0 + w(0, -1, 0)));
        DV.at(i, j, k) = -duv * ix - dvv * iy - dwv * iz;
        const T duw = dflux(W, 2, 0, half * (u(1, 0, 0) + u(1, 0, -1)), half * (u0 + u(0, 0, -1)));
        const T dvw = dflux(W, 2, 1, half * (v(0, 1, 0) + v(0, 1, -1)), half * (v0 + v(0, 0, -1)));
        const T dww = dflux(W, 2, 2, half * (w0 + w(0, 0, 1)), half * (w0 + w(0, 0, -1)));
        DW.at(i, j, k) = -duw * ix - dvw * iy - dww * iz;
      }
}

/* Sol_PassiveAdvection3D_apply_interp, sol_passiveadvection3ddev.cu:31-78 */
template <typename T>
void passive_advect(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                    const ocu_grid3d &phig, const ocu_grid3d &dg, int nx, int ny, int nz,
                    double hx, double hy, double hz, int scheme, int pmask = 0) {
  View<T> u(ug), v(vg), w(wg), phi(phig), d(dg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
  if (scheme == OCU_IT_SECOND_ORDER_UPWIND) {
    // pmask bit a: axis a periodic.  Each face flux takes its two upstream
    // cells; only the second one can leave the ghost layer.
    const int n[3] = {nx, ny, nz};
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          const int c[3] = {i, j, k};
          auto at = [&](int a, int dd) {
            return phi.at(i + (a == 0) * dd, j + (a == 1) * dd, k + (a == 2) * dd);
          };
          T fl[3];
          for (int a = 0; a < 3; a++) {
            const T vh = a == 0 ? u.at(i + 1, j, k) : a == 1 ? v.at(i, j + 1, k) : w.at(i, j, k + 1);
            const T vl = a == 0 ? u.at(i, j, k) : a == 1 ? v.at(i, j, k) : w.at(i, j, k);
            const bool per = (pmask >> a) & 1;
            int ohi = 0, olo = 0;
            const bool hhi = far_cell(c[a], 2, n[a], n[a], per, ohi);   // high face: xRR = c+2
            const bool hlo = far_cell(c[a], -2, n[a], n[a], per, olo);  // low face: xLL = c-2
            const T p = at(a, 0);
            const T fh = vh * up2<T>(vh, at(a, -1), true, p, at(a, 1), hhi ? at(a, ohi) : (T)0, hhi);
            const T flo = vl * up2<T>(vl, hlo ? at(a, olo) : (T)0, hlo, at(a, -1), p, at(a, 1), true);
            fl[a] = fh - flo;
          }
          d.at(i, j, k) = (-fl[0] * ix) - (fl[1] * iy) - (fl[2] * iz);
        }
    return;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        T p = phi.at(i, j, k);
        T u_pi = u.at(i + 1, j, k), u0 = u.at(i, j, k);
        T du = u_pi * interp<T>(scheme, u_pi, p, phi.at(i + 1, j, k)) -
               u0 * interp<T>(scheme, u0, phi.at(i - 1, j, k), p);
        T v_pj = v.at(i, j + 1, k), v0 = v.at(i, j, k);
        T dv = v_pj * interp<T>(scheme, v_pj, p, phi.at(i, j + 1, k)) -
               v0 * interp<T>(scheme, v0, phi.at(i, j - 1, k), p);
        T w_pk = w.at(i, j, k + 1), w0 = w.at(i, j, k);
        T dw = w_pk * interp<T>(scheme, w_pk, p, phi.at(i, j, k + 1)) -
               w0 * interp<T>(scheme, w0, phi.at(i, j, k - 1), p);
        d.at(i, j, k) = (-du * ix) - (dv * iy) - (dw * iz);
      }
}

/* ---- adjoints (no reference counterpart; derived from passive_advect) ----
 *
 * passive_advect is dphi_c = -sum_axes (F_{c+1} - F_c) / h with the face
 * flux F_f = u_f * I(u_f, phi_{f-1}, phi_f); for fixed upwind selection it is
 * linear in phi and in u.  With lam the adjoint of dphi (interior; the ghost
 * layer of lam must be zero) and g_f = scale * ((lam_f - lam_{f-1}) / h):
 *   adj_u_f   += g_f * I(u_f, phi_{f-1}, phi_f)
 *   adj_phi_c += sum over the faces of c of (g_f * u_f) * dI/dphi_c
 * (dI/dx0 = u<0 ? 0 : 1, dI/dx1 = u<0 ? 1 : 0 upwind; .5 centred).
 * lam_out = lam_in + adj_phi over the whole ghost box (the identity part of
 * phi_{n+1} = phi_n + dt*dphi with scale = dt); adj_u/v/w accumulate at the
 * faces [0,n] whose other two indices are interior. */
template <typename T>
inline T interp_w0(int scheme, T u) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return (T).5;
  return u < 0 ? (T)0 : (T)1;
}
template <typename T>
inline T interp_w1(int scheme, T u) {
  if (scheme == OCU_IT_SECOND_ORDER_CENTERED) return (T).5;
  return u < 0 ? (T)1 : (T)0;
}

template <typename T>
void passive_advect_adjoint(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                            const ocu_grid3d &phig, const ocu_grid3d &lamg, const ocu_grid3d &outg,
                            const ocu_grid3d &aug, const ocu_grid3d &avg, const ocu_grid3d &awg, double scaled,
                            int nx, int ny, int nz, double hx, double hy, double hz, int scheme) {
  View<T> u(ug), v(vg), w(wg), phi(phig), lam(lamg), out(outg), au(aug), av(avg), aw(awg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz), sc = (T)scaled;
#pragma omp parallel for schedule(static)
  for (int i = -1; i <= nx; i++)
    for (int j = -1; j <= ny; j++)
      for (int k = -1; k <= nz; k++) {
        const bool ii = i >= 0 && i < nx, ji = j >= 0 && j < ny, ki = k >= 0 && k < nz;
        T acc = 0;
        if (ji && ki) {
          if (i >= 0) {  // face i (cell c is its upper side)
            T g = sc * (ix * (lam.at(i, j, k) - lam.at(i - 1, j, k)));
            T f = u.at(i, j, k);
            acc += (g * f) * interp_w1<T>(scheme, f);
            au.at(i, j, k) += g * interp<T>(scheme, f, phi.at(i - 1, j, k), phi.at(i, j, k));
          }
          if (i < nx) {  // face i+1 (cell c is its lower side)
            T g = sc * (ix * (lam.at(i + 1, j, k) - lam.at(i, j, k)));
            T f = u.at(i + 1, j, k);
            acc += (g * f) * interp_w0<T>(scheme, f);
          }
        }
        if (ii && ki) {
          if (j >= 0) {
            T g = sc * (iy * (lam.at(i, j, k) - lam.at(i, j - 1, k)));
            T f = v.at(i, j, k);
            acc += (g * f) * interp_w1<T>(scheme, f);
            av.at(i, j, k) += g * interp<T>(scheme, f, phi.at(i, j - 1, k), phi.at(i, j, k));
          }
          if (j < ny) {
            T g = sc * (iy * (lam.at(i, j + 1, k) - lam.at(i, j, k)));
            T f = v.at(i, j + 1, k);
            acc += (g * f) * interp_w0<T>(scheme, f);
          }
        }
        if (ii && ji) {
          if (k >= 0) {
            T g = sc * (iz * (lam.at(i, j, k) - lam.at(i, j, k - 1)));
            T f = w.at(i, j, k);
            acc += (g * f) * interp_w1<T>(scheme, f);
            aw.at(i, j, k) += g * interp<T>(scheme, f, phi.at(i, j, k - 1), phi.at(i, j, k));
          }
          if (k < nz) {
            T g = sc * (iz * (lam.at(i, j, k + 1) - lam.at(i, j, k)));
            T f = w.at(i, j, k + 1);
            acc += (g * f) * interp_w0<T>(scheme, f);
          }
        }
        out.at(i, j, k) = lam.at(i, j, k) + acc;
      }
}

/* Adjoint of bc_fill_nocorners_device: every face ghost is c * (interior
 * source) + const with c = 1 (periodic copy, Neumann) or -1 (Dirichlet), so
 * each interior cell gathers c * adj(ghost) from the face ghosts derived from
 * it (x, then y, then z; low face then high face), and the face ghosts'
 * adjoints are zeroed.  BC_NONE ghosts are constants: their adjoint is
 * dropped. */
template <typename T>
void bc_adjoint_nocorners(const ocu_grid3d &g, const ocu_bc_set &bc) {
  View<T> a(g);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  auto coef = [](const ocu_bc &b) -> T {
    return b.type == OCU_BC_DIRICHELET ? (T)-1 : (b.type == OCU_BC_NONE ? (T)0 : (T)1);
  };
  auto per = [](const ocu_bc &b) { return b.type == OCU_BC_PERIODIC; };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        T acc = a.at(i, j, k);
        // ghost (-1) derives from 0 (or nx-1 if periodic), ghost (n) from n-1 (or 0)
        if (i == (per(bc.xneg) ? nx - 1 : 0)) acc += coef(bc.xneg) * a.at(-1, j, k);
        if (i == (per(bc.xpos) ? 0 : nx - 1)) acc += coef(bc.xpos) * a.at(nx, j, k);
        if (j == (per(bc.yneg) ? ny - 1 : 0)) acc += coef(bc.yneg) * a.at(i, -1, k);
        if (j == (per(bc.ypos) ? 0 : ny - 1)) acc += coef(bc.ypos) * a.at(i, ny, k);
        if (k == (per(bc.zneg) ? nz - 1 : 0)) acc += coef(bc.zneg) * a.at(i, j, -1);
        if (k == (per(bc.zpos) ? 0 : nz - 1)) acc += coef(bc.zpos) * a.at(i, j, nz);
        a.at(i, j, k) = acc;
      }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) a.at(-1, j, k) = a.at(nx, j, k) = 0;
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) a.at(i, -1, k) = a.at(i, ny, k) = 0;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) a.at(i, j, -1) = a.at(i, j, nz) = 0;
}

/* Adjoint of the normal-face part of apply_3d_mac_boundary_conditions_level1
 * for the faces an advection reads (the normal components at [0,n]):
 * a periodic high face duplicates face 0 at face n, so adj(0) += adj(n); a
 * forced-inflow low face sets face 0 to the boundary value, so adj(0) is
 * dropped.  Afterwards adj(n) = 0 in both cases (face n is not a parameter). */
template <typename T>
void mac_adjoint_normal(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg, const ocu_bc_set &bc) {
  View<T> u(ug), v(vg), w(wg);
  const int nx = ug.nx - 1, ny = ug.ny, nz = ug.nz;
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xpos.type == OCU_BC_PERIODIC) u.at(0, j, k) += u.at(nx, j, k);
      if (bc.xneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) u.at(0, j, k) = 0;
      u.at(nx, j, k) = 0;
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.ypos.type == OCU_BC_PERIODIC) v.at(i, 0, k) += v.at(i, ny, k);
      if (bc.yneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) v.at(i, 0, k) = 0;
      v.at(i, ny, k) = 0;
    }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (bc.zpos.type == OCU_BC_PERIODIC) w.at(i, j, 0) += w.at(i, j, nz);
      if (bc.zneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) w.at(i, j, 0) = 0;
      w.at(i, j, nz) = 0;
    }
}

/* Adjoint of divergence (no reference counterpart).  d_c = ix*(u_{c+1} -
 * u_c) + iy*(...) + iz*(...) over the interior cells, so every normal face f
 * in [0,n] of an interior row gathers
 *   adj_u_f += ix * (lam_{f-1} - lam_f)
 * with lam taken as 0 outside the interior (lam's ghosts are not read). */
template <typename T>
void divergence_adjoint(const ocu_grid3d &lamg, const ocu_grid3d &aug, const ocu_grid3d &avg, const ocu_grid3d &awg,
                        int nx, int ny, int nz, double hx, double hy, double hz) {
  View<T> lam(lamg), au(aug), av(avg), aw(awg);
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
  auto L = [&](int i, int j, int k) -> T {
    return (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) ? lam.at(i, j, k) : (T)0;
  };
#pragma omp parallel for schedule(static)
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++) {
        const T c = L(i, j, k);
        if (j < ny && k < nz) au.at(i, j, k) += ix * (L(i - 1, j, k) - c);
        if (i < nx && k < nz) av.at(i, j, k) += iy * (L(i, j - 1, k) - c);
        if (i < nx && j < ny) aw.at(i, j, k) += iz * (L(i, j, k - 1) - c);
      }
}

/* Adjoint of gradient_sub in p (no reference counterpart).  The forward
 * step is u_c += (coef*ix) * (p_c - p_{c-1}) over the cells c of p, so with
 * cx = coef*ix and U the adjoint of u on the written faces [0,n-1] of
 * interior rows (0 elsewhere), every cell of p's interior and of its three
 * low face-ghost layers gathers
 *   adj_p_c += (cx*(U_c - U_{c+x}) + cy*(V_c - V_{c+y})) + cz*(W_c - W_{c+z}).
 * The adjoint in u is the identity (u is updated in place). */
template <typename T>
void gradient_adjoint(const ocu_grid3d &aug, const ocu_grid3d &avg, const ocu_grid3d &awg, const ocu_grid3d &apg,
                      T coefficient, double hx, double hy, double hz) {
  View<T> au(aug), av(avg), aw(awg), ap(apg);
  const int nx = apg.nx, ny = apg.ny, nz = apg.nz;
  const T ix = (T)(1 / hx), iy = (T)(1 / hy), iz = (T)(1 / hz);
  const T cx = coefficient * ix, cy = coefficient * iy, cz = coefficient * iz;
  auto in = [&](int i, int j, int k) { return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz; };
  auto U = [&](int i, int j, int k) -> T { return in(i, j, k) ? au.at(i, j, k) : (T)0; };
  auto V = [&](int i, int j, int k) -> T { return in(i, j, k) ? av.at(i, j, k) : (T)0; };
  auto W = [&](int i, int j, int k) -> T { return in(i, j, k) ? aw.at(i, j, k) : (T)0; };
#pragma omp parallel for schedule(static)
  for (int i = -1; i < nx; i++)
    for (int j = -1; j < ny; j++)
      for (int k = -1; k < nz; k++) {
        if ((i < 0) + (j < 0) + (k < 0) > 1) continue;  // edges and corners are never read
        ap.at(i, j, k) += (cx * (U(i, j, k) - U(i + 1, j, k)) + cy * (V(i, j, k) - V(i, j + 1, k))) +
                          cz * (W(i, j, k) - W(i, j, k + 1));
      }
}

/* Sol_LaplacianCentered3DDevice_stencil, sol_laplaciancent3ddev.cu:27-62:
 * dphi += c * (z, y, x second differences), 1/h^2 rounded to T (:100-102). */
template <typename T>
void laplacian_acc(const ocu_grid3d &pg, const ocu_grid3d &dg, int nx, int ny, int nz,
                   double hx, double hy, double hz, T coefficient) {
  View<T> phi(pg), d(dg);
  const T ix2 = (T)(1 / (hx * hx)), iy2 = (T)(1 / (hy * hy)), iz2 = (T)(1 / (hz * hz));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        T p = phi.at(i, j, k);
        T dp = d.at(i, j, k);
        T lap = iz2 * (phi.at(i, j, k + 1) + phi.at(i, j, k - 1) - ((T)2) * p) +
                iy2 * (phi.at(i, j + 1, k) + phi.at(i, j - 1, k) - ((T)2) * p) +
                ix2 * (phi.at(i + 1, j, k) + phi.at(i - 1, j, k) - ((T)2) * p);
        d.at(i, j, k) = dp + coefficient * lap;
      }
}

/* Eqn_IncompressibleNS3D_add_thermal_force, eqn_incompressns3d.cu:25-44,
 * launched at :119-155: d += (T).5 * c * (t + t_{-vertical}). */
template <typename T>
void thermal_force(const ocu_grid3d &dg, const ocu_grid3d &tg, int nx, int ny, int nz,
                   int64_t nbr_stride, T coefficient) {
  View<T> d(dg), t(tg);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        T *tp = &t.at(i, j, k);
        d.at(i, j, k) += ((T).5) * coefficient * (tp[0] + tp[-nbr_stride]);
      }
}

/* convert_bc_to_poisson_eqn, sol_project3d.cpp:26-43 */
inline ocu_bc poisson_bc(const ocu_bc &b) {
  ocu_bc r{OCU_BC_INVALID, 0, 0};
  if (b.type == OCU_BC_PERIODIC) r.type = OCU_BC_PERIODIC;
  else if (b.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) { r.type = OCU_BC_NEUMANN; r.value = 0; }
  return r;
}

/* Sol_ProjectDivergence3DDevice<T> (sol_project3ddev.cu:28-135): MAC BC ->
 * divergence -> MG solve (LINF, 15 iterations) -> gradient (c = -1) -> MAC
 * BC.  pressure and divergence are congruent with u. */
template <typename T>
struct Projection {
  ocu_bc_set bc;
  int nx = 0, ny = 0, nz = 0;
  double hx = 1, hy = 1, hz = 1;
  ocu_grid3d u{}, v{}, w{};
  HostGrid<T> divg, pressure;
  MGSolver<T> mg;
  double last_residual = 0;

  bool initialize_storage(int nx_, int ny_, int nz_, double hx_, double hy_, double hz_,
                          const ocu_grid3d &u_, const ocu_grid3d &v_, const ocu_grid3d &w_) {
    nx = nx_; ny = ny_; nz = nz_; hx = hx_; hy = hy_; hz = hz_;
    u = u_; v = v_; w = w_;
    divg.init(nx, ny, nz, 1, 1, 1, u.pnx - (nx + 2), u.pny - (ny + 2), u.pnz - (nz + 2));
    if (!layout_match(divg.desc, u)) return false;
    pressure.init_congruent(divg.desc);
    mg.bc.xneg = poisson_bc(bc.xneg); mg.bc.xpos = poisson_bc(bc.xpos);
    mg.bc.yneg = poisson_bc(bc.yneg); mg.bc.ypos = poisson_bc(bc.ypos);
    mg.bc.zneg = poisson_bc(bc.zneg); mg.bc.zpos = poisson_bc(bc.zpos);
    if (!mg.initialize_storage(nx, ny, nz, hx, hy, hz, divg.desc, pressure.desc)) return false;
    mg.convergence = OCU_CONVERGENCE_LINF;
    return true;
  }

  bool solve(double tolerance) {
    bool ok = bc_fill_mac_level1<T>(u, v, w, bc);
    divergence<T>(u, v, w, divg.desc, nx, ny, nz, hx, hy, hz);
    double res = 0;
    ok = mg.solve(res, tolerance, 15) && ok;
    last_residual = res;
    gradient_sub<T>(u, v, w, pressure.desc, (T)-1, hx, hy, hz);
    ok = bc_fill_mac_level1<T>(u, v, w, bc) && ok;
    return ok;
  }
};

struct NSParams {
  int nx, ny, nz;
  double hx, hy, hz;
  double max_divergence;
  ocu_bc_set temp_bc, flow_bc;
  int vertical_direction;
  double gravity, bouyancy;
  double thermal_diffusion, viscosity;
  int advection_scheme, time_step;  // time_step: 1 FE, 2 AB2
  double cfl_factor;
};

/* Eqn_IncompressibleNS3D<T> (eqn_incompressns3d.cu:119-399) and
 * Equation::advance (equation.cpp:31-56). */
template <typename T>
struct NS {
  NSParams p;
  HostGrid<T> u, v, w, du, dv, dw, ldu, ldv, ldw, temp, dtemp, ldtemp;
  Projection<T> proj;
  double lastdt = 0;
  int num_steps = 0;
  T gravity, bouyancy, viscosity, thermal_diffusion;

  /* set_parameters (:164-299); init_* are host grids laid out like
   * Eqn_IncompressibleNS3DParams::init_grids (eqn_incompressns3d.h:78-86). */
  bool set_parameters(const NSParams &prm, const ocu_grid3d &iu, const ocu_grid3d &iv,
                      const ocu_grid3d &iw, const ocu_grid3d &it) {
    p = prm;
    gravity = (T)p.gravity; bouyancy = (T)p.bouyancy;
    viscosity = (T)p.viscosity; thermal_diffusion = (T)p.thermal_diffusion;
    auto cp = [](HostGrid<T> &dst, const ocu_grid3d &src) {
      dst.init_congruent(src);
      std::memcpy(dst.buf.data(), src.base, sizeof(T) * (size_t)src.allocated);
    };
    cp(u, iu); du.init_congruent(iu); ldu.init_congruent(iu);
    cp(v, iv); dv.init_congruent(iv); ldv.init_congruent(iv);
    cp(w, iw); dw.init_congruent(iw); ldw.init_congruent(iw);
    cp(temp, it); dtemp.init_congruent(it); ldtemp.init_congruent(it);
    proj.bc = p.flow_bc;
    if (!proj.initialize_storage(p.nx, p.ny, p.nz, p.hx, p.hy, p.hz, u.desc, v.desc, w.desc)) return false;
    if (!bc_fill_mac_level1<T>(u.desc, v.desc, w.desc, p.flow_bc)) return false;
    if (!bc_fill_level1<T>(temp.desc, p.temp_bc, p.hx, p.hy, p.hz)) return false;
    return true;
  }

  void add_thermal_force() {
    int vd = p.vertical_direction;
    T direction_mult = (vd & OCU_DIR_NEGATIVE_FLAG) ? 1 : -1;
    const ocu_grid3d &tgt = (vd & OCU_DIR_XAXIS_FLAG) ? du.desc : (vd & OCU_DIR_YAXIS_FLAG) ? dv.desc : dw.desc;
    int64_t stride = (vd & OCU_DIR_XAXIS_FLAG) ? (int64_t)temp.desc.pny * temp.desc.pnz
                   : (vd & OCU_DIR_YAXIS_FLAG) ? temp.desc.pnz : 1;
    thermal_force<T>(tgt, temp.desc, p.nx, p.ny, p.nz, stride, direction_mult * gravity * bouyancy);
  }

  /* get_max_stable_timestep (:301-328) */
  double get_max_stable_timestep() const {
    T mu = (T)reduce<T>(u.desc, 0), mv = (T)reduce<T>(v.desc, 0), mw = (T)reduce<T>(w.desc, 0);
    double ut = p.hx / mu, vt = p.hy / mv, wt = p.hz / mw;
    if (!std::isfinite(ut)) ut = 1e10;
    if (!std::isfinite(vt)) vt = 1e10;
    if (!std::isfinite(wt)) wt = 1e10;
    double step = p.cfl_factor * min3(ut, vt, wt);
    double minh = min3(p.hx, p.hy, p.hz);
    if (thermal_diffusion > 0) step = std::min(step, (minh * minh) / (6 * thermal_diffusion));
    if (viscosity > 0) step = std::min(step, (minh * minh) / (6 * viscosity));
    return step;
  }

  /* advance_one_step (:331-399) */
  bool advance_one_step(double dt) {
    bool ok = true;
    num_steps++;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    self_advect<T>(u.desc, v.desc, w.desc, du.desc, dv.desc, dw.desc, nx, ny, nz, p.hx, p.hy, p.hz, p.advection_scheme,
                   periodic_axes(p.flow_bc));
    if (viscosity > 0) {
      laplacian_acc<T>(u.desc, du.desc, nx + 1, ny, nz, p.hx, p.hy, p.hz, viscosity);
      laplacian_acc<T>(v.desc, dv.desc, nx, ny + 1, nz, p.hx, p.hy, p.hz, viscosity);
      laplacian_acc<T>(w.desc, dw.desc, nx, ny, nz + 1, p.hx, p.hy, p.hz, viscosity);
    }
    add_thermal_force();
    passive_advect<T>(u.desc, v.desc, w.desc, temp.desc, dtemp.desc, nx, ny, nz, p.hx, p.hy, p.hz, p.advection_scheme,
                      periodic_axes(p.temp_bc));
    if (thermal_diffusion > 0)
      laplacian_acc<T>(temp.desc, dtemp.desc, nx, ny, nz, p.hx, p.hy, p.hz, thermal_diffusion);
    T ab_coeff = -dt * dt / (2 * lastdt);
    bool ab2 = p.time_step == 2 && lastdt > 0;
    auto adv = [&](HostGrid<T> &x, HostGrid<T> &d, HostGrid<T> &ld) {
      if (ab2) {
        ok = lincomb2<T>(x.desc, (T)1.0, x.desc, (T)(dt - ab_coeff), d.desc) && ok;
        ok = lincomb2<T>(x.desc, (T)1.0, x.desc, (T)ab_coeff, ld.desc) && ok;
      } else {
        ok = lincomb2<T>(x.desc, (T)1.0, x.desc, (T)dt, d.desc) && ok;
      }
    };
    adv(temp, dtemp, ldtemp);
    ok = bc_fill_nocorners_device<T>(temp.desc, p.temp_bc, p.hx, p.hy, p.hz) && ok;
    adv(u, du, ldu);
    adv(v, dv, ldv);
    adv(w, dw, ldw);
    if (p.time_step == 2) {
      lastdt = dt;
      ldtemp.buf = dtemp.buf; ldtemp.desc.base = ldtemp.buf.data();
      ldu.buf = du.buf; ldu.desc.base = ldu.buf.data();
      ldv.buf = dv.buf; ldv.desc.base = ldv.buf.data();
      ldw.buf = dw.buf; ldw.desc.base = ldw.buf.data();
    }
    ok = proj.solve(p.max_divergence) && ok;
    return ok;
  }

  bool advance(double dt) {
    if (!std::isfinite(dt)) return false;
    double dtleft = dt;
    while (dtleft > 0) {
      int subds = (int)std::ceil(dtleft / get_max_stable_timestep());
      double subdt = dtleft / subds;
      if (!advance_one_step(subdt)) return false;
      dtleft -= subdt;
    }
    return true;
  }
};

/* Eqn_ScalarAdvection3D<T>::advance_one_step (eqn_scalaradvection3d.cpp:
 * 88-105): phi = phi + dt * dphi (forward Euler) then the device
 * no-corner ghost fill. */
template <typename T>
bool scalar_advect_step(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg,
                        const ocu_grid3d &phig, const ocu_grid3d &dphig, const ocu_bc_set &bc,
                        int nx, int ny, int nz, double hx, double hy, double hz, int scheme, double dt) {
  passive_advect<T>(ug, vg, wg, phig, dphig, nx, ny, nz, hx, hy, hz, scheme, periodic_axes(bc));
  bool ok = lincomb2<T>(phig, (T)1, phig, (T)dt, dphig);
  ok = bc_fill_nocorners_device<T>(phig, bc, hx, hy, hz) && ok;
  return ok;
}

}  // namespace orc

#endif
