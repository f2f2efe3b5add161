/*
 * ORACLE - test infrastructure only.
 *
 * CPU restatement of the two-layer ghost fills
 * (ocustorage/grid3dboundary.cu:122-232, 350-549, 904-1016) and of point
 * sampling (ocustorage/grid3dsample.cu:24-263).  Expressions keep the
 * reference's operand order and precisions (float boundary values and
 * positions, T grid arithmetic, a float result of the trilinear lookup).
 */
#ifndef OCU_ORACLE_BC2_H
#define OCU_ORACLE_BC2_H

#include <cmath>

#include "oracle_grid.h"

namespace orc {

/* kernel_apply_3d_boundary_conditions_level2_nocorners
 * (grid3dboundary.cu:122-232): z faces over (i,j), y faces over (i,k), x
 * faces over (j,k) of the interior; every ghost reads interior cells. */
template <typename T>
bool bc_fill_level2_nocorners(const ocu_grid3d &g, const ocu_bc_set &bc, T hx, T hy, T hz) {
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (b->type != OCU_BC_NONE && b->type != OCU_BC_PERIODIC && b->type != OCU_BC_DIRICHELET &&
        b->type != OCU_BC_NEUMANN)
      return false;
  if (g.gx < 2 || g.gy < 2 || g.gz < 2) return false;
  View<T> p(g);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (bc.zneg.type == OCU_BC_PERIODIC) {
        p.at(i, j, -1) = p.at(i, j, nz - 1);
        p.at(i, j, -2) = p.at(i, j, nz - 2);
      } else if (bc.zneg.type == OCU_BC_DIRICHELET) {
        p.at(i, j, -1) = -p.at(i, j, 0) + 2 * bc.zneg.value;
        p.at(i, j, -2) = -p.at(i, j, 0) + 2 * bc.zneg.value;
      } else if (bc.zneg.type == OCU_BC_NEUMANN) {
        p.at(i, j, -1) = p.at(i, j, 0) - hz * bc.zneg.value;
        p.at(i, j, -2) = p.at(i, j, 0) - 2 * hz * bc.zneg.value;
      }
      if (bc.zpos.type == OCU_BC_PERIODIC) {
        p.at(i, j, nz) = p.at(i, j, 0);
        p.at(i, j, nz + 1) = p.at(i, j, 1);
      } else if (bc.zpos.type == OCU_BC_DIRICHELET) {
        p.at(i, j, nz) = -p.at(i, j, nz - 1) + 2 * bc.zpos.value;
        p.at(i, j, nz + 1) = -p.at(i, j, nz - 1) + 2 * bc.zpos.value;
      } else if (bc.zpos.type == OCU_BC_NEUMANN) {
        p.at(i, j, nz) = p.at(i, j, nz - 1) + hz * bc.zpos.value;
        p.at(i, j, nz + 1) = p.at(i, j, nz - 1) + 2 * hz * bc.zpos.value;
      }
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.yneg.type == OCU_BC_PERIODIC) {
        p.at(i, -1, k) = p.at(i, ny - 1, k);
        p.at(i, -2, k) = p.at(i, ny - 2, k);
      } else if (bc.yneg.type == OCU_BC_DIRICHELET) {
        p.at(i, -1, k) = -p.at(i, 0, k) + 2 * bc.yneg.value;
        p.at(i, -2, k) = -p.at(i, 0, k) + 2 * bc.yneg.value;
      } else if (bc.yneg.type == OCU_BC_NEUMANN) {
        p.at(i, -1, k) = p.at(i, 0, k) - hy * bc.yneg.value;
        p.at(i, -2, k) = p.at(i, 0, k) - 2 * hy * bc.yneg.value;
      }
      if (bc.ypos.type == OCU_BC_PERIODIC) {
        p.at(i, ny, k) = p.at(i, 0, k);
        p.at(i, ny + 1, k) = p.at(i, 1, k);
      } else if (bc.ypos.type == OCU_BC_DIRICHELET) {
        p.at(i, ny, k) = -p.at(i, ny - 1, k) + 2 * bc.ypos.value;
        p.at(i, ny + 1, k) = -p.at(i, ny - 1, k) + 2 * bc.ypos.value;
      } else if (bc.ypos.type == OCU_BC_NEUMANN) {
        p.at(i, ny, k) = p.at(i, ny - 1, k) + hy * bc.ypos.value;
        p.at(i, ny + 1, k) = p.at(i, ny - 1, k) + 2 * hy * bc.ypos.value;
      }
    }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xneg.type == OCU_BC_PERIODIC) {
        p.at(-1, j, k) = p.at(nx - 1, j, k);
        p.at(-2, j, k) = p.at(nx - 2, j, k);
      } else if (bc.xneg.type == OCU_BC_DIRICHELET) {
        p.at(-1, j, k) = -p.at(0, j, k) + 2 * bc.xneg.value;
        p.at(-2, j, k) = -p.at(0, j, k) + 2 * bc.xneg.value;
      } else if (bc.xneg.type == OCU_BC_NEUMANN) {
        p.at(-1, j, k) = p.at(0, j, k) - hx * bc.xneg.value;
        p.at(-2, j, k) = p.at(0, j, k) - 2 * hx * bc.xneg.value;
      }
      if (bc.xpos.type == OCU_BC_PERIODIC) {
        p.at(nx, j, k) = p.at(0, j, k);
        p.at(nx + 1, j, k) = p.at(1, j, k);
      } else if (bc.xpos.type == OCU_BC_DIRICHELET) {
        p.at(nx, j, k) = -p.at(nx - 1, j, k) + 2 * bc.xpos.value;
        p.at(nx + 1, j, k) = -p.at(nx - 1, j, k) + 2 * bc.xpos.value;
      } else if (bc.xpos.type == OCU_BC_NEUMANN) {
        p.at(nx, j, k) = p.at(nx - 1, j, k) + hx * bc.xpos.value;
        p.at(nx + 1, j, k) = p.at(nx - 1, j, k) + 2 * hx * bc.xpos.value;
      }
    }
  return true;
}

/* One MAC face pair of kernel_apply_3d_mac_boundary_conditions_level2_*:
 * normal N, tangentials A and B on the line through (i,j,k) with stride s
 * along the axis (n cells). */
template <typename T>
void mac2_line(T *N, T *A, T *B, int64_t s, int n, const ocu_bc &neg, const ocu_bc &pos) {
  if (neg.type == OCU_BC_PERIODIC) {
    N[-1 * s] = N[(n - 1) * s];
    N[-2 * s] = N[(n - 2) * s];
    A[-1 * s] = A[(n - 1) * s];
    A[-2 * s] = A[(n - 2) * s];
    B[-1 * s] = B[(n - 1) * s];
    B[-2 * s] = B[(n - 2) * s];
  } else if (neg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    N[-2 * s] = neg.value;
    N[-1 * s] = neg.value;
    N[0] = neg.value;
    A[-1 * s] = A[0] * (1 - 2 * neg.aux_value);
    A[-2 * s] = A[0] * (1 - 2 * neg.aux_value);
    B[-1 * s] = B[0] * (1 - 2 * neg.aux_value);
    B[-2 * s] = B[0] * (1 - 2 * neg.aux_value);
  }
  if (pos.type == OCU_BC_PERIODIC) {
    N[n * s] = N[0];
    N[(n + 1) * s] = N[1 * s];
    N[(n + 2) * s] = N[2 * s];
    A[n * s] = A[0];
    A[(n + 1) * s] = A[1 * s];
    B[n * s] = B[0];
    B[(n + 1) * s] = B[1 * s];
  } else if (pos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    N[n * s] = pos.value;
    N[(n + 1) * s] = pos.value;
    N[(n + 2) * s] = pos.value;
    A[n * s] = A[(n - 1) * s] * (1 - 2 * pos.aux_value);
    A[(n + 1) * s] = A[(n - 1) * s] * (1 - 2 * pos.aux_value);
    B[n * s] = B[(n - 1) * s] * (1 - 2 * pos.aux_value);
    B[(n + 1) * s] = B[(n - 1) * s] * (1 - 2 * pos.aux_value);
  }
}

/* apply_3d_mac_boundary_conditions_level2 (grid3dboundary.cu:940-1016):
 * x pass (j,k interior), y pass (i in [-1,nx], k interior), z pass ((i,j)
 * in [-1,nx]x[-1,ny]).  nx = u.nx() - 1 (the reference's u.nx() - u.gz()
 * rejects every congruent triple with two ghost layers). */
template <typename T>
bool bc_fill_mac_level2(const ocu_grid3d &ug, const ocu_grid3d &vg, const ocu_grid3d &wg, const ocu_bc_set &bc) {
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (b->type != OCU_BC_NONE && b->type != OCU_BC_PERIODIC && b->type != OCU_BC_FORCED_INFLOW_VARIABLE_SLIP)
      return false;
  if (!layout_match(ug, vg) || !layout_match(ug, wg)) return false;
  if (ug.gx < 2 || ug.gy < 2 || ug.gz < 2) return false;
  const int nx = ug.nx - 1, ny = ug.ny, nz = ug.nz;
  if (vg.nx != nx || vg.ny != ny + 1 || vg.nz != nz) return false;
  if (wg.nx != nx || wg.ny != ny || wg.nz != nz + 1) return false;
  View<T> u(ug), v(vg), w(wg);
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) mac2_line(&u.at(0, j, k), &v.at(0, j, k), &w.at(0, j, k), u.xs, nx, bc.xneg, bc.xpos);
  for (int i = -1; i <= nx; i++)
    for (int k = 0; k < nz; k++) mac2_line(&v.at(i, 0, k), &u.at(i, 0, k), &w.at(i, 0, k), v.ys, ny, bc.yneg, bc.ypos);
  for (int i = -1; i <= nx; i++)
    for (int j = -1; j <= ny; j++)
      mac2_line(&w.at(i, j, 0), &u.at(i, j, 0), &v.at(i, j, 0), (int64_t)1, nz, bc.zneg, bc.zpos);
  return true;
}

/* apply_boundary_conditions (grid3dsample.cu:52-88) */
inline void sample_bc(float &x, float &value_offset, float valid_nx, float period_nx, const ocu_bc &xpos,
                      const ocu_bc &xneg, float inv_hx) {
  if (x < 0) {
    if (xneg.type == OCU_BC_PERIODIC) {
      x = std::fmod(x, period_nx);
      if (x < 0) x += period_nx;
    } else if (x < -.5f) {
      if (xneg.type == OCU_BC_NEUMANN) value_offset += (-.5f - x) * xneg.value / inv_hx;
      x = -.5f;
    }
  } else if (x > period_nx) {
    if (xpos.type == OCU_BC_PERIODIC) {
      x = std::fmod(x, period_nx);
    } else if (x > valid_nx - .5f) {
      if (xpos.type == OCU_BC_NEUMANN) value_offset += (x - (valid_nx - .5f)) * xpos.value / inv_hx;
      x = valid_nx - .5f;
    }
  }
}

/* trilinear_lookup (grid3dsample.cu:24-49), T arithmetic, float result.
 * The cell index is kept in [-1, n-1] so that the lookup stays inside the
 * first ghost layer (the reference indexes one past it for a position
 * exactly at the upper periodic end). */
template <typename T>
float trilinear(const View<T> &p, float x, float y, float z) {
  int gi = (int)std::floor(x), gj = (int)std::floor(y), gk = (int)std::floor(z);
  gi = std::min(std::max(gi, -1), p.nx - 1);
  gj = std::min(std::max(gj, -1), p.ny - 1);
  gk = std::min(std::max(gk, -1), p.nz - 1);
  T sx = x - gi, sy = y - gj, sz = z - gk;
  float val = (((p.at(gi, gj, gk)) * (1 - sx) + (p.at(gi + 1, gj, gk)) * (sx)) * (1 - sy) +
               ((p.at(gi, gj + 1, gk)) * (1 - sx) + (p.at(gi + 1, gj + 1, gk)) * (sx)) * (sy)) * (1 - sz) +
              (((p.at(gi, gj, gk + 1)) * (1 - sx) + (p.at(gi + 1, gj, gk + 1)) * (sx)) * (1 - sy) +
               ((p.at(gi, gj + 1, gk + 1)) * (1 - sx) + (p.at(gi + 1, gj + 1, gk + 1)) * (sx)) * (sy)) * (sz);
  return val;
}

/* sample_points_3d (grid3dsample.cu:91-132, 137-171) */
template <typename T>
bool sample_points(T *out, const float *px, const float *py, const float *pz, int n, const ocu_grid3d &g,
                   const ocu_bc_set &bc, float period_nx, float period_ny, float period_nz, float hx, float hy,
                   float hz, float shiftx, float shifty, float shiftz) {
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (b->type != OCU_BC_PERIODIC && b->type != OCU_BC_NEUMANN && b->type != OCU_BC_DIRICHELET &&
        b->type != OCU_BC_SCALAR_SLIP)
      return false;
  if (period_nx > g.nx || period_ny > g.ny || period_nz > g.nz) return false;
  View<T> p(g);
  const float ihx = 1.0f / hx, ihy = 1.0f / hy, ihz = 1.0f / hz;
  for (int q = 0; q < n; q++) {
    float x = px[q], y = py[q], z = pz[q];
    x *= ihx;
    y *= ihy;
    z *= ihz;
    x -= shiftx;
    y -= shifty;
    z -= shiftz;
    float off = 0;
    sample_bc(x, off, (float)g.nx, period_nx, bc.xpos, bc.xneg, ihx);
    sample_bc(y, off, (float)g.ny, period_ny, bc.ypos, bc.yneg, ihy);
    sample_bc(z, off, (float)g.nz, period_nz, bc.zpos, bc.zneg, ihz);
    out[q] = trilinear(p, x, y, z) + off;
  }
  return true;
}

/* sample_points_mac_grid_3d (grid3dsample.cu:174-263), including the
 * reference's choice of the positive face's type on the negative branch and
 * its translation of forced inflow to Dirichlet for u and scalar slip for v
 * and w on every face (both clamp the same way) */
template <typename T>
bool sample_points_mac(T *vx, T *vy, T *vz, const float *px, const float *py, const float *pz, int n,
                       const ocu_grid3d &u, const ocu_grid3d &v, const ocu_grid3d &w, const ocu_bc_set &bc, float hx,
                       float hy, float hz) {
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (b->type != OCU_BC_PERIODIC && b->type != OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) return false;
  const int nx = u.nx - 1, ny = v.ny - 1, nz = w.nz - 1;
  if (v.nx != nx || w.nx != nx || u.ny != ny || w.ny != ny || u.nz != nz || v.nz != nz) return false;
  ocu_bc_set ub, vb, wb;
  ocu_bc inv{OCU_BC_INVALID, 0, 0};
  ub = vb = wb = ocu_bc_set{inv, inv, inv, inv, inv, inv};
  const int P = OCU_BC_PERIODIC, F = OCU_BC_FORCED_INFLOW_VARIABLE_SLIP, D = OCU_BC_DIRICHELET,
            S = OCU_BC_SCALAR_SLIP;
  if (bc.xpos.type == P) ub.xpos.type = vb.xpos.type = wb.xpos.type = P;
  else if (bc.xpos.type == F) ub.xpos.type = D, vb.xpos.type = S, wb.xpos.type = S;
  if (bc.xneg.type == P) ub.xneg.type = vb.xneg.type = wb.xneg.type = P;
  else if (bc.xpos.type == F) ub.xneg.type = D, vb.xneg.type = S, wb.xneg.type = S;
  if (bc.ypos.type == P) ub.ypos.type = vb.ypos.type = wb.ypos.type = P;
  else if (bc.ypos.type == F) ub.ypos.type = D, vb.ypos.type = S, wb.ypos.type = S;
  if (bc.yneg.type == P) ub.yneg.type = vb.yneg.type = wb.yneg.type = P;
  else if (bc.ypos.type == F) ub.yneg.type = D, vb.yneg.type = S, wb.yneg.type = S;
  if (bc.zpos.type == P) ub.zpos.type = vb.zpos.type = wb.zpos.type = P;
  else if (bc.zpos.type == F) ub.zpos.type = D, vb.zpos.type = S, wb.zpos.type = S;
  if (bc.zneg.type == P) ub.zneg.type = vb.zneg.type = wb.zneg.type = P;
  else if (bc.zpos.type == F) ub.zneg.type = D, vb.zneg.type = S, wb.zneg.type = S;
  return sample_points(vx, px, py, pz, n, u, ub, (float)nx, (float)ny, (float)nz, hx, hy, hz, 0, .5f, .5f) &&
         sample_points(vy, px, py, pz, n, v, vb, (float)nx, (float)ny, (float)nz, hx, hy, hz, .5f, 0, .5f) &&
         sample_points(vz, px, py, pz, n, w, wb, (float)nx, (float)ny, (float)nz, hx, hy, hz, .5f, .5f, 0);
}

}  // namespace orc

#endif
