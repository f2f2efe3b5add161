/*
 * ORACLE - test infrastructure only.
 *
 * CPU restatement of OpenCurrent's Grid3D layout, ghost-cell fills and grid
 * operations.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, to compare results against; the product never links it.
 *
 * Every function names the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order and the precision each
 * operation is carried out in (including float-literal vs double-literal
 * promotions), and the file is compiled with -ffp-contract=off so no FMA is
 * formed behind our back.
 */
#ifndef OCU_ORACLE_GRID_H
#define OCU_ORACLE_GRID_H

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "../include/ocu_types.h"

namespace orc {

/* Padded layout, ocustorage/grid3ddev.cu:358-400 (identical on the host,
 * grid3dhost.cu:318-357): pnz rounded up to a multiple of 16 for every T,
 * pre_padding = 16 - gz, shift = gx*pny*pnz + gy*pnz + gz + pre_padding. */
inline ocu_grid3d layout(int nx, int ny, int nz, int gx, int gy, int gz,
                         int padx, int pady, int padz, int dtype) {
  ocu_grid3d g;
  std::memset(&g, 0, sizeof(g));
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.gx = gx; g.gy = gy; g.gz = gz;
  g.pnx = nx + 2 * gx + padx;
  g.pny = ny + 2 * gy + pady;
  g.pnz = nz + 2 * gz + padz;
  if (g.pnz & 15) g.pnz = ((g.pnz >> 4) + 1) << 4;
  int64_t pre = 16 - gz;
  int64_t pnzpny = (int64_t)g.pnz * g.pny;
  g.allocated = pnzpny * g.pnx + pre;
  g.shift = (int64_t)gx * pnzpny + (int64_t)gy * g.pnz + gz + pre;
  g.dtype = dtype;
  g.base = nullptr;
  return g;
}

template <typename T>
struct View {
  T *p;  // cell (0,0,0)
  int nx, ny, nz, gx, gy, gz;
  int64_t xs, ys;
  explicit View(const ocu_grid3d &g)
      : p((T *)g.base + g.shift), nx(g.nx), ny(g.ny), nz(g.nz), gx(g.gx),
        gy(g.gy), gz(g.gz), xs((int64_t)g.pny * g.pnz), ys(g.pnz) {}
  View() : p(nullptr), nx(0), ny(0), nz(0), gx(0), gy(0), gz(0), xs(0), ys(0) {}
  T &at(int i, int j, int k) const { return p[i * xs + j * ys + k]; }
};

/* owning host grid with the reference layout */
template <typename T>
struct HostGrid {
  ocu_grid3d desc;
  std::vector<T> buf;
  HostGrid() { std::memset(&desc, 0, sizeof(desc)); }
  void init(int nx, int ny, int nz, int gx, int gy, int gz, int padx = 0,
            int pady = 0, int padz = 0) {
    desc = layout(nx, ny, nz, gx, gy, gz, padx, pady, padz,
                  sizeof(T) == 8 ? OCU_F64 : OCU_F32);
    buf.assign((size_t)desc.allocated, T(0));
    desc.base = buf.data();
  }
  void init_congruent(const ocu_grid3d &o) {
    init(o.nx, o.ny, o.nz, o.gx, o.gy, o.gz, o.pnx - (o.nx + 2 * o.gx),
         o.pny - (o.ny + 2 * o.gy), o.pnz - (o.nz + 2 * o.gz));
  }
  View<T> view() const { return View<T>(desc); }
  void clear_zero() { std::fill(buf.begin(), buf.end(), T(0)); }
};

inline bool layout_match(const ocu_grid3d &a, const ocu_grid3d &b) {
  return a.pnx == b.pnx && a.pny == b.pny && a.pnz == b.pnz;
}

/* ------------------------------------------------------------------ */
/* Ghost fills                                                          */
/* ------------------------------------------------------------------ */

/* Dirichlet ghost: kernel text `-phi + 2 * bc.value` - the product 2*value is
 * formed in float (int * float), the sum in T. */
template <typename T>
inline T dir_ghost(T inner, float value) { return -inner + (T)(2 * value); }

/* apply_3d_boundary_conditions_level1 (grid3dboundary.cu:758-820): three
 * passes x (j,k interior), y (i in [-1,nx], k interior), z (i in [-1,nx],
 * j in [-1,ny]); kernels _x :312-347, _y :274-309, _z :237-272.  Edges and
 * corners come from the later passes reading earlier ghosts. */
template <typename T>
bool bc_fill_level1(const ocu_grid3d &g, const ocu_bc_set &bc, double hxd,
                    double hyd, double hzd) {
  const int ok[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_DIRICHELET, OCU_BC_NEUMANN};
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (std::find(ok, ok + 4, b->type) == ok + 4) return false;
  if (g.gx < 1 || g.gy < 1 || g.gz < 1) return false;
  View<T> v(g);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T hx = (T)hxd, hy = (T)hyd, hz = (T)hzd;
  // x pass
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xneg.type == OCU_BC_PERIODIC) v.at(-1, j, k) = v.at(nx - 1, j, k);
      else if (bc.xneg.type == OCU_BC_DIRICHELET) v.at(-1, j, k) = dir_ghost<T>(v.at(0, j, k), bc.xneg.value);
      else if (bc.xneg.type == OCU_BC_NEUMANN) v.at(-1, j, k) = v.at(0, j, k) - hx * (T)bc.xneg.value;
      if (bc.xpos.type == OCU_BC_PERIODIC) v.at(nx, j, k) = v.at(0, j, k);
      else if (bc.xpos.type == OCU_BC_DIRICHELET) v.at(nx, j, k) = dir_ghost<T>(v.at(nx - 1, j, k), bc.xpos.value);
      else if (bc.xpos.type == OCU_BC_NEUMANN) v.at(nx, j, k) = v.at(nx - 1, j, k) + hx * (T)bc.xpos.value;
    }
  // y pass
  for (int i = -1; i <= nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.yneg.type == OCU_BC_PERIODIC) v.at(i, -1, k) = v.at(i, ny - 1, k);
      else if (bc.yneg.type == OCU_BC_DIRICHELET) v.at(i, -1, k) = dir_ghost<T>(v.at(i, 0, k), bc.yneg.value);
      else if (bc.yneg.type == OCU_BC_NEUMANN) v.at(i, -1, k) = v.at(i, 0, k) - hy * (T)bc.yneg.value;
      if (bc.ypos.type == OCU_BC_PERIODIC) v.at(i, ny, k) = v.at(i, 0, k);
      else if (bc.ypos.type == OCU_BC_DIRICHELET) v.at(i, ny, k) = dir_ghost<T>(v.at(i, ny - 1, k), bc.ypos.value);
      else if (bc.ypos.type == OCU_BC_NEUMANN) v.at(i, ny, k) = v.at(i, ny - 1, k) + hy * (T)bc.ypos.value;
    }
  // z pass
  for (int i = -1; i <= nx; i++)
    for (int j = -1; j <= ny; j++) {
      if (bc.zneg.type == OCU_BC_PERIODIC) v.at(i, j, -1) = v.at(i, j, nz - 1);
      else if (bc.zneg.type == OCU_BC_DIRICHELET) v.at(i, j, -1) = dir_ghost<T>(v.at(i, j, 0), bc.zneg.value);
      else if (bc.zneg.type == OCU_BC_NEUMANN) v.at(i, j, -1) = v.at(i, j, 0) - hz * (T)bc.zneg.value;
      if (bc.zpos.type == OCU_BC_PERIODIC) v.at(i, j, nz) = v.at(i, j, 0);
      else if (bc.zpos.type == OCU_BC_DIRICHELET) v.at(i, j, nz) = dir_ghost<T>(v.at(i, j, nz - 1), bc.zpos.value);
      else if (bc.zpos.type == OCU_BC_NEUMANN) v.at(i, j, nz) = v.at(i, j, nz - 1) + hz * (T)bc.zpos.value;
    }
  return true;
}

/* Device apply_3d_boundary_conditions_level1_nocorners
 * (grid3dboundary.cu:28-119, launcher :719-753): faces only.  The Dirichlet
 * ghost is written `-phi + 2.0 * (T)value`, i.e. formed in double and
 * rounded to T on the store. */
template <typename T>
bool bc_fill_nocorners_device(const ocu_grid3d &g, const ocu_bc_set &bc,
                              double hxd, double hyd, double hzd) {
  const int ok[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_DIRICHELET, OCU_BC_NEUMANN};
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (std::find(ok, ok + 4, b->type) == ok + 4) return false;
  if (g.gx < 1 || g.gy < 1 || g.gz < 1) return false;
  View<T> v(g);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T hx = (T)hxd, hy = (T)hyd, hz = (T)hzd;
  auto dg = [](T inner, float value) -> T { return (T)(-(double)inner + 2.0 * (double)(T)value); };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (bc.zneg.type == OCU_BC_PERIODIC) v.at(i, j, -1) = v.at(i, j, nz - 1);
      else if (bc.zneg.type == OCU_BC_DIRICHELET) v.at(i, j, -1) = dg(v.at(i, j, 0), bc.zneg.value);
      else if (bc.zneg.type == OCU_BC_NEUMANN) v.at(i, j, -1) = v.at(i, j, 0) - hz * (T)bc.zneg.value;
      if (bc.zpos.type == OCU_BC_PERIODIC) v.at(i, j, nz) = v.at(i, j, 0);
      else if (bc.zpos.type == OCU_BC_DIRICHELET) v.at(i, j, nz) = dg(v.at(i, j, nz - 1), bc.zpos.value);
      else if (bc.zpos.type == OCU_BC_NEUMANN) v.at(i, j, nz) = v.at(i, j, nz - 1) + hz * (T)bc.zpos.value;
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.yneg.type == OCU_BC_PERIODIC) v.at(i, -1, k) = v.at(i, ny - 1, k);
      else if (bc.yneg.type == OCU_BC_DIRICHELET) v.at(i, -1, k) = dg(v.at(i, 0, k), bc.yneg.value);
      else if (bc.yneg.type == OCU_BC_NEUMANN) v.at(i, -1, k) = v.at(i, 0, k) - hy * (T)bc.yneg.value;
      if (bc.ypos.type == OCU_BC_PERIODIC) v.at(i, ny, k) = v.at(i, 0, k);
      else if (bc.ypos.type == OCU_BC_DIRICHELET) v.at(i, ny, k) = dg(v.at(i, ny - 1, k), bc.ypos.value);
      else if (bc.ypos.type == OCU_BC_NEUMANN) v.at(i, ny, k) = v.at(i, ny - 1, k) + hy * (T)bc.ypos.value;
    }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xneg.type == OCU_BC_PERIODIC) v.at(-1, j, k) = v.at(nx - 1, j, k);
      else if (bc.xneg.type == OCU_BC_DIRICHELET) v.at(-1, j, k) = dg(v.at(0, j, k), bc.xneg.value);
      else if (bc.xneg.type == OCU_BC_NEUMANN) v.at(-1, j, k) = v.at(0, j, k) - hx * (T)bc.xneg.value;
      if (bc.xpos.type == OCU_BC_PERIODIC) v.at(nx, j, k) = v.at(0, j, k);
      else if (bc.xpos.type == OCU_BC_DIRICHELET) v.at(nx, j, k) = dg(v.at(nx - 1, j, k), bc.xpos.value);
      else if (bc.xpos.type == OCU_BC_NEUMANN) v.at(nx, j, k) = v.at(nx - 1, j, k) + hx * (T)bc.xpos.value;
    }
  return true;
}

/* Host apply_3d_boundary_conditions_level1_nocorners
 * (grid3dboundary.cu:1018-1111): z faces, then y, then x; Dirichlet ghost
 * `-phi + (T)2 * (T)value` evaluated in T. */
template <typename T>
bool bc_fill_nocorners_host(const ocu_grid3d &g, const ocu_bc_set &bc,
                            double hxd, double hyd, double hzd) {
  const int ok[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_DIRICHELET, OCU_BC_NEUMANN};
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (std::find(ok, ok + 4, b->type) == ok + 4) return false;
  View<T> v(g);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T hx = (T)hxd, hy = (T)hyd, hz = (T)hzd;
  auto dg = [](T inner, float value) -> T { return -inner + (T)2 * (T)value; };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      if (bc.zneg.type == OCU_BC_PERIODIC) v.at(i, j, -1) = v.at(i, j, nz - 1);
      else if (bc.zneg.type == OCU_BC_DIRICHELET) v.at(i, j, -1) = dg(v.at(i, j, 0), bc.zneg.value);
      else if (bc.zneg.type == OCU_BC_NEUMANN) v.at(i, j, -1) = v.at(i, j, 0) - hz * (T)bc.zneg.value;
      if (bc.zpos.type == OCU_BC_PERIODIC) v.at(i, j, nz) = v.at(i, j, 0);
      else if (bc.zpos.type == OCU_BC_DIRICHELET) v.at(i, j, nz) = dg(v.at(i, j, nz - 1), bc.zpos.value);
      else if (bc.zpos.type == OCU_BC_NEUMANN) v.at(i, j, nz) = v.at(i, j, nz - 1) + hz * (T)bc.zpos.value;
    }
  for (int i = 0; i < nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.yneg.type == OCU_BC_PERIODIC) v.at(i, -1, k) = v.at(i, ny - 1, k);
      else if (bc.yneg.type == OCU_BC_DIRICHELET) v.at(i, -1, k) = dg(v.at(i, 0, k), bc.yneg.value);
      else if (bc.yneg.type == OCU_BC_NEUMANN) v.at(i, -1, k) = v.at(i, 0, k) - hy * (T)bc.yneg.value;
      if (bc.ypos.type == OCU_BC_PERIODIC) v.at(i, ny, k) = v.at(i, 0, k);
      else if (bc.ypos.type == OCU_BC_DIRICHELET) v.at(i, ny, k) = dg(v.at(i, ny - 1, k), bc.ypos.value);
      else if (bc.ypos.type == OCU_BC_NEUMANN) v.at(i, ny, k) = v.at(i, ny - 1, k) + hy * (T)bc.ypos.value;
    }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xneg.type == OCU_BC_PERIODIC) v.at(-1, j, k) = v.at(nx - 1, j, k);
      else if (bc.xneg.type == OCU_BC_DIRICHELET) v.at(-1, j, k) = dg(v.at(0, j, k), bc.xneg.value);
      else if (bc.xneg.type == OCU_BC_NEUMANN) v.at(-1, j, k) = v.at(0, j, k) - hx * (T)bc.xneg.value;
      if (bc.xpos.type == OCU_BC_PERIODIC) v.at(nx, j, k) = v.at(0, j, k);
      else if (bc.xpos.type == OCU_BC_DIRICHELET) v.at(nx, j, k) = dg(v.at(nx - 1, j, k), bc.xpos.value);
      else if (bc.xpos.type == OCU_BC_NEUMANN) v.at(nx, j, k) = v.at(nx - 1, j, k) + hx * (T)bc.xpos.value;
    }
  return true;
}

/* apply_3d_mac_boundary_conditions_level1 (grid3dboundary.cu:825-900,
 * kernels _z :551-602, _y :604-658, _x :660-716).  nx = u.nx()-1.
 * FORCED_INFLOW_VARIABLE_SLIP: normal component = value on the ghost and the
 * boundary face; tangential ghost = interior * (1.0f - 2.0f*aux) (the factor
 * is a float).  PERIODIC wraps normal (two layers on the + side) and
 * tangential components. */
template <typename T>
bool bc_fill_mac_level1(const ocu_grid3d &ug, const ocu_grid3d &vg,
                        const ocu_grid3d &wg, const ocu_bc_set &bc) {
  const int okt[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_FORCED_INFLOW_VARIABLE_SLIP};
  const ocu_bc *all[6] = {&bc.xpos, &bc.xneg, &bc.ypos, &bc.yneg, &bc.zpos, &bc.zneg};
  for (auto b : all)
    if (std::find(okt, okt + 3, b->type) == okt + 3) return false;
  if (!layout_match(ug, vg) || !layout_match(ug, wg)) return false;
  if (ug.gx < 1 || ug.gy < 1 || ug.gz < 1) return false;
  const int nx = ug.nx - 1, ny = ug.ny, nz = ug.nz;
  if (vg.nx != nx || vg.ny != ny + 1 || vg.nz != nz) return false;
  if (wg.nx != nx || wg.ny != ny || wg.nz != nz + 1) return false;
  View<T> u(ug), v(vg), w(wg);
  auto fac = [](float aux) -> float { return 1.0f - 2.0f * aux; };
  // x pass: j in [0,ny), k in [0,nz)
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++) {
      if (bc.xneg.type == OCU_BC_PERIODIC) {
        u.at(-1, j, k) = u.at(nx - 1, j, k);
        v.at(-1, j, k) = v.at(nx - 1, j, k);
        w.at(-1, j, k) = w.at(nx - 1, j, k);
      } else if (bc.xneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        u.at(-1, j, k) = (T)bc.xneg.value;
        u.at(0, j, k) = (T)bc.xneg.value;
        float f = fac(bc.xneg.aux_value);
        v.at(-1, j, k) = v.at(0, j, k) * (T)f;
        w.at(-1, j, k) = w.at(0, j, k) * (T)f;
      }
      if (bc.xpos.type == OCU_BC_PERIODIC) {
        u.at(nx, j, k) = u.at(0, j, k);
        u.at(nx + 1, j, k) = u.at(1, j, k);
        v.at(nx, j, k) = v.at(0, j, k);
        w.at(nx, j, k) = w.at(0, j, k);
      } else if (bc.xpos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        u.at(nx, j, k) = (T)bc.xpos.value;
        u.at(nx + 1, j, k) = (T)bc.xpos.value;
        float f = fac(bc.xpos.aux_value);
        v.at(nx, j, k) = v.at(nx - 1, j, k) * (T)f;
        w.at(nx, j, k) = w.at(nx - 1, j, k) * (T)f;
      }
    }
  // y pass: i in [-1,nx], k in [0,nz)
  for (int i = -1; i <= nx; i++)
    for (int k = 0; k < nz; k++) {
      if (bc.yneg.type == OCU_BC_PERIODIC) {
        v.at(i, -1, k) = v.at(i, ny - 1, k);
        u.at(i, -1, k) = u.at(i, ny - 1, k);
        w.at(i, -1, k) = w.at(i, ny - 1, k);
      } else if (bc.yneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        v.at(i, -1, k) = (T)bc.yneg.value;
        v.at(i, 0, k) = (T)bc.yneg.value;
        float f = fac(bc.yneg.aux_value);
        u.at(i, -1, k) = u.at(i, 0, k) * (T)f;
        w.at(i, -1, k) = w.at(i, 0, k) * (T)f;
      }
      if (bc.ypos.type == OCU_BC_PERIODIC) {
        v.at(i, ny, k) = v.at(i, 0, k);
        v.at(i, ny + 1, k) = v.at(i, 1, k);
        u.at(i, ny, k) = u.at(i, 0, k);
        w.at(i, ny, k) = w.at(i, 0, k);
      } else if (bc.ypos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        v.at(i, ny, k) = (T)bc.ypos.value;
        v.at(i, ny + 1, k) = (T)bc.ypos.value;
        float f = fac(bc.ypos.aux_value);
        u.at(i, ny, k) = u.at(i, ny - 1, k) * (T)f;
        w.at(i, ny, k) = w.at(i, ny - 1, k) * (T)f;
      }
    }
  // z pass: i in [-1,nx], j in [-1,ny]
  for (int i = -1; i <= nx; i++)
    for (int j = -1; j <= ny; j++) {
      if (bc.zneg.type == OCU_BC_PERIODIC) {
        w.at(i, j, -1) = w.at(i, j, nz - 1);
        u.at(i, j, -1) = u.at(i, j, nz - 1);
        v.at(i, j, -1) = v.at(i, j, nz - 1);
      } else if (bc.zneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        w.at(i, j, -1) = (T)bc.zneg.value;
        w.at(i, j, 0) = (T)bc.zneg.value;
        float f = fac(bc.zneg.aux_value);
        u.at(i, j, -1) = u.at(i, j, 0) * (T)f;
        v.at(i, j, -1) = v.at(i, j, 0) * (T)f;
      }
      if (bc.zpos.type == OCU_BC_PERIODIC) {
        w.at(i, j, nz) = w.at(i, j, 0);
        w.at(i, j, nz + 1) = w.at(i, j, 1);
        u.at(i, j, nz) = u.at(i, j, 0);
        v.at(i, j, nz) = v.at(i, j, 0);
      } else if (bc.zpos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
        w.at(i, j, nz) = (T)bc.zpos.value;
        w.at(i, j, nz + 1) = (T)bc.zpos.value;
        float f = fac(bc.zpos.aux_value);
        u.at(i, j, nz) = u.at(i, j, nz - 1) * (T)f;
        v.at(i, j, nz) = v.at(i, j, nz - 1) * (T)f;
      }
    }
  return true;
}

/* ------------------------------------------------------------------ */
/* Whole-grid operations (ocustorage/grid3ddev.cu)                       */
/* ------------------------------------------------------------------ */

/* linear_combination with two grids (grid3ddev.cu:303-349): if more than 3%
 * of the allocation is outside the interior, interior only
 * (`g1*a1 + g2*a2`, :35-55); otherwise every allocated element (:67-75). */
template <typename T>
bool lincomb2(const ocu_grid3d &r, T a1, const ocu_grid3d &g1, T a2, const ocu_grid3d &g2) {
  if (!layout_match(r, g1) || !layout_match(r, g2)) return false;
  int64_t interior = (int64_t)r.nx * r.ny * r.nz;
  if ((r.allocated - interior) / (float)r.allocated > .03f) {
    View<T> R(r), A(g1), B(g2);
    for (int i = 0; i < r.nx; i++)
      for (int j = 0; j < r.ny; j++)
        for (int k = 0; k < r.nz; k++) R.at(i, j, k) = A.at(i, j, k) * a1 + B.at(i, j, k) * a2;
  } else {
    T *R = (T *)r.base;
    const T *A = (const T *)g1.base, *B = (const T *)g2.base;
    for (int64_t n = 0; n < r.allocated; n++) R[n] = A[n] * a1 + B[n] * a2;
  }
  return true;
}

/* linear_combination with one grid (grid3ddev.cu:283-300, kernel :57-65):
 * always over the whole allocation. */
template <typename T>
bool lincomb1(const ocu_grid3d &r, T a1, const ocu_grid3d &g1) {
  if (!layout_match(r, g1)) return false;
  T *R = (T *)r.base;
  const T *A = (const T *)g1.base;
  for (int64_t n = 0; n < r.allocated; n++) R[n] = A[n] * a1;
  return true;
}

/* reductions over the interior (grid3dops.cu:26-130, reduction_op.h:30-108).
 * op: 0 maxabs, 1 max, 2 min, 3 sum, 4 sqrsum, 5 checknan.  The oracle
 * accumulates sums in double in i,j,k order; max/min/maxabs are exact. */
template <typename T>
double reduce(const ocu_grid3d &g, int op) {
  View<T> v(g);
  double acc = 0;
  bool first = true;
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++)
      for (int k = 0; k < g.nz; k++) {
        double x = (double)v.at(i, j, k);
        switch (op) {
          case 0: acc = first ? std::fabs(x) : std::max(acc, std::fabs(x)); break;
          case 1: acc = first ? x : std::max(acc, x); break;
          case 2: acc = first ? x : std::min(acc, x); break;
          case 3: acc += x; break;
          case 4: acc += x * x; break;
          case 5: if (!std::isfinite(x)) acc = x; break;
        }
        first = false;
      }
  return acc;
}

}  // namespace orc

#endif
