// Level-1 ghost values of a cell-centred grid whose six faces are Neumann or
// Dirichlet, computed from the interior cell each ghost derives from.
//
// apply_3d_boundary_conditions_level1 (ocustorage/grid3dboundary.cu:237-347)
// fills x ghosts from the interior, then y ghosts over i in [-1,nx] from the
// x-filled values, then z ghosts over (i,j) in [-1,nx]x[-1,ny].  Without
// periodic faces every ghost therefore derives from exactly one interior
// cell: clamp each out-of-range index into the interior and apply the x, y
// and z face transforms (in that order) for the axes that were out of range.
// A kernel that produces interior cell (i,j,k) can write all of the ghosts
// that derive from it with the same operations, bit for bit.
#pragma once

#include "ocu_hip_common.h"

namespace ocu_hip {

// every face is Neumann or Dirichlet (ghosts are local functions of the interior)
inline bool gfaces_local(const ocu_bc_set *bc) {
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  for (auto b : all)
    if (b->type != OCU_BC_NEUMANN && b->type != OCU_BC_DIRICHELET) return false;
  return true;
}

// every face value is zero: the level-1 fill of the zero grid is all +0.0
inline bool gfaces_homogeneous(const ocu_bc_set *bc) {
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  for (auto b : all)
    if (b->value != 0.0f) return false;
  return true;
}

// One face transform is ghost = s*a + add with s = +-1 (exact), which is
// bit-identical to the reference's expressions: Dirichlet `-phi + 2*value`
// (2*value in float, grid3dboundary.cu:250) has s = -1, add = (T)(2*value);
// Neumann `phi - h*value` / `phi + h*value` (:256-262) has s = 1 and
// add = -(h*value) / +(h*value), since a - x == a + (-x) in IEEE arithmetic.
template <typename T>
struct GhostL1 {
  T s[6], add[6];  // xpos, xneg, ypos, yneg, zpos, zneg

  __device__ __forceinline__ T xf(int f, T a) const { return s[f] * a + add[f]; }
  // value of the ghost at offset (sx,sy,sz) in {-1,0,1}^3 from the interior
  // value v it derives from: x, then y, then z transforms
  __device__ __forceinline__ T chain(int sx, int sy, int sz, T v) const {
    if (sx) v = xf(sx < 0 ? 1 : 0, v);
    if (sy) v = xf(sy < 0 ? 3 : 2, v);
    if (sz) v = xf(sz < 0 ? 5 : 4, v);
    return v;
  }
  // write every ghost of g that derives from interior cell (i,j,k) = v
  __device__ void write(const DGrid<T> &g, int i, int j, int k, T v) const {
    const bool xl = i == 0, xh = i == g.nx - 1, yl = j == 0, yh = j == g.ny - 1, zl = k == 0, zh = k == g.nz - 1;
    if (!(xl || xh || yl || yh || zl || zh)) return;
    for (int sx = -1; sx <= 1; sx++) {
      if ((sx < 0 && !xl) || (sx > 0 && !xh)) continue;
      for (int sy = -1; sy <= 1; sy++) {
        if ((sy < 0 && !yl) || (sy > 0 && !yh)) continue;
        for (int sz = -1; sz <= 1; sz++) {
          if ((sz < 0 && !zl) || (sz > 0 && !zh)) continue;
          if (!sx && !sy && !sz) continue;
          g.at(i + sx, j + sy, k + sz) = chain(sx, sy, sz, v);
        }
      }
    }
  }
};

// Faces may also be periodic in pairs: the level-1 fill copies across a
// periodic axis (ghost -1 from cell n-1, ghost n from cell 0) and transforms
// across the others, x then y then z, so every ghost still derives from one
// interior cell.  write() stores every ghost that derives from cell (i,j,k).
template <typename T>
struct GhostP {
  GhostL1<T> t;
  int per[3];  // x, y, z axis periodic
  __device__ void write(const DGrid<T> &g, int i, int j, int k, T v) const {
    const int n[3] = {g.nx, g.ny, g.nz}, c[3] = {i, j, k};
    int off[3][3], face[3][3], cnt[3];
    bool edge = false;
    for (int a = 0; a < 3; a++) {
      off[a][0] = 0;
      face[a][0] = -1;
      cnt[a] = 1;
      if (c[a] == 0) {  // the low ghost (or, periodic, the high one) derives from here
        off[a][cnt[a]] = per[a] ? n[a] : -1;
        face[a][cnt[a]++] = per[a] ? -1 : 2 * a + 1;
      }
      if (c[a] == n[a] - 1) {
        off[a][cnt[a]] = per[a] ? -n[a] : 1;
        face[a][cnt[a]++] = per[a] ? -1 : 2 * a;
      }
      edge = edge || cnt[a] > 1;
    }
    if (!edge) return;
    for (int ax = 0; ax < cnt[0]; ax++)
      for (int ay = 0; ay < cnt[1]; ay++)
        for (int az = 0; az < cnt[2]; az++) {
          if (!ax && !ay && !az) continue;
          T w = v;
          if (face[0][ax] >= 0) w = t.xf(face[0][ax], w);
          if (face[1][ay] >= 0) w = t.xf(face[1][ay], w);
          if (face[2][az] >= 0) w = t.xf(face[2][az], w);
          g.at(i + off[0][ax], j + off[1][ay], k + off[2][az]) = w;
        }
  }
};

// GhostP applies to the level-1 fill of bc when every face is Neumann,
// Dirichlet or periodic with its opposite face
inline bool gfaces_paired(const ocu_bc_set *bc) {
  const ocu_bc *lo[3] = {&bc->xneg, &bc->yneg, &bc->zneg}, *hi[3] = {&bc->xpos, &bc->ypos, &bc->zpos};
  for (int a = 0; a < 3; a++) {
    const bool pl = lo[a]->type == OCU_BC_PERIODIC, ph = hi[a]->type == OCU_BC_PERIODIC;
    if (pl != ph) return false;
    if (pl) continue;
    for (const ocu_bc *b : {lo[a], hi[a]})
      if (b->type != OCU_BC_NEUMANN && b->type != OCU_BC_DIRICHELET) return false;
  }
  return true;
}

template <typename T>
inline GhostL1<T> ghost_l1(const ocu_bc_set *bc, double hx, double hy, double hz);

template <typename T>
inline GhostP<T> ghost_p(const ocu_bc_set *bc, double hx, double hy, double hz) {
  GhostP<T> g;
  g.t = ghost_l1<T>(bc, hx, hy, hz);
  g.per[0] = bc->xpos.type == OCU_BC_PERIODIC;
  g.per[1] = bc->ypos.type == OCU_BC_PERIODIC;
  g.per[2] = bc->zpos.type == OCU_BC_PERIODIC;
  return g;
}

template <typename T>
inline GhostL1<T> ghost_l1(const ocu_bc_set *bc, double hx, double hy, double hz) {
  GhostL1<T> g;
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  const T h[3] = {(T)hx, (T)hy, (T)hz};
  for (int f = 0; f < 6; f++) {
    const ocu_bc &b = *all[f];
    if (b.type == OCU_BC_DIRICHELET) {
      g.s[f] = (T)-1;
      g.add[f] = (T)(2 * b.value);
    } else {
      const T hv = h[f / 2] * (T)b.value;
      g.s[f] = (T)1;
      g.add[f] = (f & 1) ? -hv : hv;
    }
  }
  return g;
}

}  // namespace ocu_hip
