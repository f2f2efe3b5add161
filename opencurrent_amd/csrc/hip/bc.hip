// Ghost-cell fills (ocustorage/grid3dboundary.cu).
//
// apply_3d_boundary_conditions_level1 runs three dependent passes in the
// reference (x, then y over i in [-1,nx], then z over i in [-1,nx] x
// j in [-1,ny]), so edges and corners are produced by the later passes from
// ghosts written by the earlier ones.  Here one launch covers the whole
// first ghost layer: each ghost cell evaluates the same chain of per-axis
// transforms from the interior (or untouched) cell it ultimately derives
// from, with the same operations in the same order, so the values are
// identical and no cell reads a value written by the same launch.
#include "ocu_hip_common.h"

using namespace ocu_hip;

namespace {

struct Face {
  int type;
  float value;
  float aux;
};
struct Faces {
  Face xpos, xneg, ypos, yneg, zpos, zneg;
};

inline Faces faces(const ocu_bc_set *b) {
  Faces f;
  f.xpos = {b->xpos.type, b->xpos.value, b->xpos.aux_value};
  f.xneg = {b->xneg.type, b->xneg.value, b->xneg.aux_value};
  f.ypos = {b->ypos.type, b->ypos.value, b->ypos.aux_value};
  f.yneg = {b->yneg.type, b->yneg.value, b->yneg.aux_value};
  f.zpos = {b->zpos.type, b->zpos.value, b->zpos.aux_value};
  f.zneg = {b->zneg.type, b->zneg.value, b->zneg.aux_value};
  return f;
}

__device__ __forceinline__ bool writes(const Face &f) {
  return f.type == OCU_BC_PERIODIC || f.type == OCU_BC_DIRICHELET || f.type == OCU_BC_NEUMANN;
}

// value of the ghost on face f from the interior value a it derives from.
// Level-1 Dirichlet: `-phi + 2 * bc.value` (2*value in float,
// grid3dboundary.cu:250); no-corner Dirichlet: `-phi + 2.0 * (T)value`
// (double, :44).  Neumann: phi -/+ h*value in T.
template <typename T, bool DBL_DIRICHLET>
__device__ __forceinline__ T xform(const Face &f, bool neg, T h, T a) {
  if (f.type == OCU_BC_DIRICHELET) {
    if (DBL_DIRICHLET) return (T)(-(double)a + 2.0 * (double)(T)f.value);
    return -a + (T)(2 * f.value);
  }
  if (f.type == OCU_BC_NEUMANN) return neg ? a - h * (T)f.value : a + h * (T)f.value;
  return a;  // periodic: the source index already wrapped
}

// interior source index along one axis for the ghost at c (c == -1 or n)
__device__ __forceinline__ int src_index(const Face &f, bool neg, int n) {
  if (f.type == OCU_BC_PERIODIC) return neg ? n - 1 : 0;
  return neg ? 0 : n - 1;
}

template <typename T>
struct L1 {
  DGrid<T> g;
  Faces f;
  T hx, hy, hz;

  // value after the x pass at (i,j,k), j,k interior
  __device__ T vx(int i, int j, int k) const {
    if (i >= 0 && i < g.nx) return g.at(i, j, k);
    bool neg = i < 0;
    const Face &F = neg ? f.xneg : f.xpos;
    if (!writes(F)) return g.at(i, j, k);
    return xform<T, false>(F, neg, hx, g.at(src_index(F, neg, g.nx), j, k));
  }
  // value after the x and y passes at (i,j,k), k interior
  __device__ T vxy(int i, int j, int k) const {
    if (j >= 0 && j < g.ny) return vx(i, j, k);
    bool neg = j < 0;
    const Face &F = neg ? f.yneg : f.ypos;
    if (!writes(F)) return g.at(i, j, k);
    return xform<T, false>(F, neg, hy, vx(i, src_index(F, neg, g.ny), k));
  }
};

// linear thread index -> ghost cell of the first layer:
//   [0, nZ)        z faces, (i,j) in [-1,nx]x[-1,ny], k in {-1, nz}
//   [nZ, nZ+nY)    y faces, i in [-1,nx], k in [0,nz), j in {-1, ny}
//   [.., +nX)      x faces, j in [0,ny), k in [0,nz), i in {-1, nx}
template <typename T>
__global__ void k_bc_level1(L1<T> s, int64_t nZ, int64_t nY, int64_t nX) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = s.g.nx, ny = s.g.ny, nz = s.g.nz;
  if (t < nZ) {
    int64_t per = (int64_t)(nx + 2) * (ny + 2);
    bool neg = t < per;
    int64_t q = neg ? t : t - per;
    int i = (int)(q / (ny + 2)) - 1, j = (int)(q % (ny + 2)) - 1;
    const Face &F = neg ? s.f.zneg : s.f.zpos;
    if (!writes(F)) return;
    int k = neg ? -1 : nz;
    s.g.at(i, j, k) = xform<T, false>(F, neg, s.hz, s.vxy(i, j, src_index(F, neg, nz)));
    return;
  }
  t -= nZ;
  if (t < nY) {
    int64_t per = (int64_t)(nx + 2) * nz;
    bool neg = t < per;
    int64_t q = neg ? t : t - per;
    int i = (int)(q / nz) - 1, k = (int)(q % nz);
    const Face &F = neg ? s.f.yneg : s.f.ypos;
    if (!writes(F)) return;
    int j = neg ? -1 : ny;
    s.g.at(i, j, k) = xform<T, false>(F, neg, s.hy, s.vx(i, src_index(F, neg, ny), k));
    return;
  }
  t -= nY;
  if (t < nX) {
    int64_t per = (int64_t)ny * nz;
    bool neg = t < per;
    int64_t q = neg ? t : t - per;
    int j = (int)(q / nz), k = (int)(q % nz);
    const Face &F = neg ? s.f.xneg : s.f.xpos;
    if (!writes(F)) return;
    int i = neg ? -1 : nx;
    s.g.at(i, j, k) = xform<T, false>(F, neg, s.hx, s.g.at(src_index(F, neg, nx), j, k));
  }
}

// device no-corner fill (grid3dboundary.cu:28-119): faces only, no chains
template <typename T>
__global__ void k_bc_nocorners(DGrid<T> g, Faces f, T hx, T hy, T hz, int64_t nZ, int64_t nY, int64_t nX) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (t < nZ) {
    int i = (int)(t / ny), j = (int)(t % ny);
    if (writes(f.zneg)) g.at(i, j, -1) = xform<T, true>(f.zneg, true, hz, g.at(i, j, src_index(f.zneg, true, nz)));
    if (writes(f.zpos)) g.at(i, j, nz) = xform<T, true>(f.zpos, false, hz, g.at(i, j, src_index(f.zpos, false, nz)));
    return;
  }
  t -= nZ;
  if (t < nY) {
    int i = (int)(t / nz), k = (int)(t % nz);
    if (writes(f.yneg)) g.at(i, -1, k) = xform<T, true>(f.yneg, true, hy, g.at(i, src_index(f.yneg, true, ny), k));
    if (writes(f.ypos)) g.at(i, ny, k) = xform<T, true>(f.ypos, false, hy, g.at(i, src_index(f.ypos, false, ny), k));
    return;
  }
  t -= nY;
  if (t < nX) {
    int j = (int)(t / nz), k = (int)(t % nz);
    if (writes(f.xneg)) g.at(-1, j, k) = xform<T, true>(f.xneg, true, hx, g.at(src_index(f.xneg, true, nx), j, k));
    if (writes(f.xpos)) g.at(nx, j, k) = xform<T, true>(f.xpos, false, hx, g.at(src_index(f.xpos, false, nx), j, k));
  }
}

// ---- MAC fill: the reference's three dependent passes (x, y, z) ----

// tangential factor (1.0f - 2.0f * aux) is a float
template <typename T>
__global__ void k_mac_x(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces f, int nx, int ny, int nz) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (k >= nz || j >= ny) return;
  if (f.xneg.type == OCU_BC_PERIODIC) {
    u.at(-1, j, k) = u.at(nx - 1, j, k);
    v.at(-1, j, k) = v.at(nx - 1, j, k);
    w.at(-1, j, k) = w.at(nx - 1, j, k);
  } else if (f.xneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    u.at(-1, j, k) = (T)f.xneg.value;
    u.at(0, j, k) = (T)f.xneg.value;
    float fac = 1.0f - 2.0f * f.xneg.aux;
    v.at(-1, j, k) = v.at(0, j, k) * (T)fac;
    w.at(-1, j, k) = w.at(0, j, k) * (T)fac;
  }
  if (f.xpos.type == OCU_BC_PERIODIC) {
    u.at(nx, j, k) = u.at(0, j, k);
    u.at(nx + 1, j, k) = u.at(1, j, k);
    v.at(nx, j, k) = v.at(0, j, k);
    w.at(nx, j, k) = w.at(0, j, k);
  } else if (f.xpos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    u.at(nx, j, k) = (T)f.xpos.value;
    u.at(nx + 1, j, k) = (T)f.xpos.value;
    float fac = 1.0f - 2.0f * f.xpos.aux;
    v.at(nx, j, k) = v.at(nx - 1, j, k) * (T)fac;
    w.at(nx, j, k) = w.at(nx - 1, j, k) * (T)fac;
  }
}

template <typename T>
__global__ void k_mac_y(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces f, int nx, int ny, int nz) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int i = (int)blockIdx.y - 1;
  if (k >= nz || i > nx) return;
  if (f.yneg.type == OCU_BC_PERIODIC) {
    v.at(i, -1, k) = v.at(i, ny - 1, k);
    u.at(i, -1, k) = u.at(i, ny - 1, k);
    w.at(i, -1, k) = w.at(i, ny - 1, k);
  } else if (f.yneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    v.at(i, -1, k) = (T)f.yneg.value;
    v.at(i, 0, k) = (T)f.yneg.value;
    float fac = 1.0f - 2.0f * f.yneg.aux;
    u.at(i, -1, k) = u.at(i, 0, k) * (T)fac;
    w.at(i, -1, k) = w.at(i, 0, k) * (T)fac;
  }
  if (f.ypos.type == OCU_BC_PERIODIC) {
    v.at(i, ny, k) = v.at(i, 0, k);
    v.at(i, ny + 1, k) = v.at(i, 1, k);
    u.at(i, ny, k) = u.at(i, 0, k);
    w.at(i, ny, k) = w.at(i, 0, k);
  } else if (f.ypos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    v.at(i, ny, k) = (T)f.ypos.value;
    v.at(i, ny + 1, k) = (T)f.ypos.value;
    float fac = 1.0f - 2.0f * f.ypos.aux;
    u.at(i, ny, k) = u.at(i, ny - 1, k) * (T)fac;
    w.at(i, ny, k) = w.at(i, ny - 1, k) * (T)fac;
  }
}

template <typename T>
__global__ void k_mac_z(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces f, int nx, int ny, int nz) {
  int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  int i = (int)blockIdx.y - 1;
  if (j > ny || i > nx) return;
  if (f.zneg.type == OCU_BC_PERIODIC) {
    w.at(i, j, -1) = w.at(i, j, nz - 1);
    u.at(i, j, -1) = u.at(i, j, nz - 1);
    v.at(i, j, -1) = v.at(i, j, nz - 1);
  } else if (f.zneg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    w.at(i, j, -1) = (T)f.zneg.value;
    w.at(i, j, 0) = (T)f.zneg.value;
    float fac = 1.0f - 2.0f * f.zneg.aux;
    u.at(i, j, -1) = u.at(i, j, 0) * (T)fac;
    v.at(i, j, -1) = v.at(i, j, 0) * (T)fac;
  }
  if (f.zpos.type == OCU_BC_PERIODIC) {
    w.at(i, j, nz) = w.at(i, j, 0);
    w.at(i, j, nz + 1) = w.at(i, j, 1);
    u.at(i, j, nz) = u.at(i, j, 0);
    v.at(i, j, nz) = v.at(i, j, 0);
  } else if (f.zpos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    w.at(i, j, nz) = (T)f.zpos.value;
    w.at(i, j, nz + 1) = (T)f.zpos.value;
    float fac = 1.0f - 2.0f * f.zpos.aux;
    u.at(i, j, nz) = u.at(i, j, nz - 1) * (T)fac;
    v.at(i, j, nz) = v.at(i, j, nz - 1) * (T)fac;
  }
}

int check_types(const ocu_bc_set *bc, const int *ok, int nok, const char *what) {
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  for (auto b : all) {
    bool good = false;
    for (int q = 0; q < nok; q++) good |= b->type == ok[q];
    if (!good) {
      set_error("%s - invalid boundary condition types %d, %d, %d, %d, %d, %d", what, bc->xpos.type,
                bc->xneg.type, bc->ypos.type, bc->yneg.type, bc->zpos.type, bc->zneg.type);
      return OCU_ERR_INVALID;
    }
  }
  return OCU_OK;
}

const int kCellTypes[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_DIRICHELET, OCU_BC_NEUMANN};
const int kMacTypes[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_FORCED_INFLOW_VARIABLE_SLIP};

}  // namespace

extern "C" {

int ocu_bc_fill_level1(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz,
                       void *stream) {
  if (int e = check_grid(g, "apply_3d_boundary_conditions_level1")) return e;
  OCU_REQUIRE(bc, "apply_3d_boundary_conditions_level1: null bc");
  if (int e = check_types(bc, kCellTypes, 4, "apply_3d_boundary_conditions_level1")) return e;
  OCU_REQUIRE(g->gx >= 1 && g->gy >= 1 && g->gz >= 1,
              "apply_3d_boundary_conditions_level1 - must have at least 1 ghost point");
  OCU_REQUIRE(g->dtype != OCU_I32, "apply_3d_boundary_conditions_level1: int grid");
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return OCU_OK;
  int64_t nZ = 2 * (int64_t)(g->nx + 2) * (g->ny + 2);
  int64_t nY = 2 * (int64_t)(g->nx + 2) * g->nz;
  int64_t nX = 2 * (int64_t)g->ny * g->nz;
  int64_t n = nZ + nY + nX;
  int nb = (int)((n + 255) / 256);
  if (g->dtype == OCU_F64) {
    L1<double> s{dgrid<double>(g), faces(bc), hx, hy, hz};
    k_bc_level1<double><<<nb, 256, 0, S(stream)>>>(s, nZ, nY, nX);
  } else {
    L1<float> s{dgrid<float>(g), faces(bc), (float)hx, (float)hy, (float)hz};
    k_bc_level1<float><<<nb, 256, 0, S(stream)>>>(s, nZ, nY, nX);
  }
  OCU_HIP_LAUNCHED("ocu_bc_fill_level1");
}

int ocu_bc_fill_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz,
                          void *stream) {
  if (int e = check_grid(g, "apply_3d_boundary_conditions_level1_nocorners")) return e;
  OCU_REQUIRE(bc, "apply_3d_boundary_conditions_level1_nocorners: null bc");
  if (int e = check_types(bc, kCellTypes, 4, "apply_3d_boundary_conditions_level1_nocorners")) return e;
  OCU_REQUIRE(g->gx >= 1 && g->gy >= 1 && g->gz >= 1,
              "apply_3d_boundary_conditions_level1_nocorners - must have at least 1 ghost point");
  OCU_REQUIRE(g->dtype != OCU_I32, "apply_3d_boundary_conditions_level1_nocorners: int grid");
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return OCU_OK;
  int64_t nZ = (int64_t)g->nx * g->ny, nY = (int64_t)g->nx * g->nz, nX = (int64_t)g->ny * g->nz;
  int nb = (int)((nZ + nY + nX + 255) / 256);
  if (g->dtype == OCU_F64)
    k_bc_nocorners<double><<<nb, 256, 0, S(stream)>>>(dgrid<double>(g), faces(bc), hx, hy, hz, nZ, nY, nX);
  else
    k_bc_nocorners<float><<<nb, 256, 0, S(stream)>>>(dgrid<float>(g), faces(bc), (float)hx, (float)hy,
                                                    (float)hz, nZ, nY, nX);
  OCU_HIP_LAUNCHED("ocu_bc_fill_nocorners");
}

int ocu_bc_fill_mac_level1(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w,
                           const ocu_bc_set *bc, void *stream) {
  const char *what = "apply_3d_mac_boundary_conditions_level1";
  if (int e = check_grid(u, what)) return e;
  if (int e = check_grid(v, what)) return e;
  if (int e = check_grid(w, what)) return e;
  OCU_REQUIRE(bc, "%s: null bc", what);
  if (int e = check_types(bc, kMacTypes, 3, what)) return e;
  OCU_REQUIRE(layout_match(u, v) && layout_match(u, w), "%s - layout mismatch", what);
  OCU_REQUIRE(u->dtype == v->dtype && u->dtype == w->dtype && u->dtype != OCU_I32, "%s - dtype mismatch", what);
  OCU_REQUIRE(u->gx >= 1 && u->gy >= 1 && u->gz >= 1, "%s - must have at least 1 ghost point", what);
  int nx = u->nx - 1, ny = u->ny, nz = u->nz;
  OCU_REQUIRE(v->nx == nx && v->ny == ny + 1 && v->nz == nz, "%s - v_grid dimension mismatch", what);
  OCU_REQUIRE(w->nx == nx && w->ny == ny && w->nz == nz + 1, "%s - w_grid dimension mismatch", what);
  if (nx <= 0 || ny <= 0 || nz <= 0) return OCU_OK;
  Faces f = faces(bc);
  dim3 gx((nz + 63) / 64, ny), gy((nz + 63) / 64, nx + 2), gz((ny + 2 + 63) / 64, nx + 2);
  if (u->dtype == OCU_F64) {
    auto U = dgrid<double>(u), V = dgrid<double>(v), W = dgrid<double>(w);
    k_mac_x<double><<<gx, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac_y<double><<<gy, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac_z<double><<<gz, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
  } else {
    auto U = dgrid<float>(u), V = dgrid<float>(v), W = dgrid<float>(w);
    k_mac_x<float><<<gx, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac_y<float><<<gy, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac_z<float><<<gz, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
  }
  OCU_HIP_LAUNCHED("ocu_bc_fill_mac_level1");
}

}  // extern "C"
