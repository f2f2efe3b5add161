// Two-layer ghost fills and point sampling (ocustorage/grid3dboundary.cu,
// ocustorage/grid3dsample.cu): the ghost layers second-order schemes and
// particle tracing read.
#include <cmath>

#include "ocu_hip_common.h"

using namespace ocu_hip;

namespace {

struct Face2 {
  int type;
  float value, aux;
};
struct Faces2 {
  Face2 f[6];  // xpos, xneg, ypos, yneg, zpos, zneg
};

Faces2 faces2(const ocu_bc_set *b) {
  Faces2 r;
  const ocu_bc *all[6] = {&b->xpos, &b->xneg, &b->ypos, &b->yneg, &b->zpos, &b->zneg};
  for (int q = 0; q < 6; q++) r.f[q] = Face2{all[q]->type, all[q]->value, all[q]->aux_value};
  return r;
}

// The two ghosts beyond one face from the interior line p (stride s, n
// cells): kernel_apply_3d_boundary_conditions_level2_nocorners
// (grid3dboundary.cu:122-232).  Dirichlet gives both ghosts -phi + 2*value
// from the boundary cell; Neumann phi -/+ h*value and phi -/+ 2*h*value.
template <typename T>
__device__ __forceinline__ void fill2(T *p, int64_t s, int n, const Face2 &neg, const Face2 &pos, T h) {
  if (neg.type == OCU_BC_PERIODIC) {
    p[-1 * s] = p[(int64_t)(n - 1) * s];
    p[-2 * s] = p[(int64_t)(n - 2) * s];
  } else if (neg.type == OCU_BC_DIRICHELET) {
    p[-1 * s] = -p[0] + (T)(2 * neg.value);
    p[-2 * s] = -p[0] + (T)(2 * neg.value);
  } else if (neg.type == OCU_BC_NEUMANN) {
    p[-1 * s] = p[0] - h * (T)neg.value;
    p[-2 * s] = p[0] - (T)2 * h * (T)neg.value;
  }
  if (pos.type == OCU_BC_PERIODIC) {
    p[(int64_t)n * s] = p[0];
    p[(int64_t)(n + 1) * s] = p[s];
  } else if (pos.type == OCU_BC_DIRICHELET) {
    p[(int64_t)n * s] = -p[(int64_t)(n - 1) * s] + (T)(2 * pos.value);
    p[(int64_t)(n + 1) * s] = -p[(int64_t)(n - 1) * s] + (T)(2 * pos.value);
  } else if (pos.type == OCU_BC_NEUMANN) {
    p[(int64_t)n * s] = p[(int64_t)(n - 1) * s] + h * (T)pos.value;
    p[(int64_t)(n + 1) * s] = p[(int64_t)(n - 1) * s] + (T)2 * h * (T)pos.value;
  }
}

// one thread per boundary line: z lines (i,j), then y lines (i,k), then x
// lines (j,k); every ghost is computed from interior cells only
template <typename T>
__global__ void k_bc_level2_nocorners(DGrid<T> g, Faces2 f, T hx, T hy, T hz, int64_t nZ, int64_t nY, int64_t nX) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nZ) {
    const int i = (int)(t / g.ny), j = (int)(t % g.ny);
    fill2(&g.at(i, j, 0), (int64_t)1, g.nz, f.f[5], f.f[4], hz);
    return;
  }
  t -= nZ;
  if (t < nY) {
    const int i = (int)(t / g.nz), k = (int)(t % g.nz);
    fill2(&g.at(i, 0, k), (int64_t)g.ys, g.ny, f.f[3], f.f[2], hy);
    return;
  }
  t -= nY;
  if (t < nX) {
    const int j = (int)(t / g.nz), k = (int)(t % g.nz);
    fill2(&g.at(0, j, k), g.xs, g.nx, f.f[1], f.f[0], hx);
  }
}

// One MAC face pair along an axis of stride s for the normal component N
// and the tangential components A, B at the same line (n cells):
// kernel_apply_3d_mac_boundary_conditions_level2_x/_y/_z
// (grid3dboundary.cu:350-549).  Periodic: the normal wraps two ghosts below
// and three faces above (face n is computed by the caller), the tangentials
// two each way; forced inflow: the normal takes the value on the face and
// its ghosts, the tangentials (1 - 2*aux) times the boundary cell.
template <typename T>
__device__ __forceinline__ void mac2(T *N, T *A, T *B, int64_t s, int n, const Face2 &neg, const Face2 &pos) {
  if (neg.type == OCU_BC_PERIODIC) {
    N[-1 * s] = N[(int64_t)(n - 1) * s];
    N[-2 * s] = N[(int64_t)(n - 2) * s];
    A[-1 * s] = A[(int64_t)(n - 1) * s];
    A[-2 * s] = A[(int64_t)(n - 2) * s];
    B[-1 * s] = B[(int64_t)(n - 1) * s];
    B[-2 * s] = B[(int64_t)(n - 2) * s];
  } else if (neg.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    N[-2 * s] = (T)neg.value;
    N[-1 * s] = (T)neg.value;
    N[0] = (T)neg.value;
    const float fac = 1 - 2 * neg.aux;
    A[-1 * s] = A[0] * (T)fac;
    A[-2 * s] = A[0] * (T)fac;
    B[-1 * s] = B[0] * (T)fac;
    B[-2 * s] = B[0] * (T)fac;
  }
  if (pos.type == OCU_BC_PERIODIC) {
    N[(int64_t)n * s] = N[0];
    N[(int64_t)(n + 1) * s] = N[s];
    N[(int64_t)(n + 2) * s] = N[2 * s];
    A[(int64_t)n * s] = A[0];
    A[(int64_t)(n + 1) * s] = A[s];
    B[(int64_t)n * s] = B[0];
    B[(int64_t)(n + 1) * s] = B[s];
  } else if (pos.type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
    N[(int64_t)n * s] = (T)pos.value;
    N[(int64_t)(n + 1) * s] = (T)pos.value;
    N[(int64_t)(n + 2) * s] = (T)pos.value;
    const float fac = 1 - 2 * pos.aux;
    A[(int64_t)n * s] = A[(int64_t)(n - 1) * s] * (T)fac;
    A[(int64_t)(n + 1) * s] = A[(int64_t)(n - 1) * s] * (T)fac;
    B[(int64_t)n * s] = B[(int64_t)(n - 1) * s] * (T)fac;
    B[(int64_t)(n + 1) * s] = B[(int64_t)(n - 1) * s] * (T)fac;
  }
}

// x pass over (j,k) interior
template <typename T>
__global__ void k_mac2_x(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces2 f, int nx, int ny, int nz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (k >= nz || j >= ny) return;
  mac2(&u.at(0, j, k), &v.at(0, j, k), &w.at(0, j, k), u.xs, nx, f.f[1], f.f[0]);
}
// y pass over i in [-1,nx], k interior
template <typename T>
__global__ void k_mac2_y(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces2 f, int nx, int ny, int nz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, i = (int)blockIdx.y - 1;
  if (k >= nz || i > nx) return;
  mac2(&v.at(i, 0, k), &u.at(i, 0, k), &w.at(i, 0, k), (int64_t)v.ys, ny, f.f[3], f.f[2]);
}
// z pass over (i,j) in [-1,nx]x[-1,ny]
template <typename T>
__global__ void k_mac2_z(DGrid<T> u, DGrid<T> v, DGrid<T> w, Faces2 f, int nx, int ny, int nz) {
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1, i = (int)blockIdx.y - 1;
  if (j > ny || i > nx) return;
  mac2(&w.at(i, j, 0), &u.at(i, j, 0), &v.at(i, j, 0), (int64_t)1, nz, f.f[5], f.f[4]);
}

// ---- point sampling (grid3dsample.cu:24-132) ----

struct SampleAxis {
  int tpos, tneg;
  float vpos, vneg;
  float valid_n, period_n, inv_h, shift;
  int lo, hi;  // the trilinear cell index stays in [lo, hi]
};

// apply_boundary_conditions (grid3dsample.cu:52-88): periodic axes wrap the
// position (fmodf), others clamp it half a cell beyond the last centre and,
// for Neumann, extrapolate the value with the face gradient
__device__ __forceinline__ void sample_bc(float &x, float &value_offset, const SampleAxis &a) {
  if (x < 0) {
    if (a.tneg == OCU_BC_PERIODIC) {
      x = fmodf(x, a.period_n);
      if (x < 0) x += a.period_n;
    } else if (x < -.5f) {
      if (a.tneg == OCU_BC_NEUMANN) value_offset += (-.5f - x) * a.vneg / a.inv_h;
      x = -.5f;
    }
  } else if (x > a.period_n) {
    if (a.tpos == OCU_BC_PERIODIC) {
      x = fmodf(x, a.period_n);
    } else if (x > a.valid_n - .5f) {
      if (a.tpos == OCU_BC_NEUMANN) value_offset += (x - (a.valid_n - .5f)) * a.vpos / a.inv_h;
      x = a.valid_n - .5f;
    }
  }
}

// trilinear_lookup (grid3dsample.cu:24-49): weights and products in T, the
// result rounded to float.  The cell index is kept inside the first ghost
// layer (a position exactly on the upper periodic end, or NaN, would
// otherwise index past it).
template <typename T>
__device__ __forceinline__ float trilinear(const DGrid<T> &g, float x, float y, float z, const SampleAxis *ax) {
  int gi = (int)floorf(x), gj = (int)floorf(y), gk = (int)floorf(z);
  gi = min(max(gi, ax[0].lo), ax[0].hi);
  gj = min(max(gj, ax[1].lo), ax[1].hi);
  gk = min(max(gk, ax[2].lo), ax[2].hi);
  const T sx = x - gi, sy = y - gj, sz = z - gk;
  auto P = [&](int a, int b, int c) { return g.at(gi + a, gj + b, gk + c); };
  const float val = (((P(0, 0, 0)) * (1 - sx) + (P(1, 0, 0)) * (sx)) * (1 - sy) +
                     ((P(0, 1, 0)) * (1 - sx) + (P(1, 1, 0)) * (sx)) * (sy)) * (1 - sz) +
                    (((P(0, 0, 1)) * (1 - sx) + (P(1, 0, 1)) * (sx)) * (1 - sy) +
                     ((P(0, 1, 1)) * (1 - sx) + (P(1, 1, 1)) * (sx)) * (sy)) * (sz);
  return val;
}

struct SampleAxes {
  SampleAxis a[3];
};

// kernel_sample_points_3d (grid3dsample.cu:91-132)
template <typename T>
__global__ void k_sample(T *__restrict__ out, const float *__restrict__ px, const float *__restrict__ py,
                         const float *__restrict__ pz, int n, DGrid<T> g, SampleAxes ax) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  float x = px[idx], y = py[idx], z = pz[idx];
  x *= ax.a[0].inv_h;
  y *= ax.a[1].inv_h;
  z *= ax.a[2].inv_h;
  x -= ax.a[0].shift;
  y -= ax.a[1].shift;
  z -= ax.a[2].shift;
  float value_offset = 0;
  sample_bc(x, value_offset, ax.a[0]);
  sample_bc(y, value_offset, ax.a[1]);
  sample_bc(z, value_offset, ax.a[2]);
  out[idx] = trilinear(g, x, y, z, ax.a) + value_offset;
}

const int kSampleTypes[] = {OCU_BC_PERIODIC, OCU_BC_NEUMANN, OCU_BC_DIRICHELET, OCU_BC_SCALAR_SLIP};

bool types_in(const ocu_bc_set *bc, const int *ok, int nok) {
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  for (auto b : all) {
    bool good = false;
    for (int q = 0; q < nok; q++) good |= b->type == ok[q];
    if (!good) return false;
  }
  return true;
}

SampleAxes sample_axes(const ocu_grid3d *g, const ocu_bc_set *bc, const float period[3], const float h[3],
                       const float shift[3]) {
  SampleAxes s;
  const ocu_bc *pos[3] = {&bc->xpos, &bc->ypos, &bc->zpos}, *neg[3] = {&bc->xneg, &bc->yneg, &bc->zneg};
  const int n[3] = {g->nx, g->ny, g->nz};
  for (int a = 0; a < 3; a++) {
    SampleAxis &x = s.a[a];
    x.tpos = pos[a]->type, x.tneg = neg[a]->type;
    x.vpos = pos[a]->value, x.vneg = neg[a]->value;
    x.valid_n = (float)n[a];
    x.period_n = period[a];
    x.inv_h = 1.0f / h[a];
    x.shift = shift[a];
    x.lo = -1;
    x.hi = n[a] - 1;
  }
  return s;
}

int launch_sample(void *out, const float *px, const float *py, const float *pz, int n, const ocu_grid3d *phi,
                  const ocu_bc_set *bc, const float period[3], const float h[3], const float shift[3],
                  hipStream_t stream) {
  const SampleAxes ax = sample_axes(phi, bc, period, h, shift);
  const int nb = (n + 127) / 128;
  if (phi->dtype == OCU_F64)
    k_sample<double><<<nb, 128, 0, stream>>>((double *)out, px, py, pz, n, dgrid<double>(phi), ax);
  else
    k_sample<float><<<nb, 128, 0, stream>>>((float *)out, px, py, pz, n, dgrid<float>(phi), ax);
  return OCU_OK;
}

int check_sample(const ocu_grid3d *phi, const ocu_bc_set *bc, const float *px, const float *py, const float *pz,
                 void *out, int n, const char *what) {
  if (int e = check_grid(phi, what)) return e;
  OCU_REQUIRE(bc, "%s: null bc", what);
  OCU_REQUIRE(phi->dtype != OCU_I32, "%s: int grid", what);
  OCU_REQUIRE(n >= 0 && (n == 0 || (px && py && pz && out)), "%s - position arrays do not have matching sizes",
              what);
  OCU_REQUIRE(phi->gx >= 1 && phi->gy >= 1 && phi->gz >= 1, "%s - needs a ghost layer", what);
  OCU_REQUIRE(phi->nx >= 1 && phi->ny >= 1 && phi->nz >= 1, "%s - empty grid", what);
  return OCU_OK;
}

}  // namespace

extern "C" {

int ocu_bc_fill_level2_nocorners(const ocu_grid3d *g, const ocu_bc_set *bc, double hx, double hy, double hz,
                                 void *stream) {
  const char *what = "apply_3d_boundary_conditions_level2_nocorners";
  if (int e = check_grid(g, what)) return e;
  OCU_REQUIRE(bc, "%s: null bc", what);
  const int ok[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_DIRICHELET, OCU_BC_NEUMANN};
  OCU_REQUIRE(types_in(bc, ok, 4), "%s - invalid boundary condition types %d, %d, %d, %d, %d,%d", what,
              bc->xpos.type, bc->xneg.type, bc->ypos.type, bc->yneg.type, bc->zpos.type, bc->zneg.type);
  OCU_REQUIRE(g->gx >= 2 && g->gy >= 2 && g->gz >= 2, "%s - must have at least 2 ghost points", what);
  OCU_REQUIRE(g->dtype != OCU_I32, "%s: int grid", what);
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return OCU_OK;
  // a periodic pair copies the last two cells, so each periodic axis needs two
  OCU_REQUIRE((bc->xneg.type != OCU_BC_PERIODIC || g->nx >= 2) && (bc->yneg.type != OCU_BC_PERIODIC || g->ny >= 2) &&
                  (bc->zneg.type != OCU_BC_PERIODIC || g->nz >= 2) &&
                  (bc->xpos.type != OCU_BC_PERIODIC || g->nx >= 2) && (bc->ypos.type != OCU_BC_PERIODIC || g->ny >= 2) &&
                  (bc->zpos.type != OCU_BC_PERIODIC || g->nz >= 2),
              "%s - a periodic axis needs at least two cells", what);
  const int64_t nZ = (int64_t)g->nx * g->ny, nY = (int64_t)g->nx * g->nz, nX = (int64_t)g->ny * g->nz;
  const int nb = (int)((nZ + nY + nX + 255) / 256);
  const Faces2 f = faces2(bc);
  if (g->dtype == OCU_F64)
    k_bc_level2_nocorners<double><<<nb, 256, 0, S(stream)>>>(dgrid<double>(g), f, hx, hy, hz, nZ, nY, nX);
  else
    k_bc_level2_nocorners<float><<<nb, 256, 0, S(stream)>>>(dgrid<float>(g), f, (float)hx, (float)hy, (float)hz, nZ,
                                                           nY, nX);
  OCU_HIP_LAUNCHED(what);
}

int ocu_bc_fill_mac_level2(const ocu_grid3d *u, const ocu_grid3d *v, const ocu_grid3d *w, const ocu_bc_set *bc,
                           void *stream) {
  const char *what = "apply_3d_mac_boundary_conditions_level2";
  if (int e = check_grid(u, what)) return e;
  if (int e = check_grid(v, what)) return e;
  if (int e = check_grid(w, what)) return e;
  OCU_REQUIRE(bc, "%s: null bc", what);
  const int ok[] = {OCU_BC_NONE, OCU_BC_PERIODIC, OCU_BC_FORCED_INFLOW_VARIABLE_SLIP};
  OCU_REQUIRE(types_in(bc, ok, 3), "%s - invalid boundary condition types %d, %d, %d, %d, %d,%d", what,
              bc->xpos.type, bc->xneg.type, bc->ypos.type, bc->yneg.type, bc->zpos.type, bc->zneg.type);
  OCU_REQUIRE(layout_match(u, v) && layout_match(u, w), "%s - layout mismatch", what);
  OCU_REQUIRE(u->dtype == v->dtype && u->dtype == w->dtype && u->dtype != OCU_I32, "%s - dtype mismatch", what);
  OCU_REQUIRE(u->gx >= 2 && u->gy >= 2 && u->gz >= 2, "%s - must have at least 2 ghost points", what);
  // the reference takes nx = u.nx() - u.gz(), which with two ghost layers
  // rejects every congruent MAC triple; the MAC dimension is u.nx() - 1
  const int nx = u->nx - 1, ny = u->ny, nz = u->nz;
  OCU_REQUIRE(v->nx == nx && v->ny == ny + 1 && v->nz == nz, "%s - v_grid dimension mismatch", what);
  OCU_REQUIRE(w->nx == nx && w->ny == ny && w->nz == nz + 1, "%s - w_grid dimension mismatch", what);
  if (nx <= 0 || ny <= 0 || nz <= 0) return OCU_OK;
  OCU_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "%s - periodic copies need two cells per axis", what);
  const Faces2 f = faces2(bc);
  const dim3 gx((nz + 63) / 64, ny), gy((nz + 63) / 64, nx + 2), gz((ny + 2 + 63) / 64, nx + 2);
  if (u->dtype == OCU_F64) {
    auto U = dgrid<double>(u), V = dgrid<double>(v), W = dgrid<double>(w);
    k_mac2_x<double><<<gx, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac2_y<double><<<gy, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac2_z<double><<<gz, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
  } else {
    auto U = dgrid<float>(u), V = dgrid<float>(v), W = dgrid<float>(w);
    k_mac2_x<float><<<gx, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac2_y<float><<<gy, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
    k_mac2_z<float><<<gz, 64, 0, S(stream)>>>(U, V, W, f, nx, ny, nz);
  }
  OCU_HIP_LAUNCHED(what);
}

int ocu_sample_points(void *phi_sampled, const float *position_x, const float *position_y, const float *position_z,
                      int n, const ocu_grid3d *phi, const ocu_bc_set *bc, float period_nx, float period_ny,
                      float period_nz, float hx, float hy, float hz, float shiftx, float shifty, float shiftz,
                      void *stream) {
  const char *what = "sample_points_3d";
  if (int e = check_sample(phi, bc, position_x, position_y, position_z, phi_sampled, n, what)) return e;
  OCU_REQUIRE(types_in(bc, kSampleTypes, 4), "%s - invalid boundary condition types %d, %d, %d, %d, %d,%d", what,
              bc->xpos.type, bc->xneg.type, bc->ypos.type, bc->yneg.type, bc->zpos.type, bc->zneg.type);
  OCU_REQUIRE(!(period_nx > phi->nx || period_ny > phi->ny || period_nz > phi->nz),
              "%s - periods cannot be greater than the array sizes (%f %f %f > %d %d %d)", what, period_nx,
              period_ny, period_nz, phi->nx, phi->ny, phi->nz);
  if (n == 0) return OCU_OK;
  const float period[3] = {period_nx, period_ny, period_nz}, h[3] = {hx, hy, hz}, sh[3] = {shiftx, shifty, shiftz};
  launch_sample(phi_sampled, position_x, position_y, position_z, n, phi, bc, period, h, sh, S(stream));
  OCU_HIP_LAUNCHED(what);
}

// sample_points_mac_grid_3d (grid3dsample.cu:134-263): the MAC faces are
// translated to scalar conditions per component - periodic stays periodic;
// forced inflow becomes Dirichlet for u and scalar slip for v and w (on
// every axis; both clamp the position alike).  The reference decides each negative face by
// the positive face's type (`else if (bc.xpos.type == ...)` on the neg
// branch); that choice is kept, so a face pair that mixes the two kinds
// leaves the negative face undefined and the call fails.
int ocu_sample_points_mac(void *vx, void *vy, void *vz, const float *position_x, const float *position_y,
                          const float *position_z, int n, const ocu_grid3d *u, const ocu_grid3d *v,
                          const ocu_grid3d *w, const ocu_bc_set *bc, float hx, float hy, float hz, void *stream) {
  const char *what = "sample_points_mac_grid_3d";
  OCU_REQUIRE(u && v && w && bc, "%s: null argument", what);
  const int ok[] = {OCU_BC_PERIODIC, OCU_BC_FORCED_INFLOW_VARIABLE_SLIP};
  OCU_REQUIRE(types_in(bc, ok, 2), "%s - invalid boundary condition types %d, %d, %d, %d, %d,%d", what,
              bc->xpos.type, bc->xneg.type, bc->ypos.type, bc->yneg.type, bc->zpos.type, bc->zneg.type);
  const int nx = u->nx - 1, ny = v->ny - 1, nz = w->nz - 1;
  OCU_REQUIRE(!(v->nx != nx || w->nx != nx || u->ny != ny || w->ny != ny || u->nz != nz || v->nz != nz),
              "%s - dimension mismatch between u,v,w", what);
  ocu_bc_set cb[3];  // u, v, w
  for (auto &c : cb) {
    ocu_bc *all[6] = {&c.xpos, &c.xneg, &c.ypos, &c.yneg, &c.zpos, &c.zneg};
    for (auto b : all) *b = ocu_bc{OCU_BC_INVALID, 0.0f, 0.0f};
  }
  const ocu_bc *pos[3] = {&bc->xpos, &bc->ypos, &bc->zpos}, *neg[3] = {&bc->xneg, &bc->yneg, &bc->zneg};
  for (int a = 0; a < 3; a++) {
    ocu_bc *cpos[3] = {a == 0 ? &cb[0].xpos : a == 1 ? &cb[0].ypos : &cb[0].zpos,
                       a == 0 ? &cb[1].xpos : a == 1 ? &cb[1].ypos : &cb[1].zpos,
                       a == 0 ? &cb[2].xpos : a == 1 ? &cb[2].ypos : &cb[2].zpos};
    ocu_bc *cneg[3] = {a == 0 ? &cb[0].xneg : a == 1 ? &cb[0].yneg : &cb[0].zneg,
                       a == 0 ? &cb[1].xneg : a == 1 ? &cb[1].yneg : &cb[1].zneg,
                       a == 0 ? &cb[2].xneg : a == 1 ? &cb[2].yneg : &cb[2].zneg};
    // forced inflow: u Dirichlet, v and w scalar slip, on every axis as in
    // the reference (the two clamp alike)
    if (pos[a]->type == OCU_BC_PERIODIC) {
      for (int c = 0; c < 3; c++) cpos[c]->type = OCU_BC_PERIODIC;
    } else if (pos[a]->type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
      for (int c = 0; c < 3; c++) cpos[c]->type = c == 0 ? OCU_BC_DIRICHELET : OCU_BC_SCALAR_SLIP;
    }
    if (neg[a]->type == OCU_BC_PERIODIC) {
      for (int c = 0; c < 3; c++) cneg[c]->type = OCU_BC_PERIODIC;
    } else if (pos[a]->type == OCU_BC_FORCED_INFLOW_VARIABLE_SLIP) {
      for (int c = 0; c < 3; c++) cneg[c]->type = c == 0 ? OCU_BC_DIRICHELET : OCU_BC_SCALAR_SLIP;
    }
  }
  const float period[3] = {(float)nx, (float)ny, (float)nz}, h[3] = {hx, hy, hz};
  const float shifts[3][3] = {{0, .5f, .5f}, {.5f, 0, .5f}, {.5f, .5f, 0}};
  const ocu_grid3d *g[3] = {u, v, w};
  void *outs[3] = {vx, vy, vz};
  const char *names[3] = {"u", "v", "w"};
  for (int c = 0; c < 3; c++) {
    if (int e = check_sample(g[c], &cb[c], position_x, position_y, position_z, outs[c], n, what)) return e;
    if (!types_in(&cb[c], kSampleTypes, 4)) {
      set_error("%s - failed on %s sampling: invalid boundary condition types", what, names[c]);
      return OCU_ERR_INVALID;
    }
    OCU_REQUIRE(!(period[0] > g[c]->nx || period[1] > g[c]->ny || period[2] > g[c]->nz),
                "%s - failed on %s sampling: periods exceed the array sizes", what, names[c]);
  }
  if (n == 0) return OCU_OK;
  for (int c = 0; c < 3; c++)
    launch_sample(outs[c], position_x, position_y, position_z, n, g[c], &cb[c], period, h, shifts[c], S(stream));
  OCU_HIP_LAUNCHED(what);
}

}  // extern "C"
