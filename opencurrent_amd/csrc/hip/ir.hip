// Mixed-precision iterative refinement: the FP64 outer residual with a
// scaled cast to FP32, the FP64 correction, and interior sums.  No reference
// counterpart: OpenCurrent's mixed solver (sol_mgpressuremixed3ddevd.cpp:
// 79-100) runs an FP32 FMG and then a full FP64 FMG instead.
#include <algorithm>

#include "ocu_hip_common.h"
#include "rows.h"

using namespace ocu_hip;

namespace {

struct ResC64 {
  double fx, fy, fz, diag;
};

// r = (b - bshift) - A x in double with the reference residual's operand
// order; r32 = (float)(r * scale); partials (sum r^2, max |r|, sum r)
__global__ __launch_bounds__(256) void k_ir_residual(DGrid<double> x, DGrid<double> b, double bshift, DGrid<float> r32,
                                                     double scale, ResC64 c, int tiles_k, int tiles_j, int ntiles,
                                                     double *parts) {
  double s2 = 0, mx = 0, sm = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int tk = t % tiles_k, rem = t / tiles_k, tj = rem % tiles_j, i = rem / tiles_j;
    int k = tk * TK + threadIdx.x;
    int j = tj * TJ + threadIdx.y;
    if (j >= x.ny || k >= x.nz) continue;
    const double *p = x.p;
    int64_t id = x.idx(i, j, k);
    double r = (b.p[b.idx(i, j, k)] - bshift) -
               (c.fz * (p[id - 1] + p[id + 1]) + c.fy * (p[id - x.ys] + p[id + x.ys]) +
                c.fx * (p[id - x.xs] + p[id + x.xs]) - c.diag * p[id]);
    r32.p[r32.idx(i, j, k)] = (float)(r * scale);
    s2 += r * r;
    mx = fmax(mx, fabs(r));
    sm += r;
  }
  __shared__ double sh[3][4];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  if ((tid & 63) == 0) { sh[0][tid >> 6] = s2; sh[1][tid >> 6] = mx; sh[2][tid >> 6] = sm; }
  __syncthreads();
  if (tid == 0) {
    double a = 0, m = 0, s = 0;
    for (int w = 0; w < 4; w++) { a += sh[0][w]; m = fmax(m, sh[1][w]); s += sh[2][w]; }
    parts[3 * blockIdx.x] = a;
    parts[3 * blockIdx.x + 1] = m;
    parts[3 * blockIdx.x + 2] = s;
  }
}

// The same residual with whole rows: each wave owns a row j (V cells per
// lane) and marches along the planes of its i-chunk with planes i-1, i, i+1
// of its row in registers; rows j-1 and j+1 of plane i are read beside them
// (L2 hits: other waves' rows) and k-neighbours come by DPP with the row's z
// ghosts in g.  Exactly NORM_PARTS blocks (idle ones write zero partials);
// tiles are dealt to the XCDs in contiguous runs.
constexpr int IRR_NW = 4;
template <int V>
__global__ __launch_bounds__(64 * IRR_NW) void k_ir_residual_rows(DGrid<double> x, DGrid<double> b, double bshift,
                                                                 DGrid<float> r32, double scale, ResC64 c,
                                                                 int tiles_j, int ichunk, int ntiles,
                                                                 double *parts) {
  typedef Vec<double, V> VT;
  typedef Vec<float, V> VF;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int nz = x.nz;
  const int k0 = lane * V;
  const bool active = k0 < nz;
  const int kc = active ? k0 : 0;
  const int hi_lane = (nz == 64 * V) ? 63 : nz / V;
  const int gk = lane == 0 ? -1 : nz;
  const bool g_in_data = lane == hi_lane && nz < 64 * V;
  double s2 = 0, mx = 0, sm = 0;
  const int per = gridDim.x / 8;
  const int first = (blockIdx.x % 8) * per + blockIdx.x / 8;
  for (int t = first; t < ntiles; t += gridDim.x) {
    const int tj = t % tiles_j, ti = t / tiles_j;
    const int j = tj * IRR_NW + w;
    if (j >= x.ny) continue;  // per wave; the only barrier is after the loop
    const int i0 = ti * ichunk, i1 = min(x.nx, i0 + ichunk);
    const double *xr = x.p + (int64_t)j * x.ys;
    const double *br = b.p + (int64_t)j * b.ys + kc;
    float *rr = r32.p + (int64_t)j * r32.ys + k0;
    auto row = [&](int i, int dj) {
      return *reinterpret_cast<const VT *>(xr + (int64_t)i * x.xs + (int64_t)dj * x.ys + kc);
    };
    auto rowg = [&](int i, VT &d, double &g) {
      const double *p = xr + (int64_t)i * x.xs;
      d = *reinterpret_cast<const VT *>(p + kc);
      g = p[gk];
      if (g_in_data) d.v[0] = g;
    };
    VT prev = row(i0 - 1, 0), cur, nxt;
    double gc, gn;
    rowg(i0, cur, gc);
    for (int i = i0; i < i1; ++i) {
      rowg(i + 1, nxt, gn);
      const VT ym = row(i, -1), yp = row(i, 1);
      const VT bb = *reinterpret_cast<const VT *>(br + (int64_t)i * b.xs);
      const double lo = from_lower(cur.v[V - 1], gc), hi = from_upper(cur.v[0], gc);
      VF o;
#pragma unroll
      for (int e = 0; e < V; e++) {
        const double zm = e > 0 ? cur.v[e - 1] : lo, zp = e < V - 1 ? cur.v[e + 1] : hi;
        const double r = (bb.v[e] - bshift) - (c.fz * (zm + zp) + c.fy * (ym.v[e] + yp.v[e]) +
                                               c.fx * (prev.v[e] + nxt.v[e]) - c.diag * cur.v[e]);
        o.v[e] = (float)(r * scale);
        if (active) {
          s2 += r * r;
          mx = fmax(mx, fabs(r));
          sm += r;
        }
      }
      if (active) *reinterpret_cast<VF *>(rr + (int64_t)i * r32.xs) = o;
      prev = cur;
      cur = nxt;
      gc = gn;
    }
  }
  __shared__ double sh[3][IRR_NW];
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  if (lane == 0) { sh[0][w] = s2; sh[1][w] = mx; sh[2][w] = sm; }
  __syncthreads();
  if (lane == 0 && w == 0) {
    double a = 0, m = 0, q = 0;
    for (int v = 0; v < IRR_NW; v++) { a += sh[0][v]; m = fmax(m, sh[1][v]); q += sh[2][v]; }
    parts[3 * blockIdx.x] = a;
    parts[3 * blockIdx.x + 1] = m;
    parts[3 * blockIdx.x + 2] = q;
  }
}

// interior (sum g^2, max|g|, sum g) partials per block, for ocu_norm3_finalize
__global__ __launch_bounds__(256) void k_stats3(DGrid<double> g, int tiles_k, int tiles_j, int ntiles,
                                                double *parts) {
  double s2 = 0, mx = 0, sm = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int tk = t % tiles_k, rem = t / tiles_k, tj = rem % tiles_j, i = rem / tiles_j;
    int k = tk * TK + threadIdx.x;
    int j = tj * TJ + threadIdx.y;
    if (j >= g.ny || k >= g.nz) continue;
    const double v = g.p[g.idx(i, j, k)];
    s2 += v * v;
    mx = fmax(mx, fabs(v));
    sm += v;
  }
  __shared__ double sh[3][4];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  if ((tid & 63) == 0) { sh[0][tid >> 6] = s2; sh[1][tid >> 6] = mx; sh[2][tid >> 6] = sm; }
  __syncthreads();
  if (tid == 0) {
    double a = 0, m = 0, q = 0;
    for (int w = 0; w < 4; w++) { a += sh[0][w]; m = fmax(m, sh[1][w]); q += sh[2][w]; }
    parts[3 * blockIdx.x] = a;
    parts[3 * blockIdx.x + 1] = m;
    parts[3 * blockIdx.x + 2] = q;
  }
}

__global__ void k_norm3_finalize(const double *parts, int n, double *out) {
  __shared__ double sh[3][8];
  double s = 0, m = 0, q = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    s += parts[3 * t];
    m = fmax(m, parts[3 * t + 1]);
    q += parts[3 * t + 2];
  }
  s = wave_sum(s);
  m = wave_max(m);
  q = wave_sum(q);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s; sh[1][threadIdx.x >> 6] = m; sh[2][threadIdx.x >> 6] = q; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { a += sh[0][w]; b = fmax(b, sh[1][w]); c += sh[2][w]; }
    out[0] = a;
    out[1] = b;
    out[2] = c;
  }
}

__global__ __launch_bounds__(256) void k_ir_correct(DGrid<double> x, DGrid<float> e, double inv_scale) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= x.ny || k >= x.nz) return;
  int64_t id = x.idx(i, j, k);
  x.p[id] = x.p[id] + (double)e.p[e.idx(i, j, k)] * inv_scale;
}

}  // namespace

extern "C" {

int ocu_ir_residual(const ocu_grid3d *x, const ocu_grid3d *b, double bshift, const ocu_grid3d *r32, double scale,
                    double h, double fx, double fy, double fz, double *ws, void *stream) {
  if (int e = check_grid(x, "ocu_ir_residual(x)")) return e;
  if (int e = check_grid(b, "ocu_ir_residual(b)")) return e;
  if (int e = check_grid(r32, "ocu_ir_residual(r32)")) return e;
  OCU_REQUIRE(ws, "ocu_ir_residual: null workspace");
  OCU_REQUIRE(x->dtype == OCU_F64 && b->dtype == OCU_F64 && r32->dtype == OCU_F32,
              "ocu_ir_residual: x, b must be FP64 and r32 FP32");
  OCU_REQUIRE(x->nx == b->nx && x->ny == b->ny && x->nz == b->nz && x->nx == r32->nx && x->ny == r32->ny &&
                  x->nz == r32->nz,
              "ocu_ir_residual: dimension mismatch");
  OCU_REQUIRE(x->gx >= 1 && x->gy >= 1 && x->gz >= 1, "ocu_ir_residual: x needs ghosts");
  ResC64 c{fx / (h * h), fy / (h * h), fz / (h * h), 2 * (fx + fy + fz) / (h * h)};
  const int nz = x->nz;
  const int V = nz <= 64 ? 1 : nz <= 128 ? 2 : nz <= 256 ? 4 : 0;
  if (V && nz % V == 0 && row_aligned(x, V) && row_aligned(b, V) && row_aligned(r32, V) &&
      (int64_t)x->nx * x->ny * nz > 0) {
    const int tiles_j = (x->ny + IRR_NW - 1) / IRR_NW;
    // one tile per block where the grid allows
    const int ichunk = std::max(1, (int)(((int64_t)x->nx * tiles_j + NORM_PARTS - 1) / NORM_PARTS));
    const int ntiles = tiles_j * ((x->nx + ichunk - 1) / ichunk);
#define IRR(VV)                                                                                                      \
  k_ir_residual_rows<VV><<<NORM_PARTS, dim3(64, IRR_NW), 0, S(stream)>>>(dgrid<double>(x), dgrid<double>(b), bshift, \
                                                                        dgrid<float>(r32), scale, c, tiles_j,      \
                                                                        ichunk, ntiles, ws)
    if (V == 1) IRR(1);
    else if (V == 2) IRR(2);
    else IRR(4);
#undef IRR
    OCU_HIP_LAUNCHED("ocu_ir_residual");
  }
  int tiles_k = (x->nz + TK - 1) / TK, tiles_j = (x->ny + TJ - 1) / TJ;
  int ntiles = tiles_k * tiles_j * x->nx;
  k_ir_residual<<<NORM_PARTS, stencil_block(), 0, S(stream)>>>(dgrid<double>(x), dgrid<double>(b), bshift,
                                                                dgrid<float>(r32), scale, c, tiles_k, tiles_j, ntiles, ws);
  OCU_HIP_LAUNCHED("ocu_ir_residual");
}

int ocu_grid_stats3(const ocu_grid3d *g, double *ws, void *stream) {
  if (int e = check_grid(g, "ocu_grid_stats3")) return e;
  OCU_REQUIRE(g->dtype == OCU_F64, "ocu_grid_stats3: FP64 grids only");
  OCU_REQUIRE(ws, "ocu_grid_stats3: null workspace");
  const int tiles_k = (g->nz + TK - 1) / TK, tiles_j = (g->ny + TJ - 1) / TJ;
  const int ntiles = tiles_k * tiles_j * g->nx;
  k_stats3<<<NORM_PARTS, stencil_block(), 0, S(stream)>>>(dgrid<double>(g), tiles_k, tiles_j, ntiles, ws);
  OCU_HIP_LAUNCHED("ocu_grid_stats3");
}

int ocu_norm3_finalize(const double *ws, double *out, void *stream) {
  OCU_REQUIRE(ws && out, "ocu_norm3_finalize: null pointer");
  k_norm3_finalize<<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, out);
  OCU_HIP_LAUNCHED("ocu_norm3_finalize");
}

int ocu_ir_correct(const ocu_grid3d *x, const ocu_grid3d *e, double inv_scale, void *stream) {
  if (int er = check_grid(x, "ocu_ir_correct(x)")) return er;
  if (int er = check_grid(e, "ocu_ir_correct(e)")) return er;
  OCU_REQUIRE(x->dtype == OCU_F64 && e->dtype == OCU_F32, "ocu_ir_correct: x must be FP64 and e FP32");
  OCU_REQUIRE(x->nx == e->nx && x->ny == e->ny && x->nz == e->nz, "ocu_ir_correct: dimension mismatch");
  if ((int64_t)x->nx * x->ny * x->nz == 0) return OCU_OK;
  k_ir_correct<<<stencil_grid(x->nx, x->ny, x->nz), stencil_block(), 0, S(stream)>>>(dgrid<double>(x),
                                                                                     dgrid<float>(e), inv_scale);
  OCU_HIP_LAUNCHED("ocu_ir_correct");
}

}  // extern "C"
