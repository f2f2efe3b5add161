// Mixed-precision iterative refinement: the FP64 outer residual with a
// scaled cast to FP32, the FP64 correction, and interior sums.  No reference
// counterpart: OpenCurrent's mixed solver (sol_mgpressuremixed3ddevd.cpp:
// 79-100) runs an FP32 FMG and then a full FP64 FMG instead.
#include "ocu_hip_common.h"

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
  int tiles_k = (x->nz + TK - 1) / TK, tiles_j = (x->ny + TJ - 1) / TJ;
  int ntiles = tiles_k * tiles_j * x->nx;
  ResC64 c{fx / (h * h), fy / (h * h), fz / (h * h), 2 * (fx + fy + fz) / (h * h)};
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
