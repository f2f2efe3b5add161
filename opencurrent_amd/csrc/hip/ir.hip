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
// order; r32 = (float)(r * scale); partials (sum r^2, max |r|, sum r, sum x)
__global__ __launch_bounds__(256) void k_ir_residual(DGrid<double> x, DGrid<double> b, double bshift, DGrid<float> r32,
                                                     double scale, ResC64 c, int tiles_k, int tiles_j, int ntiles,
                                                     double *parts) {
  double s2 = 0, mx = 0, sm = 0, sx = 0;
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
    sx += p[id];
  }
  __shared__ double sh[4][4];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  sx = wave_sum(sx);
  if ((tid & 63) == 0) { sh[0][tid >> 6] = s2; sh[1][tid >> 6] = mx; sh[2][tid >> 6] = sm; sh[3][tid >> 6] = sx; }
  __syncthreads();
  if (tid == 0) {
    double a = 0, m = 0, s = 0, q = 0;
    for (int w = 0; w < 4; w++) { a += sh[0][w]; m = fmax(m, sh[1][w]); s += sh[2][w]; q += sh[3][w]; }
    parts[4 * blockIdx.x] = a;
    parts[4 * blockIdx.x + 1] = m;
    parts[4 * blockIdx.x + 2] = s;
    parts[4 * blockIdx.x + 3] = q;
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
  double s2 = 0, mx = 0, sm = 0, sx = 0;
  // tiles dealt over the XCDs in contiguous runs: the first nb blocks (the
  // tiles rounded up to the XCDs, at most the grid) take them
  const int nb = min((int)gridDim.x, ((ntiles + 7) / 8) * 8);
  const int per = nb / 8;
  const int first = blockIdx.x < nb ? (blockIdx.x % 8) * per + blockIdx.x / 8 : ntiles;
  for (int t = first; t < ntiles; t += nb) {
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
        // selects, not branches: a branch here splits the loop body and
        // costs the loads their overlap
        const double ra = active ? r : 0.0, xa = active ? cur.v[e] : 0.0;
        s2 += ra * ra;
        mx = fmax(mx, fabs(ra));
        sm += ra;
        sx += xa;
      }
      if (active) *reinterpret_cast<VF *>(rr + (int64_t)i * r32.xs) = o;
      prev = cur;
      cur = nxt;
      gc = gn;
    }
  }
  __shared__ double sh[4][IRR_NW];
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  sm = wave_sum(sm);
  sx = wave_sum(sx);
  if (lane == 0) { sh[0][w] = s2; sh[1][w] = mx; sh[2][w] = sm; sh[3][w] = sx; }
  __syncthreads();
  if (lane == 0 && w == 0) {
    double a = 0, m = 0, q = 0, t = 0;
    for (int v = 0; v < IRR_NW; v++) { a += sh[0][v]; m = fmax(m, sh[1][v]); q += sh[2][v]; t += sh[3][v]; }
    parts[4 * blockIdx.x] = a;
    parts[4 * blockIdx.x + 1] = m;
    parts[4 * blockIdx.x + 2] = q;
    parts[4 * blockIdx.x + 3] = t;
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

// Reduces NW-tuples of per-block partials: slot 1 by max, the others by sum
// (sum r^2, max |r|, sum r [, sum x]).  512 threads, fixed order.  tag != 0:
// after the results, with a system-scope fence between, the tag goes into
// out[NW] (host memory that the host polls).
template <int NW>
__global__ void k_norm_finalize(const double *parts, int n, double *out, unsigned long long tag) {
  __shared__ double sh[NW][8];
  double a[NW];
#pragma unroll
  for (int q = 0; q < NW; q++) a[q] = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x)
#pragma unroll
    for (int q = 0; q < NW; q++) a[q] = q == 1 ? fmax(a[q], parts[NW * t + q]) : a[q] + parts[NW * t + q];
#pragma unroll
  for (int q = 0; q < NW; q++) a[q] = q == 1 ? wave_max(a[q]) : wave_sum(a[q]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < NW; q++) sh[q][threadIdx.x >> 6] = a[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NW; q++) {
      double r = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = q == 1 ? fmax(r, sh[q][w]) : r + sh[q][w];
      out[q] = r;
    }
    if (tag) {
      __threadfence_system();
      *reinterpret_cast<volatile unsigned long long *>(out + NW) = tag;
    }
  }
}

// Interior statistics by whole rows: a wave per row (V cells per lane), two
// rows in flight per wave, 4 waves per block, exactly NORM_PARTS blocks.
// NS = 3: (sum g^2, max |g|, sum g) per block; NS = 1: sum g per block.
// Accumulation in double.
constexpr int RS_NW = 4;
template <typename T, int V, int NS>
__global__ __launch_bounds__(64 * RS_NW) void k_row_stats(DGrid<T> g, double *parts) {
  typedef Vec<T, V> VT;
  const int lane = threadIdx.x, w = threadIdx.y;
  const int k0 = lane * V;
  const bool active = k0 < g.nz;
  const int kc = active ? k0 : 0;
  const int64_t rows = (int64_t)g.nx * g.ny;
  const int64_t stride = (int64_t)gridDim.x * RS_NW;
  double s2 = 0, mx = 0, sm = 0;
  auto load = [&](int64_t row) {
    const int i = (int)(row / g.ny), j = (int)(row - (int64_t)i * g.ny);
    return *reinterpret_cast<const VT *>(g.p + (int64_t)i * g.xs + (int64_t)j * g.ys + kc);
  };
  auto add = [&](const VT &v) {
    if (!active) return;
#pragma unroll
    for (int e = 0; e < V; e++) {
      const double d = (double)v.v[e];
      if (NS == 3) {
        s2 += d * d;
        mx = fmax(mx, fabs(d));
      }
      sm += d;
    }
  };
  int64_t row = (int64_t)blockIdx.x * RS_NW + w;
  for (; row + stride < rows; row += 2 * stride) {
    const VT a = load(row), b = load(row + stride);
    add(a);
    add(b);
  }
  if (row < rows) add(load(row));
  __shared__ double sh[3][RS_NW];
  sm = wave_sum(sm);
  if (NS == 3) {
    s2 = wave_sum(s2);
    mx = wave_max(mx);
  }
  if (lane == 0) { sh[0][w] = s2; sh[1][w] = mx; sh[2][w] = sm; }
  __syncthreads();
  if (lane == 0 && w == 0) {
    double a = 0, m = 0, q = 0;
    for (int v = 0; v < RS_NW; v++) { a += sh[0][v]; m = fmax(m, sh[1][v]); q += sh[2][v]; }
    if (NS == 3) {
      parts[3 * blockIdx.x] = a;
      parts[3 * blockIdx.x + 1] = m;
      parts[3 * blockIdx.x + 2] = q;
    } else {
      parts[blockIdx.x] = q;
    }
  }
}

// the tiled form of the sum for grids whose rows do not fit k_row_stats
template <typename T>
__global__ __launch_bounds__(256) void k_sum_tiled(DGrid<T> g, int tiles_k, int tiles_j, int ntiles, double *parts) {
  double sm = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int tk = t % tiles_k, rem = t / tiles_k, tj = rem % tiles_j, i = rem / tiles_j;
    int k = tk * TK + threadIdx.x;
    int j = tj * TJ + threadIdx.y;
    if (j >= g.ny || k >= g.nz) continue;
    sm += (double)g.p[g.idx(i, j, k)];
  }
  __shared__ double sh[4];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  sm = wave_sum(sm);
  if ((tid & 63) == 0) sh[tid >> 6] = sm;
  __syncthreads();
  if (tid == 0) parts[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// shift = (xsum + inv_scale * sum(parts)) / ncells: the mean of x + e*inv_scale
__global__ void k_ir_shift_finalize(const double *parts, int n, double xsum, double inv_scale, double ncells,
                                    double *out) {
  __shared__ double sh[8];
  double a = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) a += parts[t];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double esum = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) esum += sh[w];
    out[0] = (xsum + inv_scale * esum) / ncells;
  }
}

// The correction by whole rows: a wave per row (V cells per lane).  With
// FACES the rows run over i in [-1, nx], j in [-1, ny] less the four edge
// rows, and an interior row's lanes 0 and 63 also update its z ghosts
// (k = -1 and k = nz), loaded beside the row so they cost no extra round
// trip.  shift may be null (0).
template <int V, bool FACES>
__global__ __launch_bounds__(256) void k_ir_correct_rows(DGrid<double> x, DGrid<float> e, double inv_scale,
                                                         const double *shift, int rows) {
  typedef Vec<double, V> VD;
  typedef Vec<float, V> VF;
  const int lane = threadIdx.x;
  const int r = blockIdx.x * 4 + threadIdx.y;
  if (r >= rows) return;
  const int o = FACES ? 1 : 0, wj = x.ny + 2 * o;
  const int i = r / wj - o, j = r % wj - o;
  const bool gi = i < 0 || i >= x.nx, gj = j < 0 || j >= x.ny;
  if (gi && gj) return;
  const double sh = shift ? *shift : 0.0;
  const int k0 = lane * V;
  const bool active = k0 < x.nz;
  const int kc = active ? k0 : 0;
  double *xr = x.p + (int64_t)i * x.xs + (int64_t)j * x.ys;
  const float *er = e.p + (int64_t)i * e.xs + (int64_t)j * e.ys;
  const VD xv = *reinterpret_cast<const VD *>(xr + kc);
  const VF ev = *reinterpret_cast<const VF *>(er + kc);
  const bool zg = FACES && !gi && !gj && (lane == 0 || lane == 63);
  const int gk = lane == 0 ? -1 : x.nz;
  double xg = 0;
  float eg = 0;
  if (zg) {
    xg = xr[gk];
    eg = er[gk];
  }
  VD out;
#pragma unroll
  for (int q = 0; q < V; q++) {
    const double v = xv.v[q] + (double)ev.v[q] * inv_scale;
    out.v[q] = shift ? v - sh : v;
  }
  if (active) *reinterpret_cast<VD *>(xr + k0) = out;
  if (zg) {
    const double v = xg + (double)eg * inv_scale;
    xr[gk] = shift ? v - sh : v;
  }
}

// x = (x + (double)e * inv_scale) - shift over the interior, and with FACES
// over the six face-ghost layers too: the thread of a boundary cell also
// updates the ghost cell beside it across each face it touches (so edges
// and corners are left alone).  shift may be null (0).
template <bool FACES>
__global__ __launch_bounds__(256) void k_ir_correct(DGrid<double> x, DGrid<float> e, double inv_scale,
                                                    const double *shift) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= x.ny || k >= x.nz) return;
  const double sh = shift ? *shift : 0.0;
  auto upd = [&](int a, int b, int c) {
    const int64_t id = x.idx(a, b, c);
    const double v = x.p[id] + (double)e.p[e.idx(a, b, c)] * inv_scale;
    x.p[id] = shift ? v - sh : v;
  };
  upd(i, j, k);
  if (FACES) {
    if (k == 0) upd(i, j, -1);
    if (k == x.nz - 1) upd(i, j, x.nz);
    if (j == 0) upd(i, -1, k);
    if (j == x.ny - 1) upd(i, x.ny, k);
    if (i == 0) upd(-1, j, k);
    if (i == x.nx - 1) upd(x.nx, j, k);
  }
}

// the per-lane vector width of a whole-row kernel over g (0: rows do not fit)
int row_vec(const ocu_grid3d *g) {
  const int nz = g->nz;
  const int V = nz <= 64 ? 1 : nz <= 128 ? 2 : nz <= 256 ? 4 : 0;
  if (!V || nz % V != 0 || !row_aligned(g, V) || (int64_t)g->nx * g->ny * nz == 0) return 0;
  return V;
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
  const int V = row_vec(g);
  if (V) {
    if (V == 1) k_row_stats<double, 1, 3><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<double>(g), ws);
    else if (V == 2) k_row_stats<double, 2, 3><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<double>(g), ws);
    else k_row_stats<double, 4, 3><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<double>(g), ws);
    OCU_HIP_LAUNCHED("ocu_grid_stats3");
  }
  const int tiles_k = (g->nz + TK - 1) / TK, tiles_j = (g->ny + TJ - 1) / TJ;
  const int ntiles = tiles_k * tiles_j * g->nx;
  k_stats3<<<NORM_PARTS, stencil_block(), 0, S(stream)>>>(dgrid<double>(g), tiles_k, tiles_j, ntiles, ws);
  OCU_HIP_LAUNCHED("ocu_grid_stats3");
}

int ocu_norm3_finalize(const double *ws, double *out, void *stream) {
  OCU_REQUIRE(ws && out, "ocu_norm3_finalize: null pointer");
  k_norm_finalize<3><<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, out, 0ull);
  OCU_HIP_LAUNCHED("ocu_norm3_finalize");
}

int ocu_norm4_finalize(const double *ws, double *out, void *stream) {
  OCU_REQUIRE(ws && out, "ocu_norm4_finalize: null pointer");
  k_norm_finalize<4><<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, out, 0ull);
  OCU_HIP_LAUNCHED("ocu_norm4_finalize");
}

int ocu_norm3_finalize_tagged(const double *ws, double *out, unsigned long long tag, void *stream) {
  OCU_REQUIRE(ws && out && tag, "ocu_norm3_finalize_tagged: null pointer or zero tag");
  k_norm_finalize<3><<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, out, tag);
  OCU_HIP_LAUNCHED("ocu_norm3_finalize_tagged");
}

int ocu_norm4_finalize_tagged(const double *ws, double *out, unsigned long long tag, void *stream) {
  OCU_REQUIRE(ws && out && tag, "ocu_norm4_finalize_tagged: null pointer or zero tag");
  k_norm_finalize<4><<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, out, tag);
  OCU_HIP_LAUNCHED("ocu_norm4_finalize_tagged");
}

int ocu_ir_shift(const ocu_grid3d *e, double inv_scale, double xsum, double *ws, double *shift, void *stream) {
  if (int er = check_grid(e, "ocu_ir_shift(e)")) return er;
  OCU_REQUIRE(e->dtype == OCU_F32, "ocu_ir_shift: e must be FP32");
  OCU_REQUIRE(ws && shift, "ocu_ir_shift: null pointer");
  const double ncells = (double)e->nx * e->ny * e->nz;
  OCU_REQUIRE(ncells > 0, "ocu_ir_shift: empty grid");
  const int V = row_vec(e);
  if (V == 1) k_row_stats<float, 1, 1><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<float>(e), ws);
  else if (V == 2) k_row_stats<float, 2, 1><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<float>(e), ws);
  else if (V == 4) k_row_stats<float, 4, 1><<<NORM_PARTS, dim3(64, RS_NW), 0, S(stream)>>>(dgrid<float>(e), ws);
  else {
    const int tiles_k = (e->nz + TK - 1) / TK, tiles_j = (e->ny + TJ - 1) / TJ;
    k_sum_tiled<float><<<NORM_PARTS, stencil_block(), 0, S(stream)>>>(dgrid<float>(e), tiles_k, tiles_j,
                                                                        tiles_k * tiles_j * e->nx, ws);
  }
  k_ir_shift_finalize<<<1, 512, 0, S(stream)>>>(ws, NORM_PARTS, xsum, inv_scale, ncells, shift);
  OCU_HIP_LAUNCHED("ocu_ir_shift");
}

int ocu_ir_correct(const ocu_grid3d *x, const ocu_grid3d *e, double inv_scale, void *stream) {
  return ocu_ir_correct_ex(x, e, inv_scale, nullptr, 0, stream);
}

int ocu_ir_correct_ex(const ocu_grid3d *x, const ocu_grid3d *e, double inv_scale, const double *shift, int faces,
                      void *stream) {
  if (int er = check_grid(x, "ocu_ir_correct(x)")) return er;
  if (int er = check_grid(e, "ocu_ir_correct(e)")) return er;
  OCU_REQUIRE(x->dtype == OCU_F64 && e->dtype == OCU_F32, "ocu_ir_correct: x must be FP64 and e FP32");
  OCU_REQUIRE(x->nx == e->nx && x->ny == e->ny && x->nz == e->nz, "ocu_ir_correct: dimension mismatch");
  OCU_REQUIRE(!faces || (x->gx >= 1 && x->gy >= 1 && x->gz >= 1 && e->gx >= 1 && e->gy >= 1 && e->gz >= 1),
              "ocu_ir_correct: face mode needs a ghost layer on x and e");
  if ((int64_t)x->nx * x->ny * x->nz == 0) return OCU_OK;
  const int V = row_vec(x);
  if (V && V == row_vec(e)) {
    const int rows = faces ? (x->nx + 2) * (x->ny + 2) : x->nx * x->ny;
    const int nb = (rows + 3) / 4;
#define IRC(VV, FF)                                                                                            \
  k_ir_correct_rows<VV, FF><<<nb, dim3(64, 4), 0, S(stream)>>>(dgrid<double>(x), dgrid<float>(e), inv_scale, \
                                                              shift, rows)
    if (faces) {
      if (V == 1) IRC(1, true);
      else if (V == 2) IRC(2, true);
      else IRC(4, true);
    } else {
      if (V == 1) IRC(1, false);
      else if (V == 2) IRC(2, false);
      else IRC(4, false);
    }
#undef IRC
    OCU_HIP_LAUNCHED("ocu_ir_correct");
  }
  if (faces)
    k_ir_correct<true><<<stencil_grid(x->nx, x->ny, x->nz), stencil_block(), 0, S(stream)>>>(
        dgrid<double>(x), dgrid<float>(e), inv_scale, shift);
  else
    k_ir_correct<false><<<stencil_grid(x->nx, x->ny, x->nz), stencil_block(), 0, S(stream)>>>(
        dgrid<double>(x), dgrid<float>(e), inv_scale, shift);
  OCU_HIP_LAUNCHED("ocu_ir_correct");
}

}  // extern "C"
