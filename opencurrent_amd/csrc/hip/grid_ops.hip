// Whole-grid operations of Grid3DDevice<T> (ocustorage/grid3ddev.cu) and the
// interior reductions (ocustorage/grid3dops.cu, ocuutil/reduction_op.h).
#include "ocu_hip_common.h"

#include <climits>
#include <mutex>

using namespace ocu_hip;

namespace {

// ---- element-wise over the allocation (vectorised grid-stride loops) ----

template <typename T>
__global__ void k_fill(T *__restrict__ p, T v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T, typename S>
__global__ void k_convert(T *__restrict__ to, const S *__restrict__ from, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) to[i] = (T)from[i];
}

// grid3ddev.cu:57-75: result = g1*a1 (+ g2*a2) on every allocated element
template <typename T>
__global__ void k_lincomb_flat(T *__restrict__ r, T a1, const T *__restrict__ g1, T a2,
                               const T *__restrict__ g2, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (g2) {
    for (; i < n; i += stride) r[i] = g1[i] * a1 + g2[i] * a2;
  } else {
    for (; i < n; i += stride) r[i] = g1[i] * a1;
  }
}

// grid3ddev.cu:35-55: interior only
template <typename T>
__global__ void k_lincomb_interior(DGrid<T> r, T a1, DGrid<T> g1, T a2, DGrid<T> g2) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= r.ny || k >= r.nz) return;
  int64_t id = r.idx(i, j, k);
  r.p[id] = g1.p[id] * a1 + g2.p[id] * a2;
}

// zero the first ghost layer: one thread per (i,j) column of the ghost box;
// ghost columns are zeroed whole, interior columns at k = -1 and k = nz
template <typename T>
__global__ void k_clear_ghosts(DGrid<T> g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (g.nx + 2) * (g.ny + 2)) return;
  const int i = t / (g.ny + 2) - 1, j = t % (g.ny + 2) - 1;
  if (i < 0 || i >= g.nx || j < 0 || j >= g.ny) {
    for (int k = -1; k <= g.nz; k++) g.at(i, j, k) = (T)0;
  } else {
    g.at(i, j, -1) = (T)0;
    g.at(i, j, g.nz) = (T)0;
  }
}

template <typename T>
__global__ void k_add_scalar(DGrid<T> g, T v) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= g.ny || k >= g.nz) return;
  int64_t id = g.idx(i, j, k);
  g.p[id] = g.p[id] + v;
}

inline int flat_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

// ---- reductions: stage 1 writes NORM_PARTS partials, stage 2 folds them ----

// op: 0 maxabs, 1 max, 2 min, 3 sum, 4 sqrsum, 5 checknan
__device__ __forceinline__ double red_identity(int op) {
  return op == 1 ? -HUGE_VAL : op == 2 ? HUGE_VAL : 0.0;
}
__device__ __forceinline__ double red_process(int op, double x) {
  return op == 0 ? fabs(x) : op == 4 ? x * x : x;
}
__device__ __forceinline__ double red_combine(int op, double a, double b) {
  switch (op) {
    case 0:
    case 1: return fmax(a, b);
    case 2: return fmin(a, b);
    case 5: return isfinite(a) ? b : a;  // ReduceDevCheckNan (reduction_op.h:101-106)
    default: return a + b;
  }
}

__device__ double block_reduce(int op, double v, double *sh) {
  // wave level, fixed butterfly order
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = red_combine(op, v, __shfl_xor(v, off, 64));
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = red_identity(op);
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; w++) r = red_combine(op, r, sh[w]);
  }
  __syncthreads();
  return r;
}

// rows (i,j) are distributed round-robin over the NORM_PARTS blocks; each
// block walks its rows with 256 threads striding along k.
template <typename T>
__global__ void k_reduce_stage1(DGrid<T> g, int op, double *parts) {
  __shared__ double sh[8];
  double acc = red_identity(op);
  int64_t rows = (int64_t)g.nx * g.ny;
  for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
    int i = (int)(row / g.ny), j = (int)(row % g.ny);
    const T *p = &g.at(i, j, 0);
    for (int k = threadIdx.x; k < g.nz; k += blockDim.x)
      acc = red_combine(op, acc, red_process(op, (double)p[k]));
  }
  double r = block_reduce(op, acc, sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = r;
}

// maxabs, max and min (OP 0..2) give the same value in any order, so they
// walk the grid for bandwidth: each wave takes whole rows (rows dealt
// round-robin over the waves), every lane keeps four loads of a row in
// flight; one partial per block as in k_reduce_stage1
template <typename T, int OP>
__device__ __forceinline__ double max_rows_acc(const DGrid<T> &g, double acc) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * 4, w0 = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int rows = g.nx * g.ny, nz = g.nz;
  for (int r = w0; r < rows; r += nw) {
    const int i = r / g.ny, j = r - i * g.ny;
    const T *p = &g.at(i, j, 0);
    int k = lane;
    for (; k + 192 < nz; k += 256) {
      const double a = (double)p[k], b = (double)p[k + 64], c = (double)p[k + 128], d = (double)p[k + 192];
      acc = red_combine(OP, red_combine(OP, acc, red_process(OP, a)), red_process(OP, b));
      acc = red_combine(OP, red_combine(OP, acc, red_process(OP, c)), red_process(OP, d));
    }
    for (; k < nz; k += 64) acc = red_combine(OP, acc, red_process(OP, (double)p[k]));
  }
  return acc;
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_reduce_max_rows(DGrid<T> g, double *parts) {
  __shared__ double sh[8];
  double r = block_reduce(OP, max_rows_acc<T, OP>(g, red_identity(OP)), sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = r;
}

// max |a|, max |b|, max |c| of three grids in one pass: block b's partials
// go to parts[b], parts[NORM_PARTS + b], parts[2 * NORM_PARTS + b]
template <typename T>
__global__ __launch_bounds__(256) void k_maxabs3_rows(DGrid<T> a, DGrid<T> b, DGrid<T> c, double *parts) {
  __shared__ double sh[8];
  const double ra = block_reduce(0, max_rows_acc<T, 0>(a, 0.0), sh);
  const double rb = block_reduce(0, max_rows_acc<T, 0>(b, 0.0), sh);
  const double rc = block_reduce(0, max_rows_acc<T, 0>(c, 0.0), sh);
  if (threadIdx.x == 0) {
    parts[blockIdx.x] = ra;
    parts[NORM_PARTS + blockIdx.x] = rb;
    parts[2 * NORM_PARTS + blockIdx.x] = rc;
  }
}

// the three maxima into out[0..2]; then, with a system-scope fence
// between, the tag into out[3] (host memory that the host polls)
__global__ void k_maxabs3_stage2(const double *parts, double *out, unsigned long long tag) {
  __shared__ double sh[8];
  double r[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    double acc = 0.0;
    for (int t = threadIdx.x; t < NORM_PARTS; t += blockDim.x) acc = red_combine(0, acc, parts[q * NORM_PARTS + t]);
    r[q] = block_reduce(0, acc, sh);
  }
  if (threadIdx.x == 0) {
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
    __threadfence_system();
    *reinterpret_cast<volatile unsigned long long *>(out + 3) = tag;
  }
}

// tag != 0: after the result, with a system-scope fence between, the tag
// goes into out[1] (host memory that the host polls)
__global__ void k_reduce_stage2(int op, const double *parts, int n, double *out, unsigned long long tag) {
  __shared__ double sh[8];
  double acc = red_identity(op);
  for (int q = threadIdx.x; q < n; q += blockDim.x) acc = red_combine(op, acc, parts[q]);
  double r = block_reduce(op, acc, sh);
  if (threadIdx.x == 0) {
    *out = r;
    if (tag) {
      __threadfence_system();
      *reinterpret_cast<volatile unsigned long long *>(out + 1) = tag;
    }
  }
}

}  // namespace

// The partials of ocu_grid_reduce_noalloc: NORM_PARTS doubles (16 KiB) per
// GPU, held until the process ends.  On the host this name is only a
// stand-in for the symbol; kernels get the device address that
// scratch_on_current_device() looked up.  Every call on one GPU reuses the
// same array, so calls that overlap on that GPU (two streams, two host
// threads) race on it.
static __device__ __align__(256) double scratch[NORM_PARTS];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[SCRATCH_MAX_DEVICES];

// device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy)
int scratch_on_current_device(double **out) {
  int dev = 0;
  OCU_HIP_TRY(hipGetDevice(&dev));
  OCU_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "no-allocation scratch: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
      return OCU_ERR_RUNTIME;
    }
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return OCU_OK;
}

// the two launches of an interior reduction: NORM_PARTS partials into
// `parts`, folded into *out (device memory)
int check_reduce(const ocu_grid3d *g, int op, const double *out, const char *what) {
  if (int e = check_grid(g, what)) return e;
  OCU_REQUIRE(op >= 0 && op <= 5, "%s: bad op %d", what, op);
  OCU_REQUIRE(out, "%s: null out", what);
  OCU_REQUIRE(g->dtype != OCU_I32, "%s: int grids are not supported", what);
  OCU_REQUIRE((int64_t)g->nx * g->ny < INT_MAX, "%s: too many rows", what);
  return OCU_OK;
}

int launch_reduce(const ocu_grid3d *g, int op, double *parts, double *out, void *stream, const char *what,
                  unsigned long long tag = 0) {
#define RMAX(TT, OP) k_reduce_max_rows<TT, OP><<<NORM_PARTS, 256, 0, S(stream)>>>(dgrid<TT>(g), parts)
  if (op <= 2 && g->dtype == OCU_F64) {
    if (op == 0) RMAX(double, 0);
    else if (op == 1) RMAX(double, 1);
    else RMAX(double, 2);
  } else if (op <= 2) {
    if (op == 0) RMAX(float, 0);
    else if (op == 1) RMAX(float, 1);
    else RMAX(float, 2);
  } else if (g->dtype == OCU_F64) {
    k_reduce_stage1<double><<<NORM_PARTS, 256, 0, S(stream)>>>(dgrid<double>(g), op, parts);
  } else {
    k_reduce_stage1<float><<<NORM_PARTS, 256, 0, S(stream)>>>(dgrid<float>(g), op, parts);
  }
#undef RMAX
  k_reduce_stage2<<<1, 512, 0, S(stream)>>>(op, parts, NORM_PARTS, out, tag);
  OCU_HIP_LAUNCHED(what);
}

}  // namespace

extern "C" {

int ocu_grid_clear(const ocu_grid3d *g, double value, void *stream) {
  if (int e = check_grid(g, "ocu_grid_clear")) return e;
  int nb = flat_blocks(g->allocated);
  if (g->dtype == OCU_F64) k_fill<double><<<nb, 256, 0, S(stream)>>>((double *)g->base, value, g->allocated);
  else if (g->dtype == OCU_F32) k_fill<float><<<nb, 256, 0, S(stream)>>>((float *)g->base, (float)value, g->allocated);
  else k_fill<int><<<nb, 256, 0, S(stream)>>>((int *)g->base, (int)value, g->allocated);
  OCU_HIP_LAUNCHED("ocu_grid_clear");
}

int ocu_grid_clear_ghosts(const ocu_grid3d *g, void *stream) {
  if (int e = check_grid(g, "ocu_grid_clear_ghosts")) return e;
  OCU_REQUIRE(g->gx >= 1 && g->gy >= 1 && g->gz >= 1, "ocu_grid_clear_ghosts: needs one ghost layer");
  OCU_REQUIRE(g->dtype != OCU_I32, "ocu_grid_clear_ghosts: int grid");
  const int n = (g->nx + 2) * (g->ny + 2), nb = (n + 255) / 256;
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return OCU_OK;
  if (g->dtype == OCU_F64) k_clear_ghosts<double><<<nb, 256, 0, S(stream)>>>(dgrid<double>(g));
  else k_clear_ghosts<float><<<nb, 256, 0, S(stream)>>>(dgrid<float>(g));
  OCU_HIP_LAUNCHED("ocu_grid_clear_ghosts");
}

int ocu_grid_copy(const ocu_grid3d *dst, const ocu_grid3d *src, void *stream) {
  if (int e = check_grid(dst, "ocu_grid_copy(dst)")) return e;
  if (int e = check_grid(src, "ocu_grid_copy(src)")) return e;
  OCU_REQUIRE(layout_match(dst, src), "ocu_grid_copy: layout mismatch (%d,%d,%d) != (%d,%d,%d)",
              dst->pnx, dst->pny, dst->pnz, src->pnx, src->pny, src->pnz);
  int64_t n = dst->allocated < src->allocated ? dst->allocated : src->allocated;
  if (dst->dtype == src->dtype) {
    size_t w = dst->dtype == OCU_F64 ? 8 : 4;
    OCU_HIP_TRY(hipMemcpyAsync(dst->base, src->base, n * w, hipMemcpyDeviceToDevice, S(stream)));
    return OCU_OK;
  }
  int nb = flat_blocks(n);
  if (dst->dtype == OCU_F64 && src->dtype == OCU_F32)
    k_convert<double, float><<<nb, 256, 0, S(stream)>>>((double *)dst->base, (const float *)src->base, n);
  else if (dst->dtype == OCU_F32 && src->dtype == OCU_F64)
    k_convert<float, double><<<nb, 256, 0, S(stream)>>>((float *)dst->base, (const double *)src->base, n);
  else if (dst->dtype == OCU_F32 && src->dtype == OCU_I32)
    k_convert<float, int><<<nb, 256, 0, S(stream)>>>((float *)dst->base, (const int *)src->base, n);
  else if (dst->dtype == OCU_F64 && src->dtype == OCU_I32)
    k_convert<double, int><<<nb, 256, 0, S(stream)>>>((double *)dst->base, (const int *)src->base, n);
  else if (dst->dtype == OCU_I32 && src->dtype == OCU_F32)
    k_convert<int, float><<<nb, 256, 0, S(stream)>>>((int *)dst->base, (const float *)src->base, n);
  else
    k_convert<int, double><<<nb, 256, 0, S(stream)>>>((int *)dst->base, (const double *)src->base, n);
  OCU_HIP_LAUNCHED("ocu_grid_copy");
}

int ocu_grid_lincomb(const ocu_grid3d *r, double a1, const ocu_grid3d *g1, double a2,
                     const ocu_grid3d *g2, void *stream) {
  if (int e = check_grid(r, "ocu_grid_lincomb(result)")) return e;
  if (int e = check_grid(g1, "ocu_grid_lincomb(g1)")) return e;
  OCU_REQUIRE(layout_match(r, g1), "ocu_grid_lincomb: g1 layout mismatch");
  OCU_REQUIRE(r->dtype == g1->dtype && r->dtype != OCU_I32, "ocu_grid_lincomb: dtype mismatch");
  if (g2) {
    if (int e = check_grid(g2, "ocu_grid_lincomb(g2)")) return e;
    OCU_REQUIRE(layout_match(r, g2), "ocu_grid_lincomb: g2 layout mismatch");
    OCU_REQUIRE(r->dtype == g2->dtype, "ocu_grid_lincomb: dtype mismatch");
    int64_t interior = (int64_t)r->nx * r->ny * r->nz;
    // the reference's rule (grid3ddev.cu:316): > 3% waste -> interior only
    if ((r->allocated - interior) / (float)r->allocated > .03f) {
      if (interior == 0) return OCU_OK;
      dim3 grid = stencil_grid(r->nx, r->ny, r->nz);
      if (r->dtype == OCU_F64)
        k_lincomb_interior<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(r), a1, dgrid<double>(g1), a2, dgrid<double>(g2));
      else
        k_lincomb_interior<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(r), (float)a1, dgrid<float>(g1), (float)a2, dgrid<float>(g2));
      OCU_HIP_LAUNCHED("ocu_grid_lincomb");
    }
  }
  int nb = flat_blocks(r->allocated);
  if (r->dtype == OCU_F64)
    k_lincomb_flat<double><<<nb, 256, 0, S(stream)>>>((double *)r->base, a1, (const double *)g1->base, a2,
                                                      g2 ? (const double *)g2->base : nullptr, r->allocated);
  else
    k_lincomb_flat<float><<<nb, 256, 0, S(stream)>>>((float *)r->base, (float)a1, (const float *)g1->base, (float)a2,
                                                     g2 ? (const float *)g2->base : nullptr, r->allocated);
  OCU_HIP_LAUNCHED("ocu_grid_lincomb");
}

int ocu_grid_add_scalar(const ocu_grid3d *g, double value, void *stream) {
  if (int e = check_grid(g, "ocu_grid_add_scalar")) return e;
  if ((int64_t)g->nx * g->ny * g->nz == 0) return OCU_OK;
  dim3 grid = stencil_grid(g->nx, g->ny, g->nz);
  if (g->dtype == OCU_F64) k_add_scalar<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(g), value);
  else k_add_scalar<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(g), (float)value);
  OCU_HIP_LAUNCHED("ocu_grid_add_scalar");
}

size_t ocu_reduce_workspace_bytes(void) { return sizeof(double) * NORM_PARTS; }

int ocu_grid_reduce(const ocu_grid3d *g, int op, void *workspace, double *out, void *stream) {
  if (int e = check_reduce(g, op, out, "ocu_grid_reduce")) return e;
  OCU_REQUIRE(workspace, "ocu_grid_reduce: null workspace");
  return launch_reduce(g, op, (double *)workspace, out, stream, "ocu_grid_reduce");
}

int ocu_grid_reduce_tagged(const ocu_grid3d *g, int op, void *workspace, double *out, unsigned long long tag,
                           void *stream) {
  if (int e = check_reduce(g, op, out, "ocu_grid_reduce_tagged")) return e;
  OCU_REQUIRE(workspace, "ocu_grid_reduce_tagged: null workspace");
  OCU_REQUIRE(tag, "ocu_grid_reduce_tagged: the tag must not be 0");
  return launch_reduce(g, op, (double *)workspace, out, stream, "ocu_grid_reduce_tagged", tag);
}

int ocu_grid_maxabs3_tagged(const ocu_grid3d *a, const ocu_grid3d *b, const ocu_grid3d *c, void *workspace,
                            double *out, unsigned long long tag, void *stream) {
  const char *what = "ocu_grid_maxabs3_tagged";
  for (const ocu_grid3d *g : {a, b, c})
    if (int e = check_reduce(g, 0, out, what)) return e;
  OCU_REQUIRE(a->dtype == b->dtype && a->dtype == c->dtype, "%s: the grids' types differ", what);
  OCU_REQUIRE(workspace, "%s: null workspace", what);
  OCU_REQUIRE(tag, "%s: the tag must not be 0", what);
  double *parts = (double *)workspace;
  if (a->dtype == OCU_F64)
    k_maxabs3_rows<double><<<NORM_PARTS, 256, 0, S(stream)>>>(dgrid<double>(a), dgrid<double>(b), dgrid<double>(c),
                                                              parts);
  else
    k_maxabs3_rows<float><<<NORM_PARTS, 256, 0, S(stream)>>>(dgrid<float>(a), dgrid<float>(b), dgrid<float>(c),
                                                             parts);
  k_maxabs3_stage2<<<1, 512, 0, S(stream)>>>(parts, out, tag);
  OCU_HIP_LAUNCHED(what);
}

int ocu_grid_reduce_noalloc(const ocu_grid3d *g, int op, double *out, void *stream) {
  if (int e = check_reduce(g, op, out, "ocu_grid_reduce_noalloc")) return e;
  double *parts = nullptr;
  if (int e = scratch_on_current_device(&parts)) return e;
  return launch_reduce(g, op, parts, out, stream, "ocu_grid_reduce_noalloc");
}

int ocu_grid_sum(const ocu_grid3d *g, double *ws, double *out, void *stream) {
  return ocu_grid_reduce(g, 3, ws, out, stream);
}

}  // extern "C"
