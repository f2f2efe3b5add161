// Kernels of Sol_PCGPressure3DDevice<T> (ocuequation/sol_pcgpressure3d.cu):
// the variable-coefficient operator -div(c grad p) with face coefficients
// (c_i + c_nb)/2, its diagonal, and dot products of grid interiors.
//
// The operator launch also produces the partial sums of <src, dst>, the
// quantity CG needs right after it (d . Ld), so the reference's separate
// point-wise product grid and reduction (sol_pcgpressure3d.cu:262-293)
// disappear.  Partial sums are FP64 over T products, one per block in a
// fixed row order, folded by one block: deterministic.
#include "ocu_hip_common.h"

using namespace ocu_hip;

namespace {

// rows (i, j) of the interior are dealt round-robin to the waves of the
// NORM_PARTS blocks; lanes stride along k
constexpr int PCG_WAVES = 4;

__device__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < PCG_WAVES; w++) r += sh[w];
  return r;
}

// Sol_PCGPressure3DDevice_apply_laplacian (sol_pcgpressure3d.cu:88-126):
// dst = -(sum_nb f/h^2 * (src_nb - src) * .5 * (c_nb + c)), the f/h^2 being
// T values in double parameters (so the sum is carried in double), plus the
// partials of sum (T)(src * dst) when parts is given
template <typename T>
__global__ __launch_bounds__(64 * PCG_WAVES) void k_pcg_laplacian(DGrid<T> dst, DGrid<T> src, DGrid<T> cf, double fx,
                                                                  double fy, double fz, double *parts) {
  __shared__ double sh[PCG_WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rows = src.nx * src.ny, stride = gridDim.x * PCG_WAVES;
  double acc = 0;
  for (int r = blockIdx.x * PCG_WAVES + w; r < rows; r += stride) {
    const int i = r / src.ny, j = r - i * src.ny;
    const int64_t row = (int64_t)i * src.xs + (int64_t)j * src.ys;
    const T *s = src.p + row, *c = cf.p + row;
    T *d = dst.p + row;
    for (int k = lane; k < src.nz; k += 64) {
      const T sc = s[k], cc = c[k];
      const double v = fz * (s[k - 1] - sc) * (T)0.5 * (c[k - 1] + cc) +
                       fz * (s[k + 1] - sc) * (T)0.5 * (c[k + 1] + cc) +
                       fy * (s[k - src.ys] - sc) * (T)0.5 * (c[k - cf.ys] + cc) +
                       fy * (s[k + src.ys] - sc) * (T)0.5 * (c[k + cf.ys] + cc) +
                       fx * (s[k - src.xs] - sc) * (T)0.5 * (c[k - cf.xs] + cc) +
                       fx * (s[k + src.xs] - sc) * (T)0.5 * (c[k + cf.xs] + cc);
      const T out = (T)(-v);
      d[k] = out;
      acc += (double)(T)(sc * out);
    }
  }
  if (parts) {
    const double b = block_sum(acc, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = b;
  }
}

// Sol_PCGPressure3DDevice_diag_preconditioner (sol_pcgpressure3d.cu:45-83):
// dst = src / diagonal, the diagonal summed in double; dst may equal src
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_diag(DGrid<T> dst, DGrid<T> src, DGrid<T> cf, double fx, double fy,
                                                  double fz) {
  const int k = blockIdx.x * TK + threadIdx.x, j = blockIdx.y * TJ + threadIdx.y, i = blockIdx.z;
  if (j >= src.ny || k >= src.nz) return;
  const int64_t id = src.idx(i, j, k);
  const T *c = cf.p + id;
  const T cc = c[0];
  const double diag = fz * (T)0.5 * (c[-1] + cc) + fz * (T)0.5 * (c[1] + cc) + fy * (T)0.5 * (c[-cf.ys] + cc) +
                      fy * (T)0.5 * (c[cf.ys] + cc) + fx * (T)0.5 * (c[-cf.xs] + cc) +
                      fx * (T)0.5 * (c[cf.xs] + cc);
  dst.p[id] = (T)(src.p[id] / diag);
}

// partials of sum (T)(a * b) over the interior
template <typename T>
__global__ __launch_bounds__(64 * PCG_WAVES) void k_dot(DGrid<T> a, DGrid<T> b, double *parts) {
  __shared__ double sh[PCG_WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rows = a.nx * a.ny, stride = gridDim.x * PCG_WAVES;
  double acc = 0;
  for (int r = blockIdx.x * PCG_WAVES + w; r < rows; r += stride) {
    const int i = r / a.ny, j = r - i * a.ny;
    const T *pa = &a.at(i, j, 0), *pb = &b.at(i, j, 0);
    for (int k = lane; k < a.nz; k += 64) acc += (double)(T)(pa[k] * pb[k]);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

__global__ __launch_bounds__(512) void k_sum_parts(const double *parts, int n, double *out) {
  __shared__ double sh[8];
  double acc = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) acc += parts[q];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
    *out = r;
  }
}

int check_three(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, const char *what) {
  if (int e = check_grid(dst, what)) return e;
  if (int e = check_grid(src, what)) return e;
  if (int e = check_grid(coeff, what)) return e;
  OCU_REQUIRE(layout_match(dst, src) && layout_match(dst, coeff), "%s - layout mismatch", what);
  OCU_REQUIRE(dst->dtype == src->dtype && dst->dtype == coeff->dtype && dst->dtype != OCU_I32, "%s - dtype mismatch",
              what);
  OCU_REQUIRE(src->nx == dst->nx && src->ny == dst->ny && src->nz == dst->nz && coeff->nx == src->nx &&
                  coeff->ny == src->ny && coeff->nz == src->nz,
              "%s - dimension mismatch", what);
  OCU_REQUIRE(src->gx >= 1 && src->gy >= 1 && src->gz >= 1, "%s - needs one ghost layer", what);
  OCU_REQUIRE((int64_t)src->nx * src->ny < INT32_MAX, "%s - too many rows", what);
  return OCU_OK;
}

}  // namespace

extern "C" {

int ocu_pcg_apply_laplacian(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fx_div_hsq,
                            double fy_div_hsq, double fz_div_hsq, void *workspace, double *dot_out, void *stream) {
  const char *what = "Sol_PCGPressure3DDevice::invoke_kernel_apply_laplacian";
  if (int e = check_three(dst, src, coeff, what)) return e;
  OCU_REQUIRE(dst->base != src->base, "%s - dst must not alias src", what);
  OCU_REQUIRE(!dot_out || workspace, "%s - the dot product needs a workspace", what);
  double *parts = dot_out ? (double *)workspace : nullptr;
  if ((int64_t)src->nx * src->ny * src->nz == 0) {
    if (dot_out) OCU_HIP_TRY(hipMemsetAsync(dot_out, 0, sizeof(double), S(stream)));
    return OCU_OK;
  }
  // T values of f/h^2, carried in double as the reference's kernel parameters
  if (src->dtype == OCU_F64)
    k_pcg_laplacian<double><<<NORM_PARTS, 64 * PCG_WAVES, 0, S(stream)>>>(
        dgrid<double>(dst), dgrid<double>(src), dgrid<double>(coeff), fx_div_hsq, fy_div_hsq, fz_div_hsq, parts);
  else
    k_pcg_laplacian<float><<<NORM_PARTS, 64 * PCG_WAVES, 0, S(stream)>>>(
        dgrid<float>(dst), dgrid<float>(src), dgrid<float>(coeff), (double)(float)fx_div_hsq,
        (double)(float)fy_div_hsq, (double)(float)fz_div_hsq, parts);
  if (dot_out) k_sum_parts<<<1, 512, 0, S(stream)>>>(parts, NORM_PARTS, dot_out);
  OCU_HIP_LAUNCHED(what);
}

int ocu_pcg_diag_precond(const ocu_grid3d *dst, const ocu_grid3d *src, const ocu_grid3d *coeff, double fx_div_hsq,
                         double fy_div_hsq, double fz_div_hsq, void *stream) {
  const char *what = "Sol_PCGPressure3DDevice::invoke_kernel_diag_preconditioner";
  if (int e = check_three(dst, src, coeff, what)) return e;
  if ((int64_t)src->nx * src->ny * src->nz == 0) return OCU_OK;
  const dim3 g = stencil_grid(src->nx, src->ny, src->nz);
  if (src->dtype == OCU_F64)
    k_pcg_diag<double><<<g, stencil_block(), 0, S(stream)>>>(dgrid<double>(dst), dgrid<double>(src),
                                                             dgrid<double>(coeff), fx_div_hsq, fy_div_hsq,
                                                             fz_div_hsq);
  else
    k_pcg_diag<float><<<g, stencil_block(), 0, S(stream)>>>(dgrid<float>(dst), dgrid<float>(src), dgrid<float>(coeff),
                                                            (double)(float)fx_div_hsq, (double)(float)fy_div_hsq,
                                                            (double)(float)fz_div_hsq);
  OCU_HIP_LAUNCHED(what);
}

int ocu_grid_dot(const ocu_grid3d *a, const ocu_grid3d *b, void *workspace, double *out, void *stream) {
  const char *what = "ocu_grid_dot";
  if (int e = check_grid(a, what)) return e;
  if (int e = check_grid(b, what)) return e;
  OCU_REQUIRE(workspace && out, "%s: null workspace/out", what);
  OCU_REQUIRE(layout_match(a, b) && a->dtype == b->dtype && a->dtype != OCU_I32, "%s - layout mismatch", what);
  OCU_REQUIRE(a->nx == b->nx && a->ny == b->ny && a->nz == b->nz, "%s - dimension mismatch", what);
  OCU_REQUIRE((int64_t)a->nx * a->ny < INT32_MAX, "%s - too many rows", what);
  double *parts = (double *)workspace;
  if (a->dtype == OCU_F64)
    k_dot<double><<<NORM_PARTS, 64 * PCG_WAVES, 0, S(stream)>>>(dgrid<double>(a), dgrid<double>(b), parts);
  else
    k_dot<float><<<NORM_PARTS, 64 * PCG_WAVES, 0, S(stream)>>>(dgrid<float>(a), dgrid<float>(b), parts);
  k_sum_parts<<<1, 512, 0, S(stream)>>>(parts, NORM_PARTS, out);
  OCU_HIP_LAUNCHED(what);
}

}  // extern "C"
