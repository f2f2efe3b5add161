// Shared helpers of the gfx950 kernel library: error reporting, argument
// checks, device grid views and the launch-shape helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../../include/ocu_hip.h"

namespace ocu_hip {

void set_error(const char *fmt, ...);
void clear_error();

#define OCU_HIP_TRY(expr)                                                       \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::ocu_hip::set_error("%s failed: %s", #expr, hipGetErrorString(_e));      \
      return OCU_ERR_RUNTIME;                                                   \
    }                                                                           \
  } while (0)

// check the launch that was just enqueued
#define OCU_HIP_LAUNCHED(name)                                                  \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess) {                                                     \
      ::ocu_hip::set_error("%s launch failed: %s", name, hipGetErrorString(_e)); \
      return OCU_ERR_RUNTIME;                                                   \
    }                                                                           \
    return OCU_OK;                                                              \
  } while (0)

#define OCU_REQUIRE(cond, ...)                                                  \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::ocu_hip::set_error(__VA_ARGS__);                                        \
      return OCU_ERR_INVALID;                                                   \
    }                                                                           \
  } while (0)

inline hipStream_t S(void *s) { return (hipStream_t)s; }

// Device view of a grid: pointer to cell (0,0,0) and strides in elements.
template <typename T>
struct DGrid {
  T *p;
  int nx, ny, nz;
  int64_t xs;  // pny*pnz
  int32_t ys;  // pnz
  __device__ __forceinline__ T &at(int i, int j, int k) const {
    return p[(int64_t)i * xs + (int64_t)j * ys + k];
  }
  __device__ __forceinline__ int64_t idx(int i, int j, int k) const {
    return (int64_t)i * xs + (int64_t)j * ys + k;
  }
};

template <typename T>
inline DGrid<T> dgrid(const ocu_grid3d *g) {
  DGrid<T> d;
  d.p = (T *)g->base + g->shift;
  d.nx = g->nx; d.ny = g->ny; d.nz = g->nz;
  d.xs = (int64_t)g->pny * g->pnz;
  d.ys = g->pnz;
  return d;
}

inline bool layout_match(const ocu_grid3d *a, const ocu_grid3d *b) {
  return a->pnx == b->pnx && a->pny == b->pny && a->pnz == b->pnz;
}

// a descriptor is usable: non-null, known dtype, consistent layout numbers
int check_grid(const ocu_grid3d *g, const char *what);

inline bool any_periodic(const ocu_bc_set *bc) {
  return bc->xpos.type == OCU_BC_PERIODIC || bc->xneg.type == OCU_BC_PERIODIC ||
         bc->ypos.type == OCU_BC_PERIODIC || bc->yneg.type == OCU_BC_PERIODIC ||
         bc->zpos.type == OCU_BC_PERIODIC || bc->zneg.type == OCU_BC_PERIODIC;
}

// Interior stencil launch: 64 lanes along k (one wave per k-row segment),
// 4 rows along j, one i-plane per block z index.
constexpr int TK = 64, TJ = 4;
inline dim3 stencil_grid(int nx, int ny, int nz) {
  return dim3((nz + TK - 1) / TK, (ny + TJ - 1) / TJ, nx);
}
inline dim3 stencil_block() { return dim3(TK, TJ, 1); }

// number of per-block partials written by the norm-producing kernels
constexpr int NORM_PARTS = 2048;

// wave64 + block reductions (fixed order => deterministic)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace ocu_hip
