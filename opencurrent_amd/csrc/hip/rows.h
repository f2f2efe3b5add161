// Whole-row building blocks of the row kernels: a wave holds a k-row as V
// consecutive cells per lane in an aligned vector, and a cell's k-neighbours
// in the lanes on either side come by DPP row shifts.
#pragma once

#include "ocu_hip_common.h"

namespace ocu_hip {

template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
  T v[V];
};

// DPP wave_shr:1 / wave_shl:1; a lane whose source is off the wave keeps old
__device__ __forceinline__ int dpp_lower_i(int src, int old) {
  return __builtin_amdgcn_update_dpp(old, src, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_upper_i(int src, int old) {
  return __builtin_amdgcn_update_dpp(old, src, 0x130, 0xf, 0xf, false);
}
// value of lane l-1 (lane 0: old)
__device__ __forceinline__ float from_lower(float src, float old) {
  return __int_as_float(dpp_lower_i(__float_as_int(src), __float_as_int(old)));
}
__device__ __forceinline__ double from_lower(double src, double old) {
  long long s = __double_as_longlong(src), o = __double_as_longlong(old);
  int lo = dpp_lower_i((int)s, (int)o);
  int hi = dpp_lower_i((int)(s >> 32), (int)(o >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// value of lane l+1 (lane 63: old)
__device__ __forceinline__ float from_upper(float src, float old) {
  return __int_as_float(dpp_upper_i(__float_as_int(src), __float_as_int(old)));
}
__device__ __forceinline__ double from_upper(double src, double old) {
  long long s = __double_as_longlong(src), o = __double_as_longlong(old);
  int lo = dpp_upper_i((int)s, (int)o);
  int hi = dpp_upper_i((int)(s >> 32), (int)(o >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// rows start on vector boundaries of v elements of g's type in every plane
inline bool row_aligned(const ocu_grid3d *g, int v) {
  const size_t w = g->dtype == OCU_F64 ? 8 : 4;
  return g->pnz % v == 0 && ((int64_t)g->pny * g->pnz) % v == 0 &&
         ((uintptr_t)g->base + (uintptr_t)g->shift * w) % (w * v) == 0;
}

}  // namespace ocu_hip
