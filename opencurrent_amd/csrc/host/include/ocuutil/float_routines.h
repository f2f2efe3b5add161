// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuutil/float_routines.h: check_float, min3, max3 (max2 added)
#ifndef __OCU_UTIL_FLOAT_ROUTINES_H__
#define __OCU_UTIL_FLOAT_ROUTINES_H__

#include <cmath>

namespace ocu {

inline bool check_float(double f) { return std::isfinite(f); }

template <typename T>
T min3(T a, T b, T c) {
  return a < b ? (a < c ? a : c) : (b < c ? b : c);
}

template <typename T>
T max2(T a, T b) {
  return a > b ? a : b;
}

template <typename T>
T max3(T a, T b, T c) {
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

}  // namespace ocu

#endif
