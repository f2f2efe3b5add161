// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuutil/interpolation.h:26-31 (the functors themselves live in the kernels)
#ifndef __OCU_UTIL_INTERPOLATION_H__
#define __OCU_UTIL_INTERPOLATION_H__

namespace ocu {

enum InterpolationType {
  IT_ERROR = 0,
  IT_FIRST_ORDER_UPWIND,
  IT_SECOND_ORDER_CENTERED,
  IT_THIRD_ORDER_VAN_LEER_FLUX_LIMITED,
  // no reference counterpart: second-order (linear) upwind, see ocu_types.h
  IT_SECOND_ORDER_UPWIND
};

}  // namespace ocu

#endif
