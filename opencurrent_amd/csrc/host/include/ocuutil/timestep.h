// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuutil/timestep.h:26-30
#ifndef __OCU_UTIL_TIMESTEP_H__
#define __OCU_UTIL_TIMESTEP_H__

namespace ocu {

enum TimeStepType {
  TS_ERROR = 0,
  TS_FORWARD_EULER,
  TS_ADAMS_BASHFORD2,
};

}  // namespace ocu

#endif
