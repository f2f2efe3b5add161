// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/parameters.h
#ifndef __OCU_EQUATION_PARAMETERS_H__
#define __OCU_EQUATION_PARAMETERS_H__

namespace ocu {
class Parameters {};
}  // namespace ocu

#endif
