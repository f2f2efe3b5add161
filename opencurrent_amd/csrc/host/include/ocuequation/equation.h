// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/equation.h: Equation::advance substeps by the stable step
#ifndef __OCU_EQUATION_EQUATION_H__
#define __OCU_EQUATION_EQUATION_H__

#include "ocuequation/error_handler.h"

namespace ocu {

class Equation : public ErrorHandler {
 public:
  virtual ~Equation();
  virtual double get_max_stable_timestep() const = 0;
  virtual bool advance_one_step(double dt) = 0;
  // equation.cpp:31-56
  bool advance(double dt);
};

}  // namespace ocu

#endif
