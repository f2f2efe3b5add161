// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/solver.h: the Solver base.  Launch checking goes through the
// C ABI status codes; PreKernel/PostKernel are kept for source compatibility
// and never synchronise.
#ifndef __OCU_EQUATION_SOLVER_H__
#define __OCU_EQUATION_SOLVER_H__

#include "ocuequation/error_handler.h"
#include "ocuutil/kernel_wrapper.h"

namespace ocu {

class Solver : public ErrorHandler {
 protected:
  void PreKernel() {}
  bool PostKernel(const char *) { return true; }
  // record a C ABI status: prints the library message on failure
  bool ok_status(int status, const char *what) {
    bool ok = ocu_status_ok(status, what);
    check_ok(ok);
    return ok;
  }

 public:
  Solver() {}
  virtual ~Solver() {}
};

}  // namespace ocu

#endif
