// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ocuequation/error_handler.h:27-59: bool returns plus a sticky error flag
#ifndef __OCU_EQUATION_ERROR_HANDLER_H__
#define __OCU_EQUATION_ERROR_HANDLER_H__

#include <cstdio>

namespace ocu {

class ErrorHandler {
  bool _any_error;

 protected:
  void check_ok(bool ok) { _any_error = _any_error || !ok; }
  void check_ok(bool ok, const char *msg) {
    _any_error = _any_error || !ok;
    if (!ok) printf("[WARNING] Failure: %s\n", msg);
  }
  void check_ok(bool ok, const char *filename, int lineno) {
    _any_error = _any_error || !ok;
    if (!ok) printf("[WARNING] Failure in file \"%s\" at line %d\n", filename, lineno);
  }
  void add_error() { _any_error = true; }

 public:
  ErrorHandler() : _any_error(false) {}
  void clear_error() { _any_error = false; }
  bool any_error() const { return _any_error; }
};

#define CHECK_OK(X) check_ok((X), __FILE__, __LINE__)

}  // namespace ocu

#endif
