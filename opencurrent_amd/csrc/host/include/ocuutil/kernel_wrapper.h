// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Kernel error checking.  The reference's KernelWrapper synchronises the
// device after every launch (ocuutil/kernel_wrapper.cu:94-103); here a launch
// is only checked for an immediate error status (the C ABI's return code)
// and nothing waits, so work stays queued on the compute stream.
#ifndef __OCU_UTIL_KERNEL_WRAPPER_H__
#define __OCU_UTIL_KERNEL_WRAPPER_H__

namespace ocu {

// Print the C ABI's last error and return false when status != 0.
bool ocu_status_ok(int status, const char *what);

}  // namespace ocu

#endif
