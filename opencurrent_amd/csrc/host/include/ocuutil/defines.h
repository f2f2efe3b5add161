// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Host build of the drop-in API: no device qualifiers are needed because no
// device code is compiled against these headers (kernels live behind the C
// ABI of libocu_hip.so).  Mirrors ocuutil/defines.h.
#ifndef __OCU_UTIL_DEFINES_H__
#define __OCU_UTIL_DEFINES_H__

#define OCU_HOST
#define OCU_DEVICE
#define OCU_HOSTDEVICE
#define OCU_MAX_IMAGES 1
#define OCU_DOUBLESUPPORT 1

#endif
