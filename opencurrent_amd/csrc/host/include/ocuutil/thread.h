// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// ThreadManager for one process per GPU (the reference's co-array images,
// ocuutil/thread.h:33-103, are not part of this build): it owns the compute
// stream that every solver enqueues on.  The stream is created on first use
// with default priority; set_compute_stream() lets a caller (for example a
// torch stream) substitute its own.
#ifndef __OCU_UTIL_THREAD_H__
#define __OCU_UTIL_THREAD_H__

namespace ocu {

class ThreadManager {
 public:
  static int this_image() { return 0; }
  static int num_images() { return 1; }
  static void barrier() {}
  static bool initialize(int num_images_val);
  static bool initialize_image(int gpu_id = 0);
  static void *get_compute_stream();
  static void set_compute_stream(void *stream);
  static void compute_fence();
  static void io_fence() { compute_fence(); }
  static void barrier_and_fence() { compute_fence(); }
};

// Results that a kernel writes into pinned host memory followed by a
// completion tag (the *_tagged entry points of ocu_hip.h): a fresh nonzero
// tag per read, and the wait for it - polling the 8 bytes at `slot` (host
// address), or after a bounded poll synchronising `stream`.  False when the
// stream reports an error or the tag never arrived.
unsigned long long next_completion_tag();
bool wait_completion_tag(const void *slot, unsigned long long tag, void *stream);

}  // namespace ocu

#endif
