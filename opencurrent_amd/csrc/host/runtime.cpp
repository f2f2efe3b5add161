// Host runtime pieces of the drop-in API: the compute stream, status
// reporting, Equation::advance.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "ocu_hip.h"
#include "ocuequation/equation.h"
#include "ocuutil/float_routines.h"
#include "ocuutil/kernel_wrapper.h"
#include "ocuutil/thread.h"

namespace ocu {

namespace {
void *g_compute_stream = nullptr;
bool g_stream_owned = false;
}  // namespace

bool ThreadManager::initialize(int num_images_val) {
  if (num_images_val != 1) {
    printf("[ERROR] ThreadManager::initialize - this build runs one image (GPU) per process, asked for %d\n",
           num_images_val);
    return false;
  }
  return true;
}

bool ThreadManager::initialize_image(int gpu_id) { return ocu_status_ok(ocu_set_device(gpu_id), "hipSetDevice"); }

void *ThreadManager::get_compute_stream() {
  if (!g_compute_stream) {
    void *s = nullptr;
    if (ocu_stream_create(&s) == OCU_OK) {
      g_compute_stream = s;
      g_stream_owned = true;
    }
  }
  return g_compute_stream;
}

void ThreadManager::set_compute_stream(void *stream) {
  if (g_stream_owned && g_compute_stream && g_compute_stream != stream) ocu_stream_destroy(g_compute_stream);
  g_compute_stream = stream;
  g_stream_owned = false;
}

void ThreadManager::compute_fence() { ocu_stream_synchronize(get_compute_stream()); }

unsigned long long next_completion_tag() {
  static std::atomic<unsigned long long> last{0};
  return ++last;  // starts at 1: 0 means no tag
}

// Polling costs a host core while it waits but answers a few microseconds
// sooner than a stream synchronisation; a wait past 200 ms (long work ahead
// of the tag, or a failed kernel) falls back to the synchronisation.
bool wait_completion_tag(const void *slot, unsigned long long tag, void *stream) {
  const volatile unsigned long long *p = static_cast<const volatile unsigned long long *>(slot);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned n = 1;; n++) {
    if (*p == tag) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return true;
    }
    if ((n & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
  }
  if (ocu_stream_synchronize(stream) != OCU_OK) return false;
  std::atomic_thread_fence(std::memory_order_acquire);
  return *p == tag;
}

bool ocu_status_ok(int status, const char *what) {
  if (status == OCU_OK) return true;
  printf("[ERROR] %s - %s\n", what, ocu_last_error());
  return false;
}

Equation::~Equation() {}

// equation.cpp:31-56
bool Equation::advance(double dt) {
  if (!check_float(dt)) {
    printf("[ERROR] Equation::advance - bad dt value\n");
    return false;
  }
  double dtleft = dt;
  while (dtleft > 0) {
    int subds = (int)ceil(dtleft / get_max_stable_timestep());
    double subdt = dtleft / subds;
    if (!advance_one_step(subdt)) {
      printf("[ERROR] Equation::advance - failed to advance equation\n");
      return false;
    }
    dtleft -= subdt;
  }
  return true;
}

}  // namespace ocu
