// Runtime entry points: errors, memory, streams, events and graphs.
// Replaces the CUDA runtime plumbing of ocuutil/memory.cu, thread.cu and
// kernel_wrapper.cu; unlike KernelWrapper::PostKernel (kernel_wrapper.cu:
// 94-103) nothing here synchronises the device after a launch.
#include "ocu_hip_common.h"

namespace ocu_hip {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

int check_grid(const ocu_grid3d *g, const char *what) {
  if (!g) { set_error("%s: null grid descriptor", what); return OCU_ERR_INVALID; }
  if (!g->base) { set_error("%s: grid has no storage", what); return OCU_ERR_INVALID; }
  if (g->dtype != OCU_F32 && g->dtype != OCU_F64 && g->dtype != OCU_I32) {
    set_error("%s: unknown dtype %d", what, g->dtype);
    return OCU_ERR_INVALID;
  }
  if (g->nx < 0 || g->ny < 0 || g->nz < 0 || g->gx < 0 || g->gy < 0 || g->gz < 0 ||
      g->pnx < g->nx + 2 * g->gx || g->pny < g->ny + 2 * g->gy || g->pnz < g->nz + 2 * g->gz) {
    set_error("%s: inconsistent layout (%d,%d,%d) ghosts (%d,%d,%d) padded (%d,%d,%d)", what, g->nx,
              g->ny, g->nz, g->gx, g->gy, g->gz, g->pnx, g->pny, g->pnz);
    return OCU_ERR_INVALID;
  }
  int64_t need = (int64_t)g->pnx * g->pny * g->pnz;
  if (g->shift < (int64_t)g->gx * g->pny * g->pnz + (int64_t)g->gy * g->pnz + g->gz ||
      g->allocated < need) {
    set_error("%s: shift/allocation too small", what);
    return OCU_ERR_INVALID;
  }
  return OCU_OK;
}

}  // namespace ocu_hip

using namespace ocu_hip;

extern "C" {

const char *ocu_last_error(void) { return g_last_error.c_str(); }

int ocu_get_device_count(int *count) {
  OCU_REQUIRE(count, "ocu_get_device_count: null pointer");
  OCU_HIP_TRY(hipGetDeviceCount(count));
  return OCU_OK;
}

int ocu_set_device(int device) {
  OCU_HIP_TRY(hipSetDevice(device));
  return OCU_OK;
}

int ocu_device_synchronize(void) {
  OCU_HIP_TRY(hipDeviceSynchronize());
  return OCU_OK;
}

int ocu_malloc(void **ptr, size_t bytes) {
  OCU_REQUIRE(ptr, "ocu_malloc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return OCU_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    set_error("ocu_malloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return OCU_ERR_NOMEM;
  }
  return OCU_OK;
}

int ocu_free(void *ptr) {
  if (!ptr) return OCU_OK;
  OCU_HIP_TRY(hipFree(ptr));
  return OCU_OK;
}

int ocu_host_malloc(void **ptr, size_t bytes) {
  OCU_REQUIRE(ptr, "ocu_host_malloc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return OCU_OK;
  hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocMapped);
  if (e != hipSuccess) {
    set_error("ocu_host_malloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return OCU_ERR_NOMEM;
  }
  return OCU_OK;
}

int ocu_host_device_pointer(void **dev, void *host) {
  OCU_REQUIRE(dev && host, "ocu_host_device_pointer: null pointer");
  OCU_HIP_TRY(hipHostGetDevicePointer(dev, host, 0));
  return OCU_OK;
}

int ocu_host_free(void *ptr) {
  if (!ptr) return OCU_OK;
  OCU_HIP_TRY(hipHostFree(ptr));
  return OCU_OK;
}

int ocu_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
  if (bytes == 0) return OCU_OK;
  OCU_REQUIRE(dst && src, "ocu_memcpy_async: null pointer");
  hipMemcpyKind k;
  switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    case 3: k = hipMemcpyHostToHost; break;
    default: set_error("ocu_memcpy_async: bad kind %d", kind); return OCU_ERR_INVALID;
  }
  OCU_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, S(stream)));
  return OCU_OK;
}

int ocu_stream_create(void **stream) {
  OCU_REQUIRE(stream, "ocu_stream_create: null pointer");
  hipStream_t s;
  // default priority on purpose; a blocking stream, so work on the legacy
  // default stream (where callers such as torch stage their copies) stays
  // ordered with the solver's kernels
  OCU_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamDefault));
  *stream = s;
  return OCU_OK;
}

int ocu_stream_destroy(void *stream) {
  if (!stream) return OCU_OK;
  OCU_HIP_TRY(hipStreamDestroy(S(stream)));
  return OCU_OK;
}

int ocu_stream_synchronize(void *stream) {
  OCU_HIP_TRY(hipStreamSynchronize(S(stream)));
  return OCU_OK;
}

int ocu_event_create(void **event) {
  OCU_REQUIRE(event, "ocu_event_create: null pointer");
  hipEvent_t e;
  OCU_HIP_TRY(hipEventCreate(&e));
  *event = e;
  return OCU_OK;
}

int ocu_event_destroy(void *event) {
  if (!event) return OCU_OK;
  OCU_HIP_TRY(hipEventDestroy((hipEvent_t)event));
  return OCU_OK;
}

int ocu_event_record(void *event, void *stream) {
  OCU_HIP_TRY(hipEventRecord((hipEvent_t)event, S(stream)));
  return OCU_OK;
}

int ocu_event_synchronize(void *event) {
  OCU_HIP_TRY(hipEventSynchronize((hipEvent_t)event));
  return OCU_OK;
}

int ocu_event_elapsed_ms(float *ms, void *start, void *stop) {
  OCU_REQUIRE(ms, "ocu_event_elapsed_ms: null pointer");
  OCU_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return OCU_OK;
}

int ocu_stream_begin_capture(void *stream) {
  OCU_REQUIRE(stream, "ocu_stream_begin_capture: capture needs a non-default stream");
  OCU_HIP_TRY(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal));
  return OCU_OK;
}

int ocu_stream_end_capture(void *stream, void **graph_exec) {
  OCU_REQUIRE(graph_exec, "ocu_stream_end_capture: null pointer");
  hipGraph_t g;
  OCU_HIP_TRY(hipStreamEndCapture(S(stream), &g));
  hipGraphExec_t ge;
  hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);  // the executable graph no longer needs it; instantiation's status is what counts
  if (e != hipSuccess) {
    set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return OCU_ERR_RUNTIME;
  }
  *graph_exec = ge;
  return OCU_OK;
}

int ocu_graph_launch(void *graph_exec, void *stream) {
  OCU_HIP_TRY(hipGraphLaunch((hipGraphExec_t)graph_exec, S(stream)));
  return OCU_OK;
}

int ocu_graph_destroy(void *graph_exec) {
  if (!graph_exec) return OCU_OK;
  OCU_HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return OCU_OK;
}

int ocu_norm_parts(void) { return NORM_PARTS; }

}  // extern "C"
