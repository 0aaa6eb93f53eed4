// runtime.hip -- the runtime subset of era_cudart / cudart-sys that the NTT path rides on
// (crates/cudart/src/{device,memory,stream,event,error,graph}.rs), as direct HIP calls.
// No CUDA shims: these are HIP entry points with their error codes passed through as int.
#include "api_internal.h"

#define BA_RET(expr) return (int)(expr)

extern "C" {

int ba_get_device_count(int *count) { BA_RET(hipGetDeviceCount(count)); }
int ba_set_device(int device) { BA_RET(hipSetDevice(device)); }
int ba_get_device(int *device) { BA_RET(hipGetDevice(device)); }
int ba_device_synchronize(void) { BA_RET(hipDeviceSynchronize()); }
int ba_mem_get_info(size_t *free_bytes, size_t *total_bytes) { BA_RET(hipMemGetInfo(free_bytes, total_bytes)); }
int ba_device_get_attribute(int *value, int attribute, int device) {
  BA_RET(hipDeviceGetAttribute(value, (hipDeviceAttribute_t)attribute, device));
}
int ba_malloc(void **ptr, size_t bytes) { BA_RET(hipMalloc(ptr, bytes)); }
int ba_free(void *ptr) { BA_RET(hipFree(ptr)); }
int ba_malloc_host(void **ptr, size_t bytes) { BA_RET(hipHostMalloc(ptr, bytes, hipHostMallocDefault)); }
int ba_free_host(void *ptr) { BA_RET(hipHostFree(ptr)); }
int ba_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  if (kind < 0 || kind > 4) return BA_ERROR_INVALID_VALUE;
  BA_RET(hipMemcpy(dst, src, bytes, (hipMemcpyKind)kind));
}
int ba_memcpy_async(void *dst, const void *src, size_t bytes, int kind, ba_stream_t stream) {
  if (kind < 0 || kind > 4) return BA_ERROR_INVALID_VALUE;
  BA_RET(hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, (hipStream_t)stream));
}
int ba_memset_async(void *dst, int value, size_t bytes, ba_stream_t stream) {
  BA_RET(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
}
int ba_stream_create(ba_stream_t *stream) {
  BA_RET(hipStreamCreateWithFlags((hipStream_t *)stream, hipStreamNonBlocking));
}
int ba_stream_destroy(ba_stream_t stream) { BA_RET(hipStreamDestroy((hipStream_t)stream)); }
int ba_stream_synchronize(ba_stream_t stream) { BA_RET(hipStreamSynchronize((hipStream_t)stream)); }
int ba_stream_wait_event(ba_stream_t stream, ba_event_t event) {
  BA_RET(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
}
int ba_event_create(ba_event_t *event) { BA_RET(hipEventCreate((hipEvent_t *)event)); }
int ba_event_destroy(ba_event_t event) { BA_RET(hipEventDestroy((hipEvent_t)event)); }
int ba_event_record(ba_event_t event, ba_stream_t stream) {
  BA_RET(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
}
int ba_event_synchronize(ba_event_t event) { BA_RET(hipEventSynchronize((hipEvent_t)event)); }
int ba_event_query(ba_event_t event) { BA_RET(hipEventQuery((hipEvent_t)event)); }
int ba_event_elapsed_ms(float *ms, ba_event_t start, ba_event_t end) {
  BA_RET(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
}
int ba_get_last_error(void) { BA_RET(hipGetLastError()); }
const char *ba_error_name(int status) { return hipGetErrorName((hipError_t)status); }

int ba_stream_begin_capture(ba_stream_t stream) {
  BA_RET(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
}
int ba_stream_end_capture(ba_stream_t stream, ba_graph_exec_t *exec) {
  hipGraph_t graph = nullptr;
  hipError_t err = hipStreamEndCapture((hipStream_t)stream, &graph);
  if (err != hipSuccess) return (int)err;
  hipGraphExec_t e = nullptr;
  err = hipGraphInstantiate(&e, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) return (int)err;
  *exec = (ba_graph_exec_t)e;
  return BA_SUCCESS;
}
int ba_graph_launch(ba_graph_exec_t exec, ba_stream_t stream) {
  BA_RET(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
}
int ba_graph_destroy(ba_graph_exec_t exec) { BA_RET(hipGraphExecDestroy((hipGraphExec_t)exec)); }

}  // extern "C"

// ---------------------------------------------------------------- launch tracing
// The reference's profiling hooks are nvtx ranges (crates/shivini/Cargo.toml:31) and CUDA-event
// timing in criterion-cuda (crates/criterion-cuda/src/lib.rs:16-34).  Here a host thread can ask
// the library to bracket each kernel it launches with two caller-owned events, so a harness
// measures per-kernel durations on the launch stream without a profiler.
#include <vector>

namespace {
struct TraceState {
  hipEvent_t *events = nullptr;
  uint32_t capacity = 0;
  uint32_t count = 0;
  bool open = false;  // a launch has started but not finished
  std::vector<const char *> names;
};
thread_local TraceState g_trace;
}  // namespace

void ba_trace_before(hipStream_t stream, const char *kernel) {
  TraceState &t = g_trace;
  if (!t.events || t.count >= t.capacity) return;
  (void)hipEventRecord(t.events[2 * t.count], stream);
  t.names.push_back(kernel);
  t.open = true;
}

bool ba_trace_active() { return g_trace.events != nullptr; }

void ba_trace_after(hipStream_t stream) {
  TraceState &t = g_trace;
  if (!t.events || !t.open) return;
  (void)hipEventRecord(t.events[2 * t.count + 1], stream);
  t.count++;
  t.open = false;
}

extern "C" int ba_profile_begin(ba_event_t *events, uint32_t capacity) {
  return ba_guarded([&]() -> int {
    if (!events || capacity == 0) return BA_ERROR_INVALID_VALUE;
    g_trace = TraceState{};
    g_trace.events = (hipEvent_t *)events;
    g_trace.capacity = capacity;
    return BA_SUCCESS;
  });
}

extern "C" int ba_profile_end(uint32_t *launches) {
  return ba_guarded([&]() -> int {
    if (launches) *launches = g_trace.count;
    g_trace.events = nullptr;
    g_trace.capacity = 0;
    return BA_SUCCESS;
  });
}

extern "C" const char *ba_profile_launch_name(uint32_t index) {
  return index < g_trace.names.size() ? g_trace.names[index] : nullptr;
}
