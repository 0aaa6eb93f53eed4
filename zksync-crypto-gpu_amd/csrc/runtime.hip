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
