// Thin C-ABI wrappers over the HIP runtime used by hosts without their own
// GPU bindings (ctypes tests, the CLIs, bench.py).
#include "gbo_device.hpp"

extern "C" {

int gbo_device_count(int* count) {
  if (!count) return gbo::set_error(GBO_EINVAL, "gbo_device_count: null");
  GBO_HIP_CHECK(hipGetDeviceCount(count));
  return GBO_OK;
}
int gbo_set_device(int device) {
  GBO_HIP_CHECK(hipSetDevice(device));
  return GBO_OK;
}
int gbo_malloc(void** p, size_t bytes) {
  if (!p) return gbo::set_error(GBO_EINVAL, "gbo_malloc: null");
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return gbo::set_error(GBO_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return GBO_OK;
}
int gbo_free(void* p) {
  if (p) GBO_HIP_CHECK(hipFree(p));
  return GBO_OK;
}
int gbo_memcpy_h2d(void* d, const void* h, size_t bytes) {
  if (bytes) GBO_HIP_CHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  return GBO_OK;
}
int gbo_memcpy_d2h(void* h, const void* d, size_t bytes) {
  if (bytes) GBO_HIP_CHECK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return GBO_OK;
}
int gbo_memcpy_d2d(void* d, const void* s, size_t bytes) {
  if (bytes) GBO_HIP_CHECK(hipMemcpy(d, s, bytes, hipMemcpyDeviceToDevice));
  return GBO_OK;
}
int gbo_memset(void* d, int value, size_t bytes) {
  if (bytes) GBO_HIP_CHECK(hipMemset(d, value, bytes));
  return GBO_OK;
}
int gbo_synchronize(void) {
  GBO_HIP_CHECK(hipDeviceSynchronize());
  return GBO_OK;
}
int gbo_stream_create(gbo_stream_t* s) {
  if (!s) return gbo::set_error(GBO_EINVAL, "gbo_stream_create: null");
  hipStream_t st;
  GBO_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = st;
  return GBO_OK;
}
int gbo_stream_destroy(gbo_stream_t s) {
  if (s) GBO_HIP_CHECK(hipStreamDestroy(s));
  return GBO_OK;
}
int gbo_stream_synchronize(gbo_stream_t s) {
  GBO_HIP_CHECK(hipStreamSynchronize(s));
  return GBO_OK;
}
int gbo_event_create(void** ev) {
  if (!ev) return gbo::set_error(GBO_EINVAL, "gbo_event_create: null");
  hipEvent_t e;
  GBO_HIP_CHECK(hipEventCreate(&e));
  *ev = e;
  return GBO_OK;
}
int gbo_event_destroy(void* ev) {
  if (ev) GBO_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
  return GBO_OK;
}
int gbo_event_record(void* ev, gbo_stream_t s) {
  GBO_HIP_CHECK(hipEventRecord((hipEvent_t)ev, s));
  return GBO_OK;
}
int gbo_event_elapsed(float* ms, void* a, void* b) {
  if (!ms) return gbo::set_error(GBO_EINVAL, "gbo_event_elapsed: null");
  GBO_HIP_CHECK(hipEventSynchronize((hipEvent_t)b));
  GBO_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return GBO_OK;
}

}  // extern "C"
