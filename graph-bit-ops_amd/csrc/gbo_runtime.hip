// Thin C-ABI wrappers over the HIP runtime used by hosts without their own
// GPU bindings (ctypes tests, the CLIs, bench.py).
#include <algorithm>

#include "gbo_device.hpp"

namespace {

// Grid-stride 16-byte copy: the plain copy used to measure the HBM copy
// ceiling and to calibrate the FETCH_SIZE / WRITE_SIZE counters.  Each
// thread issues four loads before their stores (64 bytes in flight per
// lane); the stores are nontemporal.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    u32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

// Read-only stream: every 16-byte word is loaded once (four in flight per
// lane) and folded into one XOR per thread, stored only when it equals
// `never` (a value the caller knows the fold cannot take), so the loads
// cannot be dropped and nothing is written in practice -- the HBM read
// ceiling.
__global__ __launch_bounds__(256) void k_stream_read(const u32x4* __restrict__ src, size_t n16, uint32_t never,
                                                     uint32_t* __restrict__ sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t x = 0;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    u32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < n16; i += stride) {
    const u32x4 v = src[i];
    x ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  if (x == never) sink[0] = x;
}

// Gather probe: out-of-order reads table[idx[i]] for i in [0, n), idx
// streamed with 16-byte loads (four indices per lane), the four gathers of
// a lane in flight together, every value XOR-folded per thread and stored
// only when the fold equals `never`.  ELEM is the element size (1 or 4).
template <int ELEM>
__global__ __launch_bounds__(256) void k_gather_probe(const void* __restrict__ table, const int32_t* __restrict__ idx,
                                                      int64_t n, uint32_t never, uint32_t* __restrict__ sink) {
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t x = 0;
  const int64_t n4 = n / 4;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const i32x4 k = __builtin_nontemporal_load((const i32x4*)idx + i);
    if constexpr (ELEM == 1) {
      const uint8_t* t = (const uint8_t*)table;
      x ^= (uint32_t)t[k.x] ^ ((uint32_t)t[k.y] << 8) ^ ((uint32_t)t[k.z] << 16) ^ ((uint32_t)t[k.w] << 24);
    } else {
      const uint32_t* t = (const uint32_t*)table;
      x ^= t[k.x] ^ t[k.y] ^ t[k.z] ^ t[k.w];
    }
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x ^= ELEM == 1 ? (uint32_t)((const uint8_t*)table)[idx[i]] : ((const uint32_t*)table)[idx[i]];
  if (x == never) sink[0] = x;
}

// Row-gather probe: a wave per index reads row idx[i] of a table of
// 256-byte rows (lane l the l-th 4-byte word), eight rows in flight per
// wave -- the row reads of a K = 64 SpMM.
__global__ __launch_bounds__(256) void k_gather_rows(const uint32_t* __restrict__ table, const int32_t* __restrict__ idx,
                                                     int64_t n, uint32_t never, uint32_t* __restrict__ sink) {
  const int lane = threadIdx.x & 63;
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t x = 0;
  for (int64_t i = w * 8; i < n; i += nw * 8) {
    int32_t k[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) k[u] = i + u < n ? idx[i + u] : idx[i];
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = table[(int64_t)k[u] * 64 + lane];
#pragma unroll
    for (int u = 0; u < 8; ++u) x ^= v[u];
  }
  if (x == never) sink[0] = x;
}

}  // namespace

extern "C" {

int gbo_gather_probe(const void* d_table, int elem_bytes, const int32_t* d_idx, int64_t n, uint32_t never,
                     uint32_t* d_sink, gbo_stream_t stream) {
  if (!d_table || !d_idx || !d_sink || n < 0 || (elem_bytes != 1 && elem_bytes != 4 && elem_bytes != 256) ||
      (uintptr_t)d_idx % 16)
    return gbo::set_error(GBO_EINVAL, "gbo_gather_probe: bad argument (elem_bytes 1, 4 or 256, idx 16-byte aligned)");
  if (n == 0) return GBO_OK;
  if (elem_bytes == 256) {
    k_gather_rows<<<256 * 32, 256, 0, stream>>>((const uint32_t*)d_table, d_idx, n, never, d_sink);
    GBO_LAUNCH_CHECK();
    return GBO_OK;
  }
  const int64_t blocks = std::min<int64_t>((n / 4 + 255) / 256 + 1, 256 * 64);
  if (elem_bytes == 1) k_gather_probe<1><<<(unsigned)blocks, 256, 0, stream>>>(d_table, d_idx, n, never, d_sink);
  else k_gather_probe<4><<<(unsigned)blocks, 256, 0, stream>>>(d_table, d_idx, n, never, d_sink);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

int gbo_stream_copy(void* d_dst, const void* d_src, size_t bytes, gbo_stream_t stream) {
  if (bytes % 16 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16)
    return gbo::set_error(GBO_EINVAL, "gbo_stream_copy: size and pointers must be 16-byte aligned");
  const size_t n16 = bytes / 16;
  if (n16 == 0) return GBO_OK;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
  k_stream_copy<<<(unsigned)blocks, 256, 0, stream>>>((const u32x4*)d_src, (u32x4*)d_dst, n16);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

int gbo_stream_read(const void* d_src, size_t bytes, uint32_t never, uint32_t* d_sink, gbo_stream_t stream) {
  if (bytes % 16 || (uintptr_t)d_src % 16 || !d_sink)
    return gbo::set_error(GBO_EINVAL, "gbo_stream_read: size and pointer must be 16-byte aligned, sink non-null");
  const size_t n16 = bytes / 16;
  if (n16 == 0) return GBO_OK;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
  k_stream_read<<<(unsigned)blocks, 256, 0, stream>>>((const u32x4*)d_src, n16, never, d_sink);
  GBO_LAUNCH_CHECK();
  return GBO_OK;
}

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
