// Device runtime: device selection, the single work stream, a size-bucketed
// caching allocator and the error string behind dmrgx_last_error().
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "common.hpp"
#include "dmrgx.h"

namespace dmrgx {

static Runtime g_rt;
static thread_local std::string g_last_error;
static std::mutex g_pool_mu;
// free lists keyed by rounded size; live map for frees
static std::multimap<size_t, void*> g_free;
static std::unordered_map<void*, size_t> g_live;
static size_t g_in_use = 0, g_cached = 0;

void set_last_error(const std::string& msg) { g_last_error = msg; }
Runtime& rt() { return g_rt; }

void ensure_init() {
  if (!g_rt.ready) {
    int rc = dmrgx_init(0);
    if (rc != 0) throw Error(rc, "dmrgx: device initialisation failed: " + g_last_error);
  }
}

static size_t round_size(size_t b) {
  if (b <= 4096) return 4096;
  size_t r = 4096;
  while (r < b) r <<= 1;
  if (r - b > (b >> 2)) {  // beyond 25% waste use 1 MiB granularity
    const size_t g = size_t(1) << 20;
    r = (b + g - 1) / g * g;
  }
  return r;
}

void* dev_alloc(size_t bytes) {
  ensure_init();
  size_t r = round_size(bytes);
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_free.lower_bound(r);
    if (it != g_free.end() && it->first <= r + (r >> 2)) {
      void* p = it->second;
      size_t sz = it->first;
      g_free.erase(it);
      g_cached -= sz;
      g_live[p] = sz;
      g_in_use += sz;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    trim_pool();
    DMRGX_HIP(hipMalloc(&p, r));
  }
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_live[p] = r;
  g_in_use += r;
  return p;
}

void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return;
  size_t sz = it->second;
  g_live.erase(it);
  g_in_use -= sz;
  g_free.emplace(sz, p);
  g_cached += sz;
}

void trim_pool() {
  // the stream must be idle before cached blocks go back to the driver
  if (g_rt.stream) (void)hipStreamSynchronize(g_rt.stream);
  std::lock_guard<std::mutex> lk(g_pool_mu);
  for (auto& kv : g_free) (void)hipFree(kv.second);
  g_free.clear();
  g_cached = 0;
}

size_t pool_bytes_in_use() { return g_in_use; }

__global__ void k_fill_zero_u32(uint32_t* __restrict__ p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

void fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  // every element type here is a multiple of 4 bytes
  DMRGX_CHECK(((uintptr_t)p & 3) == 0 && (bytes & 3) == 0, "fill_zero: range not 4-byte aligned");
  size_t n4 = bytes / 4;
  size_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_zero_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, n4);
  DMRGX_HIP(hipGetLastError());
}

__global__ void k_fill_zero_u16(uint16_t* __restrict__ p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

void fill_zero_u16(uint16_t* p, size_t n, hipStream_t s) {
  if (!n) return;
  DMRGX_CHECK(((uintptr_t)p & 1) == 0, "fill_zero_u16: range not 2-byte aligned");
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_zero_u16, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  DMRGX_HIP(hipGetLastError());
}

void copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes) DMRGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

namespace {
// pinned ring for small host-to-device uploads: a copy from pageable memory
// returns only after the stream has drained up to it, one from pinned memory
// is enqueued and returns
constexpr size_t UP_RING = 32u << 20, UP_MAX = 1u << 20;
std::mutex g_up_mu;
char* g_up_ring = nullptr;
size_t g_up_head = 0;
}  // namespace

void upload_ring_init() {
  if (!g_up_ring) DMRGX_HIP(hipHostMalloc((void**)&g_up_ring, UP_RING, hipHostMallocDefault));
}

void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  if (bytes > UP_MAX) {
    DMRGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  std::lock_guard<std::mutex> lk(g_up_mu);
  if (!g_up_ring) upload_ring_init();
  size_t a = (g_up_head + 255) & ~(size_t)255;
  if (a + bytes > UP_RING) {
    // the next lap overwrites staged bytes: every copy that reads them must
    // have completed
    DMRGX_HIP(hipDeviceSynchronize());
    a = 0;
  }
  std::memcpy(g_up_ring + a, src, bytes);
  DMRGX_HIP(hipMemcpyAsync(dst, g_up_ring + a, bytes, hipMemcpyHostToDevice, s));
  g_up_head = a + bytes;
}

}  // namespace dmrgx

using namespace dmrgx;

#define CAPI_BEGIN try {
#define CAPI_END                                   \
  return 0;                                        \
  }                                                \
  catch (const dmrgx::Error& e) {                  \
    set_last_error(e.what());                      \
    return e.code;                                 \
  }                                                \
  catch (const std::exception& e) {                \
    set_last_error(e.what());                      \
    return -2;                                     \
  }

extern "C" {

int dmrgx_init(int device) {
  CAPI_BEGIN
  if (g_rt.ready) {
    DMRGX_CHECK(device == g_rt.device, "dmrgx_init: already initialised on another device");
    return 0;
  }
  int n = 0;
  DMRGX_HIP(hipGetDeviceCount(&n));
  DMRGX_CHECK(n > 0, "no HIP device visible");
  DMRGX_CHECK(device >= 0 && device < n, "device index out of range");
  DMRGX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  DMRGX_HIP(hipGetDeviceProperties(&prop, device));
  g_rt.num_cu = prop.multiProcessorCount;
  DMRGX_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  {
    std::lock_guard<std::mutex> lk(g_up_mu);
    upload_ring_init();
  }
  noalloc_setup(device);
  g_rt.device = device;
  g_rt.ready = true;
  CAPI_END
}

int dmrgx_finalize(void) {
  CAPI_BEGIN
  if (!g_rt.ready) return 0;
  DMRGX_HIP(hipStreamSynchronize(g_rt.stream));
  trim_pool();
  CAPI_END
}

const char* dmrgx_last_error(void) { return g_last_error.c_str(); }

int dmrgx_device_info(int* num_cu, char* name, int name_len) {
  CAPI_BEGIN
  ensure_init();
  hipDeviceProp_t prop;
  DMRGX_HIP(hipGetDeviceProperties(&prop, g_rt.device));
  if (num_cu) *num_cu = prop.multiProcessorCount;
  if (name && name_len > 0) {
    snprintf(name, name_len, "%s", prop.gcnArchName);
  }
  CAPI_END
}

int dmrgx_malloc(size_t bytes, void** ptr) {
  CAPI_BEGIN
  DMRGX_CHECK(ptr != nullptr, "dmrgx_malloc: null out pointer");
  *ptr = dev_alloc(bytes ? bytes : 8);
  CAPI_END
}

int dmrgx_free(void* ptr) {
  CAPI_BEGIN
  dev_free(ptr);
  CAPI_END
}

int dmrgx_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  CAPI_BEGIN
  ensure_init();
  if (bytes) DMRGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_rt.stream));
  DMRGX_HIP(hipStreamSynchronize(g_rt.stream));
  CAPI_END
}

int dmrgx_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  CAPI_BEGIN
  ensure_init();
  if (bytes) DMRGX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_rt.stream));
  DMRGX_HIP(hipStreamSynchronize(g_rt.stream));
  CAPI_END
}

int dmrgx_memset_zero(void* dst, size_t bytes) {
  CAPI_BEGIN
  ensure_init();
  fill_zero(dst, bytes, g_rt.stream);
  DMRGX_HIP(hipStreamSynchronize(g_rt.stream));  // complete on return
  CAPI_END
}

int dmrgx_synchronize(void) {
  CAPI_BEGIN
  ensure_init();
  DMRGX_HIP(hipStreamSynchronize(g_rt.stream));
  CAPI_END
}

void* dmrgx_stream(void) { return g_rt.ready ? (void*)g_rt.stream : nullptr; }

}  // extern "C"
