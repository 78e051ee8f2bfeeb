// Shared definitions for the dmrgx HIP library (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmrgx {

using cplx = std::complex<double>;

// Element type and arithmetic precision codes shared with include/dmrgx.h.
enum Dtype : int { DT_R64 = 0, DT_C64 = 1 };
enum Prec : int { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define DMRGX_HIP(expr)                                                             \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::dmrgx::Error(-1000 - (int)e_, std::string(#expr " failed: ") +        \
                                                hipGetErrorString(e_) + " at " +    \
                                                __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define DMRGX_CHECK(cond, msg)                                           \
  do {                                                                   \
    if (!(cond)) throw ::dmrgx::Error(-1, std::string("dmrgx: ") + (msg)); \
  } while (0)

// Runtime state: one device, one stream (default priority), a caching pool.
struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  bool ready = false;
};
Runtime& rt();
void ensure_init();

void* dev_alloc(size_t bytes);
void dev_free(void* p);
void trim_pool();
size_t pool_bytes_in_use();

// RAII device buffer with shared ownership (qtensor copies share storage and
// copy-on-write is explicit).
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) : bytes(b) { ptr = b ? dev_alloc(b) : nullptr; }
  ~DevBuf() { if (ptr) dev_free(ptr); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// --- launch helpers implemented in the .hip files --------------------------
void fill_zero(void* p, size_t bytes, hipStream_t s);
// n bfloat16 / u16 zeros at a 2-byte aligned address
void fill_zero_u16(uint16_t* p, size_t n, hipStream_t s);
void copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s);
// Host-to-device copy on stream s that returns without waiting for the
// stream: up to 1 MiB is first copied into a pinned ring (src may be reused at
// once), larger copies go straight.
void upload(void* dst, const void* src, size_t bytes, hipStream_t s);
// The pinned staging ring of upload(), allocated by dmrgx_init.
void upload_ring_init();
// Device address of the no-allocation entry point's scratch array on
// `device` (current), looked up once by dmrgx_init.
void noalloc_setup(int device);

inline size_t elem_bytes(int dtype) { return dtype == DT_C64 ? 16 : 8; }

}  // namespace dmrgx
