// Shared helpers for the gfx950 kernels of libcvqpg: error reporting for the C ABI,
// workspace carving and small device utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include <stdarg.h>

namespace cvqpg {

// Thread-local last error message returned by cvqpg_last_error().
void set_error(const char* fmt, ...);
const char* get_error();

#define CVQPG_CHECK_ARG(cond, ...)                                                    \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      ::cvqpg::set_error(__VA_ARGS__);                                               \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

#define CVQPG_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::cvqpg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                        \
      return 2;                                                                      \
    }                                                                                \
  } while (0)

#define CVQPG_LAUNCH_CHECK()                                                          \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess) {                                                          \
      ::cvqpg::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                        \
      return 2;                                                                      \
    }                                                                                \
  } while (0)

// Bump allocator over a caller-owned workspace; every region is 256-byte aligned.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
  size_t bytes() const { return (off + 255) & ~size_t(255); }
};

static inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace cvqpg
