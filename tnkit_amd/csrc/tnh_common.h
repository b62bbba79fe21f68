// Shared definitions of libtnkit_hip: dtype codes, status codes, error capture, vector types.
// gfx950 (CDNA4) only; wave size 64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/tnkit_hip.h"

namespace tnh {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;

// Thread-local last error text, returned by tnh_last_error().
void set_error(const char* fmt, ...);

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves a caller-provided workspace into aligned pieces.
struct Arena {
  char* base;
  size_t cap;
  size_t used = 0;
  Arena(void* p, size_t n) : base(static_cast<char*>(p)), cap(n) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = align_up(used, 256);
    used = off + count * sizeof(T);
    return reinterpret_cast<T*>(base + off);
  }
  bool ok() const { return used <= cap; }
};

}  // namespace tnh

#define TNH_CHECK_LAUNCH(what)                                                   \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      tnh::set_error("%s: launch failed: %s", what, hipGetErrorString(e_));      \
      return TNH_ERR_HIP;                                                        \
    }                                                                            \
  } while (0)

#define TNH_REQUIRE(cond, code, ...)  \
  do {                                \
    if (!(cond)) {                    \
      tnh::set_error(__VA_ARGS__);    \
      return code;                    \
    }                                 \
  } while (0)
