// hip_util.hpp -- HIP error handling and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace icarat_hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// status codes returned through the C ABI
enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kHipError = 2,
  kNotReady = 3,      // a call made before the state it needs (e.g. no make_dof)
  kOutOfMemory = 4,
  kNotConverged = 5,  // iterative method stopped at its iteration cap
  kNumerical = 6,     // breakdown: non-positive curvature, NaN, filter normaliser
  kInternal = 7,
};

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      throw ::icarat_hip::Error(_e == hipErrorOutOfMemory ? ::icarat_hip::kOutOfMemory \
                                                          : ::icarat_hip::kHipError,   \
                                std::string(#expr) + ": " + hipGetErrorString(_e));   \
    }                                                                                 \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    HIPCHK(hipMalloc((void **)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  operator T *() const { return p; }
};

template <class T>
struct PinnedBuf {
  T *p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf() { release(); }
  void alloc(size_t count) {
    release();
    HIPCHK(hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace icarat_hip
