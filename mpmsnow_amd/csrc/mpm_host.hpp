// mpm_host.hpp -- host-side helpers shared by the runtime and the learned preconditioner:
// error types, HIP status checks, growable device and pinned buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace mpmh {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
struct InvalidArg : std::runtime_error {
  explicit InvalidArg(const std::string& s) : std::runtime_error(s) {}
};

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) throw HipError(std::string("hip: ") + hipGetErrorName(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) HIPCHECK(hipFree(p));
    p = nullptr;
    // headroom: grids grow step to step as bodies fragment and spread, and every regrowth costs
    // a device-wide wait in hipFree plus a fresh hipMalloc (milliseconds for large buffers)
    size_t c = std::max<size_t>(2 * bytes, 256);
    HIPCHECK(hipMalloc(&p, c));
    cap = c;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  // grows with headroom: hipHostFree waits for the device and a fresh hipHostMalloc costs
  // hundreds of microseconds, while per-body host arrays grow a little most steps as bodies fragment
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) HIPCHECK(hipHostFree(p));
    p = nullptr;
    size_t c = std::max<size_t>(bytes + bytes / 2, 1 << 16);
    HIPCHECK(hipHostMalloc(&p, c));
    cap = c;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace mpmh
