// nphysics_amd/csrc/common.hpp -- host helpers shared by the C ABI translation units: the
// thread's last error message, HIP error checking and a growable device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <string>

#include "../../include/nphys.h"

namespace nph {

inline thread_local std::string g_err;

inline int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

inline int hip_err(hipError_t e) { return set_err(NPHYS_ERR_HIP, std::string("HIP: ") + hipGetErrorString(e)); }

#define HIPCK(expr)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::nph::set_err(NPHYS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Growable device buffer (contents are not preserved on growth unless asked).
// Bit pattern of a float (host side of __float_as_uint).
inline uint32_t bits_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  return u;
}

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t ensure(size_t need, bool keep = false, hipStream_t s = nullptr) {
    if (need <= bytes) return hipSuccess;
    size_t nb = std::max(need, bytes + bytes / 2);
    nb = (nb + 255) & ~size_t(255);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nb);
    if (e != hipSuccess) return e;
    if (keep && p && bytes) {
      e = hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return e;
      e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
    }
    if (p) (void)hipFree(p);
    p = q;
    bytes = nb;
    return hipSuccess;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace nph
