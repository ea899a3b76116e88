// Host-side helpers shared by the library's translation units: device
// buffers (RAII), the thread-local error message and HIP status checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/jaxrens_amd.h"

namespace jrh {

// message returned by jr_global_error(); defined once in jr_lib.hip
extern thread_local std::string g_err;

int set_global_err(int code, const char* fmt, ...);

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace jrh

#define HIPCHK_G(call)                                                    \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      return jrh::set_global_err(JR_E_HIP, "%s failed: %s", #call,        \
                                 hipGetErrorString(e_));                  \
  } while (0)
