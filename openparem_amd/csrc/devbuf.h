// Owning device buffer and the HIP error check shared by the host sources of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace opem {

#define HIPCHK(x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (b <= bytes && p) return;
    release();
    if (b == 0) b = 16;
    HIPCHK(hipMalloc(&p, b));
    bytes = b;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
  template <typename T>
  void upload(const std::vector<T> &v, hipStream_t st) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  }
};

}  // namespace opem
