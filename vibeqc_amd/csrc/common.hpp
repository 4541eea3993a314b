// Shared host/device infrastructure of the MI355X DF-JK engine: error
// propagation, owned device buffers and small math helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vqc {

// Failures raised inside the library.  The C ABI maps them to status codes
// (the reference maps exceptions the same way, src/api/error.cpp:12-35).
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NotImplemented : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NumericalFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct OutOfMemory : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory) throw OutOfMemory(std::string(what) + ": out of device memory");
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}
inline void blas_check(rocblas_status s, const char* what) {
  if (s == rocblas_status_memory_error) throw OutOfMemory(std::string(what) + ": rocBLAS out of memory");
  if (s != rocblas_status_success)
    throw DeviceError(std::string(what) + ": rocBLAS status " + std::to_string(static_cast<int>(s)));
}
#define VQC_HIP(call) ::vqc::hip_check((call), #call)
#define VQC_BLAS(call) ::vqc::blas_check((call), #call)
#define VQC_LAUNCH(what) ::vqc::hip_check(hipGetLastError(), what)

// Owned device allocation; moves, never copies.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = std::exchange(o.p_, nullptr);
      n_ = std::exchange(o.n_, 0);
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(std::size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      void* q = nullptr;
      hip_check(hipMalloc(&q, n * sizeof(T)), "hipMalloc");
      p_ = static_cast<T*>(q);
      n_ = n;
    }
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* data() const { return p_; }
  std::size_t size() const { return n_; }
  std::size_t bytes() const { return n_ * sizeof(T); }
  void upload(const T* host, std::size_t n, hipStream_t s) {
    if (n > n_) throw std::invalid_argument("upload larger than device buffer");
    hip_check(hipMemcpyAsync(p_, host, n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    resize(v.size());
    upload(v.data(), v.size(), s);
  }
  void download(T* host, std::size_t n, hipStream_t s) const {
    if (n > n_) throw std::invalid_argument("download larger than device buffer");
    hip_check(hipMemcpyAsync(host, p_, n * sizeof(T), hipMemcpyDeviceToHost, s), "download");
  }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

inline std::size_t pair_count(std::size_t n) { return n * (n + 1) / 2; }

}  // namespace vqc
