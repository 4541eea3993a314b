// Shared host/device infrastructure of the MI355X DF-JK engine: error
// propagation, owned device buffers and small math helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
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

// Host<->device copies go through a pinned staging buffer owned by the
// calling host thread and complete before returning, so no copy depends on
// the runtime's shared pageable-memory staging or outlives its host buffer
// (plans of a ragged batch copy from many host threads at once).
class PinnedStaging {
 public:
  ~PinnedStaging() {
    if (p_) (void)hipHostFree(p_);
  }
  void* get(std::size_t bytes) {
    if (bytes > cap_) {
      if (p_) (void)hipHostFree(p_);
      p_ = nullptr;
      cap_ = 0;
      const std::size_t want = std::max<std::size_t>(bytes, std::size_t{1} << 20);
      hip_check(hipHostMalloc(&p_, want, hipHostMallocDefault), "hipHostMalloc");
      cap_ = want;
    }
    return p_;
  }

 private:
  void* p_ = nullptr;
  std::size_t cap_ = 0;
};
inline PinnedStaging& thread_staging() {
  static thread_local PinnedStaging staging;
  return staging;
}
// Larger transfers are moved in pieces of at most this many bytes.
constexpr std::size_t kStagingChunk = std::size_t{64} << 20;
inline void copy_to_device(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
  for (std::size_t off = 0; off < bytes; off += kStagingChunk) {
    const std::size_t len = std::min(kStagingChunk, bytes - off);
    void* stage = thread_staging().get(len);
    std::memcpy(stage, static_cast<const char*>(src) + off, len);
    hip_check(hipMemcpyAsync(static_cast<char*>(dst) + off, stage, len, hipMemcpyHostToDevice, s), "upload");
    hip_check(hipStreamSynchronize(s), "upload");
  }
}
inline void copy_to_host(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
  for (std::size_t off = 0; off < bytes; off += kStagingChunk) {
    const std::size_t len = std::min(kStagingChunk, bytes - off);
    void* stage = thread_staging().get(len);
    hip_check(hipMemcpyAsync(stage, static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "download");
    std::memcpy(static_cast<char*>(dst) + off, stage, len);
  }
}

// Debug aid: VIBEQC_POISON_ALLOC=1 fills every new device allocation with
// 0xFF bytes (NaN in FP64/FP32/BF16), so results that depend on memory read
// before it is written become non-finite instead of silently varying.
inline bool poison_allocations() {
  static const bool on = [] {
    const char* v = std::getenv("VIBEQC_POISON_ALLOC");
    return v && v[0] == '1';
  }();
  return on;
}

// Debug aid: VIBEQC_GUARD_ALLOC=1 gives every device allocation a 4 KiB tail
// filled with 0xA5 and checks it when the allocation is released, so a kernel
// that writes past the end of a buffer is reported (to stderr, with the
// buffer size) instead of silently corrupting its neighbour.
inline bool guard_allocations() {
  static const bool on = [] {
    const char* v = std::getenv("VIBEQC_GUARD_ALLOC");
    return v && v[0] == '1';
  }();
  return on;
}
constexpr std::size_t kGuardBytes = 4096;

// Stream used for stream-ordered allocation by DeviceArrays created on this
// host thread (null: plain hipMalloc / hipFree).  Batch workers set it to
// their own stream, so the many small plans they create and destroy never
// hit hipFree's device-wide synchronisation and overlap on the GPU.
inline hipStream_t& allocation_stream() {
  static thread_local hipStream_t stream = nullptr;
  return stream;
}

// Device blocks owned by one batch worker, recycled only among that worker's
// own allocations.  Every buffer a worker allocates is used by work on the
// worker's stream only, so a block released to the cache and handed out again
// is touched by its next user's work strictly after its previous user's work
// (stream order): no block ever moves between streams while work may still
// use it.  Blocks are returned to the device with hipFree when the cache is
// destroyed (after the worker's stream has drained).
class BlockCache {
 public:
  BlockCache() = default;
  BlockCache(const BlockCache&) = delete;
  BlockCache& operator=(const BlockCache&) = delete;
  ~BlockCache() {
    for (auto& b : free_) (void)hipFree(b.second);
  }
  // A block of at least `bytes` (sizes rounded up to 4 KiB; a cached block is
  // reused when it is at most twice the request).
  void* take(std::size_t bytes) {
    bytes = (std::max<std::size_t>(bytes, 1) + 4095) & ~static_cast<std::size_t>(4095);
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= 2 * bytes) {
      void* p = it->second;
      sizes_[p] = it->first;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    sizes_[p] = bytes;
    return p;
  }
  void give(void* p) {
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    free_.emplace(it->second, p);
    sizes_.erase(it);
  }

 private:
  std::multimap<std::size_t, void*> free_;
  std::map<void*, std::size_t> sizes_;
};

// Cache used by DeviceArrays created on this host thread (null: none).
inline BlockCache*& allocation_cache() {
  static thread_local BlockCache* cache = nullptr;
  return cache;
}

// Owned device allocation; moves, never copies.
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(std::size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept
      : p_(std::exchange(o.p_, nullptr)),
        n_(std::exchange(o.n_, 0)),
        s_(std::exchange(o.s_, nullptr)),
        cache_(std::exchange(o.cache_, nullptr)) {}
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = std::exchange(o.p_, nullptr);
      n_ = std::exchange(o.n_, 0);
      s_ = std::exchange(o.s_, nullptr);
      cache_ = std::exchange(o.cache_, nullptr);
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void resize(std::size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      void* q = nullptr;
      s_ = allocation_stream();
      cache_ = allocation_cache();
      const std::size_t bytes = n * sizeof(T) + (guard_allocations() ? kGuardBytes : 0);
      if (cache_) q = cache_->take(bytes);
      else if (s_) hip_check(hipMallocAsync(&q, bytes, s_), "hipMallocAsync");
      else hip_check(hipMalloc(&q, bytes), "hipMalloc");
      if (guard_allocations()) {
        hip_check(hipMemsetAsync(static_cast<char*>(q) + n * sizeof(T), 0xA5, kGuardBytes, s_), "guard");
        hip_check(hipStreamSynchronize(s_), "guard");
      }
      p_ = static_cast<T*>(q);
      n_ = n;
      if (poison_allocations()) {
        hip_check(hipMemsetAsync(q, 0xFF, n * sizeof(T), s_), "poison");
        hip_check(hipStreamSynchronize(s_), "poison");
      }
    }
  }
  void release() {
    if (p_ && guard_allocations()) {
      unsigned char tail[kGuardBytes];
      if (hipStreamSynchronize(s_) == hipSuccess &&
          hipMemcpy(tail, reinterpret_cast<char*>(p_) + n_ * sizeof(T), kGuardBytes, hipMemcpyDeviceToHost) ==
              hipSuccess) {
        for (std::size_t k = 0; k < kGuardBytes; ++k)
          if (tail[k] != 0xA5) {
            std::fprintf(stderr, "{\"guard\": \"overrun\", \"bytes\": %zu, \"first_bad_offset\": %zu}\n",
                         n_ * sizeof(T), k);
            break;
          }
      }
    }
    if (p_) {
      if (cache_) cache_->give(p_);
      else if (s_) (void)hipFreeAsync(p_, s_);
      else (void)hipFree(p_);
    }
    p_ = nullptr;
    n_ = 0;
    s_ = nullptr;
    cache_ = nullptr;
  }
  T* data() const { return p_; }
  std::size_t size() const { return n_; }
  std::size_t bytes() const { return n_ * sizeof(T); }
  // Host-to-device copy, complete on return (the source may be a temporary).
  void upload(const T* host, std::size_t n, hipStream_t s) {
    if (n > n_) throw std::invalid_argument("upload larger than device buffer");
    copy_to_device(p_, host, n * sizeof(T), s);
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    resize(v.size());
    upload(v.data(), v.size(), s);
  }
  // Device-to-host copy, complete on return (orders after prior work on s).
  void download(T* host, std::size_t n, hipStream_t s) const {
    if (n > n_) throw std::invalid_argument("download larger than device buffer");
    copy_to_host(host, p_, n * sizeof(T), s);
  }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
  hipStream_t s_ = nullptr;  // stream of a stream-ordered allocation
  BlockCache* cache_ = nullptr;  // worker cache the block came from
};

inline std::size_t pair_count(std::size_t n) { return n * (n + 1) / 2; }

}  // namespace vqc
