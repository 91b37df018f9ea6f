// Library context: one device, one stream, the last error message, and a
// small pool of device scratch for reductions.
//
// The reference reads its configuration from globals (psize, max_dcut, the
// fac_* tables: tensor_su2.cpp:14-16, tensor.cpp:25); here everything a call
// needs hangs off this object, so calls on different contexts are independent.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace su2 {

// Device memory cache of one context's stream.  All device work on pool
// memory is queued on that one stream, so a block freed on the host can serve
// the next request at once (stream order runs every earlier kernel and copy
// that used it first, as in a stream-ordered allocator): a freed block serves
// any request of at least half its size, and only a request that no cached
// block fits goes to hipMalloc.  Cached blocks return to the driver when the
// pool is destroyed, when hipMalloc runs out of memory, and, largest first
// down to half the limit, when the cache grows past max(4 GiB, the most ever
// handed out at once) (each after a stream wait), so that several contexts
// on one device do not starve each other.  Moving the pool to another
// stream waits for the old one first.
// Objects that may outlive their context (tensors, GEMM and H_eff plans)
// share ownership of the pool: once the context is gone (detach), the pool
// hands back what it holds and frees later blocks with hipFree directly.
//
// The pools of a process form a registry: when hipMalloc (or a HIP graph's
// instantiation) runs out of device memory, the free blocks of every pool on
// that device go back to the driver (each after its own stream's wait), so
// that one context's cache does not starve another's allocations.  A pool
// is guarded by its own mutex, so that another thread can release it.
class DevicePool {
 public:
  DevicePool(int device, hipStream_t stream);
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;
  void* alloc(size_t bytes);   // alloc(0) is nullptr
  void free(void* p);
  void set_stream(hipStream_t s);
  void detach() noexcept;
  // bytes handed out now, the most handed out at once, bytes held by the cache
  size_t live_bytes() const {
    std::lock_guard<std::mutex> g(mu_);
    return live_bytes_;
  }
  size_t peak_bytes() const {
    std::lock_guard<std::mutex> g(mu_);
    return peak_bytes_;
  }
  size_t cached_bytes() const {
    std::lock_guard<std::mutex> g(mu_);
    return cached_bytes_;
  }
  void reset_peak() {
    std::lock_guard<std::mutex> g(mu_);
    peak_bytes_ = live_bytes_;
  }
  // free blocks of every pool on the device back to the driver (pools busy
  // in another thread at that moment are skipped)
  static void release_device(int device);

 private:
  void drain();          // wait for the stream: no queued work uses a cached block
  void release_free();   // cached blocks back to the driver (after drain)
  void trim();           // over the cache limit: largest blocks back to the driver
  int device_;
  hipStream_t stream_;
  bool detached_ = false;
  mutable std::mutex mu_;
  std::multimap<size_t, void*> free_;           // size -> block, reusable in stream order
  std::unordered_map<void*, size_t> live_;      // handed out
  size_t live_bytes_ = 0, peak_bytes_ = 0, high_water_ = 0, cached_bytes_ = 0;
};

class Context {
 public:
  explicit Context(int device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }
  void set_stream(hipStream_t s);
  void make_current() const;

  // Device memory for work on this context's stream (DevicePool).
  void* dev_alloc(size_t bytes) { return pool_->alloc(bytes); }
  void dev_free(void* p) { pool_->free(p); }
  template <class T>
  T* dev_alloc_n(size_t n) {
    return (T*)dev_alloc(n * sizeof(T));
  }
  const std::shared_ptr<DevicePool>& pool() const { return pool_; }

  std::string last_error;

  // device scratch for block-reduction partials (grown on demand, never shrunk)
  double* reduce_scratch(size_t n);
  // pinned host scratch for reading back reduction results
  double* host_scratch(size_t n);

  int max_dcut = 8000;  // reference default (main_omp.cpp:15)
  int svd_block_cols = 32;   // one-sided Jacobi column block width (16 or 32)
  int svd_inner_sweeps = 1;  // its inner sweeps per visit of a pair (1..3)
  int svd_precondition_min_k = 256;  // QR-precondition Jacobi SVD blocks from this size (0: off)

 private:
  int device_;
  hipStream_t stream_ = nullptr;
  std::shared_ptr<DevicePool> pool_;
  bool own_stream_ = false;
  double* d_red_ = nullptr;
  size_t d_red_n_ = 0;
  double* h_red_ = nullptr;
  size_t h_red_n_ = 0;
};

// Device buffer of doubles from a context's cache, returned to it when the
// holder goes out of scope (also when an exception unwinds the scope).
struct DevBuf {
  Context* c = nullptr;
  double* p = nullptr;
  DevBuf() = default;
  DevBuf(Context& ctx, long long n) : c(&ctx), p(ctx.dev_alloc_n<double>(n > 0 ? n : 0)) {}
  ~DevBuf() {
    if (c) c->dev_free(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : c(o.c), p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (c) c->dev_free(p);
      c = o.c;
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
};

}  // namespace su2
