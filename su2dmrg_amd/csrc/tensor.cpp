#include "tensor.hpp"

#include <algorithm>

#include "blas1.hpp"
#include "common.hpp"
#include "context.hpp"

namespace su2 {

Context::Context(int device) : device_(device) {
  int n = 0;
  SU2_HIP(hipGetDeviceCount(&n));
  SU2_CHECK(device >= 0 && device < n, kErrArgument, "device %d out of range (%d devices)", device,
            n);
  SU2_HIP(hipSetDevice(device));
  // default priority stream (never a prioritised one)
  SU2_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  own_stream_ = true;
  pool_ = std::make_shared<DevicePool>(device, stream_);
  // address of this device's copy of the allocation-free dot's scratch
  noalloc_scratch_init(device);
}

Context::~Context() {
  (void)hipSetDevice(device_);
  pool_->detach();
  if (d_red_) (void)hipFree(d_red_);
  if (h_red_) (void)hipHostFree(h_red_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Context::make_current() const { SU2_HIP(hipSetDevice(device_)); }

void Context::set_stream(hipStream_t s) {
  pool_->set_stream(s);
  stream_ = s;
}

namespace {
constexpr size_t kCacheFloor = size_t(4) << 30;   // cache limit floor: 4 GiB
// allocation granule: 256 B up to 1 MiB, 64 KiB up to 64 MiB, 2 MiB above
size_t pool_round(size_t b) {
  const size_t g = b <= (size_t(1) << 20) ? 256 : b <= (size_t(64) << 20) ? (size_t(64) << 10)
                                                                         : (size_t(2) << 20);
  return (b + g - 1) / g * g;
}

// every pool of the process (for releases when device memory runs out)
std::mutex& pool_registry_mutex() {
  static std::mutex m;
  return m;
}
std::vector<DevicePool*>& pool_registry() {
  static std::vector<DevicePool*> v;
  return v;
}
}  // namespace

DevicePool::DevicePool(int device, hipStream_t stream) : device_(device), stream_(stream) {
  std::lock_guard<std::mutex> g(pool_registry_mutex());
  pool_registry().push_back(this);
}

DevicePool::~DevicePool() {
  {
    std::lock_guard<std::mutex> g(pool_registry_mutex());
    auto& r = pool_registry();
    r.erase(std::remove(r.begin(), r.end(), this), r.end());
  }
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& x : free_) (void)hipFree(x.second);
  for (auto& x : live_) (void)hipFree(x.first);
}

void DevicePool::release_device(int device) {
  std::lock_guard<std::mutex> g(pool_registry_mutex());
  for (DevicePool* p : pool_registry()) {
    if (p->device_ != device) continue;
    std::unique_lock<std::mutex> l(p->mu_, std::try_to_lock);
    if (!l.owns_lock() || p->free_.empty()) continue;
    // no throw here: a failed wait leaves that pool's cache as it is
    if (p->stream_ && hipStreamSynchronize(p->stream_) != hipSuccess) continue;
    for (auto& x : p->free_) (void)hipFree(x.second);
    p->free_.clear();
    p->cached_bytes_ = 0;
  }
}

void* DevicePool::alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  std::unique_lock<std::mutex> lock(mu_);
  SU2_CHECK(!detached_, kErrArgument, "device memory requested after its context was destroyed");
  const size_t sz = pool_round(bytes);
  auto take = [&]() -> void* {
    auto it = free_.lower_bound(sz);
    if (it == free_.end() || it->first > 2 * sz) return nullptr;
    void* p = it->second;
    live_[p] = it->first;
    live_bytes_ += it->first;
    cached_bytes_ -= it->first;
    peak_bytes_ = std::max(peak_bytes_, live_bytes_);
    high_water_ = std::max(high_water_, live_bytes_);
    free_.erase(it);
    return p;
  };
  if (void* p = take()) return p;
  void* p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) {
    (void)hipGetLastError();
    release_free();
    if (hipMalloc(&p, sz) != hipSuccess) {
      // the other pools' caches on this device (without holding this one's
      // lock, so that two threads out of memory cannot wait on each other)
      (void)hipGetLastError();
      lock.unlock();
      release_device(device_);
      lock.lock();
      SU2_HIP(hipMalloc(&p, sz));
    }
  }
  live_[p] = sz;
  live_bytes_ += sz;
  peak_bytes_ = std::max(peak_bytes_, live_bytes_);
  high_water_ = std::max(high_water_, live_bytes_);
  return p;
}

void DevicePool::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(mu_);
  auto it = live_.find(p);
  if (detached_ || it == live_.end()) {   // context gone (or not a pool block)
    if (it != live_.end()) {
      live_bytes_ -= it->second;
      live_.erase(it);
    }
    (void)hipSetDevice(device_);
    (void)hipFree(p);
    return;
  }
  // reusable at once: later work on this stream runs after the work queued
  // so far, the last use of this block included
  free_.insert({it->second, p});
  cached_bytes_ += it->second;
  live_bytes_ -= it->second;
  live_.erase(it);
  if (cached_bytes_ > std::max(kCacheFloor, high_water_)) trim();
}

void DevicePool::trim() {
  // runs from free(), which destructors call: no throw; a failed wait
  // leaves the cache as it is
  const size_t target = std::max(kCacheFloor, high_water_) / 2;
  if (stream_ && hipStreamSynchronize(stream_) != hipSuccess) return;
  while (cached_bytes_ > target && !free_.empty()) {
    auto it = std::prev(free_.end());
    cached_bytes_ -= it->first;
    (void)hipFree(it->second);
    free_.erase(it);
  }
}

void DevicePool::set_stream(hipStream_t s) {
  std::lock_guard<std::mutex> lock(mu_);
  drain();   // cached blocks may still be in use by work on the old stream
  stream_ = s;
}

void DevicePool::detach() noexcept {
  // runs from ~Context: a sticky device error must not throw here, so the
  // wait's status is dropped (su2_ctx_destroy waits and reports it first)
  std::lock_guard<std::mutex> lock(mu_);
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& x : free_) (void)hipFree(x.second);
  free_.clear();
  cached_bytes_ = 0;
  stream_ = nullptr;
  detached_ = true;
}

void DevicePool::drain() {
  if (stream_ && !free_.empty()) SU2_HIP(hipStreamSynchronize(stream_));
}

void DevicePool::release_free() {
  drain();
  for (auto& x : free_) (void)hipFree(x.second);
  free_.clear();
  cached_bytes_ = 0;
}

double* Context::reduce_scratch(size_t n) {
  if (n > d_red_n_) {
    if (d_red_) SU2_HIP(hipFree(d_red_));
    size_t m = std::max(n, d_red_n_ * 2);
    SU2_HIP(hipMalloc(&d_red_, m * sizeof(double)));
    d_red_n_ = m;
  }
  return d_red_;
}

double* Context::host_scratch(size_t n) {
  if (n > h_red_n_) {
    if (h_red_) SU2_HIP(hipHostFree(h_red_));
    size_t m = std::max(n, h_red_n_ * 2);
    SU2_HIP(hipHostMalloc(&h_red_, m * sizeof(double), hipHostMallocDefault));
    h_red_n_ = m;
  }
  return h_red_;
}

static long long key_of(const int* twoj, int n) {
  long long k = 0;
  for (int i = 0; i < n; ++i) k = k * 64 + twoj[i];
  return k;
}

Tensor::Tensor(Context* ctx, std::vector<Leg> legs, int locspin, const std::vector<char>* present,
               bool zero)
    : ctx_(ctx), legs_(std::move(legs)), locspin_(locspin) {
  const int n = (int)legs_.size();
  SU2_CHECK(n >= 1 && n <= kMaxLegs, kErrArgument, "tensor: %d legs", n);
  SU2_CHECK(locspin >= 0, kErrArgument, "tensor: negative locspin");
  for (auto& l : legs_) {
    SU2_CHECK(l.dir == 1 || l.dir == -1, kErrArgument, "tensor: leg direction must be +-1");
    nten_ *= std::max(1, l.nsec());
    for (int t : l.twoj) SU2_CHECK(t < 64, kErrArgument, "tensor: 2j=%d too large", t);
  }
  for (auto& l : legs_) SU2_CHECK(l.nsec() > 0, kErrArgument, "tensor: empty leg");
  if (present) SU2_CHECK((int)present->size() == nten_, kErrArgument,
                         "tensor: %zu present flags for %d blocks", present->size(), nten_);
  long long off = 0;
  for (int idx = 0; idx < nten_; ++idx) {
    Block b;
    int r = idx;
    for (int i = 0; i < n; ++i) {
      int s = r % legs_[i].nsec();
      r /= legs_[i].nsec();
      b.sec[i] = s;
      b.twoj[i] = legs_[i].twoj[s];
      b.dim[i] = legs_[i].dim[s];
    }
    bool ok = labels_allowed(b.twoj, n, locspin);
    if (present) {
      if ((*present)[idx] && !ok)
        fail(kErrStructure, "tensor: block %d flagged present but its labels are not allowed", idx);
      ok = ok && (*present)[idx];
    }
    if (!ok) continue;
    b.index = idx;
    b.size = 1;
    for (int i = 0; i < n; ++i) b.size *= b.dim[i];
    b.offset = off;
    off += b.size;
    by_key_[key_of(b.twoj, n)] = (int)blocks_.size();
    blocks_.push_back(b);
  }
  size_ = off;
  ctx_->make_current();
  pool_ = ctx_->pool();
  if (size_ > 0) {
    d_ = ctx_->dev_alloc_n<double>(size_);
    if (zero) SU2_HIP(hipMemsetAsync(d_, 0, size_ * sizeof(double), ctx_->stream()));
  }
}

void ZeroBatch::add(Tensor& t) {
  if (t.size() > 0) z_.push_back({t.data(), t.size()});
}

void ZeroBatch::run(Context& c) {
  if (z_.empty()) return;
  std::vector<VecMove> h(z_.size());
  long long mx = 0;
  for (size_t i = 0; i < z_.size(); ++i) {
    h[i] = VecMove{nullptr, z_[i].first, z_[i].second, kZeroF64, 0};
    mx = std::max(mx, z_[i].second);
  }
  // the descriptor block returns to the cache at once; stream order keeps it
  // until the launch has read it
  VecMove* d = (VecMove*)c.dev_alloc(h.size() * sizeof(VecMove));
  try {
    SU2_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(VecMove), hipMemcpyHostToDevice,
                           c.stream()));
    for (size_t i = 0; i < h.size(); i += 65535)
      dev_move_batch(c, d + i, (int)std::min<size_t>(65535, h.size() - i), mx);
  } catch (...) {
    c.dev_free(d);
    throw;
  }
  c.dev_free(d);
  z_.clear();
}

Tensor::~Tensor() {
  if (d_) pool_->free(d_);
}

int Tensor::find(const int* twoj) const {
  auto it = by_key_.find(key_of(twoj, (int)legs_.size()));
  return it == by_key_.end() ? -1 : it->second;
}

bool Tensor::same_layout(const Tensor& o) const {
  if (legs_.size() != o.legs_.size() || locspin_ != o.locspin_) return false;
  for (size_t i = 0; i < legs_.size(); ++i)
    if (!(legs_[i] == o.legs_[i])) return false;
  if (blocks_.size() != o.blocks_.size()) return false;
  for (size_t i = 0; i < blocks_.size(); ++i)
    if (blocks_[i].index != o.blocks_[i].index) return false;
  return true;
}

}  // namespace su2
