// Staged host <-> device copies (declarations in hostcopy.hpp).
#include "hostcopy.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "cdx_util.hpp"

namespace cdx {

namespace {

// Host threads that split one memcpy between them.  One copy runs at a time
// (callers queue on call_mu_); the calling thread copies the last share.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* p = new CopyPool();  // lives until exit
    return *p;
  }

  void run(char* dst, const char* src, size_t bytes) {
    const int parts = (int)workers_.size() + 1;
    // shares of whole 4 KiB pages; small copies stay on the calling thread
    const size_t share = ((bytes / parts + 4095) / 4096) * 4096;
    if (parts == 1 || bytes < (size_t(1) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    std::lock_guard<std::mutex> call(call_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = dst;
      src_ = src;
      bytes_ = bytes;
      share_ = share;
      pending_ = (int)workers_.size();
      ++gen_;
    }
    cv_.notify_all();
    copy_part(parts - 1);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }

 private:
  CopyPool() {
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    const int n = (int)std::min(8u, std::max(1u, hc / 2)) - 1;  // plus the caller
    for (int i = 0; i < n; ++i) workers_.emplace_back([this, i] { loop(i); });
    for (auto& t : workers_) t.detach();
  }

  void copy_part(int i) {
    const size_t b = std::min(bytes_, (size_t)i * share_), e = std::min(bytes_, b + share_);
    if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
  }

  void loop(int i) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      copy_part(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex call_mu_, mu_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t bytes_ = 0, share_ = 0;
};

}  // namespace

void parallel_memcpy(void* dst, const void* src, size_t bytes) {
  CopyPool::get().run(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
}

HostStager::~HostStager() {
  GateShared g;
  for (int b = 0; b < 2; ++b) {
    if (ev_[b]) {
      (void)hipEventSynchronize(ev_[b]);
      (void)hipEventDestroy(ev_[b]);
    }
    if (pin_[b]) (void)hipHostFree(pin_[b]);
  }
}

void HostStager::ensure() {
  if (pin_[0]) return;
  GateShared g;
  for (int b = 0; b < 2; ++b) {
    CDX_HIP_CHECK(hipHostMalloc(&pin_[b], kChunk, hipHostMallocDefault));
    CDX_HIP_CHECK(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
  }
}

void HostStager::h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  GateShared g;
  if (bytes < kMinStaged) {
    CDX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    CDX_HIP_CHECK(hipStreamSynchronize(s));  // src may be reused on return
    return;
  }
  ensure();
  const char* in = static_cast<const char*>(src);
  char* out = static_cast<char*>(dst);
  for (size_t off = 0, k = 0; off < bytes; off += kChunk, ++k) {
    const size_t len = std::min(kChunk, bytes - off);
    const int b = (int)(k & 1);
    CDX_HIP_CHECK(hipEventSynchronize(ev_[b]));  // the DMA out of this block (two chunks ago) has ended
    parallel_memcpy(pin_[b], in + off, len);
    CDX_HIP_CHECK(hipMemcpyAsync(out + off, pin_[b], len, hipMemcpyHostToDevice, s));
    CDX_HIP_CHECK(hipEventRecord(ev_[b], s));
  }
  // every byte of src is in a pinned block or on the device: src may be
  // reused; the device copies complete in stream order
}

void HostStager::d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  GateShared g;
  if (bytes < kMinStaged) {
    CDX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    CDX_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  ensure();
  const char* in = static_cast<const char*>(src);
  char* out = static_cast<char*>(dst);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t k) {
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    const int b = (int)(k & 1);
    CDX_HIP_CHECK(hipMemcpyAsync(pin_[b], in + off, len, hipMemcpyDeviceToHost, s));
    CDX_HIP_CHECK(hipEventRecord(ev_[b], s));
  };
  issue(0);
  for (size_t k = 0; k < n; ++k) {
    // the other block was emptied in the previous iteration
    if (k + 1 < n) issue(k + 1);
    const int b = (int)(k & 1);
    CDX_HIP_CHECK(hipEventSynchronize(ev_[b]));
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    parallel_memcpy(out + off, pin_[b], len);
  }
}

}  // namespace cdx
