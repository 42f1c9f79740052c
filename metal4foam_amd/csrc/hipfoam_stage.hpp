// hipfoam_stage.hpp -- host <-> device copies of large pageable arrays.
//
// The caller's arrays (OpenFOAM's psi, source and coefficients) are pageable
// host memory.  The runtime copies those at 10-20 GB/s host to device and
// about 6 GB/s device to host; pinned memory moves at about 57 GB/s.  A
// HostStager copies through a ring of pinned chunks instead: the host-side
// memcpy of each chunk is split over a small pool of threads and overlaps the
// DMA of the chunks before it (about 40-46 GB/s with 8 threads and 8 MiB
// chunks, measured on the MI355X host).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace hipfoam {

// A fixed pool of helper threads that runs one job in parts: the calling
// thread runs part 0 itself, the helpers the others.  Used for the host side
// of staged copies (one buffer in equal parts) and for hashing the mesh
// addressing.
class CopyPool {
 public:
  explicit CopyPool(int helpers) {
    for (int k = 0; k < helpers; ++k) th_.emplace_back([this, k] { work(k + 1); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    go_.notify_all();
    for (auto& t : th_) t.join();
  }
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;

  int parts() const { return (int)th_.size() + 1; }

  // fn(k, parts) for k = 0 .. parts-1, returning when every part is done
  void run(const std::function<void(int, int)>& fn) {
    const int parts = (int)th_.size() + 1;
    if (parts == 1) {
      fn(0, 1);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      parts_ = parts;
      pending_ = parts - 1;
      ++job_;
    }
    go_.notify_all();
    fn(0, parts);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

  // memcpy(dst, src, bytes), returning when every part is copied
  void copy(void* dst, const void* src, size_t bytes) {
    if (th_.empty() || bytes < kSplitMin) {
      std::memcpy(dst, src, bytes);
      return;
    }
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    run([&](int k, int parts) {
      const size_t a = bytes * k / parts, e = bytes * (k + 1) / parts;
      std::memcpy(d + a, s + a, e - a);
    });
  }

 private:
  static constexpr size_t kSplitMin = 256u << 10;

  void work(int k) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int, int)>* fn = nullptr;
      int parts = 1;
      {
        std::unique_lock<std::mutex> lk(m_);
        go_.wait(lk, [&] { return stop_ || job_ != seen; });
        if (stop_) return;
        seen = job_;
        fn = fn_;
        parts = parts_;
      }
      (*fn)(k, parts);
      std::lock_guard<std::mutex> lk(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }

  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable go_, done_;
  const std::function<void(int, int)>* fn_ = nullptr;
  int parts_ = 1, pending_ = 0;
  uint64_t job_ = 0;
  bool stop_ = false;
};

// Pinned staging ring of kBuf chunks of kChunk bytes, and the copy pool.
// Copies are enqueued on the stream they are given, so they keep their order
// with the kernels around them.
class HostStager {
 public:
  static constexpr size_t kChunk = 8u << 20;
  static constexpr int kBuf = 3;
  static constexpr size_t kMin = 4u << 20;  // smaller copies: hipMemcpyAsync

  // pool: the owner's copy pool (outlives the stager)
  hipError_t init(CopyPool* pool) {
    for (int b = 0; b < kBuf; ++b) {
      hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&buf_[b]), kChunk, 0);
      if (e != hipSuccess) return e;
      e = hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    pool_ = pool;
    return hipSuccess;
  }

  ~HostStager() {
    for (int b = 0; b < kBuf; ++b) {
      if (ev_[b]) {
        (void)hipEventSynchronize(ev_[b]);
        (void)hipEventDestroy(ev_[b]);
      }
      if (buf_[b]) (void)hipHostFree(buf_[b]);
    }
  }

  // Host to device.  Returns once the host array has been read; the DMA of
  // the last chunks may still run on the stream.
  hipError_t h2d(void* dev, const void* host, size_t bytes, hipStream_t s) {
    const char* src = static_cast<const char*>(host);
    char* dst = static_cast<char*>(dev);
    for (size_t off = 0; off < bytes; off += kChunk) {
      const size_t n = std::min(kChunk, bytes - off);
      const int b = next_;
      next_ = (next_ + 1) % kBuf;
      hipError_t e = hipEventSynchronize(ev_[b]);  // the chunk's last DMA is done
      if (e != hipSuccess) return e;
      pool_->copy(buf_[b], src + off, n);
      e = hipMemcpyAsync(dst + off, buf_[b], n, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return e;
      e = hipEventRecord(ev_[b], s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }

  // Device to host, after the work already on the stream.  Returns once the
  // host array holds the data.
  hipError_t d2h(void* host, const void* dev, size_t bytes, hipStream_t s) {
    char* dst = static_cast<char*>(host);
    const char* src = static_cast<const char*>(dev);
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    int slot[kBuf];
    auto issue = [&](size_t k) -> hipError_t {
      const size_t off = k * kChunk, n = std::min(kChunk, bytes - off);
      const int b = next_;
      next_ = (next_ + 1) % kBuf;
      slot[k % kBuf] = b;
      hipError_t e = hipEventSynchronize(ev_[b]);
      if (e != hipSuccess) return e;
      e = hipMemcpyAsync(buf_[b], src + off, n, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return e;
      return hipEventRecord(ev_[b], s);
    };
    for (size_t k = 0; k < std::min<size_t>(kBuf, nchunks); ++k) {
      hipError_t e = issue(k);
      if (e != hipSuccess) return e;
    }
    for (size_t k = 0; k < nchunks; ++k) {
      const int b = slot[k % kBuf];
      hipError_t e = hipEventSynchronize(ev_[b]);
      if (e != hipSuccess) return e;
      const size_t off = k * kChunk, n = std::min(kChunk, bytes - off);
      pool_->copy(dst + off, buf_[b], n);
      if (k + kBuf < nchunks) {
        e = issue(k + kBuf);
        if (e != hipSuccess) return e;
      }
    }
    return hipSuccess;
  }

 private:
  char* buf_[kBuf] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_[kBuf] = {nullptr, nullptr, nullptr};
  CopyPool* pool_ = nullptr;
  int next_ = 0;
};

}  // namespace hipfoam
