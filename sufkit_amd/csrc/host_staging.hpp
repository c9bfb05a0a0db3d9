// Host <-> device copies of large arrays through pinned staging buffers
// (host code only).
//
// A copy from pageable host memory cannot be read by the DMA engines
// directly, so it is staged: the host copies a chunk into one of two pinned
// buffers (on several threads, which also spreads the page faults of a
// fresh destination) while the engines move the other chunk.  Each device
// has one stager shared by the process; a caller that finds it busy copies
// without staging instead of waiting.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace sfk {

// memcpy split over up to 16 threads for copies of several MiB.
inline void ParallelCopy(void* dst, const void* src, std::uint64_t bytes) {
  constexpr std::uint64_t kPerThread = 4ull << 20;
  const std::uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
  const std::uint64_t threads = std::min<std::uint64_t>(std::min<std::uint64_t>(hw, 16),
                                                        (bytes + kPerThread - 1) / kPerThread);
  if (threads <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const std::uint64_t per = (bytes + threads - 1) / threads;
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (std::uint64_t t = 0; t < threads; ++t) {
    const std::uint64_t off = t * per;
    if (off >= bytes) break;
    const std::uint64_t len = std::min(per, bytes - off);
    pool.emplace_back([=] {
      std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len);
    });
  }
  for (auto& th : pool) th.join();
}

class PinnedStager {
 public:
  static constexpr std::uint64_t kChunk = 64ull << 20;

  PinnedStager() = default;
  PinnedStager(const PinnedStager&) = delete;
  PinnedStager& operator=(const PinnedStager&) = delete;
  ~PinnedStager() {
    for (int k = 0; k < 2; ++k) {
      if (ev_[k] != nullptr) (void)hipEventSynchronize(ev_[k]);
      if (ev_[k] != nullptr) (void)hipEventDestroy(ev_[k]);
      if (buf_[k] != nullptr) (void)hipHostFree(buf_[k]);
    }
  }

  hipError_t Init() {
    for (int k = 0; k < 2; ++k) {
      if (buf_[k] == nullptr) {
        const hipError_t e = hipHostMalloc(&buf_[k], kChunk, hipHostMallocDefault);
        if (e != hipSuccess) return e;
      }
      if (ev_[k] == nullptr) {
        const hipError_t e = hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming);
        if (e != hipSuccess) return e;
      }
    }
    return hipSuccess;
  }

  // Host -> device on `stream`; returns after the data is on the device.
  hipError_t Upload(void* dst, const void* src, std::uint64_t bytes, hipStream_t stream) {
    bool used[2] = {false, false};
    for (std::uint64_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
      const int k = static_cast<int>(i & 1);
      const std::uint64_t len = std::min(kChunk, bytes - off);
      hipError_t e = used[k] ? hipEventSynchronize(ev_[k]) : hipSuccess;
      if (e != hipSuccess) return e;
      ParallelCopy(buf_[k], static_cast<const char*>(src) + off, len);
      e = hipMemcpyAsync(static_cast<char*>(dst) + off, buf_[k], len, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess) e = hipEventRecord(ev_[k], stream);
      if (e != hipSuccess) return e;
      used[k] = true;
    }
    return hipStreamSynchronize(stream);
  }

  // Device -> host after the work already queued on `stream`.
  hipError_t Download(void* dst, const void* src, std::uint64_t bytes, hipStream_t stream) {
    const std::uint64_t chunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](std::uint64_t i) {
      const int k = static_cast<int>(i & 1);
      const std::uint64_t off = i * kChunk;
      const std::uint64_t len = std::min(kChunk, bytes - off);
      hipError_t e = hipMemcpyAsync(buf_[k], static_cast<const char*>(src) + off, len, hipMemcpyDeviceToHost, stream);
      return e == hipSuccess ? hipEventRecord(ev_[k], stream) : e;
    };
    for (std::uint64_t i = 0; i < std::min<std::uint64_t>(chunks, 2); ++i) {
      const hipError_t e = issue(i);
      if (e != hipSuccess) return e;
    }
    for (std::uint64_t i = 0; i < chunks; ++i) {
      const int k = static_cast<int>(i & 1);
      const std::uint64_t off = i * kChunk;
      hipError_t e = hipEventSynchronize(ev_[k]);
      if (e != hipSuccess) return e;
      ParallelCopy(static_cast<char*>(dst) + off, buf_[k], std::min(kChunk, bytes - off));
      if (i + 2 < chunks && (e = issue(i + 2)) != hipSuccess) return e;
    }
    return hipStreamSynchronize(stream);
  }

 private:
  void* buf_[2] = {nullptr, nullptr};
  hipEvent_t ev_[2] = {nullptr, nullptr};
};

// The process's stager of `device`, locked, or an empty lock when another
// thread holds it (or pinned memory could not be had).
struct StagerLease {
  std::unique_lock<std::mutex> lock;
  PinnedStager* stager = nullptr;
};

inline StagerLease LeaseStager(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mutexes[kMaxDevices];
  // Kept until the process ends (not destroyed at exit, when the runtime
  // may already be gone): 128 MiB of pinned memory per device used.
  static PinnedStager* stagers[kMaxDevices] = {};
  StagerLease lease;
  if (device < 0 || device >= kMaxDevices) return lease;
  std::unique_lock<std::mutex> lock(mutexes[device], std::try_to_lock);
  if (!lock.owns_lock()) return lease;
  if (stagers[device] == nullptr) {
    auto s = std::make_unique<PinnedStager>();
    if (s->Init() != hipSuccess) {
      (void)hipGetLastError();
      return lease;
    }
    stagers[device] = s.release();
  }
  lease.lock = std::move(lock);
  lease.stager = stagers[device];
  return lease;
}

// Copies through the device's stager when it is free, directly otherwise
// (small copies always directly).  Synchronous on `stream`.
inline hipError_t StagedCopy(int device, void* dst, const void* src, std::uint64_t bytes, hipMemcpyKind kind,
                             hipStream_t stream) {
  if (bytes >= (4ull << 20)) {
    StagerLease lease = LeaseStager(device);
    if (lease.stager != nullptr) {
      return kind == hipMemcpyHostToDevice ? lease.stager->Upload(dst, src, bytes, stream)
                                           : lease.stager->Download(dst, src, bytes, stream);
    }
  }
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
  return e == hipSuccess ? hipStreamSynchronize(stream) : e;
}

}  // namespace sfk
