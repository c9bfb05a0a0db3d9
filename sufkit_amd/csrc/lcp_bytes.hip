// Byte-coded LCP of the Low-memory profile (lcp_storage.cpp:233-420), built
// on the device from a complete raw LCP (SA order) and its ISA.
//
// Row r keeps min(LCP[r], 255) in one byte.  A value of 255 or more is
// recovered from an overflow anchor: anchors are (text position p, PLCP[p])
// in increasing p, one for each p whose PLCP[p] >= 255 does not simply
// continue the run of position p - 1 (PLCP[p - 1] >= 255 and
// PLCP[p - 1] - 1 == PLCP[p]); PLCP[p] = LCP[ISA[p]].  A row then decodes
// as anchor value - (p - anchor position) from the last anchor at or
// before its suffix.  The anchors are few (long repeats only), so they are
// found in two passes over the text positions: per-tile counts, then a
// second pass that writes only in tiles that hold an anchor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "block_scan.hpp"
#include "device_common.hpp"
#include "sufkit_hip.h"

namespace sfk {
namespace {

constexpr std::uint32_t kByteLimit = 255;
constexpr int kAnchorThreads = 256;
constexpr int kAnchorItems = 16;
constexpr std::uint64_t kAnchorTile = static_cast<std::uint64_t>(kAnchorThreads) * kAnchorItems;

struct ByteFailure {
  int status;
  std::string message;
};

void Check(hipError_t err, const char* what) {
  if (err != hipSuccess) {
    (void)hipGetLastError();
    throw ByteFailure{SFK_HIP_BUILD_FAILURE, std::string(what) + ": " + hipGetErrorString(err)};
  }
}

template <typename T>
__device__ __forceinline__ std::uint64_t Get(const void* p, std::uint64_t i) {
  return static_cast<std::uint64_t>(static_cast<const T*>(p)[i]);
}

__global__ __launch_bounds__(256) void PrimaryBytesKernel(const void* __restrict__ lcp, bool wide, std::uint64_t n,
                                                          std::uint8_t* __restrict__ primary) {
  const std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const std::uint64_t v = wide ? Get<std::uint64_t>(lcp, r) : Get<std::uint32_t>(lcp, r);
  primary[r] = static_cast<std::uint8_t>(v < kByteLimit ? v : kByteLimit);
}

// PLCP[p] = LCP[ISA[p]] (0 for an ISA value outside the rows, which a
// validated ISA never holds).
__device__ __forceinline__ std::uint64_t Plcp(const void* lcp, const void* isa, bool lcp_wide, bool isa_wide,
                                              std::uint64_t n, std::uint64_t p) {
  const std::uint64_t row = isa_wide ? Get<std::uint64_t>(isa, p) : Get<std::uint32_t>(isa, p);
  if (row >= n) return 0;
  return lcp_wide ? Get<std::uint64_t>(lcp, row) : Get<std::uint32_t>(lcp, row);
}

// Whether position p starts an anchor, given PLCP[p] and PLCP[p - 1].
__device__ __forceinline__ bool StartsAnchor(std::uint64_t p, std::uint64_t v, std::uint64_t prev) {
  if (v < kByteLimit) return false;
  return !(p > 0 && prev >= kByteLimit && prev - 1 == v);
}

// Thread t of tile b handles positions b * tile + t * items + [0, items):
// PLCP of its run plus the one before it, so every flag is local.
template <bool kWrite>
__global__ __launch_bounds__(kAnchorThreads) void AnchorKernel(const void* __restrict__ lcp,
                                                               const void* __restrict__ isa, bool lcp_wide,
                                                               bool isa_wide, std::uint64_t n,
                                                               unsigned long long* __restrict__ tile_counts,
                                                               std::uint64_t* __restrict__ positions,
                                                               std::uint64_t* __restrict__ values,
                                                               std::uint64_t rows) {
  __shared__ std::uint64_t scan[kAnchorThreads / kWave];
  const std::uint64_t b = blockIdx.x;
  std::uint64_t base = 0;
  if constexpr (kWrite) {
    // Tiles without anchors write nothing (their count is that of the next
    // tile's start minus their own).
    base = tile_counts[b];
    if (tile_counts[b + 1] == base) return;
  }
  const std::uint64_t first = b * kAnchorTile + static_cast<std::uint64_t>(threadIdx.x) * kAnchorItems;
  std::uint64_t prev = first > 0 && first - 1 < n ? Plcp(lcp, isa, lcp_wide, isa_wide, n, first - 1) : 0;
  std::uint64_t v[kAnchorItems];
#pragma unroll
  for (int j = 0; j < kAnchorItems; ++j) {
    const std::uint64_t p = first + j;
    v[j] = p < n ? Plcp(lcp, isa, lcp_wide, isa_wide, n, p) : 0;
  }
  std::uint32_t mask = 0;
#pragma unroll
  for (int j = 0; j < kAnchorItems; ++j) {
    const std::uint64_t p = first + j;
    if (p < n && StartsAnchor(p, v[j], prev)) mask |= 1u << j;
    prev = v[j];
  }
  const std::uint64_t mine = static_cast<std::uint64_t>(__popc(mask));
  std::uint64_t total = 0;
  const std::uint64_t before = BlockExclusiveSum<kAnchorThreads>(mine, scan, total);
  if constexpr (!kWrite) {
    if (threadIdx.x == 0) tile_counts[b] = total;
  } else {
    std::uint64_t at = base + before;
#pragma unroll
    for (int j = 0; j < kAnchorItems; ++j) {
      if (mask & (1u << j)) {
        if (at < rows) {
          positions[at] = first + j;
          values[at] = v[j];
        }
        ++at;
      }
    }
  }
}

}  // namespace
}  // namespace sfk

extern "C" {

int sfk_hip_byte_coded_lcp(int device, const void* d_lcp, uint32_t lcp_bits, const void* d_isa, uint32_t isa_bits,
                           uint64_t n, uint8_t* primary, uint64_t* anchor_count, uint64_t* anchor_positions,
                           uint64_t* anchor_values, uint64_t capacity, char* error, size_t error_capacity) {
  auto fail = [&](int status, const std::string& message) {
    if (error != nullptr && error_capacity > 0) {
      const std::size_t len = std::min(message.size(), error_capacity - 1);
      std::memcpy(error, message.data(), len);
      error[len] = '\0';
    }
    return status;
  };
  if (d_lcp == nullptr || d_isa == nullptr || anchor_count == nullptr || n == 0 ||
      (lcp_bits != 32 && lcp_bits != 64) || (isa_bits != 32 && isa_bits != 64)) {
    return fail(SFK_HIP_INVALID_INPUT, "LCP and ISA (32 or 64 bits), n > 0 and a count word are required");
  }
  if (n > SFK_HIP_MAX_TEXT_SYMBOLS) {
    return fail(SFK_HIP_UNSUPPORTED_BACKEND, "texts of more than " + std::to_string(SFK_HIP_MAX_TEXT_SYMBOLS) +
                                                 " symbols are not supported by the HIP backend");
  }
  const std::uint64_t tiles = (n + sfk::kAnchorTile - 1) / sfk::kAnchorTile;
  void* d_primary = nullptr;
  unsigned long long* d_counts = nullptr;
  std::uint64_t* d_anchors = nullptr;
  hipStream_t stream = nullptr;
  int status = SFK_HIP_OK;
  try {
    using sfk::Check;
    Check(hipSetDevice(device), "hipSetDevice");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    const bool lcp_wide = lcp_bits == 64, isa_wide = isa_bits == 64;
    if (primary != nullptr) {
      Check(hipMalloc(&d_primary, n), "hipMalloc");
      sfk::PrimaryBytesKernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, stream>>>(d_lcp, lcp_wide, n,
                                                                                     static_cast<std::uint8_t*>(
                                                                                         d_primary));
      Check(hipGetLastError(), "primary bytes launch");
      Check(hipMemcpyAsync(primary, d_primary, n, hipMemcpyDeviceToHost, stream), "primary bytes copy");
    }
    Check(hipMalloc(reinterpret_cast<void**>(&d_counts), (tiles + 1) * sizeof(unsigned long long)), "hipMalloc");
    sfk::AnchorKernel<false><<<static_cast<unsigned>(tiles), sfk::kAnchorThreads, 0, stream>>>(
        d_lcp, d_isa, lcp_wide, isa_wide, n, d_counts, nullptr, nullptr, 0);
    Check(hipGetLastError(), "anchor count launch");
    std::vector<unsigned long long> counts(tiles + 1, 0);
    Check(hipMemcpyAsync(counts.data(), d_counts, tiles * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         stream),
          "anchor counts copy");
    Check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // Exclusive starts, with the total at [tiles].
    unsigned long long sum = 0;
    for (std::uint64_t t = 0; t <= tiles; ++t) {
      const unsigned long long c = t < tiles ? counts[t] : 0;
      counts[t] = sum;
      sum += c;
    }
    *anchor_count = sum;
    if (anchor_positions != nullptr && anchor_values != nullptr && sum > 0) {
      if (capacity < sum) throw sfk::ByteFailure{SFK_HIP_INVALID_INPUT, "anchor arrays are smaller than the count"};
      Check(hipMemcpyAsync(d_counts, counts.data(), (tiles + 1) * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, stream),
            "anchor starts copy");
      Check(hipMalloc(reinterpret_cast<void**>(&d_anchors), 2 * sum * sizeof(std::uint64_t)), "hipMalloc");
      sfk::AnchorKernel<true><<<static_cast<unsigned>(tiles), sfk::kAnchorThreads, 0, stream>>>(
          d_lcp, d_isa, lcp_wide, isa_wide, n, d_counts, d_anchors, d_anchors + sum, sum);
      Check(hipGetLastError(), "anchor write launch");
      Check(hipMemcpyAsync(anchor_positions, d_anchors, sum * sizeof(std::uint64_t), hipMemcpyDeviceToHost, stream),
            "anchor positions copy");
      Check(hipMemcpyAsync(anchor_values, d_anchors + sum, sum * sizeof(std::uint64_t), hipMemcpyDeviceToHost,
                           stream),
            "anchor values copy");
    }
    Check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  } catch (const sfk::ByteFailure& f) {
    status = fail(f.status, f.message);
  }
  if (stream != nullptr) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (d_primary != nullptr) (void)hipFree(d_primary);
  if (d_counts != nullptr) (void)hipFree(d_counts);
  if (d_anchors != nullptr) (void)hipFree(d_anchors);
  return status;
}

}  // extern "C"
