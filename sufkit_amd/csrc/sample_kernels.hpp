// Text-position sampling of a complete index (sampling rate K > 1), on the
// device: the sampled SA keeps the rows whose text position is a multiple of
// K, in SA order; the sampled ISA maps position p (a multiple of K) to its
// sampled row at p / K; the sampled LCP of a kept row is the LCP of its
// suffix and the previous kept row's, i.e. the minimum complete LCP over the
// rows after the previous kept row up to this one (0 for the first kept
// row).  The same arrays the reference's divsufsort adapter computes for K
// (divsufsort_backend.cpp:62-150: residue-zero filter, isa[SA/K] = row,
// Kasai with carry step K).
//
// Three phases over chunks of rows, as the other scans: per-chunk summaries,
// a one-block exclusive scan of the summaries, then each chunk re-walks its
// rows from its carry.  The summary of a run of rows is (kept rows, whether
// one is kept, minimum LCP after the last kept row or over all rows when
// none is kept); summaries compose left to right.
#pragma once

#include "device_common.hpp"

namespace sfk {

struct SampleSummary {
  unsigned long long rows;
  unsigned long long tail;
  unsigned int has;
  unsigned int pad;
};

constexpr unsigned long long kNoLcpYet = ~0ull;
constexpr int kSampleThreads = 256;
constexpr int kSampleItems = 8;
constexpr int kSampleTile = kSampleThreads * kSampleItems;

__device__ __forceinline__ SampleSummary SampleThen(const SampleSummary& a, const SampleSummary& b) {
  SampleSummary c;
  c.rows = a.rows + b.rows;
  c.has = a.has | b.has;
  c.tail = b.has ? b.tail : (a.tail < b.tail ? a.tail : b.tail);
  c.pad = 0;
  return c;
}

__device__ __forceinline__ SampleSummary SampleNone() { return SampleSummary{0ull, kNoLcpYet, 0u, 0u}; }

// Exclusive composition over the threads of a block (Hillis-Steele in LDS);
// `total` receives the whole block's summary.
template <int kBlock>
__device__ SampleSummary BlockExclusiveSample(SampleSummary mine, SampleSummary* lds, SampleSummary& total) {
  const int t = static_cast<int>(threadIdx.x);
  lds[t] = mine;
  __syncthreads();
  for (int d = 1; d < kBlock; d <<= 1) {
    SampleSummary cur = lds[t];
    if (t >= d) cur = SampleThen(lds[t - d], cur);
    __syncthreads();
    lds[t] = cur;
    __syncthreads();
  }
  total = lds[kBlock - 1];
  const SampleSummary out = t == 0 ? SampleNone() : lds[t - 1];
  __syncthreads();
  return out;
}

// Summary of rows [first, first + kSampleItems) below end.
template <typename Index>
__device__ __forceinline__ SampleSummary ThreadSample(const Index* __restrict__ sa, const Index* __restrict__ lcp,
                                                      std::uint64_t first, std::uint64_t end, std::uint64_t k,
                                                      std::uint64_t (&pos)[kSampleItems],
                                                      std::uint64_t (&val)[kSampleItems]) {
  SampleSummary s = SampleNone();
#pragma unroll
  for (int i = 0; i < kSampleItems; ++i) {
    const std::uint64_t r = first + i;
    pos[i] = r < end ? static_cast<std::uint64_t>(sa[r]) : 1;
    val[i] = (r < end && lcp != nullptr) ? static_cast<std::uint64_t>(lcp[r]) : kNoLcpYet;
  }
#pragma unroll
  for (int i = 0; i < kSampleItems; ++i) {
    if (first + i >= end) break;
    s.tail = val[i] < s.tail ? val[i] : s.tail;
    if (pos[i] % k == 0) {
      s.rows += 1;
      s.has = 1;
      s.tail = kNoLcpYet;
    }
  }
  return s;
}

template <typename Index>
__global__ __launch_bounds__(kSampleThreads) void SampleReduceKernel(const Index* __restrict__ sa,
                                                                     const Index* __restrict__ lcp, std::uint64_t n,
                                                                     std::uint64_t k, std::uint64_t chunk,
                                                                     SampleSummary* __restrict__ partial) {
  __shared__ SampleSummary lds[kSampleThreads];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  SampleSummary run = SampleNone();
  for (std::uint64_t base = begin; base < end; base += kSampleTile) {
    std::uint64_t pos[kSampleItems], val[kSampleItems];
    const SampleSummary mine =
        ThreadSample(sa, lcp, base + static_cast<std::uint64_t>(threadIdx.x) * kSampleItems, end, k, pos, val);
    SampleSummary total;
    BlockExclusiveSample<kSampleThreads>(mine, lds, total);
    run = SampleThen(run, total);
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = run;
}

// Exclusive scan of the chunk summaries (grid <= 1024 chunks).
__global__ __launch_bounds__(1024) void SampleSpineKernel(SampleSummary* partial, int grid) {
  __shared__ SampleSummary lds[1024];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  SampleSummary total;
  const SampleSummary excl = BlockExclusiveSample<1024>(valid ? partial[threadIdx.x] : SampleNone(), lds, total);
  if (valid) partial[threadIdx.x] = excl;
}

// Writes the kept rows: out_sa[j] = SA[r], out_isa[SA[r] / k] = j and
// out_lcp[j] (when lcp is given) for the j-th kept row r.
template <typename Index>
__global__ __launch_bounds__(kSampleThreads) void SampleApplyKernel(
    const Index* __restrict__ sa, const Index* __restrict__ lcp, std::uint64_t n, std::uint64_t k,
    std::uint64_t chunk, const SampleSummary* __restrict__ partial, Index* __restrict__ out_sa,
    Index* __restrict__ out_isa, Index* __restrict__ out_lcp) {
  __shared__ SampleSummary lds[kSampleThreads];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  SampleSummary run = partial[blockIdx.x];
  for (std::uint64_t base = begin; base < end; base += kSampleTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kSampleItems;
    std::uint64_t pos[kSampleItems], val[kSampleItems];
    const SampleSummary mine = ThreadSample(sa, lcp, first, end, k, pos, val);
    SampleSummary total;
    const SampleSummary before = SampleThen(run, BlockExclusiveSample<kSampleThreads>(mine, lds, total));
    std::uint64_t row = before.rows;
    std::uint64_t low = before.tail;
#pragma unroll
    for (int i = 0; i < kSampleItems; ++i) {
      if (first + i >= end) break;
      low = val[i] < low ? val[i] : low;
      if (pos[i] % k == 0) {
        out_sa[row] = static_cast<Index>(pos[i]);
        // A position past the text (which the permutation check flags)
        // scatters nothing.
        if (pos[i] < n) out_isa[pos[i] / k] = static_cast<Index>(row);
        if (out_lcp != nullptr) out_lcp[row] = static_cast<Index>(row == 0 ? 0 : low);
        ++row;
        low = kNoLcpYet;
      }
    }
    run = SampleThen(run, total);
  }
}

}  // namespace sfk
