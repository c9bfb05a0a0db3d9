// LCP of a complete suffix array on the device: Φ/PLCP (Kärkkäinen,
// Manzini and Puglisi, "Permuted longest-common-prefix array", CPM 2009)
// with 64-bit XOR+clz window compares, then LCP[r] = PLCP[SA[r]].
//
//   BucketStarts   the suffixes that start a first-symbol bucket (at most 7
//                  per reference, found by a binary search per reference
//                  and symbol over the SA), marked in a bitmap over text
//                  positions.  Their PLCP is 0: the predecessor starts with
//                  another symbol (or there is none: the sentinel).
//   Plcp           text positions in order, one tile of 63 consecutive
//                  positions per wave.  Φ(i) = SA[ISA[i] - 1] is gathered
//                  (one random SA line, which also holds SA[ISA[i]] for the
//                  inverse check); a lane reads Φ(i-1) from its neighbour.
//                  Position i is reducible when Φ(i-1) + 1 == Φ(i) and
//                  i - 1 does not start a bucket: then the suffixes i-1 and
//                  Φ(i) - 1 are adjacent and start with the same symbol, so
//                  T[i-1] == T[Φ(i)-1] and PLCP[i] = PLCP[i-1] - 1.  This is
//                  Kärkkäinen's reducibility test decided without reading the
//                  text at random; only irreducible positions compare
//                  32-symbol windows at i and Φ(i).
//   PlcpScan       PLCP[i] + i never decreases along the text, so the
//                  inclusive max over irreducible j <= i of PLCP[j] + j is
//                  PLCP[i] + i at every i (three phases over chunks).
//   LcpGather      LCP[r] = PLCP[SA[r]].
//
// Random writes cost far more than random reads on this memory system (a
// Φ built by scattering Φ[SA[r]] = SA[r-1] took as long as the whole PLCP
// pass), so every random access here is a read.  PLCP values are 32-bit:
// every text the backend takes has fewer than 2^32 symbols.  The same pass
// checks ISA[i] < n and SA[ISA[i]] == i for every i, which makes ISA a
// bijection and SA its inverse (the permutation check of RepackSuffixArray,
// suffix_array.cpp:852-881).
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"
#include "sa_kernels.hpp"

namespace sfk {

constexpr std::uint32_t kPlcpReducible = 0xFFFFFFFFu;
constexpr int kPlcpThreads = 256;

struct LongLce {
  unsigned long long i;
  unsigned long long j;
  unsigned long long l;
};

template <int kBlock>
__device__ __forceinline__ void BlockAddTo(std::uint64_t v, std::uint64_t* scratch, unsigned long long* out) {
  std::uint64_t total;
  BlockExclusiveSum<kBlock>(v, scratch, total);
  if (threadIdx.x == 0 && total != 0) atomicAdd(out, static_cast<unsigned long long>(total));
}

__device__ __forceinline__ bool Bit(const std::uint32_t* bits, std::uint64_t p) {
  return (bits[p >> 5] >> (p & 31)) & 1u;
}

// One thread per (reference, symbol 0..6): the first row of the reference's
// row block whose suffix starts with a symbol >= c starts a bucket (c = 0:
// the block's first row, its sentinel); its suffix gets bit SA[row] in the
// position bitmap.  refs == 0: one text of n symbols.
template <typename Index>
__global__ __launch_bounds__(64) void BucketStartsKernel(const std::uint8_t* __restrict__ text,
                                                         const Index* __restrict__ sa, std::uint64_t n,
                                                         const std::uint64_t* __restrict__ offsets,
                                                         std::uint32_t refs, std::uint32_t* __restrict__ bits) {
  const std::uint64_t t = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const std::uint64_t blocks = refs == 0 ? 1 : refs;
  if (t >= blocks * 7) return;
  const std::uint64_t ref = t / 7;
  const std::uint32_t c = static_cast<std::uint32_t>(t % 7);
  const std::uint64_t lo0 = refs == 0 ? 0 : offsets[ref];
  const std::uint64_t hi0 = refs == 0 ? n : offsets[ref + 1];
  std::uint64_t lo = lo0;
  if (c > 0) {
    std::uint64_t hi = hi0;
    while (lo < hi) {
      const std::uint64_t mid = lo + (hi - lo) / 2;
      const std::uint64_t p = static_cast<std::uint64_t>(sa[mid]);
      const std::uint32_t sym = p < n ? text[p] : 0xFFu;  // an out-of-range SA value fails the checks later
      if (sym >= c) {
        hi = mid;
      } else {
        lo = mid + 1;
      }
    }
  }
  if (lo < hi0) {
    const std::uint64_t p = static_cast<std::uint64_t>(sa[lo]);
    if (p < n) atomicOr(&bits[p >> 5], 1u << (p & 31));
  }
}

// Φ(i), or kPhiNone when suffix i starts a bucket; flags a row outside the
// text or an SA that does not invert the ISA.
constexpr std::uint64_t kPhiNone = ~0ull;

template <typename Index>
__device__ __forceinline__ std::uint64_t PhiOf(const Index* __restrict__ sa, const Index* __restrict__ isa,
                                               const std::uint32_t* __restrict__ starts, std::uint64_t n,
                                               std::uint64_t i, bool& bad) {
  const std::uint64_t r = static_cast<std::uint64_t>(__builtin_nontemporal_load(&isa[i]));
  if (r >= n) {
    bad = true;
    return kPhiNone;
  }
  const std::uint64_t self = static_cast<std::uint64_t>(sa[r]);
  if (self != i) bad = true;
  if (Bit(starts, i) || r == 0) return kPhiNone;
  const std::uint64_t j = static_cast<std::uint64_t>(sa[r - 1]);
  if (j >= n) {
    bad = true;
    return kPhiNone;
  }
  return j;
}

// PLCP of irreducible positions (kPlcpReducible at reducible ones) with the
// permutation/inverse check.  One wave per tile of kPlcpTile = 63 positions:
// lane 0 gathers Φ of the position before the tile (its checks belong to the
// previous tile's last lane), lanes 1..63 those of the tile, and each lane
// takes its predecessor's Φ by a lane shuffle, so no block barrier holds the
// waves together.  LCEs longer than word_limit windows go to the long queue.
constexpr std::uint64_t kPlcpTile = kWave - 1;

template <typename Index>
__global__ __launch_bounds__(kPlcpThreads) void PlcpKernel(PackedText t, const Index* __restrict__ sa,
                                                           const Index* __restrict__ isa,
                                                           const std::uint32_t* __restrict__ starts, std::uint64_t n,
                                                           std::uint32_t word_limit, std::uint32_t* __restrict__ plcp,
                                                           LongLce* __restrict__ queue, std::uint64_t queue_capacity,
                                                           DeviceCounters* counters) {
  constexpr std::uint64_t kWavesPerBlock = kPlcpThreads / kWave;
  const std::uint64_t tiles = (n + kPlcpTile - 1) / kPlcpTile;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kWavesPerBlock;
  const std::uint32_t lane = LaneId();
  bool bad = false;
  for (std::uint64_t tile = static_cast<std::uint64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
       tile < tiles; tile += stride) {  // wave-uniform
    const std::uint64_t first = tile * kPlcpTile;
    const std::uint64_t i = first + lane - 1;  // lane 0: the position before the tile
    const bool own = lane > 0 && i < n;
    std::uint64_t phi = kPhiNone;
    bool checked = false;
    if (own || (lane == 0 && first > 0)) phi = PhiOf(sa, isa, starts, n, i, checked);
    if (own) bad |= checked;
    const std::uint64_t before = WaveShfl(phi, static_cast<int>((lane + kWave - 1) & (kWave - 1)));
    if (!own) continue;
    if (phi == kPhiNone) {
      plcp[i] = 0;
      continue;
    }
    if (before != kPhiNone && before + 1 == phi) {
      plcp[i] = kPlcpReducible;
      continue;
    }
    std::uint64_t l = 0;
    bool done = false;
    for (std::uint32_t w = 0; w < word_limit; ++w) {
      std::uint64_t ca, cb;
      std::uint32_t sa_, sb;
      LoadWindow(t, i + l, ca, sa_);
      LoadWindow(t, phi + l, cb, sb);
      const std::uint32_t mm = WindowMismatch(ca, sa_, cb, sb);
      l += mm;
      if (mm < 32) {
        done = true;
        break;
      }
    }
    if (done) {
      plcp[i] = static_cast<std::uint32_t>(l);
      continue;
    }
    const unsigned long long slot = atomicAdd(&counters->long_items, 1ull);
    if (slot < queue_capacity) {
      queue[slot].i = i;
      queue[slot].j = phi;
      queue[slot].l = l;
    }
    plcp[i] = 0;  // the long-LCE pass writes the value
  }
  if (bad) atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagNotInverse));
}

// One wave per queued long LCE: 64 windows (2048 symbols) compared per step.
__global__ __launch_bounds__(256) void PlcpLongKernel(PackedText t, std::uint64_t n,
                                                      const LongLce* __restrict__ queue, std::uint64_t count,
                                                      std::uint32_t* __restrict__ plcp, DeviceCounters* counters) {
  const std::uint64_t item = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (item >= count) return;  // wave-uniform
  const std::uint32_t lane = LaneId();
  const std::uint64_t i = queue[item].i;
  const std::uint64_t j = queue[item].j;
  std::uint64_t l = queue[item].l;
  unsigned long long steps = 0;
  while (true) {
    const std::uint64_t pa = i + l + 32ull * lane;
    const std::uint64_t pb = j + l + 32ull * lane;
    std::uint32_t mm = 0;
    if (pa < n && pb < n) {
      std::uint64_t ca, cb;
      std::uint32_t sa_, sb;
      LoadWindow(t, pa, ca, sa_);
      LoadWindow(t, pb, cb, sb);
      mm = WindowMismatch(ca, sa_, cb, sb);
    }
    const std::uint64_t vote = __ballot(mm < 32);
    ++steps;
    if (vote) {
      const int f = __ffsll(static_cast<unsigned long long>(vote)) - 1;
      const std::uint32_t mf = __shfl(mm, f, kWave);
      l += 32ull * f + mf;
      break;
    }
    l += 32ull * kWave;
  }
  if (lane == 0) {
    plcp[i] = static_cast<std::uint32_t>(l);
    atomicAdd(&counters->long_words, steps * kWave);
  }
}

// The scan key of position k: PLCP[k] + k + 1 at irreducible positions, 0
// at reducible ones (every key is at least 1 where it counts, so 0 is the
// identity of the max).
__device__ __forceinline__ std::uint64_t PlcpKey(std::uint32_t v, std::uint64_t k) {
  return v == kPlcpReducible ? 0ull : static_cast<std::uint64_t>(v) + k + 1;
}

// Eight consecutive PLCP values from slot k (k a multiple of 8), as two
// 16-byte loads when all eight lie below end.
__device__ __forceinline__ void LoadPlcp8(const std::uint32_t* __restrict__ plcp, std::uint64_t k, std::uint64_t end,
                                          std::uint32_t (&v)[8]) {
  if (k + 8 <= end) {
    const uint4 a = *reinterpret_cast<const uint4*>(plcp + k);
    const uint4 b = *reinterpret_cast<const uint4*>(plcp + k + 4);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = k + i < end ? plcp[k + i] : kPlcpReducible;
  }
}

// Per-chunk maximum key; also counts the irreducible positions (nonzero
// keys) into counters->irreducible.  Chunks start at multiples of 8 slots.
__global__ __launch_bounds__(256) void PlcpScanReduceKernel(const std::uint32_t* __restrict__ plcp,
                                                            std::uint64_t n, std::uint64_t chunk,
                                                            std::uint64_t* __restrict__ partial,
                                                            DeviceCounters* counters) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  std::uint64_t best = 0;
  std::uint64_t irreducible = 0;
  for (std::uint64_t k = begin + threadIdx.x * 8ull; k < end; k += kBlock * 8ull) {
    std::uint32_t v[8];
    LoadPlcp8(plcp, k, end, v);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const std::uint64_t key = PlcpKey(v[i], k + i);
      best = key > best ? key : best;
      irreducible += key != 0;
    }
  }
  std::uint64_t total;
  BlockExclusiveMax<kBlock>(best, scratch, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
  BlockAddTo<kBlock>(irreducible, scratch, &counters->irreducible);
}

__global__ __launch_bounds__(1024) void PlcpScanSpineKernel(std::uint64_t* partial, int grid) {
  constexpr int kBlock = 1024;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  std::uint64_t total;
  const std::uint64_t excl = BlockExclusiveMax<kBlock>(valid ? partial[threadIdx.x] : 0ull, scratch, total);
  if (valid) partial[threadIdx.x] = excl;
}

// Inclusive max-scan of the keys, written back as PLCP[k] = scan - k - 1.
// Each thread owns 8 consecutive slots (two 16-byte loads and stores).
__global__ __launch_bounds__(256) void PlcpScanApplyKernel(std::uint32_t* __restrict__ plcp, std::uint64_t n,
                                                           std::uint64_t chunk,
                                                           const std::uint64_t* __restrict__ partial) {
  constexpr int kBlock = 256;
  constexpr int kItems = 8;
  constexpr int kTile = kBlock * kItems;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  std::uint64_t run = partial[blockIdx.x];
  for (std::uint64_t base = begin; base < end; base += kTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    std::uint32_t v[kItems];
    LoadPlcp8(plcp, first, end, v);
    std::uint64_t keys[kItems];
    std::uint64_t local = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      keys[i] = first + i < end ? PlcpKey(v[i], first + i) : 0ull;
      local = keys[i] > local ? keys[i] : local;
    }
    std::uint64_t total;
    std::uint64_t cur = BlockExclusiveMax<kBlock>(local, scratch, total);
    if (run > cur) cur = run;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      cur = keys[i] > cur ? keys[i] : cur;
      v[i] = static_cast<std::uint32_t>(cur - (first + i) - 1);
    }
    if (first + kItems <= end) {
      *reinterpret_cast<uint4*>(plcp + first) = make_uint4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<uint4*>(plcp + first + 4) = make_uint4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int i = 0; i < kItems; ++i) {
        if (first + i < end) plcp[first + i] = v[i];
      }
    }
    if (total > run) run = total;
  }
}

// LCP[r] = PLCP[SA[r]] in the requested output width; LCP[0] = 0.  Each
// thread takes kGatherRows consecutive rows (a grid-stride loop only past
// 2^32 threads): their SA (and LCP) values first, then every PLCP gather its
// rows need, so that several random reads are in flight per thread.  early:
// the LCP array already holds the rows the doubling fixed (EarlyLcp) and the
// marker elsewhere; only marked rows are gathered, and their number is
// counted.
constexpr int kGatherRows = 8;

template <typename Index, typename Out>
__global__ __launch_bounds__(256) void LcpGatherKernel(const std::uint32_t* __restrict__ plcp,
                                                       const Index* __restrict__ sa, std::uint64_t n,
                                                       Out* __restrict__ lcp, bool early,
                                                       DeviceCounters* counters) {
  constexpr Out kUnknown = sizeof(Out) == 4 ? static_cast<Out>(kLcpUnknown32) : static_cast<Out>(kLcpUnknown64);
  __shared__ std::uint32_t s_gathered;
  if (threadIdx.x == 0) s_gathered = 0;
  __syncthreads();
  std::uint32_t gathered = 0;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * 256 * kGatherRows;
  for (std::uint64_t r0 = (static_cast<std::uint64_t>(blockIdx.x) * 256 + threadIdx.x) * kGatherRows; r0 < n;
       r0 += stride) {
    // SA and LCP stream through once: non-temporal, so that they do not
    // evict the randomly read PLCP lines.  An out-of-range SA value (the
    // checks have flagged it) reads nothing.
    std::uint64_t v[kGatherRows];
    Out out[kGatherRows];
    bool need[kGatherRows];
#pragma unroll
    for (int i = 0; i < kGatherRows; ++i) {
      const std::uint64_t r = r0 + i;
      v[i] = r < n ? static_cast<std::uint64_t>(__builtin_nontemporal_load(&sa[r])) : n;
      out[i] = (r < n && early) ? __builtin_nontemporal_load(&lcp[r]) : kUnknown;
    }
#pragma unroll
    for (int i = 0; i < kGatherRows; ++i) {
      const std::uint64_t r = r0 + i;
      need[i] = r < n && out[i] == kUnknown;
      if (need[i]) out[i] = (r == 0 || v[i] >= n) ? Out(0) : static_cast<Out>(plcp[v[i]]);
    }
#pragma unroll
    for (int i = 0; i < kGatherRows; ++i) {
      if (need[i]) {
        __builtin_nontemporal_store(out[i], &lcp[r0 + i]);
        ++gathered;
      }
    }
  }
  if (early) {
    if (gathered != 0) atomicAdd(&s_gathered, gathered);
    __syncthreads();
    if (threadIdx.x == 0 && s_gathered != 0) {
      atomicAdd(&counters->lcp_gathered[blockIdx.x & 15u], static_cast<unsigned long long>(s_gathered));
    }
  }
}

}  // namespace sfk
