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
// That full pass serves builds without early LCP (sampled ones).  A build
// with early LCP computes PLCP only for the rows its doubling left marked
// (marked-row LCP, at the end of this file).
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

// A long LCE left to PlcpLongKernel: the suffixes at i and j agree on their
// first l symbols; the result goes to plcp[dst].
struct LongLce {
  unsigned long long i;
  unsigned long long j;
  unsigned long long l;
  unsigned long long dst;
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
      queue[slot].dst = i;
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
  const std::uint64_t dst = queue[item].dst;
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
    plcp[dst] = static_cast<std::uint32_t>(l);
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

// Exclusive max of the chunk maxima, grid <= kSpineChunks (four per thread).
constexpr int kSpineChunks = 4096;

__global__ __launch_bounds__(1024) void PlcpScanSpineKernel(std::uint64_t* partial, int grid) {
  constexpr int kBlock = 1024;
  constexpr int kPer = kSpineChunks / kBlock;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  std::uint64_t v[kPer];
  std::uint64_t local = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = static_cast<int>(threadIdx.x) * kPer + j;
    v[j] = c < grid ? partial[c] : 0ull;
    local = v[j] > local ? v[j] : local;
  }
  std::uint64_t total;
  std::uint64_t run = BlockExclusiveMax<kBlock>(local, scratch, total);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = static_cast<int>(threadIdx.x) * kPer + j;
    if (c < grid) partial[c] = run;
    run = v[j] > run ? v[j] : run;
  }
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


// ------------------------------------------------------- marked-row LCP --
//
// After a build with early LCP, a row holds the marker only where the
// doubling's keys did not prove its LCP; every row whose LCP is 32 or more is
// such a row, and every other row already holds its final value.  Only the
// marked rows need Φ/PLCP, so instead of a PLCP pass over every text
// position (one random SA line per position) and a gather over every row:
//   LcpMarkCount / LcpMarkWrite   rows in order: the marked rows r >= 1 go to
//                  a list of (position SA[r], Φ = SA[r-1]), whose position
//                  digits are counted for the radix sort (the row is read
//                  back as ISA[position] after the sort).  The same pass
//                  checks that ISA inverts SA (below) and writes LCP[0].
//   (radix sort of the list by position)
//   SparsePlcp     the list in text order: Kasai's reducibility test between
//                  neighbouring entries.  Entry k (position p) is reducible
//                  when entry k-1 is position p-1, p-1 does not start a
//                  bucket, and Φ(p-1) + 1 == Φ(p); then PLCP[p] = PLCP[p-1]
//                  - 1.  Otherwise its LCE is computed from its windows (a
//                  reducible position whose predecessor is unmarked is
//                  computed too: its predecessor's LCP was proven below 32,
//                  so its own is below 31, one window).
//   SparseScan     PLCP[p] + p never decreases along the text, and the
//                  entries form a subsequence of it, so the inclusive max of
//                  PLCP + p + 1 over the irreducible entries up to k is
//                  PLCP[p_k] + p_k + 1 (as in PlcpScan).  The apply reads
//                  r = ISA[p_k] (the positions ascend: one sweep over the
//                  ISA), leaves (r, LCP[r]) pairs and counts the high digits
//                  of r.
//   (radix passes over the high digits of r, ending at its top bit)
//   LcpScatter     LCP[r] = value, pairs grouped by windows of at most 256
//                  rows: the pairs of a window are contiguous, so the stores of one
//                  instruction fall in a few lines instead of one partial
//                  line each at random (pairs grouped only by r >> 16 left
//                  every store its own 32-byte write request).
//
// The check that replaces the inverse check of the full PLCP pass: the pairs
// (r, SA[r]) over rows and (ISA[i], i) over positions are the same multiset
// exactly when SA is a permutation of [0, n) and ISA its inverse (each side
// holds one pair per row and one per position).  Each pair is hashed by two
// independent 64-bit finalizers of its packed 64-bit value (row and position
// both lie below 2^32), and the pass accumulates, per finalizer, the sum over
// rows minus the sum over positions, mod 2^64; the host requires both to be
// zero, and the pass also flags any value at or past n.  A wrong SA or ISA
// passes only if two sums of distinct pseudo-random 64-bit values collide in
// both finalizers.

constexpr int kMarkThreads = 256;
constexpr int kMarkRows = 8;
constexpr int kMarkTile = kMarkThreads * kMarkRows;
// At most this many low row bits stay unsorted before the LCP scatter.
constexpr int kScatterWindowBits = 8;

template <typename T>
__device__ __forceinline__ T LcpMarker() {
  return sizeof(T) == 4 ? static_cast<T>(kLcpUnknown32) : static_cast<T>(kLcpUnknown64);
}

// Eight consecutive elements from k (a multiple of 8) of an array read once:
// 16-byte loads when the array is 16-byte aligned and all eight lie below
// end, else one load each; elements at or past end read as `fill`.
template <typename T>
__device__ __forceinline__ void LoadRun8(const T* __restrict__ a, std::uint64_t k, std::uint64_t end, bool aligned,
                                         T fill, T (&v)[8]) {
  if (aligned && k + 8 <= end) {
    constexpr int kVec = static_cast<int>(8 * sizeof(T) / 16);
    const uint4* q = reinterpret_cast<const uint4*>(a + k);
    uint4 w[kVec];
#pragma unroll
    for (int i = 0; i < kVec; ++i) w[i] = q[i];
    __builtin_memcpy(v, w, sizeof(v));
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = k + i < end ? a[k + i] : fill;
  }
}

// The two pair finalizers: MurmurHash3's fmix64 and SplitMix64's output
// function, applied to (row << 32 | position) and (position << 32 | row).
__device__ __forceinline__ std::uint64_t PairHashA(std::uint64_t row, std::uint64_t pos) {
  std::uint64_t x = ((row << 32) | pos) ^ 0x6a09e667f3bcc909ull;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
__device__ __forceinline__ std::uint64_t PairHashB(std::uint64_t row, std::uint64_t pos) {
  std::uint64_t x = ((pos << 32) | row) + 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Marked rows r >= 1 of each chunk (chunks are whole kMarkTile tiles).
template <typename Out>
__global__ __launch_bounds__(kMarkThreads) void LcpMarkCountKernel(const Out* __restrict__ lcp, std::uint64_t n,
                                                                   std::uint64_t chunk, bool aligned,
                                                                   std::uint64_t* __restrict__ partial) {
  __shared__ std::uint64_t scratch[kMarkThreads / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  const Out marker = LcpMarker<Out>();
  std::uint64_t count = 0;
  for (std::uint64_t r0 = begin + static_cast<std::uint64_t>(threadIdx.x) * kMarkRows; r0 < end;
       r0 += kMarkTile) {
    Out v[8];
    LoadRun8(lcp, r0, end, aligned, Out(0), v);
#pragma unroll
    for (int i = 0; i < 8; ++i) count += (v[i] == marker) & (r0 + i > 0);
  }
  std::uint64_t total;
  BlockExclusiveSum<kMarkThreads>(count, scratch, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive sum of the chunk counts (grid <= kSpineChunks, four per thread),
// the total into *total.
__global__ __launch_bounds__(1024) void MarkSpineKernel(std::uint64_t* partial, int grid,
                                                        unsigned long long* total) {
  constexpr int kBlock = 1024;
  constexpr int kPer = kSpineChunks / kBlock;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  std::uint64_t v[kPer];
  std::uint64_t local = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = static_cast<int>(threadIdx.x) * kPer + j;
    v[j] = c < grid ? partial[c] : std::uint64_t{0};
    local += v[j];
  }
  std::uint64_t sum;
  std::uint64_t run = BlockExclusiveSum<kBlock>(local, scratch, sum);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int c = static_cast<int>(threadIdx.x) * kPer + j;
    if (c < grid) partial[c] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) *total = sum;
}

// The list of marked rows at the chunk's offset, in row order: position
// SA[r] (out_pos) and Φ = SA[r-1] (out_phi), with the position digits
// counted for the sort, the inverse check of SA and ISA (rows and positions
// share the index), and LCP[0] = 0 when row 0 is marked.  The row itself is
// not carried: after the sort by position it is ISA[position].
template <typename R, typename Out>
__global__ __launch_bounds__(kMarkThreads) void LcpMarkWriteKernel(
    Out* __restrict__ lcp, const R* __restrict__ sa, const R* __restrict__ isa, std::uint64_t n,
    std::uint64_t chunk, bool lcp_aligned, bool sa_aligned, bool isa_aligned,
    const std::uint64_t* __restrict__ partial, std::uint32_t* __restrict__ out_pos,
    std::uint32_t* __restrict__ out_phi, RadixPasses passes, RadixSet set, DeviceCounters* counters) {
  constexpr int kWaves = kMarkThreads / kWave;
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][RadixTile::kRadix];
  __shared__ std::uint32_t scratch32[kWaves];
  __shared__ std::uint64_t scratch[kWaves];
  ClearNextRadixSet(set);
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * RadixTile::kRadix; i += kMarkThreads) {
    (&local[0][0][0])[i] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < n ? begin + chunk : n;
  const Out marker = LcpMarker<Out>();
  std::uint64_t at = partial[blockIdx.x];
  std::uint64_t hash_a = 0, hash_b = 0;
  std::uint32_t flags = 0;
  // Every thread runs the same number of tile iterations (block scans).
  for (std::uint64_t base = begin; base < end; base += kMarkTile) {
    const std::uint64_t r0 = base + static_cast<std::uint64_t>(threadIdx.x) * kMarkRows;
    Out l[8];
    R s[8], q[8];
    LoadRun8(lcp, r0, end, lcp_aligned, Out(0), l);
    LoadRun8(sa, r0, end, sa_aligned, R(0), s);
    LoadRun8(isa, r0, end, isa_aligned, R(0), q);
    std::uint32_t marks = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const std::uint64_t r = r0 + i;
      if (r < end) {
        const std::uint64_t pos = static_cast<std::uint64_t>(s[i]);
        const std::uint64_t row = static_cast<std::uint64_t>(q[i]);
        flags |= pos >= n ? static_cast<std::uint32_t>(kFlagNotPermutation) : 0u;
        flags |= row >= n ? static_cast<std::uint32_t>(kFlagNotInverse) : 0u;
        hash_a += PairHashA(r, pos) - PairHashA(row, r);
        hash_b += PairHashB(r, pos) - PairHashB(row, r);
        if (l[i] == marker) {
          if (r > 0) {
            marks |= 1u << i;
          } else {
            lcp[0] = Out(0);
          }
        }
      }
    }
    // Φ of the thread's first row: SA[r0 - 1] (read only when that row is
    // marked; the other rows take their predecessor from s[]).
    const std::uint64_t before = (marks & 1u) ? static_cast<std::uint64_t>(sa[r0 - 1]) : 0ull;
    std::uint32_t total;
    const std::uint32_t off = BlockExclusiveSum<kMarkThreads>(static_cast<std::uint32_t>(__popc(marks)), scratch32,
                                                              total);
    std::uint64_t slot = at + off;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bool m = (marks >> i) & 1u;
      const std::uint32_t pos = static_cast<std::uint32_t>(s[i]);
      if (m) {
        const std::uint64_t phi = i == 0 ? before : static_cast<std::uint64_t>(s[i - 1]);
        out_pos[slot] = pos;
        out_phi[slot] = static_cast<std::uint32_t>(phi);
        ++slot;
      }
      CountDigits<std::uint32_t>(local[wave % kDigitCopies], passes, pos, m, false);
    }
    at += total;
  }
  __syncthreads();
  FlushDigits<kDigitCopies>(local, passes, set.hist);
  std::uint64_t sum_a, sum_b;
  BlockExclusiveSum<kMarkThreads>(hash_a, scratch, sum_a);
  BlockExclusiveSum<kMarkThreads>(hash_b, scratch, sum_b);
  if (threadIdx.x == 0) {
    if (sum_a) atomicAdd(&counters->inverse_hash[0], static_cast<unsigned long long>(sum_a));
    if (sum_b) atomicAdd(&counters->inverse_hash[1], static_cast<unsigned long long>(sum_b));
  }
  if (flags) atomicOr(&counters->error_flags, flags);
}

// PLCP of the sorted list's entries (kPlcpReducible at reducible ones), one
// wave per tile of kPlcpTile entries: lane 0 reads the entry before the tile
// and each lane takes its predecessor by a lane shuffle.  LCEs longer than
// word_limit windows go to the long queue (dst: the entry).
__global__ __launch_bounds__(kPlcpThreads) void SparsePlcpKernel(
    PackedText t, const std::uint32_t* __restrict__ pos, const std::uint32_t* __restrict__ phi, std::uint64_t m,
    const std::uint32_t* __restrict__ starts, std::uint32_t word_limit, std::uint32_t* __restrict__ out,
    LongLce* __restrict__ queue, std::uint64_t queue_capacity, DeviceCounters* counters) {
  constexpr std::uint64_t kWavesPerBlock = kPlcpThreads / kWave;
  const std::uint64_t tiles = (m + kPlcpTile - 1) / kPlcpTile;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kWavesPerBlock;
  const std::uint32_t lane = LaneId();
  for (std::uint64_t tile = static_cast<std::uint64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
       tile < tiles; tile += stride) {  // wave-uniform
    const std::uint64_t first = tile * kPlcpTile;
    const std::uint64_t k = first + lane - 1;  // lane 0: the entry before the tile
    const bool own = lane > 0 && k < m;
    std::uint32_t p = 0xFFFFFFFFu, f = 0xFFFFFFFFu;
    if (own || (lane == 0 && first > 0)) {
      p = pos[k];
      f = phi[k];
    }
    const std::uint32_t pb = WaveShfl(p, static_cast<int>((lane + kWave - 1) & (kWave - 1)));
    const std::uint32_t fb = WaveShfl(f, static_cast<int>((lane + kWave - 1) & (kWave - 1)));
    if (!own) continue;
    if (Bit(starts, p)) {
      out[k] = 0;
      continue;
    }
    if (k > 0 && pb + 1u == p && fb + 1u == f && !Bit(starts, pb)) {
      out[k] = kPlcpReducible;
      continue;
    }
    std::uint64_t l = 0;
    bool done = false;
    for (std::uint32_t w = 0; w < word_limit; ++w) {
      std::uint64_t ca, cb;
      std::uint32_t sa_, sb;
      LoadWindow(t, static_cast<std::uint64_t>(p) + l, ca, sa_);
      LoadWindow(t, static_cast<std::uint64_t>(f) + l, cb, sb);
      const std::uint32_t mm = WindowMismatch(ca, sa_, cb, sb);
      l += mm;
      if (mm < 32) {
        done = true;
        break;
      }
    }
    if (done) {
      out[k] = static_cast<std::uint32_t>(l);
      continue;
    }
    const unsigned long long slot = atomicAdd(&counters->long_items, 1ull);
    if (slot < queue_capacity) {
      queue[slot].i = p;
      queue[slot].j = f;
      queue[slot].l = l;
      queue[slot].dst = k;
    }
    out[k] = 0;  // the long-LCE pass writes the value
  }
}

// Per-chunk maximum scan key of the list (PLCP + position + 1 at irreducible
// entries, 0 at reducible ones); counts the irreducible entries.
__global__ __launch_bounds__(256) void SparseScanReduceKernel(const std::uint32_t* __restrict__ plcp,
                                                              const std::uint32_t* __restrict__ pos, std::uint64_t m,
                                                              std::uint64_t chunk, std::uint64_t* __restrict__ partial,
                                                              DeviceCounters* counters) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t best = 0;
  std::uint64_t irreducible = 0;
  for (std::uint64_t k = begin + threadIdx.x * 8ull; k < end; k += kBlock * 8ull) {
    std::uint32_t v[8], p[8];
    LoadRun8(plcp, k, end, true, kPlcpReducible, v);
    LoadRun8(pos, k, end, true, 0u, p);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const std::uint64_t key = PlcpKey(v[i], p[i]);
      best = key > best ? key : best;
      irreducible += key != 0;
    }
  }
  std::uint64_t total;
  BlockExclusiveMax<kBlock>(best, scratch, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
  BlockAddTo<kBlock>(irreducible, scratch, &counters->irreducible);
}

// Inclusive max-scan of the list's keys: entry k's LCP row r_k = ISA[p_k]
// (the positions ascend, so these reads sweep the ISA once) goes to rows[k]
// and its value, scan - p_k - 1, back to plcp[k]; the row digits of the
// scatter's passes are counted.
template <typename R>
__global__ __launch_bounds__(256) void SparseScanApplyKernel(std::uint32_t* __restrict__ plcp,
                                                             const std::uint32_t* __restrict__ pos,
                                                             const R* __restrict__ isa, std::uint64_t m,
                                                             std::uint64_t chunk,
                                                             const std::uint64_t* __restrict__ partial,
                                                             std::uint32_t* __restrict__ rows, RadixPasses passes,
                                                             RadixSet set) {
  constexpr int kBlock = 256;
  constexpr int kItems = 8;
  constexpr int kTile = kBlock * kItems;
  constexpr int kWaves = kBlock / kWave;
  __shared__ std::uint64_t scratch[kWaves];
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][RadixTile::kRadix];
  ClearNextRadixSet(set);
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * RadixTile::kRadix; i += kBlock) {
    (&local[0][0][0])[i] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t run = partial[blockIdx.x];
  for (std::uint64_t base = begin; base < end; base += kTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    std::uint32_t v[kItems], p[kItems], rv[kItems];
    LoadRun8(plcp, first, end, true, kPlcpReducible, v);
    LoadRun8(pos, first, end, true, 0u, p);
#pragma unroll
    for (int i = 0; i < kItems; ++i) rv[i] = first + i < end ? static_cast<std::uint32_t>(isa[p[i]]) : 0u;
    std::uint64_t keys[kItems];
    std::uint64_t local_max = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      keys[i] = first + i < end ? PlcpKey(v[i], p[i]) : 0ull;
      local_max = keys[i] > local_max ? keys[i] : local_max;
    }
    std::uint64_t total;
    std::uint64_t cur = BlockExclusiveMax<kBlock>(local_max, scratch, total);
    if (run > cur) cur = run;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      cur = keys[i] > cur ? keys[i] : cur;
      v[i] = static_cast<std::uint32_t>(cur - p[i] - 1);
    }
    if (first + kItems <= end) {
      *reinterpret_cast<uint4*>(plcp + first) = make_uint4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<uint4*>(plcp + first + 4) = make_uint4(v[4], v[5], v[6], v[7]);
      *reinterpret_cast<uint4*>(rows + first) = make_uint4(rv[0], rv[1], rv[2], rv[3]);
      *reinterpret_cast<uint4*>(rows + first + 4) = make_uint4(rv[4], rv[5], rv[6], rv[7]);
    } else {
#pragma unroll
      for (int i = 0; i < kItems; ++i) {
        if (first + i < end) {
          plcp[first + i] = v[i];
          rows[first + i] = rv[i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      CountDigits<std::uint32_t>(local[wave % kDigitCopies], passes, rv[i], first + i < end, false);
    }
    if (total > run) run = total;
  }
  __syncthreads();
  FlushDigits<kDigitCopies>(local, passes, set.hist);
}

// LCP[rows[k]] = values[k] for the m pairs, in order (grouped by windows of
// at most 2^kScatterWindowBits rows, so the stores of one instruction fall in
// a few lines).
template <typename Out>
__global__ __launch_bounds__(256) void LcpScatterKernel(const std::uint32_t* __restrict__ rows,
                                                        const std::uint32_t* __restrict__ values, std::uint64_t m,
                                                        std::uint64_t n, Out* __restrict__ lcp,
                                                        DeviceCounters* counters) {
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * 256 * 4;
  bool bad = false;
  for (std::uint64_t k0 = static_cast<std::uint64_t>(blockIdx.x) * 256 * 4 + threadIdx.x; k0 < m; k0 += stride) {
    std::uint32_t r[4], v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const std::uint64_t k = k0 + static_cast<std::uint64_t>(j) * 256;
      r[j] = k < m ? __builtin_nontemporal_load(&rows[k]) : 0xFFFFFFFFu;
      v[j] = k < m ? __builtin_nontemporal_load(&values[k]) : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const std::uint64_t k = k0 + static_cast<std::uint64_t>(j) * 256;
      if (k >= m) continue;
      if (r[j] < n) {
        lcp[r[j]] = static_cast<Out>(v[j]);
      } else {
        bad = true;
      }
    }
  }
  if (bad) atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
}

}  // namespace sfk
