// Kernels of the suffix-index construction path (gfx950).
//
// Prefix doubling over "segments": a segment is a run of SA rows whose
// suffixes share their first h symbols (a group) and still hold more than
// one suffix.  Active suffixes live in one of two lists:
//
//   large list  segments of more than kLocalMaxSegment suffixes.  Per suffix
//               its text position (act_idx) and dense 32-bit segment id
//               (act_seg); per segment seg_delta = (first SA row) - (first
//               slot), so the row of slot k is seg_delta[seg] + k, computed
//               modulo the index width (a segment placed after segments with
//               larger rows has a "negative" delta).  One step:
//     Keygen        key = (segment id << sec_bits) | secondary(suffix)
//     radix sort    stable LSD over the digits of the key
//     RefineApply   group boundaries (key change), singletons written to SA
//                   and rank (= their final ISA entry), survivors compacted
//                   into a new list with renumbered segments, ranks of
//                   survivors set to the SA row where their group starts
//     SplitCount/SplitSpine/SplitAssign/SplitMove
//                   segments of at most kLocalMaxSegment suffixes move to the
//                   small list (while it has room), the others are renumbered
//                   in slot order
//   small list  segments of at most kLocalMaxSegment suffixes; per suffix its
//               text position and the first SA row of its segment.  One
//               LocalRefine kernel ranks every suffix inside its segment in
//               LDS (count of smaller and equal keys), writes singletons,
//               and appends surviving groups to the next small list.
//
// Within a doubling round every rank[suffix + h] is gathered (Keygen,
// QKeygen) before any rank is rewritten (LocalRefine, RefineApply).
// The rank array is indexed by text position; when no suffix is left active
// it equals the ISA.
#pragma once

#include <type_traits>

#include "block_scan.hpp"
#include "device_common.hpp"
#include "radix_sort.hpp"

namespace sfk {

enum SecondaryKind : int {
  kSecFillHi = 0,   // high 32 bits of FillKey (symbols 0..15)
  kSecFillLo = 1,   // low 32 bits of FillKey (symbols 16..31)
  kSecExact0 = 2,   // ExactKeyWord 0..2 (96-bit exact 32-symbol key)
  kSecExact1 = 3,
  kSecExact2 = 4,
  kSecRank = 5,     // rank[suffix + h]
};

enum ErrorFlag : std::uint32_t {
  kFlagBadSymbol = 1u,
  kFlagBadSentinel = 2u,
  kFlagNotPermutation = 4u,
  kFlagNotInverse = 8u,
  kFlagRange = 16u,
  kFlagLookback = 32u,  // a radix look-back wait passed its bound
  kFlagLocalSegment = 64u,  // a small-list segment exceeded kLocalMaxSegment
};

// Device counters read back by the host after each step.
// Sizes of one list of segments: suffixes and segments.
struct ListSize {
  unsigned long long survivors;
  unsigned long long segments;
};

struct DeviceCounters {
  ListSize list;      // the list RefineApply last wrote
  ListSize special;   // partition: segments holding a special window
  ListSize stash;     // partition: the other segments, set aside
  unsigned long long special_active;
  unsigned long long long_items;
  unsigned long long long_words;
  unsigned long long zero_count;  // sentinels seen by the batch packer
  unsigned long long q_next;      // suffixes appended to the next small list
  unsigned long long l_next;      // suffixes in the next large list
  unsigned long long l_next_segs; // segments in the next large list
  unsigned long long irreducible;  // PLCP positions computed directly
  unsigned long long lcp_gathered[16];  // LCP rows gathered from PLCP, over 16 words (block % 16)
  unsigned long long lcp_marked;        // rows the marked-row LCP pass computes
  unsigned long long inverse_hash[2];   // SA/ISA pair-multiset hash difference (0 when ISA inverts SA)
  unsigned int error_flags;
  unsigned int pad;
};

// Largest segment the small list holds, and the span/ownership geometry of
// one LocalRefine block: a block reads kLocalSpan consecutive slots and owns
// the segments that start in its first kLocalOwn slots (which therefore end
// inside the span).
constexpr int kLocalThreads = 1024;
constexpr int kLocalItems = 2;
constexpr int kLocalSpan = kLocalThreads * kLocalItems;
constexpr std::uint32_t kLocalMaxSegment = 256;
constexpr int kLocalOwn = kLocalSpan - static_cast<int>(kLocalMaxSegment);
constexpr int kKeygenBlock = 512;
constexpr std::uint32_t kSmallSegment = 0xFFFFFFFFu;

// ------------------------------------------------------------ K0 pack -----

// One thread per 32-symbol entry.  Encoded bytes 0..6 -> (code, spec).
// special_blocks (cleared before, or null): the entry's special-block bit
// (device_common.hpp) when it holds a special symbol.
__global__ __launch_bounds__(256) void PackTextKernel(
    const std::uint8_t* __restrict__ text, std::uint64_t n, std::uint64_t words,
    uint4* __restrict__ packed, bool aligned, bool batch, DeviceCounters* counters,
    std::uint32_t* __restrict__ special_blocks) {
  const std::uint64_t w = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= words) return;
  std::uint64_t c = 0;
  std::uint32_t s = 0;
  std::uint32_t flags = 0;
  std::uint32_t zeros = 0;
  const std::uint64_t p0 = w * 32;
  std::uint8_t bytes[32];
  if (aligned && p0 + 32 <= n) {
    const uint4* src = reinterpret_cast<const uint4*>(text + p0);
    const uint4 a = src[0];
    const uint4 b = src[1];
    const std::uint32_t parts[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
#pragma unroll
      for (int r = 0; r < 4; ++r) bytes[q * 4 + r] = static_cast<std::uint8_t>(parts[q] >> (8 * r));
    }
  } else {
#pragma unroll
    for (int j = 0; j < 32; ++j) bytes[j] = (p0 + j < n) ? text[p0 + j] : 0xFFu;
  }
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const std::uint8_t v = bytes[j];
    if (v == 0xFFu && p0 + j >= n) continue;  // padding
    std::uint32_t code, sp;
    if (v >= 2 && v <= 5) {
      code = v - 2u;
      sp = 0;
    } else if (v == 6) {
      code = 3;
      sp = 1;
    } else if (v <= 1) {
      code = v;
      sp = 1;
    } else {
      flags |= kFlagBadSymbol;
      code = 3;
      sp = 1;
    }
    if (batch) {
      zeros += v == 0;
    } else if ((v == 0) != (p0 + j == n - 1)) {
      flags |= kFlagBadSentinel;
    }
    c |= static_cast<std::uint64_t>(code) << (62 - 2 * j);
    s |= sp << (31 - j);
  }
  packed[w] = make_uint4(static_cast<std::uint32_t>(c), static_cast<std::uint32_t>(c >> 32), s, 0u);
  // A wave's 64 entries (entry-aligned: w0 a multiple of 64) lie in one
  // special block of 128 entries, so one lane sets the block's bit.
  static_assert(kSpecialBlockShift == 12, "an entry of 32 symbols is 1/128 of a special block");
  if (special_blocks != nullptr && __ballot(s != 0) != 0 && LaneId() == 0) {
    atomicOr(&special_blocks[w >> 12], 1u << ((w >> 7) & 31u));
  }
  if (flags) atomicOr(&counters->error_flags, flags);
  if (zeros) atomicAdd(&counters->zero_count, static_cast<unsigned long long>(zeros));
}

// Batch layout: reference r owns positions (and SA rows) [offsets[r],
// offsets[r+1]); its last symbol must be its sentinel.
__device__ inline std::uint32_t ReferenceOf(const std::uint64_t* offsets, std::uint32_t refs,
                                            std::uint64_t k) {
  std::uint32_t lo = 0, hi = refs;  // offsets[lo] <= k < offsets[hi]
  while (hi - lo > 1) {
    const std::uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void CheckReferenceEndsKernel(
    const std::uint8_t* __restrict__ text, const std::uint64_t* __restrict__ offsets,
    std::uint32_t refs, DeviceCounters* counters) {
  const std::uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= refs) return;
  if (offsets[r + 1] <= offsets[r] + 1 || text[offsets[r + 1] - 1] != 0) {
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagBadSentinel));
  }
}

// Initial segments of a batch: one per reference.
__global__ __launch_bounds__(256) void InitSegmentsKernel(
    const std::uint64_t* __restrict__ offsets, std::uint32_t refs, std::uint64_t n,
    std::uint32_t* __restrict__ act_seg) {
  const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) act_seg[k] = ReferenceOf(offsets, refs, k);
}

// Per-reference coordinates: SA values and ISA rows minus the reference's
// first position, each checked to stay inside its own reference.  Block b
// handles kLocalizeTile consecutive rows; the reference of a row follows from
// the block's first and last reference (usually one), and a value or ISA row
// only has to be checked against that reference's bounds.
constexpr int kLocalizeItems = 16;
constexpr int kLocalizeTile = 256 * kLocalizeItems;

template <typename Index, typename OutSa, typename OutIsa>
__global__ __launch_bounds__(256) void LocalizeKernel(
    const Index* sa, const Index* rank, std::uint64_t n, const std::uint64_t* __restrict__ offsets,
    std::uint32_t refs, OutSa* out_sa, OutIsa* out_isa, DeviceCounters* counters) {
  __shared__ std::uint32_t s_ref[2];
  const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kLocalizeTile;
  const std::uint64_t last = base + kLocalizeTile < n ? base + kLocalizeTile - 1 : n - 1;
  if (threadIdx.x < 2) s_ref[threadIdx.x] = SegmentOfSlot(offsets, 0, refs - 1, threadIdx.x == 0 ? base : last);
  __syncthreads();
  const std::uint32_t r0 = s_ref[0], r1 = s_ref[1];
  std::uint32_t flags = 0;
#pragma unroll 4
  for (int j = 0; j < kLocalizeItems; ++j) {
    const std::uint64_t k = base + static_cast<std::uint64_t>(j) * 256 + threadIdx.x;
    if (k > last) break;
    const std::uint32_t ref = r0 == r1 ? r0 : SegmentOfSlot(offsets, r0, r1, k);
    const std::uint64_t begin = offsets[ref], end = offsets[ref + 1];
    const std::uint64_t v = sa[k];
    const bool sa_ok = v >= begin && v < end;
    flags |= sa_ok ? 0u : static_cast<std::uint32_t>(kFlagNotPermutation);
    if (sa_ok) out_sa[k] = static_cast<OutSa>(v - begin);
    if (out_isa != nullptr) {
      const std::uint64_t isa_v = rank[k];
      const bool isa_ok = isa_v >= begin && isa_v < end;
      flags |= isa_ok ? 0u : static_cast<std::uint32_t>(kFlagNotInverse);
      if (isa_ok) out_isa[k] = static_cast<OutIsa>(isa_v - begin);
    }
  }
  if (flags) atomicOr(&counters->error_flags, flags);
}

// ------------------------------------------------------------ keys --------

// Secondary key of suffix v for a refine step (see SecondaryKind).
template <typename R>
__device__ inline std::uint64_t SecondaryKey(int sec_kind, PackedText t, const R* __restrict__ rank,
                                             std::uint64_t v, std::uint64_t h, DeviceCounters* counters) {
  if (sec_kind == kSecRank) {
    // An unsorted suffix never reaches the sentinel within h symbols, so
    // v + h < n; the guard only protects against a corrupted state.
    const std::uint64_t p = v + h;
    if (p < t.n) return static_cast<std::uint64_t>(rank[p]);
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
    return 0;
  }
  std::uint64_t c;
  std::uint32_t s;
  LoadWindow(t, v, c, s);
  if (sec_kind == kSecFillHi) return FillKey(c, s) >> 32;
  if (sec_kind == kSecFillLo) return FillKey(c, s) & 0xFFFFFFFFull;
  return ExactKeyWord(c, s, sec_kind - kSecExact0);
}

// Sort keys of the large list, key = (segment id << sec_bits) | secondary,
// and the digit counts of every radix pass over them.
// act_idx may be the same buffer as vals: slot k is read before slot k is
// written, by the same thread (hence no __restrict__ on either).  kRank: the
// secondary key is rank[suffix + h] (sec_kind is kSecRank); otherwise it
// comes from the suffix's 32-symbol window (FillHi, FillLo, Exact0..2).
// lo_from_value (FillLo of a 64-bit list whose values carry it above the
// position): the key is that payload and the text is not read.  vals receive
// the positions alone.
template <typename Index, typename K, bool kRank, typename R = Index>
__global__ __launch_bounds__(kKeygenBlock) void KeygenKernel(
    const Index* act_idx, const std::uint32_t* __restrict__ act_seg,
    std::uint64_t m, int sec_kind, int sec_bits, int seg_bits, PackedText t,
    const R* __restrict__ rank, std::uint64_t h, bool iota,
    K* __restrict__ keys, Index* vals, RadixPasses passes, RadixSet set, DeviceCounters* counters,
    bool lo_from_value) {
  constexpr int kWaves = kKeygenBlock / kWave;
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][RadixTile::kRadix];
  ClearNextRadixSet(set);
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * RadixTile::kRadix; i += kKeygenBlock) {
    (&local[0][0][0])[i] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kKeygenBlock;
  // kUnroll wave spans per step, every load of the step before its stores,
  // so that the step's gathers are in flight together.  Each slot is still
  // read before it is written, by the same thread.
  constexpr int kUnroll = 4;
  for (std::uint64_t wbase = (static_cast<std::uint64_t>(blockIdx.x) * kWaves + wave) * kWave; wbase < m;
       wbase += stride * kUnroll) {
    Index v[kUnroll];
    std::uint64_t seg[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const std::uint64_t k = wbase + u * stride + LaneId();
      const bool valid = k < m;
      v[u] = valid ? (iota ? static_cast<Index>(k) : act_idx[k]) : Index(0);
      seg[u] = valid && seg_bits ? static_cast<std::uint64_t>(act_seg[k]) : 0ull;
    }
    std::uint64_t sec[kUnroll];
    if constexpr (kRank) {
      // An unsorted suffix never reaches the sentinel within h symbols, so
      // v + h < n; the guard only protects against a corrupted state.
      bool bad = false;
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const bool valid = wbase + u * stride + LaneId() < m;
        const std::uint64_t p = PosOf(v[u]) + h;
        const bool ok = p < t.n;
        bad |= valid & !ok;
        sec[u] = (valid && ok) ? static_cast<std::uint64_t>(rank[ok ? p : 0]) : 0ull;
      }
      if (bad) atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
    } else if (lo_from_value) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) sec[u] = static_cast<std::uint64_t>(v[u]) >> 32;
    } else {
      uint4 e0[kUnroll], e1[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const std::uint64_t w = PosOf(v[u]) >> 5;
        e0[u] = t.words[w];
        e1[u] = t.words[w + 1];
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        std::uint64_t c;
        std::uint32_t sp;
        WindowOf(e0[u], e1[u], static_cast<std::uint32_t>(v[u] & 31), c, sp);
        sec[u] = sec_kind == kSecFillHi   ? FillKey(c, sp) >> 32
                 : sec_kind == kSecFillLo ? FillKey(c, sp) & 0xFFFFFFFFull
                                          : ExactKeyWord(c, sp, sec_kind - kSecExact0);
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const std::uint64_t kbase = wbase + u * stride;
      const std::uint64_t k = kbase + LaneId();
      const bool valid = k < m;
      const K key = valid ? static_cast<K>((seg[u] << sec_bits) | sec[u]) : K(0);
      if (valid) {
        keys[k] = key;
        vals[k] = static_cast<Index>(PosOf(v[u]));
      }
      CountDigits<K>(local[wave % kDigitCopies], passes, key, valid, kbase + kWave <= m);
    }
  }
  __syncthreads();
  FlushDigits<kDigitCopies>(local, passes, set.hist);
}

// Digit counts of the first step's keys (high half of FillKey of every
// suffix of a single text), four consecutive positions per lane from one pair
// of 16-byte entries (four positions never straddle a 32-symbol entry).
__global__ __launch_bounds__(kKeygenBlock) void TextFillHiCountKernel(PackedText t, RadixPasses passes,
                                                                      RadixSet set) {
  constexpr int kWaves = kKeygenBlock / kWave;
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][RadixTile::kRadix];
  ClearNextRadixSet(set);
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * RadixTile::kRadix; i += kKeygenBlock) {
    (&local[0][0][0])[i] = 0;
  }
  __syncthreads();
  const std::uint64_t m = t.n;
  const std::uint64_t groups = (m + 3) / 4;
  const int wave = threadIdx.x / kWave;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kKeygenBlock;
  for (std::uint64_t wbase = (static_cast<std::uint64_t>(blockIdx.x) * kWaves + wave) * kWave; wbase < groups;
       wbase += stride) {
    const std::uint64_t g = wbase + LaneId();
    uint4 e0 = make_uint4(0, 0, 0, 0), e1 = e0;
    if (g < groups) {
      e0 = t.words[g >> 3];
      e1 = t.words[(g >> 3) + 1];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const std::uint64_t p = g * 4 + j;
      std::uint64_t c;
      std::uint32_t sp;
      WindowOf(e0, e1, static_cast<std::uint32_t>(p & 31), c, sp);
      const std::uint32_t key = static_cast<std::uint32_t>(FillKey(c, sp) >> 32);
      // No test for a digit the whole wave shares: text windows rarely do,
      // and the test cost more than the colliding atomics it saves.
      CountDigits<std::uint32_t>(local[wave % kDigitCopies], passes, key, p < m, false);
    }
  }
  __syncthreads();
  FlushDigits<kDigitCopies>(local, passes, set.hist);
}

// Digit counts of the first step's keys per reference of a batch:
// counts[(r * kFillHiPasses + p) * kRadix + d].  Block b counts positions
// [b * chunk, (b + 1) * chunk), one reference at a time (a chunk usually lies
// in one reference), four consecutive positions per lane.  passes: the four
// 8-bit digits of the 32-bit key.
constexpr int kFillHiPasses = 4;

__global__ __launch_bounds__(kKeygenBlock) void TextFillHiRefCountKernel(PackedText t,
                                                                         const std::uint64_t* __restrict__ offsets,
                                                                         std::uint32_t refs, std::uint64_t chunk,
                                                                         RadixPasses passes,
                                                                         unsigned int* __restrict__ counts) {
  constexpr int kWaves = kKeygenBlock / kWave;
  constexpr int kRadix = RadixTile::kRadix;
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][kRadix];
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * kRadix; i += kKeygenBlock) (&local[0][0][0])[i] = 0;
  const std::uint64_t c0 = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t c1 = c0 + chunk < t.n ? c0 + chunk : t.n;
  if (c0 >= c1) return;
  const int wave = threadIdx.x / kWave;
  std::uint32_t r = SegmentOfSlot(offsets, 0, refs - 1, c0);
  __syncthreads();
  for (;;) {
    const std::uint64_t lo = offsets[r] > c0 ? offsets[r] : c0;
    const std::uint64_t hi = offsets[r + 1] < c1 ? offsets[r + 1] : c1;
    // Groups of four aligned positions overlapping [lo, hi).
    const std::uint64_t g0 = lo / 4, g1 = (hi + 3) / 4;
    for (std::uint64_t wbase = g0 + static_cast<std::uint64_t>(wave) * kWave; wbase < g1;
         wbase += kKeygenBlock) {
      const std::uint64_t g = wbase + LaneId();
      uint4 e0 = make_uint4(0, 0, 0, 0), e1 = e0;
      if (g < g1) {
        e0 = t.words[g >> 3];
        e1 = t.words[(g >> 3) + 1];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const std::uint64_t p = g * 4 + j;
        std::uint64_t c;
        std::uint32_t sp;
        WindowOf(e0, e1, static_cast<std::uint32_t>(p & 31), c, sp);
        const std::uint32_t key = static_cast<std::uint32_t>(FillKey(c, sp) >> 32);
        // No shared-digit test, as in TextFillHiCountKernel.
        CountDigits<std::uint32_t>(local[wave % kDigitCopies], passes, key, g < g1 && p >= lo && p < hi, false);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kFillHiPasses * kRadix; i += kKeygenBlock) {
      const int p = i / kRadix, d = i % kRadix;
      std::uint32_t sum = 0;
#pragma unroll
      for (int w = 0; w < kDigitCopies; ++w) {
        sum += local[w][p][d];
        local[w][p][d] = 0;
      }
      if (sum) atomicAdd(&counts[(static_cast<std::uint64_t>(r) * kFillHiPasses + p) * kRadix + d], sum);
    }
    __syncthreads();
    if (hi >= c1) break;
    ++r;
  }
}

// adj[(p * refs + r) * kRadix + d] = offsets[r] + (elements of r with digit
// below d in pass p); one block per (reference, pass).
__global__ __launch_bounds__(RadixTile::kRadix) void RefDigitStartsKernel(const unsigned int* __restrict__ counts,
                                                                          const std::uint64_t* __restrict__ offsets,
                                                                          std::uint32_t refs,
                                                                          long long* __restrict__ adj) {
  constexpr int kRadix = RadixTile::kRadix;
  __shared__ std::uint64_t scratch[kRadix / kWave];
  const std::uint32_t r = blockIdx.x;
  const int p = static_cast<int>(blockIdx.y);
  const std::uint64_t c = counts[(static_cast<std::uint64_t>(r) * kFillHiPasses + p) * kRadix + threadIdx.x];
  std::uint64_t total;
  const std::uint64_t ex = BlockExclusiveSum<kRadix>(c, scratch, total);
  adj[(static_cast<std::uint64_t>(p) * refs + r) * kRadix + threadIdx.x] = static_cast<long long>(offsets[r] + ex);
}

// adj[p][r][d] -= (elements with digit d in references before r); one block
// per (pass, digit), scanning the references 256 at a time.
__global__ __launch_bounds__(RadixTile::kRadix) void RefDigitBeforeKernel(const unsigned int* __restrict__ counts,
                                                                          std::uint32_t refs,
                                                                          long long* __restrict__ adj) {
  constexpr int kRadix = RadixTile::kRadix;
  __shared__ std::uint64_t scratch[kRadix / kWave];
  const int p = static_cast<int>(blockIdx.x) / kRadix;
  const int d = static_cast<int>(blockIdx.x) % kRadix;
  std::uint64_t before = 0;
  for (std::uint32_t r0 = 0; r0 < refs; r0 += kRadix) {
    const std::uint32_t r = r0 + threadIdx.x;
    const std::uint64_t c =
        r < refs ? counts[(static_cast<std::uint64_t>(r) * kFillHiPasses + p) * kRadix + d] : 0ull;
    std::uint64_t total;
    const std::uint64_t ex = BlockExclusiveSum<kRadix>(c, scratch, total);
    if (r < refs) adj[(static_cast<std::uint64_t>(p) * refs + r) * kRadix + d] -= static_cast<long long>(before + ex);
    before += total;
  }
}

// ------------------------------------------------------------ tuning ----

// Synthetic input of the radix autotuner: key i = splitmix64(seed + i)
// (truncated to K), value i = i, and the digit counts of `passes` over them.
__device__ inline std::uint64_t SplitMix64(std::uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename K, typename V>
__global__ __launch_bounds__(kKeygenBlock) void TuneKeysKernel(K* __restrict__ keys, V* __restrict__ vals,
                                                               std::uint64_t m, std::uint64_t seed,
                                                               RadixPasses passes, RadixSet set) {
  constexpr int kWaves = kKeygenBlock / kWave;
  __shared__ std::uint32_t local[kDigitCopies][RadixPasses::kMax][RadixTile::kRadix];
  ClearNextRadixSet(set);
  for (int i = threadIdx.x; i < kDigitCopies * RadixPasses::kMax * RadixTile::kRadix; i += kKeygenBlock) {
    (&local[0][0][0])[i] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kKeygenBlock;
  for (std::uint64_t wbase = (static_cast<std::uint64_t>(blockIdx.x) * kWaves + wave) * kWave; wbase < m;
       wbase += stride) {
    const std::uint64_t i = wbase + LaneId();
    const bool valid = i < m;
    const K key = static_cast<K>(SplitMix64(seed + i));
    if (valid) {
      keys[i] = key;
      vals[i] = static_cast<V>(i);
    }
    CountDigits<K>(local[wave % kDigitCopies], passes, key, valid, wbase + kWave <= m);
  }
  __syncthreads();
  FlushDigits<kDigitCopies>(local, passes, set.hist);
}

// Placement tables of a one-segment list (bucket starts of each pass), for
// timing the in-place sorts: adj[p][d] = digits below d in pass p.
__global__ __launch_bounds__(RadixTile::kRadix) void TuneAdjKernel(const unsigned long long* __restrict__ hist,
                                                                   long long* __restrict__ adj) {
  constexpr int kRadix = RadixTile::kRadix;
  __shared__ std::uint64_t scratch[kRadix / kWave];
  const int p = static_cast<int>(blockIdx.x);
  std::uint64_t total;
  const std::uint64_t ex = BlockExclusiveSum<kRadix>(static_cast<std::uint64_t>(hist[p * kRadix + threadIdx.x]),
                                                     scratch, total);
  adj[p * kRadix + threadIdx.x] = static_cast<long long>(ex);
}

// First pass of the first step: key = high half of FillKey of suffix i,
// value = i, generated from the packed text instead of read.  pack_lo (64-bit
// values only): the value also carries the low FillKey half above the
// position (PosOf), for the steps that order the first step's groups by it.
template <typename Index>
struct TextFillHiSource {
  static constexpr int kFetchGroup = 8;  // two 12-byte entry loads per item
  PackedText t;
  bool pack_lo = false;
  struct Raw {
    uint4 e0, e1;  // the packed entries holding the window
  };
  __device__ Raw Fetch(std::uint64_t i) const {
    const std::uint64_t w = i >> 5;
    return Raw{t.words[w], t.words[w + 1]};
  }
  __device__ void Make(const Raw& r, std::uint64_t i, std::uint32_t& key, Index& val) const {
    std::uint64_t c;
    std::uint32_t s;
    WindowOf(r.e0, r.e1, static_cast<std::uint32_t>(i & 31), c, s);
    const std::uint64_t fill = FillKey(c, s);
    key = static_cast<std::uint32_t>(fill >> 32);
    if constexpr (sizeof(Index) == 8) {
      val = static_cast<Index>(pack_lo ? (fill << 32) | i : i);
    } else {
      val = static_cast<Index>(i);
    }
  }
};

// Over the list RefineApply just wrote (its size is read from the counters;
// the grid covers an upper bound): counts the suffixes whose 32-symbol window
// holds a special symbol and marks their segments (flags[segment] = 1, the
// only value ever stored, so racing stores agree).  The list is in segment
// order, so one lane per run of special suffixes of one segment in a wave
// stores, and only while the flag still reads 0.  Each wave takes
// kMarkSpans consecutive spans of 64 slots, all their loads issued before
// any is used; windows outside special blocks are not read.
constexpr int kMarkSpans = 4;

template <typename Index>
__global__ __launch_bounds__(256) void MarkSpecialSegmentsKernel(
    const Index* __restrict__ act_idx, const std::uint32_t* __restrict__ act_seg, PackedText t,
    const std::uint32_t* __restrict__ special_blocks, std::uint32_t* __restrict__ flags, DeviceCounters* counters) {
  const std::uint64_t m = counters->list.survivors;
  const std::uint32_t lane = LaneId();
  const std::uint64_t base =
      (static_cast<std::uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave) * kWave * kMarkSpans;
  std::uint64_t pos[kMarkSpans];
  std::uint32_t seg[kMarkSpans];
  bool maybe[kMarkSpans];
#pragma unroll
  for (int u = 0; u < kMarkSpans; ++u) {
    const std::uint64_t k = base + static_cast<std::uint64_t>(u) * kWave + lane;
    pos[u] = k < m ? PosOf(act_idx[k]) : 0;
    seg[u] = k < m ? act_seg[k] : 0;
  }
#pragma unroll
  for (int u = 0; u < kMarkSpans; ++u) {
    const std::uint64_t k = base + static_cast<std::uint64_t>(u) * kWave + lane;
    maybe[u] = k < m && MaybeSpecialWindow(special_blocks, pos[u]);
  }
  uint4 e0[kMarkSpans], e1[kMarkSpans];
#pragma unroll
  for (int u = 0; u < kMarkSpans; ++u) {
    if (maybe[u]) {
      e0[u] = t.words[pos[u] >> 5];
      e1[u] = t.words[(pos[u] >> 5) + 1];
    }
  }
  unsigned long long specials = 0;
#pragma unroll
  for (int u = 0; u < kMarkSpans; ++u) {
    bool special = false;
    if (maybe[u]) {
      std::uint64_t c;
      std::uint32_t sp;
      WindowOf(e0[u], e1[u], static_cast<std::uint32_t>(pos[u] & 31), c, sp);
      special = sp != 0;
    }
    const std::uint64_t vote = __ballot(special);
    const std::uint32_t prev_seg = __shfl_up(seg[u], 1, kWave);
    const bool prev_special = lane > 0 && ((vote >> (lane - 1)) & 1ull) != 0;
    if (special && !(prev_special && prev_seg == seg[u]) && flags[seg[u]] == 0) flags[seg[u]] = 1;
    specials += static_cast<unsigned long long>(__popcll(vote));
  }
  if (lane == 0 && specials) atomicAdd(&counters->special_active, specials);
}

// ------------------------------------------------------------ refine ------

// Boundaries, singletons and compaction of a sorted large list.  Slot k of
// the sorted list starts a group when its key differs from slot k-1's; a
// group of one is a singleton (its SA row is final: sa[row] = v, rank[v] =
// row), the others survive into the compacted list (suffix, segment id) with
// one new segment per surviving group (delta = first row - first compacted
// slot, start = first compacted slot).  Two schedules share the per-tile code:
//   - single pass: one block of kRefineThreads per tile; the survivor count,
//     new-segment count and last group start of the tiles before come from
//     wave-wide decoupled look-backs (three values, status words interleaved
//     by 3); the last tile writes the totals (survivors, segments) to the
//     counters.  Tiles retire roughly in list order, so when sorted order
//     follows text order (segments that are the references of a batch) the
//     rank writes in flight stay within a few references;
//   - reduce / spine / apply: per-chunk totals, their scan, then each block
//     walks its chunk of tiles in order with no waits between blocks.  Used
//     for a one-segment list (the first step of a single text) beyond
//     kRefineResidentTiles tiles: its rank writes have no locality to keep,
//     and a look-back over thousands of tiles, which the chip's dies take
//     round-robin, waits on the slowest die.
// Per tile: keys staged in LDS with coalesced loads (plus the key before and
// after it); flags and per-slot offsets from a blocked scan (thread t owns
// slots kItems*t ...); the writes go out striped (values read coalesced from
// global), so that consecutive lanes write consecutive SA rows and
// consecutive compacted slots.  survivor_rank: also set rank[v] of surviving
// suffixes to their group's first row (a step whose successor rewrites every
// survivor's rank before any rank is gathered skips it).

// Early LCP values.  The LCP of a row is fixed when the row becomes the
// first of its group: its predecessor row then belongs to the group just
// before, and any two suffixes of the two groups share exactly the prefix
// their sort keys share.  The first step (16-symbol keys) writes that value
// for every group start whose keys prove it (see EarlyKeyLcp) and a marker
// in every other row; the exact local step writes its group starts (LCP 16
// to 31, or the exact mismatch in segments with special symbols).  The LCP
// pass then reads PLCP only for rows that still hold the marker.  Both
// pointers null: off.
struct EarlyLcp {
  std::uint32_t* lcp32;
  std::uint64_t* lcp64;
};
constexpr std::uint32_t kLcpUnknown32 = 0xFFFFFFFFu;
constexpr std::uint64_t kLcpUnknown64 = ~0ull;

__device__ __forceinline__ void WriteEarlyLcp(const EarlyLcp& e, std::uint64_t row, std::uint32_t value,
                                              bool known) {
  if (e.lcp32 != nullptr) {
    e.lcp32[row] = known ? value : kLcpUnknown32;
  } else if (e.lcp64 != nullptr) {
    e.lcp64[row] = known ? static_cast<std::uint64_t>(value) : kLcpUnknown64;
  }
}

// LCP of two suffixes from their 16-symbol FillKey high halves a != b (2-bit
// codes, first symbol in the top bits), when the keys prove it: the first
// differing symbol m, provided neither window holds a special symbol at or
// before m.  FillKey fills a window from its first special symbol on with
// all-0 or all-1 codes, so a key whose symbols m..15 are not all 00 and not
// all 11 has no special symbol at or before m; then symbols 0..m-1 are equal
// bases and symbol m is a base in both, so the LCP is m.  Returns 32 when the
// keys do not prove it.
__device__ __forceinline__ std::uint32_t EarlyKeyLcp(std::uint32_t a, std::uint32_t b) {
  const std::uint32_t m = static_cast<std::uint32_t>(__clz(a ^ b)) >> 1;
  const std::uint32_t tail_bits = 32u - 2u * m;                   // symbols m..15
  const std::uint32_t mask = tail_bits >= 32u ? 0xFFFFFFFFu : ((1u << tail_bits) - 1u);
  const std::uint32_t ta = a & mask, tb = b & mask;
  const bool proven = ta != 0u && ta != mask && tb != 0u && tb != mask;
  return proven ? m : 32u;
}

struct RefinePartial {
  std::uint64_t survivors;
  std::uint64_t newsegs;
  std::uint64_t last_boundary;  // 1 + index of the last group start, 0 if none
  std::uint64_t pad;
};

constexpr int kRefineThreads = 512;  // single pass
constexpr int kRefineItems = 8;
constexpr int kRefineTile = kRefineThreads * kRefineItems;
constexpr int kRefineResidentTiles = 512;
constexpr int kRefineChunkThreads = 256;  // reduce / spine / apply
constexpr int kRefineChunkTile = kRefineChunkThreads * kRefineItems;

template <typename K, int kTile>
struct RefineTileLds {
  static_assert(kTile <= 32767, "per-slot offsets are 16-bit");
  K key[kTile + 2];             // key[i + 1] = tile slot i
  std::uint16_t pos[kTile];     // survivor slot, relative to the tile's first
  std::uint16_t segc[kTile];    // segments opened through this slot (relative)
  std::int16_t group[kTile];    // slot of the group start, -1: before the tile
  std::uint8_t flags[kTile];    // bit0 starts, bit1 single
  std::uint32_t bound[kTile / 32 + 1];  // bit i: slot i starts a reference (batch first step)
};

// Totals of one tile: survivors, new segments, 1 + last group start (0: none).
struct RefineTotals {
  std::uint64_t surv, segs, last;
};

// One tile's keys and values in registers (item j of thread t is slot
// j * kBlock + t) and, in thread 0, the keys just before and after the tile:
// loaded ahead of the tile's scan when a block walks tiles in turn.
template <typename K, typename Index, int kItems>
struct RefineTileRegs {
  K key[kItems];
  Index v[kItems];
  K before, after;
};

template <int kBlock, int kItems, typename Index, typename K>
__device__ __forceinline__ void LoadRefineTile(const K* __restrict__ keys, const Index* __restrict__ vals,
                                               std::uint64_t m, std::uint64_t base, int len,
                                               RefineTileRegs<K, Index, kItems>& r) {
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int i = j * kBlock + static_cast<int>(threadIdx.x);
    if (i < len) {
      r.key[j] = keys[base + i];
      r.v[j] = vals[base + i];
    }
  }
  if (threadIdx.x == 0) {
    r.before = base > 0 ? keys[base - 1] : K(0);
    r.after = base + len < m ? keys[base + len] : K(0);
  }
}

// Stages tile [base, base + len): keys into LDS, values into registers
// (v[j] = slot j * kBlock + thread, the striped order of the writes), and
// fills the per-slot arrays.  ref_offsets (or null): the slots are a batch's
// references laid out in place (refs + 1 entries), and a reference start is
// a group start whatever the keys.  The caller puts a barrier between this
// and RefineTileWrite.
template <int kBlock, int kItems, typename Index, typename K>
__device__ __forceinline__ RefineTotals RefineTileScanRegs(const RefineTileRegs<K, Index, kItems>& r,
                                                           std::uint64_t m, std::uint64_t base, int len,
                                                           const std::uint64_t* __restrict__ ref_offsets,
                                                           std::uint32_t refs, RefineTileLds<K, kBlock * kItems>& lds,
                                                           std::uint64_t* scratch, Index (&v)[kItems]) {
  constexpr int kTile = kBlock * kItems;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int i = j * kBlock + static_cast<int>(threadIdx.x);
    if (i < len) {
      lds.key[i + 1] = r.key[j];
      v[j] = r.v[j];
    }
  }
  if (threadIdx.x == 0) {
    lds.key[0] = r.before;
    lds.key[len + 1] = r.after;
  }
  if (ref_offsets) {
    for (int i = threadIdx.x; i < kTile / 32 + 1; i += kBlock) lds.bound[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
      // From the reference holding the tile's first slot (a start only when
      // it begins there) through the one starting right after the tile.
      for (std::uint32_t r = SegmentOfSlot(ref_offsets, 0, refs - 1, base);
           r < refs && ref_offsets[r] <= base + static_cast<std::uint64_t>(len); ++r) {
        if (ref_offsets[r] < base) continue;
        const std::uint32_t i = static_cast<std::uint32_t>(ref_offsets[r] - base);
        lds.bound[i >> 5] |= 1u << (i & 31);
      }
    }
  }
  __syncthreads();
  const int first = static_cast<int>(threadIdx.x) * kItems;
  std::uint64_t surv = 0, segs = 0, last = 0;
  std::uint8_t fl[kItems];
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const int i = first + q;
    fl[q] = 0;
    if (i < len) {
      // Branch-free (all three LDS reads are in bounds), like the segment
      // starts of LocalRefineKernel.
      const std::uint64_t k = base + i;
      const K mine = lds.key[i + 1];
      const K before = lds.key[i];
      const K after = lds.key[i + 2];
      bool starts = (k == 0) | Differ(before, mine);
      bool ends = (k + 1 == m) | Differ(after, mine);
      if (ref_offsets) {
        starts |= (lds.bound[i >> 5] >> (i & 31)) & 1u;
        ends |= (lds.bound[(i + 1) >> 5] >> ((i + 1) & 31)) & 1u;
      }
      const bool single = starts & ends;
      fl[q] = static_cast<std::uint8_t>((starts ? 1 : 0) | (single ? 2 : 0));
      surv += !single;
      segs += starts && !single;
      if (starts) last = static_cast<std::uint64_t>(i) + 1;
    }
  }
  RefineTotals tot;
  std::uint64_t pos = BlockExclusiveSum<kBlock>(surv, scratch, tot.surv);
  std::uint64_t segc = BlockExclusiveSum<kBlock>(segs, scratch, tot.segs);
  std::uint64_t cur_last = BlockExclusiveMax<kBlock>(last, scratch, tot.last);  // 0: none in tile yet
#pragma unroll
  for (int q = 0; q < kItems; ++q) {
    const int i = first + q;
    if (i < len) {
      const bool starts = fl[q] & 1, single = fl[q] & 2;
      if (starts) cur_last = static_cast<std::uint64_t>(i) + 1;
      if (!single && starts) ++segc;
      lds.pos[i] = static_cast<std::uint16_t>(pos);
      lds.segc[i] = static_cast<std::uint16_t>(segc);
      lds.group[i] = static_cast<std::int16_t>(static_cast<int>(cur_last) - 1);
      lds.flags[i] = fl[q];
      if (!single) ++pos;
    }
  }
  return tot;
}

template <int kBlock, int kItems, typename Index, typename K>
__device__ __forceinline__ RefineTotals RefineTileScan(const K* __restrict__ keys, const Index* __restrict__ vals,
                                                       std::uint64_t m, std::uint64_t base, int len,
                                                       const std::uint64_t* __restrict__ ref_offsets,
                                                       std::uint32_t refs, RefineTileLds<K, kBlock * kItems>& lds,
                                                       std::uint64_t* scratch, Index (&v)[kItems]) {
  RefineTileRegs<K, Index, kItems> r;
  LoadRefineTile<kBlock, kItems>(keys, vals, m, base, len, r);
  return RefineTileScanRegs<kBlock, kItems>(r, m, base, len, ref_offsets, refs, lds, scratch, v);
}

// Writes of one tile, given the survivors, segments and 1 + last group start
// of every slot before the tile.  kOneSegment: the list is one segment whose
// delta is one_delta (no per-item delta or key loads, so that a block's loads
// of its next tile stay the only ones in flight).
template <int kBlock, int kItems, bool kOneSegment, typename Index, typename K, typename R>
__device__ __forceinline__ void RefineTileWrite(const Index (&vals)[kItems], const K* __restrict__ keys,
                                       std::uint64_t base, int len, int sec_bits,
                                       const Index* __restrict__ seg_delta, Index one_delta, std::uint64_t n,
                                       R* __restrict__ sa, R* __restrict__ rank,
                                       Index* __restrict__ out_idx, std::uint32_t* __restrict__ out_seg,
                                       Index* __restrict__ out_delta, Index* __restrict__ out_start,
                                       bool survivor_rank, bool ranks_current, std::uint64_t run_surv,
                                       std::uint64_t run_segs, std::uint64_t run_last,
                                       const RefineTileLds<K, kBlock * kItems>& lds, const EarlyLcp early) {
  // ranks_current: every suffix's rank is already its segment's first row
  // (the step that formed the segments wrote it, and rows do not move), so
  // the suffixes of a group that opens its segment keep their rank and their
  // rank stores are skipped.  A group opens its segment when its first slot
  // is the list's first or follows a slot of another segment.
  auto opens_segment = [sec_bits](std::uint64_t slot, K at, K before) {
    return slot == 0 || (static_cast<std::uint64_t>(at) >> sec_bits) != (static_cast<std::uint64_t>(before) >> sec_bits);
  };
  // The group running in from the tiles before (its first slot run_last - 1).
  bool carried_opens = false;
  if (ranks_current && run_last > 0) {
    const std::uint64_t s = run_last - 1;
    if constexpr (kOneSegment) {
      carried_opens = s == 0;
    } else {
      carried_opens = opens_segment(s, keys[s], s > 0 ? keys[s - 1] : K(0));
    }
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int i = j * kBlock + static_cast<int>(threadIdx.x);
    if (i >= len) break;
    const std::uint64_t k = base + i;
    const K key = lds.key[i + 1];
    const Index v = vals[j];
    const std::uint8_t f = lds.flags[i];
    const std::uint64_t seg = static_cast<std::uint64_t>(key) >> sec_bits;
    const Index delta = kOneSegment ? one_delta : seg_delta[seg];
    const Index row = static_cast<Index>(delta + static_cast<Index>(k));
    const Index pos = static_cast<Index>(PosOf(v));  // in the list's own width
    if (static_cast<std::uint64_t>(row) >= n || static_cast<std::uint64_t>(pos) >= n) continue;  // corrupted state only
    if (early.lcp32 != nullptr || early.lcp64 != nullptr) {
      // The first step: every row gets its early LCP or the marker.
      std::uint32_t value = 32u;
      if ((f & 1) && k > 0) {
        const K before = lds.key[i];
        if ((static_cast<std::uint64_t>(before) >> sec_bits) == seg) {
          value = EarlyKeyLcp(static_cast<std::uint32_t>(before), static_cast<std::uint32_t>(key));
        }
      }
      WriteEarlyLcp(early, static_cast<std::uint64_t>(row), value, value < 32u);
    }
    if (f & 2) {
      sa[row] = static_cast<R>(pos);
      if (!(ranks_current && opens_segment(k, key, lds.key[i]))) rank[pos] = static_cast<R>(row);
    } else {
      // The group start lies in the same segment: every segment start is a
      // key change, hence a group start.
      const int g = lds.group[i];
      const std::uint64_t group_slot = g >= 0 ? base + static_cast<std::uint64_t>(g) : run_last - 1;
      const bool kept = ranks_current && (g >= 0 ? opens_segment(group_slot, lds.key[g + 1], lds.key[g])
                                                 : carried_opens);
      if (survivor_rank && !kept) rank[pos] = static_cast<R>(delta + static_cast<Index>(group_slot));
      const std::uint64_t at = run_surv + lds.pos[i];
      const std::uint64_t sc = run_segs + lds.segc[i];  // segments opened through this slot
      if (f & 1) {
        out_delta[sc - 1] = static_cast<Index>(row - static_cast<Index>(at));
        out_start[sc - 1] = static_cast<Index>(at);
      }
      out_idx[at] = v;
      out_seg[at] = static_cast<std::uint32_t>(sc - 1);
    }
  }
}

// Single pass: one block per tile (tiles numbered by LookBackTile).
template <typename Index, typename K, typename R = Index>
__global__ __launch_bounds__(kRefineThreads) void RefineApplyKernel(
    const K* __restrict__ keys, const Index* __restrict__ vals, std::uint64_t m, int sec_bits,
    const Index* __restrict__ seg_delta, std::uint64_t n, R* __restrict__ sa, R* __restrict__ rank,
    Index* __restrict__ out_idx, std::uint32_t* __restrict__ out_seg, Index* __restrict__ out_delta,
    Index* __restrict__ out_start, bool survivor_rank, bool ranks_current, unsigned long long* status,
    std::uint32_t epoch,
    unsigned int* tickets, DeviceCounters* counters,
    const std::uint64_t* __restrict__ ref_offsets, std::uint32_t refs, const EarlyLcp early) {
  constexpr int kBlock = kRefineThreads;
  constexpr int kTile = kRefineTile;
  __shared__ RefineTileLds<K, kTile> lds;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  __shared__ std::uint64_t s_prefix[3];
  __shared__ std::uint32_t s_tile;
  const std::uint64_t tile = LookBackTile(tickets, epoch, &s_tile);
  const std::uint64_t base = tile * kTile;
  const int len = static_cast<int>(m - base < static_cast<std::uint64_t>(kTile) ? m - base : kTile);
  Index v[kRefineItems];
  const RefineTotals tot =
      RefineTileScan<kBlock, kRefineItems>(keys, vals, m, base, len, ref_offsets, refs, lds, scratch, v);
  // Three wave-wide look-backs, one per wave, side by side.
  unsigned int* error_flags = &counters->error_flags;
  const int wave = static_cast<int>(threadIdx.x) / kWave;
  if (wave == 0) {
    const std::uint64_t p = LookBackWave<false>(status, 3, tile, tot.surv, 0, epoch, error_flags, kFlagLookback);
    if (LaneId() == 0) s_prefix[0] = p;
  } else if (wave == 1) {
    const std::uint64_t p = LookBackWave<false>(status + 1, 3, tile, tot.segs, 0, epoch, error_flags, kFlagLookback);
    if (LaneId() == 0) s_prefix[1] = p;
  } else if (wave == 2) {
    const std::uint64_t p = LookBackWave<true>(status + 2, 3, tile, tot.last ? base + tot.last : 0, 0, epoch,
                                               error_flags, kFlagLookback);
    if (LaneId() == 0) s_prefix[2] = p;
  }
  __syncthreads();
  const std::uint64_t run_surv = s_prefix[0];
  const std::uint64_t run_segs = s_prefix[1];
  RefineTileWrite<kBlock, kRefineItems, false>(v, keys, base, len, sec_bits, seg_delta, Index(0), n, sa, rank,
                                               out_idx, out_seg, out_delta, out_start, survivor_rank, ranks_current,
                                               run_surv, run_segs, s_prefix[2], lds, early);
  if (threadIdx.x == 0 && base + len == m) {
    counters->list.survivors = run_surv + tot.surv;
    counters->list.segments = run_segs + tot.segs;
  }
}

// Reduce: totals of chunk b = [b * chunk, (b + 1) * chunk); chunk is a
// multiple of the tile, so every thread's 16-byte group of keys is aligned.
template <typename K>
__global__ __launch_bounds__(kRefineChunkThreads) void RefineReduceKernel(const K* __restrict__ keys,
                                                                        std::uint64_t m, std::uint64_t chunk,
                                                                        RefinePartial* __restrict__ partial) {
  constexpr int kBlock = kRefineChunkThreads;
  constexpr int kVec = 16 / static_cast<int>(sizeof(K));
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t surv = 0, segs = 0, last = 0;
  for (std::uint64_t k0 = begin + threadIdx.x * kVec; k0 < end; k0 += kBlock * kVec) {
    K w[kVec + 2];  // w[q + 1] = key of slot k0 + q
    if (k0 + kVec <= m) {
      const uint4 raw = *reinterpret_cast<const uint4*>(keys + k0);
      __builtin_memcpy(&w[1], &raw, sizeof(raw));
    } else {
#pragma unroll
      for (int q = 0; q < kVec; ++q) w[q + 1] = k0 + q < m ? keys[k0 + q] : K(0);
    }
    w[0] = k0 > 0 ? keys[k0 - 1] : K(0);
    w[kVec + 1] = k0 + kVec < m ? keys[k0 + kVec] : K(0);
#pragma unroll
    for (int q = 0; q < kVec; ++q) {
      const std::uint64_t k = k0 + q;
      const bool in = k < end;
      const bool starts = (k == 0) | Differ(w[q], w[q + 1]);
      const bool ends = (k + 1 == m) | Differ(w[q + 2], w[q + 1]);
      const bool single = starts & ends;
      surv += in & !single;
      segs += in & starts & !single;
      last = (in & starts) ? k + 1 : last;
    }
  }
  std::uint64_t total_surv, total_segs, total_last;
  BlockExclusiveSum<kBlock>(surv, scratch, total_surv);
  BlockExclusiveSum<kBlock>(segs, scratch, total_segs);
  BlockExclusiveMax<kBlock>(last, scratch, total_last);
  if (threadIdx.x == 0) {
    partial[blockIdx.x].survivors = total_surv;
    partial[blockIdx.x].newsegs = total_segs;
    partial[blockIdx.x].last_boundary = total_last;
  }
}

// Spine: exclusive scan of at most kRefineMaxChunks chunk totals by one
// block of 1024 (thread t owns chunks kSpineItems*t ...); writes the list
// totals to the counters.
constexpr int kSpineItems = 8;
constexpr int kRefineMaxChunks = 1024 * kSpineItems;

// One field of the chunk totals: sum (kMax false) or max, scanned in place.
template <bool kMax>
__device__ __forceinline__ std::uint64_t SpineField(RefinePartial* partial, int grid, std::uint64_t RefinePartial::*field,
                                                    std::uint64_t* scratch) {
  constexpr int kBlock = 1024;
  const int first = static_cast<int>(threadIdx.x) * kSpineItems;
  std::uint64_t v[kSpineItems];
  std::uint64_t acc = 0;
#pragma unroll
  for (int q = 0; q < kSpineItems; ++q) {
    v[q] = first + q < grid ? partial[first + q].*field : 0ull;
    acc = kMax ? (v[q] > acc ? v[q] : acc) : acc + v[q];
  }
  std::uint64_t total;
  acc = kMax ? BlockExclusiveMax<kBlock>(acc, scratch, total) : BlockExclusiveSum<kBlock>(acc, scratch, total);
#pragma unroll
  for (int q = 0; q < kSpineItems; ++q) {
    if (first + q < grid) partial[first + q].*field = acc;
    acc = kMax ? (v[q] > acc ? v[q] : acc) : acc + v[q];
  }
  return total;
}

__global__ __launch_bounds__(1024) void RefineSpineKernel(RefinePartial* partial, int grid,
                                                          DeviceCounters* counters) {
  __shared__ std::uint64_t scratch[1024 / kWave];
  const std::uint64_t tot_surv = SpineField<false>(partial, grid, &RefinePartial::survivors, scratch);
  const std::uint64_t tot_segs = SpineField<false>(partial, grid, &RefinePartial::newsegs, scratch);
  SpineField<true>(partial, grid, &RefinePartial::last_boundary, scratch);
  if (threadIdx.x == 0) {
    counters->list.survivors = tot_surv;
    counters->list.segments = tot_segs;
  }
}

// Apply: block b walks the tiles of its chunk (a multiple of the tile) in
// order, starting from the spine's exclusive totals.
template <typename Index, typename K, typename R = Index>
__global__ __launch_bounds__(kRefineChunkThreads) void RefineApplyChunkedKernel(
    const K* __restrict__ keys, const Index* __restrict__ vals, std::uint64_t m, std::uint64_t chunk,
    int sec_bits, const Index* __restrict__ seg_delta, const RefinePartial* __restrict__ partial,
    std::uint64_t n, R* __restrict__ sa, R* __restrict__ rank, Index* __restrict__ out_idx,
    std::uint32_t* __restrict__ out_seg, Index* __restrict__ out_delta, Index* __restrict__ out_start,
    bool survivor_rank, bool ranks_current, const std::uint64_t* __restrict__ ref_offsets, std::uint32_t refs,
    const EarlyLcp early) {
  constexpr int kBlock = kRefineChunkThreads;
  constexpr int kTile = kRefineChunkTile;
  __shared__ RefineTileLds<K, kTile> lds;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t run_surv = partial[blockIdx.x].survivors;
  std::uint64_t run_segs = partial[blockIdx.x].newsegs;
  std::uint64_t run_last = partial[blockIdx.x].last_boundary;
  // A one-segment list: one delta for every row.  The next tile's keys and
  // values are loaded while this one is scanned and written.
  const Index one_delta = seg_delta[0];
  auto tile_len = [end](std::uint64_t b) {
    return static_cast<int>(end - b < static_cast<std::uint64_t>(kTile) ? end - b : kTile);
  };
  RefineTileRegs<K, Index, kRefineItems> next;
  if (begin < end) LoadRefineTile<kBlock, kRefineItems>(keys, vals, m, begin, tile_len(begin), next);
  for (std::uint64_t base = begin; base < end; base += kTile) {
    const int len = tile_len(base);
    const RefineTileRegs<K, Index, kRefineItems> cur = next;
    if (base + kTile < end) {
      LoadRefineTile<kBlock, kRefineItems>(keys, vals, m, base + kTile, tile_len(base + kTile), next);
    }
    Index v[kRefineItems];
    const RefineTotals tot =
        RefineTileScanRegs<kBlock, kRefineItems>(cur, m, base, len, ref_offsets, refs, lds, scratch, v);
    __syncthreads();
    RefineTileWrite<kBlock, kRefineItems, true>(v, keys, base, len, sec_bits, seg_delta, one_delta, n, sa, rank,
                                                out_idx, out_seg, out_delta, out_start, survivor_rank, ranks_current,
                                                run_surv, run_segs, run_last, lds, early);
    run_surv += tot.surv;
    run_segs += tot.segs;
    if (tot.last > 0) run_last = base + tot.last;
    __syncthreads();
  }
}

// ------------------------------------------------------------ split -----

// Split of a list of segments (the one RefineApply just wrote, or one set
// aside by a partition; its segment and suffix counts are read from `size`,
// first slots from seg_start) into the next small list and the next large
// list.  Block b handles segments [b * chunk, (b + 1) * chunk), kSplitItems
// consecutive segments per thread per tile.  The small list takes a block's
// segments of at most small_max suffixes when all of them fit under q_cap
// (one bounded reservation per block, so the small list has no holes); every
// other segment goes to the large list, whose ids and slots follow segment
// order (prefix sums over blocks) after those of earlier splits of the round
// (the round counters l_next, l_next_segs).
constexpr int kSplitItems = 8;
constexpr int kSplitTile = 256 * kSplitItems;

struct SplitPartial {
  unsigned long long q_base;       // small-list slot of the block's first small segment
  unsigned long long fits;         // the block's small segments went to the small list
  unsigned long long large_segs;   // large-list segments (exclusive prefix after the spine)
  unsigned long long large_elems;  // large-list suffixes (exclusive prefix after the spine)
};

template <typename Index>
__device__ inline std::uint64_t SegmentSize(const Index* __restrict__ seg_start, std::uint64_t s,
                                            std::uint64_t segs, std::uint64_t m) {
  const std::uint64_t start = static_cast<std::uint64_t>(seg_start[s]);
  const std::uint64_t end = s + 1 < segs ? static_cast<std::uint64_t>(seg_start[s + 1]) : m;
  return end - start;
}

template <typename Index>
__global__ __launch_bounds__(256) void SplitCountKernel(
    const Index* __restrict__ seg_start, std::uint32_t small_max, std::uint64_t chunk,
    SplitPartial* __restrict__ partial, const ListSize* size) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t segs = size->segments;
  const std::uint64_t m = size->survivors;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < segs ? begin + chunk : segs;
  std::uint64_t small_elems = 0, small_segs = 0, large_segs = 0, large_elems = 0;
  for (std::uint64_t s = begin + threadIdx.x; s < end; s += kBlock) {
    const std::uint64_t size = SegmentSize(seg_start, s, segs, m);
    if (size <= small_max) {
      small_elems += size;
      ++small_segs;
    } else {
      large_elems += size;
      ++large_segs;
    }
  }
  std::uint64_t t_small, t_small_segs, t_large_segs, t_large_elems;
  BlockExclusiveSum<kBlock>(small_elems, scratch, t_small);
  BlockExclusiveSum<kBlock>(small_segs, scratch, t_small_segs);
  BlockExclusiveSum<kBlock>(large_segs, scratch, t_large_segs);
  BlockExclusiveSum<kBlock>(large_elems, scratch, t_large_elems);
  if (threadIdx.x == 0) {
    SplitPartial out;
    out.q_base = t_small;  // small suffixes and segments, until the spine
    out.fits = t_small_segs;
    out.large_segs = t_large_segs;
    out.large_elems = t_large_elems;
    partial[blockIdx.x] = out;
  }
}

// Over at most 1024 blocks: a block's small segments go to the small list
// when the small suffixes of all blocks up to it fit under q_cap (the blocks
// that fit are a prefix of the blocks, so the list stays contiguous); the
// others join the large segments, whose counts are then scanned.
__global__ __launch_bounds__(1024) void SplitSpineKernel(SplitPartial* partial, int grid, std::uint64_t q_cap,
                                                         DeviceCounters* counters) {
  constexpr int kBlock = 1024;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  SplitPartial p = valid ? partial[threadIdx.x] : SplitPartial{0, 0, 0, 0};
  const std::uint64_t small = p.q_base, small_segs = p.fits;
  const std::uint64_t q0 = counters->q_next;  // local survivors appended so far
  const std::uint64_t l0 = counters->l_next, ls0 = counters->l_next_segs;  // earlier splits of the round
  std::uint64_t tot_small;
  const std::uint64_t before = BlockExclusiveSum<kBlock>(small, scratch, tot_small);
  const bool fits = q0 + before + small <= q_cap;
  const std::uint64_t lsegs_in = p.large_segs + (fits ? 0 : small_segs);
  const std::uint64_t lelems_in = p.large_elems + (fits ? 0 : small);
  std::uint64_t tot_segs, tot_elems;
  const std::uint64_t segs = ls0 + BlockExclusiveSum<kBlock>(lsegs_in, scratch, tot_segs);
  const std::uint64_t elems = l0 + BlockExclusiveSum<kBlock>(lelems_in, scratch, tot_elems);
  // Small suffixes taken = those of the blocks that fit.
  std::uint64_t taken_total;
  BlockExclusiveSum<kBlock>(fits ? small : std::uint64_t{0}, scratch, taken_total);
  if (valid) {
    p.q_base = q0 + before;
    p.fits = fits ? 1 : 0;
    p.large_segs = segs;
    p.large_elems = elems;
    partial[threadIdx.x] = p;
  }
  if (threadIdx.x == 0) {
    counters->q_next = q0 + taken_total;
    counters->l_next = l0 + tot_elems;
    counters->l_next_segs = ls0 + tot_segs;
  }
}

template <typename Index>
__global__ __launch_bounds__(256) void SplitAssignKernel(
    const Index* __restrict__ seg_start, const Index* __restrict__ seg_delta, std::uint32_t small_max,
    std::uint64_t chunk, const SplitPartial* __restrict__ partial, Index* __restrict__ dest,
    std::uint32_t* __restrict__ new_id, Index* __restrict__ large_delta, const ListSize* size) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t segs = size->segments;
  const std::uint64_t m = size->survivors;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < segs ? begin + chunk : segs;
  const SplitPartial part = partial[blockIdx.x];
  const bool fits = part.fits != 0;
  std::uint64_t run_q = part.q_base, run_lsegs = part.large_segs, run_lelems = part.large_elems;
  for (std::uint64_t base = begin; base < end; base += kSplitTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kSplitItems;
    // The first slots of the thread's segments and of the one after them
    // (m past the last segment), loaded once for both loops.
    std::uint64_t st[kSplitItems + 1];
#pragma unroll
    for (int i = 0; i <= kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      st[i] = s < segs ? static_cast<std::uint64_t>(seg_start[s]) : m;
    }
    std::uint64_t q_elems = 0, l_segs = 0, l_elems = 0;
#pragma unroll
    for (int i = 0; i < kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      if (s < end) {
        const std::uint64_t size = st[i + 1] - st[i];
        if (fits && size <= small_max) {
          q_elems += size;
        } else {
          l_elems += size;
          ++l_segs;
        }
      }
    }
    std::uint64_t t_q, t_lsegs, t_lelems;
    std::uint64_t q_off = run_q + BlockExclusiveSum<kBlock>(q_elems, scratch, t_q);
    std::uint64_t lseg_off = run_lsegs + BlockExclusiveSum<kBlock>(l_segs, scratch, t_lsegs);
    std::uint64_t lelem_off = run_lelems + BlockExclusiveSum<kBlock>(l_elems, scratch, t_lelems);
#pragma unroll
    for (int i = 0; i < kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      if (s >= end) break;
      const std::uint64_t start = st[i];
      const std::uint64_t size = st[i + 1] - st[i];
      if (fits && size <= small_max) {
        dest[s] = static_cast<Index>(q_off);
        new_id[s] = kSmallSegment;
        q_off += size;
      } else {
        const Index row0 = static_cast<Index>(seg_delta[s] + static_cast<Index>(start));
        dest[s] = static_cast<Index>(lelem_off);
        new_id[s] = static_cast<std::uint32_t>(lseg_off);
        large_delta[lseg_off] = static_cast<Index>(row0 - static_cast<Index>(lelem_off));
        lelem_off += size;
        ++lseg_off;
      }
    }
    run_q += t_q;
    run_lsegs += t_lsegs;
    run_lelems += t_lelems;
  }
}

// One thread per suffix of that list (grid: an upper bound): copy it to its
// segment's destination.
template <typename Index>
__global__ __launch_bounds__(256) void SplitMoveKernel(
    const Index* __restrict__ act_idx, const std::uint32_t* __restrict__ act_seg,
    const Index* __restrict__ seg_start, const Index* __restrict__ seg_delta,
    const Index* __restrict__ dest, const std::uint32_t* __restrict__ new_id,
    Index* __restrict__ q_idx, Index* __restrict__ q_row0, Index* __restrict__ l_idx,
    std::uint32_t* __restrict__ l_seg, const ListSize* size) {
  const std::uint64_t p = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= size->survivors) return;
  const std::uint32_t s = act_seg[p];
  const Index start = seg_start[s];
  const std::uint64_t at = static_cast<std::uint64_t>(dest[s]) + (p - static_cast<std::uint64_t>(start));
  const Index v = act_idx[p];
  const std::uint32_t id = new_id[s];
  if (id == kSmallSegment) {
    q_idx[at] = v;
    q_row0[at] = static_cast<Index>(seg_delta[s] + start);
  } else {
    l_idx[at] = v;
    l_seg[at] = id;
  }
}

// ------------------------------------------------------------ partition -

// Partition of the list RefineApply just wrote by a per-segment flag: flagged
// segments form the "special" list (renumbered in segment order, to x_idx /
// x_seg with deltas x_delta), the others the "stash" list (to y_idx / y_seg,
// first slots y_start, deltas y_delta).  Rows stay those of the input list.
// Same geometry as the split: block b handles segments [b * chunk, (b + 1) *
// chunk).
struct PartitionPartial {
  unsigned long long x_segs, x_elems, y_segs, y_elems;  // exclusive prefixes after the spine
};

template <typename Index>
__global__ __launch_bounds__(256) void PartitionCountKernel(
    const Index* __restrict__ seg_start, const std::uint32_t* __restrict__ flags, std::uint64_t chunk,
    PartitionPartial* __restrict__ partial, const DeviceCounters* counters) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t segs = counters->list.segments;
  const std::uint64_t m = counters->list.survivors;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < segs ? begin + chunk : segs;
  std::uint64_t x_segs = 0, x_elems = 0, y_segs = 0, y_elems = 0;
  for (std::uint64_t s = begin + threadIdx.x; s < end; s += kBlock) {
    const std::uint64_t size = SegmentSize(seg_start, s, segs, m);
    if (flags[s] != 0) {
      x_elems += size;
      ++x_segs;
    } else {
      y_elems += size;
      ++y_segs;
    }
  }
  std::uint64_t t_xs, t_xe, t_ys, t_ye;
  BlockExclusiveSum<kBlock>(x_segs, scratch, t_xs);
  BlockExclusiveSum<kBlock>(x_elems, scratch, t_xe);
  BlockExclusiveSum<kBlock>(y_segs, scratch, t_ys);
  BlockExclusiveSum<kBlock>(y_elems, scratch, t_ye);
  if (threadIdx.x == 0) partial[blockIdx.x] = PartitionPartial{t_xs, t_xe, t_ys, t_ye};
}

// Over at most 1024 blocks: exclusive prefixes of the four counts, totals to
// counters->special and counters->stash.
__global__ __launch_bounds__(1024) void PartitionSpineKernel(PartitionPartial* partial, int grid,
                                                             DeviceCounters* counters) {
  __shared__ std::uint64_t scratch[1024 / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  std::uint64_t totals[4] = {0, 0, 0, 0};
  unsigned long long PartitionPartial::*fields[4] = {&PartitionPartial::x_segs, &PartitionPartial::x_elems,
                                                     &PartitionPartial::y_segs, &PartitionPartial::y_elems};
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {  // one count at a time keeps the block's registers low
    const std::uint64_t v = valid ? static_cast<std::uint64_t>(partial[threadIdx.x].*fields[f]) : 0ull;
    std::uint64_t total;
    const std::uint64_t before = BlockExclusiveSum<1024>(v, scratch, total);
    if (valid) partial[threadIdx.x].*fields[f] = before;
    if (f == 0) totals[0] = total;
    if (f == 1) totals[1] = total;
    if (f == 2) totals[2] = total;
    if (f == 3) totals[3] = total;
  }
  if (threadIdx.x == 0) {
    counters->special = ListSize{totals[1], totals[0]};
    counters->stash = ListSize{totals[3], totals[2]};
  }
}

// Per segment: its destination slot and new id in its list, and that list's
// delta (and, for the stash, first slot).
template <typename Index>
__global__ __launch_bounds__(256) void PartitionAssignKernel(
    const Index* __restrict__ seg_start, const Index* __restrict__ seg_delta,
    const std::uint32_t* __restrict__ flags, std::uint64_t chunk, const PartitionPartial* __restrict__ partial,
    Index* __restrict__ dest, std::uint32_t* __restrict__ new_id, Index* __restrict__ x_delta,
    Index* __restrict__ y_start, Index* __restrict__ y_delta, const DeviceCounters* counters) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t segs = counters->list.segments;
  const std::uint64_t m = counters->list.survivors;
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < segs ? begin + chunk : segs;
  const PartitionPartial part = partial[blockIdx.x];
  std::uint64_t run_xs = part.x_segs, run_xe = part.x_elems, run_ys = part.y_segs, run_ye = part.y_elems;
  for (std::uint64_t base = begin; base < end; base += kSplitTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kSplitItems;
    // As in SplitAssignKernel: first slots loaded once for both loops.
    std::uint64_t st[kSplitItems + 1];
    std::uint32_t fl[kSplitItems];
#pragma unroll
    for (int i = 0; i <= kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      st[i] = s < segs ? static_cast<std::uint64_t>(seg_start[s]) : m;
      if (i < kSplitItems) fl[i] = s < end ? flags[s] : 0u;
    }
    std::uint64_t xs = 0, xe = 0, ys = 0, ye = 0;
#pragma unroll
    for (int i = 0; i < kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      if (s < end) {
        const std::uint64_t size = st[i + 1] - st[i];
        if (fl[i] != 0) {
          xe += size;
          ++xs;
        } else {
          ye += size;
          ++ys;
        }
      }
    }
    std::uint64_t t_xs, t_xe, t_ys, t_ye;
    std::uint64_t xs_off = run_xs + BlockExclusiveSum<kBlock>(xs, scratch, t_xs);
    std::uint64_t xe_off = run_xe + BlockExclusiveSum<kBlock>(xe, scratch, t_xe);
    std::uint64_t ys_off = run_ys + BlockExclusiveSum<kBlock>(ys, scratch, t_ys);
    std::uint64_t ye_off = run_ye + BlockExclusiveSum<kBlock>(ye, scratch, t_ye);
#pragma unroll
    for (int i = 0; i < kSplitItems; ++i) {
      const std::uint64_t s = first + i;
      if (s >= end) break;
      const std::uint64_t start = st[i];
      const std::uint64_t size = st[i + 1] - st[i];
      const Index row0 = static_cast<Index>(seg_delta[s] + static_cast<Index>(start));
      if (fl[i] != 0) {
        dest[s] = static_cast<Index>(xe_off);
        new_id[s] = static_cast<std::uint32_t>(xs_off);
        x_delta[xs_off] = static_cast<Index>(row0 - static_cast<Index>(xe_off));
        xe_off += size;
        ++xs_off;
      } else {
        dest[s] = static_cast<Index>(ye_off);
        new_id[s] = static_cast<std::uint32_t>(ys_off);
        y_start[ys_off] = static_cast<Index>(ye_off);
        y_delta[ys_off] = static_cast<Index>(row0 - static_cast<Index>(ye_off));
        ye_off += size;
        ++ys_off;
      }
    }
    run_xs += t_xs;
    run_xe += t_xe;
    run_ys += t_ys;
    run_ye += t_ye;
  }
}

// One thread per suffix of the input list (grid: an upper bound).
template <typename Index>
__global__ __launch_bounds__(256) void PartitionMoveKernel(
    const Index* __restrict__ act_idx, const std::uint32_t* __restrict__ act_seg,
    const Index* __restrict__ seg_start, const std::uint32_t* __restrict__ flags,
    const Index* __restrict__ dest, const std::uint32_t* __restrict__ new_id, Index* __restrict__ x_idx,
    std::uint32_t* __restrict__ x_seg, Index* __restrict__ y_idx, std::uint32_t* __restrict__ y_seg,
    const DeviceCounters* counters) {
  const std::uint64_t p = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= counters->list.survivors) return;
  const std::uint32_t s = act_seg[p];
  const std::uint64_t at =
      static_cast<std::uint64_t>(dest[s]) + (p - static_cast<std::uint64_t>(seg_start[s]));
  const Index v = act_idx[p];
  if (flags[s] != 0) {
    x_idx[at] = v;
    x_seg[at] = new_id[s];
  } else {
    y_idx[at] = v;
    y_seg[at] = new_id[s];
  }
}

// ------------------------------------------------------------ small list --

// rank[suffix + h] of every small-list suffix, gathered before any rank of
// the round is rewritten.
template <typename Index, typename R = Index>
__global__ __launch_bounds__(256) void QKeygenKernel(
    const Index* __restrict__ q_idx, std::uint64_t qm, const R* __restrict__ rank,
    std::uint64_t n, std::uint64_t h, Index* __restrict__ q_key, DeviceCounters* counters) {
  const std::uint64_t p = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= qm) return;
  const std::uint64_t at = PosOf(q_idx[p]) + h;
  Index key = 0;
  if (at < n) {
    key = static_cast<Index>(rank[at]);
  } else {
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
  }
  q_key[p] = key;
}

struct ExactKey {
  std::uint32_t w0, w1, w2;
};

__device__ inline bool KeyLess(const ExactKey& a, const ExactKey& b) {
  if (a.w0 != b.w0) return a.w0 < b.w0;
  if (a.w1 != b.w1) return a.w1 < b.w1;
  return a.w2 < b.w2;
}
__device__ inline bool KeyEqual(const ExactKey& a, const ExactKey& b) {
  return a.w0 == b.w0 && a.w1 == b.w1 && a.w2 == b.w2;
}
// The same orders without short-circuit evaluation.
__device__ __forceinline__ bool KeyLessFlat(const ExactKey& a, const ExactKey& b) {
  return (a.w0 < b.w0) | ((a.w0 == b.w0) & ((a.w1 < b.w1) | ((a.w1 == b.w1) & (a.w2 < b.w2))));
}
__device__ __forceinline__ bool KeyEqualFlat(const ExactKey& a, const ExactKey& b) {
  return (a.w0 == b.w0) & (a.w1 == b.w1) & (a.w2 == b.w2);
}
template <typename T>
__device__ inline bool KeyLess(T a, T b) {
  return a < b;
}
template <typename T>
__device__ inline bool KeyEqual(T a, T b) {
  return a == b;
}

// First differing symbol of two different exact 96-bit keys (3 bits per
// symbol, first symbol in the top bits of w0).
__device__ inline std::uint32_t ExactKeyMismatch(const ExactKey& a, const ExactKey& b) {
  std::uint32_t bit;
  if (a.w0 != b.w0) {
    bit = static_cast<std::uint32_t>(__clz(a.w0 ^ b.w0));
  } else if (a.w1 != b.w1) {
    bit = 32u + static_cast<std::uint32_t>(__clz(a.w1 ^ b.w1));
  } else {
    bit = 64u + static_cast<std::uint32_t>(__clz(a.w2 ^ b.w2));
  }
  return bit / 3u;
}

// All three words of the exact 96-bit key of a 32-symbol window.
__device__ inline ExactKey ExactKeyOf(std::uint64_t codes, std::uint32_t spec) {
  std::uint64_t hi = 0;
  std::uint32_t lo = 0;
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const std::uint32_t c = static_cast<std::uint32_t>(codes >> (62 - 2 * j)) & 3u;
    const bool special = (spec >> (31 - j)) & 1u;
    const std::uint32_t v = special ? (c == 3u ? 6u : c) : c + 2u;
    hi = (hi << 3) | (lo >> 29);
    lo = (lo << 3) | v;
  }
  return ExactKey{static_cast<std::uint32_t>(hi >> 32), static_cast<std::uint32_t>(hi), lo};
}

// One refine step of the small list.  A block reads kLocalSpan slots,
// kLocalItems consecutive slots per thread, and owns the segments that start
// in its first kLocalOwn slots.  Segment starts come from a block max-scan of
// "row0 changes here" flags, sizes from LDS counters.  An owned suffix counts,
// inside its segment, the keys below its own (its group's first row is row0 +
// that count) and the keys equal to it (its group's size); the first suffix
// of each surviving group reserves the group's slots in the next small list
// (one atomic per block).  After the exact step (not kExact) every suffix's
// rank is already its segment's row0 (the step that formed the segment wrote
// it), so the suffixes with no smaller key in their segment keep their rank
// and only the others store one.
//
// Keys.  kExact (the small list's first step, on segments of equal 16-symbol
// FillKey high halves): the low half of FillKey, which orders windows without
// a separator, N or sentinel exactly; a segment holding such a window compares
// the exact 96-bit keys instead.  Otherwise q_key (rank[suffix + h]).
// packed_lo (kExact, 64-bit values): the values carry the low FillKey half
// above the position, and only windows in special blocks are read; the other
// members of a segment holding a special window rebuild their window codes
// from the high half every member of the segment shares and their own low
// half (their windows hold no special symbol, so FillKey is their codes).
template <typename Index, bool kExact, typename R = Index>
__global__ __launch_bounds__(kLocalThreads) void LocalRefineKernel(
    PackedText t, const Index* __restrict__ q_idx, const Index* __restrict__ q_row0,
    const Index* __restrict__ q_key, std::uint64_t qm, std::uint64_t n,
    R* __restrict__ sa, R* __restrict__ rank, Index* __restrict__ out_idx,
    Index* __restrict__ out_row0, DeviceCounters* counters, const EarlyLcp early,
    bool packed_lo, const std::uint32_t* __restrict__ special_blocks) {
  using KeyT = typename std::conditional<kExact, std::uint32_t, Index>::type;
  constexpr int kItems = kLocalItems;
  constexpr std::uint32_t kSizeMask = 0xFFFFu;  // s_count: size | (specials << 16)
  // Rows (below 2^32 for every text) as 32-bit values: s_row0[k + 1] =
  // slot k, s_row0[0] = slot -1.
  __shared__ std::uint32_t s_row0[kLocalSpan + 1];
  __shared__ KeyT s_key[kLocalSpan];  // slot order, then sorted order in each segment
  __shared__ ExactKey s_exact[kExact ? kLocalSpan : 1];
  __shared__ std::uint32_t s_count[kLocalSpan];  // per segment start: size | (specials << 16)
  __shared__ std::uint32_t s_group[kLocalSpan];  // per group: members counted so far
  __shared__ std::uint32_t s_off[kLocalSpan];
  __shared__ std::uint64_t scratch[kLocalThreads / kWave];
  __shared__ std::uint32_t scratch32[kLocalThreads / kWave];
  __shared__ std::uint64_t s_base;
  const int tid = threadIdx.x;
  const std::uint64_t s0 = static_cast<std::uint64_t>(blockIdx.x) * kLocalOwn;
  const int valid_len =
      static_cast<int>(qm - s0 < static_cast<std::uint64_t>(kLocalSpan) ? qm - s0 : kLocalSpan);
  Index v[kItems], r0[kItems];
  std::uint64_t codes[kItems];
  std::uint32_t spec[kItems];
  // Loads of every item first (then the dependent window loads of every
  // item), so that one thread's loads are in flight together.
  KeyT key[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = tid * kItems + i;
    const bool in = k < valid_len;
    const std::uint64_t at = s0 + (in ? k : 0);
    v[i] = in ? q_idx[at] : Index(0);
    r0[i] = in ? q_row0[at] : Index(0);
    key[i] = 0;
    codes[i] = 0;
    spec[i] = 0;
    if constexpr (!kExact) key[i] = in ? q_key[at] : KeyT(0);
  }
  bool loaded[kItems];  // kExact: the window was read from the text
  if constexpr (kExact) {
    uint4 e0[kItems], e1[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int k = tid * kItems + i;
      const std::uint64_t p = PosOf(v[i]);
      loaded[i] = k < valid_len && (!packed_lo || MaybeSpecialWindow(special_blocks, p));
      if (loaded[i]) {
        e0[i] = t.words[p >> 5];
        e1[i] = t.words[(p >> 5) + 1];
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int k = tid * kItems + i;
      if (loaded[i]) {
        WindowOf(e0[i], e1[i], static_cast<std::uint32_t>(v[i] & 31), codes[i], spec[i]);
        key[i] = static_cast<std::uint32_t>(FillKey(codes[i], spec[i]));
      } else if (k < valid_len) {
        key[i] = static_cast<std::uint32_t>(static_cast<std::uint64_t>(v[i]) >> 32);
        codes[i] = key[i];  // the high half follows if the segment needs it
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = tid * kItems + i;
    s_row0[k + 1] = static_cast<std::uint32_t>(r0[i]);
    s_key[k] = key[i];
    s_count[k] = 0;
    s_group[k] = 0;
  }
  if (tid == 0) s_row0[0] = s0 > 0 ? static_cast<std::uint32_t>(q_row0[s0 - 1]) : 0u;
  __syncthreads();

  // st + 1 of every slot (0: its segment starts before the span).
  std::uint64_t mark[kItems];
  std::uint64_t run = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    // Branch-free on purpose (both LDS reads are in bounds): hipcc for
    // gfx950 dropped the update of `run` when this was written with
    // short-circuit && / || around the LDS reads.
    const int k = tid * kItems + i;
    const std::uint32_t here = s_row0[k + 1];
    const std::uint32_t before = s_row0[k];
    const bool first_slot = (k == 0) & (s0 == 0);
    const bool starts = (k < valid_len) & (first_slot | Differ(here, before));
    run = starts ? static_cast<std::uint64_t>(k) + 1 : run;
    mark[i] = run;
  }
  // Largest mark of the earlier threads (their last slot's mark).
  std::uint64_t all_max;
  const std::uint64_t prev = BlockExclusiveMax<kLocalThreads>(run, scratch, all_max);
  int st[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = tid * kItems + i;
    st[i] = -1;
    const std::uint64_t mk = mark[i] > prev ? mark[i] : prev;
    if (k < valid_len && mk > 0) {
      st[i] = static_cast<int>(mk) - 1;
      atomicAdd(&s_count[st[i]], 1u + ((spec[i] != 0) ? (1u << 16) : 0u));
    }
  }
  __syncthreads();
  if constexpr (kExact) {
    // Exact keys only where a segment holds a special window.
    if (packed_lo) {
      // The segment's shared high half, from a member whose window was read
      // (the special one is); s_off is free until the survivors' offsets.
#pragma unroll
      for (int i = 0; i < kItems; ++i) {
        if (st[i] >= 0 && loaded[i] && (s_count[st[i]] >> 16) != 0) {
          s_off[st[i]] = static_cast<std::uint32_t>(FillKey(codes[i], spec[i]) >> 32);
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kItems; ++i) {
        if (st[i] >= 0 && !loaded[i] && (s_count[st[i]] >> 16) != 0) {
          codes[i] = (static_cast<std::uint64_t>(s_off[st[i]]) << 32) | (codes[i] & 0xFFFFFFFFull);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (st[i] >= 0 && (s_count[st[i]] >> 16) != 0) {
        s_exact[tid * kItems + i] = ExactKeyOf(codes[i], spec[i]);
      }
    }
    __syncthreads();
  }

  // Ranks.  A suffix's rank in its segment is the number of smaller keys
  // (less), scanned kScanUnroll keys per step with their LDS loads issued
  // together (the index clamped to the segment's last slot, the excess keys
  // not counted).  The members of a group share less, so slot st + less of
  // the segment is the group's own: an LDS counter there gives each member
  // its index in the group (tie) and, after a barrier, the group's size; and
  // each member writing its key to slot st + less + tie leaves the segment's
  // keys in sorted order, where slot st + less - 1 holds the largest smaller
  // key (the early LCP of a group start).
  constexpr int kScanUnroll = 4;
  bool owned[kItems], special[kItems];
  std::uint32_t less[kItems], tie[kItems], equal[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int k = tid * kItems + i;
    owned[i] = false;
    special[i] = false;
    less[i] = 0;
    tie[i] = 0;
    equal[i] = 0;
    if (st[i] < 0 || st[i] >= kLocalOwn) continue;  // not owned by this block
    const std::uint32_t packed = s_count[st[i]];
    const std::uint32_t size = packed & kSizeMask;
    if (size > kLocalMaxSegment) {
      atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagLocalSegment));
      continue;
    }
    owned[i] = true;
    const int en = st[i] + static_cast<int>(size);
    std::uint32_t count = 0;
    if (kExact && (packed >> 16) != 0) {
      special[i] = true;
      const ExactKey key = s_exact[k];
      for (int j = st[i]; j < en; j += kScanUnroll) {
        ExactKey kj[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) kj[u] = s_exact[j + u < en ? j + u : en - 1];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) count += (j + u < en) & KeyLessFlat(kj[u], key);
      }
    } else {
      const KeyT key = s_key[k];
      for (int j = st[i]; j < en; j += kScanUnroll) {
        KeyT kj[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) kj[u] = s_key[j + u < en ? j + u : en - 1];
#pragma unroll
        for (int u = 0; u < kScanUnroll; ++u) count += (j + u < en) & (kj[u] < key);
      }
    }
    less[i] = count;
    tie[i] = atomicAdd(&s_group[st[i] + static_cast<int>(count)], 1u);
  }
  __syncthreads();  // every scan read its keys and every group is counted
  ExactKey my_exact[kItems];
  KeyT my_key[kItems];
  if constexpr (kExact) {
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int k = tid * kItems + i;
      if (special[i]) my_exact[i] = s_exact[k];
      my_key[i] = s_key[k];
    }
    __syncthreads();  // keys read before any slot is overwritten
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (!owned[i]) continue;
    const int g = st[i] + static_cast<int>(less[i]);
    equal[i] = s_group[g];
    if constexpr (kExact) {
      if (special[i]) {
        s_exact[g + static_cast<int>(tie[i])] = my_exact[i];
      } else {
        s_key[g + static_cast<int>(tie[i])] = my_key[i];
      }
    }
  }
  if constexpr (kExact) __syncthreads();  // sorted keys in place

  bool survivor[kItems];
  Index head[kItems];
  std::uint32_t leader_total = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    survivor[i] = false;
    head[i] = r0[i];
    if (!owned[i]) continue;
    const int g = st[i] + static_cast<int>(less[i]);
    head[i] = static_cast<Index>(r0[i] + static_cast<Index>(less[i]));
    if constexpr (kExact) {
      // The early LCP of a group start against the group before it: 16 equal
      // symbols and the low halves of FillKey (exact 2-bit codes) in a segment
      // without a special window, the exact 96-bit keys otherwise.
      if (less[i] > 0 && tie[i] == 0 && static_cast<std::uint64_t>(head[i]) < n) {
        std::uint32_t early_lcp;
        if (special[i]) {
          early_lcp = ExactKeyMismatch(s_exact[g - 1], my_exact[i]);
        } else {
          early_lcp = 16u + (static_cast<std::uint32_t>(
                                 __clz(static_cast<std::uint32_t>(s_key[g - 1] ^ my_key[i]))) >> 1);
        }
        if (early_lcp < 32u) WriteEarlyLcp(early, static_cast<std::uint64_t>(head[i]), early_lcp, true);
      }
    }
    const std::uint64_t pos = PosOf(v[i]);
    v[i] = static_cast<Index>(pos);  // survivors go on as plain positions
    if (pos >= n || static_cast<std::uint64_t>(head[i]) >= n) {
      atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
    } else if (equal[i] == 1) {
      sa[head[i]] = static_cast<R>(pos);
      if (kExact || less[i] != 0) rank[pos] = static_cast<R>(head[i]);
    } else {
      if (kExact || less[i] != 0) rank[pos] = static_cast<R>(head[i]);
      survivor[i] = true;
      if (tie[i] == 0) leader_total += equal[i];
    }
  }
  std::uint32_t total;
  std::uint32_t off = BlockExclusiveSum<kLocalThreads>(leader_total, scratch32, total);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (survivor[i] && tie[i] == 0) {
      s_off[st[i] + static_cast<int>(less[i])] = off;
      off += equal[i];
    }
  }
  if (tid == 0) {
    s_base = total ? static_cast<std::uint64_t>(atomicAdd(&counters->q_next, static_cast<unsigned long long>(total))) : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (survivor[i]) {
      const std::uint64_t at = s_base + s_off[st[i] + static_cast<int>(less[i])] + tie[i];
      out_idx[at] = v[i];
      out_row0[at] = head[i];
    }
  }
}

// ------------------------------------------------------------ K9/K10 ------

// SA is a permutation of [0, n) and ISA its inverse: ISA[SA[r]] == r for all
// r with SA[r] < n.  (Injectivity of SA follows from the inverse identity.)
template <typename Index>
__global__ __launch_bounds__(256) void ValidateKernel(
    const Index* __restrict__ sa, const Index* __restrict__ isa, std::uint64_t n,
    DeviceCounters* counters) {
  const std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const std::uint64_t v = sa[r];
  std::uint32_t flags = 0;
  if (v >= n) {
    flags = kFlagNotPermutation;
  } else if (static_cast<std::uint64_t>(isa[v]) != r) {
    flags = kFlagNotInverse;
  }
  if (flags) atomicOr(&counters->error_flags, flags);
}

// Narrowing copy with a range check (construction width -> stored width).
template <typename In, typename Out>
__global__ __launch_bounds__(256) void NarrowKernel(const In* __restrict__ in,
                                                    std::uint64_t n,
                                                    Out* __restrict__ out,
                                                    DeviceCounters* counters) {
  const std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const std::uint64_t v = in[r];
  if (sizeof(Out) < 8 && (v >> (sizeof(Out) < 8 ? 8 * sizeof(Out) : 0)) != 0) {
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
  }
  out[r] = static_cast<Out>(v);
}

template <typename Index>
__global__ __launch_bounds__(256) void IotaKernel(Index* out, std::uint64_t n) {
  const std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n) out[r] = static_cast<Index>(r);
}

}  // namespace sfk
