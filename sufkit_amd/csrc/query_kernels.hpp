// Exact-match count queries on a resident index (packed text + SA).
//
// A pattern is up to 32 A/C/G/T symbols as 2-bit codes, most significant
// first, with its length.  Counting is two binary searches over the SA
// (SuffixArray::Count semantics, suffix_array.cpp:4413-4430): the first row
// whose suffix is >= the pattern on its first L symbols and the first row
// whose suffix is > it.  One 32-symbol window load answers each comparison.
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"

namespace sfk {

// Sign of (suffix at s, first L symbols) - pattern.  A special symbol in the
// window decides by its kind: sentinel and separator sort below every base,
// N above; so a suffix holding one inside the first L symbols never matches.
__device__ inline int ComparePattern(const PackedText& t, std::uint64_t s, std::uint64_t pattern,
                                     std::uint32_t length) {
  std::uint64_t c;
  std::uint32_t sp;
  LoadWindow(t, s, c, sp);
  const std::uint64_t mask = length >= 32 ? ~0ull : ~(~0ull >> (2 * length));
  const std::uint32_t smask = length >= 32 ? ~0u : ~(~0u >> length);
  const std::uint32_t specials = sp & smask;
  if (specials == 0) {
    const std::uint64_t cw = c & mask;
    return cw < pattern ? -1 : (cw > pattern ? 1 : 0);
  }
  const int q = __clz(specials);  // first special within the pattern length
  const std::uint64_t qmask = ~(~0ull >> (2 * q));
  const std::uint64_t cw = c & qmask, pw = pattern & qmask;
  if (cw != pw) return cw < pw ? -1 : 1;
  const std::uint32_t kind = static_cast<std::uint32_t>(c >> (62 - 2 * q)) & 3u;
  return kind == 3u ? 1 : -1;
}

// ------------------------------------------------- fast prefix directory
//
// The exact k-mer directory of a Fast index (fast_prefix_index.cpp:322-382,
// built by SuffixArray::Build and Load, suffix_array.cpp:2087-2107): for each
// k-mer over A/C/G/T (key = 2-bit codes, first symbol most significant) the
// SA interval [begin, end) of the suffixes that start with it; an absent
// k-mer holds (n, n).  Suffixes starting with one k-mer are consecutive SA
// rows, so one pass over the SA finds every interval: the row where a k-mer
// starts writes begin, the row where it ends writes end (each entry is
// written by exactly two rows, no atomics).  Rows up to 2^32 - 1 pack an
// entry into one word (begin low, end high); larger ones use two planes of
// 4^k words (begin, then end).
constexpr std::uint64_t kNoDirectoryKey = ~0ull;
constexpr int kDirectoryItems = 8;
constexpr int kDirectoryTile = 256 * kDirectoryItems;

__device__ inline std::uint64_t DirectoryKey(const PackedText& t, std::uint64_t p, int k) {
  std::uint64_t c;
  std::uint32_t sp;
  LoadWindow(t, p, c, sp);
  // The sentinel is special, so no canonical window reaches past the text.
  const std::uint32_t smask = ~(~0u >> k);
  return (sp & smask) ? kNoDirectoryKey : c >> (64 - 2 * k);
}

__global__ __launch_bounds__(256) void PrefixDirectoryFillKernel(std::uint64_t* __restrict__ table,
                                                                 std::uint64_t words, std::uint64_t value) {
  const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < words) table[i] = value;
}

// counts[0] += rows with a canonical k-mer (indexed k-mers), counts[1] +=
// distinct k-mers (non-empty entries); counts[2 ..] is a sink of
// kDirectorySinkWords words.
//
// Every row stores twice, unconditionally: a row that starts (ends) its
// k-mer's run stores into the table, any other row into its thread's sink
// slot.  The key compares use Differ64 (see device_common.hpp), and the
// table layout is a compile-time choice: with it a run-time branch inside
// the store loop, this kernel lost and misplaced entries on gfx950, while
// the same code specialised per layout was exact in every run.  The two rows
// of the tile's halo are keyed by two lanes of the first wave.
constexpr int kDirectorySinkWords = 2 * 256;

template <typename SaT, bool kWide>
__global__ __launch_bounds__(256) void PrefixDirectoryKernel(PackedText t, const SaT* __restrict__ sa,
                                                             std::uint64_t n, int k,
                                                             std::uint64_t* __restrict__ table,
                                                             unsigned long long* __restrict__ counts) {
  __shared__ std::uint64_t s_key[kDirectoryTile + 2];  // s_key[i + 1] = row base + i
  __shared__ std::uint64_t scratch[256 / kWave];
  const std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * kDirectoryTile;
  const int len = static_cast<int>(n - base < static_cast<std::uint64_t>(kDirectoryTile) ? n - base
                                                                                          : kDirectoryTile);
  std::uint64_t p[kDirectoryItems];
#pragma unroll
  for (int j = 0; j < kDirectoryItems; ++j) {
    const int i = j * 256 + static_cast<int>(threadIdx.x);
    p[j] = i < len ? static_cast<std::uint64_t>(sa[base + i]) : 0ull;
  }
#pragma unroll
  for (int j = 0; j < kDirectoryItems; ++j) {
    const int i = j * 256 + static_cast<int>(threadIdx.x);
    if (i < len) s_key[i + 1] = p[j] < n ? DirectoryKey(t, p[j], k) : kNoDirectoryKey;
  }
  if (threadIdx.x < 2) {  // lane 0: the row before the tile, lane 1: the row after it
    const bool after = threadIdx.x == 1;
    const bool inside = after ? base + len < n : base > 0;
    std::uint64_t key = kNoDirectoryKey;
    if (inside) {
      const std::uint64_t q = static_cast<std::uint64_t>(sa[after ? base + len : base - 1]);
      if (q < n) key = DirectoryKey(t, q, k);
    }
    s_key[after ? len + 1 : 0] = key;
  }
  __syncthreads();
  const std::uint64_t entries = 1ull << (2 * k);
  std::uint64_t* sink = reinterpret_cast<std::uint64_t*>(counts + 2);
  std::uint32_t* table32 = reinterpret_cast<std::uint32_t*>(table);
  std::uint32_t* sink32 = reinterpret_cast<std::uint32_t*>(sink);
  std::uint64_t indexed = 0, distinct = 0;
#pragma unroll
  for (int j = 0; j < kDirectoryItems; ++j) {
    const int i = j * 256 + static_cast<int>(threadIdx.x);
    const bool in = i < len;
    const std::uint64_t key = in ? s_key[i + 1] : kNoDirectoryKey;
    const bool valid = Differ64(key, kNoDirectoryKey);
    const bool starts = valid && Differ64(s_key[i], key);
    const bool ends = valid && Differ64(s_key[i + 2], key);
    const std::uint64_t row = base + static_cast<std::uint64_t>(i);
    indexed += valid ? 1 : 0;
    distinct += starts ? 1 : 0;
    const std::uint64_t slot = valid ? key : 0;
    if constexpr (kWide) {
      std::uint64_t* b = starts ? &table[slot] : &sink[2 * threadIdx.x];
      std::uint64_t* e = ends ? &table[entries + slot] : &sink[2 * threadIdx.x + 1];
      *b = row;
      *e = row + 1;
    } else {
      std::uint32_t* b = starts ? &table32[2 * slot] : &sink32[2 * threadIdx.x];
      std::uint32_t* e = ends ? &table32[2 * slot + 1] : &sink32[2 * threadIdx.x + 1];
      *b = static_cast<std::uint32_t>(row);
      *e = static_cast<std::uint32_t>(row + 1);
    }
  }
  std::uint64_t total_indexed, total_distinct;
  BlockExclusiveSum<256>(indexed, scratch, total_indexed);
  BlockExclusiveSum<256>(distinct, scratch, total_distinct);
  if (threadIdx.x == 0) {
    if (total_indexed) atomicAdd(&counts[0], static_cast<unsigned long long>(total_indexed));
    if (total_distinct) atomicAdd(&counts[1], static_cast<unsigned long long>(total_distinct));
  }
}

// Count queries; a directory (k > 0) narrows both binary searches to the
// SA interval of the pattern's first k symbols when the pattern has that
// many (an absent k-mer counts 0 without searching).
template <typename SaT>
__global__ __launch_bounds__(256) void CountPatternsKernel(
    PackedText t, const SaT* __restrict__ sa, std::uint64_t n,
    const std::uint64_t* __restrict__ patterns, const std::uint8_t* __restrict__ lengths,
    std::uint64_t count, std::uint32_t* __restrict__ out, const std::uint64_t* __restrict__ directory,
    int dir_k, bool dir_wide) {
  const std::uint64_t q = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const std::uint64_t p = patterns[q];
  const std::uint32_t len = lengths[q];
  if (len == 0 || len > 32) {
    out[q] = 0;
    return;
  }
  std::uint64_t lo = 0, hi = n;
  if (directory != nullptr && len >= static_cast<std::uint32_t>(dir_k)) {
    const std::uint64_t key = p >> (64 - 2 * dir_k);
    if (dir_wide) {
      lo = directory[key];
      hi = directory[(1ull << (2 * dir_k)) + key];
    } else {
      const std::uint64_t e = directory[key];
      lo = e & 0xFFFFFFFFull;
      hi = e >> 32;
    }
    if (lo >= n) {
      out[q] = 0;
      return;
    }
  }
  const std::uint64_t range_end = hi;
  while (lo < hi) {
    const std::uint64_t mid = lo + ((hi - lo) >> 1);
    if (ComparePattern(t, sa[mid], p, len) < 0) lo = mid + 1; else hi = mid;
  }
  const std::uint64_t first = lo;
  hi = range_end;
  while (lo < hi) {
    const std::uint64_t mid = lo + ((hi - lo) >> 1);
    if (ComparePattern(t, sa[mid], p, len) <= 0) lo = mid + 1; else hi = mid;
  }
  out[q] = static_cast<std::uint32_t>(lo - first);
}

__device__ inline std::uint64_t SplitMix(std::uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// Deterministic query set: even queries are windows sampled from the text
// (special symbols replaced by random bases), odd queries are uniformly
// random A/C/G/T; lengths uniform in [min_len, max_len].
__global__ __launch_bounds__(256) void GenerateQueriesKernel(
    PackedText t, std::uint64_t seed, std::uint64_t count, std::uint32_t min_len,
    std::uint32_t max_len, std::uint64_t* __restrict__ patterns, std::uint8_t* __restrict__ lengths) {
  const std::uint64_t q = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const std::uint64_t h1 = SplitMix(seed ^ (q * 0xD1B54A32D192ED03ULL));
  const std::uint64_t h2 = SplitMix(h1);
  const std::uint64_t h3 = SplitMix(h2);
  const std::uint32_t len = min_len + static_cast<std::uint32_t>(h1 % (max_len - min_len + 1));
  const std::uint64_t mask = len >= 32 ? ~0ull : ~(~0ull >> (2 * len));
  std::uint64_t codes;
  if ((q & 1) == 0 && t.n > len + 1) {
    std::uint64_t c;
    std::uint32_t sp;
    LoadWindow(t, h2 % (t.n - len - 1), c, sp);
    std::uint64_t special2 = 0;
    for (int j = 0; j < 32; ++j) {
      if ((sp >> (31 - j)) & 1u) special2 |= 3ull << (62 - 2 * j);
    }
    codes = ((c & ~special2) | (h3 & special2)) & mask;
  } else {
    codes = h2 & mask;
  }
  patterns[q] = codes;
  lengths[q] = static_cast<std::uint8_t>(len);
}

}  // namespace sfk
