// Exact-match count queries on a resident index (packed text + SA).
//
// A pattern is up to 32 A/C/G/T symbols as 2-bit codes, most significant
// first, with its length.  Counting is two binary searches over the SA
// (SuffixArray::Count semantics, suffix_array.cpp:4413-4430): the first row
// whose suffix is >= the pattern on its first L symbols and the first row
// whose suffix is > it.  One 32-symbol window load answers each comparison.
#pragma once

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

template <typename SaT>
__global__ __launch_bounds__(256) void CountPatternsKernel(
    PackedText t, const SaT* __restrict__ sa, std::uint64_t n,
    const std::uint64_t* __restrict__ patterns, const std::uint8_t* __restrict__ lengths,
    std::uint64_t count, std::uint32_t* __restrict__ out) {
  const std::uint64_t q = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const std::uint64_t p = patterns[q];
  const std::uint32_t len = lengths[q];
  if (len == 0 || len > 32) {
    out[q] = 0;
    return;
  }
  std::uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const std::uint64_t mid = lo + ((hi - lo) >> 1);
    if (ComparePattern(t, sa[mid], p, len) < 0) lo = mid + 1; else hi = mid;
  }
  const std::uint64_t first = lo;
  hi = n;
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
