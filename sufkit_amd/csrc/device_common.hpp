// Device-side building blocks shared by every suffix-index kernel (gfx950).
//
// Text layout in HBM (one copy per build): one 16-byte entry per 32 symbols,
//   entry[w].x,y  uint64 codes of symbols 32w..32w+31, symbol 32w+j in bits
//                 [63-2j, 62-2j] (most significant first, so integer order =
//                 lexicographic order); .x holds the low 32 bits
//   entry[w].z    uint32 special mask, bit (31-j) set when symbol 32w+j is
//                 not a base
//   entry[w].w    unused
// A base A/C/G/T is code 0..3 with its special bit clear.  A special symbol
// has its bit set and its kind in the code slot: 0 sentinel, 1 separator,
// 3 N.  (code, special) is a lossless encoding of the 7-symbol alphabet of the
// encoded text: 0 sentinel < 1 separator < A < C < G < T < N.  Codes and
// mask share an entry so that a 32-symbol window starting anywhere is two
// adjacent 16-byte loads.  Two zero entries follow the last symbol.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sfk {

constexpr int kWave = 64;

struct PackedText {
  const uint4* words;
  std::uint64_t n;
};

// Number of 32-symbol entries for a text of n symbols, padding included.
__host__ __device__ inline std::uint64_t PackedWords(std::uint64_t n) {
  return (n + 31) / 32 + 2;
}

__device__ inline std::uint64_t EntryCodes(const uint4& e) {
  return (static_cast<std::uint64_t>(e.y) << 32) | e.x;
}

// The 32-symbol window starting at offset o (0..31) of entry e0, continued
// in entry e1: 2-bit codes and spec mask.
__device__ inline void WindowOf(const uint4& e0, const uint4& e1, std::uint32_t o,
                                std::uint64_t& codes, std::uint32_t& spec) {
  const std::uint64_t c0 = EntryCodes(e0);
  const std::uint64_t c1 = EntryCodes(e1);
  codes = o ? (c0 << (2 * o)) | (c1 >> (64 - 2 * o)) : c0;
  const std::uint64_t s01 = (static_cast<std::uint64_t>(e0.z) << 32) | e1.z;
  spec = static_cast<std::uint32_t>((s01 << o) >> 32);
}

// The 32-symbol window starting at position i: 2-bit codes and spec mask.
__device__ inline void LoadWindow(const PackedText& t, std::uint64_t i,
                                  std::uint64_t& codes, std::uint32_t& spec) {
  const std::uint64_t w = i >> 5;
  WindowOf(t.words[w], t.words[w + 1], static_cast<std::uint32_t>(i & 31), codes, spec);
}

// Code and spec bit of the single symbol at position p.
__device__ inline std::uint32_t SymbolAt(const PackedText& t, std::uint64_t p) {
  const std::uint64_t w = p >> 5;
  const std::uint32_t o = static_cast<std::uint32_t>(p & 31);
  const uint4 e = t.words[w];
  const std::uint32_t code =
      static_cast<std::uint32_t>(EntryCodes(e) >> (62 - 2 * o)) & 3u;
  const std::uint32_t sp = (e.z >> (31 - o)) & 1u;
  return code | (sp << 2);
}

// 64-bit key of a 32-symbol window: the 2-bit codes, exact when the window
// holds only bases.  From the first special symbol on, the key is filled with
// zeros (sentinel, separator: below every base) or ones (N: above every
// base).  The map is monotone: window x < window y implies key(x) <= key(y).
// Equal keys are exact ties unless the group holds a window with a special.
__device__ inline std::uint64_t FillKey(std::uint64_t codes, std::uint32_t spec) {
  if (spec == 0) return codes;
  const int p = __clz(spec);  // first special symbol in the window
  const std::uint32_t kind =
      static_cast<std::uint32_t>(codes >> (62 - 2 * p)) & 3u;
  const std::uint64_t tail = ~0ull >> (2 * p);
  return kind == 3u ? (codes | tail) : (codes & ~tail);
}

// Exact 96-bit key of a 32-symbol window, 3 bits per symbol with the
// 7-symbol order (0 sentinel, 1 separator, 2..5 bases, 6 N), returned as the
// big-endian 32-bit word `word` (0, 1 or 2).
__device__ inline std::uint32_t ExactKeyWord(std::uint64_t codes,
                                             std::uint32_t spec, int word) {
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
  if (word == 0) return static_cast<std::uint32_t>(hi >> 32);
  if (word == 1) return static_cast<std::uint32_t>(hi);
  return lo;
}

// Position held by a list value.  Every text the backend takes is shorter
// than 2^32 symbols, so a position fits the low 32 bits, and a 64-bit list
// value may carry a 32-bit payload above it (the first step of a 64-bit
// build carries the low FillKey half of the suffix's window there).
template <typename Index>
__device__ __forceinline__ std::uint64_t PosOf(Index v) {
  if constexpr (sizeof(Index) == 8) {
    return static_cast<std::uint64_t>(v) & 0xFFFFFFFFull;
  } else {
    return static_cast<std::uint64_t>(v);
  }
}

// Special blocks: one bit per kSpecialBlock text positions, set when the
// block holds a separator, an N or the sentinel (written by PackTextKernel).
// The 32-symbol window at p can hold a special symbol only when the block of
// p or of p + 31 is special; the bitmap has a spare word past the last block.
constexpr int kSpecialBlockShift = 12;
constexpr std::uint64_t kSpecialBlock = 1ull << kSpecialBlockShift;

__host__ __device__ inline std::uint64_t SpecialBlockWords(std::uint64_t n) {
  return (n >> kSpecialBlockShift) / 32 + 2;
}

__device__ __forceinline__ bool MaybeSpecialWindow(const std::uint32_t* __restrict__ blocks, std::uint64_t p) {
  const std::uint64_t b0 = p >> kSpecialBlockShift;
  const std::uint64_t b1 = (p + 31) >> kSpecialBlockShift;
  return ((blocks[b0 >> 5] >> (b0 & 31)) | (blocks[b1 >> 5] >> (b1 & 31))) & 1u;
}

// Length of the common prefix of two 32-symbol windows (32 if equal).
__device__ inline std::uint32_t WindowMismatch(std::uint64_t ca, std::uint32_t sa,
                                               std::uint64_t cb, std::uint32_t sb) {
  const std::uint64_t xc = ca ^ cb;
  const std::uint32_t xs = sa ^ sb;
  const std::uint32_t mc = xc ? static_cast<std::uint32_t>(__clzll(xc)) >> 1 : 32u;
  const std::uint32_t ms = xs ? static_cast<std::uint32_t>(__clz(xs)) : 32u;
  return mc < ms ? mc : ms;
}

__device__ inline std::uint32_t LaneId() { return __lane_id(); }

// a != b for 64-bit values, decided by a 32-bit compare of the folded
// difference.  Background: the prefix-directory kernel once wrote wrong
// table entries at random on large inputs.  The change that removed the
// fault rewrote that kernel in several ways at once (this compare form, a
// compile-time table layout, unconditional stores, a halo computed by two
// lanes) and also ordered the host's uploads before the device work, so
// which of those removed it is not known; a miscompiled 64-bit compare is
// one unconfirmed candidate.  Since the form costs one 32-bit OR, every
// per-lane 64-bit compare that decides a store (group boundaries of the
// refine steps, segment starts of the small-list step, the directory and
// FM search) uses it, and the full-size order/LCP check of the tests
// (sfk_hip_check_index) guards the result.  The empty asm statement only
// keeps the optimiser from widening the compare back to 64 bits; it emits
// no instruction.
__device__ __forceinline__ bool Differ64(std::uint64_t a, std::uint64_t b) {
  const std::uint64_t x = a ^ b;
  std::uint32_t folded = static_cast<std::uint32_t>(x) | static_cast<std::uint32_t>(x >> 32);
  asm volatile("" : "+v"(folded));
  return folded != 0u;
}

// a != b for keys of either width (64-bit ones through Differ64).
template <typename T>
__device__ __forceinline__ bool Differ(T a, T b) {
  if constexpr (sizeof(T) == 8) {
    return Differ64(static_cast<std::uint64_t>(a), static_cast<std::uint64_t>(b));
  } else {
    return a != b;
  }
}

__device__ inline std::uint64_t LaneMaskLt() {
  return (1ull << LaneId()) - 1ull;
}

// Lanes of this wave whose `bits`-bit digit equals this lane's, among lanes
// with `valid` set.  Every lane of the wave must call this together.  One
// ballot per digit bit; a lane collects, for every other lane, whether any
// bit of that lane's digit differs from its own (the ballot XOR this lane's
// bit spread over 32 bits), so each bit costs a few 32-bit operations and no
// per-lane select.
template <int kBits>
__device__ inline std::uint64_t MatchDigit(std::uint32_t digit, bool valid) {
  const std::uint64_t invalid = ~__ballot(valid);
  std::uint32_t diff_lo = static_cast<std::uint32_t>(invalid);
  std::uint32_t diff_hi = static_cast<std::uint32_t>(invalid >> 32);
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    const std::uint32_t spread = static_cast<std::uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(digit), b, 1));
    const std::uint64_t vote = __builtin_amdgcn_uicmp(spread, 0u, 33);
    diff_lo |= static_cast<std::uint32_t>(vote) ^ spread;
    diff_hi |= static_cast<std::uint32_t>(vote >> 32) ^ spread;
  }
  return ~((static_cast<std::uint64_t>(diff_hi) << 32) | diff_lo);
}

template <typename T>
__device__ inline T WaveShfl(T value, int src) {
  return __shfl(value, src, kWave);
}

__device__ inline std::uint64_t WaveShfl(std::uint64_t value, int src) {
  const std::uint32_t lo = __shfl(static_cast<std::uint32_t>(value), src, kWave);
  const std::uint32_t hi =
      __shfl(static_cast<std::uint32_t>(value >> 32), src, kWave);
  return (static_cast<std::uint64_t>(hi) << 32) | lo;
}

// Inclusive scan across the 64 lanes of a wave (sum).
template <typename T>
__device__ inline T WaveInclusiveSum(T value) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T other = __shfl_up(value, d, kWave);
    if (static_cast<int>(LaneId()) >= d) value += other;
  }
  return value;
}

__device__ inline std::uint64_t WaveInclusiveSum(std::uint64_t value) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const std::uint32_t lo = __shfl_up(static_cast<std::uint32_t>(value), d, kWave);
    const std::uint32_t hi =
        __shfl_up(static_cast<std::uint32_t>(value >> 32), d, kWave);
    const std::uint64_t other = (static_cast<std::uint64_t>(hi) << 32) | lo;
    if (static_cast<int>(LaneId()) >= d) value += other;
  }
  return value;
}

__device__ inline std::uint64_t WaveInclusiveMax(std::uint64_t value) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const std::uint32_t lo = __shfl_up(static_cast<std::uint32_t>(value), d, kWave);
    const std::uint32_t hi =
        __shfl_up(static_cast<std::uint32_t>(value >> 32), d, kWave);
    const std::uint64_t other = (static_cast<std::uint64_t>(hi) << 32) | lo;
    if (static_cast<int>(LaneId()) >= d && other > value) value = other;
  }
  return value;
}

}  // namespace sfk
