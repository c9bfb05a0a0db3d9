// FM-index of a resident index, built on the device from its SA.
//
// The BWT row r holds the symbol before suffix SA[r] (the sentinel for the
// row of the whole text).  Exact search needs, for a base c and a row r,
// occ(c, r) = number of rows below r whose BWT symbol is c, and C[c] = the
// first row of the suffixes starting with c.  Rows are grouped in blocks of
// 64 (one 64-byte record each): per base a 64-bit membership mask of the
// block's rows and the count of that base in all rows before the block, so
// occ is one record read plus a popcount.  Only A/C/G/T get masks: query
// symbols are bases, and every other symbol (sentinel, separator, N) is a
// hard boundary.
//
// Backward search (the reference's SDSL csa_wt search, fm_index.cpp:298-309)
// narrows [lo, hi) from the last pattern symbol to the first:
//   lo = C[c] + occ(c, lo), hi = C[c] + occ(c, hi).
// It starts from the SA interval of the pattern's last k symbols when the
// index has a k-mer prefix directory (suffixes starting with those k symbols
// are exactly that interval), else from C of the last symbol.
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"
#include "sa_kernels.hpp"

namespace sfk {

struct alignas(64) FmBlock {
  unsigned long long mask[4];   // bit j: BWT[64 b + j] is base c (A, C, G, T)
  unsigned long long count[4];  // rows before the block whose BWT symbol is c
};

struct FmHeader {
  unsigned long long first[5];  // C[A], C[C], C[G], C[T], and the end of T's rows
  unsigned long long low;       // symbols below A in the text (sentinels, separators)
};

struct FmCounts {
  unsigned long long c[4];
};

constexpr int kFmItems = 8;                  // records per thread in the count scan
constexpr int kFmTile = 256 * kFmItems;

// One wave per block of 64 rows: each lane finds the BWT symbol of its row
// (a gather from the packed text at SA[r] - 1), the wave ballots one mask per
// base, and lanes 0..7 write the record (masks, then the block's own counts,
// turned into prefix counts by the scan).  An SA value outside the text
// (not a valid SA) sets kFlagRange and counts as no base.
template <typename SaT>
__global__ __launch_bounds__(256) void FmBlocksKernel(PackedText t, const SaT* __restrict__ sa, std::uint64_t n,
                                                      FmBlock* __restrict__ blocks, std::uint64_t nblocks,
                                                      unsigned int* __restrict__ error_flags) {
  const std::uint64_t b = (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  if (b >= nblocks) return;  // wave-uniform
  const std::uint32_t lane = LaneId();
  const std::uint64_t r = b * 64 + lane;
  std::uint32_t sym = 8;
  if (r < n) {
    const std::uint64_t p = static_cast<std::uint64_t>(__builtin_nontemporal_load(&sa[r]));
    if (p >= n) {
      atomicOr(error_flags, static_cast<unsigned int>(kFlagRange));
    } else {
      sym = p == 0 ? 4u : SymbolAt(t, p - 1);
    }
  }
  std::uint64_t mask[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) mask[c] = __ballot(sym == static_cast<std::uint32_t>(c));
  if (lane < 8) {
    const std::uint32_t c = lane & 3u;
    const std::uint64_t m = c == 0 ? mask[0] : c == 1 ? mask[1] : c == 2 ? mask[2] : mask[3];
    unsigned long long* out = lane < 4 ? &blocks[b].mask[c] : &blocks[b].count[c];
    *out = lane < 4 ? m : static_cast<unsigned long long>(__popcll(m));
  }
}

// Symbols below A in the text: a special symbol (spec bit) whose 2-bit code
// is 0 (sentinel) or 1 (separator).  One thread per packed entry.
__global__ __launch_bounds__(256) void FmLowCountKernel(PackedText t, FmHeader* __restrict__ header) {
  __shared__ std::uint64_t scratch[256 / kWave];
  const std::uint64_t entries = (t.n + 31) / 32;
  const std::uint64_t w = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  std::uint64_t low = 0;
  if (w < entries) {
    const uint4 e = t.words[w];
    const std::uint64_t codes = EntryCodes(e);
    std::uint32_t high = 0;  // bit (31 - j): high bit of symbol j's code
#pragma unroll
    for (int j = 0; j < 32; ++j) high |= static_cast<std::uint32_t>((codes >> (63 - 2 * j)) & 1ull) << (31 - j);
    low = static_cast<std::uint64_t>(__popc(e.z & ~high));
  }
  std::uint64_t total;
  BlockExclusiveSum<256>(low, scratch, total);
  if (threadIdx.x == 0 && total > 0) atomicAdd(&header->low, static_cast<unsigned long long>(total));
}

// Count scan over the records: per-chunk totals, one spine block (exclusive
// chunk offsets, C from the totals and the low count), then each chunk
// rewrites its records' counts as exclusive prefixes.
__global__ __launch_bounds__(256) void FmReduceKernel(const FmBlock* __restrict__ blocks, std::uint64_t nblocks,
                                                      std::uint64_t chunk, FmCounts* __restrict__ partial) {
  __shared__ std::uint64_t scratch[256 / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < nblocks ? begin + chunk : nblocks;
  std::uint64_t sum[4] = {0, 0, 0, 0};
  for (std::uint64_t b = begin + threadIdx.x; b < end; b += 256) {
#pragma unroll
    for (int c = 0; c < 4; ++c) sum[c] += blocks[b].count[c];
  }
  FmCounts out;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    std::uint64_t total;
    BlockExclusiveSum<256>(sum[c], scratch, total);
    out.c[c] = total;
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = out;
}

__global__ __launch_bounds__(1024) void FmSpineKernel(FmCounts* __restrict__ partial, int grid,
                                                      FmHeader* __restrict__ header) {
  __shared__ std::uint64_t scratch[1024 / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  std::uint64_t row = 0;
  if (threadIdx.x == 0) row = header->low;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {  // one base at a time keeps the block's registers low
    const std::uint64_t v = valid ? static_cast<std::uint64_t>(partial[threadIdx.x].c[c]) : 0ull;
    std::uint64_t total;
    const std::uint64_t before = BlockExclusiveSum<1024>(v, scratch, total);
    if (valid) partial[threadIdx.x].c[c] = before;
    if (threadIdx.x == 0) {
      header->first[c] = row;
      row += total;
    }
  }
  if (threadIdx.x == 0) header->first[4] = row;
}

__global__ __launch_bounds__(256) void FmApplyKernel(FmBlock* __restrict__ blocks, std::uint64_t nblocks,
                                                     std::uint64_t chunk, const FmCounts* __restrict__ partial) {
  __shared__ std::uint64_t scratch[256 / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < nblocks ? begin + chunk : nblocks;
  std::uint64_t run[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) run[c] = partial[blockIdx.x].c[c];
  for (std::uint64_t base = begin; base < end; base += kFmTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kFmItems;
    std::uint64_t own[kFmItems][4];
    std::uint64_t local[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < kFmItems; ++i) {
      const bool in = first + i < end;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        own[i][c] = in ? static_cast<std::uint64_t>(blocks[first + i].count[c]) : 0ull;
        local[c] += own[i][c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      std::uint64_t total;
      std::uint64_t at = run[c] + BlockExclusiveSum<256>(local[c], scratch, total);
#pragma unroll
      for (int i = 0; i < kFmItems; ++i) {
        if (first + i < end) blocks[first + i].count[c] = at;
        at += own[i][c];
      }
      run[c] += total;
    }
  }
}

// occ(c, r): rows below r whose BWT symbol is base c.
__device__ inline std::uint64_t FmOcc(const FmBlock* __restrict__ blocks, std::uint32_t c, std::uint64_t r) {
  const FmBlock* blk = blocks + (r >> 6);
  const std::uint64_t below = (1ull << (r & 63)) - 1ull;
  return blk->count[c] + static_cast<std::uint64_t>(__popcll(blk->mask[c] & below));
}

// Backward search of a pattern of `len` >= 1 bases whose symbol i (A=0 ..
// T=3) is sym(i).  Returns the SA interval in lo/hi (lo >= hi: no match).
template <typename Sym>
__device__ inline void FmSearch(const FmBlock* __restrict__ blocks, const FmHeader& h, std::uint64_t n,
                                const std::uint64_t* __restrict__ directory, int dir_k, bool dir_wide,
                                std::uint32_t len, Sym sym, std::uint64_t& lo, std::uint64_t& hi) {
  std::int64_t i;
  if (directory != nullptr && len >= static_cast<std::uint32_t>(dir_k)) {
    std::uint64_t key = 0;
    for (std::uint32_t j = len - dir_k; j < len; ++j) key = (key << 2) | sym(j);
    if (dir_wide) {
      lo = directory[key];
      hi = directory[(1ull << (2 * dir_k)) + key];
    } else {
      const std::uint64_t e = directory[key];
      lo = e & 0xFFFFFFFFull;
      hi = e >> 32;
    }
    if (lo >= n) {
      lo = hi = 0;
      return;
    }
    i = static_cast<std::int64_t>(len) - dir_k - 1;
  } else {
    const std::uint32_t c = sym(len - 1);
    lo = h.first[c];
    hi = h.first[c + 1];
    i = static_cast<std::int64_t>(len) - 2;
  }
  // lo <= hi throughout (occ is monotone), so lo < hi is lo != hi.
  for (; i >= 0 && Differ64(lo, hi); --i) {
    const std::uint32_t c = sym(static_cast<std::uint32_t>(i));
    lo = h.first[c] + FmOcc(blocks, c, lo);
    hi = h.first[c] + FmOcc(blocks, c, hi);
  }
}

// Patterns of any length as byte codes (A=0 .. T=3): pattern q is
// codes[offsets[q] .. offsets[q + 1]), searched forward (orientation 0) or as
// its reverse complement (1).  Writes [begin, end) to ranges[2q .. 2q + 1],
// (0, 0) when nothing matches (SDSL's empty result as the reference maps it).
__global__ __launch_bounds__(256) void FmRangesKernel(
    const FmBlock* __restrict__ blocks, const FmHeader* __restrict__ header, std::uint64_t n,
    const std::uint64_t* __restrict__ directory, int dir_k, bool dir_wide, const std::uint8_t* __restrict__ codes,
    const std::uint64_t* __restrict__ offsets, const std::uint8_t* __restrict__ orientation, std::uint64_t count,
    std::uint64_t* __restrict__ ranges) {
  const std::uint64_t q = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const FmHeader h = *header;
  const std::uint64_t at = offsets[q];
  const std::uint32_t len = static_cast<std::uint32_t>(offsets[q + 1] - at);
  std::uint64_t lo = 0, hi = 0;
  if (len > 0) {
    const std::uint8_t* p = codes + at;
    if (orientation[q] == 0) {
      FmSearch(blocks, h, n, directory, dir_k, dir_wide, len, [&](std::uint32_t i) { return p[i] & 3u; }, lo, hi);
    } else {
      FmSearch(blocks, h, n, directory, dir_k, dir_wide, len,
               [&](std::uint32_t i) { return 3u - (p[len - 1 - i] & 3u); }, lo, hi);
    }
  }
  if (lo >= hi) lo = hi = 0;
  ranges[2 * q] = lo;
  ranges[2 * q + 1] = hi;
}

// Count queries in the packed format of CountPatternsKernel (up to 32 codes,
// first symbol in the top bits, one length byte each).
__global__ __launch_bounds__(256) void FmCountPackedKernel(
    const FmBlock* __restrict__ blocks, const FmHeader* __restrict__ header, std::uint64_t n,
    const std::uint64_t* __restrict__ directory, int dir_k, bool dir_wide, const std::uint64_t* __restrict__ patterns,
    const std::uint8_t* __restrict__ lengths, std::uint64_t count, std::uint32_t* __restrict__ out) {
  const std::uint64_t q = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const std::uint64_t p = patterns[q];
  const std::uint32_t len = lengths[q];
  if (len == 0 || len > 32) {
    out[q] = 0;
    return;
  }
  const FmHeader h = *header;
  std::uint64_t lo, hi;
  FmSearch(blocks, h, n, directory, dir_k, dir_wide, len,
           [&](std::uint32_t i) { return static_cast<std::uint32_t>(p >> (62 - 2 * i)) & 3u; }, lo, hi);
  out[q] = lo < hi ? static_cast<std::uint32_t>(hi - lo) : 0u;
}

}  // namespace sfk
