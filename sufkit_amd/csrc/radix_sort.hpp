// LSD radix sort of (key, value) pairs, one digit per pass, for gfx950.
//
// Each pass is reduce-then-scan over a fixed grid of blocks, each block
// owning one contiguous chunk of whole tiles:
//   upsweep    per-block digit histogram of its chunk (wave ballots collapse
//              equal digits so one LDS atomic serves a whole peer group)
//   spine      one block scans the [digit][block] count matrix into global
//              output offsets
//   downsweep  each block walks its chunk tile by tile: ranks every key of
//              the tile stably with wave ballots (64-lane peer masks), stages
//              the tile in LDS in digit order, and writes it out so that
//              consecutive lanes store consecutive addresses of one bucket.
// Per pass a block reads keys twice and values once, and writes both once:
// HBM traffic per element = key + 2 * (key + value).
// Digits whose bits are constant over all keys are skipped by the caller.
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"

namespace sfk {

struct RadixTile {
  static constexpr int kBlock = 256;
  static constexpr int kDigitBits = 8;
  static constexpr int kRadix = 1 << kDigitBits;
  static constexpr int kMaxGrid = 1024;
};

// Items per thread so that a tile's keys and values fit in ~48 KiB of LDS.
template <typename K, typename V>
struct RadixItems {
  static constexpr int kValue = (sizeof(K) + sizeof(V)) <= 8 ? 16 : ((sizeof(K) + sizeof(V)) <= 12 ? 16 : 12);
};

template <typename K, int kItems>
__global__ __launch_bounds__(RadixTile::kBlock) void RadixUpsweep(
    const K* __restrict__ keys, std::uint64_t m, std::uint64_t chunk,
    int shift, std::uint64_t* __restrict__ counts, int grid) {
  constexpr int kBlock = RadixTile::kBlock;
  constexpr int kRadix = RadixTile::kRadix;
  __shared__ std::uint32_t hist[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kBlock) hist[d] = 0;
  __syncthreads();
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  const std::uint32_t lane = LaneId();
  for (std::uint64_t base = begin; base < end; base += kBlock * kItems) {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const std::uint64_t idx = base + static_cast<std::uint64_t>(j) * kBlock + threadIdx.x;
      const bool valid = idx < end;
      const std::uint32_t d =
          valid ? static_cast<std::uint32_t>(keys[idx] >> shift) & (kRadix - 1) : 0u;
      const std::uint64_t peers = MatchDigit<RadixTile::kDigitBits>(d, valid);
      if (valid && lane == static_cast<std::uint32_t>(__ffsll(static_cast<unsigned long long>(peers)) - 1)) {
        atomicAdd(&hist[d], static_cast<std::uint32_t>(__popcll(peers)));
      }
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kBlock) {
    counts[static_cast<std::uint64_t>(d) * grid + blockIdx.x] = hist[d];
  }
}

// Spine, one block per digit: exclusive scan (in place) of that digit's row
// of per-block counts, and the digit's total in digit_total[digit].  The
// downsweep adds the exclusive prefix of digit totals itself.
__global__ __launch_bounds__(256) void RadixSpine(std::uint64_t* __restrict__ counts,
                                                  int grid,
                                                  std::uint64_t* __restrict__ digit_total) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  std::uint64_t* row = counts + static_cast<std::uint64_t>(blockIdx.x) * grid;
  const int per = (grid + kBlock - 1) / kBlock;  // grid <= kMaxGrid = 1024 -> per <= 4
  const int begin = threadIdx.x * per;
  std::uint64_t vals[4];
  std::uint64_t local = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    vals[i] = (i < per && begin + i < grid) ? row[begin + i] : 0ull;
    local += vals[i];
  }
  std::uint64_t total;
  std::uint64_t running = BlockExclusiveSum<kBlock>(local, scratch, total);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i < per && begin + i < grid) {
      row[begin + i] = running;
      running += vals[i];
    }
  }
  if (threadIdx.x == 0) digit_total[blockIdx.x] = total;
}

template <typename K, typename V, int kItems>
__global__ __launch_bounds__(RadixTile::kBlock) void RadixDownsweep(
    const K* __restrict__ keys_in, const V* __restrict__ vals_in,
    K* __restrict__ keys_out, V* __restrict__ vals_out, std::uint64_t m,
    std::uint64_t chunk, int shift, const std::uint64_t* __restrict__ offsets,
    const std::uint64_t* __restrict__ digit_total, int grid) {
  constexpr int kBlock = RadixTile::kBlock;
  constexpr int kRadix = RadixTile::kRadix;
  constexpr int kWaves = kBlock / kWave;
  constexpr int kTile = kBlock * kItems;
  __shared__ std::uint64_t run_off[kRadix];
  __shared__ std::uint32_t wave_cnt[kWaves][kRadix];
  __shared__ std::uint32_t local_start[kRadix];
  __shared__ std::uint32_t tile_count[kRadix];
  __shared__ std::uint32_t scan_scratch[kBlock / kWave];
  __shared__ K s_keys[kTile];
  __shared__ V s_vals[kTile];

  const int wave = threadIdx.x / kWave;
  const std::uint32_t lane = LaneId();
  {
    static_assert(kRadix == kBlock, "one digit per thread");
    __shared__ std::uint64_t base_scratch[kBlock / kWave];
    std::uint64_t total;
    const std::uint64_t digit_base =
        BlockExclusiveSum<kBlock>(digit_total[threadIdx.x], base_scratch, total);
    run_off[threadIdx.x] =
        digit_base + offsets[static_cast<std::uint64_t>(threadIdx.x) * grid + blockIdx.x];
  }
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  for (std::uint64_t base = begin; base < end; base += kTile) {
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) {
      (&wave_cnt[0][0])[i] = 0;
    }
    __syncthreads();
    K key[kItems];
    V val[kItems];
    std::uint32_t rank[kItems];
    std::uint32_t digit[kItems];
    const std::uint64_t wave_base =
        base + static_cast<std::uint64_t>(wave) * kItems * kWave;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(j) * kWave + lane;
      const bool valid = idx < end;
      key[j] = valid ? keys_in[idx] : K(0);
      val[j] = valid ? vals_in[idx] : V(0);
      const std::uint32_t d =
          static_cast<std::uint32_t>(key[j] >> shift) & (kRadix - 1);
      digit[j] = d;
      const std::uint64_t peers = MatchDigit<RadixTile::kDigitBits>(d, valid);
      const int leader = peers ? __ffsll(static_cast<unsigned long long>(peers)) - 1 : static_cast<int>(lane);
      std::uint32_t old = 0;
      if (valid && static_cast<int>(lane) == leader) {
        old = wave_cnt[wave][d];
        wave_cnt[wave][d] = old + static_cast<std::uint32_t>(__popcll(peers));
      }
      __builtin_amdgcn_wave_barrier();
      old = __shfl(old, leader, kWave);
      rank[j] = old + static_cast<std::uint32_t>(__popcll(peers & LaneMaskLt()));
    }
    __syncthreads();
    // Per digit: prefix over waves (in place) and the tile total.
    for (int d = threadIdx.x; d < kRadix; d += kBlock) {
      std::uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const std::uint32_t c = wave_cnt[w][d];
        wave_cnt[w][d] = s;
        s += c;
      }
      tile_count[d] = s;
    }
    __syncthreads();
    {
      static_assert(kRadix == kBlock, "one digit per thread");
      std::uint32_t total;
      local_start[threadIdx.x] = BlockExclusiveSum<kBlock>(
          tile_count[threadIdx.x], scan_scratch, total);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(j) * kWave + lane;
      if (idx < end) {
        const std::uint32_t d = digit[j];
        const std::uint32_t pos = local_start[d] + wave_cnt[wave][d] + rank[j];
        s_keys[pos] = key[j];
        s_vals[pos] = val[j];
      }
    }
    __syncthreads();
    const std::uint32_t tile_n = static_cast<std::uint32_t>(
        end - base < static_cast<std::uint64_t>(kTile) ? end - base : kTile);
    for (std::uint32_t p = threadIdx.x; p < tile_n; p += kBlock) {
      const K k = s_keys[p];
      const std::uint32_t d = static_cast<std::uint32_t>(k >> shift) & (kRadix - 1);
      const std::uint64_t out = run_off[d] + (p - local_start[d]);
      if (out < m) {
        keys_out[out] = k;
        vals_out[out] = s_vals[p];
      }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadix; d += kBlock) run_off[d] += tile_count[d];
    __syncthreads();
  }
}

}  // namespace sfk
