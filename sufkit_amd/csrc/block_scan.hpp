// Block-wide scans built from wave scans and a small LDS exchange.
#pragma once

#include "device_common.hpp"

namespace sfk {

// Exclusive prefix sum over the block; returns this thread's exclusive
// prefix and writes the block total to `total`.  `scratch` must hold
// kBlock/64 elements.  Every thread of the block must call it.
template <int kBlock, typename T>
__device__ inline T BlockExclusiveSum(T value, T* scratch, T& total) {
  constexpr int kWaves = kBlock / kWave;
  const int wave = threadIdx.x / kWave;
  const T incl = WaveInclusiveSum(value);
  if (LaneId() == kWave - 1) scratch[wave] = incl;
  __syncthreads();
  T before = 0;
  T sum = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const T s = scratch[w];
    if (w < wave) before += s;
    sum += s;
  }
  total = sum;
  __syncthreads();
  return before + incl - value;
}

// Inclusive max over the block in thread order (values are unsigned).
template <int kBlock>
__device__ inline std::uint64_t BlockInclusiveMax(std::uint64_t value,
                                                  std::uint64_t* scratch,
                                                  std::uint64_t& total) {
  constexpr int kWaves = kBlock / kWave;
  const int wave = threadIdx.x / kWave;
  const std::uint64_t incl = WaveInclusiveMax(value);
  if (LaneId() == kWave - 1) scratch[wave] = incl;
  __syncthreads();
  std::uint64_t before = 0;
  std::uint64_t all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const std::uint64_t s = scratch[w];
    if (w < wave && s > before) before = s;
    if (s > all) all = s;
  }
  total = all;
  __syncthreads();
  return before > incl ? before : incl;
}

// Exclusive max over the block in thread order (0 for thread 0); writes the
// block-wide max to `total`.
template <int kBlock>
__device__ inline std::uint64_t BlockExclusiveMax(std::uint64_t value,
                                                  std::uint64_t* scratch,
                                                  std::uint64_t& total) {
  constexpr int kWaves = kBlock / kWave;
  const int wave = threadIdx.x / kWave;
  const std::uint64_t incl = WaveInclusiveMax(value);
  std::uint64_t excl = WaveShfl(incl, static_cast<int>((LaneId() + kWave - 1) & (kWave - 1)));
  if (LaneId() == 0) excl = 0;
  if (LaneId() == kWave - 1) scratch[wave] = incl;
  __syncthreads();
  std::uint64_t before = 0;
  std::uint64_t all = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const std::uint64_t s = scratch[w];
    if (w < wave && s > before) before = s;
    if (s > all) all = s;
  }
  total = all;
  __syncthreads();
  return before > excl ? before : excl;
}

}  // namespace sfk
