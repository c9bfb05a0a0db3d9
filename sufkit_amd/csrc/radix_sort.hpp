// Stable LSD radix sort of (key, value) pairs, 8-bit digits, for gfx950.
//
// The kernel that writes the keys also counts every pass's digits
// (CountDigits / FlushDigits); then one onesweep kernel per digit pass: each
// block ranks its tile (tile = block index) stably in LDS with wave ballots
// (64-lane peer masks), publishes its per-digit counts, finds the counts of
// all earlier tiles by looking back over their published status words (tile
// 0 adds the pass's bucket starts, the exclusive scan of its digit counts),
// and writes the tile out in digit order so that consecutive lanes store
// consecutive addresses of one bucket.  HBM traffic per element and pass: one
// read and one write of key and value (the first pass may generate its input
// instead of reading it, see Source).
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"

namespace sfk {

struct RadixTile {
  static constexpr int kBlock = 256;
  static constexpr int kDigitBits = 8;
  static constexpr int kRadix = 1 << kDigitBits;
};

// ------------------------------------------------------------- onesweep ---
//
// A tile waits only on tiles with smaller block indices, which were
// dispatched before it: the look-back always ends.
//
// Status word per (tile, digit), 64 bits: [63:62] flag (1 = this tile's count,
// 2 = inclusive prefix through this tile), [61:48] epoch of the pass, [47:0]
// count.  Words of other epochs read as "not yet published", so the array is
// cleared only when the 14-bit epoch wraps.  A word is self-contained (no
// other memory is published with it), so its agent-scope atomics are relaxed:
// release/acquire would write back or invalidate the XCD's L2 on every access.

namespace onesweep {
constexpr std::uint64_t kFlagAggregate = 1ull << 62;
constexpr std::uint64_t kFlagInclusive = 2ull << 62;
constexpr std::uint64_t kCountMask = (1ull << 48) - 1;
constexpr std::uint32_t kEpochMask = (1u << 14) - 1;
}  // namespace onesweep

// Tile number of this workgroup in a look-back pass of epoch `epoch`, taken
// from a counter when the workgroup starts: every tile with a lower number
// has started, so every tile a tile waits on is resident or finished,
// whatever else runs on the device.  Numbering by block index is not safe
// on MI355X: the 8 XCDs take workgroups round-robin and each dispatches its
// share only as it has room, so a tile could wait on a predecessor its XCD
// has not dispatched (two such passes in flight, from two builds on one GPU
// in one process or in two, could wait on each other).  The pass of epoch e
// counts on tickets[e & 1]; the tile with ticket 0 clears the other counter,
// which the next pass uses.  Call from every thread; a barrier follows.
__device__ inline std::uint64_t LookBackTile(unsigned int* tickets, std::uint32_t epoch, std::uint32_t* s_tile) {
  if (threadIdx.x == 0) {
    const unsigned int t = atomicAdd(&tickets[epoch & 1u], 1u);
    if (t == 0) tickets[(epoch + 1u) & 1u] = 0u;
    *s_tile = t;
  }
  __syncthreads();
  return *s_tile;
}

// Decoupled look-back of one value over tiles in ticket order
// (LookBackTile): the status word of tile t is status[t * stride].  Publishes this tile's value,
// walks back to the nearest tile whose inclusive value is published, then
// publishes its own inclusive value; returns the exclusive prefix.  Tile 0's
// prefix is `first`.  kMax: values combine by max, else by sum.  The wait is
// bounded: past kSpinLimit polls of one word the tile raises error_bit and
// stops waiting, and every waiting tile stops once the bit is up, so a broken
// protocol ends the kernel with an error instead of hanging it.  Status words
// are self-contained, so their agent-scope atomics are relaxed.
template <bool kMax>
__device__ inline std::uint64_t LookBack(unsigned long long* status, std::uint64_t stride, std::uint64_t tile,
                                         std::uint64_t value, std::uint64_t first, std::uint32_t epoch,
                                         unsigned int* error_flags, std::uint32_t error_bit) {
  constexpr std::uint32_t kSpinLimit = 1u << 20;
  const std::uint64_t ep = static_cast<std::uint64_t>(epoch & onesweep::kEpochMask) << 48;
  auto combine = [](std::uint64_t a, std::uint64_t b) { return kMax ? (a > b ? a : b) : a + b; };
  unsigned long long* slot = status + tile * stride;
  if (tile == 0) {
    __hip_atomic_store(slot, onesweep::kFlagInclusive | ep | combine(first, value), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return first;
  }
  __hip_atomic_store(slot, onesweep::kFlagAggregate | ep | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  std::uint64_t prefix = 0;
  std::int64_t t = static_cast<std::int64_t>(tile) - 1;
  std::uint32_t spins = 0;
  while (true) {
    const std::uint64_t word = __hip_atomic_load(status + static_cast<std::uint64_t>(t) * stride,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const std::uint64_t flag = word & (3ull << 62);
    if (flag == 0 || ((word >> 48) & onesweep::kEpochMask) != (epoch & onesweep::kEpochMask)) {
      // tile t has not published yet in this epoch
      if (++spins >= kSpinLimit) {
        atomicOr(error_flags, error_bit);
        break;
      }
      if ((spins & 63u) == 0 &&
          (__hip_atomic_load(error_flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & error_bit)) {
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    spins = 0;
    prefix = combine(prefix, word & onesweep::kCountMask);
    if (flag == onesweep::kFlagInclusive) break;
    --t;
  }
  __hip_atomic_store(slot, onesweep::kFlagInclusive | ep | combine(prefix, value), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// The same look-back run by a whole wave (every lane calls it with the same
// arguments): each step reads the status words of 64 predecessors at once,
// finds the nearest inclusive one with a ballot and reduces the values up to
// it across the wave, so a tile behind many unfinished tiles moves back 64
// tiles per round trip instead of one.
template <bool kMax>
__device__ __forceinline__ std::uint64_t LookBackWave(unsigned long long* status, std::uint64_t stride, std::uint64_t tile,
                                             std::uint64_t value, std::uint64_t first, std::uint32_t epoch,
                                             unsigned int* error_flags, std::uint32_t error_bit) {
  constexpr std::uint32_t kSpinLimit = 1u << 20;
  const std::uint32_t lane = LaneId();
  const std::uint64_t ep = static_cast<std::uint64_t>(epoch & onesweep::kEpochMask) << 48;
  auto combine = [](std::uint64_t a, std::uint64_t b) { return kMax ? (a > b ? a : b) : a + b; };
  unsigned long long* slot = status + tile * stride;
  if (tile == 0) {
    if (lane == 0) {
      __hip_atomic_store(slot, onesweep::kFlagInclusive | ep | combine(first, value), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    return first;
  }
  if (lane == 0) {
    __hip_atomic_store(slot, onesweep::kFlagAggregate | ep | value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  std::uint64_t prefix = 0;
  std::int64_t hi = static_cast<std::int64_t>(tile) - 1;  // window: tiles hi, hi - 1, ..., hi - 63
  std::uint32_t spins = 0;
  while (true) {
    const std::int64_t t = hi - static_cast<std::int64_t>(lane);
    std::uint64_t word = 0;
    if (t >= 0) {
      word = __hip_atomic_load(status + static_cast<std::uint64_t>(t) * stride, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    const std::uint64_t flag = word & (3ull << 62);
    const bool ready = t >= 0 && flag != 0 &&
                       ((word >> 48) & onesweep::kEpochMask) == (epoch & onesweep::kEpochMask);
    const std::uint64_t inclusive = __ballot(ready && flag == onesweep::kFlagInclusive);
    const std::uint64_t not_ready = __ballot(t >= 0 && !ready);
    // Lanes up to the nearest inclusive tile (or the whole window) must be ready.
    const int last = inclusive ? __ffsll(static_cast<unsigned long long>(inclusive)) - 1 : kWave - 1;
    const std::uint64_t needed = last == kWave - 1 ? ~0ull : ((2ull << last) - 1);
    if (not_ready & needed) {  // wave-uniform, as are all the exits below
      if (++spins >= kSpinLimit) {
        if (lane == 0) atomicOr(error_flags, error_bit);
        break;
      }
      const bool given_up = (spins & 63u) == 0 &&
          (__hip_atomic_load(error_flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & error_bit) != 0;
      if (__ballot(given_up)) break;
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    spins = 0;
    std::uint64_t v = (t >= 0 && static_cast<int>(lane) <= last) ? (word & onesweep::kCountMask) : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = combine(v, WaveShfl(v, static_cast<int>(lane ^ d)));
    prefix = combine(prefix, v);
    if (inclusive) break;
    hi -= kWave;
  }
  if (lane == 0) {
    __hip_atomic_store(slot, onesweep::kFlagInclusive | ep | combine(prefix, value), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  return prefix;
}

// Publishes digit d's count of a tile (the first of its kW lanes stores):
// its inclusive value first + value for tile 0, its aggregate otherwise.  A
// tile publishes as soon as its counts are known, so that later tiles can
// look past it while it does other work before its own look-back.
template <int kW>
__device__ inline void PublishDigit(unsigned long long* status, std::uint64_t tile, int d, std::uint64_t value,
                                    std::uint64_t first, std::uint32_t epoch) {
  constexpr int kRadix = 256;
  if (static_cast<int>(threadIdx.x) % kW != 0) return;
  const std::uint64_t ep = static_cast<std::uint64_t>(epoch & onesweep::kEpochMask) << 48;
  unsigned long long* slot = status + tile * kRadix + static_cast<std::uint64_t>(d);
  const std::uint64_t word = tile == 0 ? (onesweep::kFlagInclusive | ep | (first + value))
                                       : (onesweep::kFlagAggregate | ep | value);
  __hip_atomic_store(slot, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The onesweep look-back of all kRadix digits of a tile by a whole block of
// kW * kRadix threads, after PublishDigit: the kW consecutive lanes of digit
// d = threadIdx.x / kW each read kReach status words of different
// predecessor tiles (lane w's q-th word is tile hi - (q * kW + w)), so one
// round trip covers kW * kReach tiles; ballots find the nearest inclusive
// word and the group sums the counts up to it.  Every thread calls it;
// returns the digit's exclusive prefix (in all kW lanes) and publishes the
// inclusive value.  Epochs and the bounded wait are those of LookBack.  A
// probe of the 1024 x 16 tile found the nearest inclusive tile about 13
// tiles back: 3.3 round trips with one word per lane, 2.0 with four, but the
// four-word form took 5.7 instead of 4.1 us per tile (the agent-scope loads
// themselves are the cost), so one word per lane is the default.
template <int kW, int kReach = 1>
__device__ inline std::uint64_t LookBackDigits(unsigned long long* status, std::uint64_t tile, int d,
                                               std::uint64_t value, std::uint64_t first, std::uint32_t epoch,
                                               unsigned int* error_flags, std::uint32_t error_bit) {
  constexpr int kRadix = 256;
  constexpr std::uint32_t kSpinLimit = 1u << 20;
  static_assert(kW >= 1 && kW <= 64 && (kW & (kW - 1)) == 0, "digit groups within a wave");
  static_assert(kReach >= 1 && kReach <= 8, "status words per lane and round trip");
  constexpr int kSpan = kW * kReach;
  const std::uint64_t ep = static_cast<std::uint64_t>(epoch & onesweep::kEpochMask) << 48;
  const int w = static_cast<int>(threadIdx.x) % kW;
  const std::uint32_t lane = LaneId();
  const std::uint32_t group_base = lane - static_cast<std::uint32_t>(w);
  constexpr std::uint64_t kGroupMask = kW == 64 ? ~0ull : ((1ull << kW) - 1);
  unsigned long long* slot = status + tile * kRadix + static_cast<std::uint64_t>(d);
  if (tile == 0) return first;  // published inclusive by PublishDigit
  std::uint64_t prefix = 0;
  std::int64_t hi = static_cast<std::int64_t>(tile) - 1;
  std::uint32_t spins = 0;
  bool done = false;
  while (true) {
    if (!done) {
      std::uint64_t word[kReach];
#pragma unroll
      for (int q = 0; q < kReach; ++q) {
        const std::int64_t t = hi - (q * kW + w);
        word[q] = onesweep::kFlagInclusive | ep;  // before tile 0: an inclusive zero
        if (t >= 0) {
          word[q] = __hip_atomic_load(status + static_cast<std::uint64_t>(t) * kRadix + static_cast<std::uint64_t>(d),
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // Nearest inclusive word in distance order q * kW + w, and whether
      // every word up to it (or the whole span when there is none) is ready.
      int limit = kSpan;  // distance of the nearest inclusive word, kSpan: none
      bool all_ready = true;
#pragma unroll
      for (int q = 0; q < kReach; ++q) {
        const std::uint64_t flag = word[q] & (3ull << 62);
        const bool ready =
            flag != 0 && ((word[q] >> 48) & onesweep::kEpochMask) == (epoch & onesweep::kEpochMask);
        const bool incl = ready && flag == onesweep::kFlagInclusive;
        const std::uint64_t rmask = (__ballot(ready) >> group_base) & kGroupMask;
        const std::uint64_t imask = (__ballot(incl) >> group_base) & kGroupMask;
        if (limit == kSpan) {
          const int fi = imask ? __ffsll(static_cast<unsigned long long>(imask)) - 1 : kW;
          const std::uint64_t need = fi < kW ? ((2ull << fi) - 1) : kGroupMask;
          if ((rmask & need) != need) all_ready = false;
          if (fi < kW) limit = q * kW + fi;
        }
      }
      if (!all_ready) {
        if (++spins >= kSpinLimit) {
          atomicOr(error_flags, error_bit);
          done = true;
        } else if ((spins & 63u) == 0 &&
                   (__hip_atomic_load(error_flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & error_bit)) {
          done = true;
        } else {
          __builtin_amdgcn_s_sleep(1);
        }
      } else {
        spins = 0;
        std::uint64_t v = 0;
#pragma unroll
        for (int q = 0; q < kReach; ++q) {
          if (q * kW + w <= limit) v += word[q] & onesweep::kCountMask;
        }
#pragma unroll
        for (int o = 1; o < kW; o <<= 1) v += WaveShfl(v, static_cast<int>(lane ^ static_cast<std::uint32_t>(o)));
        prefix += v;
        if (limit < kSpan) {
          done = true;
        } else {
          hi -= kSpan;
        }
      }
    }
    if (__all(done)) break;
  }
  if (w == 0) {
    __hip_atomic_store(slot, onesweep::kFlagInclusive | ep | (prefix + value), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  return prefix;
}

// Digit passes of one sort: the bit offset of each digit that varies, and
// the digit width (8 bits, or 7 where the tuner measured 7-bit passes to be
// cheaper for the key's width).
struct RadixPasses {
  static constexpr int kMax = 8;
  int count;
  int shift[kMax];
  int bits = RadixTile::kDigitBits;
};

// Digit counts of one sort, hist[pass][digit], and the tile counters of its
// passes live in one of two alternating sets; the kernel that counts a sort's
// digits clears the other set for the next sort (which runs after this one on
// the same stream).
struct RadixSet {
  unsigned long long* hist;   // [RadixPasses::kMax][kRadix]
  unsigned int* tiles;        // [RadixPasses::kMax + 1]
  unsigned long long* next_hist;
  unsigned int* next_tiles;
};

constexpr int kRadixSetTickets = RadixPasses::kMax + 1;
constexpr std::uint64_t kRadixSetBytes =
    static_cast<std::uint64_t>(RadixPasses::kMax) * RadixTile::kRadix * 8 + kRadixSetTickets * 4;

// Clears the other set (block 0 only; call before any block's __syncthreads).
__device__ inline void ClearNextRadixSet(const RadixSet& set) {
  if (blockIdx.x != 0) return;
  for (int i = threadIdx.x; i < RadixPasses::kMax * RadixTile::kRadix; i += blockDim.x) set.next_hist[i] = 0;
  if (threadIdx.x < kRadixSetTickets) set.next_tiles[threadIdx.x] = 0;
}

// Adds one key per lane to this wave's LDS digit counts (local[pass][digit]).
// full: every lane of the wave holds a key (wave-uniform), which allows one
// atomic of 64 when all lanes share a digit (keys sorted by their high bits
// often do).
template <typename K>
__device__ inline void CountDigits(std::uint32_t (*local)[RadixTile::kRadix], const RadixPasses& passes,
                                   K key, bool valid, bool full) {
  const std::uint32_t mask = (1u << passes.bits) - 1u;
  for (int p = 0; p < passes.count; ++p) {
    const std::uint32_t d = static_cast<std::uint32_t>(key >> passes.shift[p]) & mask;
    if (full) {
      const std::uint32_t first = __builtin_amdgcn_readfirstlane(d);
      if (__all(d == first)) {
        if (LaneId() == 0) atomicAdd(&local[p][first], static_cast<std::uint32_t>(kWave));
      } else {
        atomicAdd(&local[p][d], 1u);
      }
    } else if (valid) {
      atomicAdd(&local[p][d], 1u);
    }
  }
}

// Copies of the digit counters a counting block keeps in LDS.  LDS atomics
// execute one wave instruction at a time, so waves sharing one copy collide
// no more than one wave alone, and a single copy (8 KiB at eight passes)
// leaves the LDS to occupancy: per-wave copies held a 512-thread counting
// block to two blocks per CU.
constexpr int kDigitCopies = 1;

// Sums the block's copies of the counts into the sort's digit counts (call
// after a __syncthreads that follows the last CountDigits).
template <int kCopies>
__device__ inline void FlushDigits(std::uint32_t (*local)[RadixPasses::kMax][RadixTile::kRadix],
                                   const RadixPasses& passes, unsigned long long* hist) {
  constexpr int kRadix = RadixTile::kRadix;
  for (int i = threadIdx.x; i < passes.count * kRadix; i += blockDim.x) {
    const int p = i / kRadix, d = i % kRadix;
    std::uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < kCopies; ++w) sum += local[w][p][d];
    if (sum) atomicAdd(&hist[p * kRadix + d], static_cast<unsigned long long>(sum));
  }
}

// The largest divisor of `items` that is at most `limit`.
constexpr int FetchGroup(int items, int limit) {
  int g = items < limit ? items : limit;
  while (items % g != 0) --g;
  return g;
}

// Input of a pass read from arrays.
template <typename K, typename V>
struct ArraySource {
  static constexpr int kFetchGroup = 64;  // every item of a tile at once
  const K* __restrict__ keys;
  const V* __restrict__ vals;
  // A source is read in two phases: Fetch issues an item's loads, Make turns
  // what they returned into the item's key and value.  A tile fetches all
  // of its items before making any, so their loads are in flight together.
  struct Raw {
    K key;
    V val;
  };
  __device__ Raw Fetch(std::uint64_t i) const { return Raw{keys[i], vals[i]}; }
  __device__ void Make(const Raw& r, std::uint64_t, K& key, V& val) const {
    key = r.key;
    val = r.val;
  }
};

// Segment of slot k when segment s owns slots [offsets[s], offsets[s+1]),
// searched among segments [lo, hi] (offsets[lo] <= k < offsets[hi + 1]).
__device__ inline std::uint32_t SegmentOfSlot(const std::uint64_t* offsets, std::uint32_t lo, std::uint32_t hi,
                                              std::uint64_t k) {
  hi += 1;
  while (hi - lo > 1) {
    const std::uint32_t mid = (lo + hi) >> 1;
    if (offsets[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

// Placement of a pass that sorts within segments of the slot range: segment
// s owns slots [offsets[s], offsets[s+1]) before and after the pass (the
// input is grouped by segment and each segment stays in place), and an
// element of segment s with digit d goes to
//   adj[s * kRadix + d] + (elements with digit d before it in the input)
// where adj[s][d] = offsets[s] + (elements of s with a digit below d)
//                 - (elements with digit d in segments before s).
// No segment digits are sorted: a batch of references sorts by its 32-bit
// keys in four passes, as one text does.
struct SegmentedPass {
  const std::uint64_t* offsets;  // segments + 1 entries, ending at m
  std::uint32_t segments;
  const long long* adj;          // [segments][kRadix] of this pass
};

// LDS of RadixOnesweep besides its staging buffer, and whether a tile stages
// keys and values one after the other through one buffer (when both at once
// would not fit the 160 KiB a gfx950 workgroup may hold).  A split stage
// costs one more LDS round but lets 12- and 16-byte pairs use tiles of up to
// 16384 elements, whose bucket runs span whole cache lines.
constexpr std::uint64_t kLdsLimit = 160 * 1024;
constexpr std::uint64_t OnesweepFixedLds(int block, int items, bool segmented) {
  return static_cast<std::uint64_t>(block / kWave) * RadixTile::kRadix * 4 + RadixTile::kRadix * (4 + 4 + 8 + 8) +
         static_cast<std::uint64_t>(block / kWave) * (4 + 8) +
         (segmented ? static_cast<std::uint64_t>(block) * items * 2 : 2) + 8;
}
constexpr bool OnesweepSplitStage(int block, int items, int key_bytes, int val_bytes, bool segmented) {
  return static_cast<std::uint64_t>(block) * items * (key_bytes + val_bytes) +
             OnesweepFixedLds(block, items, segmented) > kLdsLimit;
}
// A shape fits when its stage fits the LDS and its per-thread keys and
// values (and segment ids) fit the registers a wave of that block size gets
// without spilling (1024-thread blocks: 128 VGPRs, of which at most 36 for
// the items; the compiler's resource report shows no scratch for any shape
// this admits).
constexpr bool OnesweepFits(int block, int items, int key_bytes, int val_bytes, bool segmented) {
  const int widest = key_bytes > val_bytes ? key_bytes : val_bytes;
  const int item_budget = block >= 1024 ? 144 : 384;
  return static_cast<std::uint64_t>(block) * items * widest + OnesweepFixedLds(block, items, segmented) <= kLdsLimit &&
         items * (key_bytes + val_bytes + (segmented ? 4 : 0)) <= item_budget;
}

// digit_count: this pass's counts per digit (from CountDigits); unused when
// kSegmented (seg gives the placement instead).  kBits: the digit width (the
// status words keep a stride of 256 digits per tile).
template <typename K, typename V, int kBlock, int kItems, typename Source, bool kSegmented = false,
          int kBits = RadixTile::kDigitBits>
__global__ __launch_bounds__(kBlock) void RadixOnesweep(
    Source src, K* __restrict__ keys_out, V* __restrict__ vals_out, std::uint64_t m, int shift,
    const unsigned long long* __restrict__ digit_count, unsigned long long* status,
    std::uint32_t epoch, unsigned int* error_flags,
    std::uint32_t error_bit, unsigned int* tickets,
    SegmentedPass seg = SegmentedPass{nullptr, 0, nullptr}) {
  constexpr int kRadix = 1 << kBits;
  constexpr int kWaves = kBlock / kWave;
  constexpr int kTile = kBlock * kItems;
  static_assert(kBits == 7 || kBits == 8, "7- or 8-bit digits");
  static_assert(!kSegmented || kBits == RadixTile::kDigitBits, "per-reference placement tables hold 8-bit digits");
  static_assert(kBlock >= kRadix, "one digit per thread for the look-back");
  static_assert(!kSegmented || kTile <= 65536, "segment offsets within a tile are 16-bit");
  __shared__ std::uint32_t wave_cnt[kWaves][kRadix];
  __shared__ std::uint32_t local_start[kRadix];
  __shared__ std::uint32_t tile_cnt[kRadix];
  __shared__ std::uint64_t global_base[kRadix];
  __shared__ std::uint64_t bucket_start[kRadix];
  __shared__ std::uint32_t scan_scratch[kBlock / kWave];
  __shared__ std::uint64_t scan64[kBlock / kWave];
  constexpr bool kSplit = OnesweepSplitStage(kBlock, kItems, sizeof(K), sizeof(V), kSegmented);
  constexpr int kStageBytes = kSplit ? kTile * static_cast<int>(sizeof(K) > sizeof(V) ? sizeof(K) : sizeof(V))
                                     : kTile * static_cast<int>(sizeof(K) + sizeof(V));
  __shared__ alignas(16) unsigned char s_stage[kStageBytes];
  K* const s_keys = reinterpret_cast<K*>(s_stage);
  V* const s_vals = reinterpret_cast<V*>(kSplit ? s_stage : s_stage + kTile * sizeof(K));
  __shared__ std::uint16_t s_seg[kSegmented ? kTile : 1];  // segment - first segment of the tile
  __shared__ std::uint32_t s_seg_range[2];

  __shared__ std::uint32_t s_tile;
  const int wave = threadIdx.x / kWave;
  const std::uint32_t lane = LaneId();
  const std::uint64_t tile = LookBackTile(tickets, epoch, &s_tile);
  const std::uint64_t base = tile * kTile;
  const std::uint64_t end = base + kTile < m ? base + kTile : m;
  for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&wave_cnt[0][0])[i] = 0;
  if constexpr (kSegmented) {
    if (threadIdx.x < 2 && base < end) {
      s_seg_range[threadIdx.x] =
          SegmentOfSlot(seg.offsets, 0, seg.segments - 1, threadIdx.x == 0 ? base : end - 1);
    }
  }
  __syncthreads();
  if (!kSegmented && tile == 0) {
    // Bucket starts of this pass (exclusive scan of its digit counts) enter
    // the look-back through tile 0's inclusive values.
    std::uint64_t total;
    const std::uint64_t cnt = threadIdx.x < kRadix ? digit_count[threadIdx.x] : 0ull;
    const std::uint64_t ex = BlockExclusiveSum<kBlock>(cnt, scan64, total);
    if (threadIdx.x < kRadix) bucket_start[threadIdx.x] = ex;
  }
  K key[kItems];
  V val[kItems];
  std::uint32_t rank[kItems];
  std::uint32_t digit[kItems];
  std::uint32_t seg_first = 0, seg_last = 0;
  if constexpr (kSegmented) {
    if (base < end) {
      seg_first = s_seg_range[0];
      seg_last = s_seg_range[1];
    }
  }
  const std::uint64_t wave_base = base + static_cast<std::uint64_t>(wave) * kItems * kWave;
  // All of the tile's loads first, so that they are in flight together (a
  // load next to its item's LDS ranking, or next to the key computed from
  // it, waits out one memory round trip per item).  Every lane loads (slot 0
  // when out of range) and then selects: the branchy form made hipcc spill
  // the 1024 x 16 tile's registers.
  // Items are fetched in groups of at most Source::kFetchGroup (a source
  // whose raw form is wide would otherwise run out of registers).
  constexpr int kGroup = FetchGroup(kItems, Source::kFetchGroup);
#pragma unroll
  for (int g = 0; g < kItems; g += kGroup) {
    typename Source::Raw raw[kGroup];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(g + j) * kWave + lane;
      raw[j] = src.Fetch(idx < end ? idx : 0);
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(g + j) * kWave + lane;
      const bool valid = idx < end;
      K kk;
      V vv;
      src.Make(raw[j], valid ? idx : 0, kk, vv);
      key[g + j] = valid ? kk : K(0);
      val[g + j] = valid ? vv : V(0);
    }
  }
  // Stable rank of every item among the wave's items of its digit so far:
  // every lane reads its digit's count, and the last lane of each group of
  // equal digits writes the count back raised by the group's size.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(j) * kWave + lane;
    const bool valid = idx < end;
    const std::uint32_t d = static_cast<std::uint32_t>(key[j] >> shift) & (kRadix - 1);
    digit[j] = d;
    const std::uint64_t peers = MatchDigit<kBits>(d, valid);
    const std::uint32_t p_lo = static_cast<std::uint32_t>(peers), p_hi = static_cast<std::uint32_t>(peers >> 32);
    const std::uint32_t below = __builtin_amdgcn_mbcnt_hi(p_hi, __builtin_amdgcn_mbcnt_lo(p_lo, 0u));
    const std::uint32_t count = static_cast<std::uint32_t>(__popc(p_lo) + __popc(p_hi));
    const std::uint32_t old = wave_cnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && below + 1 == count) wave_cnt[wave][d] = old + count;
    __builtin_amdgcn_wave_barrier();
    rank[j] = old + below;
  }
  __syncthreads();
  std::uint32_t my_count = 0;
  if (threadIdx.x < kRadix) {
    const int d = threadIdx.x;
    std::uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const std::uint32_t c = wave_cnt[w][d];
      wave_cnt[w][d] = s;
      s += c;
    }
    my_count = s;
    tile_cnt[d] = s;
  }
  __syncthreads();
  // Publish this tile's counts, then (while predecessors finish theirs) scan
  // them into the tile's digit starts and fold those into every wave's
  // offsets, then look back for the exclusive prefixes (the whole block,
  // kBlock / kRadix lanes per digit).
  constexpr int kW = kBlock / kRadix;
  const int dg = static_cast<int>(threadIdx.x) / kW;
  const std::uint64_t first = (!kSegmented && tile == 0) ? bucket_start[dg] : 0;
  PublishDigit<kW>(status, tile, dg, tile_cnt[dg], first, epoch);
  {
    // Exclusive scan of the tile's digit counts (threads >= kRadix add 0).
    std::uint32_t total;
    const std::uint32_t v = threadIdx.x < kRadix ? my_count : 0u;
    const std::uint32_t ex = BlockExclusiveSum<kBlock>(v, scan_scratch, total);
    if (threadIdx.x < kRadix) {
      local_start[threadIdx.x] = ex;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) wave_cnt[w][threadIdx.x] += ex;
    }
  }
  {
    const std::uint64_t p = LookBackDigits<kW>(status, tile, dg, tile_cnt[dg], first, epoch, error_flags, error_bit);
    if (static_cast<int>(threadIdx.x) % kW == 0) global_base[dg] = p;
  }
  __syncthreads();
  if (threadIdx.x < kRadix) {
    // Output slot of staged slot p with digit d: global_base[d] + p from now on.
    std::uint64_t b = global_base[threadIdx.x] - local_start[threadIdx.x];
    if constexpr (kSegmented) {
      // A tile within one segment (the usual case) folds the segment's
      // placement into its digit bases once, instead of per element.
      if (seg_first == seg_last) {
        b += static_cast<std::uint64_t>(seg.adj[static_cast<std::uint64_t>(seg_first) * kRadix + threadIdx.x]);
      }
    }
    global_base[threadIdx.x] = b;
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(j) * kWave + lane;
    if (idx < end) {
      const std::uint32_t d = digit[j];
      const std::uint32_t pos = wave_cnt[wave][d] + rank[j];
      s_keys[pos] = key[j];
      if constexpr (!kSplit) s_vals[pos] = val[j];
      if constexpr (kSegmented) {
        // Segment of the input slot, for a tile that spans segments.
        if (seg_first != seg_last) {
          const std::uint32_t sg = SegmentOfSlot(seg.offsets, seg_first, seg_last, idx);
          s_seg[pos] = static_cast<std::uint16_t>(sg - seg_first);
        }
      }
    }
  }
  __syncthreads();
  const std::uint32_t tile_n = static_cast<std::uint32_t>(end > base ? end - base : 0);
  auto place = [&](std::uint32_t p, std::uint32_t d) {
    std::uint64_t out = global_base[d] + p;
    if constexpr (kSegmented) {
      if (seg_first != seg_last) {
        const std::uint64_t sg = seg_first + s_seg[p];
        out += static_cast<std::uint64_t>(seg.adj[sg * kRadix + d]);
      }
    }
    return out;
  };
  if constexpr (!kSplit) {
    for (std::uint32_t p = threadIdx.x; p < tile_n; p += kBlock) {
      const K k = s_keys[p];
      const std::uint32_t d = static_cast<std::uint32_t>(k >> shift) & (kRadix - 1);
      const std::uint64_t out = place(p, d);
      if (out < m) {
        keys_out[out] = k;
        vals_out[out] = s_vals[p];
      }
    }
  } else {
    // Keys out first, remembering each staged slot's digit; then the values
    // through the same buffer, to the same places.
    std::uint32_t slot_digit[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint32_t p = threadIdx.x + static_cast<std::uint32_t>(i) * kBlock;
      slot_digit[i] = 0;
      if (p < tile_n) {
        const K k = s_keys[p];
        const std::uint32_t d = static_cast<std::uint32_t>(k >> shift) & (kRadix - 1);
        slot_digit[i] = d;
        const std::uint64_t out = place(p, d);
        if (out < m) keys_out[out] = k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const std::uint64_t idx = wave_base + static_cast<std::uint64_t>(j) * kWave + lane;
      if (idx < end) {
        const std::uint32_t d = digit[j];
        s_vals[wave_cnt[wave][d] + rank[j]] = val[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint32_t p = threadIdx.x + static_cast<std::uint32_t>(i) * kBlock;
      if (p < tile_n) {
        const std::uint64_t out = place(p, slot_digit[i]);
        if (out < m) vals_out[out] = s_vals[p];
      }
    }
  }
}

}  // namespace sfk
