// Kernels of the suffix-index construction path (gfx950).
//
// Prefix doubling over "segments": a segment is a run of SA rows whose
// suffixes share their first h symbols (a group) and still hold more than
// one suffix.  The active list keeps, per still-unsorted suffix, its text
// position (act_idx) and its dense 32-bit segment id (act_seg); the segment
// table keeps one value per segment, seg_delta = (first SA row of the
// segment) - (its first active slot), so the SA row of active slot k is
// seg_delta[seg] + k (slots only ever move left when survivors are
// compacted, rows never do, so the delta is never negative).  One step:
//   Keygen        key = (segment id << sec_bits) | secondary(suffix)
//   radix sort    stable LSD over the varying digits of the key
//   RefineReduce/RefineSpine/RefineApply
//                 group boundaries (key change), singletons written to SA and
//                 to rank (= their final ISA entry), survivors compacted into
//                 a new active list with renumbered segments, ranks of
//                 survivors set to the SA row where their group starts.
// The rank array is indexed by text position; when no suffix is left active
// it equals the ISA.
#pragma once

#include "block_scan.hpp"
#include "device_common.hpp"

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
};

// Device counters read back by the host after each step.
struct DeviceCounters {
  unsigned long long key_or;
  unsigned long long key_and;
  unsigned long long survivors;
  unsigned long long segments;
  unsigned long long special_active;
  unsigned long long long_items;
  unsigned long long long_words;
  unsigned long long zero_count;  // sentinels seen by the batch packer
  unsigned int error_flags;
  unsigned int pad;
};

// ------------------------------------------------------------ K0 pack -----

// One thread per 32-symbol entry.  Encoded bytes 0..6 -> (code, spec).
__global__ __launch_bounds__(256) void PackTextKernel(
    const std::uint8_t* __restrict__ text, std::uint64_t n, std::uint64_t words,
    uint4* __restrict__ packed, bool aligned, bool batch, DeviceCounters* counters) {
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
// first position, each checked to stay inside its own reference.
template <typename Index, typename OutSa, typename OutIsa>
__global__ __launch_bounds__(256) void LocalizeKernel(
    const Index* sa, const Index* rank, std::uint64_t n, const std::uint64_t* __restrict__ offsets,
    std::uint32_t refs, OutSa* out_sa, OutIsa* out_isa, DeviceCounters* counters) {
  const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const std::uint32_t ref = ReferenceOf(offsets, refs, k);
  const std::uint64_t begin = offsets[ref], end = offsets[ref + 1];
  const std::uint64_t v = sa[k];
  std::uint32_t flags = 0;
  if (v < begin || v >= end) flags |= kFlagNotPermutation;
  std::uint64_t isa_v = 0;
  if (out_isa != nullptr) {
    isa_v = rank[k];
    if (isa_v < begin || isa_v >= end) flags |= kFlagNotInverse;
  }
  if (flags) {
    atomicOr(&counters->error_flags, flags);
    return;
  }
  out_sa[k] = static_cast<OutSa>(v - begin);
  if (out_isa != nullptr) out_isa[k] = static_cast<OutIsa>(isa_v - begin);
}

// ------------------------------------------------------------ keys --------

template <typename Index, typename K>
__global__ __launch_bounds__(256) void KeygenKernel(
    const Index* __restrict__ act_idx, const std::uint32_t* __restrict__ act_seg,
    std::uint64_t m, int sec_kind, int sec_bits, int seg_bits, PackedText t,
    const Index* __restrict__ rank, std::uint64_t h, bool iota,
    K* __restrict__ keys, Index* __restrict__ vals, DeviceCounters* counters) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t s_or[kBlock / kWave];
  __shared__ std::uint64_t s_and[kBlock / kWave];
  std::uint64_t kor = 0, kand = ~0ull;
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kBlock;
  for (std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * kBlock + threadIdx.x; k < m;
       k += stride) {
    const Index v = iota ? static_cast<Index>(k) : act_idx[k];
    std::uint64_t sec;
    if (sec_kind == kSecRank) {
      // An unsorted suffix never reaches the sentinel within h symbols, so
      // v + h < n; the guard only protects against a corrupted state.
      const std::uint64_t p = static_cast<std::uint64_t>(v) + h;
      if (p < t.n) {
        sec = static_cast<std::uint64_t>(rank[p]);
      } else {
        sec = 0;
        atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagRange));
      }
    } else {
      std::uint64_t c;
      std::uint32_t s;
      LoadWindow(t, v, c, s);
      if (sec_kind == kSecFillHi) {
        sec = FillKey(c, s) >> 32;
      } else if (sec_kind == kSecFillLo) {
        sec = FillKey(c, s) & 0xFFFFFFFFull;
      } else {
        sec = ExactKeyWord(c, s, sec_kind - kSecExact0);
      }
    }
    const std::uint64_t seg = seg_bits ? static_cast<std::uint64_t>(act_seg[k]) : 0ull;
    const std::uint64_t key = (seg << sec_bits) | sec;
    keys[k] = static_cast<K>(key);
    vals[k] = v;
    kor |= key;
    kand &= key;
  }
  // One atomic pair per block: wave butterfly, then LDS across waves.
  for (int d = 32; d >= 1; d >>= 1) {
    kor |= WaveShfl(kor, static_cast<int>(LaneId() ^ d));
    kand &= WaveShfl(kand, static_cast<int>(LaneId() ^ d));
  }
  const int wave = threadIdx.x / kWave;
  if (LaneId() == 0) {
    s_or[wave] = kor;
    s_and[wave] = kand;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    std::uint64_t o = 0, a = ~0ull;
    for (int w = 0; w < kBlock / kWave; ++w) {
      o |= s_or[w];
      a &= s_and[w];
    }
    atomicOr(&counters->key_or, static_cast<unsigned long long>(o));
    atomicAnd(&counters->key_and, static_cast<unsigned long long>(a));
  }
}

// Counts active suffixes whose 32-symbol window holds a special symbol.
template <typename Index>
__global__ __launch_bounds__(256) void CountSpecialActiveKernel(
    const Index* __restrict__ act_idx, std::uint64_t m, PackedText t,
    DeviceCounters* counters) {
  const std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool special = false;
  if (k < m) {
    std::uint64_t c;
    std::uint32_t s;
    LoadWindow(t, act_idx[k], c, s);
    special = s != 0;
  }
  const std::uint64_t vote = __ballot(special);
  if (LaneId() == 0 && vote) {
    atomicAdd(&counters->special_active, static_cast<unsigned long long>(__popcll(vote)));
  }
}

// ------------------------------------------------------------ refine ------

struct RefinePartial {
  std::uint64_t survivors;
  std::uint64_t newsegs;
  std::uint64_t last_boundary;  // 1 + index of the last group start, 0 if none
  std::uint64_t pad;
};

template <typename K>
__device__ inline void BoundaryFlags(const K* keys, std::uint64_t k, std::uint64_t m,
                                     bool& starts, bool& single) {
  const K mine = keys[k];
  starts = (k == 0) || keys[k - 1] != mine;
  const bool ends = (k + 1 == m) || keys[k + 1] != mine;
  single = starts && ends;
}

template <typename K>
__global__ __launch_bounds__(256) void RefineReduceKernel(
    const K* __restrict__ keys, std::uint64_t m, std::uint64_t chunk,
    RefinePartial* __restrict__ partial) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t surv = 0, segs = 0, last = 0;
  for (std::uint64_t k = begin + threadIdx.x; k < end; k += kBlock) {
    bool starts, single;
    BoundaryFlags(keys, k, m, starts, single);
    surv += !single;
    segs += starts && !single;
    if (starts) last = k + 1;
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

// Exclusive scan of at most 1024 block partials by one block of 1024.
__global__ __launch_bounds__(1024) void RefineSpineKernel(RefinePartial* partial, int grid,
                                                          DeviceCounters* counters) {
  constexpr int kBlock = 1024;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  const RefinePartial p = valid ? partial[threadIdx.x] : RefinePartial{0, 0, 0, 0};
  std::uint64_t tot_surv, tot_segs, tot_last;
  const std::uint64_t surv = BlockExclusiveSum<kBlock>(p.survivors, scratch, tot_surv);
  const std::uint64_t segs = BlockExclusiveSum<kBlock>(p.newsegs, scratch, tot_segs);
  const std::uint64_t last = BlockExclusiveMax<kBlock>(p.last_boundary, scratch, tot_last);
  if (valid) {
    partial[threadIdx.x].survivors = surv;
    partial[threadIdx.x].newsegs = segs;
    partial[threadIdx.x].last_boundary = last;
  }
  if (threadIdx.x == 0) {
    counters->survivors = tot_surv;
    counters->segments = tot_segs;
  }
}

// Each thread owns kItems consecutive sorted slots of one tile.
template <typename Index, typename K, int kItems>
__global__ __launch_bounds__(256) void RefineApplyKernel(
    const K* __restrict__ keys, const Index* __restrict__ vals, std::uint64_t m,
    std::uint64_t chunk, int sec_bits, const Index* __restrict__ seg_delta,
    const RefinePartial* __restrict__ partial,
    std::uint64_t n, Index* __restrict__ sa, Index* __restrict__ rank, Index* __restrict__ out_idx,
    std::uint32_t* __restrict__ out_seg, Index* __restrict__ out_delta) {
  constexpr int kBlock = 256;
  constexpr int kTile = kBlock * kItems;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t run_surv = partial[blockIdx.x].survivors;
  std::uint64_t run_segs = partial[blockIdx.x].newsegs;
  std::uint64_t run_last = partial[blockIdx.x].last_boundary;
  for (std::uint64_t base = begin; base < end; base += kTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    bool starts[kItems], single[kItems];
    std::uint64_t surv = 0, segs = 0, last = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint64_t k = first + i;
      starts[i] = false;
      single[i] = true;
      if (k < end) {
        BoundaryFlags(keys, k, m, starts[i], single[i]);
        surv += !single[i];
        segs += starts[i] && !single[i];
        if (starts[i]) last = k + 1;
      }
    }
    std::uint64_t tot_surv, tot_segs, tot_last;
    std::uint64_t pos = run_surv + BlockExclusiveSum<kBlock>(surv, scratch, tot_surv);
    std::uint64_t segc = run_segs + BlockExclusiveSum<kBlock>(segs, scratch, tot_segs);
    std::uint64_t cur_last = BlockExclusiveMax<kBlock>(last, scratch, tot_last);
    if (run_last > cur_last) cur_last = run_last;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint64_t k = first + i;
      if (k >= end) break;
      const K key = keys[k];
      const Index v = vals[k];
      const std::uint64_t seg = static_cast<std::uint64_t>(key) >> sec_bits;
      const std::uint64_t delta = static_cast<std::uint64_t>(seg_delta[seg]);
      const std::uint64_t row = delta + k;
      if (starts[i]) cur_last = k + 1;
      if (row >= n || static_cast<std::uint64_t>(v) >= n) continue;  // corrupted state only
      if (single[i]) {
        sa[row] = v;
        rank[v] = static_cast<Index>(row);
      } else {
        // The group start lies in the same segment: every segment start is a
        // key change, hence a group start.
        const std::uint64_t head_row = delta + (cur_last - 1);
        rank[v] = static_cast<Index>(head_row);
        if (starts[i]) {
          out_delta[segc] = static_cast<Index>(row - pos);
          ++segc;
        }
        out_idx[pos] = v;
        out_seg[pos] = static_cast<std::uint32_t>(segc - 1);
        ++pos;
      }
    }
    run_surv += tot_surv;
    run_segs += tot_segs;
    if (tot_last > run_last) run_last = tot_last;
  }
}

// ------------------------------------------------------------ LCP ---------

struct LongLce {
  unsigned long long i;
  unsigned long long j;
  unsigned long long l;
};

// PLCP by irreducible positions.  Position i (with phi = SA[ISA[i]-1]) is
// reducible when T[i-1] == T[phi-1]; then PLCP[i] = PLCP[i-1] - 1.  For every
// other position the LCE is computed directly with 32-symbol word compares.
// Output val[i] = PLCP[i] + i for irreducible i, 0 for reducible i; the
// inclusive max-scan of val then gives PLCP[i] + i everywhere, because
// PLCP[i] + i never decreases along the text.
template <typename Index>
__global__ __launch_bounds__(256) void PlcpIrreducibleKernel(
    PackedText t, const Index* __restrict__ sa, const Index* __restrict__ isa,
    std::uint64_t n, std::uint32_t word_limit, Index* __restrict__ val,
    LongLce* __restrict__ queue, std::uint64_t queue_capacity,
    DeviceCounters* counters) {
  const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const std::uint64_t r = isa[i];
  // Validation fused here: ISA[i] < n and SA[ISA[i]] == i for every i makes
  // ISA a bijection and SA its inverse (the permutation check of the SA).
  if (r >= n || static_cast<std::uint64_t>(sa[r]) != i) {
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagNotInverse));
    val[i] = 0;
    return;
  }
  if (SymbolAt(t, i) == 4u) {  // the sentinel: first row of its reference, LCP 0
    val[i] = static_cast<Index>(i);
    return;
  }
  const std::uint64_t j = sa[r - 1];
  if (j >= n) {  // only possible for a corrupted SA; keeps every load in bounds
    atomicOr(&counters->error_flags, static_cast<unsigned int>(kFlagNotPermutation));
    val[i] = 0;
    return;
  }
  if (i > 0 && j > 0 && SymbolAt(t, i - 1) == SymbolAt(t, j - 1)) {
    val[i] = 0;
    return;
  }
  std::uint64_t l = 0;
  for (std::uint32_t w = 0; w < word_limit; ++w) {
    std::uint64_t ca, cb;
    std::uint32_t sa_, sb;
    LoadWindow(t, i + l, ca, sa_);
    LoadWindow(t, j + l, cb, sb);
    const std::uint32_t mm = WindowMismatch(ca, sa_, cb, sb);
    l += mm;
    if (mm < 32) {
      val[i] = static_cast<Index>(l + i);
      return;
    }
  }
  const unsigned long long slot = atomicAdd(&counters->long_items, 1ull);
  if (slot < queue_capacity) {
    queue[slot].i = i;
    queue[slot].j = j;
    queue[slot].l = l;
  }
  val[i] = 0;
}

// One wave per queued long LCE: 64 windows (2048 symbols) compared per step.
template <typename Index>
__global__ __launch_bounds__(256) void PlcpLongKernel(
    PackedText t, std::uint64_t n, const LongLce* __restrict__ queue,
    std::uint64_t count, Index* __restrict__ val, DeviceCounters* counters) {
  const std::uint64_t item =
      (static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
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
    val[i] = static_cast<Index>(l + i);
    atomicAdd(&counters->long_words, steps * kWave);
  }
}

// Inclusive max-scan, reduce phase: per-block maximum of its chunk.
template <typename Index>
__global__ __launch_bounds__(256) void MaxReduceKernel(
    const Index* __restrict__ in, std::uint64_t m, std::uint64_t chunk,
    std::uint64_t* __restrict__ partial) {
  constexpr int kBlock = 256;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t best = 0;
  for (std::uint64_t k = begin + threadIdx.x; k < end; k += kBlock) {
    const std::uint64_t v = in[k];
    best = v > best ? v : best;
  }
  std::uint64_t total;
  BlockExclusiveMax<kBlock>(best, scratch, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void MaxSpineKernel(std::uint64_t* partial, int grid) {
  constexpr int kBlock = 1024;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const bool valid = static_cast<int>(threadIdx.x) < grid;
  std::uint64_t total;
  const std::uint64_t excl =
      BlockExclusiveMax<kBlock>(valid ? partial[threadIdx.x] : 0ull, scratch, total);
  if (valid) partial[threadIdx.x] = excl;
}

// Inclusive max-scan apply phase; each thread owns kItems consecutive items,
// and the result is turned into PLCP in place: out[i] = scan[i] - i.
template <typename Index, int kItems>
__global__ __launch_bounds__(256) void MaxScanToPlcpKernel(
    Index* __restrict__ data, std::uint64_t m, std::uint64_t chunk,
    const std::uint64_t* __restrict__ partial) {
  constexpr int kBlock = 256;
  constexpr int kTile = kBlock * kItems;
  __shared__ std::uint64_t scratch[kBlock / kWave];
  const std::uint64_t begin = static_cast<std::uint64_t>(blockIdx.x) * chunk;
  const std::uint64_t end = begin + chunk < m ? begin + chunk : m;
  std::uint64_t run = partial[blockIdx.x];
  for (std::uint64_t base = begin; base < end; base += kTile) {
    const std::uint64_t first = base + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    std::uint64_t vals[kItems];
    std::uint64_t local = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint64_t k = first + i;
      vals[i] = k < end ? static_cast<std::uint64_t>(data[k]) : 0ull;
      local = vals[i] > local ? vals[i] : local;
    }
    std::uint64_t total;
    std::uint64_t cur = BlockExclusiveMax<kBlock>(local, scratch, total);
    if (run > cur) cur = run;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const std::uint64_t k = first + i;
      if (k >= end) break;
      cur = vals[i] > cur ? vals[i] : cur;
      data[k] = static_cast<Index>(cur - k);
    }
    if (total > run) run = total;
  }
}

// LCP[r] = PLCP[SA[r]], written in the requested output width.
template <typename Index, typename Out>
__global__ __launch_bounds__(256) void LcpPermuteKernel(
    const Index* __restrict__ plcp, const Index* __restrict__ sa, std::uint64_t n,
    Out* __restrict__ lcp) {
  const std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  lcp[r] = r == 0 ? Out(0) : static_cast<Out>(plcp[sa[r]]);
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
