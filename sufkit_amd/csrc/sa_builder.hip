// MI355X suffix-index construction backend: host orchestration + C ABI.
//
// Build pipeline for one text of n symbols (encoded bytes + sentinel):
//   K0  PackTextKernel           2-bit codes + special bitmap
//   round 0 (exact 32-symbol ranks from 64-bit packed keys):
//       refine by FillKey bits 63..32 (all n suffixes)
//       refine by FillKey bits 31..0  (suffixes tied on the high half)
//       refine by the exact 96-bit key, 32 bits at a time, only when a tied
//       group holds a window with a separator, N or the sentinel
//   rounds h = 32, 64, 128, ...: refine by rank[suffix + h]
//   K9  ValidateKernel           ISA[SA[r]] == r (rank array == ISA)
//   K6-K8 Φ/PLCP/LCP             irreducible LCEs + max-scan + permutation
// Every refine step is one stable LSD radix sort of (segment, secondary) keys
// over the still-active suffixes, then boundary detection and compaction
// (sa_kernels.hpp).  The host reads back three counters per step to size the
// next one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "host_staging.hpp"
#include "query_kernels.hpp"
#include "radix_sort.hpp"
#include "sa_kernels.hpp"
#include "lcp_kernels.hpp"
#include "sample_kernels.hpp"
#include "fm_kernels.hpp"  // after sa_kernels.hpp: uses its error flags
#include "sufkit_hip.h"

namespace sfk {
namespace {

struct Failure {
  int status;
  std::string message;
};

[[noreturn]] void Fail(int status, const std::string& message) {
  throw Failure{status, message};
}

#define SFK_HIP_CHECK(expr)                                                  \
  do {                                                                       \
    const hipError_t sfk_err_ = (expr);                                      \
    if (sfk_err_ != hipSuccess) {                                            \
      ::sfk::Fail(SFK_HIP_BUILD_FAILURE,                                     \
                  std::string(#expr) + ": " + hipGetErrorString(sfk_err_));  \
    }                                                                        \
  } while (0)

inline std::uint64_t CeilDiv(std::uint64_t a, std::uint64_t b) { return (a + b - 1) / b; }

int BitsFor(std::uint64_t max_value) {
  int bits = 0;
  while (bits < 64 && (max_value >> bits) != 0) ++bits;
  return bits;
}



// Fault injection for tests of the allocation-failure path: when armed with
// k >= 0, the (k+1)-th device workspace allocation from then on fails as an
// out-of-memory hipMalloc would, once, and the hook disarms.
std::atomic<long long> g_fail_allocation{-1};

bool InjectedAllocationFailure() {
  long long left = g_fail_allocation.load();
  while (left >= 0) {
    if (g_fail_allocation.compare_exchange_weak(left, left - 1)) return left == 0;
  }
  return false;
}

// Fault injection for tests of the permutation check: when armed, the next
// build swaps SA rows 1 and 2 after the doubling (once).
std::atomic<bool> g_swap_sa_rows{false};

template <typename R>
__global__ void SwapSaRowsKernel(R* sa) {
  const R a = sa[1];
  sa[1] = sa[2];
  sa[2] = a;
}

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { Release(); }

  // Returns true when the buffer was (re)allocated, i.e. its contents are new.
  // A borrowed buffer (a slice of a caller's workspace) never reallocates.
  bool Reserve(std::uint64_t bytes) {
    if (bytes <= bytes_) return false;
    if (borrowed_) {
      Fail(SFK_HIP_BUILD_FAILURE, "workspace slice of " + std::to_string(bytes_) + " bytes is smaller than the " +
                                      std::to_string(bytes) + " bytes the build needs");
    }
    Release();
    if (InjectedAllocationFailure()) {
      Fail(SFK_HIP_BUILD_FAILURE, "device allocation of " + std::to_string(bytes) + " bytes failed (injected)");
    }
    void* ptr = nullptr;
    const hipError_t err = hipMalloc(&ptr, bytes);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      Fail(SFK_HIP_BUILD_FAILURE, "device allocation of " + std::to_string(bytes) +
                                      " bytes failed: " + hipGetErrorString(err));
    }
    ptr_ = ptr;
    bytes_ = bytes;
    return true;
  }
  void Release() {
    if (ptr_ != nullptr && !borrowed_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
    borrowed_ = false;
  }
  // Uses [ptr, ptr + bytes) of memory the buffer does not own.
  void Borrow(void* ptr, std::uint64_t bytes) {
    Release();
    ptr_ = ptr;
    bytes_ = bytes;
    borrowed_ = true;
  }
  template <typename T>
  T* As() const {
    return static_cast<T*>(ptr_);
  }
  std::uint64_t Bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  std::uint64_t bytes_ = 0;
  bool borrowed_ = false;
};

struct EventPair {
  hipEvent_t begin = nullptr;
  hipEvent_t end = nullptr;
  std::uint64_t bytes = 0;
};

}  // namespace
}  // namespace sfk

struct sfk_hip_builder {
  int device = 0;
  hipStream_t stream = nullptr;
  sfk::DeviceBuffer packed, rank, sa, act_idx, act_seg, keys0, keys1,
      vals0, vals1, seg_delta0, seg_delta1, seg_start, radix_counts, radix_status,
      partial, counters, narrow, ref_offsets, ref_counts, ref_adj, q_idx0, q_idx1, q_row0_0, q_row0_1, q_key,
      special_blocks;
  // Tile tickets of the look-back kernels: two counters, the pass of epoch e
  // numbering its tiles from tickets[e & 1] (see LookBackTile).
  sfk::DeviceBuffer tickets;
  std::uint32_t sort_epoch = 0;  // epoch of the last onesweep pass
  int radix_set = 0;             // radix count set of the next sort
  sfk::DeviceCounters* host_counters = nullptr;  // pinned mirror
  std::vector<sfk::EventPair> events;             // per-kernel profiling
  std::size_t events_used = 0;
  // Phase timing events (all null: the build is not timed).
  hipEvent_t phase[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // Workspace borrowed from a fixed region instead of allocated: buffers
  // never grow and the small lists take their slice as it is.
  bool borrowed_workspace = false;
};

namespace sfk {
namespace {

constexpr int kScanItems = 8;

// The large list's positions; after the doubling the same buffer holds the
// LCP pass's bucket-start bitmap (n / 8 bytes) and long-LCE queue, hence the
// floor for short texts.
std::uint64_t ActIdxBytes(std::uint64_t n, std::uint64_t index_bytes) {
  return std::max<std::uint64_t>(n * index_bytes, 4096);
}

struct BuildContext {
  sfk_hip_builder* b;
  std::uint64_t n;
  bool profile;
  sfk_hip_stats stats;
  // Batch build: reference r owns [offsets[r], offsets[r+1]) (device copy).
  const std::uint64_t* d_offsets = nullptr;
  std::uint32_t refs = 0;
  // Text-position sampling rate K of the outputs (1: complete arrays).
  std::uint32_t sampling = 1;
};

void ResetCounters(sfk_hip_builder* b) {
  DeviceCounters init{};
  *b->host_counters = init;
  SFK_HIP_CHECK(hipMemcpyAsync(b->counters.As<DeviceCounters>(), b->host_counters,
                               sizeof(DeviceCounters), hipMemcpyHostToDevice, b->stream));
}

// Reset of the special-window count (error flags are kept).
__global__ void ResetSpecialCountKernel(DeviceCounters* c) { c->special_active = 0; }

const DeviceCounters& ReadCounters(sfk_hip_builder* b) {
  SFK_HIP_CHECK(hipMemcpyAsync(b->host_counters, b->counters.As<DeviceCounters>(),
                               sizeof(DeviceCounters), hipMemcpyDeviceToHost, b->stream));
  SFK_HIP_CHECK(hipStreamSynchronize(b->stream));
  return *b->host_counters;
}

EventPair* BeginProfiled(BuildContext& ctx, std::uint64_t bytes) {
  if (!ctx.profile) return nullptr;
  sfk_hip_builder* b = ctx.b;
  if (b->events_used == b->events.size()) {
    EventPair p;
    SFK_HIP_CHECK(hipEventCreate(&p.begin));
    SFK_HIP_CHECK(hipEventCreate(&p.end));
    b->events.push_back(p);
  }
  EventPair* p = &b->events[b->events_used++];
  p->bytes = bytes;
  SFK_HIP_CHECK(hipEventRecord(p->begin, b->stream));
  return p;
}

void EndProfiled(BuildContext& ctx, EventPair* p) {
  if (p != nullptr) SFK_HIP_CHECK(hipEventRecord(p->end, ctx.b->stream));
}

struct SortGeometry {
  int grid;
  std::uint64_t chunk;
};

SortGeometry Geometry(std::uint64_t m, int tile, int max_grid) {
  const std::uint64_t tiles = std::max<std::uint64_t>(1, CeilDiv(m, tile));
  std::uint64_t grid = std::min<std::uint64_t>(tiles, static_cast<std::uint64_t>(max_grid));
  const std::uint64_t per = CeilDiv(tiles, grid);
  grid = CeilDiv(tiles, per);
  return SortGeometry{static_cast<int>(grid), per * static_cast<std::uint64_t>(tile)};
}

// ------------------------------------------------------------ radix shapes

// Onesweep shapes (threads x keys per thread).  A build picks one per kind of
// sort (key width, value width, in place per reference or not) and size
// class of the list from a per-device table.  The table starts from
// measured defaults and sfk_hip_autotune replaces it with the fastest
// compiled shape per (kind, class) measured on the device itself.
struct ShapeDesc {
  int block;
  int items;
};
constexpr ShapeDesc kShapes[] = {{256, 16}, {512, 8}, {512, 16}, {1024, 8}, {1024, 12}, {1024, 16}, {512, 24}};
constexpr int kShapeCount = static_cast<int>(sizeof(kShapes) / sizeof(kShapes[0]));
static_assert(kShapeCount == 7, "SortPairs dispatches shapes 0..6");
constexpr std::uint64_t kMinOnesweepTile = 256 * 16;

// Kinds: 0 u32/u32, 1 u64/u32, 2 u32/u64, 3 u64/u64 (key/value), 4 and 5
// the u32/u32 and u32/u64 sorts in place per reference (batch first step).
constexpr int kSortKinds = 6;
// Size classes: class c holds lists of [2^(16 + 2c), 2^(18 + 2c)) elements
// (class 0 also the smaller ones, class 6 the larger ones).
constexpr int kSizeClasses = 7;

int SizeClass(std::uint64_t m) {
  int c = 0;
  while (c + 1 < kSizeClasses && m >= (1ull << (18 + 2 * c))) ++c;
  return c;
}

template <typename K, typename V, bool kSeg>
constexpr int SortKind() {
  if constexpr (kSeg) return sizeof(V) == 4 ? 4 : 5;
  if constexpr (sizeof(K) == 4) return sizeof(V) == 4 ? 0 : 2;
  return sizeof(V) == 4 ? 1 : 3;
}

// Static LDS of RadixOnesweep (its __shared__ arrays) against the 160 KiB a
// gfx950 workgroup may hold (OnesweepFits).
template <typename K, typename V, int kBlock, int kItems, bool kSeg>
constexpr bool ShapeFits() {
  return OnesweepFits(kBlock, kItems, sizeof(K), sizeof(V), kSeg);
}

constexpr bool KindShapeFits(int kind, int shape) {
  const ShapeDesc d = kShapes[shape];
  const int key = kind == 1 || kind == 3 ? 8 : 4;
  const int val = kind == 2 || kind == 3 || kind == 5 ? 8 : 4;
  return OnesweepFits(d.block, d.items, key, val, kind >= 4);
}

// Defaults (measured on MI355X): large tiles write longer runs per bucket and
// look back over fewer tiles; below ~2M elements smaller tiles keep more
// blocks in flight.
int DefaultShape(int kind, std::uint64_t m) {
  if (kind == 0 && m >= (4ull << 20)) return 5;  // 1024 x 16
  if (m >= (2ull << 20)) return 3;               // 1024 x 8
  return 1;                                      // 512 x 8
}

constexpr int kMaxTuneDevices = 64;
// Digit widths a sort may use: index 0 for 8-bit digits, 1 for 7-bit ones
// (the per-reference kinds only have 8-bit placement tables).
constexpr int kDigitWidths = 2;
constexpr int DigitBitsOf(int width) { return width == 0 ? 8 : 7; }
constexpr int WidthOf(int bits) { return bits == 8 ? 0 : 1; }
// -1: the default rule; otherwise an index into kShapes.  Beside each shape,
// its measured time per pass in microseconds (0: not measured), from which a
// sort picks its digit width.
std::atomic<int> g_shape_table[kMaxTuneDevices][kSortKinds][kSizeClasses][kDigitWidths];
std::atomic<double> g_pass_us[kMaxTuneDevices][kSortKinds][kSizeClasses][kDigitWidths];
std::once_flag g_shape_table_once;

void InitShapeTable() {
  std::call_once(g_shape_table_once, [] {
    for (auto& dev : g_shape_table)
      for (auto& kind : dev)
        for (auto& c : kind)
          for (auto& w : c) w.store(-1);
    for (auto& dev : g_pass_us)
      for (auto& kind : dev)
        for (auto& c : kind)
          for (auto& w : c) w.store(0.0);
  });
}

int ShapeFor(int device, int kind, std::uint64_t m, int bits = RadixTile::kDigitBits) {
  InitShapeTable();
  const int tuned = device >= 0 && device < kMaxTuneDevices
                        ? g_shape_table[device][kind][SizeClass(m)][WidthOf(bits)].load(std::memory_order_relaxed)
                        : -1;
  return tuned >= 0 ? tuned : DefaultShape(kind, m);
}

// Digit width of a sort of m keys of key_bits bits: 7 when both widths were
// measured for this kind and size class and ceil(key_bits / 7) 7-bit passes
// take less time than ceil(key_bits / 8) 8-bit ones, else 8.
int DigitBitsFor(int device, int kind, std::uint64_t m, int key_bits) {
  InitShapeTable();
  if (device < 0 || device >= kMaxTuneDevices || kind >= 4) return RadixTile::kDigitBits;
  const int c = SizeClass(m);
  const double us8 = g_pass_us[device][kind][c][0].load(std::memory_order_relaxed);
  const double us7 = g_pass_us[device][kind][c][1].load(std::memory_order_relaxed);
  const int passes8 = (key_bits + 7) / 8, passes7 = (key_bits + 6) / 7;
  if (us8 <= 0.0 || us7 <= 0.0 || passes7 > RadixPasses::kMax) return RadixTile::kDigitBits;
  return passes7 * us7 < passes8 * us8 ? 7 : RadixTile::kDigitBits;
}

std::uint64_t OnesweepStatusBytes(std::uint64_t n) {
  return (CeilDiv(std::max<std::uint64_t>(n, 1), kMinOnesweepTile) + 1) * RadixTile::kRadix * 8;
}

// Reserves the look-back status array; a new array is cleared (its words
// then read as "not published" in every epoch).
void ReserveOnesweep(sfk_hip_builder* b, std::uint64_t n) {
  const bool new_status = b->radix_status.Reserve(OnesweepStatusBytes(n));
  const bool new_tickets = b->tickets.Reserve(2 * sizeof(unsigned int));
  if (new_status || new_tickets) {
    SFK_HIP_CHECK(hipMemsetAsync(b->radix_status.As<void>(), 0, b->radix_status.Bytes(), b->stream));
    SFK_HIP_CHECK(hipMemsetAsync(b->tickets.As<void>(), 0, b->tickets.Bytes(), b->stream));
    b->sort_epoch = 0;
  }
}

// Epoch of the next onesweep pass.  The status words carry 14 epoch bits;
// when they wrap, the array is cleared so no stale word can match.
std::uint32_t NextEpoch(sfk_hip_builder* b) {
  b->sort_epoch += 1;
  if (b->sort_epoch > onesweep::kEpochMask) {
    // Epochs restart at 1, so the ticket parity restarts too: clear both.
    SFK_HIP_CHECK(hipMemsetAsync(b->radix_status.As<void>(), 0, b->radix_status.Bytes(), b->stream));
    SFK_HIP_CHECK(hipMemsetAsync(b->tickets.As<void>(), 0, b->tickets.Bytes(), b->stream));
    b->sort_epoch = 1;
  }
  return b->sort_epoch;
}

// The radix count sets of the builder alternate between consecutive sorts.
constexpr std::uint64_t kRadixSetStride = (kRadixSetBytes + 255) / 256 * 256;

RadixSet TakeRadixSet(sfk_hip_builder* b) {
  char* base = b->radix_counts.As<char>();
  char* cur = base + b->radix_set * kRadixSetStride;
  char* next = base + (1 - b->radix_set) * kRadixSetStride;
  b->radix_set ^= 1;
  constexpr std::uint64_t hist_bytes = static_cast<std::uint64_t>(RadixPasses::kMax) * RadixTile::kRadix * 8;
  return RadixSet{reinterpret_cast<unsigned long long*>(cur), reinterpret_cast<unsigned int*>(cur + hist_bytes),
                  reinterpret_cast<unsigned long long*>(next), reinterpret_cast<unsigned int*>(next + hist_bytes)};
}

// Every digit of `bits` bits below total_bits (keys are not inspected for
// constant digits: that would take a host round trip per step).
RadixPasses AllDigits(int total_bits, int bits = RadixTile::kDigitBits) {
  RadixPasses passes{};
  passes.bits = bits;
  for (int shift = 0; shift < total_bits; shift += bits) passes.shift[passes.count++] = shift;
  return passes;
}

template <typename K, typename V, int kBlock, int kItems, bool kSeg, int kBits, typename Source>
void OnesweepPassesT(BuildContext& ctx, const Source& first, std::uint64_t first_read_bytes, K*& keys, V*& vals,
                     K*& keys_alt, V*& vals_alt, std::uint64_t m, const RadixPasses& passes, const RadixSet& set,
                     const long long* adj) {
  sfk_hip_builder* b = ctx.b;
  constexpr std::uint64_t kTile = static_cast<std::uint64_t>(kBlock) * kItems;
  static_assert(kTile >= kMinOnesweepTile, "status array sized for the smallest tile");
  const std::uint64_t tiles = CeilDiv(m, kTile);
  if ((tiles + 1) * RadixTile::kRadix * 8 > b->radix_status.Bytes()) {
    Fail(SFK_HIP_BUILD_FAILURE, "radix status array smaller than the sort");
  }
  unsigned long long* status = b->radix_status.As<unsigned long long>();
  if (b->tickets.Bytes() < 2 * sizeof(unsigned int)) Fail(SFK_HIP_BUILD_FAILURE, "look-back tickets not reserved");
  unsigned int* error_flags = &b->counters.As<DeviceCounters>()->error_flags;
  for (int p = 0; p < passes.count; ++p) {
    const std::uint32_t epoch = NextEpoch(b);
    const std::uint64_t read = p == 0 ? first_read_bytes : m * (sizeof(K) + sizeof(V));
    const std::uint64_t bytes = read + m * (sizeof(K) + sizeof(V));
    SegmentedPass seg{nullptr, 0, nullptr};
    if constexpr (kSeg) {
      seg = SegmentedPass{ctx.d_offsets, ctx.refs,
                          adj + static_cast<std::uint64_t>(p) * ctx.refs * RadixTile::kRadix};
    }
    EventPair* ev = BeginProfiled(ctx, bytes);
    if (p == 0) {
      RadixOnesweep<K, V, kBlock, kItems, Source, kSeg, kBits><<<static_cast<unsigned>(tiles), kBlock, 0, b->stream>>>(
          first, keys_alt, vals_alt, m, passes.shift[p], set.hist + p * RadixTile::kRadix, status, epoch,
          error_flags, kFlagLookback, b->tickets.As<unsigned int>(), seg);
    } else {
      RadixOnesweep<K, V, kBlock, kItems, ArraySource<K, V>, kSeg, kBits>
          <<<static_cast<unsigned>(tiles), kBlock, 0, b->stream>>>(
              ArraySource<K, V>{keys, vals}, keys_alt, vals_alt, m, passes.shift[p],
              set.hist + p * RadixTile::kRadix, status, epoch, error_flags, kFlagLookback,
              b->tickets.As<unsigned int>(), seg);
    }
    EndProfiled(ctx, ev);
    SFK_HIP_CHECK(hipGetLastError());
    std::swap(keys, keys_alt);
    std::swap(vals, vals_alt);
    ctx.stats.sort_passes += 1;
    ctx.stats.algorithmic_bytes_sa += bytes;
  }
}

// 7-bit passes exist for the sorts of arrays (not for the per-reference
// kinds, whose placement tables hold 8-bit digits, nor for a first pass that
// generates its input from the text, whose keys are 32 bits wide).
template <typename K, typename V, bool kSeg, typename Source>
constexpr bool kSevenBitSort = !kSeg && std::is_same<Source, ArraySource<K, V>>::value;

template <typename K, typename V, int kShape, bool kSeg, typename Source>
void OnesweepPassesShape(BuildContext& ctx, const Source& first, std::uint64_t first_read_bytes, K*& keys,
                         V*& vals, K*& keys_alt, V*& vals_alt, std::uint64_t m, const RadixPasses& passes,
                         const RadixSet& set, const long long* adj) {
  constexpr ShapeDesc d = kShapes[kShape];
  if constexpr (ShapeFits<K, V, d.block, d.items, kSeg>()) {
    if (passes.bits == RadixTile::kDigitBits) {
      OnesweepPassesT<K, V, d.block, d.items, kSeg, RadixTile::kDigitBits>(ctx, first, first_read_bytes, keys, vals,
                                                                          keys_alt, vals_alt, m, passes, set, adj);
      return;
    }
    if constexpr (kSevenBitSort<K, V, kSeg, Source>) {
      if (passes.bits == 7) {
        OnesweepPassesT<K, V, d.block, d.items, kSeg, 7>(ctx, first, first_read_bytes, keys, vals, keys_alt,
                                                         vals_alt, m, passes, set, adj);
        return;
      }
    }
    Fail(SFK_HIP_BUILD_FAILURE, "radix digit width not compiled for this sort");
  } else {
    Fail(SFK_HIP_BUILD_FAILURE, "radix shape does not fit the LDS of a workgroup");
  }
}

// Stable LSD sort of m (key, value) pairs over `passes`, whose digit counts
// are already in set.hist (or, kSeg, whose placement tables are in adj).
// The first pass reads `first` (the arrays keys/vals, or input generated on
// the fly); on return keys/vals point at the sorted buffers.  Each pass is
// one onesweep kernel of the shape the device's table gives.
template <typename K, typename V, bool kSeg = false, typename Source>
void SortPairs(BuildContext& ctx, const Source& first, std::uint64_t first_read_bytes, K*& keys, V*& vals,
               K* keys_alt, V* vals_alt, std::uint64_t m, const RadixPasses& passes, const RadixSet& set,
               const long long* adj = nullptr, int shape = -1) {
  if (passes.count == 0 || m == 0) return;
  if (shape < 0) shape = ShapeFor(ctx.b->device, SortKind<K, V, kSeg>(), m, passes.bits);
  switch (shape) {
    case 0: OnesweepPassesShape<K, V, 0, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 1: OnesweepPassesShape<K, V, 1, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 2: OnesweepPassesShape<K, V, 2, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 3: OnesweepPassesShape<K, V, 3, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 4: OnesweepPassesShape<K, V, 4, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 5: OnesweepPassesShape<K, V, 5, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    case 6: OnesweepPassesShape<K, V, 6, kSeg>(ctx, first, first_read_bytes, keys, vals, keys_alt, vals_alt, m, passes, set, adj); break;
    default: Fail(SFK_HIP_BUILD_FAILURE, "unknown radix shape");
  }
}

// Clears the count set a sort did not count into (the next sort's set; the
// counting kernels of the other sorts do this themselves).
__global__ void ClearNextRadixSetKernel(RadixSet set) { ClearNextRadixSet(set); }

// The first step of a batch of references whose layout fits 32-bit counts:
// every reference's suffixes sorted by the high half of FillKey, in place in
// the reference's own slot range, in four passes whose first generates its
// input from the text.  Placement tables come from per-reference digit
// counts (SegmentedPass).
constexpr std::uint32_t kMaxSegmentedRefs = 1u << 16;

template <typename Index>
void SortFillHiSegmented(BuildContext& ctx, PackedText text, std::uint32_t*& keys, Index*& vals,
                         std::uint32_t* keys_alt, Index* vals_alt, std::uint64_t m, const RadixSet& set) {
  sfk_hip_builder* b = ctx.b;
  const std::uint32_t refs = ctx.refs;
  constexpr int kRadix = RadixTile::kRadix;
  const std::uint64_t table = static_cast<std::uint64_t>(refs) * kFillHiPasses * kRadix;
  b->ref_counts.Reserve(table * sizeof(unsigned int));
  b->ref_adj.Reserve(table * sizeof(long long));
  unsigned int* counts = b->ref_counts.As<unsigned int>();
  long long* adj = b->ref_adj.As<long long>();
  SFK_HIP_CHECK(hipMemsetAsync(counts, 0, table * sizeof(unsigned int), b->stream));
  ClearNextRadixSetKernel<<<1, 256, 0, b->stream>>>(set);
  constexpr std::uint64_t kChunk = 1ull << 16;
  const RadixPasses passes = AllDigits(32);
  TextFillHiRefCountKernel<<<static_cast<unsigned>(CeilDiv(m, kChunk)), kKeygenBlock, 0, b->stream>>>(
      text, ctx.d_offsets, refs, kChunk, passes, counts);
  RefDigitStartsKernel<<<dim3(refs, kFillHiPasses), kRadix, 0, b->stream>>>(counts, ctx.d_offsets, refs, adj);
  RefDigitBeforeKernel<<<kFillHiPasses * kRadix, kRadix, 0, b->stream>>>(counts, refs, adj);
  SFK_HIP_CHECK(hipGetLastError());
  const std::uint64_t text_bytes = PackedWords(m) * 16;
  ctx.stats.algorithmic_bytes_sa += text_bytes;
  SortPairs<std::uint32_t, Index, true>(ctx, TextFillHiSource<Index>{text}, text_bytes, keys, vals, keys_alt,
                                        vals_alt, m, passes, set, adj);
}

// ------------------------------------------------------------ autotuner

struct TuneResult {
  int kind;
  int size_class;
  int shape;
  int bits;
  double us_per_pass;
};

// Times every compiled shape of one sort kind at every digit width it has on
// uniformly random keys, for each size class up to max_elements: two passes
// per sort, the best of two timed sorts after one warm-up, on the tuning
// builder's stream.
template <typename K, typename V, bool kSeg>
void TuneKind(sfk_hip_builder* b, std::uint64_t max_elements, hipEvent_t e0, hipEvent_t e1,
              std::vector<TuneResult>& out) {
  constexpr int kind = SortKind<K, V, kSeg>();
  BuildContext ctx{};
  ctx.b = b;
  ctx.profile = false;
  long long* adj = b->ref_adj.As<long long>();
  if constexpr (kSeg) {
    ctx.d_offsets = b->ref_offsets.As<std::uint64_t>();
    ctx.refs = 1;
  }
  for (int c = 0; c < kSizeClasses; ++c) {
    const std::uint64_t lower = c == 0 ? 0 : (1ull << (16 + 2 * c));
    if (lower > max_elements) break;
    const std::uint64_t m = std::min<std::uint64_t>(1ull << (17 + 2 * c), max_elements);
    if (m < kMinOnesweepTile) continue;
    ctx.n = m;
    if constexpr (kSeg) {
      const std::uint64_t offsets[2] = {0, m};
      SFK_HIP_CHECK(hipMemcpyAsync(b->ref_offsets.As<void>(), offsets, sizeof(offsets), hipMemcpyHostToDevice,
                                   b->stream));
      SFK_HIP_CHECK(hipStreamSynchronize(b->stream));
    }
    const unsigned grid = static_cast<unsigned>(
        std::max<std::uint64_t>(1, std::min<std::uint64_t>(CeilDiv(m, kKeygenBlock * 4), 1024)));
    for (int width = 0; width < (kSeg ? 1 : kDigitWidths); ++width)
    for (int shape = 0; shape < kShapeCount; ++shape) {
      if (!KindShapeFits(kind, shape)) continue;
      RadixPasses passes{};
      passes.bits = DigitBitsOf(width);
      passes.count = 2;
      passes.shift[0] = 0;
      passes.shift[1] = passes.bits;
      float best = 0;
      for (int rep = 0; rep < 3; ++rep) {
        const RadixSet set = TakeRadixSet(b);
        K* keys = b->keys0.As<K>();
        V* vals = b->vals0.As<V>();
        TuneKeysKernel<K, V><<<grid, kKeygenBlock, 0, b->stream>>>(keys, vals, m, 0x5eed0000ull * (rep + 1),
                                                                    passes, set);
        if constexpr (kSeg) TuneAdjKernel<<<passes.count, RadixTile::kRadix, 0, b->stream>>>(set.hist, adj);
        SFK_HIP_CHECK(hipGetLastError());
        SFK_HIP_CHECK(hipEventRecord(e0, b->stream));
        SortPairs<K, V, kSeg>(ctx, ArraySource<K, V>{keys, vals}, m * (sizeof(K) + sizeof(V)), keys, vals,
                              b->keys1.As<K>(), b->vals1.As<V>(), m, passes, set, adj, shape);
        SFK_HIP_CHECK(hipEventRecord(e1, b->stream));
        SFK_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        SFK_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (rep == 1 || (rep > 1 && ms < best)) best = ms;
      }
      out.push_back(TuneResult{kind, c, shape, passes.bits, best * 1000.0 / passes.count});
    }
  }
}

// Reset of the per-round list counters (small-list appends, large-list
// allocation).
__global__ void ResetRoundCountersKernel(DeviceCounters* c) {
  c->q_next = 0;
  c->l_next = 0;
  c->l_next_segs = 0;
}

// Index: suffix positions and rows in the lists; R: the SA and rank (ISA)
// arrays, which may be narrower (the stored width of a 64-bit build whose
// values fit 32 bits).
template <typename Index, typename R>
class Doubling {
 public:
  // early: where the first step and the exact local step write early LCP
  // values (the caller's LCP buffer), or off.
  Doubling(BuildContext& ctx, R* sa, R* rank, PackedText text, EarlyLcp early)
      : ctx_(ctx), b_(ctx.b), n_(ctx.n), sa_(sa), rank_(rank), text_(text), early_(early) {}

  void Run() {
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    lm_ = n_;
    // A batch sorts its first step within each reference's slot range
    // (boundaries at reference starts) when its counts fit the placement
    // tables; otherwise each reference starts as a segment of its own.
    segmented_first_ = ctx_.refs > 0 && ctx_.refs <= kMaxSegmentedRefs && n_ <= 0xFFFFFFFFull;
    l_segs_ = ctx_.refs > 0 && !segmented_first_ ? ctx_.refs : 1;
    l_idx_ = b_->act_idx.As<Index>();
    l_seg_ = b_->act_seg.As<std::uint32_t>();
    l_delta_ = b_->seg_delta0.As<Index>();
    SFK_HIP_CHECK(hipMemsetAsync(l_delta_, 0, l_segs_ * sizeof(Index), b_->stream));
    if (ctx_.refs > 0 && !segmented_first_) {
      // One initial segment per reference: its rows are its own positions.
      InitSegmentsKernel<<<static_cast<unsigned>(CeilDiv(n_, 256)), 256, 0, b_->stream>>>(
          ctx_.d_offsets, ctx_.refs, n_, b_->act_seg.As<std::uint32_t>());
      SFK_HIP_CHECK(hipGetLastError());
    }
    // First 16 symbols over every suffix (one segment per reference).
    // Its survivors' ranks are not written: the next step writes the rank of
    // every survivor (small list and large list alike) before any gather.
    GlobalStep(kSecFillHi, 32, 0, /*iota=*/true, /*local_between=*/false, /*readback=*/true,
               /*survivor_rank=*/false);
    if (ctx_.refs > 0 && b_->host_counters->zero_count != ctx_.refs) {
      Fail(SFK_HIP_INVALID_INPUT, "every reference of a batch must end with its only 0 (sentinel)");
    }
    if (lm_ == 0) return;
    ReserveSmallLists(lm_);
    ResetRoundCountersKernel<<<1, 1, 0, b_->stream>>>(dc);
    Split(OtherDelta(l_delta_));
    EndRound(true);

    // Exact 32-symbol groups: small segments in one local step on the exact
    // 96-bit window key; large segments by the low 16 symbols, then the
    // segments holding a window with a separator, N or the sentinel by the
    // exact key (the others are exact groups already and wait aside).
    ResetRoundCountersKernel<<<1, 1, 0, b_->stream>>>(dc);
    if (qm_ > 0) LocalRefine<true>();
    bool split = false;
    if (lm_ > 0) {
      GlobalStep(kSecFillLo, 32, 0, false, false, /*readback=*/false);
      std::uint32_t* flags = SpecialFlags();
      SFK_HIP_CHECK(hipMemsetAsync(flags, 0, (lm_bound_ / 2 + 1) * sizeof(std::uint32_t), b_->stream));
      ResetSpecialCountKernel<<<1, 1, 0, b_->stream>>>(dc);
      MarkSpecialSegmentsKernel<Index><<<static_cast<unsigned>(CeilDiv(lm_bound_, 256 * kMarkSpans)), 256, 0,
                                         b_->stream>>>(
          b_->act_idx.As<Index>(), b_->act_seg.As<std::uint32_t>(), text_, b_->special_blocks.As<std::uint32_t>(),
          flags, dc);
      SFK_HIP_CHECK(hipGetLastError());
      ctx_.stats.algorithmic_bytes_sa += lm_bound_ * (sizeof(Index) + 4 + 16);
      const bool special = ApplySizes(ReadCounters(b_)).special_active > 0;
      bool stashed = false;
      if (special && lm_ > 0) {
        stashed = Partition(flags);
        GlobalStep(kSecExact0, 32, 0, false, false, true, /*survivor_rank=*/false);
        if (lm_ > 0) GlobalStep(kSecExact1, 32, 0, false, false, true, /*survivor_rank=*/false);
        if (lm_ > 0) GlobalStep(kSecExact2, 32, 0, false, false, true);
      }
      Index* large_delta = OtherDelta(l_delta_);
      if (lm_ > 0) {
        Split(large_delta);
        split = true;
      }
      if (stashed) {
        SplitStash(large_delta);
        split = true;
      }
    }
    packed_lo_ = false;  // every list holds plain positions from here on
    EndRound(split);

    const int rank_bits = std::max(1, BitsFor(n_ - 1));
    for (std::uint64_t h = 32; qm_ + lm_ > 0; h <<= 1) {
      if (h >= n_) Fail(SFK_HIP_BUILD_FAILURE, "prefix doubling did not converge");
      ResetRoundCountersKernel<<<1, 1, 0, b_->stream>>>(dc);
      if (qm_ > 0) {
        QKeygenKernel<Index><<<static_cast<unsigned>(CeilDiv(qm_, 256)), 256, 0, b_->stream>>>(
            q_idx_cur_, qm_, rank_, n_, h, b_->q_key.As<Index>(), dc);
        SFK_HIP_CHECK(hipGetLastError());
        ctx_.stats.algorithmic_bytes_sa += qm_ * (2 * sizeof(Index) + sizeof(R));
      }
      split = false;
      if (lm_ > 0) {
        GlobalStep(kSecRank, rank_bits, h, false, /*local_between=*/qm_ > 0, /*readback=*/false);
        Split(OtherDelta(l_delta_));
        split = true;
      } else {
        LocalRefine<false>();
      }
      EndRound(split);
      ctx_.stats.doubling_rounds += 1;
    }
  }

 private:
  // Small lists hold at most the survivors of the first step, and no more
  // than free device memory allows (segments that do not fit stay in the
  // large list).
  void ReserveSmallLists(std::uint64_t want) {
    constexpr std::uint64_t kArrays = 5;
    const std::uint64_t held = b_->q_idx0.Bytes() / sizeof(Index);
    std::uint64_t cap = std::max<std::uint64_t>(want, 1);
    if (b_->borrowed_workspace) {
      cap = std::min(cap, held);  // the slices are sized for the whole text
    } else if (cap > held) {
      std::size_t free_bytes = 0, total_bytes = 0;
      SFK_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
      const std::uint64_t margin = 2ull << 30;
      const std::uint64_t spare = free_bytes > margin ? free_bytes - margin : 0;
      const std::uint64_t affordable = held + spare / (kArrays * sizeof(Index));
      // Reallocating gigabytes costs more than it saves unless the lists then
      // hold everything: a partial growth keeps what is held.
      if (affordable < cap && held > 0) {
        cap = held;
      } else {
        cap = std::max<std::uint64_t>(std::min(cap, affordable), 1);
      }
    }
    const std::uint64_t bytes = cap * sizeof(Index);
    b_->q_idx0.Reserve(bytes);
    b_->q_idx1.Reserve(bytes);
    b_->q_row0_0.Reserve(bytes);
    b_->q_row0_1.Reserve(bytes);
    b_->q_key.Reserve(bytes);
    q_idx_cur_ = b_->q_idx0.As<Index>();
    q_row0_cur_ = b_->q_row0_0.As<Index>();
    q_idx_next_ = b_->q_idx1.As<Index>();
    q_row0_next_ = b_->q_row0_1.As<Index>();
    q_cap_ = cap;
  }

  Index* OtherDelta(Index* d) const {
    return d == b_->seg_delta0.As<Index>() ? b_->seg_delta1.As<Index>() : b_->seg_delta0.As<Index>();
  }

  // Fails on any error flag raised so far.
  void CheckFlags(const DeviceCounters& c) {
    const unsigned f = c.error_flags;
    if (f & kFlagBadSymbol) Fail(SFK_HIP_INVALID_INPUT, "text holds a symbol outside 0..6");
    if (f & kFlagBadSentinel) Fail(SFK_HIP_INVALID_INPUT, "text must end with its only 0 (sentinel)");
    if (f & kFlagLookback) Fail(SFK_HIP_BUILD_FAILURE, "radix look-back wait passed its bound");
    if (f & kFlagLocalSegment) Fail(SFK_HIP_BUILD_FAILURE, "small-list segment larger than its bound");
    if (f & kFlagRange) Fail(SFK_HIP_BUILD_FAILURE, "rank gather past the end of the text");
    if (f != 0) {
      Fail(SFK_HIP_BUILD_FAILURE, "device consistency check failed (flags " + std::to_string(f) + ")");
    }
  }

  // Takes the sizes of the list the last RefineApply wrote from counters
  // read back after it, and books its traffic.
  const DeviceCounters& ApplySizes(const DeviceCounters& c) {
    CheckFlags(c);
    if (pending_apply_) {
      constexpr std::uint64_t w = sizeof(Index), wr = sizeof(R);
      const std::uint64_t m = pending_m_, survivors = c.list.survivors;
      ctx_.stats.algorithmic_bytes_sa +=
          m * (2 * pending_key_bytes_ + w) + (m - survivors) * 2 * wr + survivors * (2 * w + wr + 4);
      pending_apply_ = false;
    }
    lm_ = c.list.survivors;
    l_segs_ = c.list.segments;
    lm_bound_ = lm_;
    return c;
  }

  // Reads the round's list sizes back, checks the error flags and makes the
  // next small list current.
  void EndRound(bool split) {
    const DeviceCounters& c = ReadCounters(b_);
    CheckFlags(c);
    if (pending_apply_) ApplySizes(c);
    qm_ = c.q_next;
    if (split) {
      lm_ = c.l_next;
      l_segs_ = c.l_next_segs;
    } else {
      lm_ = 0;
      l_segs_ = 0;
    }
    lm_bound_ = lm_;
    if (qm_ > q_cap_) Fail(SFK_HIP_BUILD_FAILURE, "small list overflow");
    std::swap(q_idx_cur_, q_idx_next_);
    std::swap(q_row0_cur_, q_row0_next_);
  }

  // Routes the segments of the list RefineApply just wrote: small ones to
  // the next small list, large ones (renumbered) to the next large list in
  // vals0 (positions) and vals1 (segment ids), with deltas in large_delta.
  void Split(Index* large_delta) {
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    // Sizes come from the counters RefineApply wrote; grids cover bounds
    // (a list of m suffixes has at most m / 2 segments).
    const std::uint64_t m_bound = std::max<std::uint64_t>(lm_bound_, 1);
    const std::uint64_t seg_bound = pending_apply_ ? m_bound / 2 + 1 : std::max<std::uint64_t>(l_segs_, 1);
    SplitList(b_->act_idx.As<Index>(), b_->act_seg.As<std::uint32_t>(), b_->seg_start.As<Index>(), l_delta_,
              &dc->list, m_bound, seg_bound, large_delta);
  }

  // Routes the stash a partition set aside, after the special list's split,
  // into the same next lists.
  void SplitStash(Index* large_delta) {
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    SplitList(q_idx_cur_, reinterpret_cast<const std::uint32_t*>(b_->q_key.As<Index>()), q_row0_cur_,
              q_row0_cur_ + stash_seg_room_, &dc->stash, std::max<std::uint64_t>(stash_m_, 1),
              std::max<std::uint64_t>(stash_segs_, 1), large_delta);
  }

  void SplitList(const Index* idx, const std::uint32_t* seg, const Index* start, const Index* delta,
                 const ListSize* size, std::uint64_t m_bound, std::uint64_t seg_bound, Index* large_delta) {
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    Index* dest = b_->keys0.As<Index>();
    std::uint32_t* new_id = b_->keys1.As<std::uint32_t>();
    const SortGeometry g = Geometry(seg_bound, kSplitTile, 1024);
    SplitPartial* partial = b_->partial.As<SplitPartial>();
    SplitCountKernel<Index><<<g.grid, 256, 0, b_->stream>>>(start, kLocalMaxSegment, g.chunk, partial, size);
    SplitSpineKernel<<<1, 1024, 0, b_->stream>>>(partial, g.grid, q_cap_, dc);
    SplitAssignKernel<Index><<<g.grid, 256, 0, b_->stream>>>(start, delta, kLocalMaxSegment, g.chunk, partial,
                                                             dest, new_id, large_delta, size);
    SFK_HIP_CHECK(hipGetLastError());
    Index* l_idx_next = b_->vals0.As<Index>();
    std::uint32_t* l_seg_next = b_->vals1.As<std::uint32_t>();
    SplitMoveKernel<Index><<<static_cast<unsigned>(CeilDiv(m_bound, 256)), 256, 0, b_->stream>>>(
        idx, seg, start, delta, dest, new_id, q_idx_next_, q_row0_next_, l_idx_next, l_seg_next, size);
    SFK_HIP_CHECK(hipGetLastError());
    constexpr std::uint64_t w = sizeof(Index);
    ctx_.stats.algorithmic_bytes_sa += (m_bound / 2) * (4 * w + 4) + m_bound * (2 * w + 8 + 2 * w);
    l_idx_ = l_idx_next;
    l_seg_ = l_seg_next;
    l_delta_ = large_delta;
  }

  // Per-segment flags of the partition: above the split's new ids in keys1.
  std::uint32_t* SpecialFlags() const { return b_->keys1.As<std::uint32_t>() + n_; }

  // Partition of the list RefineApply just wrote (sizes read back) by the
  // segment flags: flagged segments become the large list (vals0 / vals1,
  // deltas in the other delta array), the others wait in the current small
  // list's buffers, which the small list's step has consumed (positions in
  // q_idx_cur, segment ids in q_key, first slots and deltas in q_row0_cur).
  // Returns false, leaving the list whole, when they do not fit there.
  bool Partition(const std::uint32_t* flags) {
    if (lm_ > q_cap_ || l_segs_ == 0) return false;
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    const SortGeometry g = Geometry(l_segs_, kSplitTile, 1024);
    PartitionPartial* partial = b_->partial.As<PartitionPartial>();
    Index* dest = b_->keys0.As<Index>();
    std::uint32_t* new_id = b_->keys1.As<std::uint32_t>();
    Index* x_delta = OtherDelta(l_delta_);
    stash_seg_room_ = l_segs_;
    PartitionCountKernel<Index><<<g.grid, 256, 0, b_->stream>>>(b_->seg_start.As<Index>(), flags, g.chunk,
                                                                partial, dc);
    PartitionSpineKernel<<<1, 1024, 0, b_->stream>>>(partial, g.grid, dc);
    PartitionAssignKernel<Index><<<g.grid, 256, 0, b_->stream>>>(
        b_->seg_start.As<Index>(), l_delta_, flags, g.chunk, partial, dest, new_id, x_delta, q_row0_cur_,
        q_row0_cur_ + stash_seg_room_, dc);
    SFK_HIP_CHECK(hipGetLastError());
    Index* x_idx = b_->vals0.As<Index>();
    std::uint32_t* x_seg = b_->vals1.As<std::uint32_t>();
    PartitionMoveKernel<Index><<<static_cast<unsigned>(CeilDiv(lm_, 256)), 256, 0, b_->stream>>>(
        b_->act_idx.As<Index>(), b_->act_seg.As<std::uint32_t>(), b_->seg_start.As<Index>(), flags, dest, new_id,
        x_idx, x_seg, q_idx_cur_, reinterpret_cast<std::uint32_t*>(b_->q_key.As<Index>()), dc);
    SFK_HIP_CHECK(hipGetLastError());
    constexpr std::uint64_t w = sizeof(Index);
    ctx_.stats.algorithmic_bytes_sa += l_segs_ * (4 * w + 8) + lm_ * (2 * w + 8 + 2 * w);
    const DeviceCounters& c = ReadCounters(b_);
    CheckFlags(c);
    lm_ = c.special.survivors;
    l_segs_ = c.special.segments;
    lm_bound_ = lm_;
    stash_m_ = c.stash.survivors;
    stash_segs_ = c.stash.segments;
    l_idx_ = x_idx;
    l_seg_ = x_seg;
    l_delta_ = x_delta;
    return true;
  }

  template <bool kExact>
  void LocalRefine() {
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    const unsigned grid = static_cast<unsigned>(CeilDiv(qm_, kLocalOwn));
    LocalRefineKernel<Index, kExact><<<grid, kLocalThreads, 0, b_->stream>>>(
        text_, q_idx_cur_, q_row0_cur_, b_->q_key.As<Index>(), qm_, n_, sa_, rank_, q_idx_next_,
        q_row0_next_, dc, kExact ? early_ : EarlyLcp{nullptr, nullptr}, kExact && packed_lo_,
        b_->special_blocks.As<std::uint32_t>());
    SFK_HIP_CHECK(hipGetLastError());
    constexpr std::uint64_t w = sizeof(Index);
    ctx_.stats.refine_steps += 1;
    ctx_.stats.active_total += qm_;
    ctx_.stats.algorithmic_bytes_sa += qm_ * (2 * w + (kExact ? 32 : w) + 2 * sizeof(R) + 2 * w);
    StepLog(kExact ? "local-exact" : "local", qm_, 0, 0);
  }

  // Diagnostics: with SFK_HIP_STEP_LOG set, one line per refine step on
  // stderr (kind, suffixes, segments, key bits).
  static bool StepLogOn() {
    static const bool on = std::getenv("SFK_HIP_STEP_LOG") != nullptr;
    return on;
  }
  void StepLog(const char* kind, std::uint64_t m, std::uint64_t segs, int key_bits) const {
    if (!StepLogOn()) return;
    std::fprintf(stderr, "sfk step %-12s m=%llu segs=%llu key_bits=%d small=%llu\n", kind,
                 static_cast<unsigned long long>(m), static_cast<unsigned long long>(segs), key_bits,
                 static_cast<unsigned long long>(qm_));
  }

  // readback: read the new list's sizes back right after RefineApply (the
  // host needs them before the next step); otherwise only lm_bound_ (the
  // input size) is known until the next read.
  void GlobalStep(int sec_kind, int sec_bits, std::uint64_t h, bool iota, bool local_between,
                  bool readback, bool survivor_rank = true) {
    const int seg_bits = l_segs_ > 1 ? BitsFor(l_segs_ - 1) : 0;
    const int total_bits = seg_bits + sec_bits;
    if (total_bits > 64) {
      Fail(SFK_HIP_UNSUPPORTED_BACKEND,
           "segment and rank bits exceed one 64-bit sort key");
    }
    ctx_.stats.refine_steps += 1;
    ctx_.stats.active_total += lm_;
    static const char* const kKindNames[] = {"fill-hi", "fill-lo", "exact0", "exact1", "exact2", "rank"};
    StepLog(kKindNames[sec_kind], lm_, l_segs_, total_bits);
    if (total_bits <= 32) {
      GlobalStepTyped<std::uint32_t>(sec_kind, sec_bits, seg_bits, h, iota, local_between, survivor_rank);
    } else {
      GlobalStepTyped<std::uint64_t>(sec_kind, sec_bits, seg_bits, h, iota, local_between, survivor_rank);
    }
    if (readback) ApplySizes(ReadCounters(b_));
  }

  // One refine step of the large list: keygen, (the small list's step, whose
  // rank gathers are done), radix sort, boundaries/singletons/compaction.
  template <typename K>
  void GlobalStepTyped(int sec_kind, int sec_bits, int seg_bits, std::uint64_t h, bool iota,
                       bool local_between, bool survivor_rank) {
    const std::uint64_t m = lm_;
    const int total_bits = seg_bits + sec_bits;
    K* keys = b_->keys0.As<K>();
    K* keys_alt = b_->keys1.As<K>();
    Index* vals = b_->vals0.As<Index>();
    Index* vals_alt = b_->vals1.As<Index>();
    DeviceCounters* dc = b_->counters.As<DeviceCounters>();
    constexpr std::uint64_t w = sizeof(Index);

    const bool rank_step = sec_kind == kSecRank;
    const RadixSet set = TakeRadixSet(b_);
    const unsigned keygen_grid = static_cast<unsigned>(
        std::max<std::uint64_t>(1, std::min<std::uint64_t>(CeilDiv(m, kKeygenBlock * 4), 1024)));
    // The first step of a single text generates its first pass's input from
    // the text (keys of 16 symbols, values = positions); a counting kernel reads
    // the text for the digit counts.
    constexpr bool kNarrowKey = sizeof(K) == 4;
    const bool segmented = segmented_first_ && sec_kind == kSecFillHi;
    // Sorts of keygen arrays take the digit width the tuner's times favour
    // for their key width; the first step's (from the text, or in place per
    // reference) takes 8-bit digits.
    const bool from_text = kNarrowKey && sec_kind == kSecFillHi && (segmented || (iota && seg_bits == 0));
    const RadixPasses passes =
        AllDigits(total_bits, from_text ? RadixTile::kDigitBits
                                        : DigitBitsFor(b_->device, SortKind<K, Index, false>(), m, total_bits));
    if (kNarrowKey && segmented) {
      if (m != n_ || !iota || seg_bits != 0) Fail(SFK_HIP_BUILD_FAILURE, "segmented first step out of place");
      if constexpr (kNarrowKey) {
        SortFillHiSegmented<Index>(ctx_, text_, keys, vals, keys_alt, vals_alt, m, set);
      }
    } else if (kNarrowKey && iota && seg_bits == 0 && sec_kind == kSecFillHi) {
      const unsigned count_grid = static_cast<unsigned>(
          std::max<std::uint64_t>(1, std::min<std::uint64_t>(CeilDiv(m, kKeygenBlock * 16), 1024)));
      TextFillHiCountKernel<<<count_grid, kKeygenBlock, 0, b_->stream>>>(text_, passes, set);
      SFK_HIP_CHECK(hipGetLastError());
      const std::uint64_t text_bytes = PackedWords(n_) * 16;
      ctx_.stats.algorithmic_bytes_sa += text_bytes;
      if constexpr (kNarrowKey) {
        // A 64-bit list carries each suffix's low FillKey half above its
        // position until the exact steps have used it.
        packed_lo_ = sizeof(Index) == 8;
        SortPairs<K, Index>(ctx_, TextFillHiSource<Index>{text_, packed_lo_}, text_bytes, keys, vals, keys_alt,
                            vals_alt, m, passes, set);
      }
    } else {
      // The large list may sit in vals0/vals1 (after a split): Keygen reads
      // slot k of it before writing slot k of vals0, in the same thread.
      const bool lo_from_value = packed_lo_ && sec_kind == kSecFillLo;
      if (rank_step) {
        KeygenKernel<Index, K, true><<<keygen_grid, kKeygenBlock, 0, b_->stream>>>(
            l_idx_, l_seg_, m, sec_kind, sec_bits, seg_bits, text_, rank_, h, iota, keys, vals, passes, set, dc,
            false);
      } else {
        KeygenKernel<Index, K, false><<<keygen_grid, kKeygenBlock, 0, b_->stream>>>(
            l_idx_, l_seg_, m, sec_kind, sec_bits, seg_bits, text_, rank_, h, iota, keys, vals, passes, set, dc,
            lo_from_value);
      }
      SFK_HIP_CHECK(hipGetLastError());
      ctx_.stats.algorithmic_bytes_sa +=
          m * ((iota ? 0 : w) + (seg_bits ? w : 0) + (rank_step ? sizeof(R) : lo_from_value ? 0 : 24) + sizeof(K) +
               w);
      if (local_between) LocalRefine<false>();
      SortPairs<K, Index>(ctx_, ArraySource<K, Index>{keys, vals}, m * (sizeof(K) + w), keys, vals, keys_alt,
                          vals_alt, m, passes, set);
    }

    // Boundaries, singletons, compaction: reduce / spine / apply for a long
    // one-segment list, one look-back kernel otherwise.  In a doubling round
    // every suffix's rank is its segment's first row: the step before wrote
    // the ranks of all survivors (the exact steps before the first round end
    // with one that does), so unchanged ranks are not stored again.
    const bool ranks_current = rank_step;
    Index* delta_next = OtherDelta(l_delta_);
    const std::uint64_t tiles = CeilDiv(m, kRefineTile);
    if (seg_bits > 0 || segmented || tiles <= static_cast<std::uint64_t>(kRefineResidentTiles)) {
      if ((tiles + 1) * 3 * 8 > b_->radix_status.Bytes()) {
        Fail(SFK_HIP_BUILD_FAILURE, "look-back status array smaller than the list");
      }
      if (b_->tickets.Bytes() < 2 * sizeof(unsigned int)) Fail(SFK_HIP_BUILD_FAILURE, "look-back tickets not reserved");
      RefineApplyKernel<Index, K><<<static_cast<unsigned>(tiles), kRefineThreads, 0, b_->stream>>>(
          keys, vals, m, sec_bits, l_delta_, n_, sa_, rank_, b_->act_idx.As<Index>(),
          b_->act_seg.As<std::uint32_t>(), delta_next, b_->seg_start.As<Index>(), survivor_rank, ranks_current,
          b_->radix_status.As<unsigned long long>(), NextEpoch(b_), b_->tickets.As<unsigned int>(), dc,
          segmented ? ctx_.d_offsets : nullptr, segmented ? ctx_.refs : 0u,
          sec_kind == kSecFillHi ? early_ : EarlyLcp{nullptr, nullptr});
    } else {
      const SortGeometry g = Geometry(m, kRefineChunkTile, kRefineMaxChunks);
      RefinePartial* partial = b_->partial.As<RefinePartial>();
      RefineReduceKernel<K><<<g.grid, kRefineChunkThreads, 0, b_->stream>>>(keys, m, g.chunk, partial);
      RefineSpineKernel<<<1, 1024, 0, b_->stream>>>(partial, g.grid, dc);
      RefineApplyChunkedKernel<Index, K><<<g.grid, kRefineChunkThreads, 0, b_->stream>>>(
          keys, vals, m, g.chunk, sec_bits, l_delta_, partial, n_, sa_, rank_, b_->act_idx.As<Index>(),
          b_->act_seg.As<std::uint32_t>(), delta_next, b_->seg_start.As<Index>(), survivor_rank, ranks_current,
          nullptr, 0u,
          sec_kind == kSecFillHi ? early_ : EarlyLcp{nullptr, nullptr});
    }
    SFK_HIP_CHECK(hipGetLastError());
    if (sec_kind == kSecFillHi) {
      ctx_.stats.algorithmic_bytes_sa += m * (early_.lcp32 ? 4 : early_.lcp64 ? 8 : 0);  // early LCP or marker per row
    }
    pending_apply_ = true;
    pending_m_ = m;
    pending_key_bytes_ = sizeof(K);
    lm_bound_ = m;
    l_idx_ = b_->act_idx.As<Index>();
    l_seg_ = b_->act_seg.As<std::uint32_t>();
    l_delta_ = delta_next;
  }

  BuildContext& ctx_;
  sfk_hip_builder* b_;
  std::uint64_t n_;
  R* sa_;
  R* rank_;
  PackedText text_;
  EarlyLcp early_;
  // The large and small lists hold 64-bit values with the low FillKey half
  // above the position (from the first step until the exact steps).
  bool packed_lo_ = false;
  // Large list: suffixes, segment ids, per-segment deltas.  After a
  // RefineApply whose sizes are not read back yet, lm_bound_ bounds lm_.
  std::uint64_t lm_ = 0;
  std::uint64_t lm_bound_ = 0;
  bool pending_apply_ = false;
  bool segmented_first_ = false;
  std::uint64_t pending_m_ = 0;
  std::uint64_t pending_key_bytes_ = 0;
  std::uint64_t l_segs_ = 0;
  // The stash of a partition: sizes, and where its deltas start in
  // q_row0_cur (after room for the input list's segment count).
  std::uint64_t stash_m_ = 0;
  std::uint64_t stash_segs_ = 0;
  std::uint64_t stash_seg_room_ = 0;
  const Index* l_idx_ = nullptr;
  const std::uint32_t* l_seg_ = nullptr;
  Index* l_delta_ = nullptr;
  // Small lists (current and next): suffixes and their segments' first rows.
  std::uint64_t qm_ = 0;
  std::uint64_t q_cap_ = 0;
  Index* q_idx_cur_ = nullptr;
  Index* q_row0_cur_ = nullptr;
  Index* q_idx_next_ = nullptr;
  Index* q_row0_next_ = nullptr;
};

// Grid of a kernel that takes `per_block` items per 256-thread block: one
// block per group of items, capped where the dispatch would pass 2^32
// work-items (the kernel's grid-stride loop takes the rest).
unsigned FlatGrid(std::uint64_t m, std::uint64_t per_block) {
  constexpr std::uint64_t kMaxBlocks = 0xFFFFFFFFull / 256;
  return static_cast<unsigned>(std::max<std::uint64_t>(1, std::min<std::uint64_t>(CeilDiv(m, per_block), kMaxBlocks)));
}

// LCP pass over every text position (lcp_kernels.hpp), for builds without
// early LCP (sampled ones): bucket-start suffixes, PLCP of irreducible
// positions (with the permutation/inverse check; long LCEs by whole waves),
// PLCP max-scan, LCP gather.  keys0 holds PLCP (32-bit), keys1 the
// bucket-start position bitmap and the long-LCE queue.
template <typename Index, typename Out>
void LaunchLcp(BuildContext& ctx, const std::uint8_t* d_text, PackedText text, const Index* sa, const Index* isa,
               Out* lcp) {
  sfk_hip_builder* b = ctx.b;
  const std::uint64_t n = ctx.n;
  DeviceCounters* dc = b->counters.As<DeviceCounters>();
  std::uint32_t* plcp = b->keys0.As<std::uint32_t>();
  if (b->keys0.Bytes() < n * 4) Fail(SFK_HIP_BUILD_FAILURE, "LCP pass workspace smaller than the PLCP");
  std::uint32_t* starts = b->keys1.As<std::uint32_t>();
  const std::uint64_t bit_words = n / 32 + 1;
  const std::uint64_t bit_bytes = (bit_words * 4 + 63) / 64 * 64;
  if (b->keys1.Bytes() < bit_bytes) Fail(SFK_HIP_BUILD_FAILURE, "LCP pass workspace too small for the bitmap");
  LongLce* queue = reinterpret_cast<LongLce*>(b->keys1.As<char>() + bit_bytes);
  const std::uint64_t capacity = (b->keys1.Bytes() - bit_bytes) / sizeof(LongLce);
  constexpr std::uint64_t w = sizeof(Index);
  constexpr std::uint32_t kWordLimit = 8;
  SFK_HIP_CHECK(hipMemsetAsync(starts, 0, bit_words * sizeof(std::uint32_t), b->stream));
  const std::uint64_t bucket_threads = 7ull * (ctx.refs > 0 ? ctx.refs : 1);
  BucketStartsKernel<Index><<<static_cast<unsigned>(CeilDiv(bucket_threads, 64)), 64, 0, b->stream>>>(
      d_text, sa, n, ctx.d_offsets, ctx.refs, starts);
  const std::uint64_t plcp_tiles = CeilDiv(n, kPlcpTile);
  PlcpKernel<Index><<<FlatGrid(plcp_tiles, kPlcpThreads / kWave), kPlcpThreads, 0, b->stream>>>(
      text, sa, isa, starts, n, kWordLimit, plcp, queue, capacity, dc);
  SFK_HIP_CHECK(hipGetLastError());
  const DeviceCounters& c = ReadCounters(b);
  if (c.error_flags & (kFlagNotPermutation | kFlagNotInverse)) {
    Fail(SFK_HIP_BUILD_FAILURE, "suffix array is not a permutation with an inverse");
  }
  const std::uint64_t items = c.long_items;
  if (items > capacity) Fail(SFK_HIP_BUILD_FAILURE, "long-LCE queue overflow");
  if (items > 0) {
    PlcpLongKernel<<<static_cast<unsigned>(CeilDiv(items * kWave, 256)), 256, 0, b->stream>>>(text, n, queue, items,
                                                                                              plcp, dc);
    SFK_HIP_CHECK(hipGetLastError());
  }
  const SortGeometry g = Geometry(n, 256 * kScanItems, 1024);
  std::uint64_t* partial = b->partial.As<std::uint64_t>();
  PlcpScanReduceKernel<<<g.grid, 256, 0, b->stream>>>(plcp, n, g.chunk, partial, dc);
  PlcpScanSpineKernel<<<1, 1024, 0, b->stream>>>(partial, g.grid);
  PlcpScanApplyKernel<<<g.grid, 256, 0, b->stream>>>(plcp, n, g.chunk, partial);
  LcpGatherKernel<Index, Out><<<FlatGrid(n, 256 * kGatherRows), 256, 0, b->stream>>>(plcp, sa, n, lcp, false, dc);
  SFK_HIP_CHECK(hipGetLastError());
  ctx.stats.algorithmic_bytes_lcp +=
      n / 8                            // bucket-start bitmap
      + n * (w + w + 4) + n / 2        // PLCP: ISA read, SA[ISA[i] - 1] (with SA[ISA[i]]), PLCP write, text at i
      + n * (3 * 4)                    // max-scan: two reads, one write
      + n * (w + 4 + sizeof(Out));     // LCP = PLCP[SA]: SA, PLCP, LCP
}

// LCP of a build with early LCP (lcp_kernels.hpp, marked-row LCP): the
// rows the doubling left marked get Φ/PLCP through a list sorted by
// position, and their values go back through a scatter sorted by row; the
// pass that builds the list checks that ISA inverts SA.  Workspace (every
// buffer holds n 32-bit words, keys0/keys1 n 64-bit ones): vals0/vals1 the
// list's positions and their radix ping-pong, keys0/keys1 its (Φ, row)
// values and theirs, act_idx the bucket-start bitmap and the long-LCE
// queue, act_seg the list's PLCP; the (row, value) pairs, sorted by row for
// the scatter, reuse the buffers the first sort left free.
static_assert(kSpineChunks * sizeof(std::uint64_t) <= kRefineMaxChunks * sizeof(RefinePartial),
              "the builder's partial buffer holds every chunk total of the LCP scans");

template <typename R, typename Out>
void LaunchLcpMarked(BuildContext& ctx, const std::uint8_t* d_text, PackedText text, const R* sa, const R* isa,
                     Out* lcp) {
  sfk_hip_builder* b = ctx.b;
  const std::uint64_t n = ctx.n;
  DeviceCounters* dc = b->counters.As<DeviceCounters>();
  for (const DeviceBuffer* buf : {&b->vals0, &b->vals1, &b->act_idx, &b->act_seg}) {
    if (buf->Bytes() < n * 4) Fail(SFK_HIP_BUILD_FAILURE, "LCP pass workspace too small for the marked-row list");
  }
  for (const DeviceBuffer* buf : {&b->keys0, &b->keys1}) {
    if (buf->Bytes() < n * 8) Fail(SFK_HIP_BUILD_FAILURE, "LCP pass workspace too small for the marked-row list");
  }
  std::uint32_t* pos = b->vals0.As<std::uint32_t>();
  std::uint32_t* pos_alt = b->vals1.As<std::uint32_t>();
  std::uint32_t* phi = b->keys0.As<std::uint32_t>();
  std::uint32_t* phi_alt = b->keys1.As<std::uint32_t>();
  std::uint32_t* list_plcp = b->act_seg.As<std::uint32_t>();
  std::uint32_t* starts = b->act_idx.As<std::uint32_t>();
  const std::uint64_t bit_words = n / 32 + 1;
  const std::uint64_t bit_bytes = (bit_words * 4 + 63) / 64 * 64;
  if (b->act_idx.Bytes() < bit_bytes) Fail(SFK_HIP_BUILD_FAILURE, "LCP pass workspace too small for the bitmap");
  LongLce* queue = reinterpret_cast<LongLce*>(b->act_idx.As<char>() + bit_bytes);
  const std::uint64_t capacity = (b->act_idx.Bytes() - bit_bytes) / sizeof(LongLce);
  constexpr std::uint32_t kWordLimit = 8;
  auto aligned = [](const void* p) { return reinterpret_cast<std::uintptr_t>(p) % 16 == 0; };

  SFK_HIP_CHECK(hipMemsetAsync(starts, 0, bit_words * sizeof(std::uint32_t), b->stream));
  const std::uint64_t bucket_threads = 7ull * (ctx.refs > 0 ? ctx.refs : 1);
  BucketStartsKernel<R><<<static_cast<unsigned>(CeilDiv(bucket_threads, 64)), 64, 0, b->stream>>>(
      d_text, sa, n, ctx.d_offsets, ctx.refs, starts);
  const SortGeometry g = Geometry(n, kMarkTile, kSpineChunks);
  std::uint64_t* partial = b->partial.As<std::uint64_t>();
  LcpMarkCountKernel<Out><<<g.grid, kMarkThreads, 0, b->stream>>>(lcp, n, g.chunk, aligned(lcp), partial);
  MarkSpineKernel<<<1, 1024, 0, b->stream>>>(partial, g.grid, &dc->lcp_marked);
  const int bits = std::max(1, BitsFor(n - 1));
  // The marked count is known only once the write kernel has run; the digit
  // width of the position sort takes its size class from n (m <= n).
  const RadixPasses passes =
      AllDigits(bits, DigitBitsFor(b->device, SortKind<std::uint32_t, std::uint32_t, false>(), n, bits));
  const RadixSet set = TakeRadixSet(b);
  LcpMarkWriteKernel<R, Out><<<g.grid, kMarkThreads, 0, b->stream>>>(
      lcp, sa, isa, n, g.chunk, aligned(lcp), aligned(sa), aligned(isa), partial, pos, phi, passes, set, dc);
  SFK_HIP_CHECK(hipGetLastError());
  constexpr std::uint64_t r = sizeof(R);
  std::uint64_t bytes = n / 8 + n * sizeof(Out) + n * (sizeof(Out) + 2 * r);  // bitmap; two LCP reads, SA and ISA
  const DeviceCounters& c = ReadCounters(b);
  if ((c.error_flags & (kFlagNotPermutation | kFlagNotInverse)) || c.inverse_hash[0] != 0 || c.inverse_hash[1] != 0) {
    Fail(SFK_HIP_BUILD_FAILURE, "suffix array is not a permutation with an inverse");
  }
  const std::uint64_t m = c.lcp_marked;
  if (m > 0) {
    bytes += m * 8;  // the list written
    const std::uint64_t sort_before = ctx.stats.algorithmic_bytes_sa;
    SortPairs<std::uint32_t, std::uint32_t>(ctx, ArraySource<std::uint32_t, std::uint32_t>{pos, phi}, m * 8, pos,
                                            phi, pos_alt, phi_alt, m, passes, set);
    const std::uint64_t tiles = CeilDiv(m, kPlcpTile);
    SparsePlcpKernel<<<FlatGrid(tiles, kPlcpThreads / kWave), kPlcpThreads, 0, b->stream>>>(
        text, pos, phi, m, starts, kWordLimit, list_plcp, queue, capacity, dc);
    SFK_HIP_CHECK(hipGetLastError());
    const DeviceCounters& c2 = ReadCounters(b);
    const std::uint64_t items = c2.long_items;
    if (items > capacity) Fail(SFK_HIP_BUILD_FAILURE, "long-LCE queue overflow");
    if (items > 0) {
      PlcpLongKernel<<<static_cast<unsigned>(CeilDiv(items * kWave, 256)), 256, 0, b->stream>>>(text, n, queue, items,
                                                                                                list_plcp, dc);
      SFK_HIP_CHECK(hipGetLastError());
    }
    // The scatter's pairs: rows into the position buffer the sort left free
    // (pos_alt is now the other one), values in place.
    std::uint32_t* rows = pos == b->vals0.As<std::uint32_t>() ? b->vals1.As<std::uint32_t>()
                                                              : b->vals0.As<std::uint32_t>();
    // Rows sorted down to windows of at most 2^kScatterWindowBits rows: the
    // pairs of one window are contiguous, so the stores of one instruction
    // fall in a few lines, for one radix pass fewer than a full sort.  The
    // passes end at the row's top bit, so the spare bits of their digits
    // narrow the windows further.
    const int span = std::max(0, bits - kScatterWindowBits);
    RadixPasses scatter_passes{};
    scatter_passes.bits = DigitBitsFor(b->device, SortKind<std::uint32_t, std::uint32_t, false>(), m, span);
    const int scatter_count = (span + scatter_passes.bits - 1) / scatter_passes.bits;
    for (int p = 0; p < scatter_count; ++p) {
      scatter_passes.shift[scatter_passes.count++] = std::max(0, bits - scatter_count * scatter_passes.bits) +
                                                      p * scatter_passes.bits;
    }
    const RadixSet scatter_set = TakeRadixSet(b);
    const SortGeometry gs = Geometry(m, 256 * kScanItems, kSpineChunks);
    SparseScanReduceKernel<<<gs.grid, 256, 0, b->stream>>>(list_plcp, pos, m, gs.chunk, partial, dc);
    PlcpScanSpineKernel<<<1, 1024, 0, b->stream>>>(partial, gs.grid);
    SparseScanApplyKernel<R><<<gs.grid, 256, 0, b->stream>>>(list_plcp, pos, isa, m, gs.chunk, partial, rows,
                                                             scatter_passes, scatter_set);
    SFK_HIP_CHECK(hipGetLastError());
    std::uint32_t* values = list_plcp;
    std::uint32_t* rows_alt = pos;  // free once the scan has read it
    std::uint32_t* values_alt = phi == b->keys0.As<std::uint32_t>() ? b->keys1.As<std::uint32_t>()
                                                                 : b->keys0.As<std::uint32_t>();
    SortPairs<std::uint32_t, std::uint32_t>(ctx, ArraySource<std::uint32_t, std::uint32_t>{rows, values}, m * 8,
                                            rows, values, rows_alt, values_alt, m, scatter_passes, scatter_set);
    bytes += ctx.stats.algorithmic_bytes_sa - sort_before;  // both sorts belong to the LCP phase
    ctx.stats.algorithmic_bytes_sa = sort_before;
    LcpScatterKernel<Out><<<FlatGrid(m, 256 * 4), 256, 0, b->stream>>>(rows, values, m, n, lcp, dc);
    SFK_HIP_CHECK(hipGetLastError());
    bytes += m * (8 + 4 + 1)          // PLCP pass: the entry read, PLCP written, text window at the position
             + m * (8 + 8 + sizeof(R) + 8)  // scan: PLCP and position read twice, the row ISA[position] read,
                                            // (row, value) written
             + m * (8 + sizeof(Out));  // scatter: the pair read, LCP written
  }
  ctx.stats.algorithmic_bytes_lcp += bytes;
}

template <typename In, typename Out>
void Convert(sfk_hip_builder* b, const In* in, std::uint64_t n, Out* out) {
  NarrowKernel<In, Out><<<static_cast<unsigned>(CeilDiv(n, 256)), 256, 0, b->stream>>>(
      in, n, out, b->counters.As<DeviceCounters>());
  SFK_HIP_CHECK(hipGetLastError());
}

template <typename In>
void ConvertTo(sfk_hip_builder* b, const In* in, std::uint64_t n, void* out, std::uint32_t bits) {
  if (bits == 32) {
    Convert<In, std::uint32_t>(b, in, n, static_cast<std::uint32_t*>(out));
  } else {
    Convert<In, std::uint64_t>(b, in, n, static_cast<std::uint64_t*>(out));
  }
}

void RecordPhase(sfk_hip_builder* b, int k) {
  if (b->phase[k] != nullptr) SFK_HIP_CHECK(hipEventRecord(b->phase[k], b->stream));
}

// Rows a sampled index keeps: the text positions that are multiples of K.
std::uint64_t SampledRows(std::uint64_t n, std::uint32_t k) { return (n + k - 1) / k; }

// Sampled SA, ISA and LCP (K > 1) from the complete SA and LCP: written in
// the construction width to free workspace (vals1, act_idx, keys0), then
// narrowed into the caller's buffers with the range check.
template <typename Index>
void SampleOutputs(BuildContext& ctx, const Index* sa, const Index* lcp, void* d_sa, std::uint32_t sa_bits,
                   void* d_isa, std::uint32_t isa_bits, void* d_lcp, std::uint32_t lcp_bits) {
  sfk_hip_builder* b = ctx.b;
  const std::uint64_t n = ctx.n;
  const std::uint64_t k = ctx.sampling;
  const std::uint64_t rows = SampledRows(n, ctx.sampling);
  Index* s_sa = b->vals1.As<Index>();
  Index* s_isa = b->act_idx.As<Index>();
  Index* s_lcp = lcp != nullptr ? b->keys0.As<Index>() : nullptr;
  static_assert(sizeof(SampleSummary) * 1024 <= kRefineMaxChunks * sizeof(RefinePartial) ||
                    sizeof(SampleSummary) * 1024 <= 1024 * sizeof(SplitPartial),
                "the builder's partial buffer holds the sampling summaries");
  const SortGeometry g = Geometry(n, kSampleTile, 1024);
  SampleSummary* partial = b->partial.As<SampleSummary>();
  SampleReduceKernel<Index><<<g.grid, kSampleThreads, 0, b->stream>>>(sa, lcp, n, k, g.chunk, partial);
  SampleSpineKernel<<<1, 1024, 0, b->stream>>>(partial, g.grid);
  SampleApplyKernel<Index><<<g.grid, kSampleThreads, 0, b->stream>>>(sa, lcp, n, k, g.chunk, partial, s_sa, s_isa,
                                                                     s_lcp);
  SFK_HIP_CHECK(hipGetLastError());
  constexpr std::uint64_t w = sizeof(Index);
  ctx.stats.algorithmic_bytes_isa += 2 * n * (w + (lcp != nullptr ? w : 0)) + rows * (2 * w + (lcp ? w : 0));
  ConvertTo<Index>(b, s_sa, rows, d_sa, sa_bits);
  if (d_isa != nullptr) ConvertTo<Index>(b, s_isa, rows, d_isa, isa_bits);
  if (d_lcp != nullptr && s_lcp != nullptr) ConvertTo<Index>(b, s_lcp, rows, d_lcp, lcp_bits);
  ctx.stats.algorithmic_bytes_isa += rows * (w + sa_bits / 8) + (d_isa ? rows * (w + isa_bits / 8) : 0) +
                                     (d_lcp ? rows * (w + lcp_bits / 8) : 0);
}

// Index: the construction width of the lists; R: that of the SA and rank
// arrays (Index, or 32 bits for a 64-bit build of a single text whose SA and
// ISA are stored in 32 bits: the doubling then writes them in their stored
// width, and the values, rows and positions below n <= 2^32, fit).
template <typename Index, typename R>
void BuildTyped(BuildContext& ctx, const std::uint8_t* d_text, std::uint32_t flags,
                void* d_sa, std::uint32_t sa_bits, void* d_isa, std::uint32_t isa_bits,
                void* d_lcp, std::uint32_t lcp_bits) {
  sfk_hip_builder* b = ctx.b;
  const std::uint64_t n = ctx.n;
  constexpr std::uint32_t kIndexBits = 8 * sizeof(R);
  constexpr bool kNarrowRank = sizeof(R) < sizeof(Index);
  if (kNarrowRank && (ctx.sampling > 1 || ctx.refs > 0 || n > (1ull << 32))) {
    Fail(SFK_HIP_BUILD_FAILURE, "stored-width SA and rank arrays taken for a build that cannot use them");
  }
  const std::uint64_t words = PackedWords(n);

  // Workspace (grow-only).
  b->packed.Reserve(words * 16);
  b->special_blocks.Reserve(SpecialBlockWords(n) * sizeof(std::uint32_t));
  // Sampled outputs (K > 1) are shorter than the complete arrays the build
  // works on, so those stay in the workspace.
  const bool sampled = ctx.sampling > 1;
  const bool sa_direct = !sampled && sa_bits == kIndexBits;
  const bool isa_direct = !sampled && d_isa != nullptr && isa_bits == kIndexBits;
  if (!sa_direct) b->sa.Reserve(n * sizeof(R));
  if (!isa_direct) b->rank.Reserve(n * sizeof(R));
  b->act_idx.Reserve(ActIdxBytes(n, sizeof(Index)));
  b->act_seg.Reserve(n * sizeof(std::uint32_t));
  b->keys0.Reserve(std::max<std::uint64_t>(n * 8, 64));
  b->keys1.Reserve(std::max<std::uint64_t>(n * 8, 64));
  b->vals0.Reserve(n * sizeof(Index));
  b->vals1.Reserve(n * sizeof(Index));
  const std::uint64_t seg_cap = n / 2 + 2;
  b->seg_delta0.Reserve(seg_cap * sizeof(Index));
  b->seg_delta1.Reserve(seg_cap * sizeof(Index));
  b->seg_start.Reserve(seg_cap * sizeof(Index));
  b->radix_counts.Reserve(2 * kRadixSetStride);
  ReserveOnesweep(b, n);
  b->partial.Reserve(std::max(kRefineMaxChunks * sizeof(RefinePartial), 1024 * sizeof(SplitPartial)));

  R* sa = sa_direct ? static_cast<R*>(d_sa) : b->sa.As<R>();
  R* rank = isa_direct ? static_cast<R*>(d_isa) : b->rank.As<R>();
  PackedText text{b->packed.As<uint4>(), n};

  ResetCounters(b);
  // Both radix count sets start cleared (each sort then clears the other).
  SFK_HIP_CHECK(hipMemsetAsync(b->radix_counts.As<void>(), 0, 2 * kRadixSetStride, b->stream));
  b->radix_set = 0;
  RecordPhase(b, 0);
  SFK_HIP_CHECK(hipMemsetAsync(b->special_blocks.As<void>(), 0, SpecialBlockWords(n) * sizeof(std::uint32_t),
                               b->stream));
  PackTextKernel<<<static_cast<unsigned>(CeilDiv(words, 256)), 256, 0, b->stream>>>(
      d_text, n, words, b->packed.As<uint4>(),
      (reinterpret_cast<std::uintptr_t>(d_text) % 16) == 0, ctx.refs > 0,
      b->counters.As<DeviceCounters>(), b->special_blocks.As<std::uint32_t>());
  SFK_HIP_CHECK(hipGetLastError());
  if (ctx.refs > 0) {
    CheckReferenceEndsKernel<<<static_cast<unsigned>(CeilDiv(ctx.refs, 256)), 256, 0, b->stream>>>(
        d_text, ctx.d_offsets, ctx.refs, b->counters.As<DeviceCounters>());
    SFK_HIP_CHECK(hipGetLastError());
  }
  ctx.stats.algorithmic_bytes_sa += n + words * 16;

  // Early LCP into the caller's LCP buffer (complete builds only: a sampled
  // build computes its LCP in the workspace).
  const bool want_lcp = d_lcp != nullptr && (flags & SFK_HIP_RETAIN_LCP);
  EarlyLcp early{nullptr, nullptr};
  if (want_lcp && !sampled) {
    if (lcp_bits == 32) {
      early.lcp32 = static_cast<std::uint32_t*>(d_lcp);
    } else {
      early.lcp64 = static_cast<std::uint64_t*>(d_lcp);
    }
  }
  Doubling<Index, R>(ctx, sa, rank, text, early).Run();
  if (n >= 3 && g_swap_sa_rows.exchange(false)) {
    SwapSaRowsKernel<R><<<1, 1, 0, b->stream>>>(sa);
    SFK_HIP_CHECK(hipGetLastError());
  }
  RecordPhase(b, 1);

  // K9: permutation + inverse check on the device (fused into the PLCP pass
  // when the LCP is built).
  if (!want_lcp) {
    ValidateKernel<R><<<static_cast<unsigned>(CeilDiv(n, 256)), 256, 0, b->stream>>>(
        sa, rank, n, b->counters.As<DeviceCounters>());
    SFK_HIP_CHECK(hipGetLastError());
    ctx.stats.algorithmic_bytes_isa += n * 2 * sizeof(R);
  }
  RecordPhase(b, 2);
  if (want_lcp) {
    if (sampled) {
      // The complete LCP into the workspace (vals0 is free after the
      // doubling), for the interval minima of the sampled rows.
      if constexpr (!kNarrowRank) LaunchLcp<Index, Index>(ctx, d_text, text, sa, rank, b->vals0.As<Index>());
    } else if (lcp_bits == 32) {
      LaunchLcpMarked<R, std::uint32_t>(ctx, d_text, text, sa, rank, static_cast<std::uint32_t*>(d_lcp));
    } else {
      LaunchLcpMarked<R, std::uint64_t>(ctx, d_text, text, sa, rank, static_cast<std::uint64_t*>(d_lcp));
    }
  }
  RecordPhase(b, 3);
  // SA and ISA in their stored widths (and, for a batch, per-reference
  // coordinates).
  if constexpr (!kNarrowRank) {
    if (ctx.refs > 0) {
      const unsigned grid = static_cast<unsigned>(CeilDiv(n, kLocalizeTile));
      auto localize = [&](auto* out_sa) {
        using OutSa = std::remove_pointer_t<decltype(out_sa)>;
        if (d_isa == nullptr) {
          LocalizeKernel<Index, OutSa, std::uint32_t><<<grid, 256, 0, b->stream>>>(
              sa, rank, n, ctx.d_offsets, ctx.refs, out_sa, static_cast<std::uint32_t*>(nullptr),
              b->counters.As<DeviceCounters>());
        } else if (isa_bits == 32) {
          LocalizeKernel<Index, OutSa, std::uint32_t><<<grid, 256, 0, b->stream>>>(
              sa, rank, n, ctx.d_offsets, ctx.refs, out_sa, static_cast<std::uint32_t*>(d_isa),
              b->counters.As<DeviceCounters>());
        } else {
          LocalizeKernel<Index, OutSa, std::uint64_t><<<grid, 256, 0, b->stream>>>(
              sa, rank, n, ctx.d_offsets, ctx.refs, out_sa, static_cast<std::uint64_t*>(d_isa),
              b->counters.As<DeviceCounters>());
        }
      };
      if (sa_bits == 32) {
        localize(static_cast<std::uint32_t*>(d_sa));
      } else {
        localize(static_cast<std::uint64_t*>(d_sa));
      }
      SFK_HIP_CHECK(hipGetLastError());
      ctx.stats.algorithmic_bytes_isa += n * (2 * sizeof(Index) + sa_bits / 8 + isa_bits / 8);
    } else if (sampled) {
      SampleOutputs<Index>(ctx, sa, want_lcp ? b->vals0.As<Index>() : nullptr, d_sa, sa_bits, d_isa, isa_bits,
                           want_lcp ? d_lcp : nullptr, lcp_bits);
    }
  }
  if (ctx.refs == 0 && !sampled) {
    if (d_isa != nullptr && !isa_direct) {
      ConvertTo<R>(b, rank, n, d_isa, isa_bits);
      ctx.stats.algorithmic_bytes_isa += n * (sizeof(R) + isa_bits / 8);
    }
    if (!sa_direct) {
      ConvertTo<R>(b, sa, n, d_sa, sa_bits);
      ctx.stats.algorithmic_bytes_isa += n * (sizeof(R) + sa_bits / 8);
    }
  }
  RecordPhase(b, 4);
  const DeviceCounters& c = ReadCounters(b);
  if (c.error_flags & (kFlagNotPermutation | kFlagNotInverse)) {
    Fail(SFK_HIP_BUILD_FAILURE, "suffix array is not a permutation with an inverse");
  }
  if (c.error_flags & kFlagRange) {
    Fail(SFK_HIP_BUILD_FAILURE, "value does not fit the requested stored width");
  }
  ctx.stats.lce_long_items = c.long_items;
  if (want_lcp) ctx.stats.algorithmic_bytes_lcp += c.irreducible * 32;  // window pair at Φ(i) per irreducible i
  if (early.lcp32 != nullptr || early.lcp64 != nullptr) {
    ctx.stats.lcp_gathered_rows = c.lcp_marked;  // rows whose LCP came from PLCP (the rest were early)
  } else if (want_lcp) {
    ctx.stats.lcp_gathered_rows = n;
  }
  ctx.stats.lce_long_words = c.long_words;
  if (b->phase[0] == nullptr) return;  // untimed build
  float ms = 0.f;
  SFK_HIP_CHECK(hipEventElapsedTime(&ms, b->phase[0], b->phase[1]));
  ctx.stats.sa_seconds = ms * 1e-3;
  SFK_HIP_CHECK(hipEventElapsedTime(&ms, b->phase[2], b->phase[3]));
  ctx.stats.lcp_seconds = ms * 1e-3;
  float ms_isa = 0.f;
  SFK_HIP_CHECK(hipEventElapsedTime(&ms, b->phase[1], b->phase[2]));
  SFK_HIP_CHECK(hipEventElapsedTime(&ms_isa, b->phase[3], b->phase[4]));
  ctx.stats.isa_seconds = (ms + ms_isa) * 1e-3;
  SFK_HIP_CHECK(hipEventElapsedTime(&ms, b->phase[0], b->phase[4]));
  ctx.stats.device_seconds = ms * 1e-3;
  if (ctx.profile) {
    for (std::size_t e = 0; e < b->events_used; ++e) {
      SFK_HIP_CHECK(hipEventElapsedTime(&ms, b->events[e].begin, b->events[e].end));
      ctx.stats.sort_scatter_seconds += ms * 1e-3;
      ctx.stats.sort_scatter_launches += 1;
      ctx.stats.sort_scatter_bytes += b->events[e].bytes;
    }
  }
}

void CopyError(const std::string& message, char* error, size_t capacity) {
  if (error == nullptr || capacity == 0) return;
  const std::size_t len = std::min(message.size(), capacity - 1);
  std::memcpy(error, message.data(), len);
  error[len] = '\0';
}

int ValidateArgs(std::uint64_t n, std::uint32_t construction_bits, std::uint32_t sa_bits,
                 const void* isa, std::uint32_t isa_bits, const void* lcp,
                 std::uint32_t lcp_bits, std::uint32_t flags, std::string& message,
                 std::uint32_t sampling = 1) {
  if (construction_bits != 32 && construction_bits != 64) {
    message = "construction width must be 32 or 64 bits";
    return SFK_HIP_INVALID_INPUT;
  }
  if (sampling == 0) {
    message = "sampling rate must be greater than zero";
    return SFK_HIP_INVALID_INPUT;
  }
  if (n < 1) {
    message = "reference text is empty";
    return SFK_HIP_INVALID_INPUT;
  }
  if (construction_bits == 32 && n > 0xFFFFFFFFull) {
    message = "reference is too large for the 32-bit HIP backend";
    return SFK_HIP_INVALID_INPUT;
  }
  if (n > SFK_HIP_MAX_TEXT_SYMBOLS) {
    message = "texts of more than " + std::to_string(SFK_HIP_MAX_TEXT_SYMBOLS) +
              " symbols are not supported by the HIP backend";
    return SFK_HIP_UNSUPPORTED_BACKEND;
  }
  if (sa_bits != 32 && sa_bits != 64) {
    message = "SA output width must be 32 or 64 bits";
    return SFK_HIP_INVALID_INPUT;
  }
  if (isa != nullptr && isa_bits != 32 && isa_bits != 64) {
    message = "ISA output width must be 32 or 64 bits";
    return SFK_HIP_INVALID_INPUT;
  }
  if (lcp != nullptr && (flags & SFK_HIP_RETAIN_LCP) && lcp_bits != 32 && lcp_bits != 64) {
    message = "LCP output width must be 32 or 64 bits";
    return SFK_HIP_INVALID_INPUT;
  }
  const std::uint64_t max_value = n - 1;
  if ((sa_bits == 32 && max_value > 0xFFFFFFFFull) ||
      (isa != nullptr && isa_bits == 32 && max_value > 0xFFFFFFFFull) ||
      (lcp != nullptr && (flags & SFK_HIP_RETAIN_LCP) && lcp_bits == 32 && max_value > 0xFFFFFFFFull)) {
    message = "requested output width cannot represent the text coordinates";
    return SFK_HIP_INVALID_INPUT;
  }
  return SFK_HIP_OK;
}

int RunDeviceBuild(sfk_hip_builder* b, const std::uint8_t* d_text, std::uint64_t n,
                   std::uint32_t construction_bits, std::uint32_t flags, void* d_sa,
                   std::uint32_t sa_bits, void* d_isa, std::uint32_t isa_bits, void* d_lcp,
                   std::uint32_t lcp_bits, sfk_hip_stats* stats, std::string& message,
                   const std::uint64_t* offsets = nullptr, std::uint32_t refs = 0,
                   std::uint32_t sampling = 1) {
  BuildContext ctx{b, n, (flags & SFK_HIP_PROFILE_KERNELS) != 0, sfk_hip_stats{}};
  ctx.sampling = sampling;
  b->events_used = 0;
  SFK_HIP_CHECK(hipSetDevice(b->device));
  if (refs > 0) {
    b->ref_offsets.Reserve((static_cast<std::uint64_t>(refs) + 1) * 8);
    SFK_HIP_CHECK(hipMemcpyAsync(b->ref_offsets.As<void>(), offsets,
                                 (static_cast<std::uint64_t>(refs) + 1) * 8, hipMemcpyHostToDevice,
                                 b->stream));
    ctx.d_offsets = b->ref_offsets.As<std::uint64_t>();
    ctx.refs = refs;
  }
  // A 64-bit build of one text whose SA and ISA are stored in 32 bits
  // writes them in 32 bits during the doubling (no 64-bit copies to narrow
  // afterwards, and the memory they would take goes to the small lists).
  const bool narrow_rank = refs == 0 && sampling == 1 && n <= (1ull << 32) && sa_bits == 32 &&
                           (d_isa == nullptr || isa_bits == 32);
  if (construction_bits == 32) {
    BuildTyped<std::uint32_t, std::uint32_t>(ctx, d_text, flags, d_sa, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits);
  } else if (narrow_rank) {
    BuildTyped<std::uint64_t, std::uint32_t>(ctx, d_text, flags, d_sa, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits);
  } else {
    BuildTyped<std::uint64_t, std::uint64_t>(ctx, d_text, flags, d_sa, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits);
  }
  if (stats != nullptr) {
    const double h2d = stats->h2d_seconds, d2h = stats->d2h_seconds;
    *stats = ctx.stats;
    stats->h2d_seconds = h2d;
    stats->d2h_seconds = d2h;
  }
  (void)message;
  return SFK_HIP_OK;
}

}  // namespace
}  // namespace sfk

extern "C" {

int sfk_hip_abi_version(void) { return SFK_HIP_ABI_VERSION; }

void sfk_hip_inject_allocation_failure(long long after) {
  sfk::g_fail_allocation.store(after < 0 ? -1 : after);
}

void sfk_hip_inject_sa_swap(int armed) { sfk::g_swap_sa_rows.store(armed != 0); }

int sfk_hip_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return count;
}

int sfk_hip_builder_create(int device, sfk_hip_builder** out, char* error,
                           size_t error_capacity) {
  if (out == nullptr) return SFK_HIP_INVALID_INPUT;
  *out = nullptr;
  sfk_hip_builder* b = nullptr;
  try {
    if (sfk_hip_device_count() <= device || device < 0) {
      sfk::Fail(SFK_HIP_UNSUPPORTED_BACKEND, "no HIP device " + std::to_string(device));
    }
    b = new sfk_hip_builder();
    b->device = device;
    SFK_HIP_CHECK(hipSetDevice(device));
    SFK_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    SFK_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b->host_counters),
                                sizeof(sfk::DeviceCounters), hipHostMallocDefault));
    b->counters.Reserve(sizeof(sfk::DeviceCounters));
    for (auto& e : b->phase) SFK_HIP_CHECK(hipEventCreate(&e));
    *out = b;
    return SFK_HIP_OK;
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    sfk_hip_builder_destroy(b);
    return f.status;
  } catch (const std::bad_alloc&) {
    sfk::CopyError("host allocation failed", error, error_capacity);
    sfk_hip_builder_destroy(b);
    return SFK_HIP_BUILD_FAILURE;
  }
}

void sfk_hip_builder_destroy(sfk_hip_builder* b) {
  if (b == nullptr) return;
  (void)hipSetDevice(b->device);
  if (b->stream != nullptr) (void)hipStreamSynchronize(b->stream);
  for (auto& e : b->events) {
    (void)hipEventDestroy(e.begin);
    (void)hipEventDestroy(e.end);
  }
  for (auto& e : b->phase) {
    if (e != nullptr) (void)hipEventDestroy(e);
  }
  if (b->host_counters != nullptr) (void)hipHostFree(b->host_counters);
  if (b->stream != nullptr) (void)hipStreamDestroy(b->stream);
  delete b;
}

void* sfk_hip_builder_stream(sfk_hip_builder* b) {
  return b == nullptr ? nullptr : static_cast<void*>(b->stream);
}

uint64_t sfk_hip_workspace_bytes(uint64_t n, uint32_t construction_bits) {
  const uint64_t w = construction_bits == 64 ? 8 : 4;
  const uint64_t words = sfk::PackedWords(n);
  return words * 16 + n * (w + 4) /*act*/ + n * 16 /*keys*/ + n * w * 2 /*vals*/ +
         (n / 2 + 2) * w * 3 /*segment deltas, starts*/ + n * w * 2 /*sa, rank when not direct*/ +
         n * w * 5 /*small lists (at most the first step's survivors)*/ +
         2 * sfk::kRadixSetStride +
         std::max(sfk::kRefineMaxChunks * sizeof(sfk::RefinePartial), 1024 * sizeof(sfk::SplitPartial)) +
         sfk::OnesweepStatusBytes(n);
}

int sfk_hip_build_device(sfk_hip_builder* b, const uint8_t* d_text, uint64_t n,
                         uint32_t construction_bits, uint32_t sampling_rate, uint32_t flags,
                         void* d_sa, uint32_t sa_bits, void* d_isa, uint32_t isa_bits,
                         void* d_lcp, uint32_t lcp_bits, sfk_hip_stats* stats, char* error,
                         size_t error_capacity) {
  std::string message;
  if (b == nullptr || d_text == nullptr || d_sa == nullptr) {
    sfk::CopyError("builder, text and SA buffers are required", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  int status = sfk::ValidateArgs(n, construction_bits, sa_bits, d_isa, isa_bits, d_lcp,
                                 lcp_bits, flags, message, sampling_rate);
  if (status != SFK_HIP_OK) {
    sfk::CopyError(message, error, error_capacity);
    return status;
  }
  try {
    return sfk::RunDeviceBuild(b, d_text, n, construction_bits, flags, d_sa, sa_bits, d_isa,
                               isa_bits, d_lcp, lcp_bits, stats, message, nullptr, 0, sampling_rate);
  } catch (const sfk::Failure& f) {
    (void)hipStreamSynchronize(b->stream);
    sfk::CopyError(f.message, error, error_capacity);
    return f.status;
  } catch (const std::bad_alloc&) {
    sfk::CopyError("host allocation failed", error, error_capacity);
    return SFK_HIP_BUILD_FAILURE;
  }
}

int sfk_hip_build_batch_device(sfk_hip_builder* b, const uint8_t* d_text, uint64_t n,
                               const uint64_t* offsets, uint32_t refs, uint32_t construction_bits,
                               uint32_t flags, void* d_sa, uint32_t sa_bits, void* d_isa,
                               uint32_t isa_bits, void* d_lcp, uint32_t lcp_bits,
                               sfk_hip_stats* stats, char* error, size_t error_capacity) {
  std::string message;
  if (b == nullptr || d_text == nullptr || d_sa == nullptr || offsets == nullptr || refs == 0) {
    sfk::CopyError("builder, text, SA buffer and at least one reference are required", error,
                   error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  if (offsets[0] != 0 || offsets[refs] != n) {
    sfk::CopyError("reference offsets must start at 0 and end at n", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  for (uint32_t r = 0; r < refs; ++r) {
    if (offsets[r + 1] < offsets[r] + 2) {
      sfk::CopyError("every reference needs at least one symbol and its sentinel", error,
                     error_capacity);
      return SFK_HIP_INVALID_INPUT;
    }
  }
  int status = sfk::ValidateArgs(n, construction_bits, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits,
                                 flags, message);
  if (status != SFK_HIP_OK) {
    sfk::CopyError(message, error, error_capacity);
    return status;
  }
  try {
    return sfk::RunDeviceBuild(b, d_text, n, construction_bits, flags, d_sa, sa_bits, d_isa,
                               isa_bits, d_lcp, lcp_bits, stats, message, offsets, refs);
  } catch (const sfk::Failure& f) {
    (void)hipStreamSynchronize(b->stream);
    sfk::CopyError(f.message, error, error_capacity);
    return f.status;
  } catch (const std::bad_alloc&) {
    sfk::CopyError("host allocation failed", error, error_capacity);
    return SFK_HIP_BUILD_FAILURE;
  }
}

// ------------------------------------------------- fast prefix directory
//
// Selection of k follows FastPrefixIndex::SelectK with its default options
// (fast_prefix_index.cpp:286-307, .hpp FastPrefixIndexOptions): no directory
// below 2^16 suffixes; otherwise the largest k in [8, 10] whose table
// (4^k entries of two row-width coordinates) fits 25% of the resident core
// bytes, else none.
namespace sfk {
namespace {

constexpr std::uint32_t kDirectoryMinK = 8;
constexpr std::uint32_t kDirectoryMaxK = 10;
constexpr std::uint64_t kDirectoryMinSuffixes = 1ull << 16;
constexpr std::uint64_t kDirectoryBudgetBasisPoints = 2500;

bool DirectoryWide(std::uint64_t suffix_count) { return suffix_count > 0xFFFFFFFFull; }

std::uint64_t DirectoryBytesFor(std::uint32_t k, std::uint64_t suffix_count) {
  if (k == 0 || k > kDirectoryMaxK) return 0;
  return (1ull << (2 * k)) * 2 * (DirectoryWide(suffix_count) ? 8 : 4);
}

std::uint32_t SelectDirectoryK(std::uint64_t suffix_count, std::uint64_t resident_core_bytes) {
  if (suffix_count < kDirectoryMinSuffixes) return 0;
  // 25% of the core bytes, split around the division as the policy does.
  const std::uint64_t budget = (resident_core_bytes / 10000) * kDirectoryBudgetBasisPoints +
                               ((resident_core_bytes % 10000) * kDirectoryBudgetBasisPoints) / 10000;
  for (std::uint32_t k = kDirectoryMaxK; k >= kDirectoryMinK; --k) {
    if (DirectoryBytesFor(k, suffix_count) <= budget) return k;
  }
  return 0;
}

// Fills the table and runs the directory pass on `stream`; counts (2 words
// of device memory followed by kDirectorySinkWords of sink) receive the
// indexed k-mers and the non-empty entries.
static_assert((2 + kDirectorySinkWords) * sizeof(unsigned long long) <= 1024 * sizeof(SplitPartial),
              "the builder's partial buffer holds the directory counts and sink");
void LaunchPrefixDirectory(hipStream_t stream, PackedText t, const void* d_sa, std::uint32_t sa_bits,
                           std::uint64_t n, std::uint32_t k, std::uint64_t* table, unsigned long long* counts) {
  const bool wide = DirectoryWide(n);
  const std::uint64_t entries = 1ull << (2 * k);
  const std::uint64_t words = wide ? 2 * entries : entries;
  const std::uint64_t empty = wide ? n : (n | (n << 32));
  PrefixDirectoryFillKernel<<<static_cast<unsigned>(CeilDiv(words, 256)), 256, 0, stream>>>(table, words, empty);
  SFK_HIP_CHECK(hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), stream));
  const unsigned grid = static_cast<unsigned>(CeilDiv(n, kDirectoryTile));
  const int kk = static_cast<int>(k);
  if (sa_bits == 32) {
    const auto* sa = static_cast<const std::uint32_t*>(d_sa);
    if (wide) {
      PrefixDirectoryKernel<std::uint32_t, true><<<grid, 256, 0, stream>>>(t, sa, n, kk, table, counts);
    } else {
      PrefixDirectoryKernel<std::uint32_t, false><<<grid, 256, 0, stream>>>(t, sa, n, kk, table, counts);
    }
  } else {
    const auto* sa = static_cast<const std::uint64_t*>(d_sa);
    if (wide) {
      PrefixDirectoryKernel<std::uint64_t, true><<<grid, 256, 0, stream>>>(t, sa, n, kk, table, counts);
    } else {
      PrefixDirectoryKernel<std::uint64_t, false><<<grid, 256, 0, stream>>>(t, sa, n, kk, table, counts);
    }
  }
  SFK_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace sfk

int sfk_hip_build(const uint8_t* text, uint64_t n, uint32_t construction_bits, uint32_t sampling_rate,
                  uint32_t flags, void* sa, uint32_t sa_bits, void* isa, uint32_t isa_bits, void* lcp,
                  uint32_t lcp_bits, sfk_hip_stats* stats, char* error, size_t error_capacity) {
  return sfk_hip_build_fast(text, n, construction_bits, sampling_rate, flags, sa, sa_bits, isa, isa_bits, lcp,
                            lcp_bits, 0, nullptr, nullptr, stats, error, error_capacity);
}

int sfk_hip_build_fast(const uint8_t* text, uint64_t n, uint32_t construction_bits, uint32_t sampling_rate,
                       uint32_t flags, void* sa, uint32_t sa_bits, void* isa, uint32_t isa_bits, void* lcp,
                       uint32_t lcp_bits, uint32_t prefix_k, uint64_t* prefix_table, uint64_t prefix_counts[2],
                       sfk_hip_stats* stats, char* error, size_t error_capacity) {
  std::string message;
  if (text == nullptr || sa == nullptr) {
    sfk::CopyError("text and SA buffers are required", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  if (prefix_k > sfk::kDirectoryMaxK || (prefix_k > 0 && prefix_table == nullptr) ||
      (prefix_k > 0 && sampling_rate != 1)) {
    sfk::CopyError("prefix directory k must be 0..10 with a table for k > 0, of a complete (K = 1) index", error,
                   error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  int status = sfk::ValidateArgs(n, construction_bits, sa_bits, isa, isa_bits, lcp, lcp_bits,
                                 flags, message, sampling_rate);
  if (status != SFK_HIP_OK) {
    sfk::CopyError(message, error, error_capacity);
    return status;
  }
  const auto t_begin = std::chrono::steady_clock::now();
  sfk_hip_builder* b = nullptr;
  status = sfk_hip_builder_create(0, &b, error, error_capacity);
  if (status != SFK_HIP_OK) return status;
  sfk::DeviceBuffer d_text, d_sa, d_isa, d_lcp;
  const bool want_isa = isa != nullptr;
  const bool want_lcp = lcp != nullptr && (flags & SFK_HIP_RETAIN_LCP);
  const std::uint64_t rows = sfk::SampledRows(n, sampling_rate);
  try {
    d_text.Reserve(n + 64);
    d_sa.Reserve(rows * (sa_bits / 8));
    if (want_isa) d_isa.Reserve(rows * (isa_bits / 8));
    if (want_lcp) d_lcp.Reserve(rows * (lcp_bits / 8));
    // Copies go through pinned staging on the builder's stream; each one
    // returns after its data has landed, so the times are the copies'.
    const auto t0 = std::chrono::steady_clock::now();
    SFK_HIP_CHECK(sfk::StagedCopy(b->device, d_text.As<void>(), text, n, hipMemcpyHostToDevice, b->stream));
    sfk_hip_stats local{};
    local.h2d_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    status = sfk::RunDeviceBuild(b, d_text.As<uint8_t>(), n, construction_bits, flags,
                                 d_sa.As<void>(), sa_bits, want_isa ? d_isa.As<void>() : nullptr,
                                 isa_bits, want_lcp ? d_lcp.As<void>() : nullptr, lcp_bits,
                                 &local, message, nullptr, 0, sampling_rate);
    sfk::DeviceBuffer d_table;
    if (status == SFK_HIP_OK && prefix_k > 0) {
      // The directory from the device SA while it is resident (the packed
      // text of the build is still in the builder).
      const std::uint64_t words = sfk_prefix_directory_words(prefix_k, n);
      d_table.Reserve(words * 8);
      b->partial.Reserve(1024 * sizeof(sfk::SplitPartial));
      auto* dcounts = b->partial.As<unsigned long long>();
      const auto td = std::chrono::steady_clock::now();
      sfk::LaunchPrefixDirectory(b->stream, sfk::PackedText{b->packed.As<uint4>(), n}, d_sa.As<void>(), sa_bits, n,
                                 prefix_k, d_table.As<std::uint64_t>(), dcounts);
      unsigned long long host[2] = {0, 0};
      SFK_HIP_CHECK(hipMemcpyAsync(host, dcounts, sizeof(host), hipMemcpyDeviceToHost, b->stream));
      SFK_HIP_CHECK(hipStreamSynchronize(b->stream));
      local.prefix_directory_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - td).count();
      SFK_HIP_CHECK(sfk::StagedCopy(b->device, prefix_table, d_table.As<void>(), words * 8, hipMemcpyDeviceToHost,
                                    b->stream));
      if (prefix_counts != nullptr) {
        prefix_counts[0] = host[0];
        prefix_counts[1] = host[1];
      }
    }
    const auto t2 = std::chrono::steady_clock::now();
    SFK_HIP_CHECK(sfk::StagedCopy(b->device, sa, d_sa.As<void>(), rows * (sa_bits / 8), hipMemcpyDeviceToHost,
                                  b->stream));
    if (want_isa) {
      SFK_HIP_CHECK(sfk::StagedCopy(b->device, isa, d_isa.As<void>(), rows * (isa_bits / 8), hipMemcpyDeviceToHost,
                                    b->stream));
    }
    if (want_lcp) {
      SFK_HIP_CHECK(sfk::StagedCopy(b->device, lcp, d_lcp.As<void>(), rows * (lcp_bits / 8), hipMemcpyDeviceToHost,
                                    b->stream));
    }
    local.d2h_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
    local.call_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    if (stats != nullptr) *stats = local;
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    status = f.status;
  }
  d_text.Release();
  d_sa.Release();
  d_isa.Release();
  d_lcp.Release();
  sfk_hip_builder_destroy(b);
  return status;
}

uint32_t sfk_prefix_directory_select_k(uint64_t suffix_count, uint64_t resident_core_bytes) {
  return sfk::SelectDirectoryK(suffix_count, resident_core_bytes);
}

uint64_t sfk_prefix_directory_words(uint32_t k, uint64_t suffix_count) {
  if (k == 0 || k > sfk::kDirectoryMaxK) return 0;
  return (1ull << (2 * k)) * (sfk::DirectoryWide(suffix_count) ? 2 : 1);
}

int sfk_hip_prefix_directory(sfk_hip_builder* b, const uint8_t* d_text, uint64_t n, const void* d_sa,
                             uint32_t sa_bits, uint32_t k, uint64_t* d_table, uint64_t counts[2], char* error,
                             size_t error_capacity) {
  if (b == nullptr || d_text == nullptr || d_sa == nullptr || d_table == nullptr || n < 2 || n > SFK_HIP_MAX_TEXT_SYMBOLS || k < 1 ||
      k > sfk::kDirectoryMaxK || (sa_bits != 32 && sa_bits != 64)) {
    sfk::CopyError("prefix directory needs a builder, a text, its SA (32/64-bit), a table and 1 <= k <= 10",
                   error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  try {
    SFK_HIP_CHECK(hipSetDevice(b->device));
    const std::uint64_t words = sfk::PackedWords(n);
    b->packed.Reserve(words * 16);
    b->counters.Reserve(sizeof(sfk::DeviceCounters));
    sfk::ResetCounters(b);
    sfk::PackTextKernel<<<static_cast<unsigned>(sfk::CeilDiv(words, 256)), 256, 0, b->stream>>>(
        d_text, n, words, b->packed.As<uint4>(), (reinterpret_cast<std::uintptr_t>(d_text) % 16) == 0, false,
        b->counters.As<sfk::DeviceCounters>(), nullptr);
    SFK_HIP_CHECK(hipGetLastError());
    b->partial.Reserve(1024 * sizeof(sfk::SplitPartial));
    auto* dcounts = b->partial.As<unsigned long long>();
    sfk::LaunchPrefixDirectory(b->stream, sfk::PackedText{b->packed.As<uint4>(), n}, d_sa, sa_bits, n, k, d_table,
                               dcounts);
    unsigned long long host[2] = {0, 0};
    SFK_HIP_CHECK(hipMemcpyAsync(host, dcounts, sizeof(host), hipMemcpyDeviceToHost, b->stream));
    const sfk::DeviceCounters& c = sfk::ReadCounters(b);
    if (c.error_flags != 0) sfk::Fail(SFK_HIP_INVALID_INPUT, "text must hold symbols 0..6 with its only 0 last");
    if (counts != nullptr) {
      counts[0] = host[0];
      counts[1] = host[1];
    }
    return SFK_HIP_OK;
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    return f.status;
  }
}

// ------------------------------------------------------ resident index ---

struct sfk_hip_index {
  int device = 0;
  hipStream_t stream = nullptr;
  const uint8_t* d_text = nullptr;
  uint64_t n = 0;
  const void* d_sa = nullptr;
  uint32_t sa_bits = 32;
  sfk::DeviceBuffer packed, counters, directory, directory_counts;
  uint32_t directory_k = 0;  // 0: no directory
  sfk::DeviceCounters* host_counters = nullptr;
  // FM-index (sfk_hip_index_build_fm): occurrence records and C.
  sfk::DeviceBuffer fm_blocks, fm_header, fm_partial;
  bool fm_built = false;
};

void sfk_hip_index_destroy(sfk_hip_index* index) {
  if (index == nullptr) return;
  (void)hipSetDevice(index->device);
  if (index->stream != nullptr) (void)hipStreamSynchronize(index->stream);
  if (index->host_counters != nullptr) (void)hipHostFree(index->host_counters);
  if (index->stream != nullptr) (void)hipStreamDestroy(index->stream);
  delete index;
}

int sfk_hip_index_create(int device, const uint8_t* d_text, uint64_t n, const void* d_sa,
                         uint32_t sa_bits, sfk_hip_index** out, char* error,
                         size_t error_capacity) {
  if (out == nullptr || d_text == nullptr || d_sa == nullptr || n < 1 || n > SFK_HIP_MAX_TEXT_SYMBOLS ||
      (sa_bits != 32 && sa_bits != 64)) {
    sfk::CopyError("index needs a text, its SA and a 32/64-bit SA width", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  *out = nullptr;
  sfk_hip_index* index = nullptr;
  try {
    index = new sfk_hip_index();
    index->device = device;
    index->d_text = d_text;
    index->n = n;
    index->d_sa = d_sa;
    index->sa_bits = sa_bits;
    SFK_HIP_CHECK(hipSetDevice(device));
    SFK_HIP_CHECK(hipStreamCreateWithFlags(&index->stream, hipStreamNonBlocking));
    SFK_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&index->host_counters),
                                sizeof(sfk::DeviceCounters), hipHostMallocDefault));
    const uint64_t words = sfk::PackedWords(n);
    index->packed.Reserve(words * 16);
    index->counters.Reserve(sizeof(sfk::DeviceCounters));
    *index->host_counters = sfk::DeviceCounters{};
    SFK_HIP_CHECK(hipMemcpyAsync(index->counters.As<void>(), index->host_counters,
                                 sizeof(sfk::DeviceCounters), hipMemcpyHostToDevice, index->stream));
    sfk::PackTextKernel<<<static_cast<unsigned>(sfk::CeilDiv(words, 256)), 256, 0, index->stream>>>(
        d_text, n, words, index->packed.As<uint4>(),
        (reinterpret_cast<std::uintptr_t>(d_text) % 16) == 0, false,
        index->counters.As<sfk::DeviceCounters>(), nullptr);
    SFK_HIP_CHECK(hipGetLastError());
    SFK_HIP_CHECK(hipMemcpyAsync(index->host_counters, index->counters.As<void>(),
                                 sizeof(sfk::DeviceCounters), hipMemcpyDeviceToHost, index->stream));
    SFK_HIP_CHECK(hipStreamSynchronize(index->stream));
    if (index->host_counters->error_flags != 0) {
      sfk::Fail(SFK_HIP_INVALID_INPUT, "index text must hold symbols 0..6 with its only 0 last");
    }
    // The Fast profile's prefix directory, sized for the resident core
    // arrays of a served index (text, SA, ISA and LCP of the SA's width).
    index->directory_k = sfk::SelectDirectoryK(n, n * (1 + 3 * (sa_bits / 8)));
    if (index->directory_k > 0) {
      index->directory.Reserve(sfk_prefix_directory_words(index->directory_k, n) * 8);
      index->directory_counts.Reserve((2 + sfk::kDirectorySinkWords) * sizeof(unsigned long long));
      sfk::LaunchPrefixDirectory(index->stream, sfk::PackedText{index->packed.As<uint4>(), n}, d_sa, sa_bits, n,
                                 index->directory_k, index->directory.As<std::uint64_t>(),
                                 index->directory_counts.As<unsigned long long>());
      SFK_HIP_CHECK(hipStreamSynchronize(index->stream));
    }
    *out = index;
    return SFK_HIP_OK;
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    sfk_hip_index_destroy(index);
    return f.status;
  }
}

int sfk_hip_index_generate_queries(sfk_hip_index* index, uint64_t seed, uint64_t count,
                                   uint32_t min_len, uint32_t max_len, uint64_t* d_patterns,
                                   uint8_t* d_lengths, void* stream) {
  if (index == nullptr || min_len < 1 || max_len > 32 || min_len > max_len) {
    return SFK_HIP_INVALID_INPUT;
  }
  if (count == 0) return SFK_HIP_OK;
  hipStream_t s = stream != nullptr ? static_cast<hipStream_t>(stream) : index->stream;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  sfk::PackedText t{index->packed.As<uint4>(), index->n};
  sfk::GenerateQueriesKernel<<<static_cast<unsigned>(sfk::CeilDiv(count, 256)), 256, 0, s>>>(
      t, seed, count, min_len, max_len, d_patterns, d_lengths);
  if (hipGetLastError() != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  if (stream == nullptr && hipStreamSynchronize(s) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return SFK_HIP_OK;
}

int sfk_hip_index_count(sfk_hip_index* index, const uint64_t* d_patterns, const uint8_t* d_lengths,
                        uint64_t count, uint32_t* d_counts, void* stream) {
  if (index == nullptr || (count > 0 && (d_patterns == nullptr || d_lengths == nullptr ||
                                         d_counts == nullptr))) {
    return SFK_HIP_INVALID_INPUT;
  }
  if (count == 0) return SFK_HIP_OK;
  hipStream_t s = stream != nullptr ? static_cast<hipStream_t>(stream) : index->stream;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  sfk::PackedText t{index->packed.As<uint4>(), index->n};
  const unsigned grid = static_cast<unsigned>(sfk::CeilDiv(count, 256));
  const std::uint64_t* dir = index->directory_k > 0 ? index->directory.As<std::uint64_t>() : nullptr;
  const int dir_k = static_cast<int>(index->directory_k);
  const bool dir_wide = sfk::DirectoryWide(index->n);
  if (index->sa_bits == 32) {
    sfk::CountPatternsKernel<uint32_t><<<grid, 256, 0, s>>>(
        t, static_cast<const uint32_t*>(index->d_sa), index->n, d_patterns, d_lengths, count, d_counts, dir,
        dir_k, dir_wide);
  } else {
    sfk::CountPatternsKernel<uint64_t><<<grid, 256, 0, s>>>(
        t, static_cast<const uint64_t*>(index->d_sa), index->n, d_patterns, d_lengths, count, d_counts, dir,
        dir_k, dir_wide);
  }
  if (hipGetLastError() != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  if (stream == nullptr && hipStreamSynchronize(s) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return SFK_HIP_OK;
}

uint64_t sfk_hip_index_device_bytes(const sfk_hip_index* index) {
  return index == nullptr ? 0 : index->packed.Bytes() + index->directory.Bytes() + index->fm_blocks.Bytes();
}

int sfk_hip_index_build_fm(sfk_hip_index* index, char* error, size_t error_capacity) {
  if (index == nullptr) {
    sfk::CopyError("no index", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  try {
    SFK_HIP_CHECK(hipSetDevice(index->device));
    const std::uint64_t n = index->n;
    const std::uint64_t nblocks = n / 64 + 1;  // occ(c, n) reads record n / 64
    index->fm_blocks.Reserve(nblocks * sizeof(sfk::FmBlock));
    index->fm_header.Reserve(sizeof(sfk::FmHeader));
    index->fm_partial.Reserve(1024 * sizeof(sfk::FmCounts));
    hipStream_t s = index->stream;
    sfk::FmBlock* blocks = index->fm_blocks.As<sfk::FmBlock>();
    sfk::FmHeader* header = index->fm_header.As<sfk::FmHeader>();
    sfk::FmCounts* partial = index->fm_partial.As<sfk::FmCounts>();
    sfk::DeviceCounters* dc = index->counters.As<sfk::DeviceCounters>();
    const sfk::PackedText t{index->packed.As<uint4>(), n};
    index->fm_built = false;
    SFK_HIP_CHECK(hipMemsetAsync(header, 0, sizeof(sfk::FmHeader), s));
    SFK_HIP_CHECK(hipMemsetAsync(&dc->error_flags, 0, sizeof(unsigned int), s));
    sfk::FmLowCountKernel<<<static_cast<unsigned>(sfk::CeilDiv(sfk::CeilDiv(n, 32), 256)), 256, 0, s>>>(t, header);
    const unsigned grid = static_cast<unsigned>(sfk::CeilDiv(nblocks * sfk::kWave, 256));
    if (index->sa_bits == 32) {
      sfk::FmBlocksKernel<uint32_t><<<grid, 256, 0, s>>>(t, static_cast<const uint32_t*>(index->d_sa), n, blocks,
                                                          nblocks, &dc->error_flags);
    } else {
      sfk::FmBlocksKernel<uint64_t><<<grid, 256, 0, s>>>(t, static_cast<const uint64_t*>(index->d_sa), n, blocks,
                                                          nblocks, &dc->error_flags);
    }
    const sfk::SortGeometry g = sfk::Geometry(nblocks, sfk::kFmTile, 1024);
    sfk::FmReduceKernel<<<g.grid, 256, 0, s>>>(blocks, nblocks, g.chunk, partial);
    sfk::FmSpineKernel<<<1, 1024, 0, s>>>(partial, g.grid, header);
    sfk::FmApplyKernel<<<g.grid, 256, 0, s>>>(blocks, nblocks, g.chunk, partial);
    SFK_HIP_CHECK(hipGetLastError());
    SFK_HIP_CHECK(hipMemcpyAsync(index->host_counters, dc, sizeof(sfk::DeviceCounters), hipMemcpyDeviceToHost, s));
    SFK_HIP_CHECK(hipStreamSynchronize(s));
    if (index->host_counters->error_flags != 0) {
      sfk::Fail(SFK_HIP_INVALID_INPUT, "index SA holds a value outside the text");
    }
    index->fm_built = true;
    return SFK_HIP_OK;
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    return f.status;
  }
}

uint64_t sfk_hip_index_fm_bytes(const sfk_hip_index* index) {
  return index == nullptr || !index->fm_built ? 0 : index->fm_blocks.Bytes();
}

int sfk_hip_index_fm_header(sfk_hip_index* index, uint64_t out[6]) {
  if (index == nullptr || !index->fm_built || out == nullptr) return SFK_HIP_INVALID_INPUT;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  sfk::FmHeader h{};
  if (hipMemcpy(&h, index->fm_header.As<void>(), sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
    return SFK_HIP_BUILD_FAILURE;
  }
  for (int c = 0; c < 5; ++c) out[c] = h.first[c];
  out[5] = h.low;
  return SFK_HIP_OK;
}

int sfk_hip_index_fm_ranges(sfk_hip_index* index, const uint8_t* d_codes, const uint64_t* d_offsets,
                            const uint8_t* d_orientation, uint64_t count, uint64_t* d_ranges, void* stream) {
  if (index == nullptr || !index->fm_built ||
      (count > 0 && (d_codes == nullptr || d_offsets == nullptr || d_orientation == nullptr || d_ranges == nullptr))) {
    return SFK_HIP_INVALID_INPUT;
  }
  if (count == 0) return SFK_HIP_OK;
  hipStream_t s = stream != nullptr ? static_cast<hipStream_t>(stream) : index->stream;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  const std::uint64_t* dir = index->directory_k > 0 ? index->directory.As<std::uint64_t>() : nullptr;
  sfk::FmRangesKernel<<<static_cast<unsigned>(sfk::CeilDiv(count, 256)), 256, 0, s>>>(
      index->fm_blocks.As<sfk::FmBlock>(), index->fm_header.As<sfk::FmHeader>(), index->n, dir,
      static_cast<int>(index->directory_k), sfk::DirectoryWide(index->n), d_codes, d_offsets, d_orientation, count,
      d_ranges);
  if (hipGetLastError() != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  if (stream == nullptr && hipStreamSynchronize(s) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return SFK_HIP_OK;
}

int sfk_hip_index_fm_count(sfk_hip_index* index, const uint64_t* d_patterns, const uint8_t* d_lengths,
                           uint64_t count, uint32_t* d_counts, void* stream) {
  if (index == nullptr || !index->fm_built ||
      (count > 0 && (d_patterns == nullptr || d_lengths == nullptr || d_counts == nullptr))) {
    return SFK_HIP_INVALID_INPUT;
  }
  if (count == 0) return SFK_HIP_OK;
  hipStream_t s = stream != nullptr ? static_cast<hipStream_t>(stream) : index->stream;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  const std::uint64_t* dir = index->directory_k > 0 ? index->directory.As<std::uint64_t>() : nullptr;
  sfk::FmCountPackedKernel<<<static_cast<unsigned>(sfk::CeilDiv(count, 256)), 256, 0, s>>>(
      index->fm_blocks.As<sfk::FmBlock>(), index->fm_header.As<sfk::FmHeader>(), index->n, dir,
      static_cast<int>(index->directory_k), sfk::DirectoryWide(index->n), d_patterns, d_lengths, count, d_counts);
  if (hipGetLastError() != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  if (stream == nullptr && hipStreamSynchronize(s) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return SFK_HIP_OK;
}

int sfk_hip_index_sa_rows(sfk_hip_index* index, uint64_t begin, uint64_t end, uint64_t* out) {
  if (index == nullptr || begin > end || end > index->n || (end > begin && out == nullptr)) {
    return SFK_HIP_INVALID_INPUT;
  }
  if (end == begin) return SFK_HIP_OK;
  if (hipSetDevice(index->device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  if (hipStreamSynchronize(index->stream) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  const std::uint64_t rows = end - begin;
  if (index->sa_bits == 64) {
    const auto* src = static_cast<const std::uint64_t*>(index->d_sa) + begin;
    if (hipMemcpy(out, src, rows * 8, hipMemcpyDeviceToHost) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
    return SFK_HIP_OK;
  }
  // 32-bit rows land in the upper half of the output, then widen in place
  // from the front (element i is read before any write can reach it).
  const auto* src = static_cast<const std::uint32_t*>(index->d_sa) + begin;
  auto* staged = reinterpret_cast<std::uint32_t*>(out) + rows;
  if (hipMemcpy(staged, src, rows * 4, hipMemcpyDeviceToHost) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  for (std::uint64_t i = 0; i < rows; ++i) out[i] = staged[i];
  return SFK_HIP_OK;
}

uint32_t sfk_hip_index_prefix_k(const sfk_hip_index* index) {
  return index == nullptr ? 0 : index->directory_k;
}

int sfk_hip_malloc(int device, uint64_t bytes, void** out) {
  if (out == nullptr) return SFK_HIP_INVALID_INPUT;
  if (hipSetDevice(device) != hipSuccess) return SFK_HIP_UNSUPPORTED_BACKEND;
  if (hipMalloc(out, bytes == 0 ? 1 : bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return SFK_HIP_BUILD_FAILURE;
  }
  return SFK_HIP_OK;
}

int sfk_hip_memcpy_d2d(void* dst, const void* src, uint64_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? SFK_HIP_OK
                                                                           : SFK_HIP_BUILD_FAILURE;
}

int sfk_hip_device_info(int device, uint64_t out[8]) {
  hipDeviceProp_t prop;
  if (out == nullptr || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    (void)hipGetLastError();
    return SFK_HIP_UNSUPPORTED_BACKEND;
  }
  out[0] = prop.sharedMemPerBlock;
  out[1] = prop.maxSharedMemoryPerMultiProcessor;
  out[2] = static_cast<uint64_t>(prop.multiProcessorCount);
  out[3] = static_cast<uint64_t>(prop.l2CacheSize);
  out[4] = prop.totalGlobalMem;
  out[5] = static_cast<uint64_t>(prop.clockRate);
  out[6] = static_cast<uint64_t>(prop.memoryClockRate);
  out[7] = static_cast<uint64_t>(prop.memoryBusWidth);
  return SFK_HIP_OK;
}

int sfk_hip_free(void* ptr) {
  return hipFree(ptr) == hipSuccess ? SFK_HIP_OK : SFK_HIP_BUILD_FAILURE;
}

// Synchronous copies between pageable host memory and the device, through
// the device's pinned stager when it is free.  Both return after the data
// has landed, so work on any stream launched after the call sees it.
int sfk_hip_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return sfk::StagedCopy(device, dst, src, bytes, hipMemcpyHostToDevice, nullptr) == hipSuccess
             ? SFK_HIP_OK
             : SFK_HIP_BUILD_FAILURE;
}

int sfk_hip_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return SFK_HIP_BUILD_FAILURE;
  return sfk::StagedCopy(device, dst, src, bytes, hipMemcpyDeviceToHost, nullptr) == hipSuccess
             ? SFK_HIP_OK
             : SFK_HIP_BUILD_FAILURE;
}

}  // extern "C"

namespace {

void TuneEntryOf(int kind, int size_class, int shape, int bits, double us, bool chosen, bool tuned,
                 sfk_hip_tune_entry* e) {
  e->kind = static_cast<uint32_t>(kind);
  e->size_class = static_cast<uint32_t>(size_class);
  e->block = static_cast<uint32_t>(sfk::kShapes[shape].block);
  e->items = static_cast<uint32_t>(sfk::kShapes[shape].items);
  e->us_per_pass = us;
  e->chosen = chosen ? 1u : 0u;
  e->tuned = tuned ? 1u : 0u;
  e->digit_bits = static_cast<uint32_t>(bits);
}

}  // namespace

int sfk_hip_autotune(int device, uint64_t max_elements, sfk_hip_tune_entry* entries, uint32_t capacity,
                     uint32_t* count, char* error, size_t error_capacity) {
  if (count != nullptr) *count = 0;
  if (device < 0 || device >= sfk::kMaxTuneDevices) {
    sfk::CopyError("device index out of range", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  sfk_hip_builder* b = nullptr;
  const int status = sfk_hip_builder_create(device, &b, error, error_capacity);
  if (status != SFK_HIP_OK) return status;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int result = SFK_HIP_OK;
  try {
    std::size_t free_bytes = 0, total_bytes = 0;
    SFK_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    const std::uint64_t spare = free_bytes > (4ull << 30) ? free_bytes - (4ull << 30) : 0;
    max_elements = std::min<std::uint64_t>(max_elements, spare / 40);
    max_elements = std::max<std::uint64_t>(max_elements, sfk::kMinOnesweepTile);
    const std::uint64_t top = std::min<std::uint64_t>(max_elements, 1ull << 29);
    b->keys0.Reserve(top * 8);
    b->keys1.Reserve(top * 8);
    b->vals0.Reserve(top * 8);
    b->vals1.Reserve(top * 8);
    b->ref_offsets.Reserve(2 * sizeof(std::uint64_t));
    b->ref_adj.Reserve(sfk::RadixPasses::kMax * sfk::RadixTile::kRadix * sizeof(long long));
    b->radix_counts.Reserve(2 * sfk::kRadixSetStride);
    SFK_HIP_CHECK(hipMemsetAsync(b->radix_counts.As<void>(), 0, b->radix_counts.Bytes(), b->stream));
    b->radix_set = 0;
    sfk::ReserveOnesweep(b, top);
    sfk::ResetCounters(b);
    SFK_HIP_CHECK(hipEventCreate(&e0));
    SFK_HIP_CHECK(hipEventCreate(&e1));
    std::vector<sfk::TuneResult> results;
    sfk::TuneKind<std::uint32_t, std::uint32_t, false>(b, top, e0, e1, results);
    sfk::TuneKind<std::uint64_t, std::uint32_t, false>(b, top, e0, e1, results);
    sfk::TuneKind<std::uint32_t, std::uint64_t, false>(b, top, e0, e1, results);
    sfk::TuneKind<std::uint64_t, std::uint64_t, false>(b, top, e0, e1, results);
    sfk::TuneKind<std::uint32_t, std::uint32_t, true>(b, top, e0, e1, results);
    sfk::TuneKind<std::uint32_t, std::uint64_t, true>(b, top, e0, e1, results);
    const sfk::DeviceCounters& c = sfk::ReadCounters(b);
    if (c.error_flags != 0) sfk::Fail(SFK_HIP_BUILD_FAILURE, "radix autotuning failed (look-back did not converge)");
    // Fastest shape per (kind, class, digit width), with its time per pass;
    // classes above max_elements keep the table's current entries.
    constexpr int W = sfk::kDigitWidths;
    int best[sfk::kSortKinds][sfk::kSizeClasses][W];
    double best_us[sfk::kSortKinds][sfk::kSizeClasses][W];
    for (int k = 0; k < sfk::kSortKinds; ++k)
      for (int cl = 0; cl < sfk::kSizeClasses; ++cl)
        for (int w = 0; w < W; ++w) best[k][cl][w] = -1, best_us[k][cl][w] = 0;
    for (const auto& r : results) {
      const int w = sfk::WidthOf(r.bits);
      if (best[r.kind][r.size_class][w] < 0 || r.us_per_pass < best_us[r.kind][r.size_class][w]) {
        best[r.kind][r.size_class][w] = r.shape;
        best_us[r.kind][r.size_class][w] = r.us_per_pass;
      }
    }
    sfk::InitShapeTable();
    for (int k = 0; k < sfk::kSortKinds; ++k)
      for (int cl = 0; cl < sfk::kSizeClasses; ++cl)
        for (int w = 0; w < W; ++w)
          if (best[k][cl][w] >= 0) {
            sfk::g_shape_table[device][k][cl][w].store(best[k][cl][w]);
            sfk::g_pass_us[device][k][cl][w].store(best_us[k][cl][w]);
          }
    std::uint32_t written = 0;
    for (const auto& r : results) {
      if (entries != nullptr && written < capacity) {
        TuneEntryOf(r.kind, r.size_class, r.shape, r.bits, r.us_per_pass,
                    best[r.kind][r.size_class][sfk::WidthOf(r.bits)] == r.shape, true, &entries[written]);
        ++written;
      }
    }
    if (count != nullptr) *count = static_cast<uint32_t>(results.size());
  } catch (const sfk::Failure& f) {
    sfk::CopyError(f.message, error, error_capacity);
    result = f.status;
  } catch (const std::bad_alloc&) {
    sfk::CopyError("host allocation failed", error, error_capacity);
    result = SFK_HIP_BUILD_FAILURE;
  }
  if (e0 != nullptr) (void)hipEventDestroy(e0);
  if (e1 != nullptr) (void)hipEventDestroy(e1);
  sfk_hip_builder_destroy(b);
  return result;
}

int sfk_hip_tune_get(int device, sfk_hip_tune_entry* entries, uint32_t capacity, uint32_t* count) {
  if (count != nullptr) *count = 0;
  if (device < 0 || device >= sfk::kMaxTuneDevices) return SFK_HIP_INVALID_INPUT;
  sfk::InitShapeTable();
  std::uint32_t written = 0;
  for (int k = 0; k < sfk::kSortKinds; ++k) {
    for (int cl = 0; cl < sfk::kSizeClasses; ++cl) {
      for (int w = 0; w < sfk::kDigitWidths; ++w) {
        if (k >= 4 && w > 0) continue;  // per-reference kinds: 8-bit digits only
        const int tuned = sfk::g_shape_table[device][k][cl][w].load();
        const std::uint64_t rep = 1ull << (17 + 2 * cl);
        const int shape = tuned >= 0 ? tuned : sfk::DefaultShape(k, rep);
        if (entries != nullptr && written < capacity) {
          TuneEntryOf(k, cl, shape, sfk::DigitBitsOf(w), sfk::g_pass_us[device][k][cl][w].load(), true, tuned >= 0,
                      &entries[written]);
        }
        ++written;
      }
    }
  }
  if (count != nullptr) *count = written;
  return SFK_HIP_OK;
}

int sfk_hip_tune_set(int device, const sfk_hip_tune_entry* entries, uint32_t count, char* error,
                     size_t error_capacity) {
  if (device < 0 || device >= sfk::kMaxTuneDevices || (entries == nullptr && count > 0)) {
    sfk::CopyError("device index out of range or no entries", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  constexpr int W = sfk::kDigitWidths;
  int shapes[sfk::kSortKinds][sfk::kSizeClasses][W];
  double us[sfk::kSortKinds][sfk::kSizeClasses][W];
  for (auto& k : shapes)
    for (auto& c : k)
      for (auto& w : c) w = -2;  // untouched
  for (uint32_t i = 0; i < count; ++i) {
    const sfk_hip_tune_entry& e = entries[i];
    if (e.chosen == 0) continue;
    int shape = -1;
    for (int j = 0; j < sfk::kShapeCount; ++j) {
      if (static_cast<int>(e.block) == sfk::kShapes[j].block && static_cast<int>(e.items) == sfk::kShapes[j].items) {
        shape = j;
      }
    }
    const std::uint32_t bits = e.digit_bits == 0 ? 8u : e.digit_bits;  // 0: a table saved before widths
    if (e.kind >= static_cast<uint32_t>(sfk::kSortKinds) || e.size_class >= static_cast<uint32_t>(sfk::kSizeClasses) ||
        shape < 0 || !sfk::KindShapeFits(static_cast<int>(e.kind), shape) || (bits != 8 && bits != 7) ||
        (bits == 7 && e.kind >= 4)) {
      sfk::CopyError("entry " + std::to_string(i) + ": no compiled radix shape " + std::to_string(e.block) + "x" +
                         std::to_string(e.items) + " with " + std::to_string(bits) + "-bit digits for kind " +
                         std::to_string(e.kind) + ", class " + std::to_string(e.size_class),
                     error, error_capacity);
      return SFK_HIP_INVALID_INPUT;
    }
    const int w = sfk::WidthOf(static_cast<int>(bits));
    shapes[e.kind][e.size_class][w] = shape;
    us[e.kind][e.size_class][w] = e.us_per_pass;
  }
  sfk::InitShapeTable();
  for (int k = 0; k < sfk::kSortKinds; ++k)
    for (int cl = 0; cl < sfk::kSizeClasses; ++cl)
      for (int w = 0; w < W; ++w)
        if (shapes[k][cl][w] != -2) {
          sfk::g_shape_table[device][k][cl][w].store(shapes[k][cl][w]);
          sfk::g_pass_us[device][k][cl][w].store(us[k][cl][w]);
        }
  return SFK_HIP_OK;
}

int sfk_hip_tune_reset(int device) {
  if (device < 0 || device >= sfk::kMaxTuneDevices) return SFK_HIP_INVALID_INPUT;
  sfk::InitShapeTable();
  for (auto& k : sfk::g_shape_table[device])
    for (auto& c : k)
      for (auto& w : c) w.store(-1);
  for (auto& k : sfk::g_pass_us[device])
    for (auto& c : k)
      for (auto& w : c) w.store(0.0);
  return SFK_HIP_OK;
}

// ------------------------------------------------ allocation-free build
//
// sfk_hip_build_device_noalloc: sfk_hip_build_device for callers that cannot
// allocate device memory.  Its workspace is `scratch`, a fixed array in
// device memory that every GPU the library runs on holds from module load
// until the process ends; each buffer of the build borrows a slice of it.
// Host code never uses the array's name as an address: the device address
// comes from hipGetSymbolAddress, once per device with that device current,
// and only that pointer is handed to kernels and copies.  Every call on a
// GPU shares that GPU's array (and one host word block for the step
// counters), so calls that overlap on one GPU race; the call clears what a
// build needs cleared (counters, radix digit counts, look-back status and
// tickets) before it starts.

static __device__ __align__(256) unsigned char scratch[SFK_HIP_NOALLOC_SCRATCH_BYTES];

namespace sfk {
namespace {

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxTuneDevices] = {};
DeviceCounters g_noalloc_counters[kMaxTuneDevices];

// The device address of `scratch` on `device` (which must be current).
void* ScratchAddress(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      Fail(SFK_HIP_BUILD_FAILURE, std::string("address of the allocation-free workspace: ") + hipGetErrorString(err));
    }
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

// Slices of the workspace for a 32-bit build of n symbols, in the sizes
// BuildTyped and Doubling reserve (the small lists sized for every suffix).
// Returns the bytes used; borrows the slices into `b` when base is given.
std::uint64_t NoAllocLayout(sfk_hip_builder* b, std::uint64_t n, bool need_rank, char* base) {
  constexpr std::uint64_t w = sizeof(std::uint32_t);
  std::uint64_t used = 0;
  auto take = [&](DeviceBuffer& buf, std::uint64_t bytes) {
    bytes = std::max<std::uint64_t>(bytes, 1);
    if (base != nullptr) buf.Borrow(base + used, bytes);
    used += (bytes + 255) / 256 * 256;
  };
  const std::uint64_t seg_cap = n / 2 + 2;
  take(b->counters, sizeof(DeviceCounters));
  take(b->packed, PackedWords(n) * 16);
  take(b->special_blocks, SpecialBlockWords(n) * sizeof(std::uint32_t));
  if (need_rank) take(b->rank, n * w);
  take(b->act_idx, ActIdxBytes(n, w));
  take(b->act_seg, n * sizeof(std::uint32_t));
  take(b->keys0, std::max<std::uint64_t>(n * 8, 64));
  take(b->keys1, std::max<std::uint64_t>(n * 8, 64));
  take(b->vals0, n * w);
  take(b->vals1, n * w);
  take(b->seg_delta0, seg_cap * w);
  take(b->seg_delta1, seg_cap * w);
  take(b->seg_start, seg_cap * w);
  take(b->radix_counts, 2 * kRadixSetStride);
  take(b->radix_status, OnesweepStatusBytes(n));
  take(b->tickets, 2 * sizeof(unsigned int));
  take(b->partial, std::max(kRefineMaxChunks * sizeof(RefinePartial), 1024 * sizeof(SplitPartial)));
  for (DeviceBuffer* q : {&b->q_idx0, &b->q_idx1, &b->q_row0_0, &b->q_row0_1, &b->q_key}) take(*q, n * w);
  return used;
}

}  // namespace
}  // namespace sfk

int sfk_hip_build_device_noalloc(int device, const uint8_t* d_text, uint64_t n, uint32_t construction_bits,
                                 uint32_t sampling_rate, uint32_t flags, void* d_sa, uint32_t sa_bits, void* d_isa,
                                 uint32_t isa_bits, void* d_lcp, uint32_t lcp_bits, void* stream,
                                 sfk_hip_stats* stats, char* error, size_t error_capacity) {
  std::string message;
  if (d_text == nullptr || d_sa == nullptr || device < 0 || device >= sfk::kMaxTuneDevices) {
    sfk::CopyError("device, text and SA buffers are required", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  if (construction_bits != 32 || sampling_rate != 1 || sa_bits != 32 || (d_isa != nullptr && isa_bits != 32) ||
      (d_lcp != nullptr && (flags & SFK_HIP_RETAIN_LCP) && lcp_bits != 32) || (flags & SFK_HIP_PROFILE_KERNELS)) {
    sfk::CopyError("the allocation-free build takes 32-bit construction, sampling rate 1, 32-bit outputs and no "
                   "kernel profiling",
                   error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  if (n > SFK_HIP_NOALLOC_MAX_SYMBOLS) {
    sfk::CopyError("the allocation-free build takes at most " + std::to_string(SFK_HIP_NOALLOC_MAX_SYMBOLS) +
                       " symbols (its workspace is a fixed " + std::to_string(SFK_HIP_NOALLOC_SCRATCH_BYTES) +
                       "-byte array); " + std::to_string(n) + " given",
                   error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  int status = sfk::ValidateArgs(n, construction_bits, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits, flags, message);
  if (status != SFK_HIP_OK) {
    sfk::CopyError(message, error, error_capacity);
    return status;
  }
  sfk_hip_builder b;
  b.device = device;
  b.stream = static_cast<hipStream_t>(stream);
  b.borrowed_workspace = true;
  b.host_counters = &sfk::g_noalloc_counters[device];
  try {
    const auto t0 = std::chrono::steady_clock::now();
    SFK_HIP_CHECK(hipSetDevice(device));
    char* base = static_cast<char*>(sfk::ScratchAddress(device));
    const bool need_rank = d_isa == nullptr;
    if (sfk::NoAllocLayout(&b, n, need_rank, nullptr) > SFK_HIP_NOALLOC_SCRATCH_BYTES) {
      sfk::Fail(SFK_HIP_BUILD_FAILURE, "allocation-free workspace layout exceeds the fixed array");
    }
    sfk::NoAllocLayout(&b, n, need_rank, base);
    // The look-back status words and tile tickets of the last call are
    // stale for this one's epochs.
    SFK_HIP_CHECK(hipMemsetAsync(b.radix_status.As<void>(), 0, b.radix_status.Bytes(), b.stream));
    SFK_HIP_CHECK(hipMemsetAsync(b.tickets.As<void>(), 0, b.tickets.Bytes(), b.stream));
    b.sort_epoch = 0;
    status = sfk::RunDeviceBuild(&b, d_text, n, construction_bits, flags, d_sa, sa_bits, d_isa, isa_bits, d_lcp,
                                 lcp_bits, stats, message);
    if (stats != nullptr) {
      stats->call_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  } catch (const sfk::Failure& f) {
    (void)hipStreamSynchronize(b.stream);
    sfk::CopyError(f.message, error, error_capacity);
    status = f.status;
  } catch (const std::bad_alloc&) {
    sfk::CopyError("host allocation failed", error, error_capacity);
    status = SFK_HIP_BUILD_FAILURE;
  }
  return status;
}

uint64_t sfk_hip_noalloc_workspace_bytes(uint64_t n, int with_isa) {
  sfk_hip_builder b;
  return sfk::NoAllocLayout(&b, n, with_isa == 0, nullptr);
}
