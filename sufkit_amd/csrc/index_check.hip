// Full-size checks of a built suffix index on the device (gfx950).
//
//   sfk_hip_digest       position-keyed sum digest of an SA / ISA / LCP array,
//                        n*G + sum_r fmix64(v[r] + G*(r+1)) mod 2^64, the
//                        number the full-size golden digests hold (computed
//                        by a reduction; the terms commute)
//   sfk_hip_check_index  proves SA, ISA and LCP of a text without a second
//                        construction:
//     - range and inverse: every SA value < n and ISA[SA[r]] == r, so SA is a
//       permutation and ISA its inverse (the checks of RepackSuffixArray,
//       suffix_array.cpp:852-881, and of Load, :4999-5010,5120-5131);
//     - order (Burkhardt and Kaerkkaeinen, CPM 2003): SA[0] == n-1 and for
//       every adjacent pair a = SA[r-1], b = SA[r]: T[a] < T[b], or T[a] ==
//       T[b] and ISA[a+1] < ISA[b+1].  With the permutation this makes SA the
//       suffix array of T (byte order, the unique sentinel last);
//     - LCP: LCP[0] == 0 and for r >= 1 with l = LCP[r]: T[a+l] != T[b+l]
//       (so lcp(a, b) <= l; a value too small fails here) and the first l
//       symbols of a and b hash equal (a value too large fails here) under
//       a Karp-Rabin hash modulo 2^61-1 with a seeded random base.  A wrong
//       LCP value passes with probability at most n / 2^61 per row.
//   The checks read every value they index with before using it, so a
//   corrupted array gives failure counts, never an access outside the
//   text or the arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sufkit_hip.h"

namespace sfk {
namespace check {
namespace {

constexpr std::uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr std::uint64_t kPrime = (1ull << 61) - 1;  // Mersenne prime 2^61 - 1
constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr std::uint64_t kChunk = static_cast<std::uint64_t>(kThreads) * kItems;
constexpr int kSpineThreads = 1024;
constexpr unsigned kMaxGrid = 1u << 16;
constexpr int kPowBits = 16;
constexpr std::uint64_t kPowSize = 1ull << kPowBits;

__host__ __device__ inline std::uint64_t Fmix64(std::uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// a * b mod 2^61-1 for a, b < 2^61-1: the 122-bit product hi*2^64 + lo is
// congruent to 8*hi + (lo >> 61) + (lo & p), because 2^61 == 1 (mod p).
__device__ __forceinline__ std::uint64_t MulMod(std::uint64_t a, std::uint64_t b) {
  const std::uint64_t lo = a * b;
  const std::uint64_t hi = __umul64hi(a, b);
  std::uint64_t r = (lo & kPrime) + (lo >> 61) + (hi << 3);
  r = (r & kPrime) + (r >> 61);
  return r >= kPrime ? r - kPrime : r;
}

__device__ __forceinline__ std::uint64_t AddMod(std::uint64_t a, std::uint64_t b) {
  const std::uint64_t r = a + b;
  return r >= kPrime ? r - kPrime : r;
}

__device__ __forceinline__ std::uint64_t SubMod(std::uint64_t a, std::uint64_t b) {
  return a >= b ? a - b : a + kPrime - b;
}

// Powers of the hash base: lo[k] = B^k, mid[k] = B^(k * 2^16),
// hi[k] = B^(k * 2^32), k < 2^16, so B^e for e < 2^48 is three lookups.
struct Powers {
  const std::uint64_t* lo;
  const std::uint64_t* mid;
  const std::uint64_t* hi;
};

__device__ __forceinline__ std::uint64_t Pow(const Powers& pw, std::uint64_t e) {
  std::uint64_t r = pw.lo[e & (kPowSize - 1)];
  const std::uint64_t m = (e >> kPowBits) & (kPowSize - 1);
  const std::uint64_t h = e >> (2 * kPowBits);
  if (m != 0) r = MulMod(r, pw.mid[m]);
  if (h != 0) r = MulMod(r, pw.hi[h]);
  return r;
}

// An affine map x -> x * m + a (mod p); Then(f, g) applies f, then g.
struct Affine {
  std::uint64_t m;
  std::uint64_t a;
};

__device__ __forceinline__ Affine Then(const Affine& f, const Affine& g) {
  return Affine{MulMod(f.m, g.m), AddMod(MulMod(f.a, g.m), g.a)};
}

// Exclusive scan of one affine map per thread over a block of kBlock
// threads (Hillis-Steele in LDS); returns the composition of the maps of
// the threads before this one (identity for thread 0).
template <int kBlock>
__device__ Affine BlockExclusiveCompose(Affine mine, Affine* lds) {
  const int t = static_cast<int>(threadIdx.x);
  lds[t] = mine;
  __syncthreads();
  for (int d = 1; d < kBlock; d <<= 1) {
    Affine cur = lds[t];
    if (t >= d) cur = Then(lds[t - d], cur);
    __syncthreads();
    lds[t] = cur;
    __syncthreads();
  }
  const Affine out = t == 0 ? Affine{1, 0} : lds[t - 1];
  __syncthreads();
  return out;
}

// Symbol weight in the hash (never 0).
__device__ __forceinline__ std::uint64_t Weight(std::uint8_t s) { return static_cast<std::uint64_t>(s) + 1; }

// Thread t's items of chunk c: symbols [c*kChunk + t*kItems, +kItems) below n.
// Returns the number of valid items and their hash (Horner).
__device__ __forceinline__ int ThreadHash(const std::uint8_t* __restrict__ text, std::uint64_t n,
                                          std::uint64_t first, std::uint64_t base, std::uint64_t* hash,
                                          std::uint8_t (&sym)[kItems]) {
  int count = 0;
  std::uint64_t h = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const std::uint64_t i = first + k;
    sym[k] = i < n ? text[i] : 0;
  }
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (first + k < n) {
      h = AddMod(MulMod(h, base), Weight(sym[k]));
      ++count;
    }
  }
  *hash = h;
  return count;
}

// Hash of every chunk of kChunk symbols (the last may be shorter).
__global__ __launch_bounds__(kThreads) void HashChunkKernel(const std::uint8_t* __restrict__ text, std::uint64_t n,
                                                            std::uint64_t chunks, std::uint64_t base,
                                                            const std::uint64_t* __restrict__ pow_lo,
                                                            std::uint64_t* __restrict__ chunk_hash) {
  __shared__ std::uint64_t sums[kThreads];
  for (std::uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
    const std::uint64_t first = c * kChunk + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    std::uint8_t sym[kItems];
    std::uint64_t h = 0;
    ThreadHash(text, n, first, base, &h, sym);
    // Symbols of this chunk after this thread's items.
    const std::uint64_t chunk_end = std::min<std::uint64_t>(n, (c + 1) * kChunk);
    const std::uint64_t after = chunk_end > first + kItems ? chunk_end - first - kItems : 0;
    sums[threadIdx.x] = MulMod(h, pow_lo[after]);
    __syncthreads();
    for (int d = kThreads / 2; d > 0; d >>= 1) {
      if (static_cast<int>(threadIdx.x) < d) sums[threadIdx.x] = AddMod(sums[threadIdx.x], sums[threadIdx.x + d]);
      __syncthreads();
    }
    if (threadIdx.x == 0) chunk_hash[c] = sums[0];
    __syncthreads();
  }
}

// prefix[c] = hash of all symbols before chunk c: prefix[c+1] = prefix[c] *
// B^kChunk + chunk_hash[c] (every chunk but the last is full).
__global__ __launch_bounds__(kSpineThreads) void HashSpineKernel(const std::uint64_t* __restrict__ chunk_hash,
                                                                 std::uint64_t chunks, std::uint64_t chunk_pow,
                                                                 std::uint64_t* __restrict__ prefix) {
  __shared__ Affine lds[kSpineThreads];
  const std::uint64_t per = (chunks + kSpineThreads - 1) / kSpineThreads;
  const std::uint64_t begin = std::min<std::uint64_t>(chunks, threadIdx.x * per);
  const std::uint64_t end = std::min<std::uint64_t>(chunks, begin + per);
  Affine mine{1, 0};
  for (std::uint64_t c = begin; c < end; ++c) mine = Then(mine, Affine{chunk_pow, chunk_hash[c]});
  const Affine before = BlockExclusiveCompose<kSpineThreads>(mine, lds);
  std::uint64_t x = before.a;  // the map applied to the empty prefix (0)
  for (std::uint64_t c = begin; c < end; ++c) {
    prefix[c] = x;
    x = AddMod(MulMod(x, chunk_pow), chunk_hash[c]);
  }
}

// H[i + 1] = hash of T[0..i]; H[0] = 0.
__global__ __launch_bounds__(kThreads) void HashApplyKernel(const std::uint8_t* __restrict__ text, std::uint64_t n,
                                                            std::uint64_t chunks, std::uint64_t base,
                                                            const std::uint64_t* __restrict__ pow_lo,
                                                            const std::uint64_t* __restrict__ prefix,
                                                            std::uint64_t* __restrict__ hashes) {
  __shared__ Affine lds[kThreads];
  if (blockIdx.x == 0 && threadIdx.x == 0) hashes[0] = 0;
  for (std::uint64_t c = blockIdx.x; c < chunks; c += gridDim.x) {
    const std::uint64_t first = c * kChunk + static_cast<std::uint64_t>(threadIdx.x) * kItems;
    std::uint8_t sym[kItems];
    std::uint64_t h = 0;
    const int count = ThreadHash(text, n, first, base, &h, sym);
    const Affine before = BlockExclusiveCompose<kThreads>(Affine{pow_lo[count], h}, lds);
    std::uint64_t x = AddMod(MulMod(prefix[c], before.m), before.a);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      if (first + k < n) {
        x = AddMod(MulMod(x, base), Weight(sym[k]));
        hashes[first + k + 1] = x;
      }
    }
  }
}

// Failure categories counted by CheckRowsKernel.
enum Check : int { kRange = 0, kInverse = 1, kOrder = 2, kLcp = 3, kLcpPrefix = 4, kChecks = 5 };

struct Result {
  unsigned long long failures[kChecks];
  unsigned long long first_failed_row;
};

__device__ __forceinline__ void Count(Result* res, int check, std::uint64_t r) {
  atomicAdd(&res->failures[check], 1ull);
  atomicMin(&res->first_failed_row, static_cast<unsigned long long>(r));
}

template <typename S, typename I, typename L>
__global__ __launch_bounds__(kThreads) void CheckRowsKernel(const std::uint8_t* __restrict__ text, std::uint64_t n,
                                                            const S* __restrict__ sa, const I* __restrict__ isa,
                                                            const L* __restrict__ lcp,
                                                            const std::uint64_t* __restrict__ hashes, Powers pw,
                                                            Result* res) {
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kThreads;
  for (std::uint64_t r = static_cast<std::uint64_t>(blockIdx.x) * kThreads + threadIdx.x; r < n; r += stride) {
    const std::uint64_t b = static_cast<std::uint64_t>(sa[r]);
    const std::uint64_t a = r > 0 ? static_cast<std::uint64_t>(sa[r - 1]) : 0;
    if (b >= n || a >= n) {
      Count(res, kRange, r);
      continue;
    }
    const std::uint64_t inv = static_cast<std::uint64_t>(isa[b]);
    if (inv >= n) {
      Count(res, kRange, r);
    } else if (inv != r) {
      Count(res, kInverse, r);
    }
    if (r == 0) {
      if (b != n - 1) Count(res, kOrder, r);
      if (lcp != nullptr && lcp[0] != 0) Count(res, kLcp, r);
      continue;
    }
    const std::uint8_t ta = text[a], tb = text[b];
    bool ordered = ta < tb;
    if (ta == tb) {
      ordered = a + 1 < n && b + 1 < n &&
                static_cast<std::uint64_t>(isa[a + 1]) < static_cast<std::uint64_t>(isa[b + 1]);
    }
    if (!ordered) Count(res, kOrder, r);
    if (lcp == nullptr) continue;
    const std::uint64_t l = static_cast<std::uint64_t>(lcp[r]);
    const std::uint64_t room = n - (a > b ? a : b);  // symbols left in the later suffix
    if (l >= room || text[a + l] == text[b + l]) {
      Count(res, kLcp, r);
      continue;
    }
    if (l > 0) {
      const std::uint64_t pl = Pow(pw, l);
      const std::uint64_t ha = SubMod(hashes[a + l], MulMod(hashes[a], pl));
      const std::uint64_t hb = SubMod(hashes[b + l], MulMod(hashes[b], pl));
      if (ha != hb) Count(res, kLcpPrefix, r);
    }
  }
}

// Sum digest of an array (value_bits 8 / 32 / 64), without the n*G term.
template <typename T>
__global__ __launch_bounds__(kThreads) void DigestKernel(const T* __restrict__ values, std::uint64_t count,
                                                         unsigned long long* out) {
  __shared__ std::uint64_t sums[kThreads];
  const std::uint64_t stride = static_cast<std::uint64_t>(gridDim.x) * kThreads;
  std::uint64_t s = 0;
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < count; i += stride) {
    s += Fmix64(static_cast<std::uint64_t>(values[i]) + kGolden * (i + 1));
  }
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int d = kThreads / 2; d > 0; d >>= 1) {
    if (static_cast<int>(threadIdx.x) < d) sums[threadIdx.x] += sums[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, static_cast<unsigned long long>(sums[0]));
}

unsigned GridFor(std::uint64_t items, std::uint64_t per_block) {
  const std::uint64_t blocks = std::max<std::uint64_t>(1, (items + per_block - 1) / per_block);
  return static_cast<unsigned>(std::min<std::uint64_t>(blocks, kMaxGrid));
}

struct Status {
  int code = SFK_HIP_OK;
  std::string message;
};

bool Ok(hipError_t err, const char* what, Status& st) {
  if (err == hipSuccess) return true;
  (void)hipGetLastError();
  st.code = SFK_HIP_BUILD_FAILURE;
  st.message = std::string(what) + ": " + hipGetErrorString(err);
  return false;
}

void CopyError(const Status& st, char* error, size_t capacity) {
  if (error == nullptr || capacity == 0) return;
  const std::size_t len = std::min(st.message.size(), capacity - 1);
  std::memcpy(error, st.message.data(), len);
  error[len] = '\0';
}

// Device buffers of one check call, released on every path.
struct Scratch {
  std::vector<void*> held;
  ~Scratch() {
    for (void* p : held) (void)hipFree(p);
  }
  template <typename T>
  T* Alloc(std::uint64_t count, Status& st) {
    void* p = nullptr;
    if (!Ok(hipMalloc(&p, std::max<std::uint64_t>(count, 1) * sizeof(T)), "device allocation", st)) return nullptr;
    held.push_back(p);
    return static_cast<T*>(p);
  }
};

std::uint64_t HostMulMod(std::uint64_t a, std::uint64_t b) {
  const unsigned __int128 p = static_cast<unsigned __int128>(a) * b;
  std::uint64_t r = static_cast<std::uint64_t>(p & kPrime) + static_cast<std::uint64_t>(p >> 61);
  r = (r & kPrime) + (r >> 61);
  return r >= kPrime ? r - kPrime : r;
}

template <typename S, typename I, typename L>
void LaunchRows(const std::uint8_t* text, std::uint64_t n, const void* sa, const void* isa, const void* lcp,
                const std::uint64_t* hashes, Powers pw, Result* res, hipStream_t stream) {
  CheckRowsKernel<S, I, L><<<GridFor(n, kThreads * 4), kThreads, 0, stream>>>(
      text, n, static_cast<const S*>(sa), static_cast<const I*>(isa), static_cast<const L*>(lcp), hashes, pw, res);
}

template <typename S, typename I>
void LaunchRowsLcp(std::uint32_t lcp_bits, const std::uint8_t* text, std::uint64_t n, const void* sa, const void* isa,
                   const void* lcp, const std::uint64_t* hashes, Powers pw, Result* res, hipStream_t stream) {
  if (lcp_bits == 64) {
    LaunchRows<S, I, std::uint64_t>(text, n, sa, isa, lcp, hashes, pw, res, stream);
  } else {
    LaunchRows<S, I, std::uint32_t>(text, n, sa, isa, lcp, hashes, pw, res, stream);
  }
}

template <typename S>
void LaunchRowsIsa(std::uint32_t isa_bits, std::uint32_t lcp_bits, const std::uint8_t* text, std::uint64_t n,
                   const void* sa, const void* isa, const void* lcp, const std::uint64_t* hashes, Powers pw,
                   Result* res, hipStream_t stream) {
  if (isa_bits == 64) {
    LaunchRowsLcp<S, std::uint64_t>(lcp_bits, text, n, sa, isa, lcp, hashes, pw, res, stream);
  } else {
    LaunchRowsLcp<S, std::uint32_t>(lcp_bits, text, n, sa, isa, lcp, hashes, pw, res, stream);
  }
}

Status RunCheck(int device, const std::uint8_t* d_text, std::uint64_t n, const void* d_sa, std::uint32_t sa_bits,
                const void* d_isa, std::uint32_t isa_bits, const void* d_lcp, std::uint32_t lcp_bits,
                std::uint64_t seed, sfk_hip_check_report* report) {
  Status st;
  const auto t0 = std::chrono::steady_clock::now();
  if (d_text == nullptr || d_sa == nullptr || d_isa == nullptr || n == 0 || report == nullptr) {
    return Status{SFK_HIP_INVALID_INPUT, "text, SA, ISA and report are required and n must be positive"};
  }
  if ((sa_bits != 32 && sa_bits != 64) || (isa_bits != 32 && isa_bits != 64) ||
      (d_lcp != nullptr && lcp_bits != 32 && lcp_bits != 64)) {
    return Status{SFK_HIP_INVALID_INPUT, "SA, ISA and LCP widths must be 32 or 64 bits"};
  }
  if (n > SFK_HIP_MAX_TEXT_SYMBOLS) {
    return Status{SFK_HIP_UNSUPPORTED_BACKEND, "texts of more than " + std::to_string(SFK_HIP_MAX_TEXT_SYMBOLS) +
                                                   " symbols are not supported by the HIP backend"};
  }
  if (!Ok(hipSetDevice(device), "hipSetDevice", st)) return st;
  hipStream_t stream = nullptr;
  if (!Ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate", st)) return st;
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{stream};
  Scratch scratch;
  // Hash base from the seed, in [2, p - 2]; powers on the host.
  const std::uint64_t base = 2 + Fmix64(seed ^ kGolden) % (kPrime - 3);
  std::vector<std::uint64_t> powers(3 * kPowSize);
  std::uint64_t* lo = powers.data();
  std::uint64_t* mid = lo + kPowSize;
  std::uint64_t* hi = mid + kPowSize;
  lo[0] = 1;
  for (std::uint64_t k = 1; k < kPowSize; ++k) lo[k] = HostMulMod(lo[k - 1], base);
  const std::uint64_t step_mid = HostMulMod(lo[kPowSize - 1], base);  // B^(2^16)
  mid[0] = 1;
  for (std::uint64_t k = 1; k < kPowSize; ++k) mid[k] = HostMulMod(mid[k - 1], step_mid);
  const std::uint64_t step_hi = HostMulMod(mid[kPowSize - 1], step_mid);  // B^(2^32)
  hi[0] = 1;
  for (std::uint64_t k = 1; k < kPowSize; ++k) hi[k] = HostMulMod(hi[k - 1], step_hi);
  const std::uint64_t chunk_pow = lo[kChunk];

  const std::uint64_t chunks = (n + kChunk - 1) / kChunk;
  std::uint64_t* d_pow = scratch.Alloc<std::uint64_t>(powers.size(), st);
  std::uint64_t* d_chunk = d_pow ? scratch.Alloc<std::uint64_t>(chunks, st) : nullptr;
  std::uint64_t* d_prefix = d_chunk ? scratch.Alloc<std::uint64_t>(chunks, st) : nullptr;
  std::uint64_t* d_hashes = d_prefix && d_lcp ? scratch.Alloc<std::uint64_t>(n + 1, st) : nullptr;
  Result* d_res = (d_prefix && (d_hashes || !d_lcp)) ? scratch.Alloc<Result>(1, st) : nullptr;
  if (d_res == nullptr) return st;
  Result init{};
  init.first_failed_row = n;
  if (!Ok(hipMemcpyAsync(d_pow, powers.data(), powers.size() * 8, hipMemcpyHostToDevice, stream), "upload", st) ||
      !Ok(hipMemcpyAsync(d_res, &init, sizeof(init), hipMemcpyHostToDevice, stream), "upload", st)) {
    return st;
  }
  const Powers pw{d_pow, d_pow + kPowSize, d_pow + 2 * kPowSize};
  if (d_lcp != nullptr) {
    HashChunkKernel<<<GridFor(chunks, 1), kThreads, 0, stream>>>(d_text, n, chunks, base, d_pow, d_chunk);
    HashSpineKernel<<<1, kSpineThreads, 0, stream>>>(d_chunk, chunks, chunk_pow, d_prefix);
    HashApplyKernel<<<GridFor(chunks, 1), kThreads, 0, stream>>>(d_text, n, chunks, base, d_pow, d_prefix,
                                                                  d_hashes);
    if (!Ok(hipGetLastError(), "hash launch", st)) return st;
  }
  if (sa_bits == 64) {
    LaunchRowsIsa<std::uint64_t>(isa_bits, lcp_bits, d_text, n, d_sa, d_isa, d_lcp, d_hashes, pw, d_res, stream);
  } else {
    LaunchRowsIsa<std::uint32_t>(isa_bits, lcp_bits, d_text, n, d_sa, d_isa, d_lcp, d_hashes, pw, d_res, stream);
  }
  if (!Ok(hipGetLastError(), "check launch", st)) return st;
  Result res{};
  if (!Ok(hipMemcpyAsync(&res, d_res, sizeof(res), hipMemcpyDeviceToHost, stream), "readback", st) ||
      !Ok(hipStreamSynchronize(stream), "check", st)) {
    return st;
  }
  report->rows = n;
  report->range_failures = res.failures[kRange];
  report->inverse_failures = res.failures[kInverse];
  report->order_failures = res.failures[kOrder];
  report->lcp_failures = res.failures[kLcp];
  report->lcp_prefix_failures = res.failures[kLcpPrefix];
  report->first_failed_row = res.first_failed_row;
  report->lcp_checked = d_lcp != nullptr ? 1u : 0u;
  report->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::uint64_t total = 0;
  for (auto f : res.failures) total += f;
  if (total != 0) {
    st.code = SFK_HIP_BUILD_FAILURE;
    st.message = "suffix index check failed at row " + std::to_string(res.first_failed_row) + ": " +
                 std::to_string(res.failures[kRange]) + " out of range, " +
                 std::to_string(res.failures[kInverse]) + " not inverse, " +
                 std::to_string(res.failures[kOrder]) + " out of order, " +
                 std::to_string(res.failures[kLcp] + res.failures[kLcpPrefix]) + " wrong LCP";
  }
  return st;
}

}  // namespace
}  // namespace check
}  // namespace sfk

extern "C" {

int sfk_hip_digest(int device, const void* d_values, uint64_t count, uint32_t value_bits, uint64_t* out,
                   char* error, size_t error_capacity) {
  using namespace sfk::check;
  Status st;
  if (out == nullptr || (count > 0 && d_values == nullptr) ||
      (value_bits != 8 && value_bits != 32 && value_bits != 64)) {
    st = Status{SFK_HIP_INVALID_INPUT, "digest needs an output and values of 8, 32 or 64 bits"};
    CopyError(st, error, error_capacity);
    return st.code;
  }
  unsigned long long* d_sum = nullptr;
  hipStream_t stream = nullptr;
  unsigned long long sum = 0;
  if (Ok(hipSetDevice(device), "hipSetDevice", st) &&
      Ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate", st) &&
      Ok(hipMalloc(&d_sum, sizeof(unsigned long long)), "device allocation", st) &&
      Ok(hipMemcpyAsync(d_sum, &sum, sizeof(sum), hipMemcpyHostToDevice, stream), "upload", st)) {
    if (count > 0) {
      const unsigned grid = GridFor(count, kThreads * 8);
      if (value_bits == 8) {
        DigestKernel<std::uint8_t><<<grid, kThreads, 0, stream>>>(static_cast<const std::uint8_t*>(d_values), count,
                                                                  d_sum);
      } else if (value_bits == 32) {
        DigestKernel<std::uint32_t><<<grid, kThreads, 0, stream>>>(static_cast<const std::uint32_t*>(d_values),
                                                                   count, d_sum);
      } else {
        DigestKernel<std::uint64_t><<<grid, kThreads, 0, stream>>>(static_cast<const std::uint64_t*>(d_values),
                                                                   count, d_sum);
      }
    }
    if (Ok(hipGetLastError(), "digest launch", st) &&
        Ok(hipMemcpyAsync(&sum, d_sum, sizeof(sum), hipMemcpyDeviceToHost, stream), "readback", st) &&
        Ok(hipStreamSynchronize(stream), "digest", st)) {
      *out = count * kGolden + sum;
    }
  }
  if (stream != nullptr) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (d_sum != nullptr) (void)hipFree(d_sum);
  CopyError(st, error, error_capacity);
  return st.code;
}

int sfk_hip_check_index(int device, const uint8_t* d_text, uint64_t n, const void* d_sa, uint32_t sa_bits,
                        const void* d_isa, uint32_t isa_bits, const void* d_lcp, uint32_t lcp_bits,
                        uint64_t hash_seed, sfk_hip_check_report* report, char* error, size_t error_capacity) {
  if (report != nullptr) std::memset(report, 0, sizeof(*report));
  const sfk::check::Status st =
      sfk::check::RunCheck(device, d_text, n, d_sa, sa_bits, d_isa, isa_bits, d_lcp, lcp_bits, hash_seed, report);
  sfk::check::CopyError(st, error, error_capacity);
  return st.code;
}

}  // extern "C"
