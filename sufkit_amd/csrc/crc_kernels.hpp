// CRC-32 of device buffers, equal to zlib's crc32() (IEEE 802.3 polynomial,
// reflected, initial value and final xor 0xFFFFFFFF), the checksum every
// .sufidx section and header carries (serialization.cpp:47-141).
//
// Polynomials mod P are held reflected: bit 31 is the coefficient of x^0,
// bit 0 that of x^31.  For a message M with register start 0 and no final
// xor, raw(M) = M(x) * x^32 mod P, so for two pieces
//     raw(A ++ B) = raw(A) * x^(8|B|) + raw(B)          (mod P, over GF(2))
// and zlib's value continuing from crc_in is
//     crc32(crc_in, M) = ~(~crc_in * x^(8|M|) + raw(M)).
// That lets every thread checksum its own piece of the buffer and the pieces
// be joined in a tree: the device computes raw() of 256-byte pieces with
// byte tables in LDS (slicing by four words), joins 256 of them per
// workgroup (a 64 KiB block), and one workgroup joins the blocks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sfk {

constexpr std::uint32_t kCrcPoly = 0xEDB88320u;  // reflected x^32 + x^26 + ... + 1
constexpr std::uint32_t kCrcOne = 0x80000000u;   // the polynomial 1
constexpr int kCrcThreads = 256;
constexpr int kCrcPieceBytes = 256;                                   // per thread
constexpr std::uint64_t kCrcBlockBytes = std::uint64_t{kCrcThreads} * kCrcPieceBytes;  // 64 KiB
constexpr int kCrcBlockShift = 16;

// x^(8 * 2^k) mod P for k = 0..63, passed by value to the kernels.
struct CrcPowers {
  std::uint32_t x8pow2[64];
};

// a * b mod P.
__host__ __device__ inline std::uint32_t CrcMulMod(std::uint32_t a, std::uint32_t b) {
  std::uint32_t p = 0;
  for (int i = 0; i < 32; ++i) {
    if (a & (kCrcOne >> i)) p ^= b;             // a has x^i: add b * x^i
    b = (b & 1u) ? (b >> 1) ^ kCrcPoly : b >> 1;  // b *= x
  }
  return p;
}

inline CrcPowers MakeCrcPowers() {
  CrcPowers t{};
  t.x8pow2[0] = kCrcOne >> 8;  // x^8
  for (int k = 1; k < 64; ++k) t.x8pow2[k] = CrcMulMod(t.x8pow2[k - 1], t.x8pow2[k - 1]);
  return t;
}

// x^(8 * bytes) mod P.
__host__ __device__ inline std::uint32_t CrcX8Pow(const CrcPowers& t, std::uint64_t bytes) {
  std::uint32_t r = kCrcOne;
  for (int k = 0; bytes != 0; ++k, bytes >>= 1) {
    if (bytes & 1u) r = CrcMulMod(r, t.x8pow2[k]);
  }
  return r;
}

// raw(A ++ B) from raw(A), raw(B) and |B|.
__host__ __device__ inline std::uint32_t CrcJoin(const CrcPowers& t, std::uint32_t a, std::uint32_t b,
                                                 std::uint64_t b_bytes) {
  return CrcMulMod(a, CrcX8Pow(t, b_bytes)) ^ b;
}

// zlib's crc32(crc_in, M) from raw(M) and |M|.
__host__ __device__ inline std::uint32_t CrcFinish(const CrcPowers& t, std::uint32_t crc_in, std::uint32_t raw,
                                                   std::uint64_t bytes) {
  return ~(CrcMulMod(~crc_in, CrcX8Pow(t, bytes)) ^ raw);
}

// Byte tables for slicing by four: tab[0][i] = raw of the byte i, and
// tab[k][i] = raw of byte i followed by k zero bytes.  Built by the
// workgroup in LDS (256 threads, one entry per thread and table).
__device__ inline void CrcBuildTables(std::uint32_t (*tab)[256]) {
  const int i = threadIdx.x;
  std::uint32_t c = static_cast<std::uint32_t>(i);
  for (int b = 0; b < 8; ++b) c = (c & 1u) ? (c >> 1) ^ kCrcPoly : c >> 1;
  tab[0][i] = c;
  __syncthreads();
  for (int k = 1; k < 4; ++k) {
    const std::uint32_t prev = tab[k - 1][i];
    tab[k][i] = (prev >> 8) ^ tab[0][prev & 0xFFu];
    __syncthreads();
  }
}

// Four message bytes (little-endian word w) into the register.
__device__ __forceinline__ std::uint32_t CrcWord(const std::uint32_t (*tab)[256], std::uint32_t c, std::uint32_t w) {
  c ^= w;
  return tab[3][c & 0xFFu] ^ tab[2][(c >> 8) & 0xFFu] ^ tab[1][(c >> 16) & 0xFFu] ^ tab[0][c >> 24];
}

__device__ __forceinline__ std::uint32_t CrcByte(const std::uint32_t (*tab)[256], std::uint32_t c, std::uint32_t b) {
  return (c >> 8) ^ tab[0][(c ^ b) & 0xFFu];
}

// Joins the 256 per-thread raw values of consecutive pieces held in s[] (in
// piece order) into s[0].  span_bytes(first, stride) is the byte length of
// the pieces of threads [first, first + stride) (0 past the end); it is the
// right operand's length that decides each shift.
template <typename PieceBytes>
__device__ inline void CrcJoinTree(const CrcPowers& t, std::uint32_t* s, PieceBytes span_bytes) {
  for (int stride = 1; stride < kCrcThreads; stride <<= 1) {
    __syncthreads();
    const int i = threadIdx.x;
    if ((i & (2 * stride - 1)) == 0) {
      const std::uint64_t right = span_bytes(i + stride, stride);
      if (right != 0) s[i] = CrcJoin(t, s[i], s[i + stride], right);
    }
  }
  __syncthreads();
}

// raw() of every full 64 KiB block of `data` (16-byte aligned): thread t
// checksums bytes [256 t, 256 t + 256) of its block, all sixteen 16-byte
// loads issued before the table work, then the block's pieces are joined.
__global__ __launch_bounds__(kCrcThreads) void CrcBlocksKernel(const std::uint8_t* __restrict__ data,
                                                               std::uint64_t blocks, CrcPowers t,
                                                               std::uint32_t* __restrict__ out) {
  __shared__ std::uint32_t tab[4][256];
  __shared__ std::uint32_t s[kCrcThreads];
  CrcBuildTables(tab);
  for (std::uint64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
    const uint4* p = reinterpret_cast<const uint4*>(data + b * kCrcBlockBytes +
                                                    static_cast<std::uint64_t>(threadIdx.x) * kCrcPieceBytes);
    uint4 v[kCrcPieceBytes / 16];
#pragma unroll
    for (int j = 0; j < kCrcPieceBytes / 16; ++j) v[j] = p[j];
    std::uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kCrcPieceBytes / 16; ++j) {
      c = CrcWord(tab, c, v[j].x);
      c = CrcWord(tab, c, v[j].y);
      c = CrcWord(tab, c, v[j].z);
      c = CrcWord(tab, c, v[j].w);
    }
    s[threadIdx.x] = c;
    CrcJoinTree(t, s, [](int, int stride) { return static_cast<std::uint64_t>(stride) * kCrcPieceBytes; });
    if (threadIdx.x == 0) out[b] = s[0];
    __syncthreads();
  }
}

// raw() of a short range (fewer than 64 KiB, any alignment) read byte by
// byte: one workgroup per range.  Range r is [begin[r], begin[r] + bytes[r]).
struct CrcShortRanges {
  const std::uint8_t* begin[2];
  std::uint64_t bytes[2];
};

__global__ __launch_bounds__(kCrcThreads) void CrcShortKernel(CrcShortRanges ranges, CrcPowers t,
                                                              std::uint32_t* __restrict__ out) {
  __shared__ std::uint32_t tab[4][256];
  __shared__ std::uint32_t s[kCrcThreads];
  CrcBuildTables(tab);
  const std::uint8_t* data = ranges.begin[blockIdx.x];
  const std::uint64_t bytes = ranges.bytes[blockIdx.x];
  const std::uint64_t from = static_cast<std::uint64_t>(threadIdx.x) * kCrcPieceBytes;
  const std::uint64_t to = from + kCrcPieceBytes < bytes ? from + kCrcPieceBytes : bytes;
  std::uint32_t c = 0;
  for (std::uint64_t i = from; i < to; ++i) c = CrcByte(tab, c, data[i]);
  s[threadIdx.x] = c;
  CrcJoinTree(t, s, [bytes](int first, int stride) -> std::uint64_t {
    const std::uint64_t lo = static_cast<std::uint64_t>(first) * kCrcPieceBytes;
    if (lo >= bytes) return 0;
    const std::uint64_t hi = lo + static_cast<std::uint64_t>(stride) * kCrcPieceBytes;
    return (hi < bytes ? hi : bytes) - lo;
  });
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// zlib's crc32(start, head ++ blocks ++ tail) into *crc_out, where start is
// *crc_in, or crc_in_value when crc_in is null.  `short_raw`
// holds raw(head) and raw(tail) from CrcShortKernel; thread i first joins a
// contiguous run of the block values in order, then the runs are joined in
// a tree.  crc_in and crc_out are device words, so checksums of
// consecutive buffers chain on one stream.
__global__ __launch_bounds__(kCrcThreads) void CrcCombineKernel(const std::uint32_t* __restrict__ block_raw,
                                                                std::uint64_t blocks,
                                                                const std::uint32_t* __restrict__ short_raw,
                                                                std::uint64_t head_bytes, std::uint64_t tail_bytes,
                                                                CrcPowers t, const std::uint32_t* crc_in,
                                                                std::uint32_t crc_in_value,
                                                                std::uint32_t* crc_out) {
  __shared__ std::uint32_t s[kCrcThreads];
  const std::uint64_t per = (blocks + kCrcThreads - 1) / kCrcThreads;
  const std::uint64_t first = static_cast<std::uint64_t>(threadIdx.x) * per;
  const std::uint64_t last = first + per < blocks ? first + per : blocks;
  const std::uint32_t shift_block = t.x8pow2[kCrcBlockShift];  // x^(8 * 64 KiB)
  std::uint32_t c = 0;
  for (std::uint64_t b = first; b < last; ++b) c = CrcMulMod(c, shift_block) ^ block_raw[b];
  s[threadIdx.x] = c;
  CrcJoinTree(t, s, [per, blocks](int i, int stride) -> std::uint64_t {
    const std::uint64_t lo = static_cast<std::uint64_t>(i) * per;
    if (lo >= blocks) return 0;
    const std::uint64_t hi = lo + static_cast<std::uint64_t>(stride) * per;
    return ((hi < blocks ? hi : blocks) - lo) << kCrcBlockShift;
  });
  if (threadIdx.x == 0) {
    const std::uint64_t body = blocks << kCrcBlockShift;
    std::uint32_t raw = CrcJoin(t, short_raw[0], s[0], body);
    raw = CrcJoin(t, raw, short_raw[1], tail_bytes);
    const std::uint32_t start = crc_in != nullptr ? *crc_in : crc_in_value;
    *crc_out = CrcFinish(t, start, raw, head_bytes + body + tail_bytes);
  }
}

}  // namespace sfk
