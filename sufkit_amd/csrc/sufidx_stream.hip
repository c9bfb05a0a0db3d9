// Streaming of device-resident index data to a host sink, with the CRC-32
// of every streamed byte computed on the device.
//
// The reference writes a .sufidx section by pushing host arrays through a
// checksumming stream buffer (serialization.cpp:60-141, 529-619): zlib's
// crc32 over every byte on one host thread, then a second full pass when
// the file is reread for validation.  Here a section payload that lives in
// HBM (text, SA, ISA, LCP) goes out in chunks: a kernel lays the chunk out
// in its stored plane form (native32, the low 32 bits, the high 8 or 16
// bits of split40/split48, native64, or a byte-coded LCP's capped bytes)
// when that differs from the device
// form, the CRC kernels checksum the chunk where it lies, and the chunk is
// copied to one of two pinned host buffers while the host hands the other
// one to the sink (a file write).  The checksum of a plane continues from
// the caller's running section checksum, on the device, so a section's CRC
// is complete when its last plane is.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "crc_kernels.hpp"
#include "sufkit_hip.h"

namespace sfk {
namespace {

constexpr unsigned kPlaneRange = 1u;  // a value does not fit its stored plane

struct StreamFailure {
  int status;
  std::string message;
};

[[noreturn]] void StreamFail(int status, const std::string& message) { throw StreamFailure{status, message}; }

void Check(hipError_t err, const char* what) {
  if (err != hipSuccess) {
    (void)hipGetLastError();
    StreamFail(SFK_HIP_BUILD_FAILURE, std::string(what) + ": " + hipGetErrorString(err));
  }
}

void CopyMessage(const std::string& message, char* error, size_t capacity) {
  if (error == nullptr || capacity == 0) return;
  const std::size_t len = std::min(message.size(), capacity - 1);
  std::memcpy(error, message.data(), len);
  error[len] = '\0';
}

std::uint32_t PlaneElementBytes(std::uint32_t plane) {
  switch (plane) {
    case SFK_HIP_PLANE_BYTES: return 1;
    case SFK_HIP_PLANE_NATIVE32: return 4;
    case SFK_HIP_PLANE_LOW32: return 4;
    case SFK_HIP_PLANE_HIGH8: return 1;
    case SFK_HIP_PLANE_HIGH16: return 2;
    case SFK_HIP_PLANE_NATIVE64: return 8;
    case SFK_HIP_PLANE_CAPPED8: return 1;
    default: return 0;
  }
}

// The stored form of values v[i] in one plane.  Native32 and the split
// planes check the range their codec can hold (32, 40 or 48 bits).
template <typename SrcT, std::uint32_t kPlane>
__global__ __launch_bounds__(256) void PlaneKernel(const SrcT* __restrict__ src, std::uint64_t count,
                                                   void* __restrict__ dst, unsigned int* __restrict__ error) {
  const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const std::uint64_t v = static_cast<std::uint64_t>(src[i]);
  if constexpr (kPlane == SFK_HIP_PLANE_NATIVE32) {
    if ((v >> 32) != 0) atomicOr(error, kPlaneRange);
    static_cast<std::uint32_t*>(dst)[i] = static_cast<std::uint32_t>(v);
  } else if constexpr (kPlane == SFK_HIP_PLANE_LOW32) {
    static_cast<std::uint32_t*>(dst)[i] = static_cast<std::uint32_t>(v);
  } else if constexpr (kPlane == SFK_HIP_PLANE_HIGH8) {
    if ((v >> 40) != 0) atomicOr(error, kPlaneRange);
    static_cast<std::uint8_t*>(dst)[i] = static_cast<std::uint8_t>(v >> 32);
  } else if constexpr (kPlane == SFK_HIP_PLANE_HIGH16) {
    if ((v >> 48) != 0) atomicOr(error, kPlaneRange);
    static_cast<std::uint16_t*>(dst)[i] = static_cast<std::uint16_t>(v >> 32);
  } else if constexpr (kPlane == SFK_HIP_PLANE_CAPPED8) {
    static_cast<std::uint8_t*>(dst)[i] = static_cast<std::uint8_t>(v < 255 ? v : 255);
  } else {
    static_cast<std::uint64_t*>(dst)[i] = v;
  }
}

template <typename SrcT>
void LaunchPlane(std::uint32_t plane, const SrcT* src, std::uint64_t count, void* dst, unsigned int* error,
                 hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>((count + 255) / 256));
  switch (plane) {
    case SFK_HIP_PLANE_NATIVE32: PlaneKernel<SrcT, SFK_HIP_PLANE_NATIVE32><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
    case SFK_HIP_PLANE_LOW32: PlaneKernel<SrcT, SFK_HIP_PLANE_LOW32><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
    case SFK_HIP_PLANE_HIGH8: PlaneKernel<SrcT, SFK_HIP_PLANE_HIGH8><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
    case SFK_HIP_PLANE_HIGH16: PlaneKernel<SrcT, SFK_HIP_PLANE_HIGH16><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
    case SFK_HIP_PLANE_CAPPED8: PlaneKernel<SrcT, SFK_HIP_PLANE_CAPPED8><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
    default: PlaneKernel<SrcT, SFK_HIP_PLANE_NATIVE64><<<grid, 256, 0, stream>>>(src, count, dst, error); break;
  }
  Check(hipGetLastError(), "plane kernel launch");
}

// Device words of one CRC computation: the running checksum, raw() of the
// short head and tail ranges, and raw() of up to `blocks` 64 KiB blocks.
struct CrcScratch {
  std::uint32_t* running = nullptr;
  std::uint32_t* short_raw = nullptr;
  std::uint32_t* block_raw = nullptr;
  std::uint64_t blocks = 0;
};

// zlib's crc32 of `bytes` at `data`, continuing from *scratch.running (or
// from crc_in_value when use_value), into *scratch.running.  Three launches
// on `stream`: head and tail bytes, the 64 KiB blocks, the join.
void LaunchCrc(const std::uint8_t* data, std::uint64_t bytes, bool use_value, std::uint32_t crc_in_value,
               const CrcScratch& scratch, const CrcPowers& powers, hipStream_t stream) {
  const std::uint64_t misalign = reinterpret_cast<std::uintptr_t>(data) % 16;
  const std::uint64_t head = std::min<std::uint64_t>(bytes, misalign == 0 ? 0 : 16 - misalign);
  const std::uint64_t blocks = (bytes - head) / kCrcBlockBytes;
  const std::uint64_t tail = bytes - head - blocks * kCrcBlockBytes;
  if (blocks > scratch.blocks) StreamFail(SFK_HIP_BUILD_FAILURE, "CRC block scratch too small");
  CrcShortRanges ranges{};
  ranges.begin[0] = data;
  ranges.bytes[0] = head;
  ranges.begin[1] = data + head + blocks * kCrcBlockBytes;
  ranges.bytes[1] = tail;
  CrcShortKernel<<<2, kCrcThreads, 0, stream>>>(ranges, powers, scratch.short_raw);
  Check(hipGetLastError(), "CRC short-range launch");
  if (blocks != 0) {
    const unsigned grid = static_cast<unsigned>(std::min<std::uint64_t>(blocks, 4096));
    CrcBlocksKernel<<<grid, kCrcThreads, 0, stream>>>(data + head, blocks, powers, scratch.block_raw);
    Check(hipGetLastError(), "CRC block launch");
  }
  CrcCombineKernel<<<1, kCrcThreads, 0, stream>>>(scratch.block_raw, blocks, scratch.short_raw, head, tail, powers,
                                                  use_value ? nullptr : scratch.running, crc_in_value,
                                                  scratch.running);
  Check(hipGetLastError(), "CRC join launch");
}

}  // namespace
}  // namespace sfk

struct sfk_hip_plane_writer {
  int device = 0;
  std::uint64_t chunk_bytes = 0;
  hipStream_t stream = nullptr;
  void* pinned[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};
  void* staging[2] = {nullptr, nullptr};
  std::uint32_t* words = nullptr;  // device: running CRC, 2 short raws, error flag, block raws
  std::uint32_t* host_words = nullptr;  // pinned: final CRC and error flag
  sfk::CrcScratch crc;
  sfk::CrcPowers powers{};

  void Release() {
    if (stream != nullptr) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; ++i) {
      if (pinned[i] != nullptr) (void)hipHostFree(pinned[i]);
      if (staging[i] != nullptr) (void)hipFree(staging[i]);
      if (copied[i] != nullptr) (void)hipEventDestroy(copied[i]);
      pinned[i] = staging[i] = nullptr;
      copied[i] = nullptr;
    }
    if (words != nullptr) (void)hipFree(words);
    if (host_words != nullptr) (void)hipHostFree(host_words);
    if (stream != nullptr) (void)hipStreamDestroy(stream);
    words = host_words = nullptr;
    stream = nullptr;
  }
};

namespace {

constexpr std::uint64_t kDefaultChunkBytes = 64ull << 20;

// Streams one plane of `count` values through the writer; returns the
// plane's bytes.  The sink sees the chunks in order; the running CRC stays
// on the device until the last chunk has been handed over.
std::uint64_t StreamPlane(sfk_hip_plane_writer* w, const void* d_values, std::uint64_t count,
                          std::uint32_t value_bits, std::uint32_t plane, std::uint32_t crc_in, sfk_hip_sink sink,
                          void* context, std::uint32_t* crc_out) {
  using sfk::Check;
  const std::uint32_t out_bytes = sfk::PlaneElementBytes(plane);
  if (out_bytes == 0) sfk::StreamFail(SFK_HIP_INVALID_INPUT, "unknown plane kind");
  if (plane == SFK_HIP_PLANE_BYTES ? value_bits != 8 : (value_bits != 32 && value_bits != 64)) {
    sfk::StreamFail(SFK_HIP_INVALID_INPUT, "value width does not suit the plane kind");
  }
  if (count != 0 && d_values == nullptr) sfk::StreamFail(SFK_HIP_INVALID_INPUT, "device values are required");
  const std::uint32_t in_bytes = value_bits / 8;
  // The stored form equals the device form: stream the values where they lie.
  const bool direct = plane == SFK_HIP_PLANE_BYTES ||
                      (in_bytes == 4 && (plane == SFK_HIP_PLANE_NATIVE32 || plane == SFK_HIP_PLANE_LOW32)) ||
                      (in_bytes == 8 && plane == SFK_HIP_PLANE_NATIVE64);
  Check(hipSetDevice(w->device), "hipSetDevice");
  // The values may come from any stream of this device.
  Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  std::uint32_t* error = w->words + 3;
  Check(hipMemsetAsync(error, 0, sizeof(std::uint32_t), w->stream), "clear plane error flag");
  const std::uint64_t per_chunk = w->chunk_bytes / out_bytes;
  const std::uint64_t chunks = count == 0 ? 0 : (count + per_chunk - 1) / per_chunk;
  std::uint64_t pending_bytes[2] = {0, 0};
  auto drain = [&](std::uint64_t c) {
    const int b = static_cast<int>(c & 1u);
    Check(hipEventSynchronize(w->copied[b]), "wait for chunk copy");
    if (sink(context, w->pinned[b], pending_bytes[b]) != 0) {
      sfk::StreamFail(SFK_HIP_IO_ERROR, "the sink refused a chunk");
    }
  };
  bool first = true;
  for (std::uint64_t c = 0; c < chunks; ++c) {
    const int b = static_cast<int>(c & 1u);
    const std::uint64_t begin = c * per_chunk;
    const std::uint64_t n = std::min(per_chunk, count - begin);
    const std::uint64_t bytes = n * out_bytes;
    const std::uint8_t* chunk = nullptr;
    if (direct) {
      chunk = static_cast<const std::uint8_t*>(d_values) + begin * in_bytes;
    } else {
      if (in_bytes == 4) {
        sfk::LaunchPlane(plane, static_cast<const std::uint32_t*>(d_values) + begin, n, w->staging[b], error,
                         w->stream);
      } else {
        sfk::LaunchPlane(plane, static_cast<const std::uint64_t*>(d_values) + begin, n, w->staging[b], error,
                         w->stream);
      }
      chunk = static_cast<const std::uint8_t*>(w->staging[b]);
    }
    sfk::LaunchCrc(chunk, bytes, first, crc_in, w->crc, w->powers, w->stream);
    first = false;
    Check(hipMemcpyAsync(w->pinned[b], chunk, bytes, hipMemcpyDeviceToHost, w->stream), "chunk copy");
    Check(hipEventRecord(w->copied[b], w->stream), "record chunk copy");
    pending_bytes[b] = bytes;
    if (c >= 1) drain(c - 1);
  }
  if (chunks >= 1) drain(chunks - 1);
  if (chunks == 0) {
    w->host_words[0] = crc_in;
    w->host_words[1] = 0;
  } else {
    Check(hipMemcpyAsync(w->host_words, w->crc.running, sizeof(std::uint32_t), hipMemcpyDeviceToHost, w->stream),
          "read CRC");
    Check(hipMemcpyAsync(w->host_words + 1, error, sizeof(std::uint32_t), hipMemcpyDeviceToHost, w->stream),
          "read plane error flag");
    Check(hipStreamSynchronize(w->stream), "hipStreamSynchronize");
  }
  if (w->host_words[1] != 0) {
    sfk::StreamFail(SFK_HIP_INVALID_INPUT, "a value does not fit the stored coordinate width");
  }
  if (crc_out != nullptr) *crc_out = w->host_words[0];
  return count * out_bytes;
}

}  // namespace

extern "C" {

int sfk_hip_plane_writer_create(int device, uint64_t chunk_bytes, sfk_hip_plane_writer** out, char* error,
                                size_t error_capacity) {
  if (out == nullptr) return SFK_HIP_INVALID_INPUT;
  *out = nullptr;
  if (chunk_bytes == 0) chunk_bytes = kDefaultChunkBytes;
  if (chunk_bytes % 8 != 0 || chunk_bytes > (1ull << 32)) {
    sfk::CopyMessage("chunk size must be a multiple of 8 bytes and at most 4 GiB", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || device < 0 || device >= devices) {
    (void)hipGetLastError();
    sfk::CopyMessage("no HIP device " + std::to_string(device), error, error_capacity);
    return SFK_HIP_UNSUPPORTED_BACKEND;
  }
  auto* w = new (std::nothrow) sfk_hip_plane_writer();
  if (w == nullptr) return SFK_HIP_BUILD_FAILURE;
  try {
    using sfk::Check;
    w->device = device;
    w->chunk_bytes = chunk_bytes;
    w->powers = sfk::MakeCrcPowers();
    Check(hipSetDevice(device), "hipSetDevice");
    Check(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    for (int i = 0; i < 2; ++i) {
      Check(hipHostMalloc(&w->pinned[i], chunk_bytes, hipHostMallocDefault), "hipHostMalloc");
      Check(hipMalloc(&w->staging[i], chunk_bytes), "hipMalloc");
      Check(hipEventCreateWithFlags(&w->copied[i], hipEventDisableTiming), "hipEventCreateWithFlags");
    }
    w->crc.blocks = chunk_bytes / sfk::kCrcBlockBytes + 1;
    Check(hipMalloc(reinterpret_cast<void**>(&w->words), (4 + w->crc.blocks) * sizeof(std::uint32_t)), "hipMalloc");
    Check(hipHostMalloc(reinterpret_cast<void**>(&w->host_words), 2 * sizeof(std::uint32_t), hipHostMallocDefault),
          "hipHostMalloc");
    w->crc.running = w->words;
    w->crc.short_raw = w->words + 1;
    w->crc.block_raw = w->words + 4;
    *out = w;
    return SFK_HIP_OK;
  } catch (const sfk::StreamFailure& f) {
    sfk::CopyMessage(f.message, error, error_capacity);
    w->Release();
    delete w;
    return f.status;
  }
}

void sfk_hip_plane_writer_destroy(sfk_hip_plane_writer* writer) {
  if (writer == nullptr) return;
  writer->Release();
  delete writer;
}

int sfk_hip_write_plane(sfk_hip_plane_writer* writer, const void* d_values, uint64_t count, uint32_t value_bits,
                        uint32_t plane, uint32_t crc_in, sfk_hip_sink sink, void* context, uint32_t* crc_out,
                        char* error, size_t error_capacity) {
  if (writer == nullptr || sink == nullptr) {
    sfk::CopyMessage("writer and sink are required", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  try {
    (void)StreamPlane(writer, d_values, count, value_bits, plane, crc_in, sink, context, crc_out);
    return SFK_HIP_OK;
  } catch (const sfk::StreamFailure& f) {
    (void)hipStreamSynchronize(writer->stream);
    sfk::CopyMessage(f.message, error, error_capacity);
    return f.status;
  }
}

int sfk_hip_crc32(int device, const void* d_data, uint64_t bytes, uint32_t crc_in, uint32_t* crc_out, char* error,
                  size_t error_capacity) {
  if (crc_out == nullptr || (bytes != 0 && d_data == nullptr)) {
    sfk::CopyMessage("device data and an output word are required", error, error_capacity);
    return SFK_HIP_INVALID_INPUT;
  }
  if (bytes == 0) {
    *crc_out = crc_in;
    return SFK_HIP_OK;
  }
  std::uint32_t* words = nullptr;
  std::uint32_t* host = nullptr;
  hipStream_t stream = nullptr;
  int status = SFK_HIP_OK;
  try {
    using sfk::Check;
    Check(hipSetDevice(device), "hipSetDevice");
    Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    sfk::CrcScratch scratch;
    scratch.blocks = bytes / sfk::kCrcBlockBytes + 1;
    Check(hipMalloc(reinterpret_cast<void**>(&words), (4 + scratch.blocks) * sizeof(std::uint32_t)), "hipMalloc");
    Check(hipHostMalloc(reinterpret_cast<void**>(&host), sizeof(std::uint32_t), hipHostMallocDefault),
          "hipHostMalloc");
    scratch.running = words;
    scratch.short_raw = words + 1;
    scratch.block_raw = words + 4;
    const sfk::CrcPowers powers = sfk::MakeCrcPowers();
    sfk::LaunchCrc(static_cast<const std::uint8_t*>(d_data), bytes, true, crc_in, scratch, powers, stream);
    Check(hipMemcpyAsync(host, words, sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream), "read CRC");
    Check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    *crc_out = *host;
  } catch (const sfk::StreamFailure& f) {
    sfk::CopyMessage(f.message, error, error_capacity);
    status = f.status;
  }
  if (stream != nullptr) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
  }
  if (words != nullptr) (void)hipFree(words);
  if (host != nullptr) (void)hipHostFree(host);
  return status;
}

}  // extern "C"
