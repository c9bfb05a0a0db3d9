// SaveDeviceIndex: the .sufidx file of a device build, streamed from HBM.
//
// The sections, their order and bytes are those SuffixArray::Save writes
// for the same build (metadata, text, SA, ISA, LCP; suffix_array.cpp:
// 4853-4916): the stored widths follow the same policies Build applies
// (ResolveProfileStorageWidth, ResolveAuxiliaryStorageWidth, raw LCP of 32
// bits unless n - 1 needs 64), the codec headers are built on the host, and
// every payload plane is streamed from device memory by
// sfk_hip_write_plane, which checksums it on the device.  A Low-memory
// index (byte-coded LCP, no ISA) streams its LCP's capped bytes the same
// way; its few overflow anchors are found on the device and written from
// the host.
#include <chrono>
#include <limits>
#include <sstream>
#include <string>

#include "internal.hpp"
#include "sufkit_hip.h"

namespace sufkit {

namespace {

using detail::StreamPart;
using detail::StreamSection;

StreamPart HostPart(std::vector<std::uint8_t> bytes) {
  StreamPart p;
  p.host = std::move(bytes);
  return p;
}

StreamPart DevicePart(const void* values, std::uint64_t count, std::uint32_t bits, std::uint32_t plane) {
  StreamPart p;
  p.on_device = true;
  p.device_values = values;
  p.count = count;
  p.value_bits = bits;
  p.plane = plane;
  return p;
}

// SKC1 section: codec header, then the planes of the stored width.
StreamSection CoordinateSection(detail::SectionType type, CoordinateStorageWidth width, const void* values,
                                std::uint32_t bits, std::uint64_t count, std::uint64_t symbol_count) {
  StreamSection s{type, {}};
  s.parts.push_back(HostPart(detail::CoordinateSectionHeader(width, symbol_count, count)));
  switch (width) {
    case CoordinateStorageWidth::kBits32:
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_NATIVE32));
      break;
    case CoordinateStorageWidth::kBits40:
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_LOW32));
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_HIGH8));
      break;
    case CoordinateStorageWidth::kBits48:
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_LOW32));
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_HIGH16));
      break;
    case CoordinateStorageWidth::kBits64:
      s.parts.push_back(DevicePart(values, count, bits, SFK_HIP_PLANE_NATIVE64));
      break;
    default:
      detail::Throw(ErrorCode::kBuildFailure, "storage width was not resolved before packing");
  }
  return s;
}

void CheckBits(std::uint32_t bits, const char* label) {
  if (bits != 32 && bits != 64) {
    detail::Throw(ErrorCode::kInvalidInput, std::string(label) + " device values must be 32 or 64 bits wide");
  }
}

}  // namespace

void SaveDeviceIndex(const GenomeReference& reference, const DeviceIndexBuffers& buffers,
                     const std::filesystem::path& path, const SaveOptions& save,
                     const SuffixArrayBuildOptions& options, DeviceSaveStatistics* statistics) {
  const auto begin = std::chrono::steady_clock::now();
  const std::uint64_t n = reference.Encoded().size() + 1;
  if (n < 2) detail::Throw(ErrorCode::kInvalidInput, "the reference has no text");
  if (options.sampling_rate != 1) {
    detail::Throw(ErrorCode::kInvalidInput, "saving from device memory needs sampling rate 1");
  }
  // The Low-memory profile stores a byte-coded LCP and no ISA
  // (suffix_array.cpp:4142-4151); its anchors are found from the device ISA.
  const bool low_memory = options.resource_profile == SaResourceProfile::kLowMemory;
  if (!low_memory &&
      (options.acceleration == SaAcceleration::kLcpChild || options.acceleration == SaAcceleration::kFull)) {
    detail::Throw(ErrorCode::kUnsupportedBackend, "CHILD arrays are not part of this build");
  }
  if (static_cast<std::uint8_t>(options.acceleration) > static_cast<std::uint8_t>(SaAcceleration::kFull)) {
    detail::Throw(ErrorCode::kInvalidInput, "acceleration must be none, LCP, LCP+CHILD, LCP+suffix links or full");
  }
  const auto width = buffers.construction_width;
  if (width != CoordinateWidth::kBits32 && width != CoordinateWidth::kBits64) {
    detail::Throw(ErrorCode::kInvalidInput, "construction width must be auto, 32 or 64 bits");
  }
  if (!detail::SaConstructionCanRepresent(SaBackend::kHip, width, n)) {
    detail::Throw(ErrorCode::kInvalidInput, "text is longer than the chosen construction width can index");
  }
  const bool retain_lcp = low_memory || options.acceleration != SaAcceleration::kNone;
  const bool retain_isa = !low_memory && options.acceleration == SaAcceleration::kLcpSuffixLink;
  if (buffers.text == nullptr || buffers.sa == nullptr || ((retain_isa || low_memory) && buffers.isa == nullptr) ||
      (retain_lcp && buffers.lcp == nullptr)) {
    detail::Throw(ErrorCode::kInvalidInput, "device text, SA and the retained ISA / LCP are required");
  }
  CheckBits(buffers.sa_bits, "SA");
  if (retain_isa || low_memory) CheckBits(buffers.isa_bits, "ISA");
  if (retain_lcp) CheckBits(buffers.lcp_bits, "LCP");
  const auto stored = detail::ResolveProfileStorageWidth(options.storage_width, options.resource_profile, n);

  detail::ContainerSpec spec;
  spec.format_minor = 4;
  spec.kind = 1;
  spec.backend = width == CoordinateWidth::kBits32 ? detail::StoredBackend::kHip32 : detail::StoredBackend::kHip64;
  spec.coordinate_width = static_cast<std::uint8_t>(width);
  spec.profile = options.resource_profile;
  spec.sequence_count = reference.Sequences().size();
  spec.total_bases = reference.TotalBases();
  spec.text_symbols = n;
  spec.ambiguous_bases = reference.AmbiguousBases();
  spec.fingerprint = reference.Fingerprint();

  std::vector<StreamSection> sections;
  {
    std::ostringstream meta;
    detail::WriteMetadataSection(meta, reference.Sequences());
    const std::string bytes = meta.str();
    sections.push_back({detail::SectionType::kMetadata, {HostPart({bytes.begin(), bytes.end()})}});
  }
  sections.push_back({detail::SectionType::kText, {DevicePart(buffers.text, n, 8, SFK_HIP_PLANE_BYTES)}});
  sections.push_back(
      CoordinateSection(detail::SectionType::kSuffixArray, stored, buffers.sa, buffers.sa_bits, n, n));
  if (retain_isa) {
    const auto aux = detail::ResolveAuxiliaryStorageWidth(stored, n);
    sections.push_back(
        CoordinateSection(detail::SectionType::kInverseSuffixArray, aux, buffers.isa, buffers.isa_bits, n, n));
  }
  if (low_memory) {
    // SKL1 codec 1: primary bytes streamed from the raw LCP, anchors found
    // on the device and written from the host (sa_codec.cpp:569-631).
    const std::uint32_t anchor_width = n - 1U > std::numeric_limits<std::uint32_t>::max() ? 64 : 32;
    char err[512] = {0};
    // The checks BuildByteCoded makes before coding (lcp_storage.cpp:237-297):
    // LCP[0] is 0, the ISA is the inverse of a permutation SA (here with the
    // order of every adjacent pair as well), anchors fit their coordinates.
    std::uint64_t first = 0;
    if (sfk_hip_memcpy_d2h(&first, buffers.lcp, buffers.lcp_bits / 8) != SFK_HIP_OK) {
      detail::Throw(ErrorCode::kBuildFailure, "cannot read the first LCP row from the device");
    }
    if (first != 0) detail::Throw(ErrorCode::kBuildFailure, "byte-coded LCP: row 0 of the raw LCP is not 0");
    sfk_hip_check_report report{};
    if (sfk_hip_check_index(buffers.device, static_cast<const std::uint8_t*>(buffers.text), n, buffers.sa,
                            buffers.sa_bits, buffers.isa, buffers.isa_bits, nullptr, 0, 1, &report, err,
                            sizeof(err)) != SFK_HIP_OK) {
      detail::Throw(ErrorCode::kBuildFailure, std::string("byte-coded LCP: ") + err);
    }
    std::uint64_t anchors = 0;
    int status = sfk_hip_byte_coded_lcp(buffers.device, buffers.lcp, buffers.lcp_bits, buffers.isa, buffers.isa_bits,
                                        n, nullptr, &anchors, nullptr, nullptr, 0, err, sizeof(err));
    std::vector<std::uint64_t> positions(anchors), values(anchors);
    if (status == SFK_HIP_OK && anchors != 0) {
      status = sfk_hip_byte_coded_lcp(buffers.device, buffers.lcp, buffers.lcp_bits, buffers.isa, buffers.isa_bits,
                                      n, nullptr, &anchors, positions.data(), values.data(), positions.size(), err,
                                      sizeof(err));
    }
    if (status != SFK_HIP_OK) {
      detail::Throw(static_cast<ErrorCode>(status - 1), std::string("byte-coded LCP anchors: ") + err);
    }
    const std::uint64_t value_limit = anchor_width == 64 ? std::numeric_limits<std::uint64_t>::max()
                                                         : std::numeric_limits<std::uint32_t>::max();
    for (std::uint64_t a = 0; a < anchors; ++a) {
      if (positions[a] >= n || values[a] > value_limit || values[a] < detail::LcpStorage::kByteLimit) {
        detail::Throw(ErrorCode::kBuildFailure,
                      "byte-coded LCP: anchor " + std::to_string(a) + " at position " + std::to_string(positions[a]) +
                          " holds " + std::to_string(values[a]) + ", outside the " +
                          std::to_string(anchor_width) + "-bit anchor range");
      }
    }
    std::vector<std::uint8_t> planes;
    planes.reserve(2 * anchors * (anchor_width / 8));
    for (const auto* plane : {&positions, &values}) {
      for (const std::uint64_t v : *plane) {
        for (std::uint32_t b = 0; b < anchor_width / 8; ++b) planes.push_back(static_cast<std::uint8_t>(v >> (8 * b)));
      }
    }
    StreamSection s{detail::SectionType::kLcp, {}};
    s.parts.push_back(HostPart(detail::LcpSectionHeader(anchor_width, 1, n, n, true, anchors)));
    s.parts.push_back(DevicePart(buffers.lcp, n, buffers.lcp_bits, SFK_HIP_PLANE_CAPPED8));
    s.parts.push_back(HostPart(std::move(planes)));
    sections.push_back(std::move(s));
  } else if (retain_lcp) {
    const std::uint32_t lcp_width = n - 1U > std::numeric_limits<std::uint32_t>::max() ? 64 : 32;
    StreamSection s{detail::SectionType::kLcp, {}};
    s.parts.push_back(HostPart(detail::LcpSectionHeader(lcp_width, 1, n, n)));
    s.parts.push_back(DevicePart(buffers.lcp, n, buffers.lcp_bits,
                                 lcp_width == 32 ? SFK_HIP_PLANE_NATIVE32 : SFK_HIP_PLANE_NATIVE64));
    sections.push_back(std::move(s));
  }
  detail::StreamStatistics s;
  detail::WriteContainerStreamed(path, save, spec, sections, buffers.device, &s);
  if (statistics != nullptr) {
    statistics->write_seconds = s.write_seconds;
    statistics->validate_seconds = s.validate_seconds;
    statistics->bytes = s.bytes;
    statistics->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
  }
}

}  // namespace sufkit
