// SuffixArray build / persist / load around the HIP construction backend.
//
// Build follows SuffixArray::Build (suffix_array.cpp:4086-4390): the same
// stage notifications, the same backend/width/storage/LCP-width policies,
// text = encoded ++ [sentinel], then one device build that returns the
// device-validated SA (permutation + inverse checked on the GPU, so the
// host does not repeat RepackSuffixArray's O(n) random bitmap pass), the ISA
// and the raw LCP directly in their resolved widths.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

#include "internal.hpp"
#include "sufkit_hip.h"

namespace sufkit {

SuffixArrayBuildOptions FastSuffixArrayBuildOptions() { return SuffixArrayBuildOptions{}; }

const char* StoredBackendName(std::uint8_t id) noexcept {
  switch (id) {
    case 1: return "divsufsort32";
    case 2: return "divsufsort64";
    case 3: return "caps32";
    case 4: return "caps64";
    case 5: return "hip32";
    case 6: return "hip64";
    default: return "unknown";
  }
}

const char* StoredBackendSignature(std::uint8_t id) noexcept {
  switch (id) {
    case 1: return "libdivsufsort-2.0.2/saidx_t";
    case 2: return "libdivsufsort-2.0.2/saidx64_t";
    case 3: return "CaPS-SA@2597b373/uint32_t+ParlayLib@e1f1dc0";
    case 4: return "CaPS-SA@2597b373/uint64_t+ParlayLib@e1f1dc0";
    case 5: return "sufkit-amd-hip-gfx950-prefix-doubling/uint32_t";
    case 6: return "sufkit-amd-hip-gfx950-prefix-doubling/uint64_t";
    default: return "unknown";
  }
}

bool HipBackendAvailable() noexcept { return sfk_hip_device_count() > 0; }

struct SuffixArray::Impl {
  std::vector<SequenceInfo> sequences;
  std::uint64_t total_bases = 0;
  std::uint64_t ambiguous_bases = 0;
  std::uint64_t fingerprint = 0;
  std::vector<std::uint8_t> text;
  detail::StoredBackend backend = detail::StoredBackend::kHip32;
  std::uint8_t construction_width = 32;
  std::uint32_t sampling_rate = 1;
  SaAcceleration acceleration = SaAcceleration::kLcpSuffixLink;
  detail::CoordinateStorage sa;
  detail::CoordinateStorage isa;
  detail::LcpStorage lcp;
  // Fast prefix directory (k-mer -> SA interval), derived after build and
  // load, never serialized (fast_prefix_index.hpp): k = 0 when absent.
  std::uint32_t prefix_k = 0;
  std::vector<std::uint64_t> prefix_table;
  IndexInfo info;

  void RefreshInfo() {
    info.format_version = "1.4";
    info.library_version = "0.3.0";
    info.backend = StoredBackendName(static_cast<std::uint8_t>(backend));
    info.backend_signature = StoredBackendSignature(static_cast<std::uint8_t>(backend));
    info.coordinate_width = construction_width;
    info.stored_coordinate_width = static_cast<std::uint8_t>(sa.Width());
    info.sequence_count = sequences.size();
    info.total_bases = total_bases;
    info.text_symbols = text.size();
    info.suffix_count = sa.Size();
    info.sa_sampling_rate = sampling_rate;
    info.ambiguous_bases = ambiguous_bases;
    info.fingerprint = fingerprint;
    info.sa_acceleration = acceleration;
    info.lcp_encoding =
        lcp.Empty() ? SaLcpEncoding::kNone : (lcp.ByteCoded() ? SaLcpEncoding::kByteCoded : SaLcpEncoding::kRaw);
    info.text_bytes = text.size();
    info.sa_bytes = sa.Bytes();
    info.isa_bytes = isa.Bytes();
    info.lcp_bytes = lcp.Bytes();
    // The directory's k is chosen from the core arrays alone; its bytes
    // then count as auxiliary and resident (suffix_array.cpp:2097-2106).
    info.auxiliary_bytes = prefix_table.size() * sizeof(std::uint64_t);
    info.resident_core_bytes = CoreBytes() + info.auxiliary_bytes;
  }

  std::uint64_t CoreBytes() const noexcept { return text.size() + sa.Bytes() + isa.Bytes() + lcp.Bytes(); }

  // BuildFastPrefix's conditions (suffix_array.cpp:2087-2091): the Fast
  // profile over every row with an ISA (no learned model exists here).
  bool WantsPrefixDirectory(SaResourceProfile profile) const noexcept {
    return profile == SaResourceProfile::kFast && sampling_rate == 1 && !isa.Empty();
  }

  // The directory of a loaded index, computed on the device from the host
  // text and SA (FastPrefixIndex::Build runs in Load too,
  // suffix_array.cpp:5242); left absent when no device is visible.
  void BuildPrefixDirectoryOnDevice();
};

namespace {

using Clock = std::chrono::steady_clock;

double Since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

ErrorCode FromStatus(int status) {
  switch (status) {
    case SFK_HIP_INVALID_INPUT: return ErrorCode::kInvalidInput;
    case SFK_HIP_IO_ERROR: return ErrorCode::kIoError;
    case SFK_HIP_UNSUPPORTED_BACKEND: return ErrorCode::kUnsupportedBackend;
    case SFK_HIP_CORRUPT_INDEX: return ErrorCode::kCorruptIndex;
    case SFK_HIP_VERSION_MISMATCH: return ErrorCode::kVersionMismatch;
    default: return ErrorCode::kBuildFailure;
  }
}

template <class T>
detail::CoordinateStorage Store(detail::ValueVector<T>&& values, CoordinateStorageWidth width,
                                std::uint64_t domain, ErrorCode code, const char* label) {
  if constexpr (std::is_same_v<T, std::uint32_t>) {
    return detail::CoordinateStorage::FromUInt32(std::move(values), width, domain, code, label);
  } else {
    return detail::CoordinateStorage::FromUInt64(std::move(values), width, domain, code, label);
  }
}

template <class SaT, class IsaT, class LcpT>
void BuildWithHip(SuffixArray::Impl& impl, const SuffixArrayBuildOptions& options,
                  CoordinateStorageWidth stored_width, bool retain_isa, bool retain_lcp, std::uint32_t prefix_k,
                  SuffixArrayBuildStatistics& stats, const std::function<void(const char*)>& notify) {
  const std::uint64_t n = impl.text.size();
  // K > 1: the device keeps the rows of positions that are multiples of K
  // (suffix_array.cpp:912-974), with their ISA and interval-minimum LCP.
  const std::uint32_t k = options.sampling_rate;
  const std::uint64_t rows = 1 + (n - 1) / k;
  // Filled by the device build (never read before it writes them).
  detail::ValueVector<SaT> sa(rows);
  detail::ValueVector<IsaT> isa(retain_isa ? rows : 0);
  detail::ValueVector<LcpT> lcp(retain_lcp ? rows : 0);
  std::vector<std::uint64_t> table(prefix_k > 0 ? sfk_prefix_directory_words(prefix_k, n) : 0);
  sfk_hip_stats hs{};
  char err[512] = {0};
  const std::uint32_t flags = (retain_isa ? SFK_HIP_RETAIN_ISA : 0u) | (retain_lcp ? SFK_HIP_RETAIN_LCP : 0u);
  std::uint64_t prefix_counts[2] = {0, 0};
  const int status = sfk_hip_build_fast(impl.text.data(), n, impl.construction_width, k, flags, sa.data(),
                                        8 * sizeof(SaT), isa.empty() ? nullptr : isa.data(), 8 * sizeof(IsaT),
                                        lcp.empty() ? nullptr : lcp.data(), 8 * sizeof(LcpT), prefix_k,
                                        table.empty() ? nullptr : table.data(), prefix_counts, &hs, err,
                                        sizeof(err));
  if (status != SFK_HIP_OK) detail::Throw(FromStatus(status), std::string("HIP backend: ") + err);
  impl.prefix_k = prefix_k;
  impl.prefix_table = std::move(table);
  stats.prefix_directory_seconds = hs.prefix_directory_seconds;
  stats.sa_seconds = hs.sa_seconds;
  stats.isa_seconds = hs.isa_seconds;
  stats.lcp_seconds = hs.lcp_seconds;
  stats.device_seconds = hs.device_seconds;
  stats.transfer_seconds = hs.h2d_seconds + hs.d2h_seconds;

  stats.call_seconds = hs.call_seconds;
  const auto compaction = Clock::now();
  notify("index-storage-layout");
  impl.sa = Store(std::move(sa), stored_width, n, ErrorCode::kBuildFailure, "suffix array");
  stats.storage_compaction_seconds = Since(compaction);
  if (retain_isa) {
    notify("index-isa");
    const auto aux = detail::ResolveAuxiliaryStorageWidth(stored_width, rows);
    impl.isa = Store(std::move(isa), aux, rows, ErrorCode::kBuildFailure, "inverse suffix array");
  }
  if (retain_lcp) {
    notify("index-lcp-layout");
    const auto fin = Clock::now();
    if constexpr (std::is_same_v<LcpT, std::uint32_t>) {
      impl.lcp = detail::LcpStorage::FromRaw32(std::move(lcp), k, ErrorCode::kBuildFailure);
    } else {
      impl.lcp = detail::LcpStorage::FromRaw64(std::move(lcp), k, ErrorCode::kBuildFailure);
    }
    stats.lcp_finalize_seconds = Since(fin);
  }
}

// Device buffer of the Low-memory build, freed on every path.
struct DeviceBuffer {
  void* p = nullptr;
  explicit DeviceBuffer(std::uint64_t bytes) {
    if (sfk_hip_malloc(0, bytes == 0 ? 1 : bytes, &p) != SFK_HIP_OK) {
      p = nullptr;
      detail::Throw(ErrorCode::kBuildFailure, "HIP backend: device allocation failed");
    }
  }
  ~DeviceBuffer() {
    if (p != nullptr) sfk_hip_free(p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Low-memory profile: the device builds SA, ISA and raw LCP, keeps the ISA
// and raw LCP only long enough to byte-code the LCP
// (sfk_hip_byte_coded_lcp), and returns the SA and the byte-coded LCP
// (BuildByteCodedLcpDirect, suffix_array.cpp:1103-1113, with the anchor
// coordinate width of the raw LCP's, :4152-4156).
void BuildLowMemoryWithHip(SuffixArray::Impl& impl, CoordinateStorageWidth stored_width,
                           SuffixArrayBuildStatistics& stats) {
  const std::uint64_t n = impl.text.size();
  const std::uint32_t bits = impl.construction_width;
  const std::uint32_t sa_bits = stored_width == CoordinateStorageWidth::kBits32 ? 32 : 64;
  const std::uint32_t wide = n - 1U > std::numeric_limits<std::uint32_t>::max() ? 64 : 32;
  char err[512] = {0};
  auto check = [&](int status, const char* what) {
    if (status != SFK_HIP_OK) detail::Throw(FromStatus(status), std::string("HIP backend: ") + what + ": " + err);
  };
  DeviceBuffer text(n + 64), sa(n * (sa_bits / 8)), isa(n * (wide / 8)), lcp(n * (wide / 8));
  check(sfk_hip_memcpy_h2d(text.p, impl.text.data(), n), "text upload");
  sfk_hip_builder* builder = nullptr;
  check(sfk_hip_builder_create(0, &builder, err, sizeof(err)), "builder");
  sfk_hip_stats hs{};
  const int built = sfk_hip_build_device(builder, static_cast<const std::uint8_t*>(text.p), n, bits, 1,
                                         SFK_HIP_RETAIN_ISA | SFK_HIP_RETAIN_LCP, sa.p, sa_bits, isa.p, wide, lcp.p,
                                         wide, &hs, err, sizeof(err));
  sfk_hip_builder_destroy(builder);
  check(built, "build");
  stats.sa_seconds = hs.sa_seconds;
  stats.isa_seconds = hs.isa_seconds;
  stats.lcp_seconds = hs.lcp_seconds;
  stats.device_seconds = hs.device_seconds;
  std::vector<std::uint8_t> primary(n);
  std::uint64_t anchors = 0;
  const auto fin = Clock::now();
  check(sfk_hip_byte_coded_lcp(0, lcp.p, wide, isa.p, wide, n, primary.data(), &anchors, nullptr, nullptr, 0, err,
                               sizeof(err)),
        "byte-coded LCP");
  std::vector<std::uint64_t> positions(anchors), values(anchors);
  if (anchors != 0) {
    check(sfk_hip_byte_coded_lcp(0, lcp.p, wide, isa.p, wide, n, nullptr, &anchors, positions.data(),
                                 values.data(), positions.size(), err, sizeof(err)),
          "byte-coded LCP anchors");
  }
  impl.lcp = detail::LcpStorage::FromByteCoded(std::move(primary), std::move(positions), std::move(values), 1, n,
                                               static_cast<std::uint8_t>(wide), ErrorCode::kBuildFailure);
  stats.lcp_finalize_seconds = Since(fin);
  const auto compaction = Clock::now();
  if (sa_bits == 32) {
    detail::U32Values values32(n);
    check(sfk_hip_memcpy_d2h(values32.data(), sa.p, n * 4), "SA download");
    impl.sa = Store(std::move(values32), stored_width, n, ErrorCode::kBuildFailure, "suffix array");
  } else {
    detail::U64Values values64(n);
    check(sfk_hip_memcpy_d2h(values64.data(), sa.p, n * 8), "SA download");
    impl.sa = Store(std::move(values64), stored_width, n, ErrorCode::kBuildFailure, "suffix array");
  }
  stats.storage_compaction_seconds = Since(compaction);
}

}  // namespace

SuffixArray::SuffixArray(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
SuffixArray::SuffixArray(SuffixArray&&) noexcept = default;
SuffixArray& SuffixArray::operator=(SuffixArray&&) noexcept = default;
SuffixArray::~SuffixArray() = default;

SuffixArray SuffixArray::Build(const GenomeReference& reference, const SuffixArrayBuildOptions& options) {
  const auto total_begin = Clock::now();
  const std::function<void(const char*)> notify = [&](const char* stage) {
    if (options.stage_callback) options.stage_callback(stage, options.stage_context);
  };
  notify("index-text-prepare");
  if (options.threads == 0) {
    detail::Throw(ErrorCode::kInvalidInput, "threads must be 1 or more");
  }
  if (options.sampling_rate == 0) {
    detail::Throw(ErrorCode::kInvalidInput, "sampling rate must be 1 or more");
  }
  if (static_cast<std::uint8_t>(options.acceleration) > static_cast<std::uint8_t>(SaAcceleration::kFull)) {
    detail::Throw(ErrorCode::kInvalidInput, "acceleration must be none, LCP, LCP+CHILD, LCP+suffix links or full");
  }
  SuffixArrayBuildStatistics local_stats;
  SuffixArrayBuildStatistics& stats = options.statistics ? *options.statistics : local_stats;
  stats = {};
  auto impl = std::make_unique<Impl>();
  const auto& encoded = reference.Encoded();
  impl->sequences = reference.Sequences();
  impl->total_bases = reference.TotalBases();
  impl->ambiguous_bases = reference.AmbiguousBases();
  impl->fingerprint = reference.Fingerprint();
  impl->text.reserve(encoded.size() + 1);
  impl->text.insert(impl->text.end(), encoded.begin(), encoded.end());
  impl->text.push_back(detail::kSentinel);
  if (impl->text.size() < 2) detail::Throw(ErrorCode::kInvalidInput, "the reference has no text");
  const std::uint64_t n = impl->text.size();

  const auto backend = detail::ResolveSaBackend(options.backend, n, options.threads, HipBackendAvailable());
  if (backend == SaBackend::kDivsufsort || backend == SaBackend::kCaps) {
    detail::Throw(ErrorCode::kUnsupportedBackend,
                  "the divsufsort and CaPS-SA engines are not part of this build; use SaBackend::kHip");
  }
  if (backend != SaBackend::kHip) detail::Throw(ErrorCode::kInvalidInput, "backend must be auto, divsufsort, CaPS or HIP");
  if (!HipBackendAvailable()) detail::Throw(ErrorCode::kUnsupportedBackend, "no HIP device is visible");
  const auto width = detail::ResolveSaCoordinateWidth(backend, options.coordinate_width, n);
  if (width != CoordinateWidth::kBits32 && width != CoordinateWidth::kBits64) {
    detail::Throw(ErrorCode::kInvalidInput, "construction width must be auto, 32 or 64 bits");
  }
  if (!detail::SaConstructionCanRepresent(backend, width, n)) {
    detail::Throw(ErrorCode::kInvalidInput, "text is longer than the chosen construction width can index");
  }
  const auto stored_width = detail::ResolveProfileStorageWidth(options.storage_width, options.resource_profile, n);
  const bool low_memory = options.resource_profile == SaResourceProfile::kLowMemory;
  if (low_memory && options.sampling_rate != 1) {
    detail::Throw(ErrorCode::kInvalidInput, "the Low-memory profile keeps every suffix (sampling rate 1)");
  }
  // The Low-memory profile is authoritative (suffix_array.cpp:4142-4151):
  // byte-coded LCP, and no ISA or CHILD stays resident.
  if (!low_memory &&
      (options.acceleration == SaAcceleration::kLcpChild || options.acceleration == SaAcceleration::kFull)) {
    detail::Throw(ErrorCode::kUnsupportedBackend, "CHILD arrays are not part of this build");
  }
  const bool wide_lcp = n - 1U > std::numeric_limits<std::uint32_t>::max();
  const bool retain_lcp = options.acceleration != SaAcceleration::kNone;
  const bool retain_isa = options.acceleration == SaAcceleration::kLcpSuffixLink;
  impl->construction_width = static_cast<std::uint8_t>(width);
  impl->backend = width == CoordinateWidth::kBits32 ? detail::StoredBackend::kHip32 : detail::StoredBackend::kHip64;
  impl->sampling_rate = options.sampling_rate;
  impl->acceleration = low_memory ? SaAcceleration::kLcp : options.acceleration;
  stats.text_prepare_seconds = Since(total_begin);
  if (low_memory) {
    notify("index-sa-lcp");
    BuildLowMemoryWithHip(*impl, stored_width, stats);
    impl->RefreshInfo();
    impl->info.sa_resource_profile = options.resource_profile;
    stats.total_seconds = Since(total_begin);
    notify("index-build-return");
    return SuffixArray(std::move(impl));
  }

  notify("index-sa-lcp");
  const bool narrow_sa = stored_width == CoordinateStorageWidth::kBits32;
  const std::uint64_t rows = 1 + (n - 1) / options.sampling_rate;
  const bool narrow_isa =
      detail::ResolveAuxiliaryStorageWidth(stored_width, rows) == CoordinateStorageWidth::kBits32;
  // The Fast prefix directory is computed by the same device build while
  // its SA is resident; its k comes from the core bytes this build will
  // hold (text, SA, ISA, raw LCP in their resident widths), checked below.
  std::uint32_t prefix_k = 0;
  if (options.resource_profile == SaResourceProfile::kFast && options.sampling_rate == 1 && retain_isa) {
    const std::uint64_t sa_row = stored_width == CoordinateStorageWidth::kBits32   ? 4
                                 : stored_width == CoordinateStorageWidth::kBits40 ? 5
                                 : stored_width == CoordinateStorageWidth::kBits48 ? 6
                                                                                   : 8;
    const std::uint64_t core = n + n * sa_row + n * (narrow_isa ? 4 : 8) +
                               (retain_lcp ? n * (wide_lcp ? 8 : 4) : 0);
    prefix_k = sfk_prefix_directory_select_k(n, core);
  }
  // Device outputs already in the resident widths: SA 32 or 64 (40/48 are
  // split on the host from 64), ISA in the auxiliary width, LCP 32/64.
  if (narrow_sa && narrow_isa && !wide_lcp) {
    BuildWithHip<std::uint32_t, std::uint32_t, std::uint32_t>(*impl, options, stored_width, retain_isa,
                                                              retain_lcp, prefix_k, stats, notify);
  } else if (narrow_sa && !wide_lcp) {
    BuildWithHip<std::uint32_t, std::uint64_t, std::uint32_t>(*impl, options, stored_width, retain_isa,
                                                              retain_lcp, prefix_k, stats, notify);
  } else if (!wide_lcp) {
    BuildWithHip<std::uint64_t, std::uint64_t, std::uint32_t>(*impl, options, stored_width, retain_isa,
                                                              retain_lcp, prefix_k, stats, notify);
  } else {
    BuildWithHip<std::uint64_t, std::uint64_t, std::uint64_t>(*impl, options, stored_width, retain_isa,
                                                              retain_lcp, prefix_k, stats, notify);
  }
  notify("index-prefix-directory");
  if (prefix_k != 0 && prefix_k != sfk_prefix_directory_select_k(n, impl->CoreBytes())) {
    detail::Throw(ErrorCode::kBuildFailure, "Fast prefix directory sized for different resident bytes");
  }
  impl->RefreshInfo();
  impl->info.sa_resource_profile = options.resource_profile;
  stats.total_seconds = Since(total_begin);
  notify("index-build-return");
  return SuffixArray(std::move(impl));
}

void SuffixArray::Save(const std::filesystem::path& path, const SaveOptions& options) const {
  detail::ContainerSpec spec;
  spec.format_minor = 4;
  spec.kind = 1;
  spec.backend = impl_->backend;
  spec.coordinate_width = impl_->construction_width;
  spec.profile = impl_->info.sa_resource_profile;
  spec.sequence_count = impl_->sequences.size();
  spec.total_bases = impl_->total_bases;
  spec.text_symbols = impl_->text.size();
  spec.ambiguous_bases = impl_->ambiguous_bases;
  spec.fingerprint = impl_->fingerprint;
  const Impl& impl = *impl_;
  std::vector<detail::SectionWriter> writers{
      {detail::SectionType::kMetadata,
       [&](std::ostream& out) { detail::WriteMetadataSection(out, impl.sequences); }},
      {detail::SectionType::kText,
       [&](std::ostream& out) {
         out.write(reinterpret_cast<const char*>(impl.text.data()), static_cast<std::streamsize>(impl.text.size()));
       }},
      {detail::SectionType::kSuffixArray,
       [&](std::ostream& out) { detail::WriteCoordinateSectionV14(out, impl.sa, impl.text.size()); }}};
  if (!impl.isa.Empty()) {
    writers.push_back({detail::SectionType::kInverseSuffixArray, [&](std::ostream& out) {
                         detail::WriteCoordinateSectionV14(out, impl.isa, impl.sa.Size());
                       }});
  }
  if (!impl.lcp.Empty()) {
    writers.push_back({detail::SectionType::kLcp, [&](std::ostream& out) {
                         detail::WriteLcpSectionV14(out, impl.lcp, impl.text.size());
                       }});
  }
  if (impl.sampling_rate > 1) {
    writers.push_back({detail::SectionType::kSaSampling, [&](std::ostream& out) {
                         detail::WriteU32(out, impl.sampling_rate);
                         detail::WriteU64(out, impl.sa.Size());
                       }});
  }
  detail::WriteContainer(path, options, spec, writers);
}

namespace {

struct ByteReader {
  const std::vector<std::uint8_t>& bytes;
  std::size_t pos = 0;
  void Need(std::size_t n, const char* what) const {
    if (bytes.size() - pos < n) detail::Throw(ErrorCode::kCorruptIndex, std::string(what) + " is truncated");
  }
  std::uint32_t U32(const char* what) {
    Need(4, what);
    std::uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= static_cast<std::uint32_t>(bytes[pos + i]) << (8 * i);
    pos += 4;
    return v;
  }
  std::uint64_t U64(const char* what) {
    Need(8, what);
    std::uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= static_cast<std::uint64_t>(bytes[pos + i]) << (8 * i);
    pos += 8;
    return v;
  }
  std::string Str(const char* what) {
    const auto len = U32(what);
    Need(len, what);
    std::string s(reinterpret_cast<const char*>(bytes.data() + pos), len);
    pos += len;
    return s;
  }
};

// Legacy (format 1.0-1.3) coordinate sections: u64 count [+ u8 width] +
// values (suffix_array.cpp:1778-1830).
detail::U64Values ReadLegacyValues(const std::vector<std::uint8_t>& bytes, bool has_width,
                                            std::uint8_t default_width, const char* what) {
  ByteReader r{bytes};
  const std::uint64_t count = r.U64(what);
  std::uint8_t width = default_width;
  if (has_width) {
    r.Need(1, what);
    width = bytes[r.pos++];
  }
  if (width != 32 && width != 64) detail::Throw(ErrorCode::kCorruptIndex, std::string(what) + " has an invalid width");
  const std::size_t w = width / 8;
  if ((bytes.size() - r.pos) != count * w) detail::Throw(ErrorCode::kCorruptIndex, std::string(what) + " has the wrong number of entries");
  detail::U64Values out(count);
  for (std::uint64_t i = 0; i < count; ++i) {
    std::uint64_t v = 0;
    for (std::size_t b = 0; b < w; ++b) v |= static_cast<std::uint64_t>(bytes[r.pos + i * w + b]) << (8 * b);
    out[i] = v;
  }
  return out;
}

}  // namespace

SuffixArray SuffixArray::Load(const std::filesystem::path& path) {
  const auto container = detail::ReadContainer(path);
  if (container.spec.kind != 1) detail::Throw(ErrorCode::kCorruptIndex, "file holds no suffix-array payload");
  const auto backend_id = static_cast<std::uint8_t>(container.spec.backend);
  if (backend_id < 1 || backend_id > 6) {
    detail::Throw(ErrorCode::kUnsupportedBackend, "file holds a payload kind this library does not read");
  }
  auto impl = std::make_unique<Impl>();
  impl->backend = container.spec.backend;
  impl->construction_width = container.spec.coordinate_width;
  impl->fingerprint = container.spec.fingerprint;

  // Metadata (serialization.cpp:801-870).
  {
    const auto bytes = detail::ReadSection(container, detail::SectionType::kMetadata);
    ByteReader r{bytes};
    const auto count = r.U32("metadata contig count");
    if (count == 0 || count != container.spec.sequence_count) {
      detail::Throw(ErrorCode::kCorruptIndex, "contig count of the metadata and of the header differ");
    }
    std::uint64_t expected_offset = 0;
    for (std::uint32_t i = 0; i < count; ++i) {
      SequenceInfo s;
      s.id = r.U32("sequence id");
      s.global_offset = r.U64("contig offset");
      s.length = r.U64("contig length");
      s.ambiguous_bases = r.U64("contig N count");
      s.name = r.Str("sequence name");
      s.description = r.Str("contig description");
      if (s.id != i || s.name.empty() || s.length == 0 || s.global_offset != expected_offset ||
          s.ambiguous_bases > s.length) {
        detail::Throw(ErrorCode::kCorruptIndex, "contig records are inconsistent");
      }
      expected_offset += s.length + 1;
      impl->total_bases += s.length;
      impl->ambiguous_bases += s.ambiguous_bases;
      impl->sequences.push_back(std::move(s));
    }
    if (r.pos != bytes.size()) detail::Throw(ErrorCode::kCorruptIndex, "bytes remain after the metadata records");
    if (impl->total_bases != container.spec.total_bases ||
        impl->ambiguous_bases != container.spec.ambiguous_bases ||
        expected_offset + 1 != container.spec.text_symbols) {
      detail::Throw(ErrorCode::kCorruptIndex, "metadata totals and header totals differ");
    }
  }
  // Text: sentinel, separators, fingerprint (suffix_array.cpp:4945-4969).
  impl->text = detail::ReadSection(container, detail::SectionType::kText);
  const std::uint64_t n = impl->text.size();
  if (n != container.spec.text_symbols || n < 2 || impl->text.back() != detail::kSentinel ||
      std::find(impl->text.begin(), impl->text.end() - 1, detail::kSentinel) != impl->text.end() - 1) {
    detail::Throw(ErrorCode::kCorruptIndex, "text section does not end with its only 0");
  }
  for (const auto& s : impl->sequences) {
    const auto sep = s.global_offset + s.length;
    if (sep >= n || impl->text[sep] != detail::kSeparator) {
      detail::Throw(ErrorCode::kCorruptIndex, "text section has separators where the contigs do not end");
    }
  }
  if (detail::ContentFingerprint(impl->text.data(), n - 1) != impl->fingerprint) {
    detail::Throw(ErrorCode::kCorruptIndex, "text section does not hash to the header's fingerprint");
  }
  if (const auto* sampling = container.Find(detail::SectionType::kSaSampling)) {
    (void)sampling;
    const auto bytes = detail::ReadSection(container, detail::SectionType::kSaSampling);
    ByteReader r{bytes};
    impl->sampling_rate = r.U32("SA sampling rate");
    const auto recorded = r.U64("kept suffix count");
    if (impl->sampling_rate <= 1 || recorded != 1 + (n - 1) / impl->sampling_rate || r.pos != bytes.size()) {
      detail::Throw(ErrorCode::kCorruptIndex, "sampling record is inconsistent");
    }
  }
  const std::uint32_t k = impl->sampling_rate;
  const std::uint64_t count = 1 + (n - 1) / k;
  const bool backend_is_32 = backend_id == 1 || backend_id == 3 || backend_id == 5;
  if (container.spec.coordinate_width != (backend_is_32 ? 32 : 64)) {
    detail::Throw(ErrorCode::kCorruptIndex, "stored backend does not match the construction width");
  }
  const bool v14 = container.spec.format_minor >= 4;
  // SA section + sampled-permutation check (suffix_array.cpp:4999-5010).
  {
    const auto& d = container.Require(detail::SectionType::kSuffixArray);
    if (v14) {
      std::ifstream in(path, std::ios::binary);
      in.seekg(static_cast<std::streamoff>(d.offset));
      impl->sa = detail::ReadCoordinateSectionV14(in, d.size, count, n, "suffix array");
    } else {
      auto values = ReadLegacyValues(detail::ReadSection(container, detail::SectionType::kSuffixArray), false,
                                     container.spec.coordinate_width, "suffix array");
      if (values.size() != count) detail::Throw(ErrorCode::kCorruptIndex, "suffix-array length mismatch");
      impl->sa = detail::CoordinateStorage::FromUInt64(
          std::move(values), backend_is_32 ? CoordinateStorageWidth::kBits32 : CoordinateStorageWidth::kBits64, n,
          ErrorCode::kCorruptIndex, "suffix array");
    }
    std::vector<bool> seen(count, false);
    for (std::uint64_t r = 0; r < count; ++r) {
      const auto v = impl->sa.At(r);
      if (v >= n || v % k != 0) detail::Throw(ErrorCode::kCorruptIndex, "a stored suffix is not at a multiple of the sampling rate");
      if (seen[v / k]) detail::Throw(ErrorCode::kCorruptIndex, "stored suffixes repeat a position");
      seen[v / k] = true;
    }
  }
  const auto* isa_d = container.Find(detail::SectionType::kInverseSuffixArray);
  const auto* lcp_d = container.Find(detail::SectionType::kLcp);
  if (isa_d != nullptr) {
    if (v14) {
      std::ifstream in(path, std::ios::binary);
      in.seekg(static_cast<std::streamoff>(isa_d->offset));
      impl->isa = detail::ReadCoordinateSectionV14(in, isa_d->size, count, count, "inverse suffix array");
    } else {
      auto values = ReadLegacyValues(detail::ReadSection(container, detail::SectionType::kInverseSuffixArray),
                                     true, 64, "inverse suffix array");
      if (values.size() != count) detail::Throw(ErrorCode::kCorruptIndex, "ISA length mismatch");
      impl->isa = detail::CoordinateStorage::FromUInt64(std::move(values), CoordinateStorageWidth::kAutoSelect,
                                                        count, ErrorCode::kCorruptIndex, "inverse suffix array");
    }
    for (std::uint64_t r = 0; r < count; ++r) {  // suffix_array.cpp:5120-5131
      if (impl->isa.At(impl->sa.At(r) / k) != r) {
        detail::Throw(ErrorCode::kCorruptIndex, "inverse suffix array does not invert the suffix array");
      }
    }
  }
  if (lcp_d != nullptr) {
    if (v14) {
      std::ifstream in(path, std::ios::binary);
      in.seekg(static_cast<std::streamoff>(lcp_d->offset));
      impl->lcp = detail::ReadLcpSectionV14(in, lcp_d->size, count, k, n);
    } else {
      auto values = ReadLegacyValues(detail::ReadSection(container, detail::SectionType::kLcp), true, 64, "LCP");
      if (values.size() != count) detail::Throw(ErrorCode::kCorruptIndex, "LCP length mismatch");
      impl->lcp = detail::LcpStorage::FromRaw64(std::move(values), k, ErrorCode::kCorruptIndex);
    }
    impl->lcp.Validate(impl->sa, n, ErrorCode::kCorruptIndex);
  }
  impl->acceleration = lcp_d == nullptr ? SaAcceleration::kNone
                                        : (isa_d != nullptr ? SaAcceleration::kLcpSuffixLink : SaAcceleration::kLcp);
  impl->RefreshInfo();
  impl->info.sa_resource_profile = container.spec.profile;
  impl->info.format_version = "1." + std::to_string(container.spec.format_minor);
  impl->info.library_version = std::to_string(container.spec.library_major) + "." +
                               std::to_string(container.spec.library_minor) + "." +
                               std::to_string(container.spec.library_patch);
  impl->info.serialized_bytes = container.file_size;
  if (impl->WantsPrefixDirectory(container.spec.profile) && HipBackendAvailable()) {
    impl->BuildPrefixDirectoryOnDevice();
    impl->RefreshInfo();
    impl->info.sa_resource_profile = container.spec.profile;
    impl->info.format_version = "1." + std::to_string(container.spec.format_minor);
    impl->info.library_version = std::to_string(container.spec.library_major) + "." +
                                 std::to_string(container.spec.library_minor) + "." +
                                 std::to_string(container.spec.library_patch);
    impl->info.serialized_bytes = container.file_size;
  }
  return SuffixArray(std::move(impl));
}

void SuffixArray::Impl::BuildPrefixDirectoryOnDevice() {
  const std::uint64_t n = text.size();
  const std::uint32_t k = sfk_prefix_directory_select_k(n, CoreBytes());
  prefix_k = 0;
  prefix_table.clear();
  if (k == 0) return;
  const std::uint32_t bits = n - 1 <= std::numeric_limits<std::uint32_t>::max() ? 32 : 64;
  std::vector<std::uint64_t> sa64;
  std::vector<std::uint32_t> sa32;
  if (bits == 32) {
    sa32.resize(n);
    for (std::uint64_t r = 0; r < n; ++r) sa32[r] = static_cast<std::uint32_t>(sa.At(r));
  } else {
    sa64.resize(n);
    for (std::uint64_t r = 0; r < n; ++r) sa64[r] = sa.At(r);
  }
  const std::uint64_t words = sfk_prefix_directory_words(k, n);
  std::vector<std::uint64_t> table(words);
  sfk_hip_builder* b = nullptr;
  void *d_text = nullptr, *d_sa = nullptr, *d_table = nullptr;
  char err[512] = {0};
  int status = sfk_hip_builder_create(0, &b, err, sizeof(err));
  if (status == SFK_HIP_OK) status = sfk_hip_malloc(0, n + 64, &d_text);
  if (status == SFK_HIP_OK) status = sfk_hip_malloc(0, n * (bits / 8), &d_sa);
  if (status == SFK_HIP_OK) status = sfk_hip_malloc(0, words * 8, &d_table);
  if (status == SFK_HIP_OK) status = sfk_hip_memcpy_h2d(d_text, text.data(), n);
  if (status == SFK_HIP_OK) {
    status = sfk_hip_memcpy_h2d(d_sa, bits == 32 ? static_cast<const void*>(sa32.data()) : sa64.data(),
                                n * (bits / 8));
  }
  std::uint64_t counts[2] = {0, 0};
  if (status == SFK_HIP_OK) {
    status = sfk_hip_prefix_directory(b, static_cast<const std::uint8_t*>(d_text), n, d_sa, bits, k,
                                      static_cast<std::uint64_t*>(d_table), counts, err, sizeof(err));
  }
  if (status == SFK_HIP_OK) status = sfk_hip_memcpy_d2h(table.data(), d_table, words * 8);
  for (void* p : {d_text, d_sa, d_table}) {
    if (p != nullptr) sfk_hip_free(p);
  }
  sfk_hip_builder_destroy(b);
  if (status != SFK_HIP_OK) detail::Throw(FromStatus(status), std::string("Fast prefix directory: ") + err);
  prefix_k = k;
  prefix_table = std::move(table);
}

const IndexInfo& SuffixArray::Info() const noexcept { return impl_->info; }
std::uint64_t SuffixArray::Size() const noexcept { return impl_->sa.Size(); }
std::uint64_t SuffixArray::TextSymbols() const noexcept { return impl_->text.size(); }
Position SuffixArray::SuffixAt(std::uint64_t row) const {
  if (row >= impl_->sa.Size()) detail::Throw(ErrorCode::kInvalidInput, "suffix-array row out of range");
  return impl_->sa.At(row);
}
bool SuffixArray::HasIsa() const noexcept { return !impl_->isa.Empty(); }
bool SuffixArray::HasLcp() const noexcept { return !impl_->lcp.Empty(); }
std::uint64_t SuffixArray::IsaAt(std::uint64_t sample) const {
  if (sample >= impl_->isa.Size()) detail::Throw(ErrorCode::kInvalidInput, "ISA index out of range");
  return impl_->isa.At(sample);
}
std::uint64_t SuffixArray::LcpAt(std::uint64_t row) const {
  if (row >= impl_->lcp.Size()) detail::Throw(ErrorCode::kInvalidInput, "LCP row out of range");
  return impl_->lcp.Decode(row, impl_->lcp.ByteCoded() ? impl_->sa.At(row) : 0, ErrorCode::kInvalidInput);
}
const std::vector<std::uint8_t>& SuffixArray::Text() const noexcept { return impl_->text; }

std::uint32_t SuffixArray::PrefixDirectoryK() const noexcept { return impl_->prefix_k; }

std::pair<std::uint64_t, std::uint64_t> SuffixArray::PrefixDirectoryInterval(std::uint64_t key) const {
  const std::uint32_t k = impl_->prefix_k;
  if (k == 0 || key >= (std::uint64_t{1} << (2 * k))) {
    detail::Throw(ErrorCode::kInvalidInput, "no prefix directory entry for this key");
  }
  const auto& table = impl_->prefix_table;
  if (impl_->text.size() - 1 <= std::numeric_limits<std::uint32_t>::max()) {
    return {table[key] & 0xFFFFFFFFull, table[key] >> 32};
  }
  return {table[key], table[(std::uint64_t{1} << (2 * k)) + key]};
}

std::uint64_t SuffixArray::Count(std::string_view pattern) const {
  if (pattern.empty()) detail::Throw(ErrorCode::kInvalidInput, "an empty pattern was given");
  if (impl_->sampling_rate != 1) {
    detail::Throw(ErrorCode::kUnsupportedBackend, "count over a sampled suffix array is not part of this build");
  }
  std::vector<std::uint8_t> p(pattern.size());
  for (std::size_t i = 0; i < pattern.size(); ++i) {
    switch (pattern[i]) {
      case 'A': case 'a': p[i] = detail::kA; break;
      case 'C': case 'c': p[i] = detail::kC; break;
      case 'G': case 'g': p[i] = detail::kG; break;
      case 'T': case 't': p[i] = detail::kT; break;
      default: detail::Throw(ErrorCode::kInvalidInput, "patterns may hold only A, C, G and T");
    }
  }
  const auto& text = impl_->text;
  const std::uint64_t n = text.size();
  // compare suffix at position s with the pattern: <0, 0 (prefix match), >0
  auto cmp = [&](std::uint64_t s) {
    for (std::size_t i = 0; i < p.size(); ++i) {
      if (s + i >= n) return -1;
      if (text[s + i] != p[i]) return text[s + i] < p[i] ? -1 : 1;
    }
    return 0;
  };
  std::uint64_t lo = 0, hi = impl_->sa.Size();
  // The Fast prefix directory gives the interval of the first k symbols
  // (FastPrefixRange, suffix_array.cpp:2173-2190); an absent k-mer counts 0.
  const std::uint32_t k = impl_->prefix_k;
  if (k > 0 && p.size() >= k) {
    std::uint64_t key = 0;
    for (std::uint32_t i = 0; i < k; ++i) key = (key << 2) | static_cast<std::uint64_t>(p[i] - detail::kA);
    const auto& table = impl_->prefix_table;
    if (n - 1 <= std::numeric_limits<std::uint32_t>::max()) {
      lo = table[key] & 0xFFFFFFFFull;
      hi = table[key] >> 32;
    } else {
      lo = table[key];
      hi = table[(std::uint64_t{1} << (2 * k)) + key];
    }
    if (lo >= impl_->sa.Size()) return 0;
  }
  const auto range_end = hi;
  while (lo < hi) {  // first row with suffix >= pattern
    const auto mid = lo + (hi - lo) / 2;
    if (cmp(impl_->sa.At(mid)) < 0) lo = mid + 1; else hi = mid;
  }
  const auto first = lo;
  hi = range_end;
  while (lo < hi) {  // first row with suffix > pattern (prefix matches count as equal)
    const auto mid = lo + (hi - lo) / 2;
    if (cmp(impl_->sa.At(mid)) <= 0) lo = mid + 1; else hi = mid;
  }
  return lo - first;
}

}  // namespace sufkit
