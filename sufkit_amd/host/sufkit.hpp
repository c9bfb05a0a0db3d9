// sufkit_amd host API: the reference's SuffixArray build/persist surface
// (include/sufkit/types.hpp, suffix_array.hpp, genome_reference.hpp) with a
// third construction backend, SaBackend::kHip, that runs on one MI355X
// through the C ABI of include/sufkit_hip.h.
//
// Names, argument meaning and error categories follow the reference:
//   GenomeReference::FromRecords/FromFasta   genome_reference.cpp:222-281
//   SuffixArray::Build                        suffix_array.cpp:4086-4390
//   SuffixArray::Save / Load                  suffix_array.cpp:4853-5244
//   SuffixArrayBuildOptions / statistics      types.hpp:244-301
//   Error / ErrorCode                         types.hpp:708-741
#pragma once

#include <cstdint>
#include <filesystem>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace sufkit {

using Position = std::uint64_t;
using SequenceId = std::uint32_t;

enum class ErrorCode : std::uint8_t {
  kInvalidInput,
  kIoError,
  kUnsupportedBackend,
  kCorruptIndex,
  kVersionMismatch,
  kBuildFailure
};

class Error : public std::runtime_error {
 public:
  Error(ErrorCode code, const std::string& message)
      : std::runtime_error(message), code_(code) {}
  ErrorCode Code() const noexcept { return code_; }

 private:
  ErrorCode code_;
};

// types.hpp:75-85 plus kHip.  kDivsufsort and kCaps are accepted selectors
// whose engines are not part of this build (kUnsupportedBackend, as a
// reference build with a backend compiled out reports).
enum class SaBackend : std::uint8_t { kAutoSelect = 0, kDivsufsort = 1, kCaps = 2, kHip = 3 };
enum class CoordinateWidth : std::uint8_t { kAutoSelect = 0, kBits32 = 32, kBits64 = 64 };
enum class CoordinateStorageWidth : std::uint8_t {
  kAutoSelect = 0, kBits32 = 32, kBits40 = 40, kBits48 = 48, kBits64 = 64
};
enum class SaResourceProfile : std::uint8_t { kFast = 0, kLowMemory = 1 };
enum class SaAcceleration : std::uint8_t {
  kNone = 0, kLcp = 1, kLcpChild = 2, kLcpSuffixLink = 3, kFull = 4
};
enum class SaLcpEncoding : std::uint8_t { kNone = 0, kRaw = 1, kByteCoded = 2 };

struct SequenceRecord {
  std::string name;
  std::string description;
  std::string sequence;
};

struct SequenceInfo {
  SequenceId id = 0;
  std::string name;
  std::string description;
  Position length = 0;
  Position global_offset = 0;
  Position ambiguous_bases = 0;
};

struct SuffixArrayBuildStatistics {
  double sa_seconds = 0.0;
  double storage_compaction_seconds = 0.0;
  double isa_seconds = 0.0;
  double lcp_seconds = 0.0;
  double text_prepare_seconds = 0.0;
  double lcp_finalize_seconds = 0.0;
  double total_seconds = 0.0;
  double prefix_directory_seconds = 0.0;
  // Device-side detail of the HIP backend.
  double device_seconds = 0.0;
  double transfer_seconds = 0.0;
  double call_seconds = 0.0;  // the backend call: allocation, copies, build
};

struct SuffixArrayBuildOptions {
  SaBackend backend = SaBackend::kAutoSelect;
  CoordinateWidth coordinate_width = CoordinateWidth::kAutoSelect;
  std::uint32_t threads = 1;
  std::uint32_t sampling_rate = 1;
  SaAcceleration acceleration = SaAcceleration::kLcpSuffixLink;
  SuffixArrayBuildStatistics* statistics = nullptr;
  CoordinateStorageWidth storage_width = CoordinateStorageWidth::kAutoSelect;
  SaResourceProfile resource_profile = SaResourceProfile::kFast;
  void (*stage_callback)(const char* stage, void* context) = nullptr;
  void* stage_context = nullptr;
};

SuffixArrayBuildOptions FastSuffixArrayBuildOptions();

struct SaveOptions {
  bool overwrite = false;
};

struct IndexInfo {
  std::string format_version = "1.4";
  std::string library_version;
  std::string backend;
  std::string backend_signature;
  std::uint8_t coordinate_width = 0;
  std::uint8_t stored_coordinate_width = 0;
  std::uint64_t sequence_count = 0;
  std::uint64_t total_bases = 0;
  std::uint64_t text_symbols = 0;
  std::uint64_t suffix_count = 0;
  std::uint32_t sa_sampling_rate = 1;
  std::uint64_t ambiguous_bases = 0;
  std::uint64_t fingerprint = 0;
  std::uint64_t serialized_bytes = 0;
  SaAcceleration sa_acceleration = SaAcceleration::kNone;
  SaResourceProfile sa_resource_profile = SaResourceProfile::kFast;
  SaLcpEncoding lcp_encoding = SaLcpEncoding::kNone;
  std::uint64_t text_bytes = 0;
  std::uint64_t sa_bytes = 0;
  std::uint64_t isa_bytes = 0;
  std::uint64_t lcp_bytes = 0;
  std::uint64_t auxiliary_bytes = 0;  // the Fast prefix directory
  std::uint64_t resident_core_bytes = 0;
};

class GenomeReference {
 public:
  static GenomeReference FromFasta(const std::filesystem::path& path);
  static GenomeReference FromRecords(std::vector<SequenceRecord> records);
  // Encoded bytes produced elsewhere (one contig per length, separators
  // excluded), validated like FromRecords; names are "contig<i>".
  static GenomeReference FromEncoded(const std::uint8_t* bases, const std::uint64_t* lengths,
                                     std::uint64_t count);

  GenomeReference(GenomeReference&&) noexcept;
  GenomeReference& operator=(GenomeReference&&) noexcept;
  ~GenomeReference();

  std::uint64_t SequenceCount() const noexcept;
  std::uint64_t TotalBases() const noexcept;
  std::uint64_t AmbiguousBases() const noexcept;
  std::uint64_t Fingerprint() const noexcept;
  const std::vector<SequenceInfo>& Sequences() const noexcept;
  // Encoded bytes (bases + separators), without the sentinel.
  const std::vector<std::uint8_t>& Encoded() const noexcept;

  struct Impl;

 private:
  explicit GenomeReference(std::unique_ptr<Impl> impl);
  std::unique_ptr<Impl> impl_;
  friend class SuffixArray;
};

class SuffixArray {
 public:
  static SuffixArray Build(const GenomeReference& reference,
                           const SuffixArrayBuildOptions& options = {});
  static SuffixArray Load(const std::filesystem::path& path);
  void Save(const std::filesystem::path& path, const SaveOptions& options = {}) const;

  SuffixArray(SuffixArray&&) noexcept;
  SuffixArray& operator=(SuffixArray&&) noexcept;
  ~SuffixArray();

  const IndexInfo& Info() const noexcept;
  std::uint64_t Size() const noexcept;          // stored suffix rows
  std::uint64_t TextSymbols() const noexcept;
  Position SuffixAt(std::uint64_t row) const;   // SA[row]
  bool HasIsa() const noexcept;
  bool HasLcp() const noexcept;
  std::uint64_t IsaAt(std::uint64_t sample) const;
  std::uint64_t LcpAt(std::uint64_t row) const;
  const std::vector<std::uint8_t>& Text() const noexcept;  // incl. sentinel
  // Exact occurrence count of an A/C/G/T pattern (two binary searches,
  // inside the prefix directory's interval when the pattern is at least k
  // long; suffix_array.cpp:4413-4430 semantics: patterns with other
  // symbols match nothing).
  std::uint64_t Count(std::string_view pattern) const;
  // k of the Fast prefix directory (0: none), and its SA interval [begin,
  // end) for a k-mer key (2-bit codes, first symbol most significant).
  std::uint32_t PrefixDirectoryK() const noexcept;
  std::pair<std::uint64_t, std::uint64_t> PrefixDirectoryInterval(std::uint64_t key) const;

  struct Impl;

 private:
  explicit SuffixArray(std::unique_ptr<Impl> impl);
  std::unique_ptr<Impl> impl_;
};

// ----------------------------------------------- saving from device memory
//
// The arrays of a device build (sfk_hip_build_device) held in HIP device
// memory, with the device text they were built from (encoded ++ sentinel).
// Value widths are those of the device buffers (32 or 64 bits).
struct DeviceIndexBuffers {
  int device = 0;
  const void* text = nullptr;
  const void* sa = nullptr;
  std::uint32_t sa_bits = 32;
  const void* isa = nullptr;  // required when the options keep the ISA
  std::uint32_t isa_bits = 32;
  const void* lcp = nullptr;  // required when the options keep the LCP
  std::uint32_t lcp_bits = 32;
  CoordinateWidth construction_width = CoordinateWidth::kBits32;
};

struct DeviceSaveStatistics {
  double seconds = 0.0;           // whole call
  double write_seconds = 0.0;     // sections and header written
  double validate_seconds = 0.0;  // the file reread with every CRC checked
  std::uint64_t bytes = 0;
};

// Writes the .sufidx file that SuffixArray::Build(reference, options)
// followed by Save(path, save) writes, byte for byte, from the device
// arrays: each section is streamed from HBM through pinned buffers with its
// CRC32 computed on the device, then the file is reread and checked and
// published as Save does (suffix_array.cpp:4853-4916,
// serialization.cpp:529-619).  Sampling rate 1 only.  A Low-memory index
// needs the device ISA and raw LCP too: its byte-coded LCP is made from them.
void SaveDeviceIndex(const GenomeReference& reference, const DeviceIndexBuffers& buffers,
                     const std::filesystem::path& path, const SaveOptions& save = {},
                     const SuffixArrayBuildOptions& options = FastSuffixArrayBuildOptions(),
                     DeviceSaveStatistics* statistics = nullptr);

// ------------------------------------------------------------ FM index
//
// The reference's FmIndex surface (include/sufkit/fm_index.hpp,
// types.hpp:200-229,564-610) over an FM-index that this build constructs on
// the device from the GPU-built SA (BWT membership masks with per-block base
// counts, kept resident) instead of an SDSL CSA on the host.  Queries return
// what the reference returns: the SA of text ++ sentinel is unique, so row
// ranges, counts and located coordinates are those of the SDSL index.

// types.hpp:200-209.  The three SDSL types select the same device index
// here (their query results are identical); kSdslCsaSada stays unavailable.
enum class FmBackend : std::uint8_t {
  kSdslCsaWtHuff = 1,
  kSdslCsaWtBalanced = 2,
  kSdslCsaSada = 3,
  kSdslCsaWtEpr = 4
};
enum class StrandMode : std::uint8_t { kForward = 0, kReverseComplement = 1, kBoth = 2 };
enum class Strand : std::uint8_t { kForward = 0, kReverseComplement = 1, kBoth = 2 };

struct FmIndexBuildOptions {
  FmBackend backend = FmBackend::kSdslCsaWtHuff;
};

struct FmBatchOptions {
  StrandMode strands = StrandMode::kForward;
  std::uint32_t batch_width = 0;  // 0 selects 16; explicit 1..256
};

struct LocateOptions {
  StrandMode strands = StrandMode::kForward;
  std::optional<std::uint64_t> max_hits;  // absent: every hit
};

struct SuffixRange {
  std::uint64_t begin = 0;
  std::uint64_t end = 0;
  std::uint64_t Size() const noexcept { return end - begin; }
  bool Empty() const noexcept { return begin == end; }
};

struct Match {
  SequenceId sequence_id = 0;
  Position position = 0;
  std::uint64_t length = 0;
  Strand strand = Strand::kForward;
};

struct QueryResult {
  std::uint64_t total_hits = 0;
  std::vector<Match> hits;
  bool truncated = false;
};

class FmIndex {
 public:
  static FmIndex Build(const GenomeReference& reference, const FmIndexBuildOptions& options = {});
  // The SDSL payload of an FM .sufidx is not produced by this build.
  static FmIndex Load(const std::filesystem::path& path);
  void Save(const std::filesystem::path& path, const SaveOptions& options = {}) const;

  FmIndex(FmIndex&&) noexcept;
  FmIndex& operator=(FmIndex&&) noexcept;
  ~FmIndex();

  SuffixRange EqualRange(std::string_view pattern) const;
  std::vector<SuffixRange> EqualRangeBatch(const std::vector<std::string_view>& patterns,
                                           std::uint32_t batch_width = 0) const;
  std::uint64_t Count(std::string_view pattern, StrandMode strands = StrandMode::kForward) const;
  std::vector<std::uint64_t> CountBatch(const std::vector<std::string_view>& patterns,
                                        const FmBatchOptions& options = {}) const;
  QueryResult Locate(std::string_view pattern, const LocateOptions& options = {}) const;
  SequenceInfo GetSequenceInfo(SequenceId id) const;
  const IndexInfo& GetInfo() const noexcept;
  // k of the device index's prefix directory (0: none) and its FM bytes.
  std::uint32_t PrefixDirectoryK() const noexcept;
  std::uint64_t DeviceBytes() const noexcept;

  struct Impl;

 private:
  explicit FmIndex(std::unique_ptr<Impl> impl);
  std::unique_ptr<Impl> impl_;
};

// Stored backend names of the .sufidx header byte 14
// (serialization.hpp:25-34 plus the two HIP identities).
const char* StoredBackendName(std::uint8_t id) noexcept;
const char* StoredBackendSignature(std::uint8_t id) noexcept;
bool HipBackendAvailable() noexcept;

}  // namespace sufkit
