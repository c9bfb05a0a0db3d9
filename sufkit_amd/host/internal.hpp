// Internal building blocks of the host library: symbol codes, width
// policies, coordinate/LCP storage, the .sufidx 1.4 section codecs and the
// CRC-protected container.
#pragma once

#include <array>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <iosfwd>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

#include "sufkit.hpp"

namespace sufkit::detail {

// Allocator that leaves the elements of a vector sized without values
// uninitialised: the coordinate and LCP vectors a build fills from the
// device are not zero-filled first (tens of gigabytes of single-threaded
// stores); their pages are first touched by the multi-threaded copy that
// fills them.  Construction with a value still stores it.
template <class T>
struct NoInitAllocator : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAllocator<U>;
  };
  NoInitAllocator() noexcept = default;
  template <class U>
  NoInitAllocator(const NoInitAllocator<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible_v<U>) {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) {
    ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};
template <class T, class U>
bool operator==(const NoInitAllocator<T>&, const NoInitAllocator<U>&) noexcept { return true; }
template <class T, class U>
bool operator!=(const NoInitAllocator<T>&, const NoInitAllocator<U>&) noexcept { return false; }

// Coordinates and raw LCP values (see NoInitAllocator).
template <class T>
using ValueVector = std::vector<T, NoInitAllocator<T>>;
using U32Values = ValueVector<std::uint32_t>;
using U64Values = ValueVector<std::uint64_t>;

// reference_data.hpp:20-27
constexpr std::uint8_t kSentinel = 0;
constexpr std::uint8_t kSeparator = 1;
constexpr std::uint8_t kA = 2;
constexpr std::uint8_t kC = 3;
constexpr std::uint8_t kG = 4;
constexpr std::uint8_t kT = 5;
constexpr std::uint8_t kN = 6;
constexpr std::uint8_t kNormalizationId = 1;

constexpr std::uint64_t kMaxCoordinate40 = (1ULL << 40) - 1;
constexpr std::uint64_t kMaxCoordinate48 = (1ULL << 48) - 1;

// Library version written into .sufidx headers: the reference format
// version 0.3.0 this writer follows (version.hpp:10-16).
constexpr std::uint8_t kVersionMajor = 0;
constexpr std::uint8_t kVersionMinor = 3;
constexpr std::uint8_t kVersionPatch = 0;

// Stored backend IDs (serialization.hpp:25-34); 5 and 6 are the HIP builds.
enum class StoredBackend : std::uint8_t {
  kDivsufsort32 = 1,
  kDivsufsort64 = 2,
  kCaps32 = 3,
  kCaps64 = 4,
  kHip32 = 5,
  kHip64 = 6,
};

enum class SectionType : std::uint32_t {
  kMetadata = 1,
  kText = 2,
  kSuffixArray = 3,
  kSdslCsa = 4,
  kInverseSuffixArray = 5,
  kLcp = 6,
  kChild = 7,
  kLearnedSa = 8,
  kSaSampling = 9,
};

[[noreturn]] void Throw(ErrorCode code, const std::string& message);

std::uint64_t ContentFingerprint(const std::uint8_t* data, std::size_t size);

// ------------------------------------------------------------- policies ---
// caps_backend.hpp:25-81, suffix_array.cpp:806-831, coordinate_storage.cpp:
// 363-417, with the HIP backend's limits (32-bit construction up to
// UINT32_MAX symbols, like CaPS-SA's uint32_t).
SaBackend ResolveSaBackend(SaBackend requested, std::uint64_t text_symbols,
                           std::uint32_t threads, bool hip_available) noexcept;
CoordinateWidth ResolveSaCoordinateWidth(SaBackend backend, CoordinateWidth requested,
                                         std::uint64_t text_symbols) noexcept;
std::uint64_t MaximumSaConstructionTextSymbols(SaBackend backend,
                                               CoordinateWidth width) noexcept;
bool SaConstructionCanRepresent(SaBackend backend, CoordinateWidth width,
                                std::uint64_t text_symbols) noexcept;
CoordinateStorageWidth SelectCoordinateStorageWidth(std::uint64_t symbol_count,
                                                    ErrorCode error_code = ErrorCode::kInvalidInput);
CoordinateStorageWidth ResolveCoordinateStorageWidth(CoordinateStorageWidth requested,
                                                     std::uint64_t symbol_count,
                                                     ErrorCode error_code = ErrorCode::kInvalidInput);
std::uint64_t MaxCoordinateForWidth(CoordinateStorageWidth width,
                                    ErrorCode error_code = ErrorCode::kInvalidInput);
CoordinateStorageWidth ResolveProfileStorageWidth(CoordinateStorageWidth requested,
                                                  SaResourceProfile profile,
                                                  std::uint64_t symbol_count);
CoordinateStorageWidth ResolveAuxiliaryStorageWidth(CoordinateStorageWidth preferred,
                                                    std::uint64_t symbol_count);

// ----------------------------------------------------- coordinate storage ---
struct Coordinate40 {
  std::vector<std::uint32_t> low;
  std::vector<std::uint8_t> high;
};
struct Coordinate48 {
  std::vector<std::uint32_t> low;
  std::vector<std::uint16_t> high;
};

// Immutable coordinates in one of four physical widths; every value is
// range-checked against the symbol domain when built (coordinate_storage.cpp:
// 439-598 semantics).
class CoordinateStorage {
 public:
  using Storage = std::variant<std::monostate, U32Values, U64Values, Coordinate40, Coordinate48>;
  CoordinateStorage() = default;
  static CoordinateStorage FromUInt32(U32Values&& values,
                                      CoordinateStorageWidth width, std::uint64_t symbol_count,
                                      ErrorCode error_code, const char* label);
  static CoordinateStorage FromUInt64(U64Values&& values,
                                      CoordinateStorageWidth width, std::uint64_t symbol_count,
                                      ErrorCode error_code, const char* label);
  static CoordinateStorage FromPacked40(Coordinate40&& values, std::uint64_t symbol_count,
                                        ErrorCode error_code, const char* label);
  static CoordinateStorage FromPacked48(Coordinate48&& values, std::uint64_t symbol_count,
                                        ErrorCode error_code, const char* label);
  bool Empty() const noexcept { return storage_.index() == 0; }
  std::uint64_t Size() const noexcept;
  CoordinateStorageWidth Width() const noexcept { return width_; }
  std::uint64_t Bytes() const noexcept;
  std::uint64_t At(std::uint64_t index) const noexcept;
  const Storage& Raw() const noexcept { return storage_; }

 private:
  Storage storage_;
  CoordinateStorageWidth width_ = CoordinateStorageWidth::kAutoSelect;
};

// Raw LCP (the Fast profile's layout): one 32- or 64-bit value per row,
// LCP[0] == 0 (lcp_storage.cpp:508-528).  Byte-coded LCP (the Low-memory
// profile's layout, lcp_storage.cpp:55-230): one byte per row, min(LCP,
// 255); a 255 is decoded from the overflow anchors (text position, PLCP) by
// the suffix position of its row, through a guide of the first anchor of
// every 4096-position block.  Width() is the value width of a raw LCP and
// the anchor coordinate width of a byte-coded one.
class LcpStorage {
 public:
  static constexpr std::uint32_t kByteLimit = 255;
  static constexpr std::uint32_t kGuideBlockShift = 12;

  LcpStorage() = default;
  static LcpStorage FromRaw32(U32Values&& values, std::uint32_t sampling_rate,
                              ErrorCode error_code);
  static LcpStorage FromRaw64(U64Values&& values, std::uint32_t sampling_rate,
                              ErrorCode error_code);
  // Validates as ValidateByteStructure (lcp_storage.cpp:97-150) and builds
  // the guide (BuildGuide, :55-95).
  static LcpStorage FromByteCoded(std::vector<std::uint8_t>&& primary, std::vector<std::uint64_t>&& positions,
                                  std::vector<std::uint64_t>&& values, std::uint32_t sampling_rate,
                                  std::uint64_t text_symbols, std::uint8_t coordinate_width, ErrorCode error_code);
  bool Empty() const noexcept { return width_ == 0; }
  bool ByteCoded() const noexcept { return byte_coded_; }
  std::uint64_t Size() const noexcept;
  std::uint8_t Width() const noexcept { return width_; }
  std::uint32_t SamplingRate() const noexcept { return sampling_rate_; }
  // Raw LCP only.
  std::uint64_t At(std::uint64_t row) const noexcept;
  // Either layout: the value of `row`, whose suffix starts at
  // suffix_position (DecodeByte, lcp_storage.cpp:186-215).
  std::uint64_t Decode(std::uint64_t row, std::uint64_t suffix_position, ErrorCode error_code) const;
  std::uint64_t Bytes() const noexcept;
  const U32Values& Raw32() const noexcept { return raw32_; }
  const U64Values& Raw64() const noexcept { return raw64_; }
  const std::vector<std::uint8_t>& Primary() const noexcept { return primary_; }
  const std::vector<std::uint64_t>& AnchorPositions() const noexcept { return anchor_positions_; }
  const std::vector<std::uint64_t>& AnchorValues() const noexcept { return anchor_values_; }
  // lcp_storage.cpp:693-770: row count, LCP[0]==0, bounds.
  void Validate(const CoordinateStorage& sa, std::uint64_t text_symbols,
                ErrorCode error_code) const;

 private:
  U32Values raw32_;
  U64Values raw64_;
  std::vector<std::uint8_t> primary_;
  std::vector<std::uint64_t> anchor_positions_;
  std::vector<std::uint64_t> anchor_values_;
  std::vector<std::uint64_t> guide_;
  bool byte_coded_ = false;
  std::uint8_t width_ = 0;
  std::uint32_t sampling_rate_ = 1;
};

// ------------------------------------------------------------- codecs ---
void WriteCoordinateSectionV14(std::ostream& out, const CoordinateStorage& storage,
                               std::uint64_t symbol_count);
CoordinateStorage ReadCoordinateSectionV14(std::istream& in, std::uint64_t section_size,
                                           std::uint64_t expected_elements,
                                           std::uint64_t expected_symbols, const char* label);
void WriteLcpSectionV14(std::ostream& out, const LcpStorage& lcp, std::uint64_t text_symbols);
// The codec headers those sections start with (SKC1 56 bytes, SKL1 80 bytes).
std::vector<std::uint8_t> CoordinateSectionHeader(CoordinateStorageWidth width, std::uint64_t symbol_count,
                                                  std::uint64_t count);
std::vector<std::uint8_t> LcpSectionHeader(std::uint32_t width, std::uint32_t sampling_rate,
                                           std::uint64_t text_symbols, std::uint64_t rows,
                                           bool byte_coded = false, std::uint64_t anchors = 0);
// Section 1 (serialization.cpp:785-799): the sequence records.
void WriteMetadataSection(std::ostream& out, const std::vector<SequenceInfo>& sequences);
LcpStorage ReadLcpSectionV14(std::istream& in, std::uint64_t section_size,
                             std::uint64_t expected_rows, std::uint32_t sampling_rate,
                             std::uint64_t text_symbols);

// ---------------------------------------------------------- container ---
struct ContainerSpec {
  std::uint16_t format_minor = 4;
  std::uint8_t kind = 1;  // IndexKind::kSuffixArray
  StoredBackend backend = StoredBackend::kHip32;
  std::uint8_t coordinate_width = 32;
  std::uint8_t library_major = kVersionMajor;
  std::uint8_t library_minor = kVersionMinor;
  std::uint8_t library_patch = kVersionPatch;
  std::uint8_t sdsl_major = 0, sdsl_minor = 0, sdsl_patch = 0;
  SaResourceProfile profile = SaResourceProfile::kFast;
  std::uint64_t sequence_count = 0;
  std::uint64_t total_bases = 0;
  std::uint64_t text_symbols = 0;
  std::uint64_t ambiguous_bases = 0;
  std::uint64_t fingerprint = 0;
};

struct SectionDescriptor {
  SectionType type = SectionType::kMetadata;
  std::uint32_t flags = 1;
  std::uint64_t offset = 0;
  std::uint64_t size = 0;
  std::uint32_t crc32 = 0;
};

struct SectionWriter {
  SectionType type;
  std::function<void(std::ostream&)> write;
};

struct ParsedContainer {
  std::filesystem::path path;
  ContainerSpec spec;
  std::vector<SectionDescriptor> sections;
  std::uint64_t file_size = 0;
  const SectionDescriptor* Find(SectionType type) const noexcept;
  const SectionDescriptor& Require(SectionType type) const;
};

std::uint32_t Crc32(std::uint32_t crc, const void* data, std::size_t size);
std::vector<std::uint8_t> BuildHeader(const ContainerSpec& spec,
                                      const std::vector<SectionDescriptor>& sections);
void WriteContainer(const std::filesystem::path& path, const SaveOptions& options,
                    const ContainerSpec& spec, const std::vector<SectionWriter>& writers);
ParsedContainer ReadContainer(const std::filesystem::path& path);
std::vector<std::uint8_t> ReadSection(const ParsedContainer& container, SectionType type);
// zlib CRC32 of bytes [offset, offset + size) of a file, read by several
// threads in parallel and joined (the reference's CrcFileRange,
// serialization.cpp:225-244).
std::uint32_t CrcFileRange(const std::filesystem::path& path, std::uint64_t offset, std::uint64_t size);

// A section whose payload is streamed: host bytes and device planes, in
// order.  A device plane is handed to sfk_hip_write_plane.
struct StreamPart {
  std::vector<std::uint8_t> host;
  const void* device_values = nullptr;
  std::uint64_t count = 0;
  std::uint32_t value_bits = 0;
  std::uint32_t plane = 0;
  bool on_device = false;
};
struct StreamSection {
  SectionType type;
  std::vector<StreamPart> parts;
};
struct StreamStatistics {
  double write_seconds = 0.0;     // sections and header written
  double validate_seconds = 0.0;  // reread with every CRC checked
  std::uint64_t bytes = 0;
};
// WriteContainer for sections held in device memory: same file, same
// publication, CRCs of device planes computed on the device.
void WriteContainerStreamed(const std::filesystem::path& path, const SaveOptions& options,
                            const ContainerSpec& spec, const std::vector<StreamSection>& sections, int device,
                            StreamStatistics* stats);

void WriteU32(std::ostream& out, std::uint32_t value);
void WriteU64(std::ostream& out, std::uint64_t value);

}  // namespace sufkit::detail
