// .sufidx 1.4 persistence: SKC1 coordinate and SKL1 LCP section codecs
// (sa_codec.cpp:466-690) and the CRC-protected outer container with atomic,
// no-clobber publication (serialization.cpp:246-362, 529-765).  All integers
// are little-endian.
#include <fcntl.h>
#include <stdio.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <fstream>
#include <istream>
#include <map>
#include <mutex>
#include <ostream>
#include <set>
#include <system_error>
#include <thread>

#include "internal.hpp"
#include "sufkit_hip.h"

namespace sufkit::detail {
namespace {

constexpr std::uint32_t kCoordinateMagic = 0x31434b53U;  // "SKC1"
constexpr std::uint32_t kLcpMagic = 0x314c4b53U;         // "SKL1"
constexpr std::uint32_t kCodecVersion = 1;
constexpr std::uint32_t kLcpGuideBlockShift = 12;
constexpr std::uint64_t kCoordinateHeaderBytes = 56;
constexpr std::uint64_t kLcpHeaderBytes = 80;
constexpr std::array<std::uint8_t, 8> kMagic{{'S', 'U', 'F', 'K', 'I', 'D', 'X', 0}};
constexpr std::size_t kBaseHeaderSize = 80;
constexpr std::size_t kSectionEntrySize = 32;
constexpr std::size_t kHeaderCrcOffset = 72;
constexpr std::uint32_t kMaxSections = 16;

enum class CoordinateCodec : std::uint32_t { kNative32 = 0, kNative64 = 1, kSplit40 = 2, kSplit48 = 3 };

void WriteBytes(std::ostream& out, const void* data, std::size_t size) {
  out.write(static_cast<const char*>(data), static_cast<std::streamsize>(size));
  if (!out) Throw(ErrorCode::kIoError, "write to the index file failed");
}

template <class T, class A>
void WritePlane(std::ostream& out, const std::vector<T, A>& values) {
  // Little-endian host: the in-memory layout is the file layout.
  constexpr std::size_t kBlock = 1u << 18;
  for (std::size_t begin = 0; begin < values.size(); begin += kBlock) {
    const std::size_t count = std::min(kBlock, values.size() - begin);
    WriteBytes(out, values.data() + begin, count * sizeof(T));
  }
}

void ReadExact(std::istream& in, void* data, std::size_t size, const char* what) {
  in.read(static_cast<char*>(data), static_cast<std::streamsize>(size));
  if (!in || static_cast<std::size_t>(in.gcount()) != size) {
    Throw(ErrorCode::kCorruptIndex, std::string("truncated ") + what);
  }
}

std::uint32_t ReadU32(std::istream& in, const char* what) {
  std::uint8_t b[4];
  ReadExact(in, b, 4, what);
  return static_cast<std::uint32_t>(b[0]) | (static_cast<std::uint32_t>(b[1]) << 8) |
         (static_cast<std::uint32_t>(b[2]) << 16) | (static_cast<std::uint32_t>(b[3]) << 24);
}

std::uint64_t ReadU64(std::istream& in, const char* what) {
  std::uint8_t b[8];
  ReadExact(in, b, 8, what);
  std::uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v |= static_cast<std::uint64_t>(b[i]) << (8 * i);
  return v;
}

// Vec: std::vector<T>, or ValueVector<T> for coordinates and raw LCP (not
// zero-filled before the read fills it).
template <class T, class Vec = std::vector<T>>
Vec ReadPlane(std::istream& in, std::uint64_t count, const char* what) {
  Vec values;
  try {
    values.resize(static_cast<std::size_t>(count));
  } catch (const std::exception&) {
    Throw(ErrorCode::kCorruptIndex, std::string("no memory for ") + what);
  }
  if (count) ReadExact(in, values.data(), static_cast<std::size_t>(count) * sizeof(T), what);
  return values;
}

std::uint64_t CheckedMul(std::uint64_t a, std::uint64_t b, ErrorCode code) {
  if (b != 0 && a > std::numeric_limits<std::uint64_t>::max() / b) Throw(code, "section size overflows 64 bits");
  return a * b;
}

void AppendU16(std::vector<std::uint8_t>& h, std::uint16_t v) {
  h.push_back(static_cast<std::uint8_t>(v));
  h.push_back(static_cast<std::uint8_t>(v >> 8));
}
void AppendU32(std::vector<std::uint8_t>& h, std::uint32_t v) {
  for (int i = 0; i < 4; ++i) h.push_back(static_cast<std::uint8_t>(v >> (8 * i)));
}
void AppendU64(std::vector<std::uint8_t>& h, std::uint64_t v) {
  for (int i = 0; i < 8; ++i) h.push_back(static_cast<std::uint8_t>(v >> (8 * i)));
}
std::uint16_t GetU16(const std::uint8_t* p) { return static_cast<std::uint16_t>(p[0] | (p[1] << 8)); }
std::uint32_t GetU32(const std::uint8_t* p) {
  return static_cast<std::uint32_t>(p[0]) | (static_cast<std::uint32_t>(p[1]) << 8) |
         (static_cast<std::uint32_t>(p[2]) << 16) | (static_cast<std::uint32_t>(p[3]) << 24);
}
std::uint64_t GetU64(const std::uint8_t* p) {
  std::uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v |= static_cast<std::uint64_t>(p[i]) << (8 * i);
  return v;
}

// Stream buffer that forwards to a file buffer and accumulates the CRC32 of
// everything written, so no payload is reread to fill the section table.
class CrcBuffer : public std::streambuf {
 public:
  explicit CrcBuffer(std::streambuf* sink) : sink_(sink) { setp(buffer_, buffer_ + sizeof(buffer_)); }
  bool Finish() { return Flush() == 0; }
  std::uint32_t Crc() const { return crc_; }

 protected:
  int_type overflow(int_type ch) override {
    if (Flush() != 0) return traits_type::eof();
    if (!traits_type::eq_int_type(ch, traits_type::eof())) {
      *pptr() = traits_type::to_char_type(ch);
      pbump(1);
    }
    return traits_type::not_eof(ch);
  }
  int sync() override { return Flush(); }
  std::streamsize xsputn(const char* s, std::streamsize n) override {
    if (n >= static_cast<std::streamsize>(sizeof(buffer_))) {
      if (Flush() != 0) return 0;
      crc_ = Crc32(crc_, s, static_cast<std::size_t>(n));
      return sink_->sputn(s, n);
    }
    return std::streambuf::xsputn(s, n);
  }

 private:
  int Flush() {
    const std::ptrdiff_t n = pptr() - pbase();
    if (n > 0) {
      crc_ = Crc32(crc_, pbase(), static_cast<std::size_t>(n));
      if (sink_->sputn(pbase(), n) != n) return -1;
      pbump(static_cast<int>(-n));
    }
    return 0;
  }
  std::streambuf* sink_;
  char buffer_[1 << 16];
  std::uint32_t crc_ = 0;
};

// The process's file-creation mask, read once (umask() can only be read by
// setting it, so this happens before any file is written).
mode_t CreationMask() {
  static const mode_t mask = [] {
    const mode_t old = umask(0);
    umask(old);
    return old;
  }();
  return mask;
}

// A new empty file beside `target` for the index to be written into, named
// by mkostemp from a "<name>.sfk-XXXXXX" template, so its directory (and
// hence its file system) is the target's.  It gets the permissions a plain
// create would give (0666 less the creation mask), not mkostemp's 0600.
std::filesystem::path StagingFileFor(const std::filesystem::path& target) {
  std::string name = target.string() + ".sfk-XXXXXX";
  const mode_t mask = CreationMask();
  const int fd = mkostemp(name.data(), O_CLOEXEC);
  if (fd < 0) {
    Throw(ErrorCode::kIoError, "no staging file could be made beside " + target.string() + " (" +
                                   std::error_code(errno, std::generic_category()).message() + ")");
  }
  const bool ok = fchmod(fd, 0666 & ~mask) == 0;
  const int chmod_errno = errno;
  close(fd);
  if (!ok) {
    unlink(name.c_str());
    Throw(ErrorCode::kIoError, "staging file " + name + " keeps mkostemp's permissions (" +
                                   std::error_code(chmod_errno, std::generic_category()).message() + ")");
  }
  return name;
}

// Moves the finished staging file to `target` in one step.  Without
// overwrite the move must not replace a file that appeared meanwhile:
// renameat2 with RENAME_NOREPLACE, or, where the file system lacks it, a
// hard link (which fails on an existing name) and the staging name removed.
void MoveIntoPlace(const std::filesystem::path& staging, const std::filesystem::path& target, bool overwrite) {
  auto fail = [&](int err) {
    if (err == EEXIST) {
      Throw(ErrorCode::kIoError, target.string() + " appeared while the index was written; it was left as it is");
    }
    Throw(ErrorCode::kIoError, "the written index could not be moved to " + target.string() + " (" +
                                   std::error_code(err, std::generic_category()).message() + ")");
  };
  if (overwrite) {
    if (rename(staging.c_str(), target.c_str()) != 0) fail(errno);
    return;
  }
  if (renameat2(AT_FDCWD, staging.c_str(), AT_FDCWD, target.c_str(), RENAME_NOREPLACE) == 0) return;
  if (errno != EINVAL && errno != ENOSYS) fail(errno);
  if (link(staging.c_str(), target.c_str()) != 0) fail(errno);
  unlink(staging.c_str());
}

}  // namespace

std::uint32_t Crc32(std::uint32_t crc, const void* data, std::size_t size) {
  const auto* p = static_cast<const unsigned char*>(data);
  uLong c = crc;
  while (size > 0) {
    const auto chunk = static_cast<uInt>(std::min<std::size_t>(size, 1u << 30));
    c = crc32(c, p, chunk);
    p += chunk;
    size -= chunk;
  }
  return static_cast<std::uint32_t>(c);
}

void WriteU32(std::ostream& out, std::uint32_t v) {
  std::uint8_t b[4];
  for (int i = 0; i < 4; ++i) b[i] = static_cast<std::uint8_t>(v >> (8 * i));
  WriteBytes(out, b, 4);
}

void WriteU64(std::ostream& out, std::uint64_t v) {
  std::uint8_t b[8];
  for (int i = 0; i < 8; ++i) b[i] = static_cast<std::uint8_t>(v >> (8 * i));
  WriteBytes(out, b, 8);
}

// ------------------------------------------------------------- codecs ---

std::vector<std::uint8_t> CoordinateSectionHeader(CoordinateStorageWidth width, std::uint64_t symbol_count,
                                                  std::uint64_t count) {
  (void)ResolveCoordinateStorageWidth(width, symbol_count, ErrorCode::kBuildFailure);
  CoordinateCodec codec = CoordinateCodec::kNative32;
  std::uint64_t low = 0, high = 0;
  switch (width) {
    case CoordinateStorageWidth::kBits32: codec = CoordinateCodec::kNative32; low = count * 4; break;
    case CoordinateStorageWidth::kBits64: codec = CoordinateCodec::kNative64; low = count * 8; break;
    case CoordinateStorageWidth::kBits40: codec = CoordinateCodec::kSplit40; low = count * 4; high = count; break;
    case CoordinateStorageWidth::kBits48: codec = CoordinateCodec::kSplit48; low = count * 4; high = count * 2; break;
    default: Throw(ErrorCode::kBuildFailure, "storage width was not resolved before packing");
  }
  std::vector<std::uint8_t> h;
  h.reserve(kCoordinateHeaderBytes);
  AppendU32(h, kCoordinateMagic);
  AppendU32(h, kCodecVersion);
  AppendU32(h, static_cast<std::uint32_t>(codec));
  AppendU32(h, 0);
  AppendU64(h, symbol_count);
  AppendU64(h, count);
  AppendU64(h, low);
  AppendU64(h, high);
  AppendU64(h, 0);
  return h;
}

std::vector<std::uint8_t> LcpSectionHeader(std::uint32_t width, std::uint32_t sampling_rate,
                                           std::uint64_t text_symbols, std::uint64_t rows, bool byte_coded,
                                           std::uint64_t anchors) {
  std::vector<std::uint8_t> h;
  h.reserve(kLcpHeaderBytes);
  AppendU32(h, kLcpMagic);
  AppendU32(h, kCodecVersion);
  AppendU32(h, byte_coded ? 1 : 0);  // raw or byte-coded
  AppendU32(h, width);
  AppendU32(h, sampling_rate);
  AppendU32(h, kLcpGuideBlockShift);
  AppendU64(h, text_symbols);
  AppendU64(h, rows);
  AppendU64(h, byte_coded ? rows : rows * (width / 8u));
  AppendU64(h, byte_coded ? anchors : 0);
  AppendU64(h, byte_coded ? anchors * (width / 8u) : 0);
  AppendU64(h, byte_coded ? anchors * (width / 8u) : 0);
  AppendU64(h, 0);
  return h;
}

void WriteMetadataSection(std::ostream& out, const std::vector<SequenceInfo>& sequences) {
  WriteU32(out, static_cast<std::uint32_t>(sequences.size()));
  for (const auto& s : sequences) {
    WriteU32(out, s.id);
    WriteU64(out, s.global_offset);
    WriteU64(out, s.length);
    WriteU64(out, s.ambiguous_bases);
    WriteU32(out, static_cast<std::uint32_t>(s.name.size()));
    out.write(s.name.data(), static_cast<std::streamsize>(s.name.size()));
    WriteU32(out, static_cast<std::uint32_t>(s.description.size()));
    out.write(s.description.data(), static_cast<std::streamsize>(s.description.size()));
  }
}

void WriteCoordinateSectionV14(std::ostream& out, const CoordinateStorage& storage,
                               std::uint64_t symbol_count) {
  const auto header = CoordinateSectionHeader(storage.Width(), symbol_count, storage.Size());
  WriteBytes(out, header.data(), header.size());
  const auto& raw = storage.Raw();
  switch (raw.index()) {
    case 1: WritePlane(out, std::get<1>(raw)); break;
    case 2: WritePlane(out, std::get<2>(raw)); break;
    case 3: WritePlane(out, std::get<3>(raw).low); WritePlane(out, std::get<3>(raw).high); break;
    case 4: WritePlane(out, std::get<4>(raw).low); WritePlane(out, std::get<4>(raw).high); break;
    default: break;
  }
}

CoordinateStorage ReadCoordinateSectionV14(std::istream& in, std::uint64_t section_size,
                                           std::uint64_t expected_elements,
                                           std::uint64_t expected_symbols, const char* label) {
  if (ReadU32(in, "SKC1 magic") != kCoordinateMagic) {
    Throw(ErrorCode::kCorruptIndex, "SKC1 section: magic is not SKC1");
  }
  if (ReadU32(in, "SKC1 version") != kCodecVersion) {
    Throw(ErrorCode::kVersionMismatch, "SKC1 section: header version is not 1");
  }
  const std::uint32_t raw_codec = ReadU32(in, "SKC1 codec id");
  if (raw_codec > 3) Throw(ErrorCode::kVersionMismatch, "SKC1 section: codec id is not native32/native64/split40/split48");
  if (ReadU32(in, "coordinate reserved") != 0) Throw(ErrorCode::kCorruptIndex, "SKC1 section: a reserved field is set");
  const std::uint64_t symbols = ReadU64(in, "SKC1 symbol count");
  const std::uint64_t count = ReadU64(in, "SKC1 element count");
  const std::uint64_t low = ReadU64(in, "SKC1 low plane bytes");
  const std::uint64_t high = ReadU64(in, "SKC1 high plane bytes");
  if (ReadU64(in, "coordinate reserved") != 0) Throw(ErrorCode::kCorruptIndex, "SKC1 section: a reserved field is set");
  const auto codec = static_cast<CoordinateCodec>(raw_codec);
  std::uint64_t want_low = 0, want_high = 0;
  CoordinateStorageWidth width = CoordinateStorageWidth::kBits32;
  switch (codec) {
    case CoordinateCodec::kNative32: want_low = CheckedMul(count, 4, ErrorCode::kCorruptIndex); width = CoordinateStorageWidth::kBits32; break;
    case CoordinateCodec::kNative64: want_low = CheckedMul(count, 8, ErrorCode::kCorruptIndex); width = CoordinateStorageWidth::kBits64; break;
    case CoordinateCodec::kSplit40: want_low = CheckedMul(count, 4, ErrorCode::kCorruptIndex); want_high = count; width = CoordinateStorageWidth::kBits40; break;
    case CoordinateCodec::kSplit48: want_low = CheckedMul(count, 4, ErrorCode::kCorruptIndex); want_high = CheckedMul(count, 2, ErrorCode::kCorruptIndex); width = CoordinateStorageWidth::kBits48; break;
  }
  if (symbols == 0 || low != want_low || high != want_high ||
      kCoordinateHeaderBytes + low + high != section_size) {
    Throw(ErrorCode::kCorruptIndex, "SKC1 section: plane byte counts do not follow from the codec");
  }
  if (count != expected_elements || symbols != expected_symbols) {
    Throw(ErrorCode::kCorruptIndex, std::string("SKC1 section (") + label + ")" +
                                        " fields disagree with the file header");
  }
  switch (codec) {
    case CoordinateCodec::kNative32:
      return CoordinateStorage::FromUInt32(ReadPlane<std::uint32_t, U32Values>(in, count, label), width, symbols,
                                           ErrorCode::kCorruptIndex, label);
    case CoordinateCodec::kNative64:
      return CoordinateStorage::FromUInt64(ReadPlane<std::uint64_t, U64Values>(in, count, label), width, symbols,
                                           ErrorCode::kCorruptIndex, label);
    case CoordinateCodec::kSplit40: {
      Coordinate40 v;
      v.low = ReadPlane<std::uint32_t>(in, count, label);
      v.high = ReadPlane<std::uint8_t>(in, count, label);
      return CoordinateStorage::FromPacked40(std::move(v), symbols, ErrorCode::kCorruptIndex, label);
    }
    case CoordinateCodec::kSplit48: {
      Coordinate48 v;
      v.low = ReadPlane<std::uint32_t>(in, count, label);
      v.high = ReadPlane<std::uint16_t>(in, count, label);
      return CoordinateStorage::FromPacked48(std::move(v), symbols, ErrorCode::kCorruptIndex, label);
    }
  }
  Throw(ErrorCode::kCorruptIndex, "SKC1 section: codec id not handled");
}

void WriteLcpSectionV14(std::ostream& out, const LcpStorage& lcp, std::uint64_t text_symbols) {
  if (lcp.Empty()) Throw(ErrorCode::kBuildFailure, "SKL1 section: there is no LCP to write");
  if (text_symbols == 0 || lcp.Size() != 1 + (text_symbols - 1) / lcp.SamplingRate()) {
    Throw(ErrorCode::kBuildFailure, "SKL1 section: LCP fields disagree before writing");
  }
  const auto header = LcpSectionHeader(lcp.Width(), lcp.SamplingRate(), text_symbols, lcp.Size(), lcp.ByteCoded(),
                                       lcp.AnchorPositions().size());
  WriteBytes(out, header.data(), header.size());
  if (lcp.ByteCoded()) {
    WriteBytes(out, lcp.Primary().data(), lcp.Primary().size());
    for (const auto* plane : {&lcp.AnchorPositions(), &lcp.AnchorValues()}) {
      if (lcp.Width() == 32) {
        std::vector<std::uint32_t> narrow(plane->begin(), plane->end());
        WritePlane(out, narrow);
      } else {
        WritePlane(out, *plane);
      }
    }
  } else if (lcp.Width() == 32) {
    WritePlane(out, lcp.Raw32());
  } else {
    WritePlane(out, lcp.Raw64());
  }
}

LcpStorage ReadLcpSectionV14(std::istream& in, std::uint64_t section_size, std::uint64_t expected_rows,
                             std::uint32_t sampling_rate, std::uint64_t text_symbols) {
  if (ReadU32(in, "SKL1 magic") != kLcpMagic) Throw(ErrorCode::kCorruptIndex, "SKL1 section: magic is not SKL1");
  if (ReadU32(in, "SKL1 version") != kCodecVersion) {
    Throw(ErrorCode::kVersionMismatch, "SKL1 section: header version is not 1");
  }
  const std::uint32_t codec = ReadU32(in, "SKL1 codec id");
  if (codec > 1) Throw(ErrorCode::kVersionMismatch, "SKL1 section: codec id is not raw or byte-coded");
  const std::uint32_t width = ReadU32(in, "SKL1 value width");
  const std::uint32_t rate = ReadU32(in, "SKL1 sampling rate");
  if (ReadU32(in, "SKL1 guide shift") != kLcpGuideBlockShift) {
    Throw(ErrorCode::kCorruptIndex, "SKL1 section: guide shift is not 12");
  }
  const std::uint64_t symbols = ReadU64(in, "SKL1 symbol count");
  const std::uint64_t rows = ReadU64(in, "LCP row count");
  const std::uint64_t primary = ReadU64(in, "SKL1 primary bytes");
  const std::uint64_t anchors = ReadU64(in, "SKL1 anchor count");
  const std::uint64_t anchor_pos = ReadU64(in, "SKL1 anchor position bytes");
  const std::uint64_t anchor_val = ReadU64(in, "SKL1 anchor value bytes");
  if (ReadU64(in, "LCP reserved") != 0) Throw(ErrorCode::kCorruptIndex, "SKL1 section: a reserved field is set");
  if (rate == 0 || (width != 32 && width != 64) || symbols == 0) {
    Throw(ErrorCode::kCorruptIndex, "SKL1 section: header fields are inconsistent");
  }
  if (rows != expected_rows || rate != sampling_rate || symbols != text_symbols) {
    Throw(ErrorCode::kCorruptIndex, "SKL1 section: fields disagree with the file header");
  }
  if (codec == 1) {  // byte-coded: primary bytes, anchor positions, anchor values (sa_codec.cpp:353-370, 655-680)
    if (width == 32 && symbols - 1 > std::numeric_limits<std::uint32_t>::max()) {
      Throw(ErrorCode::kCorruptIndex, "SKL1 section: text positions do not fit the value width");
    }
    if (rows != 1 + (symbols - 1) / rate || primary != rows || anchors >= rows ||
        anchor_pos != CheckedMul(anchors, width / 8u, ErrorCode::kCorruptIndex) || anchor_val != anchor_pos ||
        kLcpHeaderBytes + primary + anchor_pos + anchor_val != section_size) {
      Throw(ErrorCode::kCorruptIndex, "SKL1 section: byte-coded plane sizes are inconsistent");
    }
    std::vector<std::uint8_t> bytes(static_cast<std::size_t>(rows));
    ReadExact(in, bytes.data(), bytes.size(), "SKL1 primary plane");
    std::vector<std::uint64_t> positions, values;
    if (width == 32) {
      auto p32 = ReadPlane<std::uint32_t>(in, anchors, "SKL1 anchor positions");
      auto v32 = ReadPlane<std::uint32_t>(in, anchors, "SKL1 anchor values");
      positions.assign(p32.begin(), p32.end());
      values.assign(v32.begin(), v32.end());
    } else {
      positions = ReadPlane<std::uint64_t>(in, anchors, "SKL1 anchor positions");
      values = ReadPlane<std::uint64_t>(in, anchors, "SKL1 anchor values");
    }
    return LcpStorage::FromByteCoded(std::move(bytes), std::move(positions), std::move(values), rate, symbols,
                                     static_cast<std::uint8_t>(width), ErrorCode::kCorruptIndex);
  }
  if (rows != 1 + (symbols - 1) / rate || primary != CheckedMul(rows, width / 8u, ErrorCode::kCorruptIndex) ||
      anchors != 0 || anchor_pos != 0 || anchor_val != 0 || kLcpHeaderBytes + primary != section_size) {
    Throw(ErrorCode::kCorruptIndex, "SKL1 section: raw plane size is inconsistent");
  }
  if (width == 32) return LcpStorage::FromRaw32(ReadPlane<std::uint32_t, U32Values>(in, rows, "raw32 LCP"), rate, ErrorCode::kCorruptIndex);
  return LcpStorage::FromRaw64(ReadPlane<std::uint64_t, U64Values>(in, rows, "raw64 LCP"), rate, ErrorCode::kCorruptIndex);
}

// ---------------------------------------------------------- container ---

const SectionDescriptor* ParsedContainer::Find(SectionType type) const noexcept {
  for (const auto& s : sections) {
    if (s.type == type) return &s;
  }
  return nullptr;
}

const SectionDescriptor& ParsedContainer::Require(SectionType type) const {
  const auto* s = Find(type);
  if (s == nullptr) Throw(ErrorCode::kCorruptIndex, "a section the index needs is absent");
  return *s;
}

std::vector<std::uint8_t> BuildHeader(const ContainerSpec& spec,
                                      const std::vector<SectionDescriptor>& sections) {
  std::vector<std::uint8_t> h(kMagic.begin(), kMagic.end());
  h.reserve(kBaseHeaderSize + sections.size() * kSectionEntrySize);
  AppendU16(h, 1);
  AppendU16(h, spec.format_minor);
  h.push_back(1);  // little endian
  h.push_back(spec.kind);
  h.push_back(static_cast<std::uint8_t>(spec.backend));
  h.push_back(spec.coordinate_width);
  h.push_back(kNormalizationId);
  h.push_back(spec.library_major);
  h.push_back(spec.library_minor);
  h.push_back(spec.library_patch);
  h.push_back(spec.sdsl_major);
  h.push_back(spec.sdsl_minor);
  h.push_back(spec.sdsl_patch);
  h.push_back(spec.format_minor >= 4 ? static_cast<std::uint8_t>(spec.profile) : 0);
  AppendU32(h, static_cast<std::uint32_t>(sections.size()));
  AppendU32(h, static_cast<std::uint32_t>(kBaseHeaderSize + sections.size() * kSectionEntrySize));
  AppendU64(h, spec.sequence_count);
  AppendU64(h, spec.total_bases);
  AppendU64(h, spec.text_symbols);
  AppendU64(h, spec.ambiguous_bases);
  AppendU64(h, spec.fingerprint);
  AppendU32(h, 0);  // header CRC, patched below
  AppendU32(h, 0);
  for (const auto& s : sections) {
    AppendU32(h, static_cast<std::uint32_t>(s.type));
    AppendU32(h, s.flags);
    AppendU64(h, s.offset);
    AppendU64(h, s.size);
    AppendU32(h, s.crc32);
    AppendU32(h, 0);
  }
  const std::uint32_t crc = Crc32(0, h.data(), h.size());
  for (int i = 0; i < 4; ++i) h[kHeaderCrcOffset + i] = static_cast<std::uint8_t>(crc >> (8 * i));
  return h;
}

void WriteContainer(const std::filesystem::path& path, const SaveOptions& options,
                    const ContainerSpec& spec, const std::vector<SectionWriter>& writers) {
  if (writers.empty() || writers.size() > kMaxSections) {
    Throw(ErrorCode::kBuildFailure, "an index has 1 to 16 sections");
  }
  if (path.empty()) Throw(ErrorCode::kInvalidInput, "no output path was given");
  if (!options.overwrite && std::filesystem::exists(path)) {
    Throw(ErrorCode::kIoError, path.string() + " exists; saving over it needs the overwrite option");
  }
  const auto temp = StagingFileFor(path);
  try {
    const std::size_t header_size = kBaseHeaderSize + writers.size() * kSectionEntrySize;
    std::fstream out(temp, std::ios::binary | std::ios::in | std::ios::out | std::ios::trunc);
    if (!out) Throw(ErrorCode::kIoError, "staging file could not be opened: " + temp.string());
    std::vector<char> placeholder(header_size, '\0');
    out.write(placeholder.data(), static_cast<std::streamsize>(placeholder.size()));
    std::vector<SectionDescriptor> sections;
    std::set<SectionType> types;
    for (const auto& w : writers) {
      if (!types.insert(w.type).second) Throw(ErrorCode::kBuildFailure, "a section type was written twice");
      const auto begin = out.tellp();
      CrcBuffer crc(out.rdbuf());
      std::ostream section(&crc);
      w.write(section);
      section.flush();
      if (!section || !crc.Finish() || !out) Throw(ErrorCode::kIoError, "writing a section failed");
      const auto end = out.tellp();
      SectionDescriptor d;
      d.type = w.type;
      d.flags = 1;
      d.offset = static_cast<std::uint64_t>(begin);
      d.size = static_cast<std::uint64_t>(end - begin);
      d.crc32 = crc.Crc();
      sections.push_back(d);
    }
    out.flush();
    const auto header = BuildHeader(spec, sections);
    out.seekp(0);
    out.write(reinterpret_cast<const char*>(header.data()), static_cast<std::streamsize>(header.size()));
    out.flush();
    out.close();
    if (!out) Throw(ErrorCode::kIoError, "closing the staging file failed");
    (void)ReadContainer(temp);  // self-validate before publication
    MoveIntoPlace(temp, path, options.overwrite);
  } catch (...) {
    std::error_code ignored;
    std::filesystem::remove(temp, ignored);
    throw;
  }
}

ParsedContainer ReadContainer(const std::filesystem::path& path) {
  std::error_code ec;
  const auto file_size = std::filesystem::file_size(path, ec);
  if (ec) Throw(ErrorCode::kIoError, "index file cannot be examined: " + path.string());
  if (file_size < kBaseHeaderSize) Throw(ErrorCode::kCorruptIndex, "file is shorter than an index header");
  std::ifstream in(path, std::ios::binary);
  if (!in) Throw(ErrorCode::kIoError, "index file could not be opened: " + path.string());
  std::vector<std::uint8_t> head(kBaseHeaderSize);
  ReadExact(in, head.data(), head.size(), "index header");
  if (!std::equal(kMagic.begin(), kMagic.end(), head.begin())) {
    Throw(ErrorCode::kCorruptIndex, "file does not start with the index magic");
  }
  const auto major = GetU16(&head[8]);
  const auto minor = GetU16(&head[10]);
  if (major != 1 || minor > 4) Throw(ErrorCode::kVersionMismatch, "index format version is newer than 1.4");
  if (head[12] != 1) Throw(ErrorCode::kVersionMismatch, "unsupported index endianness");
  const std::uint32_t count = GetU32(&head[24]);
  const std::uint32_t header_size = GetU32(&head[28]);
  if (count == 0 || count > kMaxSections || header_size != kBaseHeaderSize + count * kSectionEntrySize ||
      header_size > file_size) {
    Throw(ErrorCode::kCorruptIndex, "index section table is invalid");
  }
  head.resize(header_size);
  ReadExact(in, head.data() + kBaseHeaderSize, header_size - kBaseHeaderSize, "section table");
  const std::uint32_t stored_crc = GetU32(&head[kHeaderCrcOffset]);
  std::vector<std::uint8_t> zeroed = head;
  std::memset(&zeroed[kHeaderCrcOffset], 0, 4);
  if (Crc32(0, zeroed.data(), zeroed.size()) != stored_crc) {
    Throw(ErrorCode::kCorruptIndex, "index header checksum mismatch");
  }
  ParsedContainer c;
  c.path = path;
  c.file_size = file_size;
  c.spec.format_minor = minor;
  c.spec.kind = head[13];
  c.spec.backend = static_cast<StoredBackend>(head[14]);
  c.spec.coordinate_width = head[15];
  if (head[16] != kNormalizationId) Throw(ErrorCode::kVersionMismatch, "unsupported normalization");
  c.spec.library_major = head[17];
  c.spec.library_minor = head[18];
  c.spec.library_patch = head[19];
  c.spec.sdsl_major = head[20];
  c.spec.sdsl_minor = head[21];
  c.spec.sdsl_patch = head[22];
  c.spec.profile = static_cast<SaResourceProfile>(head[23]);
  c.spec.sequence_count = GetU64(&head[32]);
  c.spec.total_bases = GetU64(&head[40]);
  c.spec.text_symbols = GetU64(&head[48]);
  c.spec.ambiguous_bases = GetU64(&head[56]);
  c.spec.fingerprint = GetU64(&head[64]);
  std::set<std::uint32_t> types;
  for (std::uint32_t i = 0; i < count; ++i) {
    const std::uint8_t* e = &head[kBaseHeaderSize + i * kSectionEntrySize];
    SectionDescriptor d;
    d.type = static_cast<SectionType>(GetU32(e));
    d.flags = GetU32(e + 4);
    d.offset = GetU64(e + 8);
    d.size = GetU64(e + 16);
    d.crc32 = GetU32(e + 24);
    if (!types.insert(GetU32(e)).second) Throw(ErrorCode::kCorruptIndex, "two sections share a type");
    if (d.offset < header_size || d.offset > file_size || d.size > file_size - d.offset) {
      Throw(ErrorCode::kCorruptIndex, "index section lies outside the file");
    }
    c.sections.push_back(d);
  }
  auto sorted = c.sections;
  std::sort(sorted.begin(), sorted.end(), [](const auto& a, const auto& b) { return a.offset < b.offset; });
  for (std::size_t i = 1; i < sorted.size(); ++i) {
    if (sorted[i - 1].offset + sorted[i - 1].size > sorted[i].offset) {
      Throw(ErrorCode::kCorruptIndex, "two sections share bytes");
    }
  }
  // Every section's CRC32.
  for (const auto& d : c.sections) {
    if (CrcFileRange(path, d.offset, d.size) != d.crc32) {
      Throw(ErrorCode::kCorruptIndex, "index section checksum mismatch");
    }
  }
  return c;
}

std::uint32_t CrcFileRange(const std::filesystem::path& path, std::uint64_t offset, std::uint64_t size) {
  constexpr std::uint64_t kPiece = 16ull << 20;
  constexpr std::size_t kBuffer = 4u << 20;
  if (size == 0) return 0;
  const std::uint64_t pieces = (size + kPiece - 1) / kPiece;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned threads = static_cast<unsigned>(std::min<std::uint64_t>({pieces, hw, 16}));
  const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
  if (fd < 0) Throw(ErrorCode::kIoError, "index file could not be opened: " + path.string());
  std::vector<std::uint32_t> crcs(pieces, 0);
  std::atomic<std::uint64_t> next{0};
  std::atomic<bool> failed{false};
  auto work = [&] {
    std::vector<unsigned char> buffer(kBuffer);
    for (std::uint64_t p = next.fetch_add(1); p < pieces && !failed.load(); p = next.fetch_add(1)) {
      std::uint64_t at = offset + p * kPiece;
      std::uint64_t left = std::min(kPiece, size - p * kPiece);
      uLong crc = crc32(0L, Z_NULL, 0);
      while (left > 0) {
        const auto want = static_cast<std::size_t>(std::min<std::uint64_t>(left, kBuffer));
        const ssize_t got = ::pread(fd, buffer.data(), want, static_cast<off_t>(at));
        if (got <= 0) {
          failed = true;
          return;
        }
        crc = crc32(crc, buffer.data(), static_cast<uInt>(got));
        at += static_cast<std::uint64_t>(got);
        left -= static_cast<std::uint64_t>(got);
      }
      crcs[p] = static_cast<std::uint32_t>(crc);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  ::close(fd);
  if (failed) Throw(ErrorCode::kCorruptIndex, "index is truncated: cannot read a section");
  uLong crc = crcs[0];
  for (std::uint64_t p = 1; p < pieces; ++p) {
    crc = crc32_combine(crc, crcs[p], static_cast<z_off_t>(std::min(kPiece, size - p * kPiece)));
  }
  return static_cast<std::uint32_t>(crc);
}

namespace {

// Appends to a file with positioned writes (the header is patched at
// offset 0 at the end).
struct FileSink {
  int fd = -1;
  std::uint64_t written = 0;

  static int Write(void* context, const void* data, std::uint64_t bytes) {
    auto* sink = static_cast<FileSink*>(context);
    const auto* p = static_cast<const char*>(data);
    while (bytes > 0) {
      const ssize_t n = ::pwrite(sink->fd, p, static_cast<std::size_t>(std::min<std::uint64_t>(bytes, 1u << 30)),
                                 static_cast<off_t>(sink->written));
      if (n <= 0) return 1;
      p += n;
      bytes -= static_cast<std::uint64_t>(n);
      sink->written += static_cast<std::uint64_t>(n);
    }
    return 0;
  }
};

// One plane writer per device, created on the first save and kept for the
// life of the process (its pinned buffers cost more to set up than a
// chunk costs to stream); saves on one device take turns on it.
constexpr std::uint64_t kPlaneChunkBytes = 16ull << 20;

class PlaneWriter {
 public:
  explicit PlaneWriter(int device) : lock_(Pool().mutex) {
    auto& slot = Pool().writers[device];
    if (slot == nullptr) {
      char err[512] = {0};
      const int status = sfk_hip_plane_writer_create(device, kPlaneChunkBytes, &slot, err, sizeof(err));
      if (status != SFK_HIP_OK) {
        slot = nullptr;
        Throw(static_cast<ErrorCode>(status - 1), std::string("plane writer: ") + err);
      }
    }
    writer_ = slot;
  }
  PlaneWriter(const PlaneWriter&) = delete;
  PlaneWriter& operator=(const PlaneWriter&) = delete;
  sfk_hip_plane_writer* get() const { return writer_; }

 private:
  struct Writers {
    std::mutex mutex;
    std::map<int, sfk_hip_plane_writer*> writers;
  };
  static Writers& Pool() {
    static Writers* pool = new Writers();  // never destroyed: outlives every save
    return *pool;
  }
  std::lock_guard<std::mutex> lock_;
  sfk_hip_plane_writer* writer_ = nullptr;
};

}  // namespace

void WriteContainerStreamed(const std::filesystem::path& path, const SaveOptions& options,
                            const ContainerSpec& spec, const std::vector<StreamSection>& sections, int device,
                            StreamStatistics* stats) {
  using Clock = std::chrono::steady_clock;
  const auto begin = Clock::now();
  if (sections.empty() || sections.size() > kMaxSections) {
    Throw(ErrorCode::kBuildFailure, "an index has 1 to 16 sections");
  }
  if (path.empty()) Throw(ErrorCode::kInvalidInput, "no output path was given");
  if (!options.overwrite && std::filesystem::exists(path)) {
    Throw(ErrorCode::kIoError, path.string() + " exists; saving over it needs the overwrite option");
  }
  PlaneWriter writer(device);
  const auto temp = StagingFileFor(path);
  FileSink sink;
  try {
    sink.fd = ::open(temp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (sink.fd < 0) Throw(ErrorCode::kIoError, "staging file could not be opened: " + temp.string());
    const std::size_t header_size = kBaseHeaderSize + sections.size() * kSectionEntrySize;
    const std::vector<std::uint8_t> placeholder(header_size, 0);
    if (FileSink::Write(&sink, placeholder.data(), placeholder.size()) != 0) {
      Throw(ErrorCode::kIoError, "failed while writing index header");
    }
    std::vector<SectionDescriptor> written;
    std::set<SectionType> types;
    for (const auto& section : sections) {
      if (!types.insert(section.type).second) Throw(ErrorCode::kBuildFailure, "a section type was written twice");
      SectionDescriptor d;
      d.type = section.type;
      d.flags = 1;
      d.offset = sink.written;
      std::uint32_t crc = 0;
      for (const auto& part : section.parts) {
        if (!part.on_device) {
          if (FileSink::Write(&sink, part.host.data(), part.host.size()) != 0) {
            Throw(ErrorCode::kIoError, "writing a section failed");
          }
          crc = Crc32(crc, part.host.data(), part.host.size());
          continue;
        }
        char err[512] = {0};
        const int status = sfk_hip_write_plane(writer.get(), part.device_values, part.count, part.value_bits,
                                               part.plane, crc, &FileSink::Write, &sink, &crc, err, sizeof(err));
        if (status != SFK_HIP_OK) {
          Throw(static_cast<ErrorCode>(status - 1), std::string("streaming a device plane: ") + err);
        }
      }
      d.size = sink.written - d.offset;
      d.crc32 = crc;
      written.push_back(d);
    }
    const auto header = BuildHeader(spec, written);
    if (::pwrite(sink.fd, header.data(), header.size(), 0) != static_cast<ssize_t>(header.size())) {
      Throw(ErrorCode::kIoError, "failed to write the index header");
    }
    const int closed = ::close(sink.fd);
    sink.fd = -1;
    if (closed != 0) Throw(ErrorCode::kIoError, "closing the staging file failed");
    const auto validate = Clock::now();
    if (stats != nullptr) {
      stats->write_seconds = std::chrono::duration<double>(validate - begin).count();
      stats->bytes = sink.written;
    }
    (void)ReadContainer(temp);  // self-validate before publication
    if (stats != nullptr) stats->validate_seconds = std::chrono::duration<double>(Clock::now() - validate).count();
    MoveIntoPlace(temp, path, options.overwrite);
  } catch (...) {
    if (sink.fd >= 0) ::close(sink.fd);
    std::error_code ignored;
    std::filesystem::remove(temp, ignored);
    throw;
  }
}

std::vector<std::uint8_t> ReadSection(const ParsedContainer& container, SectionType type) {
  const auto& d = container.Require(type);
  std::ifstream in(container.path, std::ios::binary);
  if (!in) Throw(ErrorCode::kIoError, "index file could not be opened: " + container.path.string());
  in.seekg(static_cast<std::streamoff>(d.offset));
  std::vector<std::uint8_t> bytes(static_cast<std::size_t>(d.size));
  if (d.size) ReadExact(in, bytes.data(), bytes.size(), "index section");
  return bytes;
}

}  // namespace sufkit::detail
