// Reference normalization (genome_reference.cpp:30-160 of the reference):
// A/C/G/T (either case) -> 2..5, every other byte -> N (6); one separator 1
// after every contig; FNV-1a-64 over the encoded bytes; record validation
// (non-empty unique names, non-empty sequences, at least one record).
#include <zlib.h>

#include <array>
#include <limits>
#include <memory>
#include <string_view>
#include <unordered_set>
#include <utility>

#include "internal.hpp"

namespace sufkit {

struct GenomeReference::Impl {
  std::vector<std::uint8_t> encoded;
  std::vector<SequenceInfo> sequences;
  std::uint64_t total_bases = 0;
  std::uint64_t ambiguous_bases = 0;
  std::uint64_t fingerprint = 14695981039346656037ULL;
};

namespace {

constexpr std::array<std::uint8_t, 256> MakeEncodingTable() {
  std::array<std::uint8_t, 256> t{};
  for (auto& v : t) v = detail::kN;
  t['A'] = t['a'] = detail::kA;
  t['C'] = t['c'] = detail::kC;
  t['G'] = t['g'] = detail::kG;
  t['T'] = t['t'] = detail::kT;
  return t;
}
constexpr auto kEncoding = MakeEncodingTable();

class Builder {
 public:
  void Add(std::string name, std::string description, std::string_view sequence) {
    if (name.empty()) detail::Throw(ErrorCode::kInvalidInput, "a record has no name");
    if (!names_.insert(name).second) {
      detail::Throw(ErrorCode::kInvalidInput, "two records share the name " + name);
    }
    if (sequence.empty()) detail::Throw(ErrorCode::kInvalidInput, "record holds no bases: " + name);
    if (impl_->sequences.size() > std::numeric_limits<SequenceId>::max()) {
      detail::Throw(ErrorCode::kInvalidInput, "more records than a reference can hold");
    }
    SequenceInfo info;
    info.id = static_cast<SequenceId>(impl_->sequences.size());
    info.name = std::move(name);
    info.description = std::move(description);
    info.length = sequence.size();
    info.global_offset = next_offset_;
    auto& enc = impl_->encoded;
    std::uint64_t fp = impl_->fingerprint;
    std::uint64_t amb = 0;
    enc.reserve(enc.size() + sequence.size() + 1);
    for (const char raw : sequence) {
      const std::uint8_t v = kEncoding[static_cast<unsigned char>(raw)];
      amb += v == detail::kN;
      enc.push_back(v);
      fp = (fp ^ v) * 1099511628211ULL;
    }
    enc.push_back(detail::kSeparator);
    fp = (fp ^ detail::kSeparator) * 1099511628211ULL;
    impl_->fingerprint = fp;
    info.ambiguous_bases = amb;
    impl_->total_bases += info.length;
    impl_->ambiguous_bases += amb;
    next_offset_ += info.length + 1;
    impl_->sequences.push_back(std::move(info));
  }

  std::unique_ptr<GenomeReference::Impl> Finish() {
    if (impl_->sequences.empty()) {
      detail::Throw(ErrorCode::kInvalidInput, "a reference needs one record or more");
    }
    return std::move(impl_);
  }

 private:
  std::unique_ptr<GenomeReference::Impl> impl_ = std::make_unique<GenomeReference::Impl>();
  std::unordered_set<std::string> names_;
  std::uint64_t next_offset_ = 0;
};

}  // namespace

GenomeReference::GenomeReference(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
GenomeReference::GenomeReference(GenomeReference&&) noexcept = default;
GenomeReference& GenomeReference::operator=(GenomeReference&&) noexcept = default;
GenomeReference::~GenomeReference() = default;

GenomeReference GenomeReference::FromRecords(std::vector<SequenceRecord> records) {
  if (records.empty()) {
    detail::Throw(ErrorCode::kInvalidInput, "a reference needs one record or more");
  }
  Builder builder;
  for (auto& r : records) builder.Add(std::move(r.name), std::move(r.description), r.sequence);
  return GenomeReference(builder.Finish());
}

GenomeReference GenomeReference::FromEncoded(const std::uint8_t* bases, const std::uint64_t* lengths,
                                             std::uint64_t count) {
  if (count == 0) {
    detail::Throw(ErrorCode::kInvalidInput, "a reference needs one record or more");
  }
  auto impl = std::make_unique<Impl>();
  std::uint64_t offset = 0, in = 0;
  for (std::uint64_t c = 0; c < count; ++c) {
    if (lengths[c] == 0) detail::Throw(ErrorCode::kInvalidInput, "a record holds no bases");
    SequenceInfo info;
    info.id = static_cast<SequenceId>(c);
    info.name = "contig" + std::to_string(c);
    info.length = lengths[c];
    info.global_offset = offset;
    for (std::uint64_t k = 0; k < lengths[c]; ++k) {
      const std::uint8_t v = bases[in++];
      if (v < detail::kA || v > detail::kN) {
        detail::Throw(ErrorCode::kInvalidInput, "encoded base outside A/C/G/T/N");
      }
      info.ambiguous_bases += v == detail::kN;
      impl->encoded.push_back(v);
    }
    impl->encoded.push_back(detail::kSeparator);
    impl->total_bases += info.length;
    impl->ambiguous_bases += info.ambiguous_bases;
    offset += info.length + 1;
    impl->sequences.push_back(std::move(info));
  }
  impl->fingerprint = detail::ContentFingerprint(impl->encoded.data(), impl->encoded.size());
  return GenomeReference(std::move(impl));
}

GenomeReference GenomeReference::FromFasta(const std::filesystem::path& path) {
  gzFile file = gzopen(path.c_str(), "rb");
  if (file == nullptr) detail::Throw(ErrorCode::kIoError, "FASTA file could not be opened: " + path.string());
  std::unique_ptr<gzFile_s, int (*)(gzFile)> guard(file, gzclose);
  Builder builder;
  std::string name, description, sequence;
  bool in_record = false;
  std::string line;
  char buffer[1 << 16];
  bool eof = false;
  auto flush = [&]() {
    if (in_record) builder.Add(std::move(name), std::move(description), sequence);
    name.clear();
    description.clear();
    sequence.clear();
  };
  while (!eof) {
    line.clear();
    // Read one full line (lines may exceed the buffer).
    while (true) {
      if (gzgets(file, buffer, sizeof(buffer)) == nullptr) {
        int err = 0;
        gzerror(file, &err);
        if (err != Z_OK && err != Z_STREAM_END) {
          detail::Throw(ErrorCode::kInvalidInput, "not FASTA (no '>' header before the bases): " + path.string());
        }
        eof = true;
        break;
      }
      line += buffer;
      if (!line.empty() && line.back() == '\n') break;
    }
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    if (line.empty()) continue;
    if (line[0] == '>') {
      flush();
      in_record = true;
      const auto header = std::string_view(line).substr(1);
      const auto space = header.find_first_of(" \t");
      name = std::string(header.substr(0, space));
      if (space != std::string_view::npos) {
        const auto rest = header.find_first_not_of(" \t", space);
        if (rest != std::string_view::npos) description = std::string(header.substr(rest));
      }
      continue;
    }
    if (!in_record) detail::Throw(ErrorCode::kInvalidInput, "not FASTA (no '>' header before the bases): " + path.string());
    for (const char ch : line) {
      if (ch != ' ' && ch != '\t') sequence.push_back(ch);
    }
  }
  flush();
  return GenomeReference(builder.Finish());
}

std::uint64_t GenomeReference::SequenceCount() const noexcept { return impl_->sequences.size(); }
std::uint64_t GenomeReference::TotalBases() const noexcept { return impl_->total_bases; }
std::uint64_t GenomeReference::AmbiguousBases() const noexcept { return impl_->ambiguous_bases; }
std::uint64_t GenomeReference::Fingerprint() const noexcept { return impl_->fingerprint; }
const std::vector<SequenceInfo>& GenomeReference::Sequences() const noexcept { return impl_->sequences; }
const std::vector<std::uint8_t>& GenomeReference::Encoded() const noexcept { return impl_->encoded; }

}  // namespace sufkit
