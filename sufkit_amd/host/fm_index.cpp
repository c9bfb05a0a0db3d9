// FmIndex over a device-resident FM-index built from the GPU SA.
//
// Mirrors the reference's FmIndex (src/fm_index.cpp): Build validates the
// reference as fm_index.cpp:488-524 does, the queries follow
// fm_index.cpp:526-677 (pattern encoding and reverse complement of
// query.cpp:47-80, palindrome handling, batch-width and strand checks, the
// bounded smallest-hit retention and the contig mapping of
// fm_index.cpp:112-181).  The index itself lives on the device: the text,
// its SA (built by sfk_hip_build_device) and the occurrence records of
// sfk_hip_index_build_fm; row ranges come from sfk_hip_index_fm_ranges and
// located positions from sfk_hip_index_sa_rows.
#include <algorithm>
#include <cstring>
#include <limits>
#include <string>
#include <tuple>
#include <utility>

#include "internal.hpp"
#include "sufkit_hip.h"

namespace sufkit {

namespace {

constexpr std::uint32_t kAutomaticBatchWidth = 16;
constexpr std::uint32_t kMaximumBatchWidth = 256;

// Device buffer owned by the host object.
struct DevicePtr {
  void* p = nullptr;
  DevicePtr() = default;
  DevicePtr(const DevicePtr&) = delete;
  DevicePtr& operator=(const DevicePtr&) = delete;
  ~DevicePtr() {
    if (p != nullptr) sfk_hip_free(p);
  }
  void Allocate(std::uint64_t bytes) {
    if (sfk_hip_malloc(0, bytes, &p) != SFK_HIP_OK) {
      p = nullptr;
      detail::Throw(ErrorCode::kBuildFailure, "device allocation failed");
    }
  }
};

void CheckHip(int status, const char* what) {
  if (status == SFK_HIP_OK) return;
  const ErrorCode code = status == SFK_HIP_INVALID_INPUT       ? ErrorCode::kInvalidInput
                         : status == SFK_HIP_UNSUPPORTED_BACKEND ? ErrorCode::kUnsupportedBackend
                                                                 : ErrorCode::kBuildFailure;
  detail::Throw(code, what);
}

// query.cpp:47-65: A/C/G/T in either case, as codes A=0 .. T=3.
std::vector<std::uint8_t> EncodePattern(std::string_view pattern) {
  if (pattern.empty()) detail::Throw(ErrorCode::kInvalidInput, "an empty pattern was given");
  std::vector<std::uint8_t> out(pattern.size());
  for (std::size_t i = 0; i < pattern.size(); ++i) {
    switch (pattern[i]) {
      case 'A': case 'a': out[i] = 0; break;
      case 'C': case 'c': out[i] = 1; break;
      case 'G': case 'g': out[i] = 2; break;
      case 'T': case 't': out[i] = 3; break;
      default: detail::Throw(ErrorCode::kInvalidInput, "patterns may hold only A, C, G and T");
    }
  }
  return out;
}

// Complement of code c is 3 - c (A<->T, C<->G).
bool IsReverseComplementPalindrome(const std::vector<std::uint8_t>& p) {
  for (std::size_t l = 0, r = p.size(); l < r; ++l) {
    --r;
    if (p[l] != 3 - p[r]) return false;
  }
  return true;
}

std::uint32_t EffectiveBatchWidth(std::uint32_t requested) {
  if (requested == 0) return kAutomaticBatchWidth;
  if (requested > kMaximumBatchWidth) {
    detail::Throw(ErrorCode::kInvalidInput, "FM batch width is 0 (automatic) or 1 to 256");
  }
  return requested;
}

struct GlobalHit {
  Position position = 0;
  Strand strand = Strand::kForward;
};

bool HitLess(const GlobalHit& a, const GlobalHit& b) {
  return std::tie(a.position, a.strand) < std::tie(b.position, b.strand);
}

}  // namespace

struct FmIndex::Impl {
  std::vector<SequenceInfo> sequences;
  std::vector<Position> contig_starts, contig_lengths;
  std::uint64_t n = 0;  // text symbols, sentinel included
  std::uint32_t sa_bits = 32;
  DevicePtr text, sa;
  sfk_hip_index* index = nullptr;
  IndexInfo info;

  ~Impl() {
    if (index != nullptr) sfk_hip_index_destroy(index);
  }

  // One device search per (pattern, orientation) lane; (0, 0) when absent.
  std::vector<SuffixRange> Ranges(const std::vector<std::vector<std::uint8_t>>& patterns,
                                  const std::vector<std::size_t>& which,
                                  const std::vector<std::uint8_t>& orientation) const {
    const std::uint64_t count = which.size();
    std::vector<SuffixRange> out(count);
    if (count == 0) return out;
    std::vector<std::uint64_t> offsets(count + 1, 0);
    for (std::uint64_t q = 0; q < count; ++q) offsets[q + 1] = offsets[q] + patterns[which[q]].size();
    std::vector<std::uint8_t> codes(std::max<std::uint64_t>(offsets[count], 1));
    for (std::uint64_t q = 0; q < count; ++q) {
      const auto& p = patterns[which[q]];
      std::memcpy(codes.data() + offsets[q], p.data(), p.size());
    }
    DevicePtr d_codes, d_offsets, d_orientation, d_ranges;
    d_codes.Allocate(codes.size());
    d_offsets.Allocate(offsets.size() * 8);
    d_orientation.Allocate(count);
    d_ranges.Allocate(count * 16);
    CheckHip(sfk_hip_memcpy_h2d(d_codes.p, codes.data(), codes.size()), "pattern upload failed");
    CheckHip(sfk_hip_memcpy_h2d(d_offsets.p, offsets.data(), offsets.size() * 8), "pattern upload failed");
    CheckHip(sfk_hip_memcpy_h2d(d_orientation.p, orientation.data(), count), "pattern upload failed");
    CheckHip(sfk_hip_index_fm_ranges(index, static_cast<const std::uint8_t*>(d_codes.p),
                                     static_cast<const std::uint64_t*>(d_offsets.p),
                                     static_cast<const std::uint8_t*>(d_orientation.p), count,
                                     static_cast<std::uint64_t*>(d_ranges.p), nullptr),
             "FM backward search failed");
    std::vector<std::uint64_t> ranges(2 * count);
    CheckHip(sfk_hip_memcpy_d2h(ranges.data(), d_ranges.p, count * 16), "range download failed");
    for (std::uint64_t q = 0; q < count; ++q) out[q] = SuffixRange{ranges[2 * q], ranges[2 * q + 1]};
    return out;
  }

  SuffixRange Range(const std::vector<std::uint8_t>& pattern, bool reverse_complement) const {
    return Ranges({pattern}, {0}, {static_cast<std::uint8_t>(reverse_complement ? 1 : 0)})[0];
  }

  // fm_index.cpp:112-131,470-480: every row's position, or the `limit`
  // smallest (position, strand) pairs when a limit is set.
  void Collect(SuffixRange range, Strand strand, bool limited, std::uint64_t limit,
               std::vector<GlobalHit>& hits) const {
    constexpr std::uint64_t kChunk = 1u << 20;
    std::vector<std::uint64_t> rows;
    for (std::uint64_t at = range.begin; at < range.end; at += kChunk) {
      const std::uint64_t end = std::min(range.end, at + kChunk);
      rows.resize(end - at);
      CheckHip(sfk_hip_index_sa_rows(index, at, end, rows.data()), "SA row download failed");
      for (const std::uint64_t position : rows) {
        const GlobalHit hit{position, strand};
        if (!limited) {
          hits.push_back(hit);
        } else if (limit == 0) {
          return;
        } else if (hits.size() < limit) {
          hits.push_back(hit);
          std::push_heap(hits.begin(), hits.end(), HitLess);
        } else if (HitLess(hit, hits.front())) {
          std::pop_heap(hits.begin(), hits.end(), HitLess);
          hits.back() = hit;
          std::push_heap(hits.begin(), hits.end(), HitLess);
        }
      }
    }
  }

  // fm_index.cpp:133-181.
  QueryResult Finalize(std::vector<GlobalHit> hits, std::uint64_t total, std::uint64_t length) const {
    std::sort(hits.begin(), hits.end(), HitLess);
    std::size_t size = 0;
    for (std::size_t i = 0; i < hits.size(); ++i) {
      if (size != 0 && hits[size - 1].position == hits[i].position) {
        if (hits[size - 1].strand != hits[i].strand) hits[size - 1].strand = Strand::kBoth;
        continue;
      }
      hits[size++] = hits[i];
    }
    hits.resize(size);
    if (!hits.empty() && (contig_starts.empty() || contig_starts.size() != contig_lengths.size())) {
      detail::Throw(ErrorCode::kCorruptIndex, "contig table of the FM-index is inconsistent");
    }
    QueryResult result;
    result.total_hits = total;
    result.hits.reserve(hits.size());
    std::size_t contig = 0;
    for (const auto& hit : hits) {
      while (contig + 1 < contig_starts.size() && contig_starts[contig + 1] <= hit.position) ++contig;
      if (hit.position < contig_starts[contig]) {
        detail::Throw(ErrorCode::kCorruptIndex, "FM-index hit is outside a reference contig");
      }
      const Position local = hit.position - contig_starts[contig];
      if (local > contig_lengths[contig] || length > contig_lengths[contig] - local) {
        detail::Throw(ErrorCode::kCorruptIndex, "FM-index hit is outside a reference contig");
      }
      result.hits.push_back(Match{static_cast<SequenceId>(contig), local, length, hit.strand});
    }
    result.truncated = result.hits.size() < result.total_hits;
    return result;
  }
};

FmIndex::FmIndex(std::unique_ptr<Impl> impl) : impl_(std::move(impl)) {}
FmIndex::FmIndex(FmIndex&&) noexcept = default;
FmIndex& FmIndex::operator=(FmIndex&&) noexcept = default;
FmIndex::~FmIndex() = default;

FmIndex FmIndex::Build(const GenomeReference& reference, const FmIndexBuildOptions& options) {
  const auto& encoded = reference.Encoded();
  if (encoded.empty()) detail::Throw(ErrorCode::kInvalidInput, "the reference has no text");
  if (std::find(encoded.begin(), encoded.end(), detail::kSentinel) != encoded.end()) {
    detail::Throw(ErrorCode::kInvalidInput, "FM construction input contains a zero symbol");
  }
  switch (options.backend) {
    case FmBackend::kSdslCsaWtHuff:
    case FmBackend::kSdslCsaWtBalanced:
    case FmBackend::kSdslCsaWtEpr:
      break;
    case FmBackend::kSdslCsaSada:
      detail::Throw(ErrorCode::kUnsupportedBackend, "the sdsl-csa-sada FM backend is not available");
    default:
      detail::Throw(ErrorCode::kUnsupportedBackend, "FM backend id is not one this library knows");
  }
  if (!HipBackendAvailable()) detail::Throw(ErrorCode::kUnsupportedBackend, "no HIP device is visible");

  auto impl = std::make_unique<Impl>();
  const std::uint64_t n = encoded.size() + 1;
  impl->n = n;
  impl->sa_bits = n - 1 <= std::numeric_limits<std::uint32_t>::max() ? 32 : 64;
  impl->sequences = reference.Sequences();
  for (const auto& s : impl->sequences) {
    impl->contig_starts.push_back(s.global_offset);
    impl->contig_lengths.push_back(s.length);
  }
  impl->text.Allocate(n + 64);
  impl->sa.Allocate(n * (impl->sa_bits / 8));
  CheckHip(sfk_hip_memcpy_h2d(impl->text.p, encoded.data(), encoded.size()), "text upload failed");
  const std::uint8_t sentinel = detail::kSentinel;
  CheckHip(sfk_hip_memcpy_h2d(static_cast<std::uint8_t*>(impl->text.p) + encoded.size(), &sentinel, 1),
           "text upload failed");
  char error[256] = {0};
  sfk_hip_builder* builder = nullptr;
  CheckHip(sfk_hip_builder_create(0, &builder, error, sizeof(error)), error);
  sfk_hip_stats stats{};
  const int status = sfk_hip_build_device(builder, static_cast<const std::uint8_t*>(impl->text.p), n,
                                          impl->sa_bits, 1, 0, impl->sa.p, impl->sa_bits, nullptr, 0, nullptr, 0,
                                          &stats, error, sizeof(error));
  sfk_hip_builder_destroy(builder);
  CheckHip(status, error);
  CheckHip(sfk_hip_index_create(0, static_cast<const std::uint8_t*>(impl->text.p), n, impl->sa.p, impl->sa_bits,
                                &impl->index, error, sizeof(error)),
           error);
  CheckHip(sfk_hip_index_build_fm(impl->index, error, sizeof(error)), error);

  IndexInfo& info = impl->info;
  info.library_version = "0.3.0";
  info.backend = "hip-fm";
  info.backend_signature = "hip-fm: device BWT occurrence records over the GPU SA";
  info.coordinate_width = static_cast<std::uint8_t>(impl->sa_bits);
  info.stored_coordinate_width = static_cast<std::uint8_t>(impl->sa_bits);
  info.sequence_count = reference.SequenceCount();
  info.total_bases = reference.TotalBases();
  info.text_symbols = n;
  info.suffix_count = n;
  info.ambiguous_bases = reference.AmbiguousBases();
  info.fingerprint = reference.Fingerprint();
  info.text_bytes = n;
  info.sa_bytes = n * (impl->sa_bits / 8);
  info.auxiliary_bytes = sfk_hip_index_fm_bytes(impl->index);
  info.resident_core_bytes = info.text_bytes + info.sa_bytes + info.auxiliary_bytes;
  return FmIndex(std::move(impl));
}

FmIndex FmIndex::Load(const std::filesystem::path&) {
  detail::Throw(ErrorCode::kUnsupportedBackend, "FM-index persistence (SDSL CSA payload) is not part of this build");
}

void FmIndex::Save(const std::filesystem::path&, const SaveOptions&) const {
  detail::Throw(ErrorCode::kUnsupportedBackend, "FM-index persistence (SDSL CSA payload) is not part of this build");
}

SuffixRange FmIndex::EqualRange(std::string_view pattern) const {
  return impl_->Range(EncodePattern(pattern), false);
}

std::vector<SuffixRange> FmIndex::EqualRangeBatch(const std::vector<std::string_view>& patterns,
                                                  std::uint32_t batch_width) const {
  (void)EffectiveBatchWidth(batch_width);
  std::vector<std::vector<std::uint8_t>> encoded;
  encoded.reserve(patterns.size());
  for (const auto p : patterns) encoded.push_back(EncodePattern(p));
  std::vector<std::size_t> which(encoded.size());
  for (std::size_t i = 0; i < which.size(); ++i) which[i] = i;
  return impl_->Ranges(encoded, which, std::vector<std::uint8_t>(which.size(), 0));
}

std::uint64_t FmIndex::Count(std::string_view pattern, StrandMode strands) const {
  const auto encoded = EncodePattern(pattern);
  switch (strands) {
    case StrandMode::kForward:
      return impl_->Range(encoded, false).Size();
    case StrandMode::kReverseComplement:
      return impl_->Range(encoded, true).Size();
    case StrandMode::kBoth: {
      if (IsReverseComplementPalindrome(encoded)) return impl_->Range(encoded, false).Size();
      const auto r = impl_->Ranges({encoded}, {0, 0}, {0, 1});
      return r[0].Size() + r[1].Size();
    }
  }
  detail::Throw(ErrorCode::kInvalidInput, "count: strand mode must be forward, reverse complement or both");
}

std::vector<std::uint64_t> FmIndex::CountBatch(const std::vector<std::string_view>& patterns,
                                               const FmBatchOptions& options) const {
  (void)EffectiveBatchWidth(options.batch_width);
  if (options.strands != StrandMode::kForward && options.strands != StrandMode::kReverseComplement &&
      options.strands != StrandMode::kBoth) {
    detail::Throw(ErrorCode::kInvalidInput, "batch count: strand mode must be forward, reverse complement or both");
  }
  std::vector<std::vector<std::uint8_t>> encoded;
  encoded.reserve(patterns.size());
  for (const auto p : patterns) encoded.push_back(EncodePattern(p));
  std::vector<std::size_t> which;
  std::vector<std::uint8_t> orientation;
  for (std::size_t i = 0; i < encoded.size(); ++i) {
    if (options.strands != StrandMode::kReverseComplement) {
      which.push_back(i);
      orientation.push_back(0);
    }
    // A reverse-complement palindrome is one search in both-strand mode.
    if (options.strands == StrandMode::kReverseComplement ||
        (options.strands == StrandMode::kBoth && !IsReverseComplementPalindrome(encoded[i]))) {
      which.push_back(i);
      orientation.push_back(1);
    }
  }
  const auto ranges = impl_->Ranges(encoded, which, orientation);
  std::vector<std::uint64_t> counts(patterns.size(), 0);
  for (std::size_t q = 0; q < which.size(); ++q) counts[which[q]] += ranges[q].Size();
  return counts;
}

QueryResult FmIndex::Locate(std::string_view pattern, const LocateOptions& options) const {
  const auto encoded = EncodePattern(pattern);
  SuffixRange first, second;
  Strand first_strand = Strand::kForward, second_strand = Strand::kForward;
  bool has_second = false;
  switch (options.strands) {
    case StrandMode::kForward:
      first = impl_->Range(encoded, false);
      break;
    case StrandMode::kReverseComplement:
      first = impl_->Range(encoded, true);
      first_strand = Strand::kReverseComplement;
      break;
    case StrandMode::kBoth:
      if (IsReverseComplementPalindrome(encoded)) {
        first = impl_->Range(encoded, false);
        first_strand = Strand::kBoth;
      } else {
        const auto r = impl_->Ranges({encoded}, {0, 0}, {0, 1});
        first = r[0];
        second = r[1];
        second_strand = Strand::kReverseComplement;
        has_second = true;
      }
      break;
    default:
      detail::Throw(ErrorCode::kInvalidInput, "locate: strand mode must be forward, reverse complement or both");
  }
  const std::uint64_t total = first.Size() + (has_second ? second.Size() : 0);
  // max_hits = 0 asks for the exact count without coordinates.
  if (options.max_hits && *options.max_hits == 0) return impl_->Finalize({}, total, encoded.size());
  const bool limited = options.max_hits && *options.max_hits < total;
  const std::uint64_t limit = limited ? *options.max_hits : 0;
  std::vector<GlobalHit> hits;
  hits.reserve(static_cast<std::size_t>(limited ? limit : total));
  impl_->Collect(first, first_strand, limited, limit, hits);
  if (has_second) impl_->Collect(second, second_strand, limited, limit, hits);
  return impl_->Finalize(std::move(hits), total, encoded.size());
}

SequenceInfo FmIndex::GetSequenceInfo(SequenceId id) const {
  if (static_cast<std::size_t>(id) >= impl_->sequences.size()) {
    detail::Throw(ErrorCode::kInvalidInput, "no contig has this id");
  }
  return impl_->sequences[id];
}

const IndexInfo& FmIndex::GetInfo() const noexcept { return impl_->info; }

std::uint32_t FmIndex::PrefixDirectoryK() const noexcept { return sfk_hip_index_prefix_k(impl_->index); }

std::uint64_t FmIndex::DeviceBytes() const noexcept { return sfk_hip_index_device_bytes(impl_->index); }

}  // namespace sufkit
