// C entry points over the compiled reference library: build with the
// reference's own SuffixArray::Build (divsufsort backend) and Save a .sufidx
// 1.4 file, or Load one and report its SA/ISA/LCP; run the reference's
// FastPrefixIndex and FmIndex.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target ref-full)
// against the reference sources where they lie; nothing of the reference is
// copied into this repository.  Used by tests/golden/make_golden.py to write
// golden .sufidx files, so the writer of this repository can be compared
// with the reference's byte for byte.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <string_view>
#include <vector>

#include <sufkit/fm_index.hpp>
#include <sufkit/sufkit.hpp>

#include "coordinate_storage.hpp"
#include "fast_prefix_index.hpp"

// Runs f: 0, or the reference's ErrorCode + 1 when it throws sufkit::Error.
namespace {

template <class F>
int RefGuard(F&& f) {
  try {
    f();
    return 0;
  } catch (const sufkit::Error& e) {
    return static_cast<int>(e.Code()) + 1;
  } catch (const std::exception&) {
    return 100;
  }
}

}  // namespace

extern "C" {

// acceleration: SaAcceleration value; coordinate_width 0/32/64; storage
// 0/32/40/48/64; profile: SaResourceProfile (0 Fast, 1 Low-memory).
// seconds (optional): normalization, Build, Save wall time of the
// reference's own calls.
int sfk_ref_write_sufidx_profile(const char* const* names, const char* const* sequences,
                                 const std::uint64_t* lengths, std::uint64_t count, const char* path,
                                 int coordinate_width, int storage_width, int acceleration, int profile,
                                 double* seconds, char* error, std::size_t capacity) {
  try {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    std::vector<sufkit::SequenceRecord> records(count);
    for (std::uint64_t i = 0; i < count; ++i) {
      records[i].name = names[i];
      records[i].sequence.assign(sequences[i], lengths[i]);
    }
    const auto reference = sufkit::GenomeReference::FromRecords(std::move(records));
    sufkit::SuffixArrayBuildOptions options;
    options.backend = sufkit::SaBackend::kDivsufsort;
    options.coordinate_width = static_cast<sufkit::CoordinateWidth>(coordinate_width);
    options.storage_width = static_cast<sufkit::CoordinateStorageWidth>(storage_width);
    options.acceleration = static_cast<sufkit::SaAcceleration>(acceleration);
    options.resource_profile = static_cast<sufkit::SaResourceProfile>(profile);
    const auto t1 = Clock::now();
    const auto index = sufkit::SuffixArray::Build(reference, options);
    const auto t2 = Clock::now();
    sufkit::SaveOptions save;
    save.overwrite = true;
    index.Save(path, save);
    const auto t3 = Clock::now();
    if (seconds != nullptr) {
      seconds[0] = std::chrono::duration<double>(t1 - t0).count();
      seconds[1] = std::chrono::duration<double>(t2 - t1).count();
      seconds[2] = std::chrono::duration<double>(t3 - t2).count();
    }
    return 0;
  } catch (const std::exception& e) {
    if (error != nullptr && capacity > 0) {
      std::strncpy(error, e.what(), capacity - 1);
      error[capacity - 1] = '\0';
    }
    return -1;
  }
}

int sfk_ref_write_sufidx_timed(const char* const* names, const char* const* sequences,
                               const std::uint64_t* lengths, std::uint64_t count, const char* path,
                               int coordinate_width, int storage_width, int acceleration, double* seconds,
                               char* error, std::size_t capacity) {
  return sfk_ref_write_sufidx_profile(names, sequences, lengths, count, path, coordinate_width, storage_width,
                                      acceleration, 0, seconds, error, capacity);
}

int sfk_ref_write_sufidx(const char* const* names, const char* const* sequences,
                         const std::uint64_t* lengths, std::uint64_t count, const char* path,
                         int coordinate_width, int storage_width, int acceleration,
                         char* error, std::size_t capacity) {
  return sfk_ref_write_sufidx_profile(names, sequences, lengths, count, path, coordinate_width, storage_width,
                                      acceleration, 0, nullptr, error, capacity);
}

// The reference's own FastPrefixIndex (fast_prefix_index.cpp:286-382) over
// a complete SA/ISA with default options: its k (from resident_core_bytes),
// and for every k-mer key its Lookup interval ((n, n) when absent), plus
// the indexed k-mers and non-empty entries.  Returns k, or -1 on error.
int sfk_ref_prefix_directory(const std::uint8_t* text, std::uint64_t n, const std::uint64_t* sa,
                             const std::uint64_t* isa, const std::uint64_t* contig_starts,
                             const std::uint64_t* contig_lengths, std::uint64_t contigs,
                             std::uint64_t resident_core_bytes, std::uint64_t* begin, std::uint64_t* end,
                             std::uint64_t* counts) {
  try {
    namespace d = sufkit::detail;
    const std::vector<std::uint8_t> t(text, text + n);
    auto sa_store = d::CoordinateStorage::FromUInt64(std::vector<std::uint64_t>(sa, sa + n),
                                                     sufkit::CoordinateStorageWidth::kBits64, n);
    auto isa_store = d::CoordinateStorage::FromUInt64(std::vector<std::uint64_t>(isa, isa + n),
                                                      sufkit::CoordinateStorageWidth::kBits64, n);
    const std::vector<sufkit::Position> starts(contig_starts, contig_starts + contigs);
    const std::vector<sufkit::Position> lengths(contig_lengths, contig_lengths + contigs);
    const auto index = d::FastPrefixIndex::Build(t, sa_store, isa_store, starts, lengths, resident_core_bytes);
    const std::uint32_t k = index.K();
    if (k > 0) {
      std::vector<std::uint8_t> pattern(k);
      for (std::uint64_t key = 0; key < (std::uint64_t{1} << (2 * k)); ++key) {
        for (std::uint32_t i = 0; i < k; ++i) {
          pattern[i] = static_cast<std::uint8_t>(2 + ((key >> (2 * (k - 1 - i))) & 3u));  // A=2 .. T=5
        }
        const auto interval = index.Lookup(pattern);
        const bool present = interval && !interval->Empty();
        begin[key] = present ? interval->begin : n;
        end[key] = present ? interval->end : n;
      }
      counts[0] = index.IndexedKmers();
      counts[1] = index.NonemptyEntries();
    }
    return static_cast<int>(k);
  } catch (const std::exception&) {
    return -1;
  }
}

// ------------------------------------------------------------ FM index
//
// The reference's own FmIndex (fm_index.cpp: SDSL csa_wt built on the host)
// over a reference from records, for the device FM-index's golden vectors
// and error behaviour (entry points return as RefGuard does).
int sfk_ref_fm_build(const char* const* names, const char* const* sequences, const std::uint64_t* lengths,
                     std::uint64_t count, int backend, void** out) {
  return RefGuard([&] {
    std::vector<sufkit::SequenceRecord> records(count);
    for (std::uint64_t i = 0; i < count; ++i) {
      records[i].name = names[i];
      records[i].sequence.assign(sequences[i], lengths[i]);
    }
    const auto reference = sufkit::GenomeReference::FromRecords(std::move(records));
    sufkit::FmIndexBuildOptions options;
    options.backend = static_cast<sufkit::FmBackend>(backend);
    *out = new sufkit::FmIndex(sufkit::FmIndex::Build(reference, options));
  });
}

void sfk_ref_fm_free(void* fm) { delete static_cast<sufkit::FmIndex*>(fm); }

int sfk_ref_fm_equal_range(const void* fm, const char* pattern, std::uint64_t length, std::uint64_t out[2]) {
  return RefGuard([&] {
    const auto r = static_cast<const sufkit::FmIndex*>(fm)->EqualRange(std::string_view(pattern, length));
    out[0] = r.begin;
    out[1] = r.end;
  });
}

int sfk_ref_fm_count(const void* fm, const char* pattern, std::uint64_t length, int strands, std::uint64_t* out) {
  return RefGuard([&] {
    *out = static_cast<const sufkit::FmIndex*>(fm)->Count(std::string_view(pattern, length),
                                                           static_cast<sufkit::StrandMode>(strands));
  });
}

int sfk_ref_fm_count_batch(const void* fm, const char* patterns, const std::uint64_t* offsets, std::uint64_t count,
                           int strands, std::uint32_t batch_width, std::uint64_t* out) {
  return RefGuard([&] {
    std::vector<std::string_view> list(count);
    for (std::uint64_t i = 0; i < count; ++i) {
      list[i] = std::string_view(patterns + offsets[i], offsets[i + 1] - offsets[i]);
    }
    sufkit::FmBatchOptions options;
    options.strands = static_cast<sufkit::StrandMode>(strands);
    options.batch_width = batch_width;
    const auto counts = static_cast<const sufkit::FmIndex*>(fm)->CountBatch(list, options);
    std::memcpy(out, counts.data(), count * sizeof(std::uint64_t));
  });
}

// info = {total hits, retained hits, truncated}; the first `capacity` hits
// go to ids / positions / lengths / strand codes.
int sfk_ref_fm_locate(const void* fm, const char* pattern, std::uint64_t length, int strands, int has_max,
                      std::uint64_t max_hits, std::uint64_t info[3], std::uint32_t* ids, std::uint64_t* positions,
                      std::uint64_t* lengths, std::uint8_t* strand_codes, std::uint64_t capacity) {
  return RefGuard([&] {
    sufkit::LocateOptions options;
    options.strands = static_cast<sufkit::StrandMode>(strands);
    if (has_max != 0) options.max_hits = max_hits;
    const auto result = static_cast<const sufkit::FmIndex*>(fm)->Locate(std::string_view(pattern, length), options);
    info[0] = result.total_hits;
    info[1] = result.hits.size();
    info[2] = result.truncated ? 1 : 0;
    for (std::uint64_t i = 0; i < result.hits.size() && i < capacity; ++i) {
      ids[i] = result.hits[i].sequence_id;
      positions[i] = result.hits[i].position;
      lengths[i] = result.hits[i].length;
      strand_codes[i] = static_cast<std::uint8_t>(result.hits[i].strand);
    }
  });
}

// FastPrefixIndex::SelectK with default options.
std::uint32_t sfk_ref_prefix_select_k(std::uint64_t suffix_count, std::uint64_t resident_core_bytes) {
  return sufkit::detail::FastPrefixIndex::SelectK(suffix_count, resident_core_bytes);
}

}  // extern "C"
