// C entry points over the host library, for language bindings and tests.
// Every function returns 0 or ErrorCode + 1 and fills `error` on failure.
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "internal.hpp"

namespace {

int Fail(const std::exception& e, int status, char* error, size_t capacity) {
  if (error != nullptr && capacity > 0) {
    const std::string msg = e.what();
    const size_t len = std::min(msg.size(), capacity - 1);
    std::memcpy(error, msg.data(), len);
    error[len] = '\0';
  }
  return status;
}

template <class F>
int Guard(char* error, size_t capacity, F&& f) {
  try {
    f();
    return 0;
  } catch (const sufkit::Error& e) {
    return Fail(e, static_cast<int>(e.Code()) + 1, error, capacity);
  } catch (const std::bad_alloc& e) {
    return Fail(e, static_cast<int>(sufkit::ErrorCode::kBuildFailure) + 1, error, capacity);
  } catch (const std::exception& e) {
    return Fail(e, static_cast<int>(sufkit::ErrorCode::kIoError) + 1, error, capacity);
  }
}

}  // namespace

extern "C" {

struct sfk_reference {
  sufkit::GenomeReference value;
};
struct sfk_suffix_array {
  sufkit::SuffixArray value;
};

struct sfk_build_options {
  int backend;             // SaBackend
  int coordinate_width;    // 0, 32, 64
  int storage_width;       // 0, 32, 40, 48, 64
  int resource_profile;    // 0 fast, 1 low-memory
  int acceleration;        // SaAcceleration
  uint32_t threads;
  uint32_t sampling_rate;
};

struct sfk_build_stats {
  double sa_seconds, storage_compaction_seconds, isa_seconds, lcp_seconds, text_prepare_seconds,
      lcp_finalize_seconds, total_seconds, device_seconds, transfer_seconds, prefix_directory_seconds,
      call_seconds;
};

struct sfk_index_info {
  char backend[32];
  char backend_signature[96];
  char format_version[8];
  uint8_t coordinate_width, stored_coordinate_width, acceleration, lcp_encoding;
  uint32_t sa_sampling_rate;
  uint64_t sequence_count, total_bases, text_symbols, suffix_count, ambiguous_bases, fingerprint,
      serialized_bytes, text_bytes, sa_bytes, isa_bytes, lcp_bytes, resident_core_bytes, auxiliary_bytes;
  uint32_t prefix_directory_k;
};

int sfk_reference_from_records(const char* const* names, const char* const* sequences,
                               const uint64_t* lengths, uint64_t count, sfk_reference** out,
                               char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    std::vector<sufkit::SequenceRecord> records(count);
    for (uint64_t i = 0; i < count; ++i) {
      records[i].name = names[i];
      records[i].sequence.assign(sequences[i], lengths[i]);
    }
    *out = new sfk_reference{sufkit::GenomeReference::FromRecords(std::move(records))};
  });
}

int sfk_reference_from_fasta(const char* path, sfk_reference** out, char* error, size_t capacity) {
  return Guard(error, capacity,
               [&] { *out = new sfk_reference{sufkit::GenomeReference::FromFasta(path)}; });
}

int sfk_reference_from_encoded(const uint8_t* bases, const uint64_t* lengths, uint64_t count,
                               sfk_reference** out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    *out = new sfk_reference{sufkit::GenomeReference::FromEncoded(bases, lengths, count)};
  });
}

void sfk_reference_free(sfk_reference* r) { delete r; }

void sfk_reference_summary(const sfk_reference* r, uint64_t out[5]) {
  out[0] = r->value.SequenceCount();
  out[1] = r->value.TotalBases();
  out[2] = r->value.AmbiguousBases();
  out[3] = r->value.Fingerprint();
  out[4] = r->value.Encoded().size() + 1;  // text symbols incl. sentinel
}

// Logical text (encoded bytes + sentinel) into out[text_symbols].
void sfk_reference_text(const sfk_reference* r, uint8_t* out) {
  const auto& e = r->value.Encoded();
  std::memcpy(out, e.data(), e.size());
  out[e.size()] = 0;
}

// Per contig: global offset, length, ambiguous bases.
void sfk_reference_contigs(const sfk_reference* r, uint64_t* out) {
  const auto& s = r->value.Sequences();
  for (size_t i = 0; i < s.size(); ++i) {
    out[3 * i] = s[i].global_offset;
    out[3 * i + 1] = s[i].length;
    out[3 * i + 2] = s[i].ambiguous_bases;
  }
}

int sfk_sa_build(const sfk_reference* r, const sfk_build_options* o, sfk_suffix_array** out,
                 sfk_build_stats* stats, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::SuffixArrayBuildOptions options;
    sufkit::SuffixArrayBuildStatistics s;
    if (o != nullptr) {
      options.backend = static_cast<sufkit::SaBackend>(o->backend);
      options.coordinate_width = static_cast<sufkit::CoordinateWidth>(o->coordinate_width);
      options.storage_width = static_cast<sufkit::CoordinateStorageWidth>(o->storage_width);
      options.resource_profile = static_cast<sufkit::SaResourceProfile>(o->resource_profile);
      options.acceleration = static_cast<sufkit::SaAcceleration>(o->acceleration);
      options.threads = o->threads;
      options.sampling_rate = o->sampling_rate;
    }
    options.statistics = &s;
    *out = new sfk_suffix_array{sufkit::SuffixArray::Build(r->value, options)};
    if (stats != nullptr) {
      *stats = sfk_build_stats{s.sa_seconds, s.storage_compaction_seconds, s.isa_seconds,
                               s.lcp_seconds, s.text_prepare_seconds, s.lcp_finalize_seconds,
                               s.total_seconds, s.device_seconds, s.transfer_seconds,
                               s.prefix_directory_seconds, s.call_seconds};
    }
  });
}

int sfk_sa_save(const sfk_suffix_array* sa, const char* path, int overwrite, char* error,
                size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::SaveOptions opts;
    opts.overwrite = overwrite != 0;
    sa->value.Save(path, opts);
  });
}

// SaveDeviceIndex: the file Build(options) + Save would write, streamed from
// device arrays.  stats (optional): seconds, write seconds, validate
// seconds, bytes.
int sfk_save_device_index(const sfk_reference* r, int device, const void* d_text, const void* d_sa,
                          uint32_t sa_bits, const void* d_isa, uint32_t isa_bits, const void* d_lcp,
                          uint32_t lcp_bits, uint32_t construction_bits, const sfk_build_options* o,
                          const char* path, int overwrite, double stats[4], char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::SuffixArrayBuildOptions options;
    if (o != nullptr) {
      options.storage_width = static_cast<sufkit::CoordinateStorageWidth>(o->storage_width);
      options.resource_profile = static_cast<sufkit::SaResourceProfile>(o->resource_profile);
      options.acceleration = static_cast<sufkit::SaAcceleration>(o->acceleration);
      options.sampling_rate = o->sampling_rate;
    }
    sufkit::DeviceIndexBuffers b;
    b.device = device;
    b.text = d_text;
    b.sa = d_sa;
    b.sa_bits = sa_bits;
    b.isa = d_isa;
    b.isa_bits = isa_bits;
    b.lcp = d_lcp;
    b.lcp_bits = lcp_bits;
    b.construction_width = static_cast<sufkit::CoordinateWidth>(construction_bits);
    sufkit::SaveOptions save;
    save.overwrite = overwrite != 0;
    sufkit::DeviceSaveStatistics st;
    sufkit::SaveDeviceIndex(r->value, b, path, save, options, &st);
    if (stats != nullptr) {
      stats[0] = st.seconds;
      stats[1] = st.write_seconds;
      stats[2] = st.validate_seconds;
      stats[3] = static_cast<double>(st.bytes);
    }
  });
}

int sfk_file_crc32(const char* path, uint64_t offset, uint64_t size, uint32_t* out, char* error,
                   size_t capacity) {
  return Guard(error, capacity, [&] { *out = sufkit::detail::CrcFileRange(path, offset, size); });
}

int sfk_sa_load(const char* path, sfk_suffix_array** out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] { *out = new sfk_suffix_array{sufkit::SuffixArray::Load(path)}; });
}

void sfk_sa_free(sfk_suffix_array* sa) { delete sa; }

void sfk_sa_info(const sfk_suffix_array* sa, sfk_index_info* out) {
  const auto& i = sa->value.Info();
  std::memset(out, 0, sizeof(*out));
  std::strncpy(out->backend, i.backend.c_str(), sizeof(out->backend) - 1);
  std::strncpy(out->backend_signature, i.backend_signature.c_str(), sizeof(out->backend_signature) - 1);
  std::strncpy(out->format_version, i.format_version.c_str(), sizeof(out->format_version) - 1);
  out->coordinate_width = i.coordinate_width;
  out->stored_coordinate_width = i.stored_coordinate_width;
  out->acceleration = static_cast<uint8_t>(i.sa_acceleration);
  out->lcp_encoding = static_cast<uint8_t>(i.lcp_encoding);
  out->sa_sampling_rate = i.sa_sampling_rate;
  out->sequence_count = i.sequence_count;
  out->total_bases = i.total_bases;
  out->text_symbols = i.text_symbols;
  out->suffix_count = i.suffix_count;
  out->ambiguous_bases = i.ambiguous_bases;
  out->fingerprint = i.fingerprint;
  out->serialized_bytes = i.serialized_bytes;
  out->text_bytes = i.text_bytes;
  out->sa_bytes = i.sa_bytes;
  out->isa_bytes = i.isa_bytes;
  out->lcp_bytes = i.lcp_bytes;
  out->resident_core_bytes = i.resident_core_bytes;
  out->auxiliary_bytes = i.auxiliary_bytes;
  out->prefix_directory_k = sa->value.PrefixDirectoryK();
}

// Decoded arrays (uint64) of the resident index; return the element count.
uint64_t sfk_sa_copy_sa(const sfk_suffix_array* sa, uint64_t* out) {
  const auto n = sa->value.Size();
  if (out != nullptr) for (uint64_t r = 0; r < n; ++r) out[r] = sa->value.SuffixAt(r);
  return n;
}

uint64_t sfk_sa_copy_isa(const sfk_suffix_array* sa, uint64_t* out) {
  if (!sa->value.HasIsa()) return 0;
  const auto n = sa->value.Size();
  if (out != nullptr) for (uint64_t r = 0; r < n; ++r) out[r] = sa->value.IsaAt(r);
  return n;
}

uint64_t sfk_sa_copy_lcp(const sfk_suffix_array* sa, uint64_t* out) {
  if (!sa->value.HasLcp()) return 0;
  const auto n = sa->value.Size();
  if (out != nullptr) for (uint64_t r = 0; r < n; ++r) out[r] = sa->value.LcpAt(r);
  return n;
}

int sfk_sa_count(const sfk_suffix_array* sa, const char* pattern, uint64_t length, uint64_t* count,
                 char* error, size_t capacity) {
  return Guard(error, capacity,
               [&] { *count = sa->value.Count(std::string_view(pattern, length)); });
}

// ------------------------------------------------------------ FM index

struct sfk_fm_index {
  sufkit::FmIndex value;
};
struct sfk_fm_hits {
  sufkit::QueryResult value;
};

namespace {

std::vector<std::string_view> SplitPatterns(const char* patterns, const uint64_t* offsets, uint64_t count) {
  std::vector<std::string_view> out(count);
  for (uint64_t i = 0; i < count; ++i) out[i] = std::string_view(patterns + offsets[i], offsets[i + 1] - offsets[i]);
  return out;
}

sufkit::StrandMode StrandOf(int strands) {
  if (strands < 0 || strands > 2) {
    sufkit::detail::Throw(sufkit::ErrorCode::kInvalidInput, "invalid strand mode for FM query");
  }
  return static_cast<sufkit::StrandMode>(strands);
}

}  // namespace

int sfk_fm_build(const sfk_reference* r, int backend, sfk_fm_index** out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::FmIndexBuildOptions options;
    options.backend = static_cast<sufkit::FmBackend>(backend);
    *out = new sfk_fm_index{sufkit::FmIndex::Build(r->value, options)};
  });
}

void sfk_fm_free(sfk_fm_index* fm) { delete fm; }

// {text symbols, sequences, total bases, fingerprint, FM device bytes, prefix k, SA width}.
void sfk_fm_summary(const sfk_fm_index* fm, uint64_t out[7]) {
  const auto& i = fm->value.GetInfo();
  out[0] = i.text_symbols;
  out[1] = i.sequence_count;
  out[2] = i.total_bases;
  out[3] = i.fingerprint;
  out[4] = i.auxiliary_bytes;
  out[5] = fm->value.PrefixDirectoryK();
  out[6] = i.coordinate_width;
}

int sfk_fm_equal_range(const sfk_fm_index* fm, const char* pattern, uint64_t length, uint64_t out[2], char* error,
                       size_t capacity) {
  return Guard(error, capacity, [&] {
    const auto r = fm->value.EqualRange(std::string_view(pattern, length));
    out[0] = r.begin;
    out[1] = r.end;
  });
}

// Patterns concatenated; pattern i is patterns[offsets[i] .. offsets[i + 1]).
int sfk_fm_equal_range_batch(const sfk_fm_index* fm, const char* patterns, const uint64_t* offsets, uint64_t count,
                             uint32_t batch_width, uint64_t* out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    const auto ranges = fm->value.EqualRangeBatch(SplitPatterns(patterns, offsets, count), batch_width);
    for (uint64_t i = 0; i < count; ++i) {
      out[2 * i] = ranges[i].begin;
      out[2 * i + 1] = ranges[i].end;
    }
  });
}

int sfk_fm_count(const sfk_fm_index* fm, const char* pattern, uint64_t length, int strands, uint64_t* out,
                 char* error, size_t capacity) {
  return Guard(error, capacity,
               [&] { *out = fm->value.Count(std::string_view(pattern, length), StrandOf(strands)); });
}

int sfk_fm_count_batch(const sfk_fm_index* fm, const char* patterns, const uint64_t* offsets, uint64_t count,
                       int strands, uint32_t batch_width, uint64_t* out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::FmBatchOptions options;
    options.strands = StrandOf(strands);
    options.batch_width = batch_width;
    const auto counts = fm->value.CountBatch(SplitPatterns(patterns, offsets, count), options);
    std::memcpy(out, counts.data(), count * sizeof(uint64_t));
  });
}

// has_max != 0: keep the max_hits smallest hits (LocateOptions::max_hits).
int sfk_fm_locate(const sfk_fm_index* fm, const char* pattern, uint64_t length, int strands, int has_max,
                  uint64_t max_hits, sfk_fm_hits** out, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::LocateOptions options;
    options.strands = StrandOf(strands);
    if (has_max != 0) options.max_hits = max_hits;
    *out = new sfk_fm_hits{fm->value.Locate(std::string_view(pattern, length), options)};
  });
}

// {total hits, retained hits, truncated}.
void sfk_fm_hits_summary(const sfk_fm_hits* h, uint64_t out[3]) {
  out[0] = h->value.total_hits;
  out[1] = h->value.hits.size();
  out[2] = h->value.truncated ? 1 : 0;
}

void sfk_fm_hits_get(const sfk_fm_hits* h, uint32_t* sequence_ids, uint64_t* positions, uint64_t* lengths,
                     uint8_t* strands) {
  const auto& hits = h->value.hits;
  for (size_t i = 0; i < hits.size(); ++i) {
    sequence_ids[i] = hits[i].sequence_id;
    positions[i] = hits[i].position;
    lengths[i] = hits[i].length;
    strands[i] = static_cast<uint8_t>(hits[i].strand);
  }
}

void sfk_fm_hits_free(sfk_fm_hits* h) { delete h; }

int sfk_fm_save(const sfk_fm_index* fm, const char* path, int overwrite, char* error, size_t capacity) {
  return Guard(error, capacity, [&] {
    sufkit::SaveOptions opts;
    opts.overwrite = overwrite != 0;
    fm->value.Save(path, opts);
  });
}

int sfk_hip_backend_available(void) { return sufkit::HipBackendAvailable() ? 1 : 0; }

// The width/backend policy SuffixArray::Build applies, without building:
// out = {backend, construction width, stored SA width, ISA width, LCP width}.
int sfk_resolve_build_policy(int backend, int coordinate_width, int storage_width, int profile,
                             uint64_t text_symbols, uint32_t threads, int out[5], char* error,
                             size_t capacity) {
  return Guard(error, capacity, [&] {
    using namespace sufkit;
    const auto b = detail::ResolveSaBackend(static_cast<SaBackend>(backend), text_symbols, threads,
                                            HipBackendAvailable());
    const auto w = detail::ResolveSaCoordinateWidth(b, static_cast<CoordinateWidth>(coordinate_width),
                                                    text_symbols);
    if (!detail::SaConstructionCanRepresent(b, w, text_symbols)) {
      detail::Throw(ErrorCode::kInvalidInput, "text is longer than the chosen construction width can index");
    }
    const auto stored = detail::ResolveProfileStorageWidth(
        static_cast<CoordinateStorageWidth>(storage_width), static_cast<SaResourceProfile>(profile),
        text_symbols);
    out[0] = static_cast<int>(b);
    out[1] = static_cast<int>(w);
    out[2] = static_cast<int>(stored);
    out[3] = static_cast<int>(detail::ResolveAuxiliaryStorageWidth(stored, text_symbols));
    out[4] = text_symbols - 1U <= 0xFFFFFFFFull ? 32 : 64;
  });
}

// .sufidx container self-check (CRC of header and every section).
int sfk_sufidx_verify(const char* path, char* error, size_t capacity) {
  return Guard(error, capacity, [&] { (void)sufkit::detail::ReadContainer(path); });
}

}  // extern "C"
