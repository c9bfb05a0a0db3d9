/*
 * sufkit_hip.h -- C ABI of the MI355X (gfx950) suffix-index construction
 * backend.  Plain pointers and sizes only.
 *
 * This is the third construction backend beside the reference's private
 * adapters
 *   detail::BuildDivsufsort32/64  (src/divsufsort_backend.hpp:27-33)
 *   detail::BuildCaps32/64        (src/caps_backend.hpp:93-99)
 * which SuffixArray::Build selects by (SaBackend, CoordinateWidth)
 * (src/suffix_array.cpp:4182-4236).  Like them it receives the logical text
 * (encoded reference bytes followed by the sentinel 0, suffix_array.cpp:
 * 4107-4117) and returns the complete SA, its ISA and the raw LCP of the Fast
 * profile.  Unlike them it validates the permutation/inverse on the device
 * (the checks of RepackSuffixArray, suffix_array.cpp:852-881) and writes the
 * SA/ISA directly in the requested stored width, so the caller need not redo
 * O(n) random CPU work.
 *
 * Status codes are the reference's ErrorCode (include/sufkit/types.hpp:
 * 708-722) plus one: SFK_HIP_OK = 0, otherwise ErrorCode + 1.
 */
#ifndef SUFKIT_HIP_H_
#define SUFKIT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFK_HIP_ABI_VERSION 6

enum sfk_hip_status {
  SFK_HIP_OK = 0,
  SFK_HIP_INVALID_INPUT = 1,       /* ErrorCode::kInvalidInput */
  SFK_HIP_IO_ERROR = 2,            /* ErrorCode::kIoError */
  SFK_HIP_UNSUPPORTED_BACKEND = 3, /* ErrorCode::kUnsupportedBackend */
  SFK_HIP_CORRUPT_INDEX = 4,       /* ErrorCode::kCorruptIndex */
  SFK_HIP_VERSION_MISMATCH = 5,    /* ErrorCode::kVersionMismatch */
  SFK_HIP_BUILD_FAILURE = 6        /* ErrorCode::kBuildFailure */
};

/* Longest text (symbols, sentinel included) any entry point accepts:
 * 2^32 - 2^16.  Kernels that take one element per work-item launch at most
 * this many work-items, within the 32-bit grid size of a dispatch, and the
 * workspace of a 64-bit construction of a longer text (over 100 bytes per
 * symbol) would not fit one MI355X's 288 GB anyway.  Longer texts give
 * SFK_HIP_UNSUPPORTED_BACKEND. */
#define SFK_HIP_MAX_TEXT_SYMBOLS 4294901760ull

/* Build flags (retain_isa / retain_lcp of divsufsort_backend.hpp:27-33). */
#define SFK_HIP_RETAIN_ISA 1u
#define SFK_HIP_RETAIN_LCP 2u
/* Record per-kernel HIP events (for roofline accounting); slower. */
#define SFK_HIP_PROFILE_KERNELS 4u

/* Per-build statistics.  Times are HIP-event times on the build stream. */
typedef struct sfk_hip_stats {
  double sa_seconds;        /* pack + prefix doubling (SA and ranks=ISA) */
  double isa_seconds;       /* ISA validation / narrowing */
  double lcp_seconds;       /* PLCP + LCP permutation */
  double device_seconds;    /* whole device build */
  double h2d_seconds;       /* host text -> device (host entry point only) */
  double d2h_seconds;       /* device results -> host (host entry point only) */
  uint64_t refine_steps;    /* sort+compact steps */
  uint64_t doubling_rounds; /* rounds with h >= 32 */
  uint64_t active_total;    /* sum over steps of active suffixes */
  uint64_t sort_passes;     /* LSD digit passes actually run */
  uint64_t lce_long_items;  /* LCEs handed to the wave-wide comparer */
  uint64_t lce_long_words;  /* 32-symbol windows compared by those waves */
  uint64_t algorithmic_bytes_sa;  /* HBM byte model, SA phase */
  uint64_t algorithmic_bytes_isa; /* HBM byte model, ISA phase */
  uint64_t algorithmic_bytes_lcp; /* HBM byte model, LCP phase */
  /* Dominant kernel (LSD scatter) when SFK_HIP_PROFILE_KERNELS is set. */
  double sort_scatter_seconds;
  uint64_t sort_scatter_launches;
  uint64_t sort_scatter_bytes;
  double prefix_directory_seconds; /* sfk_hip_build_fast's directory pass */
  double call_seconds;      /* whole host call: allocation, copies, build,
                               directory (host entry points only) */
  uint64_t lcp_gathered_rows; /* rows whose LCP the LCP pass took from
                                 PLCP; the others were fixed when the
                                 doubling separated them */
} sfk_hip_stats;

typedef struct sfk_hip_builder sfk_hip_builder;

/* ABI version of the loaded library (SFK_HIP_ABI_VERSION). */
int sfk_hip_abi_version(void);

/* Fault injection for tests of the allocation-failure path (the kBuildFailure
 * of a failed allocation, caps_backend.cpp:68-76): after `after` more device
 * workspace allocations succeed, the next one fails as an out-of-memory
 * hipMalloc would (SFK_HIP_BUILD_FAILURE), once; a negative value disarms.
 * Process-wide.  No counterpart in the reference. */
void sfk_hip_inject_allocation_failure(long long after);

/* Fault injection for tests of the device permutation check (the
 * kBuildFailure of RepackSuffixArray's check, suffix_array.cpp:852-881): when
 * armed (non-zero), the next build of at least 3 symbols swaps SA rows 1 and
 * 2 after the doubling, before the check runs, and the hook disarms; zero
 * disarms.  Process-wide.  No counterpart in the reference. */
void sfk_hip_inject_sa_swap(int armed);

/* Number of visible HIP devices (0 when none); never fails. */
int sfk_hip_device_count(void);

/* A builder owns one HIP stream and a grow-only device workspace on
 * `device`.  Builders are independent: two builders may build concurrently
 * from two host threads (reentrancy as in test_caps_backend.cpp:373-404). */
int sfk_hip_builder_create(int device, sfk_hip_builder** out, char* error,
                           size_t error_capacity);
void sfk_hip_builder_destroy(sfk_hip_builder* builder);
/* The builder's stream (a hipStream_t), for callers that order their own
 * work after a device build. */
void* sfk_hip_builder_stream(sfk_hip_builder* builder);
/* Bytes of device workspace a build of n symbols at construction_bits needs
 * (excluding the caller's input/output buffers). */
uint64_t sfk_hip_workspace_bytes(uint64_t n, uint32_t construction_bits);

/*
 * Device-resident build.  d_text: n device bytes (values 0..6, a single 0 at
 * position n-1).  sampling_rate K >= 1 (sampling_rate of
 * BuildDivsufsort32/64, divsufsort_backend.hpp:27-33): the outputs keep the
 * rows whose text position is a multiple of K, rows = ceil(n / K) = (n + K -
 * 1) / K elements each (n when K = 1):
 *   d_sa  of sa_bits  (32 or 64), required: the kept SA values in SA order
 *   d_isa of isa_bits (32 or 64), or NULL: d_isa[p / K] = kept row of p
 *   d_lcp of lcp_bits (32 or 64), or NULL: LCP of each kept row and the
 *         previous kept row (0 for the first), the minimum complete LCP
 *         over the rows between them
 * exactly the arrays the reference's adapter returns for K
 * (divsufsort_backend.cpp:62-150); K = 0 gives SFK_HIP_INVALID_INPUT.
 * construction_bits: 32 (n <= 2^32 - 1) or 64.  A stored width narrower than
 * the construction width is checked on the device (range error ->
 * SFK_HIP_BUILD_FAILURE).  Synchronous: returns after the stream drains.
 */
int sfk_hip_build_device(sfk_hip_builder* builder, const uint8_t* d_text,
                         uint64_t n, uint32_t construction_bits,
                         uint32_t sampling_rate, uint32_t flags, void* d_sa,
                         uint32_t sa_bits, void* d_isa, uint32_t isa_bits,
                         void* d_lcp, uint32_t lcp_bits, sfk_hip_stats* stats,
                         char* error, size_t error_capacity);

/*
 * Allocation-free variant of sfk_hip_build_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  It calls no hipMalloc/hipFree or other allocating routine:
 * its workspace is a fixed array of SFK_HIP_NOALLOC_SCRATCH_BYTES in device
 * memory that each GPU the library runs on holds from the first use of the
 * library until the process ends.  Limits: texts of at most
 * SFK_HIP_NOALLOC_MAX_SYMBOLS symbols, 32-bit construction, sampling rate
 * 1, 32-bit SA/ISA/LCP, no SFK_HIP_PROFILE_KERNELS; anything else gives
 * SFK_HIP_INVALID_INPUT.  The caller passes the device, its stream (a
 * hipStream_t, NULL for the null stream) and its input and output buffers,
 * as for sfk_hip_build_device; results are identical.  Every call on one
 * GPU shares that GPU's array, so calls that overlap on one GPU (two host
 * threads, or two streams) race and must be serialized by the caller.
 * Untimed: stats get the counts and call_seconds (host wall time of the
 * synchronous call), not the per-phase event times.
 * sfk_hip_noalloc_workspace_bytes: bytes of the array a build of n symbols
 * uses (with_isa: an ISA output is given).
 */
#define SFK_HIP_NOALLOC_MAX_SYMBOLS 1000000ull
#define SFK_HIP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int sfk_hip_build_device_noalloc(int device, const uint8_t* d_text, uint64_t n,
                                 uint32_t construction_bits,
                                 uint32_t sampling_rate, uint32_t flags,
                                 void* d_sa, uint32_t sa_bits, void* d_isa,
                                 uint32_t isa_bits, void* d_lcp,
                                 uint32_t lcp_bits, void* stream,
                                 sfk_hip_stats* stats, char* error,
                                 size_t error_capacity);
uint64_t sfk_hip_noalloc_workspace_bytes(uint64_t n, int with_isa);

/*
 * Batched device build of `refs` independent references laid out back to
 * back in one device text: reference r is the logical text (its encoded
 * bytes followed by its own sentinel 0) at [offsets[r], offsets[r+1]);
 * offsets (host array of refs + 1 entries) starts at 0 and ends at n.  The
 * outputs use the same layout: rows [offsets[r], offsets[r+1]) of d_sa /
 * d_isa / d_lcp hold reference r's SA, ISA and LCP in its own coordinates,
 * identical to building it alone.  All references are refined together: the
 * segment list of still-unsorted groups is one work queue over all of them,
 * so one set of launches serves the whole batch.  No counterpart in the
 * reference, which builds one reference per SuffixArray::Build call.
 */
int sfk_hip_build_batch_device(sfk_hip_builder* builder, const uint8_t* d_text,
                               uint64_t n, const uint64_t* offsets, uint32_t refs,
                               uint32_t construction_bits, uint32_t flags,
                               void* d_sa, uint32_t sa_bits, void* d_isa,
                               uint32_t isa_bits, void* d_lcp, uint32_t lcp_bits,
                               sfk_hip_stats* stats, char* error,
                               size_t error_capacity);

/*
 * Host build: the same, from host text into host output buffers (each
 * ceil(n / K) elements of the given width; isa/lcp may be NULL).  This is
 * the entry point a SuffixArray::Build backend branch calls in place of
 * detail::BuildDivsufsort32/64(text, sampling_rate, retain_lcp, retain_isa,
 * threads) (divsufsort_backend.hpp:27-33, .cpp:24-156).  The text goes up
 * and the results come back through pinned staging buffers; stats get the
 * copy times and the whole call's time.
 */
int sfk_hip_build(const uint8_t* text, uint64_t n, uint32_t construction_bits,
                  uint32_t sampling_rate, uint32_t flags, void* sa,
                  uint32_t sa_bits, void* isa, uint32_t isa_bits, void* lcp,
                  uint32_t lcp_bits, sfk_hip_stats* stats, char* error,
                  size_t error_capacity);

/*
 * Fast prefix directory: the exact k-mer -> SA-interval table that
 * SuffixArray::Build and Load derive for a Fast index (BuildFastPrefix,
 * suffix_array.cpp:2087-2107; FastPrefixIndex::Build/SelectK,
 * fast_prefix_index.cpp:286-382), built on the device from a resident SA.
 * Entry `key` (the k-mer's 2-bit codes A=0..T=3, first symbol most
 * significant) holds the rows [begin, end) of the suffixes that start with
 * that k-mer over A/C/G/T, or (n, n) when none does.  For n <= 2^32 - 1 an
 * entry is one word (begin | end << 32); otherwise the table is two planes
 * of 4^k words (all begins, then all ends).
 *   sfk_prefix_directory_select_k: the k SelectK picks with the default
 *     options (0: no directory) for suffix_count rows and the index's
 *     resident core bytes (text + SA + ISA + LCP);
 *   sfk_prefix_directory_words: table size in 64-bit words;
 *   sfk_hip_prefix_directory: from a device text and SA into a device table
 *     (counts[0] = indexed k-mers, counts[1] = non-empty entries);
 *   sfk_hip_build_fast: sfk_hip_build plus the directory of its SA into a
 *     host table, computed while the SA is still on the device
 *     (prefix_k = 0: none; sfk_hip_build is this call with 0; k > 0 needs
 *     a complete index, sampling_rate 1, as Build derives the directory
 *     only then).
 */
uint32_t sfk_prefix_directory_select_k(uint64_t suffix_count, uint64_t resident_core_bytes);
uint64_t sfk_prefix_directory_words(uint32_t k, uint64_t suffix_count);
int sfk_hip_prefix_directory(sfk_hip_builder* builder, const uint8_t* d_text, uint64_t n,
                             const void* d_sa, uint32_t sa_bits, uint32_t k, uint64_t* d_table,
                             uint64_t counts[2], char* error, size_t error_capacity);
int sfk_hip_build_fast(const uint8_t* text, uint64_t n, uint32_t construction_bits,
                       uint32_t sampling_rate, uint32_t flags, void* sa, uint32_t sa_bits, void* isa,
                       uint32_t isa_bits, void* lcp, uint32_t lcp_bits, uint32_t prefix_k,
                       uint64_t* prefix_table, uint64_t prefix_counts[2],
                       sfk_hip_stats* stats, char* error, size_t error_capacity);

/*
 * Resident index for exact count queries (serving).  The index borrows the
 * caller's device text (n symbols, sentinel last) and SA (sa_bits 32 or 64)
 * and packs the text into its own 2-bit planes; the caller keeps d_text and
 * d_sa alive until sfk_hip_index_destroy.  Patterns are up to 32 A/C/G/T
 * symbols as 2-bit codes (A=0 .. T=3), the first symbol in the top bits of a
 * uint64, with one length byte each.  Counting follows SuffixArray::Count
 * (suffix_array.cpp:4413-4430): the number of suffixes that start with the
 * pattern; N, separators and the sentinel never match.  `stream` is a
 * hipStream_t to enqueue on, or NULL for the index's own stream with a
 * synchronous return.  No counterpart in the reference (its queries run on
 * the host).
 */
typedef struct sfk_hip_index sfk_hip_index;
int sfk_hip_index_create(int device, const uint8_t* d_text, uint64_t n,
                         const void* d_sa, uint32_t sa_bits, sfk_hip_index** out,
                         char* error, size_t error_capacity);
void sfk_hip_index_destroy(sfk_hip_index* index);
uint64_t sfk_hip_index_device_bytes(const sfk_hip_index* index);
/* Deterministic query set: even queries sampled from the text, odd ones
 * uniformly random; lengths uniform in [min_len, max_len] (1..32). */
int sfk_hip_index_generate_queries(sfk_hip_index* index, uint64_t seed,
                                   uint64_t count, uint32_t min_len,
                                   uint32_t max_len, uint64_t* d_patterns,
                                   uint8_t* d_lengths, void* stream);
/* k of the index's prefix directory (0: none); count queries of at least
 * k symbols search only the SA interval of their first k symbols. */
uint32_t sfk_hip_index_prefix_k(const sfk_hip_index* index);
int sfk_hip_index_count(sfk_hip_index* index, const uint64_t* d_patterns,
                        const uint8_t* d_lengths, uint64_t count,
                        uint32_t* d_counts, void* stream);

/*
 * FM-index of a resident index, built on the device from its SA: the BWT
 * (the symbol before each suffix) as per-base membership masks in records
 * of 64 rows, each with the base counts of all rows before it, and C (the
 * first row of each base's suffixes).  Replaces the SDSL CSA that
 * FmIndex::Build constructs on the host (fm_index.cpp:236-294,488-524);
 * backward search gives the same half-open row ranges as the reference's
 * EqualRange (fm_index.cpp:298-309), because the SA of text ++ sentinel is
 * unique.  About n bytes of device memory.
 */
int sfk_hip_index_build_fm(sfk_hip_index* index, char* error,
                           size_t error_capacity);
uint64_t sfk_hip_index_fm_bytes(const sfk_hip_index* index);
/* out = {C[A], C[C], C[G], C[T], end of T's rows, symbols below A}. */
int sfk_hip_index_fm_header(sfk_hip_index* index, uint64_t out[6]);
/* Backward search of `count` patterns of any length: pattern q is
 * d_codes[d_offsets[q] .. d_offsets[q + 1]) (A=0 .. T=3, one byte each),
 * searched as given (d_orientation[q] = 0) or as its reverse complement
 * (1).  Its row range [begin, end) goes to d_ranges[2q], d_ranges[2q + 1],
 * (0, 0) when it does not occur (FmIndex::EqualRange / EqualRangeBatch /
 * CountBatch lanes, fm_index.cpp:298-309,526-610). */
int sfk_hip_index_fm_ranges(sfk_hip_index* index, const uint8_t* d_codes,
                            const uint64_t* d_offsets,
                            const uint8_t* d_orientation, uint64_t count,
                            uint64_t* d_ranges, void* stream);
/* Count queries in sfk_hip_index_count's packed format, by backward search
 * (the served loop's exact FM count queries). */
int sfk_hip_index_fm_count(sfk_hip_index* index, const uint64_t* d_patterns,
                           const uint8_t* d_lengths, uint64_t count,
                           uint32_t* d_counts, void* stream);
/* SA values of rows [begin, end) as uint64 into host memory (the positions
 * FmIndex::Locate recovers for a row range, fm_index.cpp:470-480). */
int sfk_hip_index_sa_rows(sfk_hip_index* index, uint64_t begin, uint64_t end,
                          uint64_t* out);

/*
 * Radix autotuner.  Every onesweep pass of a build runs one of a set of
 * compiled shapes (threads per block x keys per thread) at a digit width of
 * 8 or 7 bits, chosen per kind of sort and size class of the list from a
 * per-device table:
 *   kind 0 u32 key/u32 value, 1 u64/u32, 2 u32/u64, 3 u64/u64,
 *        4 u32/u32 and 5 u32/u64 sorted in place per reference (the first
 *        step of a batch)
 *   size_class c: lists of [2^(16+2c), 2^(18+2c)) elements (class 0 also
 *        the smaller ones, class 6 the larger ones)
 * The table starts from measured defaults.  sfk_hip_autotune times every
 * shape that fits a workgroup's LDS for every kind and class up to
 * max_elements (capped by free device memory) on random keys, installs the
 * fastest, and reports each measurement (chosen = 1 for the installed one).
 * *count receives the number of measurements even when capacity is
 * smaller.  Builds on the device read the table at each sort, so a table
 * change never affects the result, only the speed.  sfk_hip_tune_get lists
 * the shape in use for every (kind, class) (tuned = 1 when it came from the
 * tuner or sfk_hip_tune_set); sfk_hip_tune_set installs the entries with
 * chosen = 1 (a saved table) and rejects shapes that are not compiled for
 * their kind; sfk_hip_tune_reset restores the defaults.  No counterpart in
 * the reference, whose CPU engines have no device shapes to pick.
 * Digit widths: the table holds a shape and its time per pass for each
 * width (kinds 4 and 5 only 8-bit); the tuner times both, and a sort of
 * b-bit keys takes 7-bit digits when ceil(b/7) 7-bit passes of its kind and
 * class take less time than ceil(b/8) 8-bit ones (8 when either width has
 * no time, so the defaults and tables saved without times sort 8-bit).
 */
typedef struct sfk_hip_tune_entry {
  uint32_t kind;
  uint32_t size_class;
  uint32_t block;
  uint32_t items;
  double us_per_pass;
  uint32_t chosen;
  uint32_t tuned;
  uint32_t digit_bits; /* 8 or 7 (0 reads as 8 in sfk_hip_tune_set) */
} sfk_hip_tune_entry;
int sfk_hip_autotune(int device, uint64_t max_elements, sfk_hip_tune_entry* entries,
                     uint32_t capacity, uint32_t* count, char* error,
                     size_t error_capacity);
int sfk_hip_tune_get(int device, sfk_hip_tune_entry* entries, uint32_t capacity,
                     uint32_t* count);
int sfk_hip_tune_set(int device, const sfk_hip_tune_entry* entries, uint32_t count,
                     char* error, size_t error_capacity);
int sfk_hip_tune_reset(int device);

/*
 * Streaming .sufidx section payloads from device buffers, with CRC-32
 * computed on the device.  Replaces, for data resident in HBM, the
 * checksumming stream buffer the reference writes every section through
 * (serialization.cpp:60-141, WriteContainer :529-619) and the plane writes
 * of WriteCoordinateSectionV14 / WriteLcpSectionV14 (sa_codec.cpp:466-631).
 *
 * A plane kind names the stored form of each value:
 *   BYTES     the bytes as they are (value_bits 8; the text section);
 *   NATIVE32  32-bit values (native32 codec, raw32 LCP); values must fit;
 *   LOW32     the low 32 bits (first plane of split40 / split48);
 *   HIGH8     bits 32..39 (second plane of split40); values < 2^40;
 *   HIGH16    bits 32..47 (second plane of split48); values < 2^48;
 *   NATIVE64  64-bit values (native64 codec, raw64 LCP);
 *   CAPPED8   min(value, 255) in one byte (the primary plane of a
 *             byte-coded LCP, lcp_storage.cpp:233-300).
 * Device values are 32- or 64-bit (value_bits) for every kind but BYTES.
 * The plane's bytes are handed to `sink` in order, in chunks of at most the
 * writer's chunk size (pinned host memory, valid during the call); a
 * non-zero return from the sink aborts with SFK_HIP_IO_ERROR.  *crc_out is
 * zlib's crc32(crc_in, plane bytes), so a section's checksum chains through
 * its header bytes (checksummed by the caller) and its planes.  A value out
 * of its plane's range gives SFK_HIP_INVALID_INPUT.  The call waits for all
 * prior work on the device before reading the values.
 */
#define SFK_HIP_PLANE_BYTES 0u
#define SFK_HIP_PLANE_NATIVE32 1u
#define SFK_HIP_PLANE_LOW32 2u
#define SFK_HIP_PLANE_HIGH8 3u
#define SFK_HIP_PLANE_HIGH16 4u
#define SFK_HIP_PLANE_NATIVE64 5u
#define SFK_HIP_PLANE_CAPPED8 6u
typedef struct sfk_hip_plane_writer sfk_hip_plane_writer;
typedef int (*sfk_hip_sink)(void* context, const void* data, uint64_t bytes);
/* chunk_bytes 0 selects 64 MiB; two pinned host and two device buffers of
 * that size are held by the writer. */
int sfk_hip_plane_writer_create(int device, uint64_t chunk_bytes,
                                sfk_hip_plane_writer** out, char* error,
                                size_t error_capacity);
void sfk_hip_plane_writer_destroy(sfk_hip_plane_writer* writer);
int sfk_hip_write_plane(sfk_hip_plane_writer* writer, const void* d_values,
                        uint64_t count, uint32_t value_bits, uint32_t plane,
                        uint32_t crc_in, sfk_hip_sink sink, void* context,
                        uint32_t* crc_out, char* error, size_t error_capacity);
/* zlib's crc32(crc_in, bytes at d_data) computed on the device (any
 * alignment); the UpdateCrc / CrcFileRange checksums of
 * serialization.cpp:47-58,200-240 for data in HBM. */
int sfk_hip_crc32(int device, const void* d_data, uint64_t bytes,
                  uint32_t crc_in, uint32_t* crc_out, char* error,
                  size_t error_capacity);

/*
 * Byte-coded LCP of the Low-memory profile from a complete raw LCP (SA
 * order) and its ISA in device memory: BuildByteData / BuildByteDataDirect
 * of lcp_storage.cpp:233-420.  primary (host, n bytes; may be NULL) gets
 * min(LCP[r], 255).  *anchor_count gets the number of overflow anchors: text
 * positions p with PLCP[p] >= 255 that do not continue the run of p - 1
 * (PLCP[p - 1] >= 255 and PLCP[p - 1] - 1 == PLCP[p]).  When
 * anchor_positions and anchor_values are given (host, capacity entries),
 * they get the anchors' positions and PLCP values in increasing position;
 * call once with NULL arrays to learn the count.
 */
int sfk_hip_byte_coded_lcp(int device, const void* d_lcp, uint32_t lcp_bits,
                           const void* d_isa, uint32_t isa_bits, uint64_t n,
                           uint8_t* primary, uint64_t* anchor_count,
                           uint64_t* anchor_positions, uint64_t* anchor_values,
                           uint64_t capacity, char* error,
                           size_t error_capacity);

/*
 * Full-size checks of a built index in device memory.
 *
 * sfk_hip_digest: the position-keyed sum digest
 *   count * G + sum_i fmix64(values[i] + G * (i + 1))  (mod 2^64)
 * (G = 0x9E3779B97F4A7C15, fmix64 the splitmix64 finaliser) of `count`
 * device values of value_bits 8, 32 or 64; equal arrays give equal digests
 * whatever their width.  No counterpart in the reference; its benchmark
 * compares whole SAs by checksum the same way (sa_build_benchmark.cpp:
 * 455-465,1130-1140).
 *
 * sfk_hip_check_index: proves SA, ISA and (when d_lcp is given) LCP of the
 * n-symbol device text without a second construction:
 *   - every SA value < n and ISA[SA[r]] == r (SA a permutation, ISA its
 *     inverse: the checks of RepackSuffixArray, suffix_array.cpp:852-881, and
 *     of Load, :4999-5010,5120-5131);
 *   - SA[0] == n - 1 and every adjacent pair a = SA[r-1], b = SA[r] ordered:
 *     T[a] < T[b], or T[a] == T[b] and ISA[a+1] < ISA[b+1] (Burkhardt and
 *     Kaerkkaeinen's suffix array checker);
 *   - LCP[0] == 0 and, for l = LCP[r], T[a+l] != T[b+l] and the first l
 *     symbols of a and b equal under a Karp-Rabin hash modulo 2^61 - 1 whose
 *     base comes from hash_seed (a wrong value passes with probability at
 *     most n / 2^61).
 * Corrupted arrays give failure counts, never an access outside the arrays.
 * Returns SFK_HIP_OK when every check passes, SFK_HIP_BUILD_FAILURE when one
 * fails (the reference's code for a suffix array that is not a permutation,
 * suffix_array.cpp:852-881), with the counts in *report either way.  Holds
 * 8 (n + 1) bytes of device memory for the prefix hashes while it runs.
 */
typedef struct sfk_hip_check_report {
  uint64_t rows;
  uint64_t range_failures;      /* SA or ISA value >= n */
  uint64_t inverse_failures;    /* ISA[SA[r]] != r */
  uint64_t order_failures;      /* SA[0] != n - 1, or an adjacent pair out of order */
  uint64_t lcp_failures;        /* LCP[0] != 0, or T[a+l] == T[b+l] / past the end */
  uint64_t lcp_prefix_failures; /* the first LCP[r] symbols differ */
  uint64_t first_failed_row;    /* n when every row passed */
  uint32_t lcp_checked;         /* 1 when d_lcp was given */
  uint32_t reserved;
  double seconds;
} sfk_hip_check_report;
int sfk_hip_digest(int device, const void* d_values, uint64_t count,
                   uint32_t value_bits, uint64_t* out, char* error,
                   size_t error_capacity);
int sfk_hip_check_index(int device, const uint8_t* d_text, uint64_t n,
                        const void* d_sa, uint32_t sa_bits, const void* d_isa,
                        uint32_t isa_bits, const void* d_lcp, uint32_t lcp_bits,
                        uint64_t hash_seed, sfk_hip_check_report* report,
                        char* error, size_t error_capacity);

/* Device memory helpers so callers without a device allocator of their own
 * can stage buffers (tests, benchmarks). */
int sfk_hip_malloc(int device, uint64_t bytes, void** out);
int sfk_hip_free(void* ptr);
int sfk_hip_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int sfk_hip_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int sfk_hip_memcpy_d2d(void* dst, const void* src, uint64_t bytes);
/* out: LDS bytes per block, LDS bytes per CU, CUs, L2 bytes, HBM bytes,
 * core clock kHz, memory clock kHz, memory bus width bits. */
int sfk_hip_device_info(int device, uint64_t out[8]);

#ifdef __cplusplus
}
#endif

#endif /* SUFKIT_HIP_H_ */
