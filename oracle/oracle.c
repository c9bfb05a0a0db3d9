/*
 * CPU oracle for the suffix-index construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (sufkit_amd/) links,
 * loads or calls this file.  It is the checker that tests/, the smoke entry
 * point and the CPU-baseline leg of bench.py compare the HIP backend against.
 *
 * Every function restates the reference algorithm it names, in plain C:
 *
 *   sfk_oracle_encode_records  genome_reference.cpp:30-46 (encode table),
 *                              :131-160 (separator 1 after EVERY contig),
 *                              suffix_array.cpp:4107-4117 (sentinel 0 appended)
 *   sfk_oracle_fingerprint     reference_data.hpp:46-54 (FNV-1a-64)
 *   sfk_oracle_sa_naive        the unique suffix order of text (byte compare),
 *                              as produced by divsufsort() (divsufsort.c:332)
 *   sfk_oracle_sa_doubling     same output, O(n log^2 n) rank doubling so
 *                              medium texts finish in seconds
 *   sfk_oracle_isa             divsufsort_backend.cpp:81-119 and
 *                              suffix_array.cpp:977-1022 (isa[SA[r]/K] = r)
 *   sfk_oracle_lcp_kasai       divsufsort_backend.cpp:121-150 (generalised
 *                              Kasai with carry common -= K, byte equality,
 *                              LCP[0] = 0)
 *   sfk_oracle_crc32           zlib crc32 (IEEE 802.3, reflected 0xEDB88320)
 *                              as used by serialization.cpp:60-141
 *   sfk_oracle_lean_digests    ISA scatter as above and the LCP through
 *                              Phi/PLCP (same values as the Kasai above) for
 *                              texts beyond the 64-bit-array entry points
 *   sfk_oracle_occurrences,    exact pattern occurrences and the first SA
 *   sfk_oracle_suffixes_below  row of a pattern by definition (the FM-index
 *                              queries, fm_index.cpp:298-309,547-677)
 *
 * Pinning (what makes each function trustworthy, checked in tests/):
 *   - naive SA, doubling SA: the reference's own golden vectors decoded
 *     from tests/data/legacy-sa-v1.{1,3}.fixture, each other on random and
 *     run-heavy texts, and the compiled reference (oracle/_ref) where built;
 *   - ISA: closed form (inverse permutation, checked both ways);
 *   - Kasai: brute-force adjacent compare on small texts + fixture LCPs;
 *   - lean digests: the digests of the compiled reference's full SA/ISA/LCP
 *     and of the Kasai entry point on the same texts;
 *   - CRC32: the standard check value crc32("123456789") = 0xCBF43926 and
 *     the header/section CRCs stored in the legacy fixtures;
 *   - FNV-1a-64: the fixture fingerprint 0x0e8d89af19be03e1;
 *   - occurrences / suffixes below: the reference's own FmIndex (SDSL,
 *     compiled into oracle/_ref) on the same references and patterns, and
 *     the SA interval of the pattern in the brute-force SA.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SFK_SENTINEL 0u
#define SFK_SEPARATOR 1u
#define SFK_A 2u
#define SFK_C 3u
#define SFK_G 4u
#define SFK_T 5u
#define SFK_N 6u

static uint8_t encode_base(uint8_t raw) {
  switch (raw) {
    case 'A': case 'a': return SFK_A;
    case 'C': case 'c': return SFK_C;
    case 'G': case 'g': return SFK_G;
    case 'T': case 't': return SFK_T;
    default: return SFK_N;
  }
}

/* Encode `count` raw records laid out back to back in `raw` with lengths in
 * `lengths`.  Writes sum(lengths) + count + 1 bytes (the sentinel included)
 * to `out` and returns that length.  `ambiguous` (may be NULL) receives the
 * number of bases normalised to N. */
uint64_t sfk_oracle_encode_records(const uint8_t* raw, const uint64_t* lengths,
                                   uint64_t count, uint8_t* out,
                                   uint64_t* ambiguous) {
  uint64_t in = 0, o = 0, amb = 0;
  for (uint64_t r = 0; r < count; ++r) {
    for (uint64_t k = 0; k < lengths[r]; ++k) {
      uint8_t v = encode_base(raw[in++]);
      amb += (v == SFK_N);
      out[o++] = v;
    }
    out[o++] = SFK_SEPARATOR;
  }
  out[o++] = SFK_SENTINEL;
  if (ambiguous) *ambiguous = amb;
  return o;
}

uint64_t sfk_oracle_fingerprint(const uint8_t* data, uint64_t size) {
  uint64_t hash = 14695981039346656037ULL;
  for (uint64_t i = 0; i < size; ++i) {
    hash ^= data[i];
    hash *= 1099511628211ULL;
  }
  return hash;
}

/* ---------------------------------------------------------------- SA ---- */

static const uint8_t* g_text;
static uint64_t g_n;

static int cmp_suffix(const void* a, const void* b) {
  uint64_t i = *(const uint64_t*)a, j = *(const uint64_t*)b;
  while (i < g_n && j < g_n) {
    if (g_text[i] != g_text[j]) return g_text[i] < g_text[j] ? -1 : 1;
    ++i;
    ++j;
  }
  /* unreachable for a text with a unique terminal sentinel */
  return i < j ? 1 : (i > j ? -1 : 0);
}

/* Brute-force suffix sort: qsort of suffix start positions under byte-wise
 * lexicographic comparison.  Quadratic on repetitive text; use for tiny
 * inputs only. */
void sfk_oracle_sa_naive(const uint8_t* text, uint64_t n, uint64_t* sa) {
  for (uint64_t i = 0; i < n; ++i) sa[i] = i;
  g_text = text;
  g_n = n;
  qsort(sa, n, sizeof(uint64_t), cmp_suffix);
}

typedef struct {
  uint64_t key0, key1, idx;
} triple_t;

static int cmp_triple(const void* a, const void* b) {
  const triple_t* x = (const triple_t*)a;
  const triple_t* y = (const triple_t*)b;
  if (x->key0 != y->key0) return x->key0 < y->key0 ? -1 : 1;
  if (x->key1 != y->key1) return x->key1 < y->key1 ? -1 : 1;
  return 0;
}

/* Manber-Myers rank doubling with comparison sorts: rank_h(i) orders the
 * h-prefixes; (rank_h(i), rank_h(i+h)) orders the 2h-prefixes.  Positions
 * past the end rank below everything (they never decide the order here
 * because the terminal sentinel is unique).  Returns 0 on success, -1 on
 * allocation failure. */
int sfk_oracle_sa_doubling(const uint8_t* text, uint64_t n, uint64_t* sa) {
  if (n == 0) return 0;
  triple_t* t = (triple_t*)malloc(n * sizeof(triple_t));
  uint64_t* rank = (uint64_t*)malloc(n * sizeof(uint64_t));
  if (!t || !rank) {
    free(t);
    free(rank);
    return -1;
  }
  for (uint64_t i = 0; i < n; ++i) rank[i] = (uint64_t)text[i] + 1;
  for (uint64_t h = 1;; h <<= 1) {
    for (uint64_t i = 0; i < n; ++i) {
      t[i].key0 = rank[i];
      t[i].key1 = (i + h < n) ? rank[i + h] : 0;
      t[i].idx = i;
    }
    qsort(t, n, sizeof(triple_t), cmp_triple);
    uint64_t groups = 1;
    rank[t[0].idx] = 1;
    for (uint64_t r = 1; r < n; ++r) {
      if (t[r].key0 != t[r - 1].key0 || t[r].key1 != t[r - 1].key1) ++groups;
      rank[t[r].idx] = groups;
    }
    if (groups == n || h >= n) break;
  }
  for (uint64_t r = 0; r < n; ++r) sa[r] = t[r].idx;
  free(t);
  free(rank);
  return 0;
}

/* ISA of a (possibly sampled) SA: isa[SA[r] / K] = r. */
void sfk_oracle_isa(const uint64_t* sa, uint64_t count, uint32_t sampling_rate,
                    uint64_t* isa) {
  for (uint64_t r = 0; r < count; ++r) isa[sa[r] / sampling_rate] = r;
}

/* Generalised Kasai over sampled positions K apart, byte equality, exactly
 * as divsufsort_backend.cpp:121-150: LCP[row] = lcp(SA[row-1], SA[row]),
 * LCP[0] = 0, carry common -= K (floored at 0) between samples. */
void sfk_oracle_lcp_kasai(const uint8_t* text, uint64_t n, const uint64_t* sa,
                          const uint64_t* isa, uint64_t count,
                          uint32_t sampling_rate, uint64_t* lcp) {
  uint64_t common = 0;
  memset(lcp, 0, count * sizeof(uint64_t));
  for (uint64_t sample = 0; sample < count; ++sample) {
    uint64_t suffix = sample * sampling_rate;
    uint64_t row = isa[sample];
    if (row != 0) {
      uint64_t previous = sa[row - 1];
      if (suffix + common < n && previous + common < n) {
        uint64_t a = suffix + common, b = previous + common;
        uint64_t remaining = (n - a) < (n - b) ? (n - a) : (n - b);
        uint64_t k = 0;
        while (k < remaining && text[a + k] == text[b + k]) ++k;
        common += k;
      }
      lcp[row] = common;
    }
    common = common > sampling_rate ? common - sampling_rate : 0;
  }
}

/* Brute-force LCP: compare every adjacent SA pair from scratch. */
void sfk_oracle_lcp_brute(const uint8_t* text, uint64_t n, const uint64_t* sa,
                          uint64_t count, uint64_t* lcp) {
  for (uint64_t r = 0; r < count; ++r) {
    if (r == 0) {
      lcp[0] = 0;
      continue;
    }
    uint64_t a = sa[r - 1], b = sa[r], k = 0;
    while (a + k < n && b + k < n && text[a + k] == text[b + k]) ++k;
    lcp[r] = k;
  }
}

/* --------------------------------------------------------------- CRC ---- */

static uint32_t g_crc_table[256];
static int g_crc_ready;

static void crc_init(void) {
  for (uint32_t i = 0; i < 256; ++i) {
    uint32_t c = i;
    for (int k = 0; k < 8; ++k) c = (c & 1) ? (0xEDB88320u ^ (c >> 1)) : (c >> 1);
    g_crc_table[i] = c;
  }
  g_crc_ready = 1;
}

/* zlib-compatible running CRC32: crc32(crc32(0, a), b) == crc32(0, a||b). */
uint32_t sfk_oracle_crc32(uint32_t crc, const uint8_t* data, uint64_t size) {
  if (!g_crc_ready) crc_init();
  crc = ~crc;
  for (uint64_t i = 0; i < size; ++i)
    crc = g_crc_table[(crc ^ data[i]) & 0xFFu] ^ (crc >> 8);
  return ~crc;
}

/* Order-sensitive 64-bit digest of an integer array (splitmix64 finaliser
 * over value and index), used to compare large SA/ISA/LCP arrays without
 * shipping them. */
uint64_t sfk_oracle_digest_u64(const uint64_t* values, uint64_t count) {
  uint64_t h = 0x9E3779B97F4A7C15ULL ^ count;
  for (uint64_t i = 0; i < count; ++i) {
    uint64_t z = values[i] + 0x9E3779B97F4A7C15ULL * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    h = (h ^ z) * 0x100000001B3ULL + (h >> 29);
  }
  return h;
}

/* Position-keyed sum digest of an integer array:
 *   count * G + sum_i fmix(values[i] + G * (i + 1))   (mod 2^64)
 * with G the golden-ratio constant and fmix the splitmix64 finaliser.  The
 * terms commute, so a device reduction computes the same number in any
 * order; a changed value or two swapped values change it.  Used for the
 * full-size golden digests (SA/ISA/LCP of whole workloads). */
static uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t sfk_oracle_sum_digest_u64(const uint64_t* values, uint64_t count) {
  uint64_t h = count * 0x9E3779B97F4A7C15ULL;
  for (uint64_t i = 0; i < count; ++i) h += fmix64(values[i] + 0x9E3779B97F4A7C15ULL * (i + 1));
  return h;
}

uint64_t sfk_oracle_sum_digest_u32(const uint32_t* values, uint64_t count) {
  uint64_t h = count * 0x9E3779B97F4A7C15ULL;
  for (uint64_t i = 0; i < count; ++i) h += fmix64((uint64_t)values[i] + 0x9E3779B97F4A7C15ULL * (i + 1));
  return h;
}

/* Sum digests of the ISA and LCP of a complete 32-bit SA, for texts too
 * large for the 64-bit arrays of the other entry points (n <= 2^32 - 1):
 * ISA by the scatter isa[SA[r]] = r (divsufsort_backend.cpp:81-119) and the
 * LCP through Phi/PLCP (Karkkainen, Manzini and Puglisi, CPM 2009): Phi[SA[r]]
 * = SA[r-1], PLCP[i] = lcp(i, Phi[i]) with PLCP[i+1] >= PLCP[i] - 1, byte
 * equality as LongestCommonPrefixBytes compares (sequence_compare.hpp:
 * 153-157), LCP[r] = PLCP[SA[r]], LCP[0] = 0.  The same LCP as the
 * generalised Kasai of divsufsort_backend.cpp:121-150 at K = 1.  Holds two
 * 4n-byte arrays.  out = {sa, isa, lcp} digests, out[3] = max LCP.
 * Returns 0, or -1 when SA is not a permutation or memory runs out. */
int sfk_oracle_lean_digests(const uint8_t* text, uint64_t n, const uint32_t* sa, uint64_t out[4]) {
  if (n == 0 || n > 0xFFFFFFFFull) return -1;
  uint32_t* isa = (uint32_t*)malloc(n * sizeof(uint32_t));
  uint32_t* plcp = (uint32_t*)malloc(n * sizeof(uint32_t));
  if (!isa || !plcp) { free(isa); free(plcp); return -1; }
  const uint32_t none = 0xFFFFFFFFu;
  memset(isa, 0xFF, n * sizeof(uint32_t));
  for (uint64_t r = 0; r < n; ++r) {
    if (sa[r] >= n || isa[sa[r]] != none) { free(isa); free(plcp); return -1; }
    isa[sa[r]] = (uint32_t)r;
  }
  out[0] = sfk_oracle_sum_digest_u32(sa, n);
  out[1] = sfk_oracle_sum_digest_u32(isa, n);
  free(isa);
  /* plcp first holds Phi, then is overwritten in text order. */
  plcp[sa[0]] = none;
  for (uint64_t r = 1; r < n; ++r) plcp[sa[r]] = sa[r - 1];
  uint64_t l = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t j = plcp[i];
    if (j == none) { plcp[i] = 0; l = 0; continue; }
    while (i + l < n && j + l < n && text[i + l] == text[j + l]) ++l;
    plcp[i] = (uint32_t)l;
    if (l > 0) --l;
  }
  uint64_t h = n * 0x9E3779B97F4A7C15ULL, lmax = 0;
  for (uint64_t r = 0; r < n; ++r) {
    const uint64_t v = r == 0 ? 0 : plcp[sa[r]];
    if (v > lmax) lmax = v;
    h += fmix64(v + 0x9E3779B97F4A7C15ULL * (r + 1));
  }
  out[2] = h;
  out[3] = lmax;
  free(plcp);
  return 0;
}

/* Fast prefix directory of a complete SA, restated from
 * fast_prefix_index.cpp:137-176 (PopulatePartition) and :90-135
 * (RecordRow): every position p whose k symbols are all A/C/G/T (codes
 * 2..5) widens the interval of its k-mer key (2-bit codes, first symbol
 * most significant) to cover row ISA[p]; keys never seen hold (n, n).
 * Returns the number of indexed positions; *nonempty the number of keys
 * seen. */
uint64_t sfk_oracle_prefix_directory(const uint8_t* text, uint64_t n, const uint64_t* isa, uint32_t k,
                                     uint64_t* begin, uint64_t* end, uint64_t* nonempty) {
  const uint64_t entries = (uint64_t)1 << (2 * k);
  uint64_t indexed = 0, seen = 0;
  for (uint64_t key = 0; key < entries; ++key) begin[key] = end[key] = n;
  for (uint64_t p = 0; p + k <= n; ++p) {
    uint64_t key = 0;
    int canonical = 1;
    for (uint32_t i = 0; i < k && canonical; ++i) {
      const uint8_t c = text[p + i];
      canonical = c >= 2 && c <= 5;
      key = (key << 2) | (uint64_t)(c - 2);
    }
    if (!canonical) continue;
    const uint64_t row = isa[p];
    if (begin[key] == n) {
      begin[key] = row;
      end[key] = row + 1;
      ++seen;
    } else {
      if (row < begin[key]) begin[key] = row;
      if (row + 1 > end[key]) end[key] = row + 1;
    }
    ++indexed;
  }
  if (nonempty) *nonempty = seen;
  return indexed;
}


/*
 * Exact pattern search by definition, for the FM-index queries
 * (fm_index.cpp:298-309,547-677): a pattern of bases (codes 2..5) occurs at
 * p when text[p .. p + len) equals it.  sfk_oracle_occurrences writes the
 * occurrence positions in increasing order (up to capacity) and returns
 * their number; sfk_oracle_suffixes_below counts the suffixes whose first
 * len symbols compare below the pattern (the sentinel ends every suffix and
 * is below every base), which is the first row of the pattern's SA range.
 * Parity of these with the reference's FmIndex is checked in tests/.
 */
uint64_t sfk_oracle_occurrences(const uint8_t* text, uint64_t n, const uint8_t* pattern, uint64_t len,
                                uint64_t* out, uint64_t capacity) {
  uint64_t found = 0;
  if (len == 0 || len > n) return 0;
  for (uint64_t p = 0; p + len <= n; ++p) {
    uint64_t i = 0;
    while (i < len && text[p + i] == pattern[i]) ++i;
    if (i == len) {
      if (found < capacity) out[found] = p;
      ++found;
    }
  }
  return found;
}

uint64_t sfk_oracle_suffixes_below(const uint8_t* text, uint64_t n, const uint8_t* pattern, uint64_t len) {
  uint64_t below = 0;
  for (uint64_t s = 0; s < n; ++s) {
    for (uint64_t i = 0; i < len; ++i) {
      /* s + i < n always holds before a difference: text[n - 1] is the
       * sentinel and no pattern symbol equals it. */
      const uint8_t c = text[s + i];
      if (c != pattern[i]) {
        if (c < pattern[i]) ++below;
        break;
      }
    }
  }
  return below;
}
