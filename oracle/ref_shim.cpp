// C entry points over the compiled reference's own divsufsort backend.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled by oracle/Makefile against
// the reference sources where they lie (src/divsufsort_backend.cpp,
// src/coordinate_storage.cpp, src/error.cpp and the vendored libdivsufsort
// 2.0.2 C files); nothing of the reference is copied into this repository.
// The resulting oracle/_ref/libsufkit_ref.so is used by tests/ to pin the
// oracle and by bench.py as the CPU baseline ("kind": "reference").
//
// sfk_ref_build_fast runs detail::BuildDivsufsort32/64
// (divsufsort_backend.hpp:27-33) with sampling K, retain_lcp, retain_isa
// exactly as SuffixArray::Build's Fast profile asks for them
// (suffix_array.cpp:4207-4236), then copies the SA/ISA/LCP out as uint64.

#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include <sys/mman.h>

#include "divsufsort_backend.hpp"

extern "C" {
#include "divsufsort.h"
#include "divsufsort64.h"
}

namespace {
template <class Result>
int CopyOut(const Result& built, std::uint64_t* sa_out, std::uint64_t* isa_out,
            std::uint64_t* lcp_out, double* seconds) {
  const auto count = built.suffix_array.size();
  if (sa_out != nullptr) {
    for (std::size_t r = 0; r < count; ++r) {
      sa_out[r] = static_cast<std::uint64_t>(built.suffix_array[r]);
    }
  }
  if (isa_out != nullptr && built.isa.size() == count) {
    std::memcpy(isa_out, built.isa.data(), count * sizeof(std::uint64_t));
  }
  if (lcp_out != nullptr && built.lcp.size() == count) {
    std::memcpy(lcp_out, built.lcp.data(), count * sizeof(std::uint64_t));
  }
  if (seconds != nullptr) {
    seconds[0] = built.sa_seconds;
    seconds[1] = built.isa_seconds;
    seconds[2] = built.lcp_seconds;
  }
  return 0;
}
}  // namespace

extern "C" {

// Returns the number of suffix rows written, or -1 on a reference error.
long long sfk_ref_build_fast(const std::uint8_t* text, std::uint64_t n,
                             int width_bits, std::uint32_t sampling_rate,
                             std::uint32_t threads, std::uint64_t* sa_out,
                             std::uint64_t* isa_out, std::uint64_t* lcp_out,
                             double* seconds) {
  try {
    std::vector<std::uint8_t> input(text, text + n);
    if (width_bits == 32) {
      auto built = sufkit::detail::BuildDivsufsort32(input, sampling_rate,
                                                     true, true, threads);
      CopyOut(built, sa_out, isa_out, lcp_out, seconds);
      return static_cast<long long>(built.suffix_array.size());
    }
    auto built = sufkit::detail::BuildDivsufsort64(input, sampling_rate, true,
                                                   true, threads);
    CopyOut(built, sa_out, isa_out, lcp_out, seconds);
    return static_cast<long long>(built.suffix_array.size());
  } catch (const std::exception&) {
    return -1;
  }
}

// Plain libdivsufsort SA (no ISA/LCP), for timing the SA share alone.
int sfk_ref_divsufsort(const std::uint8_t* text, std::uint64_t n,
                       int width_bits, std::uint64_t* sa_out) {
  if (width_bits == 32) {
    std::vector<saidx_t> sa(n);
    if (divsufsort(text, sa.data(), static_cast<saidx_t>(n)) != 0) return -1;
    for (std::uint64_t r = 0; r < n; ++r) sa_out[r] = static_cast<std::uint64_t>(sa[r]);
    return 0;
  }
  std::vector<saidx64_t> sa(n);
  if (divsufsort64(text, sa.data(), static_cast<saidx64_t>(n)) != 0) return -1;
  for (std::uint64_t r = 0; r < n; ++r) sa_out[r] = static_cast<std::uint64_t>(sa[r]);
  return 0;
}

// The reference's divsufsort64 (the SA its 64-bit construction computes,
// divsufsort_backend.cpp:46-60) for texts of up to 2^32 - 1 symbols whose
// full 64-bit ISA/LCP would not fit this host's memory: the 8n-byte SA is
// built in an anonymous mapping, narrowed into the caller's 32-bit array and
// unmapped.  Returns 0, or -1 on a divsufsort error or failed mapping.
int sfk_ref_divsufsort64_narrow(const std::uint8_t* text, std::uint64_t n,
                                std::uint32_t* sa_out, double* seconds) {
  if (n == 0 || n > 0xFFFFFFFFull) return -1;
  const std::size_t bytes = n * sizeof(saidx64_t);
  void* mem = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
  if (mem == MAP_FAILED) return -1;
  auto* sa = static_cast<saidx64_t*>(mem);
  const auto t0 = std::chrono::steady_clock::now();
  const int status = divsufsort64(text, sa, static_cast<saidx64_t>(n));
  if (seconds != nullptr) {
    seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (status == 0) {
    for (std::uint64_t r = 0; r < n; ++r) sa_out[r] = static_cast<std::uint32_t>(sa[r]);
  }
  munmap(mem, bytes);
  return status == 0 ? 0 : -1;
}

}  // extern "C"
