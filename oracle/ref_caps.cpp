// C entry point over the compiled reference's own CaPS-SA backend.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (src/caps_backend.cpp, src/error.cpp, the
// vendored CaPS-SA third_party/caps-sa/src/Suffix_Array.cpp and the
// header-only ParlayLib); nothing of the reference is copied into this
// repository.  bench.py times it as the multi-core CPU baseline next to
// divsufsort, and tests/ checks it against the oracle.
//
// sfk_ref_caps_fast runs detail::BuildCaps32/64 (caps_backend.hpp:93-99) with
// `threads` scheduler threads and retain_lcp, i.e. the SA and the merge-built
// LCP, then the ISA as the scatter isa[sa[r]] = r split over the same number
// of threads (the resident ISA that SuffixArray::Build's Fast profile adds,
// suffix_array.cpp:977-1047), and copies SA/ISA/LCP out as uint64.

#include <chrono>
#include <cstdint>
#include <exception>
#include <thread>
#include <vector>

#include "caps_backend.hpp"

namespace {
template <class Result>
void Finish(const Result& built, std::uint32_t threads, std::uint64_t* sa_out, std::uint64_t* isa_out,
            std::uint64_t* lcp_out, double* seconds) {
  const std::uint64_t n = built.suffix_array.size();
  const auto begin = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  const std::uint64_t parts = threads == 0 ? 1 : threads;
  for (std::uint64_t p = 0; p < parts; ++p) {
    pool.emplace_back([&, p] {
      const std::uint64_t lo = n * p / parts, hi = n * (p + 1) / parts;
      for (std::uint64_t r = lo; r < hi; ++r) {
        const std::uint64_t v = static_cast<std::uint64_t>(built.suffix_array[r]);
        if (isa_out != nullptr && v < n) isa_out[v] = r;
        if (sa_out != nullptr) sa_out[r] = v;
        if (lcp_out != nullptr) lcp_out[r] = static_cast<std::uint64_t>(built.lcp[r]);
      }
    });
  }
  for (auto& t : pool) t.join();
  if (seconds != nullptr) {
    seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
  }
}
}  // namespace

extern "C" {

// Returns the number of suffix rows written, or -1 on a reference error.
// seconds[0]: SA + LCP (CaPS-SA), seconds[1]: ISA scatter and copy-out.
long long sfk_ref_caps_fast(const std::uint8_t* text, std::uint64_t n, int width_bits, std::uint32_t threads,
                            std::uint64_t* sa_out, std::uint64_t* isa_out, std::uint64_t* lcp_out,
                            double* seconds) {
  try {
    std::vector<std::uint8_t> input(text, text + n);
    const auto begin = std::chrono::steady_clock::now();
    if (width_bits == 32) {
      auto built = sufkit::detail::BuildCaps32(input, threads, true, nullptr);
      if (seconds != nullptr) {
        seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
      }
      Finish(built, threads, sa_out, isa_out, lcp_out, seconds);
      return static_cast<long long>(built.suffix_array.size());
    }
    auto built = sufkit::detail::BuildCaps64(input, threads, true, nullptr);
    if (seconds != nullptr) {
      seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    }
    Finish(built, threads, sa_out, isa_out, lcp_out, seconds);
    return static_cast<long long>(built.suffix_array.size());
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
