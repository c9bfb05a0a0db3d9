// Deterministic synthetic genomes for the benchmark workloads.
//
// Output is the logical text SuffixArray::Build hands to a construction
// backend: encoded bases (A=2, C=3, G=4, T=5, N=6), a separator 1 after every
// contig and the sentinel 0 (genome_reference.cpp:131-160,
// suffix_array.cpp:4107-4117).  Random streams are splitmix64, seeded per
// (workload, seed, reference, contig), so a contig's content never depends on
// thread scheduling; the default seed 20260822 follows the reference's
// benchmark convention (benchmarks/sa_build_benchmark.cpp:56).
//
// Workloads (sizes at scale 1):
//   1  one contig, 4,000,000 i.i.d. uniform A/C/G/T
//   2  64 Mbp: 200 contigs, log-normal lengths (median 200 kbp), GC 45%,
//      1% N in runs of 10-5000, 30% repeat cover from 300 bp - 6 kbp
//      families copied with 1-15% substitutions, plus one 1 Mbp adversarial
//      contig (poly-A, period-2 and period-7 tandems)
//   3  3.1 Gbp: 24 contigs of 50-250 Mbp, GC 41%, 45% interspersed repeats
//      (300 bp families at 10-20% divergence, 6 kbp families at 2-5%), 3%
//      satellite tandem arrays, N gaps up to 10 Mbp
//   4  one bacterial-like reference of a 512-reference batch: log-normal
//      length clipped to 0.5-12 Mbp (median 4 Mbp), GC uniform 30-70%, 5%
//      mobile elements (1-20 kbp) drawn from a pool shared by the batch
//   5  one served reference: log-uniform 1-200 Mbp, workload-2 generator
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr std::uint8_t kA = 2, kC = 3, kG = 4, kT = 5, kN = 6, kSep = 1;

struct Rng {
  std::uint64_t state;
  explicit Rng(std::uint64_t seed) : state(seed) {}
  std::uint64_t Next() {
    std::uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double Uniform() { return (Next() >> 11) * (1.0 / 9007199254740992.0); }
  std::uint64_t Below(std::uint64_t bound) { return bound ? Next() % bound : 0; }
  std::uint64_t Range(std::uint64_t lo, std::uint64_t hi) { return lo + Below(hi - lo + 1); }
  double Normal() {
    const double u1 = std::max(Uniform(), 1e-300);
    const double u2 = Uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  double LogUniform(double lo, double hi) {
    return std::exp(std::log(lo) + Uniform() * (std::log(hi) - std::log(lo)));
  }
};

std::uint64_t Mix(std::uint64_t a, std::uint64_t b) {
  Rng r(a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL));
  r.Next();
  return r.Next();
}

// Four bases per 64-bit draw (16-bit thresholds), GC content gc.
void FillBases(Rng& rng, std::uint8_t* out, std::uint64_t len, double gc) {
  const std::uint32_t t_a = static_cast<std::uint32_t>((1.0 - gc) / 2.0 * 65536.0);
  const std::uint32_t t_c = static_cast<std::uint32_t>(0.5 * 65536.0);
  const std::uint32_t t_g = static_cast<std::uint32_t>((0.5 + gc / 2.0) * 65536.0);
  std::uint64_t i = 0;
  while (i < len) {
    std::uint64_t r = rng.Next();
    for (int k = 0; k < 4 && i < len; ++k, ++i) {
      const std::uint32_t v = static_cast<std::uint32_t>(r & 0xFFFF);
      r >>= 16;
      out[i] = v < t_a ? kA : (v < t_c ? kC : (v < t_g ? kG : kT));
    }
  }
}

void Mutate(Rng& rng, std::uint8_t* seq, std::uint64_t len, double rate) {
  if (rate <= 0) return;
  // Geometric gaps between substitutions.
  const double lg = std::log(1.0 - rate);
  std::uint64_t i = static_cast<std::uint64_t>(std::log(std::max(rng.Uniform(), 1e-300)) / lg);
  while (i < len) {
    if (seq[i] >= kA && seq[i] <= kT) {
      const std::uint8_t shift = static_cast<std::uint8_t>(1 + rng.Below(3));
      seq[i] = static_cast<std::uint8_t>(kA + (seq[i] - kA + shift) % 4);
    }
    i += 1 + static_cast<std::uint64_t>(std::log(std::max(rng.Uniform(), 1e-300)) / lg);
  }
}

struct Family {
  std::vector<std::uint8_t> seq;
  double div_lo, div_hi;
};

void InsertRepeats(Rng& rng, std::uint8_t* contig, std::uint64_t len,
                   const std::vector<Family>& families, double cover,
                   bool truncate) {
  if (families.empty() || len < 64) return;
  const std::uint64_t target = static_cast<std::uint64_t>(cover * static_cast<double>(len));
  std::uint64_t covered = 0;
  std::vector<std::uint8_t> copy;
  while (covered < target) {
    const Family& f = families[rng.Below(families.size())];
    std::uint64_t begin = 0, size = f.seq.size();
    if (truncate && size > 1000) {
      const std::uint64_t keep = rng.Range(500, size);
      begin = size - keep;  // 5'-truncated copies, as for LINE-like families
      size = keep;
    }
    if (size + 2 >= len) break;
    copy.assign(f.seq.begin() + static_cast<std::ptrdiff_t>(begin),
                f.seq.begin() + static_cast<std::ptrdiff_t>(begin + size));
    Mutate(rng, copy.data(), copy.size(), f.div_lo + rng.Uniform() * (f.div_hi - f.div_lo));
    const std::uint64_t pos = rng.Below(len - size);
    std::memcpy(contig + pos, copy.data(), size);
    covered += size;
  }
}

void InsertNRuns(Rng& rng, std::uint8_t* contig, std::uint64_t len, double cover,
                 std::uint64_t lo, std::uint64_t hi) {
  const std::uint64_t target = static_cast<std::uint64_t>(cover * static_cast<double>(len));
  std::uint64_t covered = 0;
  while (covered < target) {
    std::uint64_t run = rng.Range(lo, hi);
    run = std::min(run, std::max(lo, target - covered));  // do not overshoot the cover
    if (run >= len) run = len / 2;
    const std::uint64_t pos = rng.Below(len - run);
    std::memset(contig + pos, kN, run);
    covered += run;
  }
}

void Tandem(Rng& rng, std::uint8_t* out, std::uint64_t len, std::uint64_t period, double gc,
            double divergence) {
  std::vector<std::uint8_t> unit(period);
  FillBases(rng, unit.data(), period, gc);
  for (std::uint64_t i = 0; i < len; ++i) out[i] = unit[i % period];
  Mutate(rng, out, len, divergence);
}

std::vector<Family> MakeFamilies(Rng& rng, int count, std::uint64_t lo, std::uint64_t hi,
                                 double div_lo, double div_hi, double gc) {
  std::vector<Family> fams(static_cast<std::size_t>(count));
  for (auto& f : fams) {
    const std::uint64_t len = static_cast<std::uint64_t>(rng.LogUniform(static_cast<double>(lo), static_cast<double>(hi)));
    f.seq.resize(len);
    FillBases(rng, f.seq.data(), len, gc);
    f.div_lo = div_lo;
    f.div_hi = div_hi;
  }
  return fams;
}

struct Plan {
  std::vector<std::uint64_t> lengths;
  double gc = 0.5;
  double extra = 0;  // workload-specific parameter
};

std::vector<std::uint64_t> ScaleTo(std::vector<double> raw, std::uint64_t total, std::uint64_t min_len) {
  double sum = 0;
  for (double v : raw) sum += v;
  std::vector<std::uint64_t> out(raw.size());
  std::uint64_t acc = 0;
  for (std::size_t i = 0; i < raw.size(); ++i) {
    out[i] = std::max<std::uint64_t>(min_len, static_cast<std::uint64_t>(raw[i] / sum * static_cast<double>(total)));
    acc += out[i];
  }
  // Put the rounding remainder on the longest contig.
  auto longest = std::max_element(out.begin(), out.end());
  if (acc < total) *longest += total - acc;
  else if (acc > total && *longest > acc - total + min_len) *longest -= acc - total;
  return out;
}

Plan MakePlan(int workload, std::uint64_t seed, std::uint64_t ref, double scale) {
  Plan p;
  Rng rng(Mix(Mix(seed, static_cast<std::uint64_t>(workload)), ref));
  switch (workload) {
    case 1:
      p.lengths = {std::max<std::uint64_t>(1, static_cast<std::uint64_t>(4000000 * scale))};
      p.gc = 0.5;
      break;
    case 2:
    case 5: {
      double total = 64000000.0 * scale;
      if (workload == 5) total = rng.LogUniform(1e6, 2e8) * scale;
      const std::uint64_t adversarial = static_cast<std::uint64_t>(std::min(1e6 * scale, total / 64));
      std::vector<double> raw(200);
      for (auto& v : raw) v = 200000.0 * std::exp(1.0 * rng.Normal());
      p.lengths = ScaleTo(raw, static_cast<std::uint64_t>(total) - adversarial, 64);
      p.lengths.push_back(std::max<std::uint64_t>(adversarial, 64));
      p.gc = 0.45;
      break;
    }
    case 3: {
      std::vector<double> raw(24);
      for (auto& v : raw) v = 50e6 + rng.Uniform() * 200e6;
      p.lengths = ScaleTo(raw, static_cast<std::uint64_t>(3.1e9 * scale), 4096);
      p.gc = 0.41;
      break;
    }
    case 4: {
      double len = 4e6 * std::exp(0.5 * rng.Normal());
      len = std::min(12e6, std::max(0.5e6, len)) * scale;
      const std::uint64_t chromosome = std::max<std::uint64_t>(64, static_cast<std::uint64_t>(len));
      p.lengths = {chromosome};
      const std::uint64_t plasmids = rng.Below(3);
      for (std::uint64_t k = 0; k < plasmids; ++k) {
        p.lengths.push_back(std::max<std::uint64_t>(64, static_cast<std::uint64_t>(rng.Range(5000, 200000) * scale)));
      }
      p.gc = 0.30 + 0.40 * rng.Uniform();
      break;
    }
    default:
      break;
  }
  return p;
}

void FillContig(int workload, std::uint64_t seed, std::uint64_t ref, std::size_t index,
                const Plan& plan, std::uint8_t* out) {
  const std::uint64_t len = plan.lengths[index];
  Rng rng(Mix(Mix(Mix(seed, static_cast<std::uint64_t>(workload) + 100), ref), index));
  FillBases(rng, out, len, plan.gc);
  if (workload == 1) return;
  if (workload == 2 || workload == 5) {
    const bool adversarial = index + 1 == plan.lengths.size();
    if (adversarial) {
      const std::uint64_t a = len * 2 / 5, b = len * 3 / 10;
      std::memset(out, kA, a);
      Tandem(rng, out + a, b, 2, plan.gc, 0.0);
      Tandem(rng, out + a + b, len - a - b, 7, plan.gc, 0.0);
      return;
    }
    // Families are shared by the whole reference: seed them per reference.
    Rng frng(Mix(Mix(seed, 2000 + static_cast<std::uint64_t>(workload)), ref));
    const auto fams = MakeFamilies(frng, 60, 300, 6000, 0.01, 0.15, plan.gc);
    InsertRepeats(rng, out, len, fams, 0.30, false);
    InsertNRuns(rng, out, len, 0.01, 10, 5000);
    return;
  }
  if (workload == 3) {
    Rng frng(Mix(seed, 3003));
    const auto short_fams = MakeFamilies(frng, 30, 280, 320, 0.10, 0.20, plan.gc);
    const auto long_fams = MakeFamilies(frng, 12, 5500, 6500, 0.02, 0.05, plan.gc);
    InsertRepeats(rng, out, len, short_fams, 0.25, false);
    InsertRepeats(rng, out, len, long_fams, 0.20, true);
    // Satellite arrays: 3% of the contig.
    const std::uint64_t target = len * 3 / 100;
    std::uint64_t covered = 0;
    while (covered < target && len > 20000) {
      std::uint64_t arr = static_cast<std::uint64_t>(rng.LogUniform(1e4, 5e6));
      arr = std::min(arr, len / 4);
      const std::uint64_t period = static_cast<std::uint64_t>(rng.LogUniform(2, 200));
      const std::uint64_t pos = rng.Below(len - arr);
      Tandem(rng, out + pos, arr, period, plan.gc, 0.01 + 0.02 * rng.Uniform());
      covered += arr;
    }
    // N gaps: telomeric 10 kbp at both ends, one centromeric gap up to 10 Mbp.
    const std::uint64_t tel = std::min<std::uint64_t>(10000, len / 8);
    std::memset(out, kN, tel);
    std::memset(out + len - tel, kN, tel);
    std::uint64_t gap = static_cast<std::uint64_t>(rng.LogUniform(5e4, 1e7));
    gap = std::min(gap, len / 8);
    const std::uint64_t gpos = len / 3 + rng.Below(len / 3);
    std::memset(out + gpos, kN, std::min(gap, len - gpos));
    InsertNRuns(rng, out, len, 0.002, 10, 5000);
    return;
  }
  if (workload == 4) {
    Rng prng(Mix(seed, 4004));  // pool shared by every reference of the batch
    const auto pool = MakeFamilies(prng, 200, 1000, 20000, 0.0, 0.01, 0.5);
    InsertRepeats(rng, out, len, pool, 0.05, false);
  }
}

}  // namespace

extern "C" {

// Plans reference `ref` of `workload`: returns the logical text length
// (bases + one separator per contig + sentinel) and writes up to `capacity`
// contig lengths.  Returns 0 for an unknown workload.
std::uint64_t sfk_synth_plan(int workload, std::uint64_t seed, std::uint64_t ref, double scale,
                             std::uint64_t* contig_lengths, std::uint64_t capacity,
                             std::uint64_t* contig_count) {
  const Plan p = MakePlan(workload, seed, ref, scale);
  if (p.lengths.empty()) return 0;
  std::uint64_t n = 1;
  for (std::size_t i = 0; i < p.lengths.size(); ++i) {
    n += p.lengths[i] + 1;
    if (contig_lengths != nullptr && i < capacity) contig_lengths[i] = p.lengths[i];
  }
  if (contig_count != nullptr) *contig_count = p.lengths.size();
  return n;
}

// Writes the n-symbol logical text planned by sfk_synth_plan.  Contigs are
// generated on up to `threads` host threads.  Returns 0, or -1 when n does
// not match the plan.
int sfk_synth_generate(int workload, std::uint64_t seed, std::uint64_t ref, double scale,
                       std::uint8_t* text, std::uint64_t n, unsigned threads) {
  const Plan p = MakePlan(workload, seed, ref, scale);
  std::uint64_t expect = 1;
  std::vector<std::uint64_t> offsets;
  for (auto len : p.lengths) {
    offsets.push_back(expect - 1);
    expect += len + 1;
  }
  if (p.lengths.empty() || expect != n) return -1;
  threads = std::max(1u, threads);
  std::vector<std::thread> pool;
  std::size_t next = 0;
  std::vector<std::size_t> order(p.lengths.size());
  for (std::size_t i = 0; i < order.size(); ++i) order[i] = i;
  // Longest contigs first for balance.
  std::sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
    return p.lengths[a] > p.lengths[b];
  });
  std::vector<std::vector<std::size_t>> work(threads);
  std::vector<std::uint64_t> load(threads, 0);
  for (std::size_t i : order) {
    const auto t = static_cast<std::size_t>(std::min_element(load.begin(), load.end()) - load.begin());
    work[t].push_back(i);
    load[t] += p.lengths[i];
  }
  (void)next;
  for (unsigned t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (std::size_t i : work[t]) {
        FillContig(workload, seed, ref, i, p, text + offsets[i]);
        text[offsets[i] + p.lengths[i]] = kSep;
      }
    });
  }
  for (auto& th : pool) th.join();
  text[n - 1] = 0;
  return 0;
}

}  // extern "C"
