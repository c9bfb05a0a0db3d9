// Build-request serving loop with continuous batching on one GPU.
//
// Requests (a logical text and an arrival time) are replayed on a wall-clock
// timeline.  `slots` worker threads each own one HIP builder (its own stream
// and workspace); a worker admits the oldest waiting request as soon as the
// request's HBM footprint fits the budget, so requests that arrive while
// others are on the GPU start as slots and memory free up, and their kernels
// overlap on the device.  A finished index (text, SA, ISA, LCP) stays
// resident; the worker then answers the request's exact count queries on the
// device against it (by backward search over an FM-index built on the
// device from the resident SA when fm_queries is set, else by binary search
// over the SA).  Per request the loop records arrival, admission, build
// completion and query completion; the summary gives p50/p99 latency, Mbp/s
// and queries/s.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sufkit_hip.h"

extern "C" {

struct sfk_serve_request {
  double arrival_seconds;
  const uint8_t* text;  // host logical text (sentinel last)
  uint64_t n;
};

struct sfk_serve_config {
  uint64_t hbm_budget_bytes;
  uint32_t slots;
  uint64_t queries_per_index;
  uint32_t min_query_len;
  uint32_t max_query_len;
  uint64_t query_seed;
  uint32_t keep_resident;   // keep every finished index on the device until the end
  uint32_t sample_queries;  // queries per request copied back for checking
  uint32_t digest_outputs;  // copy SA/ISA/LCP back and hash them (checking)
  uint32_t fm_queries;      // answer by FM backward search (FM-index built on the resident SA)
  uint32_t strict_budget;   // refuse a request whose footprint alone exceeds the budget
                            // (status SFK_HIP_INVALID_INPUT) instead of running it alone
};

struct sfk_serve_result {
  double arrival, admitted, built, answered;  // seconds since the start of the run
  double device_build_seconds;
  uint64_t n;
  uint64_t footprint_bytes;
  uint64_t count_sum;
  uint64_t count_hash;
  uint64_t output_hash[3];  // SA, ISA, LCP (digest_outputs)
  int status;
  int slot;
};

struct sfk_serve_summary {
  double makespan_seconds;
  double latency_p50, latency_p99;       // arrival -> index built
  double end_to_end_p50, end_to_end_p99; // arrival -> queries answered
  double build_mbps;                     // bases built / makespan
  double queries_per_second;             // all queries / summed query time per slot
  uint64_t peak_reserved_bytes;
  uint64_t requests_failed;
  uint64_t max_concurrent;
  uint64_t evicted;
  char first_error[256];  // message of the first request that failed ("" if none)
};

}  // extern "C"

namespace {

using Clock = std::chrono::steady_clock;

struct Resident {
  void* text = nullptr;
  void* sa = nullptr;
  void* isa = nullptr;
  void* lcp = nullptr;
  sfk_hip_index* index = nullptr;
  uint64_t bytes = 0;
};

void FreeResident(Resident& r) {
  if (r.index != nullptr) sfk_hip_index_destroy(r.index);
  for (void* p : {r.text, r.sa, r.isa, r.lcp}) {
    if (p != nullptr) sfk_hip_free(p);
  }
  r = Resident{};
}

double Percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p * static_cast<double>(v.size() - 1);
  const auto lo = static_cast<std::size_t>(pos);
  const auto hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
}

}  // namespace

extern "C" int sfk_serve_run(const sfk_serve_request* requests, uint64_t count,
                             const sfk_serve_config* config, sfk_serve_result* results,
                             uint64_t* sample_patterns, uint8_t* sample_lengths,
                             uint32_t* sample_counts, sfk_serve_summary* summary, char* error,
                             size_t error_capacity) {
  if (requests == nullptr || config == nullptr || results == nullptr || config->slots == 0) {
    if (error != nullptr && error_capacity > 0) {
      std::strncpy(error, "requests, config (slots > 0) and results are required", error_capacity - 1);
      error[error_capacity - 1] = '\0';
    }
    return SFK_HIP_INVALID_INPUT;
  }
  const auto t0 = Clock::now();
  const auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  std::mutex mu;
  std::string first_error;  // under mu
  std::condition_variable cv;
  std::deque<uint64_t> waiting;
  std::deque<uint64_t> resident_order;  // finished indexes, oldest first
  uint64_t arrived = 0, finished = 0;
  uint64_t reserved = 0, peak = 0, concurrent = 0, max_concurrent = 0, evicted = 0;
  std::vector<Resident> resident(count);
  std::vector<double> query_seconds(config->slots, 0.0);
  std::vector<uint64_t> workspace_reserved(config->slots, 0);

  for (uint64_t i = 0; i < count; ++i) {
    results[i] = sfk_serve_result{};
    results[i].n = requests[i].n;
    results[i].arrival = requests[i].arrival_seconds;
    results[i].status = -1;
  }

  std::thread arrivals([&] {
    for (uint64_t i = 0; i < count; ++i) {
      const double wait = requests[i].arrival_seconds - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      std::lock_guard<std::mutex> lock(mu);
      waiting.push_back(i);
      ++arrived;
      cv.notify_all();
    }
  });

  auto worker = [&](uint32_t slot) {
    sfk_hip_builder* builder = nullptr;
    char err[256] = {0};
    if (sfk_hip_builder_create(0, &builder, err, sizeof(err)) != SFK_HIP_OK) builder = nullptr;
    while (true) {
      uint64_t i = 0;
      uint64_t need = 0, ws_growth = 0, own = 0;
      {
        std::unique_lock<std::mutex> lock(mu);
        bool admitted = false;
        while (!admitted) {
          if (waiting.empty()) {
            if (arrived == count) break;
            cv.wait(lock);
            continue;
          }
          const auto& r = requests[waiting.front()];
          const uint64_t ws = sfk_hip_workspace_bytes(r.n, r.n - 1 <= 0xFFFFFFFFull ? 32 : 64);
          const uint64_t growth = ws > workspace_reserved[slot] ? ws - workspace_reserved[slot] : 0;
          const uint64_t w = r.n - 1 <= 0xFFFFFFFFull ? 4 : 8;
          const uint64_t want = (r.n + 64) + 3 * r.n * w + r.n / 2 + 64 + (config->fm_queries ? r.n + 64 : 0) +
                                config->queries_per_index * 13;
          if (reserved + growth + want <= config->hbm_budget_bytes) {
            admitted = true;
          } else if (!resident_order.empty()) {
            // Make room: the oldest resident index leaves the device.
            const uint64_t victim = resident_order.front();
            resident_order.pop_front();
            reserved -= resident[victim].bytes;
            FreeResident(resident[victim]);
            ++evicted;
          } else if (concurrent == 0) {
            admitted = true;  // nothing in flight: an oversized request runs alone, or is refused
          } else {
            cv.wait(lock);
          }
        }
        if (!admitted) break;  // every request has been taken
        i = waiting.front();
        waiting.pop_front();
        const auto& r = requests[i];
        const uint64_t ws = sfk_hip_workspace_bytes(r.n, r.n - 1 <= 0xFFFFFFFFull ? 32 : 64);
        ws_growth = ws > workspace_reserved[slot] ? ws - workspace_reserved[slot] : 0;
        workspace_reserved[slot] += ws_growth;
        const uint64_t w = r.n - 1 <= 0xFFFFFFFFull ? 4 : 8;
        need = (r.n + 64) + 3 * r.n * w + r.n / 2 + 64 + (config->fm_queries ? r.n + 64 : 0) +
               config->queries_per_index * 13;
        own = ws + need;  // this request's footprint alone
        reserved += ws_growth + need;
        peak = std::max(peak, reserved);
        max_concurrent = std::max(max_concurrent, ++concurrent);
        results[i].admitted = now();
        results[i].slot = static_cast<int>(slot);
        results[i].footprint_bytes = need;
      }
      const auto& req = requests[i];
      auto& res = results[i];
      Resident& rs = resident[i];
      // This request's message: the failing call's own text, or the name of
      // the call when it leaves none.
      err[0] = '\0';
      int status = SFK_HIP_OK;
      auto step = [&](int st, const std::string& what) {
        if (st != SFK_HIP_OK && status == SFK_HIP_OK) {
          status = st;
          if (err[0] == '\0') std::snprintf(err, sizeof(err), "%s failed", what.c_str());
        }
        return status == SFK_HIP_OK;
      };
      if (builder == nullptr) step(SFK_HIP_BUILD_FAILURE, "builder creation for slot " + std::to_string(slot));
      const uint64_t n = req.n;
      const uint32_t bits = n - 1 <= 0xFFFFFFFFull ? 32 : 64;
      const uint64_t w = bits / 8;
      void* patterns = nullptr;
      void* lengths = nullptr;
      void* counts = nullptr;
      const bool refused = status == SFK_HIP_OK && config->strict_budget && own > config->hbm_budget_bytes;
      if (refused) {
        std::snprintf(err, sizeof(err), "footprint of %llu bytes exceeds the HBM budget of %llu bytes",
                      static_cast<unsigned long long>(own),
                      static_cast<unsigned long long>(config->hbm_budget_bytes));
        step(SFK_HIP_INVALID_INPUT, "admission");
      }
      if (status == SFK_HIP_OK) {
        const std::string bytes = std::to_string(n * w);
        step(sfk_hip_malloc(0, n + 64, &rs.text), "device allocation of " + std::to_string(n + 64) + " text bytes");
        if (status == SFK_HIP_OK) step(sfk_hip_malloc(0, n * w, &rs.sa), "device allocation of " + bytes + " SA bytes");
        if (status == SFK_HIP_OK) step(sfk_hip_malloc(0, n * w, &rs.isa), "device allocation of " + bytes + " ISA bytes");
        if (status == SFK_HIP_OK) step(sfk_hip_malloc(0, n * w, &rs.lcp), "device allocation of " + bytes + " LCP bytes");
        if (status == SFK_HIP_OK) step(sfk_hip_memcpy_h2d(rs.text, req.text, n), "text upload");
      }
      if (status == SFK_HIP_OK) {
        sfk_hip_stats stats{};
        step(sfk_hip_build_device(builder, static_cast<const uint8_t*>(rs.text), n, bits, 1,
                                  SFK_HIP_RETAIN_ISA | SFK_HIP_RETAIN_LCP, rs.sa, bits, rs.isa, bits, rs.lcp, bits,
                                  &stats, err, sizeof(err)),
             "device build");
        res.device_build_seconds = stats.device_seconds;
      }
      res.built = now();
      if (status == SFK_HIP_OK && config->digest_outputs) {
        void* planes[3] = {rs.sa, rs.isa, rs.lcp};
        std::vector<uint8_t> host(n * w);
        for (int k = 0; k < 3 && status == SFK_HIP_OK; ++k) {
          step(sfk_hip_memcpy_d2h(host.data(), planes[k], n * w), "output download");
          uint64_t hash = 1469598103934665603ULL;
          for (uint8_t byte : host) hash = (hash ^ byte) * 1099511628211ULL;
          res.output_hash[k] = hash;
        }
      }
      if (status == SFK_HIP_OK) {
        step(sfk_hip_index_create(0, static_cast<const uint8_t*>(rs.text), n, rs.sa, bits, &rs.index, err,
                                  sizeof(err)),
             "resident index creation");
      }
      if (status == SFK_HIP_OK && config->fm_queries) {
        step(sfk_hip_index_build_fm(rs.index, err, sizeof(err)), "FM-index build");
      }
      const uint64_t q = config->queries_per_index;
      if (status == SFK_HIP_OK && q > 0) {
        const double qs = now();
        step(sfk_hip_malloc(0, q * 8, &patterns), "device allocation of " + std::to_string(q * 8) + " query bytes");
        if (status == SFK_HIP_OK) step(sfk_hip_malloc(0, q, &lengths), "device allocation of query lengths");
        if (status == SFK_HIP_OK) step(sfk_hip_malloc(0, q * 4, &counts), "device allocation of query counts");
        if (status == SFK_HIP_OK) {
          step(sfk_hip_index_generate_queries(rs.index, config->query_seed ^ (i * 0x9E3779B97F4A7C15ULL), q,
                                              config->min_query_len, config->max_query_len,
                                              static_cast<uint64_t*>(patterns), static_cast<uint8_t*>(lengths),
                                              nullptr),
               "query generation");
        }
        if (status == SFK_HIP_OK) {
          step(config->fm_queries
                   ? sfk_hip_index_fm_count(rs.index, static_cast<uint64_t*>(patterns),
                                            static_cast<uint8_t*>(lengths), q, static_cast<uint32_t*>(counts),
                                            nullptr)
                   : sfk_hip_index_count(rs.index, static_cast<uint64_t*>(patterns), static_cast<uint8_t*>(lengths),
                                         q, static_cast<uint32_t*>(counts), nullptr),
               "count queries");
        }
        if (status == SFK_HIP_OK) {
          std::vector<uint32_t> host(q);
          step(sfk_hip_memcpy_d2h(host.data(), counts, q * 4), "count download");
          uint64_t sum = 0, hash = 1469598103934665603ULL;
          for (uint64_t k = 0; k < q; ++k) {
            sum += host[k];
            hash = (hash ^ host[k]) * 1099511628211ULL;
          }
          res.count_sum = sum;
          res.count_hash = hash;
          const uint64_t s = std::min<uint64_t>(config->sample_queries, q);
          if (status == SFK_HIP_OK && s > 0 && sample_patterns != nullptr) {
            step(sfk_hip_memcpy_d2h(sample_patterns + i * config->sample_queries, patterns, s * 8),
                 "sample query download");
            if (status == SFK_HIP_OK) {
              step(sfk_hip_memcpy_d2h(sample_lengths + i * config->sample_queries, lengths, s),
                   "sample length download");
            }
            std::memcpy(sample_counts + i * config->sample_queries, host.data(), s * 4);
          }
        }
        query_seconds[slot] += now() - qs;
      }
      for (void* p : {patterns, lengths, counts}) {
        if (p != nullptr) sfk_hip_free(p);
      }
      res.answered = now();
      res.status = status;
      {
        std::lock_guard<std::mutex> lock(mu);
        if (status != SFK_HIP_OK && first_error.empty()) {
          first_error = "request " + std::to_string(i) + " (status " + std::to_string(status) + "): " + err;
        }
        --concurrent;
        ++finished;
        if (refused) {  // its workspace growth never happened
          workspace_reserved[slot] -= ws_growth;
          reserved -= ws_growth;
        }
        // Query buffers are gone; a resident index keeps the rest.
        reserved -= config->queries_per_index * 13;
        if (!config->keep_resident || status != SFK_HIP_OK) {
          FreeResident(rs);
          reserved -= need - config->queries_per_index * 13;
        } else {
          rs.bytes = need - config->queries_per_index * 13;
          resident_order.push_back(i);
        }
        cv.notify_all();
      }
    }
    if (builder != nullptr) sfk_hip_builder_destroy(builder);
  };

  std::vector<std::thread> workers;
  for (uint32_t s = 0; s < config->slots; ++s) workers.emplace_back(worker, s);
  arrivals.join();
  for (auto& t : workers) t.join();

  std::vector<double> latency, e2e;
  double first_arrival = 1e300, last_done = 0, bases = 0;
  uint64_t failed = 0;
  for (uint64_t i = 0; i < count; ++i) {
    if (results[i].status != SFK_HIP_OK) {
      ++failed;
      continue;
    }
    latency.push_back(results[i].built - results[i].arrival);
    e2e.push_back(results[i].answered - results[i].arrival);
    first_arrival = std::min(first_arrival, results[i].arrival);
    last_done = std::max(last_done, results[i].built);
    bases += static_cast<double>(results[i].n);
  }
  for (auto& r : resident) FreeResident(r);
  if (summary != nullptr) {
    *summary = sfk_serve_summary{};
    summary->makespan_seconds = last_done > first_arrival ? last_done - first_arrival : 0.0;
    summary->latency_p50 = Percentile(latency, 0.50);
    summary->latency_p99 = Percentile(latency, 0.99);
    summary->end_to_end_p50 = Percentile(e2e, 0.50);
    summary->end_to_end_p99 = Percentile(e2e, 0.99);
    summary->build_mbps = summary->makespan_seconds > 0 ? bases / 1e6 / summary->makespan_seconds : 0;
    double qt = 0;
    for (double v : query_seconds) qt += v;
    summary->queries_per_second =
        qt > 0 ? static_cast<double>(config->queries_per_index * (count - failed)) / qt : 0;
    summary->peak_reserved_bytes = peak;
    summary->requests_failed = failed;
    std::memset(summary->first_error, 0, sizeof(summary->first_error));
    std::strncpy(summary->first_error, first_error.c_str(), sizeof(summary->first_error) - 1);
    summary->max_concurrent = max_concurrent;
    summary->evicted = evicted;
  }
  return failed == 0 ? SFK_HIP_OK : SFK_HIP_BUILD_FAILURE;
}
