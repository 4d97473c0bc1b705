// Shared parts of the drivers (bms2_main: library/src/BMS2.cpp, bms2_batch: BMS2_batch.cpp)
// written against include/sigmo.hpp: the command line of src/utils.hpp:41-86, the report
// helpers of src/utils.hpp:93-178 (formatNumber, getBytesSize, CandidatesInspector) and the two
// ways to run one (query batch, data batch) pair: the reference's stage calls, or the fused
// pipeline (bms2_pipeline_*) with --fused.
#pragma once

#include <sigmo.hpp>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace driver {

using Clock = std::chrono::high_resolution_clock;
using Dur = std::chrono::duration<double>;

inline long ms(Dur d) { return long(std::chrono::duration_cast<std::chrono::milliseconds>(d).count()); }

struct Args {
  std::string query_file, data_file;
  size_t refinement_steps = 0;
  bool find_all = false, skip_join = false, print_candidates = false, use_cs = false;
  bool skip_candidates_analysis = false, fused = false;
  std::string candidates_domain = "query";
  size_t mul_data = 1, mul_query = 1;
  size_t max_data_graphs = 1000000, max_query_graphs = 1000;
  size_t join_work_group = 128, filter_work_group = 512;
  bool query_filter = false;
  size_t filter_min = 0, filter_max = SIZE_MAX;
  size_t batch_size = 100000;  // bms2_batch: data lines per batch (BMS2_batch.cpp:13)
};

[[noreturn]] inline void usage(const char* prog, const std::string& msg, bool batch) {
  if (!msg.empty()) std::cerr << prog << ": " << msg << "\n";
  std::cerr << "usage: " << prog << " -Q query.dat -D data.dat [options]\n"
            << "  -i, --iterations K          refinement steps\n"
            << "  -p, --print-candidates      candidates per query node (stderr)\n"
            << "  -c, --candidates-domain D   query | data (reported)\n"
            << "  -m, --multiply N            replicate query and data graphs N times\n"
            << "  -d, --mul-data N / -q, --mul-query N\n"
            << "  --find-all                  count every embedding (default: first match)\n"
            << "  --skip-join                 stop after the filter\n"
            << "  --query-filter MIN[:MAX]    keep queries with MIN..MAX nodes\n"
            << "  --skip-candidates-analysis  no candidate statistics\n"
            << "  --max-data-graphs N / --max-query-graphs N\n"
            << "  --join-work-group N / --filter-work-group N (reported; the kernels size "
               "their own launches)\n"
            << "  --use-cs                    partial-match depth = first query's nodes\n"
            << "  --fused                     run the fused pipeline instead of the stage calls\n";
  if (batch) std::cerr << "  --batch-size N              data lines per batch (100000)\n";
  std::exit(msg.empty() ? 0 : 2);
}

inline Args parse(int argc, char** argv, bool batch) {
  Args a;
  bool have_q = false, have_d = false;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) usage(argv[0], "missing value for " + s, batch);
      return argv[++i];
    };
    auto num = [&]() -> size_t {
      const std::string v = next();
      char* end = nullptr;
      const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
      if (!end || *end) usage(argv[0], "not a number: " + v, batch);
      return size_t(x);
    };
    if (s == "-Q") {
      a.query_file = next();
      have_q = true;
    } else if (s == "-D") {
      a.data_file = next();
      have_d = true;
    } else if (s == "-i" || s == "--iterations") a.refinement_steps = num();
    else if (s == "-p" || s == "--print-candidates") a.print_candidates = true;
    else if (s == "-c" || s == "--candidates-domain") a.candidates_domain = next();
    else if (s == "-m" || s == "--multiply") a.mul_data = a.mul_query = num();
    else if (s == "-d" || s == "--mul-data") a.mul_data = num();
    else if (s == "-q" || s == "--mul-query") a.mul_query = num();
    else if (s == "--find-all") a.find_all = true;
    else if (s == "--skip-join") a.skip_join = true;
    else if (s == "--skip-candidates-analysis") a.skip_candidates_analysis = true;
    else if (s == "--max-data-graphs") a.max_data_graphs = num();
    else if (s == "--max-query-graphs") a.max_query_graphs = num();
    else if (s == "--join-work-group") a.join_work_group = num();
    else if (s == "--filter-work-group") a.filter_work_group = num();
    else if (s == "--use-cs") a.use_cs = true;
    else if (s == "--fused") a.fused = true;
    else if (batch && s == "--batch-size") a.batch_size = std::max<size_t>(1, num());
    else if (s == "--query-filter") {
      // min[:max] (src/utils.hpp:76-85)
      const std::string v = next();
      const size_t c = v.find(':');
      a.query_filter = true;
      a.filter_min = std::stoul(v.substr(0, c));
      if (c != std::string::npos) a.filter_max = std::stoul(v.substr(c + 1));
    } else if (s == "-h" || s == "--help") usage(argv[0], "", batch);
    else usage(argv[0], "unknown flag " + s, batch);
  }
  if (have_q != have_d) usage(argv[0], "Both query and data files must be provided", batch);
  if (!have_q) usage(argv[0], "Specify input data", batch);
  if (a.candidates_domain != "query" && a.candidates_domain != "data")
    usage(argv[0], "candidates domain must be query or data", batch);
  if (a.mul_data == 0 || a.mul_query == 0) usage(argv[0], "multiply factors start at 1", batch);
  if (a.use_cs && a.fused) usage(argv[0], "--use-cs runs the stage calls; drop --fused", batch);
  return a;
}

// formatNumber (src/utils.hpp:132-137): '.' every three digits.
inline std::string format_number(uint64_t x) {
  std::string d = std::to_string(x), out;
  for (size_t i = 0; i < d.size(); ++i) {
    if (i && (d.size() - i) % 3 == 0) out += '.';
    out += d[i];
  }
  return out;
}

// getBytesSize (src/utils.hpp:139-156).
inline std::string bytes_size(uint64_t n, bool round = true) {
  double b = double(n);
  std::string unit = "B";
  if (n >= (1ull << 30)) {
    b /= double(1ull << 30);
    unit = "GB";
  } else if (n >= (1ull << 20)) {
    b /= double(1ull << 20);
    unit = "MB";
  } else if (n >= 1024) {
    b /= 1024.0;
    unit = "KB";
  }
  if (round) return std::to_string(long(b + 0.5)) + " " + unit;
  return std::to_string(b) + " " + unit;
}

// CandidatesInspector (src/utils.hpp:158-178).
struct Inspector {
  uint64_t total = 0, avg = 0, median = 0, zero = 0;
  explicit Inspector(std::vector<uint32_t> sizes) {
    for (uint32_t s : sizes) {
      total += s;
      zero += s == 0;
    }
    if (!sizes.empty()) {
      avg = total / sizes.size();
      std::sort(sizes.begin(), sizes.end());
      median = sizes[sizes.size() / 2];
    }
  }
};

// n copies of the graphs end to end (the -d / -q factors; copied by index, since inserting a
// vector's own range into it is undefined)
inline void multiply(std::vector<sigmo::CSRGraph>& d, size_t factor) {
  const size_t n = d.size();
  d.reserve(n * factor);
  for (size_t i = 1; i < factor; ++i)
    for (size_t j = 0; j < n; ++j) d.push_back(d[j]);
}

// Query graphs as BMS2.cpp:24-50 prepares them: size filter, multiply, limit.
inline void prepare_queries(std::vector<sigmo::CSRGraph>& q, const Args& a) {
  if (a.query_filter)
    q.erase(std::remove_if(q.begin(), q.end(),
                           [&](const sigmo::CSRGraph& g) {
                             return g.getNumNodes() < a.filter_min || g.getNumNodes() > a.filter_max;
                           }),
            q.end());
  multiply(q, a.mul_query);
  if (q.size() > a.max_query_graphs) q.erase(q.begin() + long(a.max_query_graphs), q.end());
}

// GPU time per stage (profiling info of each call) and host phase times.
struct Times {
  Dur sig_data{}, sig_query{}, path_data{}, path_query{}, cycle_data{}, cycle_query{}, filter{},
      join{};
  Dur host_setup{}, host_sig{}, host_filter{}, host_mapping{}, host_join{};
  Dur gpu_total() const {
    return sig_data + sig_query + path_data + path_query + cycle_data + cycle_query + filter + join;
  }
};

struct PairResult {
  uint64_t matches = 0;
  std::vector<uint32_t> candidate_counts;  // stage calls, unless the analysis is skipped
  uint64_t partial_matches = 0;
};

// The stage sequence of BMS2.cpp:99-300 on one (query, data) pair.  `verbose` prints the
// per-stage lines of BMS2.cpp:125-199.
inline PairResult run_stages(sigmo::Queue& queue, sigmo::DeviceBatchedCSRGraph& dq,
                             sigmo::DeviceBatchedCSRGraph& dd, const Args& a, size_t depth,
                             Times& t, bool verbose) {
  PairResult r;
  auto h0 = Clock::now();
  sigmo::candidates::Candidates candidates{queue, dq.total_nodes, dd.total_nodes};
  sigmo::signature::Signature<> sig{queue, dd.total_nodes, dq.total_nodes};
  sigmo::signature::PathSignature path{queue, dd.total_nodes, dq.total_nodes};
  sigmo::signature::CycleSignature cyc{queue, dd.total_nodes, dq.total_nodes};
  auto h1 = Clock::now();
  t.host_setup += h1 - h0;
  auto step = [&](sigmo::utils::BatchedEvent e, Dur& acc, const char* what) {
    const Dur d = e.getProfilingInfo();
    acc += d;
    if (verbose) std::cout << "- " << what << " in " << ms(d) << " ms\n";
  };
  if (verbose) std::cout << "------------- Runtime Filter Phase -------------\n[*] Initialization Step:\n";
  step(sig.generateDataSignatures(dd), t.sig_data, "Data (label) signatures generated");
  step(path.generateDataPathSignatures(dd), t.path_data, "Data (path) signatures generated");
  step(cyc.generateDataCycleSignatures(dd), t.cycle_data, "Data (cycle) signatures generated");
  step(sig.generateQuerySignatures(dq), t.sig_query, "Query (label) signatures generated");
  step(path.generateQueryPathSignatures(dq), t.path_query, "Query (path) signatures generated");
  step(cyc.generateQueryCycleSignatures(dq), t.cycle_query, "Query (cycle) signatures generated");
  auto h2 = Clock::now();
  t.host_sig += h2 - h1;
  step(sigmo::isomorphism::filter::filterCandidates(queue, dq, dd, sig, path, cyc, candidates),
       t.filter, "Candidates filtered (fused label + path + cycle)");
  for (size_t k = 1; k <= a.refinement_steps; ++k) {
    if (verbose) std::cout << "[*] Refinement step " << k << ":\n";
    step(sig.refineDataSignatures(dd, k), t.sig_data, "Data signatures refined");
    step(sig.refineQuerySignatures(dq, k), t.sig_query, "Query signatures refined");
    step(sigmo::isomorphism::filter::refineCandidates(queue, dq, dd, sig, candidates), t.filter,
         "Candidates refined (label)");
  }
  auto h3 = Clock::now();
  t.host_filter += h3 - h2;
  if (!a.skip_join) {
    if (verbose) std::cout << "[*] Generating DQCR\n";
    sigmo::isomorphism::mapping::GMCR gmcr{queue};
    gmcr.generateGMCR(dq, dd, candidates);
    queue.wait();
    auto h4 = Clock::now();
    t.host_mapping += h4 - h3;
    if (verbose) std::cout << "[*] Starting Join\n";
    size_t total = 0;
    if (depth == 0) {
      t.join += sigmo::isomorphism::join::joinCandidates(queue, dq, dd, candidates, gmcr, &total,
                                                         !a.find_all)
                    .getProfilingInfo();
    } else {
      size_t found = 0;
      try {  // capacity 0: the count comes back with the capacity error
        sigmo::isomorphism::join::joinPartialCandidates2(queue, dq, dd, candidates, depth, nullptr,
                                                         &found, 0);
      } catch (const std::exception&) {
      }
      int* partial = sigmo::malloc_device<int>(found * (depth + 1), queue);
      t.join += sigmo::isomorphism::join::joinPartialCandidates2(queue, dq, dd, candidates, depth,
                                                                 partial, &found, found)
                    .getProfilingInfo();
      if (verbose) std::cout << "  > Found " << format_number(found) << " partial matches.\n";
      t.join += sigmo::isomorphism::join::joinWithPartialMatches(queue, dq, dd, candidates, gmcr,
                                                                 depth, partial, found, &total,
                                                                 !a.find_all)
                    .getProfilingInfo();
      sigmo::free(partial, queue);
      r.partial_matches = found;
    }
    t.host_join += Clock::now() - h4;
    r.matches = total;
  }
  if (!a.skip_candidates_analysis) r.candidate_counts = candidates.getCandidatesCounts();
  return r;
}

// The fused pipeline (bms2_pipeline_create/run) on one pair: signatures, filter, refinement,
// GMCR and join in one call, timed as the join; no domain matrix, so no candidate statistics.
inline PairResult run_fused(sigmo::Queue& queue, sigmo::DeviceBatchedCSRGraph& dq,
                            sigmo::DeviceBatchedCSRGraph& dd, const Args& a, Times& t) {
  PairResult r;
  auto h0 = Clock::now();
  bms2_pipeline_options o{};
  o.refinement_steps = uint32_t(a.refinement_steps);
  o.find_first = a.find_all ? 0 : 1;
  bms2_pipeline p = nullptr;
  sigmo::check(bms2_pipeline_create(queue.handle(), &dq, &dd, &o, &p));
  const uint64_t D = dd.num_graphs, Q = dq.num_graphs, VW = (Q + 63) / 64;
  void *counts = nullptr, *valid = nullptr, *stats = nullptr;
  int rc = BMS2_OK;
  if ((rc = bms2_device_malloc(queue.handle(), 8 * std::max<uint64_t>(D * Q, 1), &counts)) ||
      (rc = bms2_device_malloc(queue.handle(), 8 * std::max<uint64_t>(D * VW, 1), &valid)) ||
      (rc = bms2_device_malloc(queue.handle(), 64, &stats)) ||
      (rc = bms2_memset_zero(queue.handle(), stats, 64))) {
    bms2_pipeline_destroy(p);
    sigmo::check(rc);
  }
  t.host_setup += Clock::now() - h0;
  auto h1 = Clock::now();
  bms2_event ev = nullptr;
  rc = bms2_pipeline_run(p, static_cast<uint64_t*>(counts), static_cast<uint64_t*>(valid),
                         static_cast<uint64_t*>(stats), &ev, nullptr);
  uint64_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (rc == BMS2_OK) rc = bms2_memcpy_d2h(queue.handle(), st, stats, sizeof(st));
  if (rc == BMS2_OK && ev) {
    float e = 0;
    if (bms2_event_elapsed_ms(ev, &e) == BMS2_OK) t.join += Dur(e / 1000.0);
  }
  if (ev) bms2_event_destroy(ev);
  t.host_join += Clock::now() - h1;
  bms2_device_free(queue.handle(), counts);
  bms2_device_free(queue.handle(), valid);
  bms2_device_free(queue.handle(), stats);
  bms2_pipeline_destroy(p);
  sigmo::check(rc);
  // stats: GMCR pairs, matched pairs, embeddings, DFS nodes
  r.matches = a.find_all ? st[2] : st[1];
  return r;
}

inline void print_gpu_stats(const Times& t, const Args& a, const char* total_label = "Total time") {
  std::cout << "------------- Overall GPU Stats -------------\n"
            << "Data (label) signature time: " << ms(t.sig_data) << " ms\n"
            << "Query (label) signature time: " << ms(t.sig_query) << " ms\n"
            << "Data (path) signature time: " << ms(t.path_data) << " ms\n"
            << "Query (path) signature time: " << ms(t.path_query) << " ms\n"
            << "Data (cycle) signature time: " << ms(t.cycle_data) << " ms\n"
            << "Query (cycle) signature time: " << ms(t.cycle_query) << " ms\n"
            << "Filter (label/fused) time: " << ms(t.filter) << " ms\n";
  if (a.skip_join)
    std::cout << "Join time: skipped\n";
  else
    std::cout << "Join time: " << ms(t.join) << " ms" << (a.fused ? " (fused pipeline, all stages)" : "")
              << "\n";
  std::cout << total_label << ": " << ms(t.gpu_total()) << " ms\n";
}

inline void print_host_stats(const Times& t, const Args& a, Dur wall, const char* total_label) {
  std::cout << "------------- Overall Host Stats -------------\n"
            << "Setup Data time: " << ms(t.host_setup) << " ms (not included in total)\n"
            << "Signature Gen time: " << ms(t.host_sig) << " ms\n"
            << "Filter & Refine time: " << ms(t.host_filter) << " ms\n";
  if (a.skip_join)
    std::cout << "Mapping time: skipped\nJoin time: skipped\n";
  else
    std::cout << "Mapping time: " << ms(t.host_mapping) << " ms\nJoin time: " << ms(t.host_join)
              << " ms\n";
  std::cout << total_label << ": " << ms(wall) << " ms\n";
}

// BMS2.cpp:373-388.
inline void print_results(const std::vector<uint32_t>& counts, uint64_t matches, const Args& a,
                          bool have_counts) {
  std::cout << "------------- Results -------------\n";
  if (!a.skip_candidates_analysis && have_counts) {
    Inspector in(counts);
    if (a.print_candidates)
      for (size_t i = 0; i < counts.size(); ++i) std::cerr << "Node " << i << ": " << counts[i] << "\n";
    std::cout << "# Total candidates: " << format_number(in.total) << "\n"
              << "# Average candidates: " << format_number(in.avg) << "\n"
              << "# Median candidates: " << format_number(in.median) << "\n"
              << "# Zero candidates: " << format_number(in.zero) << "\n";
  }
  if (!a.skip_join) std::cout << "# Matches: " << format_number(matches) << "\n";
}

}  // namespace driver
