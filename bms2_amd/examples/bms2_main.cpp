// bms2_main: the single-GPU driver of the reference (library/src/BMS2.cpp) written against the
// sigmo:: facade in include/sigmo.hpp.  Same stage order (BMS2.cpp:131-234) and the same result
// lines; flags follow src/utils.hpp:41-59 for the subset implemented here:
//   -Q <query.dat> -D <data.dat> [-i K] [--find-all] [--max-data-graphs N]
//   [--max-query-graphs N] [--skip-join] [-p] [--use-cs]
// --use-cs joins as BMS2.cpp:217-300 does: partial matches of the first query (a cluster file's
// prototype) at depth = its node count, then their extensions to every GMCR query.
#include <sigmo.hpp>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <numeric>
#include <string>

namespace {

struct Args {
  std::string query_file, data_file;
  size_t refinement_steps = 0;
  bool find_all = false, skip_join = false, print_candidates = false, use_cs = false;
  size_t max_data_graphs = 1000000, max_query_graphs = 1000;
};

[[noreturn]] void usage(const char* msg) {
  std::cerr << "bms2_main: " << msg << "\nusage: bms2_main -Q query.dat -D data.dat [-i K] "
            << "[--find-all] [--max-data-graphs N] [--max-query-graphs N] [--skip-join] [-p] "
            << "[--use-cs]\n";
  std::exit(2);
}

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) usage(("missing value for " + s).c_str());
      return argv[++i];
    };
    if (s == "-Q" || s == "--query") a.query_file = next();
    else if (s == "-D" || s == "--data") a.data_file = next();
    else if (s == "-i" || s == "--iterations") a.refinement_steps = std::stoul(next());
    else if (s == "--find-all") a.find_all = true;
    else if (s == "--skip-join") a.skip_join = true;
    else if (s == "-p" || s == "--print-candidates") a.print_candidates = true;
    else if (s == "--use-cs") a.use_cs = true;
    else if (s == "--max-data-graphs") a.max_data_graphs = std::stoul(next());
    else if (s == "--max-query-graphs") a.max_query_graphs = std::stoul(next());
    else usage(("unknown flag " + s).c_str());
  }
  if (a.query_file.empty() || a.data_file.empty()) usage("Specify input data");
  return a;
}

long ms(std::chrono::duration<double> d) {
  return long(std::chrono::duration_cast<std::chrono::milliseconds>(d).count());
}

}  // namespace

int main(int argc, char** argv) {
  Args args = parse(argc, argv);
  try {
    sigmo::Queue queue(0);
    auto query_graphs = sigmo::io::loadCSRGraphsFromFile(args.query_file);
    auto data_graphs = sigmo::io::loadCSRGraphsFromFile(args.data_file);
    if (query_graphs.size() > args.max_query_graphs)
      query_graphs.erase(query_graphs.begin() + args.max_query_graphs, query_graphs.end());
    if (data_graphs.size() > args.max_data_graphs)
      data_graphs.erase(data_graphs.begin() + args.max_data_graphs, data_graphs.end());
    auto dq = sigmo::createDeviceCSRGraph(queue, query_graphs);
    auto dd = sigmo::createDeviceCSRGraph(queue, data_graphs);
    std::cout << "------------- Input Data -------------\n"
              << "# Query Nodes " << dq.total_nodes << "\n# Query Graphs " << dq.num_graphs
              << "\n# Data Nodes " << dd.total_nodes << "\n# Data Graphs " << dd.num_graphs << "\n";

    sigmo::candidates::Candidates candidates{queue, dq.total_nodes, dd.total_nodes};
    sigmo::signature::Signature<> signatures{queue, dd.total_nodes, dq.total_nodes};
    sigmo::signature::PathSignature path_signatures{queue, dd.total_nodes, dq.total_nodes};
    sigmo::signature::CycleSignature cycle_signatures{queue, dd.total_nodes, dq.total_nodes};

    std::chrono::duration<double> sig_time{0}, filter_time{0}, join_time{0};
    sig_time += signatures.generateDataSignatures(dd).getProfilingInfo();
    sig_time += path_signatures.generateDataPathSignatures(dd).getProfilingInfo();
    sig_time += cycle_signatures.generateDataCycleSignatures(dd).getProfilingInfo();
    sig_time += signatures.generateQuerySignatures(dq).getProfilingInfo();
    sig_time += path_signatures.generateQueryPathSignatures(dq).getProfilingInfo();
    sig_time += cycle_signatures.generateQueryCycleSignatures(dq).getProfilingInfo();
    filter_time += sigmo::isomorphism::filter::filterCandidates(
                       queue, dq, dd, signatures, path_signatures, cycle_signatures, candidates)
                       .getProfilingInfo();
    for (size_t step = 1; step <= args.refinement_steps; ++step) {
      sig_time += signatures.refineDataSignatures(dd, step).getProfilingInfo();
      sig_time += signatures.refineQuerySignatures(dq, step).getProfilingInfo();
      filter_time += sigmo::isomorphism::filter::refineCandidates(queue, dq, dd, signatures, candidates)
                         .getProfilingInfo();
    }
    size_t total_matches = 0;
    if (!args.skip_join) {
      sigmo::isomorphism::mapping::GMCR gmcr{queue};
      join_time += gmcr.generateGMCR(dq, dd, candidates).getProfilingInfo();
      // BMS2.cpp:217-227: --use-cs takes the first query's node count as partial-match depth
      const size_t depth = args.use_cs && !query_graphs.empty() ? query_graphs[0].getNumNodes() : 0;
      if (depth == 0) {
        join_time += sigmo::isomorphism::join::joinCandidates(queue, dq, dd, candidates, gmcr,
                                                              &total_matches, !args.find_all)
                         .getProfilingInfo();
      } else {
        std::cout << "[Config] --use-cs enabled. partial_match_depth set to " << depth << "\n";
        // count first, then a buffer of exactly that many records
        size_t found = 0;
        try {
          sigmo::isomorphism::join::joinPartialCandidates2(queue, dq, dd, candidates, depth, nullptr,
                                                           &found, 0);
        } catch (const std::exception&) {
          // capacity 0: the count is returned with the capacity error
        }
        int* partial = sigmo::malloc_device<int>(found * (depth + 1), queue);
        join_time += sigmo::isomorphism::join::joinPartialCandidates2(
                         queue, dq, dd, candidates, depth, partial, &found, found)
                         .getProfilingInfo();
        std::cout << "  > Found " << found << " partial matches.\n";
        join_time += sigmo::isomorphism::join::joinWithPartialMatches(
                         queue, dq, dd, candidates, gmcr, depth, partial, found, &total_matches,
                         !args.find_all)
                         .getProfilingInfo();
        sigmo::free(partial, queue);
      }
    }
    std::cout << "------------- Overall GPU Stats -------------\n"
              << "Signature time: " << ms(sig_time) << " ms\n"
              << "Filter (label/fused) time: " << ms(filter_time) << " ms\n"
              << "Join time: " << (args.skip_join ? std::string("skipped") : std::to_string(ms(join_time)) + " ms")
              << "\n------------- Results -------------\n";
    auto counts = candidates.getCandidatesCounts();
    uint64_t total = 0, zero = 0;
    for (size_t i = 0; i < counts.size(); ++i) {
      total += counts[i];
      zero += counts[i] == 0;
      if (args.print_candidates) std::cerr << "Node " << i << ": " << counts[i] << "\n";
    }
    std::cout << "# Total candidates: " << total << "\n# Zero candidates: " << zero << "\n";
    if (!args.skip_join) std::cout << "# Matches: " << total_matches << "\n";
    sigmo::destroyDeviceCSRGraph(dd, queue);
    sigmo::destroyDeviceCSRGraph(dq, queue);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
