// bms2_main: the single-GPU driver of the reference (library/src/BMS2.cpp) written against the
// sigmo:: facade in include/sigmo.hpp.  Same command line (src/utils.hpp:41-86), same stage order
// (BMS2.cpp:99-300) and the same report sections and result lines (BMS2.cpp:69-388); --fused runs
// the fused pipeline instead of the stage calls.
#include "driver_common.hpp"

int main(int argc, char** argv) {
  using namespace driver;
  const Args args = parse(argc, argv, false);
  try {
    const auto wall0 = Clock::now();
    sigmo::Queue queue(0);
    auto query_graphs = sigmo::io::loadCSRGraphsFromFile(args.query_file);
    auto data_graphs = sigmo::io::loadCSRGraphsFromFile(args.data_file);
    if (args.query_filter)
      query_graphs.erase(std::remove_if(query_graphs.begin(), query_graphs.end(),
                                        [&](const sigmo::CSRGraph& g) {
                                          return g.getNumNodes() < args.filter_min ||
                                                 g.getNumNodes() > args.filter_max;
                                        }),
                         query_graphs.end());
    const size_t first_query_nodes = query_graphs.empty() ? 0 : query_graphs[0].getNumNodes();
    multiply(query_graphs, args.mul_query);
    multiply(data_graphs, args.mul_data);
    if (query_graphs.size() > args.max_query_graphs)
      query_graphs.erase(query_graphs.begin() + long(args.max_query_graphs), query_graphs.end());
    if (data_graphs.size() > args.max_data_graphs)
      data_graphs.erase(data_graphs.begin() + long(args.max_data_graphs), data_graphs.end());
    auto dq = sigmo::createDeviceCSRGraph(queue, query_graphs);
    auto dd = sigmo::createDeviceCSRGraph(queue, data_graphs);

    std::cout << "------------- Input Data -------------\n"
              << "Read data graph and query graph\n"
              << "# Query Nodes " << dq.total_nodes << "\n# Query Graphs " << dq.num_graphs
              << "\n# Data Nodes " << dd.total_nodes << "\n# Data Graphs " << dd.num_graphs << "\n"
              << "------------- Configs -------------\n"
              << "Filter domain: " << args.candidates_domain << "\n"
              << "Filter Work Group Size: " << args.filter_work_group << "\n"
              << "Join Work Group Size: " << args.join_work_group << "\n"
              << "Find all: " << (args.find_all ? "Yes" : "No") << "\n";
    if (args.use_cs) std::cout << "Use CS (Partial Join based on Query Size): Yes\n";
    if (args.fused) std::cout << "Fused pipeline: Yes\n";
    const uint64_t qb = sigmo::getDeviceGraphAllocSize(dq), db = sigmo::getDeviceGraphAllocSize(dd);
    std::cout << "------------- Setup Data -------------\n"
              << "Allocated " << bytes_size(db) << " for graph data\n"
              << "Allocated " << bytes_size(qb) << " for query data\n";

    size_t depth = 0;
    if (args.use_cs && !args.fused) {
      if (first_query_nodes > 0) {
        depth = first_query_nodes;
        std::cout << "[Config] --use-cs enabled. partial_match_depth set to " << depth << "\n";
      } else {
        std::cerr << "[Warning] --use-cs enabled but no query graphs found. Keeping "
                     "partial_match_depth at 0.\n";
      }
    }
    Times t;
    PairResult r = args.fused ? run_fused(queue, dq, dd, args, t)
                              : run_stages(queue, dq, dd, args, depth, t, true);
    std::cout << "[!] End\n";
    print_gpu_stats(t, args);
    print_host_stats(t, args, Clock::now() - wall0, "Total time");
    print_results(r.candidate_counts, r.matches, args, !args.fused);
    sigmo::destroyDeviceCSRGraph(dd, queue);
    sigmo::destroyDeviceCSRGraph(dq, queue);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
