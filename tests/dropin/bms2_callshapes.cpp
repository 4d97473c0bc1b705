// A driver making the library calls of the reference's BMS2 driver (library/src/BMS2.cpp) in the
// same shapes -- the same names, argument lists, return values used and object lifetimes -- with
// the reference's `sycl::` names spelt `sigmo::`.  It compiles against include/sigmo.hpp and links
// libbms2.so; run on a GPU it prints the candidate statistics and "# Matches: N" that BMS2 prints.
//
// usage: bms2_callshapes QUERY.dat DATA.dat ITERATIONS FIND_ALL(0|1) USE_CS(0|1)
#include <sigmo.hpp>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

namespace {
// The parts of the reference's Args (src/utils.hpp:15-59) the calls below read.
struct Options {
  Options(int argc, char** argv, sigmo::device::DeviceOptions& device_options) {
    if (argc != 6) {
      std::cerr << "usage: " << argv[0] << " QUERY.dat DATA.dat ITERATIONS FIND_ALL USE_CS\n";
      std::exit(2);
    }
    query_file = argv[1];
    data_file = argv[2];
    refinement_steps = std::atoi(argv[3]);
    find_all = std::atoi(argv[4]) != 0;
    use_cs = std::atoi(argv[5]) != 0;
    device_options.join_work_group_size = 128;  // the --join-work-group default
  }
  std::string query_file, data_file;
  int refinement_steps = 0;
  bool find_all = false, use_cs = false;
};
}  // namespace

int main(int argc, char** argv) {
  Options args{argc, argv, sigmo::device::deviceOptions};  // BMS2.cpp:8

  sigmo::DeviceBatchedCSRGraph device_data_graph;
  sigmo::DeviceBatchedCSRGraph device_query_graph;
  sigmo::queue queue{sigmo::gpu_selector_v, sigmo::property::queue::enable_profiling{}};  // :14
  size_t gpu_mem = queue.get_device().get_info<sigmo::info::device::global_mem_size>();
  std::string gpu_name = queue.get_device().get_info<sigmo::info::device::name>();

  // :21-47
  auto query_graphs = sigmo::io::loadCSRGraphsFromFile(args.query_file);
  auto data_graphs = sigmo::io::loadCSRGraphsFromFile(args.data_file);
  size_t first_query_nodes = query_graphs.empty() ? 0 : query_graphs[0].getNumNodes();
  device_query_graph = sigmo::createDeviceCSRGraph(queue, query_graphs);
  device_data_graph = sigmo::createDeviceCSRGraph(queue, data_graphs);

  // :54-60
  size_t num_query_graphs = device_query_graph.num_graphs;
  size_t num_data_graphs = device_data_graph.num_graphs;
  size_t query_nodes = device_query_graph.total_nodes;
  size_t data_nodes = device_data_graph.total_nodes;
  size_t data_graph_bytes = sigmo::getDeviceGraphAllocSize(device_data_graph);
  size_t query_graphs_bytes = sigmo::getDeviceGraphAllocSize(device_query_graph);
  std::cout << "# Query Graphs " << num_query_graphs << "\n# Data Graphs " << num_data_graphs
            << "\nJoin Work Group Size: " << sigmo::device::deviceOptions.join_work_group_size
            << "\n";

  // :86-110
  sigmo::candidates::Candidates candidates{queue, query_nodes, data_nodes};
  size_t candidates_bytes = candidates.getAllocationSize();
  sigmo::signature::Signature<> signatures{queue, device_data_graph.total_nodes,
                                           device_query_graph.total_nodes};
  size_t data_signatures_bytes = signatures.getDataSignatureAllocationSize();
  size_t query_signatures_bytes = signatures.getQuerySignatureAllocationSize();
  sigmo::signature::PathSignature path_signatures{queue, device_data_graph.total_nodes,
                                                  device_query_graph.total_nodes};
  size_t data_path_signatures_bytes = path_signatures.getDataSignatureAllocationSize();
  size_t query_path_signatures_bytes = path_signatures.getQuerySignatureAllocationSize();
  sigmo::signature::CycleSignature cycle_signatures{queue, device_data_graph.total_nodes,
                                                    device_query_graph.total_nodes};
  size_t data_cycle_signatures_bytes = cycle_signatures.getDataSignatureAllocationSize();
  size_t query_cycle_signatures_bytes = cycle_signatures.getQuerySignatureAllocationSize();
  size_t total_bytes = data_graph_bytes + query_graphs_bytes + candidates_bytes +
                       data_signatures_bytes + query_signatures_bytes +
                       data_path_signatures_bytes + query_path_signatures_bytes +
                       data_cycle_signatures_bytes + query_cycle_signatures_bytes;
  std::cout << "Total allocated memory: " << total_bytes << " out of " << gpu_mem
            << " available on " << gpu_name << "\n";

  // :131-195
  std::chrono::duration<double> time{0}, total_time{0};
  auto e1 = signatures.generateDataSignatures(device_data_graph);
  queue.wait_and_throw();
  total_time += e1.getProfilingInfo();
  auto e_path_1 = path_signatures.generateDataPathSignatures(device_data_graph);
  queue.wait_and_throw();
  total_time += e_path_1.getProfilingInfo();
  auto e_cycle_1 = cycle_signatures.generateDataCycleSignatures(device_data_graph);
  queue.wait_and_throw();
  total_time += e_cycle_1.getProfilingInfo();
  auto e2 = signatures.generateQuerySignatures(device_query_graph);
  queue.wait_and_throw();
  total_time += e2.getProfilingInfo();
  auto e_path_2 = path_signatures.generateQueryPathSignatures(device_query_graph);
  queue.wait_and_throw();
  total_time += e_path_2.getProfilingInfo();
  auto e_cycle_2 = cycle_signatures.generateQueryCycleSignatures(device_query_graph);
  queue.wait_and_throw();
  total_time += e_cycle_2.getProfilingInfo();
  auto e3 = sigmo::isomorphism::filter::filterCandidates(queue, device_query_graph,
                                                         device_data_graph, signatures,
                                                         path_signatures, cycle_signatures,
                                                         candidates);
  queue.wait_and_throw();
  total_time += e3.getProfilingInfo();
  for (size_t ref_step = 1; ref_step <= size_t(args.refinement_steps); ++ref_step) {
    auto r1 = signatures.refineDataSignatures(device_data_graph, ref_step);
    queue.wait_and_throw();
    auto r2 = signatures.refineQuerySignatures(device_query_graph, ref_step);
    queue.wait_and_throw();
    auto r3 = sigmo::isomorphism::filter::refineCandidates(queue, device_query_graph,
                                                           device_data_graph, signatures,
                                                           candidates);
    queue.wait_and_throw();
    total_time += r1.getProfilingInfo() + r2.getProfilingInfo() + r3.getProfilingInfo();
  }

  // :198-290
  std::chrono::duration<double> join_time2{0};
  size_t* total_full_matches = sigmo::malloc_shared<size_t>(1, queue);
  total_full_matches[0] = 0;
  sigmo::isomorphism::mapping::GMCR gmcr{queue};
  gmcr.generateGMCR(device_query_graph, device_data_graph, candidates);
  int partial_match_depth = args.use_cs ? static_cast<int>(first_query_nodes) : 0;
  if (partial_match_depth == 0) {
    auto join_e = sigmo::isomorphism::join::joinCandidates(
        queue, device_query_graph, device_data_graph, candidates, gmcr, total_full_matches,
        !args.find_all);
    join_e.wait();
    join_time2 = join_e.getProfilingInfo();
  } else {
    const size_t MAX_PARTIAL_MATCHES = 1000000;  // the reference reserves 10^8 records
    const size_t partial_match_size = partial_match_depth + 1;
    int* partial_matches = sigmo::malloc_shared<int>(MAX_PARTIAL_MATCHES * partial_match_size, queue);
    size_t* num_partial_matches = sigmo::malloc_shared<size_t>(1, queue);
    num_partial_matches[0] = 0;
    auto join_e2 = sigmo::isomorphism::join::joinPartialCandidates2(
        queue, device_query_graph, device_data_graph, candidates, partial_match_depth,
        partial_matches, num_partial_matches);
    join_e2.wait();
    size_t partial_matches_found = num_partial_matches[0];
    std::cout << "  > Found " << partial_matches_found << " partial matches.\n";
    const uint32_t num_q = device_query_graph.num_graphs;
    const uint32_t num_d = device_data_graph.num_graphs;
    uint32_t* pair_done = sigmo::malloc_shared<uint32_t>(num_q * num_d, queue);
    queue.memset(pair_done, 0, sizeof(uint32_t) * num_q * num_d).wait();
    bool find_first = !args.find_all;
    auto full_join_e = sigmo::isomorphism::join::joinWithPartialMatches(
        queue, device_query_graph, device_data_graph, candidates, gmcr, partial_match_depth,
        partial_matches, partial_matches_found, total_full_matches, find_first, pair_done);
    full_join_e.wait();
    join_time2 = join_e2.getProfilingInfo() + full_join_e.getProfilingInfo();
    size_t flagged = 0;
    for (size_t i = 0; i < size_t(num_q) * num_d; ++i) flagged += pair_done[i];
    if (find_first) std::cout << "# Pairs done: " << flagged << "\n";
    sigmo::free(pair_done, queue);
    sigmo::free(partial_matches, queue);
    sigmo::free(num_partial_matches, queue);
  }
  total_time += join_time2;
  std::cout << "Total time: "
            << std::chrono::duration_cast<std::chrono::milliseconds>(total_time).count() << " ms\n";

  // :374-388
  auto host_candidates = candidates.getHostCandidates();
  size_t total = 0, zero = 0;
  for (size_t i = 0; i < query_nodes; ++i) {
    auto count = host_candidates.getCandidatesCount(i);
    total += count;
    zero += count == 0;
  }
  std::cout << "# Total candidates: " << total << "\n# Zero candidates: " << zero << "\n";
  std::cout << "# Matches: " << total_full_matches[0] << "\n";

  // :390-392
  sigmo::free(total_full_matches, queue);
  sigmo::destroyDeviceCSRGraph(device_data_graph, queue);
  sigmo::destroyDeviceCSRGraph(device_query_graph, queue);
}
