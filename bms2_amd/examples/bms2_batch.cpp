// bms2_batch: the batched driver of the reference (library/src/BMS2_batch.cpp) written against
// include/sigmo.hpp.  The data file is streamed in batches of --batch-size non-empty lines
// (BMS2_batch.cpp:104-146); each batch is parsed (malformed lines dropped), multiplied, packed on
// the GPU (createDeviceCSRGraph) and run against the resident query batch, with the per-batch and
// final report lines of BMS2_batch.cpp:96-390.  A reader thread parses the next batch while the
// GPU runs the current one (the reference reads and parses between batches); in first-match mode
// the run stops after the first batch with a match, as the reference does (:307-310).
#include "driver_common.hpp"

#include <atomic>
#include <condition_variable>
#include <deque>
#include <exception>
#include <fstream>
#include <mutex>
#include <thread>

namespace {

struct Batch {
  std::vector<sigmo::CSRGraph> graphs;
  size_t raw_lines = 0;
  bool last = false;  // end of input (or of max-data-graphs); carries no graphs
};

// Bounded hand-off between the reader thread and the GPU loop.
class BatchQueue {
 public:
  void push(Batch b) {
    std::unique_lock<std::mutex> l(m_);
    cv_.wait(l, [&] { return q_.size() < 2 || stop_; });
    if (stop_) return;
    q_.push_back(std::move(b));
    cv_.notify_all();
  }
  Batch pop() {
    std::unique_lock<std::mutex> l(m_);
    cv_.wait(l, [&] { return !q_.empty(); });
    Batch b = std::move(q_.front());
    q_.pop_front();
    cv_.notify_all();
    return b;
  }
  void stop() {
    std::lock_guard<std::mutex> l(m_);
    stop_ = true;
    cv_.notify_all();
  }
  bool stopped() {
    std::lock_guard<std::mutex> l(m_);
    return stop_;
  }

 private:
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<Batch> q_;
  bool stop_ = false;
};

}  // namespace

int main(int argc, char** argv) {
  using namespace driver;
  const Args args = parse(argc, argv, true);
  try {
    const auto wall0 = Clock::now();
    sigmo::Queue queue(0);
    std::cout << "[Init] Loading Query graphs...\n";
    auto query_graphs = sigmo::io::loadCSRGraphsFromFile(args.query_file);
    if (args.query_filter)
      query_graphs.erase(std::remove_if(query_graphs.begin(), query_graphs.end(),
                                        [&](const sigmo::CSRGraph& g) {
                                          return g.getNumNodes() < args.filter_min ||
                                                 g.getNumNodes() > args.filter_max;
                                        }),
                         query_graphs.end());
    const size_t first_query_nodes = query_graphs.empty() ? 0 : query_graphs[0].getNumNodes();
    multiply(query_graphs, args.mul_query);
    if (query_graphs.size() > args.max_query_graphs)
      query_graphs.erase(query_graphs.begin() + long(args.max_query_graphs), query_graphs.end());
    std::ifstream data(args.data_file);
    if (!data) throw std::runtime_error("Could not open data file");
    auto dq = sigmo::createDeviceCSRGraph(queue, query_graphs);

    std::cout << "------------- Configs -------------\n"
              << "Batch Size: " << args.batch_size << "\n"
              << "Refinement Steps: " << args.refinement_steps << "\n";
    if (args.use_cs) std::cout << "Use CS (Partial Join based on Query Size): Yes\n";
    if (args.fused) std::cout << "Fused pipeline: Yes\n";

    // reader: lines -> parsed, multiplied, limited batches
    BatchQueue bq;
    std::exception_ptr reader_error;
    std::thread reader([&] {
      try {
        size_t produced = 0;
        std::string line;
        while (!bq.stopped()) {
          Batch b;
          std::vector<std::string> lines;
          while (lines.size() < args.batch_size && std::getline(data, line))
            if (!line.empty()) lines.push_back(line);
          if (lines.empty() || produced >= args.max_data_graphs) break;
          b.raw_lines = lines.size();
          b.graphs = sigmo::io::loadCSRGraphsFromLines(lines, true);
          multiply(b.graphs, args.mul_data);
          if (produced + b.graphs.size() > args.max_data_graphs)
            b.graphs.erase(b.graphs.begin() + long(args.max_data_graphs - produced), b.graphs.end());
          produced += b.graphs.size();
          bq.push(std::move(b));
        }
      } catch (...) {
        reader_error = std::current_exception();
      }
      Batch end;
      end.last = true;
      bq.push(std::move(end));
    });

    const size_t depth = args.use_cs ? first_query_nodes : 0;
    Times t;
    uint64_t total = 0, processed = 0;
    size_t batch_idx = 0;
    while (true) {
      Batch b = bq.pop();
      if (b.last) break;
      ++batch_idx;
      std::cout << "\n>>> Processing Batch " << batch_idx << " (" << b.raw_lines << " raw lines)...\n";
      if (depth && batch_idx == 1)
        std::cout << "[Config] --use-cs enabled. partial_match_depth set to " << depth << "\n";
      auto h0 = Clock::now();
      auto dd = sigmo::createDeviceCSRGraph(queue, b.graphs);
      t.host_setup += Clock::now() - h0;
      processed += b.graphs.size();
      PairResult r = args.fused ? run_fused(queue, dq, dd, args, t)
                                : run_stages(queue, dq, dd, args, depth, t, false);
      sigmo::destroyDeviceCSRGraph(dd, queue);
      total += r.matches;
      std::cout << "    -> Matches found: " << r.matches << "\n";
      if (!args.find_all && total > 0) {
        std::cout << "[Info] Match found (find_all=false). Stopping.\n";
        break;
      }
    }
    bq.stop();
    reader.join();
    if (reader_error) std::rethrow_exception(reader_error);

    std::cout << "\n============= Final Stats (Batched) =============\n"
              << "Total Data Graphs: " << processed << "\n";
    print_gpu_stats(t, args, "Total time (GPU Sum)");
    print_host_stats(t, args, Clock::now() - wall0, "Total time (Wall Clock)");
    std::cout << "------------- Results -------------\n";
    if (!args.skip_join) std::cout << "# Matches: " << format_number(total) << "\n";
    sigmo::destroyDeviceCSRGraph(dq, queue);
    std::cout << "[!] End\n";
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
