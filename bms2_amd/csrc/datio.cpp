// .dat text format reader and writer.
//
// One graph per line: "n#N l#L" then N pairs "id label", then "e#E" then E edges "u v [label]"
// (IntermediateGraph, graph.hpp:150-202; writer generate_data.py:76-92).  Edge labels are optional
// for the whole line: 3E trailing tokens mean labelled edges, 2E unlabelled (label 0), anything
// else is a format error (graph.hpp:184-201).  Labels are read as integers and stored as u8, as
// the reference's int-to-label_t assignment does (graph.hpp:169-170).
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hostgraph.hpp"

namespace bms2 {
namespace {

bool parse_u64(const std::string& t, unsigned long long* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  unsigned long long x = std::strtoull(t.c_str(), &end, 10);
  if (errno || *end != '\0' || t[0] == '-') return false;
  *v = x;
  return true;
}

bool parse_i64(const std::string& t, long long* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  long long x = std::strtoll(t.c_str(), &end, 10);
  if (errno || *end != '\0') return false;
  *v = x;
  return true;
}

// Parses one line into `g`; returns an empty string or the reason it is malformed.
std::string parse_line(const std::string& line, EdgeListGraph* g) {
  std::istringstream iss(line);
  std::vector<std::string> tok;
  std::string t;
  while (iss >> t) tok.push_back(t);
  size_t p = 0;
  auto counted = [&](unsigned long long* v) {
    // "n#N", "l#L", "e#E": the reference skips two characters and reads a number.
    if (p >= tok.size() || tok[p].size() < 3) return false;
    return parse_u64(tok[p++].substr(2), v);
  };
  unsigned long long n = 0, l = 0, e = 0;
  if (!counted(&n)) return "missing node count";
  if (!counted(&l)) return "missing label count";
  if (n > BMS2_MAX_GRAPH_NODES * 64ull) return "node count too large";
  g->labels.assign(n, 0);
  for (unsigned long long i = 0; i < n; ++i) {
    unsigned long long id;
    long long lab;
    if (p + 1 >= tok.size() || !parse_u64(tok[p], &id) || !parse_i64(tok[p + 1], &lab))
      return "bad node entry";
    if (id >= n) return "node id out of range";
    g->labels[id] = static_cast<uint8_t>(lab);
    p += 2;
  }
  if (!counted(&e)) return "missing edge count";
  size_t rest = tok.size() - p;
  bool labelled;
  if (rest == e * 3)
    labelled = true;
  else if (rest == e * 2)
    labelled = false;
  else
    return "edge token count matches neither 2E nor 3E";
  g->edges.clear();
  g->edges.reserve(e);
  const size_t stride = labelled ? 3 : 2;
  for (unsigned long long i = 0; i < e; ++i) {
    unsigned long long u, v;
    long long lab = 0;
    if (!parse_u64(tok[p + i * stride], &u) || !parse_u64(tok[p + i * stride + 1], &v))
      return "bad edge endpoint";
    if (labelled && !parse_i64(tok[p + i * stride + 2], &lab)) return "bad edge label";
    if (u >= n || v >= n) return "edge endpoint out of range";
    g->edges.push_back({uint32_t(u), uint32_t(v), static_cast<uint8_t>(lab)});
  }
  return {};
}

}  // namespace
}  // namespace bms2

extern "C" int bms2_load_dat(const char* path, uint64_t max_graphs, int skip_malformed,
                             bms2_host_csr* out) {
  using namespace bms2;
  if (!path || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  std::ifstream f(path);
  if (!f) return set_error(BMS2_ERR_IO, "cannot open %s", path);
  std::vector<EdgeListGraph> graphs;
  std::string line;
  uint64_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (max_graphs && graphs.size() >= max_graphs) break;
    if (line.find_first_not_of(" \t\r\n") == std::string::npos) {
      if (skip_malformed) continue;
      return set_error(BMS2_ERR_FORMAT, "%s:%llu: empty line", path, (unsigned long long)lineno);
    }
    EdgeListGraph g;
    std::string why = parse_line(line, &g);
    if (why.empty()) {
      for (const auto& e : g.edges)
        if (e.u == e.v) {
          why = "self-loop edge (not supported)";
          break;
        }
    }
    if (!why.empty()) {
      if (skip_malformed) continue;
      return set_error(BMS2_ERR_FORMAT, "%s:%llu: %s", path, (unsigned long long)lineno,
                       why.c_str());
    }
    graphs.push_back(std::move(g));
  }
  return pack_graphs(graphs, out, 1);
}

extern "C" int bms2_parse_dat_lines(const char* const* lines, uint64_t num_lines,
                                    int skip_malformed, bms2_host_csr* out) {
  using namespace bms2;
  if (!out || (num_lines && !lines)) return set_error(BMS2_ERR_INVALID, "null argument");
  std::vector<EdgeListGraph> graphs;
  graphs.reserve(num_lines);
  for (uint64_t i = 0; i < num_lines; ++i) {
    const std::string line = lines[i] ? lines[i] : "";
    EdgeListGraph g;
    std::string why;
    if (line.find_first_not_of(" \t\r\n") == std::string::npos) why = "empty line";
    if (why.empty()) why = parse_line(line, &g);
    if (why.empty())
      for (const auto& e : g.edges)
        if (e.u == e.v) {
          why = "self-loop edge (not supported)";
          break;
        }
    if (!why.empty()) {
      if (skip_malformed) continue;
      return set_error(BMS2_ERR_FORMAT, "line %llu of the batch: %s", (unsigned long long)i,
                       why.c_str());
    }
    graphs.push_back(std::move(g));
  }
  return pack_graphs(graphs, out, 1);
}

extern "C" int bms2_write_dat(const char* path, const bms2_host_csr* g) {
  using namespace bms2;
  if (!path || !g) return set_error(BMS2_ERR_INVALID, "null argument");
  int rc = validate_host_csr(g, 0xffffffffu, nullptr);
  if (rc) return rc;
  std::ofstream f(path);
  if (!f) return set_error(BMS2_ERR_IO, "cannot open %s for writing", path);
  std::string buf;
  for (uint32_t gi = 0; gi < g->num_graphs; ++gi) {
    uint32_t s = g->graph_offsets[gi], e = g->graph_offsets[gi + 1];
    buf.clear();
    buf += "n#" + std::to_string(e - s) + " l#11 ";
    for (uint32_t u = s; u < e; ++u)
      buf += std::to_string(u - s) + " " + std::to_string(g->node_labels[u]) + " ";
    std::string edges;
    uint64_t ne = 0;
    for (uint32_t u = s; u < e; ++u)
      for (uint32_t k = g->row_offsets[u]; k < g->row_offsets[u + 1]; ++k) {
        uint32_t v = g->column_indices[k];
        if (u < v) {
          edges += std::to_string(u - s) + " " + std::to_string(v - s) + " " +
                   std::to_string(g->edge_labels ? g->edge_labels[k] : 0) + " ";
          ++ne;
        }
      }
    buf += "e#" + std::to_string(ne) + " " + edges;
    if (!buf.empty() && buf.back() == ' ') buf.pop_back();
    f << buf << "\n";
  }
  if (!f) return set_error(BMS2_ERR_IO, "write to %s failed", path);
  return BMS2_OK;
}
