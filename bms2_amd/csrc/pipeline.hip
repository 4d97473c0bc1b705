// Batched pipeline: signatures -> filter -> K refinement steps -> GMCR -> join over a resident
// data batch (the stage sequence of BMS2.cpp:131-234), as two kernels per chunk of data packs.
//
// screen_kernel: one wave64 per pack of whole data graphs (<= 64 nodes, <= 8 graphs).
//   1. Node signature tuples (label, degree, path Bloom, ring counts, shell histograms) are built
//      on chip, one node per lane, and stay in VGPRs; the pack's adjacency rows go to HBM for the
//      join.
//   2. Every query is screened against the pack.  Its node records arrive through scalar loads in
//      a most-selective-first order; a node's domain is one u64 ballot over the lanes, and a graph
//      drops out as soon as one query node has no candidate in it (GMCR validity,
//      gmcr.hpp:54-88).  Pairs outside the GMCR get count 0 immediately.
//   3. GMCR pairs (query, graph, domains in matching order) are staged in LDS and appended to an
//      HBM record buffer in batches.
// join_kernel: persistent; one DFS per lane over the records (isomorphism.hpp:291-472 semantics
//   with standard backtracking), lanes refilling from a global cursor as their pair finishes.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "runtime.hpp"
#include "tuples.hpp"

#define BMS2_CONST __attribute__((address_space(4)))

namespace bms2 {

constexpr int kPackGraphs = 8;   // graphs per pack
constexpr int kStageSmall = 59;  // LDS-staged small pair records per wave
constexpr int kStageLarge = 8;   // LDS-staged large pair records per wave (>= kPackGraphs)
static_assert(kStageSmall >= kPackGraphs && kStageLarge >= kPackGraphs,
              "a stage must hold one query's records over the pack's graphs");
constexpr int kSegRecs = 64;     // most small records per segment (a flush writes one)
constexpr int kRecDwords = 48;   // screening record of one query node
constexpr int kPreDwords = 20;   // prefilter record of one query
constexpr int kSumDwords = 16;   // prefilter summary of one data graph
// the screen's tuple scratch at the start of its dynamic LDS: Bloom words (u64 units), adjacency
// rows, labels
constexpr int kScreenBloomWords = 64 * 8 * 4 / 8;
constexpr int kScreenScratchBytes = kScreenBloomWords * 8 + 64 * 8 + 64;
constexpr int kSmallNq = 8;      // queries of up to 8 nodes use the small records and join
constexpr int kSmallWords = 2 + kSmallNq;

// Screening record layout (u32 words): [0] lower end of the node's key range (label << 8 |
// degree, or the degree alone for the wildcard), [1..8] path Bloom, [9..10] ring counts, [11]
// degree | label << 8, [12 .. 11 + S] the K shell levels' label counts (S = shell_dwords: 5-bit
// fields, six to a word, over 8 compact or 16 labels per level), then at A = 12 + S: [A] 255 -
// degree (the range's span), [A + 1] the data key bits compared (0xffff, 0xff for the wildcard),
// [A + 2] the node's depth in the matching order.  The screen keeps the data node's counts as
// 15 - d per field, so one 64-bit add over twelve fields gives q + 15 - d per field (at most 30:
// no carry between fields), with the field's bit 4 set exactly when q > d.
struct ScreenParams {
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* labels;
  const uint32_t* pack_gbegin;  // [P + 1] first entry of each pack in pack_gids
  const uint32_t* pack_gids;    // graph ids, pack by pack in lane order
  uint32_t pack_begin;
  uint32_t Q;
  const uint32_t* qhead;    // [Q][4]: first record, nq, valid (nq > 1, gmcr.hpp:64), 0
  const uint32_t* qpre;     // [Q][kPreDwords] prefilter records
  const uint32_t* qrec;     // [Qn][kRecDwords] records in screening order per query
  uint32_t hist_map;        // labels of the compact shell histograms (screen_kernel<K, 2>)
  const uint8_t* jorder;    // [Q][32] local query node matched at each DFS depth
  const uint64_t* qadj8;    // [Q] byte d: earlier depths adjacent to depth d (nq <= kSmallNq)
  uint64_t* adj_out;        // [Dn] adjacency rows over pack lanes
  uint64_t* counts;         // [D][Q]
  uint64_t* valid;          // [D][VW]
  uint32_t VW;
  uint64_t* recs_s;         // small-query records (nq <= kSmallNq), kSmallWords u64 each
  uint64_t cap_s;
  uint4* segs;              // one descriptor per flush of small records: {first, count, rows, nn}
  uint64_t cap_seg;
  uint64_t* recs_l;         // large-query records, rec_words_l u64 each
  uint64_t cap_l;
  uint32_t rec_words_l;     // 2 + nq_cap
  uint64_t* ctl;            // kCtl* slots
  uint64_t* domsum;         // trace (nullable): [Qn] summed domain sizes over GMCR pairs
  uint64_t* qpairs;         // trace (nullable): [Q] GMCR pairs
};

// Control words of one chunk (u64).  Each counter that many waves update at once (the screen's
// record reservations of each class and its GMCR pair count, the joins' claim cursors) has a
// 128-byte line of its own: atomics on different words of one line wait for each other (sharing
// one line cost the screen 5 ms per run on 1M molecules x 1,024 queries).
constexpr int kCtlLine = 16;                  // u64 words per 128-byte line
constexpr int kCtlSmall = 0;                  // small records reserved (low 40 bits), segments (high 24)
constexpr int kCtlOverflow = 1;               // a record or segment buffer overflowed
constexpr int kCtlLarge = kCtlLine;           // large records reserved
constexpr int kCtlPairs = 2 * kCtlLine;       // stats: GMCR pairs, matched pairs, embeddings, DFS nodes
constexpr int kCtlSegCursor = 3 * kCtlLine;   // small join: segments claimed
constexpr int kCtlLargeCursor = 4 * kCtlLine; // large join: records claimed
constexpr int kSegShift = 40;
constexpr int kCtlWords = 5 * kCtlLine;

struct JoinParams {
  const uint64_t* recs;     // large-query records
  uint64_t rec_cap;
  uint32_t rec_words;
  uint32_t Q;
  const uint64_t* adj;      // [Dn]
  const uint64_t* qadj;     // [Q][32]: bit i < 32 of word d = depths i and d adjacent; with sym,
                            // bit 37 = depth d's candidates must exceed the node at depth bits
                            // 32-36, and bits 40-45 of word 0 = the count's multiplier
  uint64_t* counts;
  uint64_t* ctl;
  uint32_t find_first;
  uint32_t sym;             // apply the symmetry breaking of qadj
  uint32_t split;           // hand a long search's remaining root candidates to idle lanes
                            // (count-all runs without traces: partial counts add up)
  uint64_t* qnodes;         // trace (nullable): [Q] DFS nodes
  uint32_t* pair_nodes;     // trace (nullable): [D][Q] DFS nodes per pair, saturating
};

struct QueryPrepParams {
  const uint32_t* graph_offsets;
  const uint32_t* row_offsets;
  const uint32_t* column_indices;
  const uint8_t* labels;
  const uint32_t* pack_gbegin;
  const uint32_t* rec_pos;  // [Qn] record slot of each query node
  uint32_t* qrec;
  uint32_t hist_map;        // 8 labels (4 bits each) of the compact shell histograms
  uint32_t compact;         // 1: shell levels over the 8 hist_map labels, 0: over all 16
  const uint8_t* jorder;    // [Q][32] local query node matched at each DFS depth
};

// Shell histogram restricted to the 8 labels of `map` (nibble i = count of label map[i]).  A
// label that no query node carries has query count 0 at every level and passes every data node,
// so dropping it keeps the dominance test exact.
__device__ inline uint32_t compact_hist(uint64_t h, uint32_t map) {
  uint32_t out = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t lab = (map >> (4 * i)) & 15u;
    out |= uint32_t((h >> (4 * lab)) & 15u) << (4 * i);
  }
  return out;
}

// Shell counts of K levels x L labels (4 bits each) as 5-bit fields, six to a u32 (bits 0-29):
// field f = level f / L, label slot f % L.  The node test adds the query's field to the data
// node's 15 - count; the sum is at most 30, so no carry leaves a field, and its bit 4 is set iff
// the query's count is the larger.  The fields run to an even number of u32 words (the test adds
// u64s); padding fields are 0 on both sides and pass.
template <int K, int L>
constexpr int shell_dwords() {
  return 2 * ((K * L + 11) / 12);
}
constexpr uint32_t kShellFieldBits = 0x21084210u;  // bit 4 of each of the six 5-bit fields

// nib[k]: level k's L nibbles (label slot i at bits 4i); invert: store 15 - count
template <int K, int L>
__device__ inline void pack_shells(const uint64_t (&nib)[K > 0 ? K : 1], bool invert,
                                   uint32_t (&out)[shell_dwords<K, L>() > 0 ? shell_dwords<K, L>() : 1]) {
#pragma unroll
  for (int w = 0; w < shell_dwords<K, L>(); ++w) out[w] = 0;
#pragma unroll
  for (int f = 0; f < K * L; ++f) {
    uint32_t v = uint32_t(nib[f / L] >> (4 * (f % L))) & 15u;
    if (invert) v = 15u - v;
    out[f / 6] |= v << (5 * (f % 6));
  }
}

__device__ inline uint32_t cload(const uint32_t* p, uint32_t i) {
  return ((const BMS2_CONST uint32_t*)(p))[i];
}

// ------------------------------------------------------------------ query signature records

template <int K>
__global__ __launch_bounds__(64) void query_prep_kernel(QueryPrepParams p) {
  __shared__ uint64_t adj[64];
  __shared__ uint8_t lbl[64];
  __shared__ uint32_t bloom[64 * 8];
  const uint32_t pack = blockIdx.x;
  const uint32_t gb = p.pack_gbegin[pack], ge = p.pack_gbegin[pack + 1];
  const uint32_t n0 = p.graph_offsets[gb], nn = p.graph_offsets[ge] - n0;
  Tuple t;
  compute_tuples<K>(n0, nn, p.row_offsets, p.column_indices, p.labels, PackScratch{adj, lbl, bloom},
                    t);
  const uint32_t lane = lane_id();
  if (lane >= nn) return;
  const uint32_t node = n0 + lane;
  // local id inside its query graph
  uint32_t g = gb;
  while (p.graph_offsets[g + 1] <= node) ++g;
  const uint32_t local = node - p.graph_offsets[g];
  const uint32_t qd = min(t.deg, 255u), lab = t.label & 255u;
  const bool wild = lab == kWildcard;
  uint32_t* r = p.qrec + size_t(p.rec_pos[node]) * kRecDwords;
  r[0] = wild ? qd : (lab << 8) | qd;  // lower end of the (label, degree) key range
#pragma unroll
  for (int w = 0; w < 8; ++w) r[1 + w] = t.path[w];
  r[9] = uint32_t(t.cycle);
  r[10] = uint32_t(t.cycle >> 32);
  r[11] = qd | (lab << 8);
  // shell levels: 8 labels (compact) or 16 per level in 5-bit fields
  uint64_t nib[K > 0 ? K : 1];
  nib[0] = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) nib[k] = p.compact ? compact_hist(t.hist[k], p.hist_map) : t.hist[k];
  uint32_t aux;
  if (p.compact) {
    uint32_t sw[shell_dwords<K, 8>() > 0 ? shell_dwords<K, 8>() : 1];
    pack_shells<K, 8>(nib, false, sw);
#pragma unroll
    for (int w = 0; w < shell_dwords<K, 8>(); ++w) r[12 + w] = sw[w];
    aux = 12 + shell_dwords<K, 8>();
  } else {
    uint32_t sw[shell_dwords<K, 16>() > 0 ? shell_dwords<K, 16>() : 1];
    pack_shells<K, 16>(nib, false, sw);
#pragma unroll
    for (int w = 0; w < shell_dwords<K, 16>(); ++w) r[12 + w] = sw[w];
    aux = 12 + shell_dwords<K, 16>();
  }
  r[aux + 0] = 255u - qd;                 // span of the key range
  r[aux + 1] = wild ? 0xffu : 0xffffu;    // key bits compared: degree only for the wildcard
  // the node's depth in the matching order: the lane of the screen's domain registers that
  // takes its domain, so the domains arrive in the order the pair records carry them
  const uint8_t* jo = p.jorder + size_t(g) * 32;
  const uint32_t nqg = p.graph_offsets[g + 1] - p.graph_offsets[g];
  uint32_t depth = 0;
  for (uint32_t d = 0; d < nqg; ++d)
    if (jo[d] == local) depth = d;
  r[aux + 2] = depth;
  for (uint32_t w = aux + 3; w < uint32_t(kRecDwords); ++w) r[w] = 0;
}

// Per-query prefilter record (kPreDwords): [0..3] presence bits every node needs (label < 16 at
// its degree threshold, capped at 7), [4..11] union of the nodes' path Bloom filters, [12..15]
// the largest of each ring-count field over the nodes, [16] screened (nq > 1, and with qrep the
// query is its class's first), [17] first record, [18] nq, [19] records the screen tests << 24 |
// for queries of up to kSmallNq nodes the depth whose domain each depth takes, 3 bits per depth
// (a node automorphic to an earlier-screened one has its domain: its record is not tested).
__global__ void query_pre_kernel(const uint32_t* __restrict__ qhead, const uint32_t* __restrict__ qrec,
                                 uint32_t Q, const uint32_t* __restrict__ qrep,
                                 const uint8_t* __restrict__ jorder, const uint8_t* __restrict__ qorb,
                                 uint32_t* __restrict__ qpre) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const uint32_t rb = qhead[4 * q], nq = qhead[4 * q + 1], ntest = qhead[4 * q + 3];
  uint32_t out[kPreDwords] = {};
  for (uint32_t i = 0; i < nq; ++i) {
    const uint32_t* r = qrec + size_t(rb + i) * kRecDwords;
    const uint32_t d = r[11] & 255u, l = (r[11] >> 8) & 255u;
    if (l < 16) {
      const uint32_t bit = l * 8 + min(d, 7u);
      out[bit >> 5] |= 1u << (bit & 31);
    }
    for (int w = 0; w < 8; ++w) out[4 + w] |= r[1 + w];
    out[12] = max(out[12], r[9] & 0xffffu);
    out[13] = max(out[13], r[9] >> 16);
    out[14] = max(out[14], r[10] & 0xffffu);
    out[15] = max(out[15], r[10] >> 16);
  }
  out[16] = nq > 1 && (!qrep || qrep[q] == q) ? 1u : 0u;
  out[17] = rb;
  out[18] = nq;
  uint32_t map = 0;
  if (nq <= uint32_t(kSmallNq)) {
    const uint8_t* jo = jorder + size_t(q) * 32;
    uint32_t inv[kSmallNq];
    for (uint32_t d = 0; d < nq; ++d) inv[jo[d]] = d;
    for (uint32_t d = 0; d < nq; ++d) map |= inv[qorb[rb + jo[d]]] << (3 * d);
  }
  out[19] = (ntest << 24) | map;
  for (int w = 0; w < kPreDwords; ++w) qpre[size_t(q) * kPreDwords + w] = out[w];
}

// ------------------------------------------------------------------ screening

// HW: 2 for the compact shell histograms (8 labels per level), 4 for all 16 labels.  Six waves
// per SIMD: the register cap (80 VGPRs) spills two 8-byte addresses used once per 64-query block
// and once per flush, and 59 staged small and 8 large records keep the LDS at 6,384 bytes per
// wave when the large records have at most 16 nodes: 24 waves per CU fit the 160 KB after the
// allocation rounds up (62 records, 6,624 bytes, held the screen at 21 waves per CU, 28.9 against
// 27.0 ms on 1M molecules).  Seven waves (72 VGPRs, 48 small records) spilled on the per-query
// path and ran 1 ms slower.
template <int K, int HW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) void screen_kernel(
    ScreenParams p) {
  __shared__ __align__(16) uint32_t gsum_lds[kPackGraphs * kSumDwords];
  // staged pair records; before them the tuples' scratch: Bloom words, adjacency rows, labels
  // (the static LDS stays at the 512-byte summaries)
  extern __shared__ __align__(16) uint64_t stage[];
  uint64_t* const adj = stage + kScreenBloomWords;
  uint8_t* const lbl = reinterpret_cast<uint8_t*>(adj + 64);
  const uint32_t lane = lane_id();
  const uint32_t pack = p.pack_begin + blockIdx.x;
  const uint32_t gb = p.pack_gbegin[pack], ge = p.pack_gbegin[pack + 1];
  const uint32_t ng = ge - gb;
  // Graph slot j (graph pack_gids[gb + j]) occupies the next run of lanes.  mytop = last lane of
  // the node's graph (64 for lanes past the pack), base = its first node - its first lane;
  // graph-slot lanes j < ng hold the graph id and its top lane.
  uint32_t nn = 0, mytop = 64, mygraph = 0, myslot = 0, slot_top = 64, slot_gid = 0, base = 0;
  for (uint32_t j = 0; j < ng; ++j) {
    const uint32_t g = p.pack_gids[gb + j];
    const uint32_t g0 = p.graph_offsets[g], e = nn + (p.graph_offsets[g + 1] - g0);
    if (lane >= nn && lane < e) {
      mytop = e - 1;
      mygraph = g;
      myslot = j;
      base = g0 - nn;
    }
    if (lane == j) {
      slot_gid = g;
      slot_top = e > nn ? e - 1 : 64;
    }
    nn = e;
  }
  if (nn > 64) return;  // a graph above 64 nodes: the staged path serves it
  const uint64_t node_mask = low_mask(nn);
  const uint32_t rows = pack * 64;  // the pack's adjacency rows, one per lane

  Tuple t;
  compute_tuples<K>(base, nn, p.row_offsets, p.column_indices, p.labels,
                    PackScratch{adj, lbl, reinterpret_cast<uint32_t*>(stage)}, t);
  if (lane < nn) p.adj_out[rows + lane] = adj[lane];

  // label << 8 | degree (capped at 255); lanes past the pack get a key above every query's range
  const uint32_t d_deg = min(t.deg, 255u);
  const uint32_t d_key = lane < nn ? ((t.label & 255u) << 8) | d_deg : 0x10000u;
  uint32_t d_pnot[8];
#pragma unroll
  for (int w = 0; w < 8; ++w) d_pnot[w] = ~t.path[w];
  // ring counts with the top bit of each u16 set: (d | 0x8000) - q has bit 15 set iff d >= q
  const uint32_t d_cyc0 = uint32_t(t.cycle) | 0x80008000u;
  const uint32_t d_cyc1 = uint32_t(t.cycle >> 32) | 0x80008000u;
  // shell levels as 15 - count in the 5-bit fields of the query records' shell words
  constexpr int SL = HW * 4;                  // labels per level
  constexpr int SD = shell_dwords<K, SL>();   // shell u32 words of a record
  constexpr int SQ = SD / 2;                  // u64 words the test adds
  uint64_t d_sh[SQ > 0 ? SQ : 1];
  {
    uint64_t nib[K > 0 ? K : 1];
    nib[0] = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) nib[k] = HW == 2 ? compact_hist(t.hist[k], p.hist_map) : t.hist[k];
    uint32_t sw[SD > 0 ? SD : 1];
    pack_shells<K, SL>(nib, true, sw);
#pragma unroll
    for (int j = 0; j < SQ; ++j) {
      d_sh[j] = (uint64_t(sw[2 * j + 1]) << 32) | sw[2 * j];
      // opaque from here on: the test adds the query word to this value as it stands (left
      // visible, the compiler rebuilds q + 15 - d from d with a borrow chain, twice the work)
      asm volatile("" : "+v"(d_sh[j]));
    }
  }

  const uint64_t T = ballot(lane < nn && lane == mytop);  // top lane of every non-empty graph
  const uint64_t S = ((T << 1) | 1) & node_mask;           // first lane of every non-empty graph

  // LDS staging of pair records: small records (queries of up to kSmallNq nodes) first, then
  // large records; each flush appends them to HBM with one atomic per class, and the small
  // records of a flush (all of this pack) get a segment descriptor for the small join.
  const uint32_t RWL = p.rec_words_l;
  uint64_t* stage_s = stage;
  uint64_t* stage_l = stage + kStageSmall * kSmallWords;
  uint32_t ns = 0, nl = 0;
  uint64_t vbits = 0, pairs = 0;

  auto flush = [&]() {
    if (!(ns | nl)) return;
    __syncthreads();
    if (ns) {
      // one atomic reserves the records and their segments of up to kSegRecs records
      const uint32_t nsg = (ns + kSegRecs - 1) / kSegRecs;
      uint64_t old = 0;
      if (lane == 0)
        old = atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlSmall),
                        (unsigned long long)(ns | (uint64_t(nsg) << kSegShift)));
      old = uniform64(readlane64(old, 0));
      const uint64_t base = old & low_mask(kSegShift), si = old >> kSegShift;
      if (base + ns > p.cap_s || si + nsg > p.cap_seg) {
        if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(p.ctl + kCtlOverflow), 1ull);
      } else {
        uint64_t* dst = p.recs_s + base * kSmallWords;
        for (uint32_t w = lane; w < ns * kSmallWords; w += 64) dst[w] = stage_s[w];
        if (lane < nsg)
          p.segs[si + lane] = make_uint4(uint32_t(base) + lane * kSegRecs,
                                         min(uint32_t(kSegRecs), ns - lane * kSegRecs), rows, nn);
      }
    }
    if (nl) {
      uint64_t base = 0;
      if (lane == 0)
        base = atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlLarge), (unsigned long long)nl);
      base = uniform64(readlane64(base, 0));
      if (base + nl > p.cap_l) {
        if (lane == 0) atomicOr(reinterpret_cast<unsigned long long*>(p.ctl + kCtlOverflow), 1ull);
      } else {
        uint64_t* dst = p.recs_l + base * RWL;
        for (uint32_t w = lane; w < nl * RWL; w += 64) dst[w] = stage_l[w];
      }
    }
    __syncthreads();
    ns = 0;
    nl = 0;
  };

  // Per-graph summaries for the query prefilter (LDS, kSumDwords per graph slot): presence bit
  // (label l < 16, degree threshold t < 8) set when some node of the graph has label l and degree
  // >= t; union of the nodes' path Bloom filters; largest value of each ring-count field.
  uint32_t* gsum = gsum_lds;
  for (uint32_t w = lane; w < uint32_t(kPackGraphs * kSumDwords); w += 64) gsum[w] = 0;
  __syncthreads();
  if (lane < nn) {
    uint32_t* gs = gsum + myslot * kSumDwords;
    if (t.label < 16) {
      const uint32_t top = min(t.deg, 7u), base = t.label * 8;
      const uint64_t bitsl = (uint64_t(2) << top) - 1;  // thresholds 0..top
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const int lo = 32 * w;
        if (base + 8 > uint32_t(lo) && base < uint32_t(lo + 32))
          atomicOr(gs + w, uint32_t(bitsl << (base - lo)));
      }
    }
#pragma unroll
    for (int w = 0; w < 8; ++w) atomicOr(gs + 4 + w, t.path[w]);
    atomicMax(gs + 12, uint32_t(t.cycle) & 0xffffu);
    atomicMax(gs + 13, uint32_t(t.cycle) >> 16);
    atomicMax(gs + 14, uint32_t(t.cycle >> 32) & 0xffffu);
    atomicMax(gs + 15, uint32_t(t.cycle >> 48));
  }
  __syncthreads();
  const uint32_t myslotbit = lane < nn ? 1u << myslot : 0u;  // this lane's graph in a gpass mask
  // the node test's masks held in VGPRs (as scalar constants they are rebuilt per test)
  uint32_t k_sh_bits = kShellFieldBits, k_ring_bits = 0x80008000u;
  asm volatile("" : "+v"(k_sh_bits), "+v"(k_ring_bits));

  for (uint32_t qb = 0; qb < p.Q; qb += 64) {
    // 64 queries at once, one per lane, against every graph of the pack: every node's (label,
    // degree) present in the graph, every node's path bits inside the graph's union and every
    // ring field within the graph's largest (all necessary for a candidate of each node).
    const uint32_t qq = qb + lane;
    uint32_t gpass = 0, q_rb = 0, q_nq = 0, q_map = 0;
    if (qq < p.Q) {
      const uint4* pr = reinterpret_cast<const uint4*>(p.qpre + size_t(qq) * kPreDwords);
      const uint4 a = pr[0], b = pr[1], c = pr[2], rg = pr[3], e = pr[4];
      q_rb = e.y;
      q_nq = e.z;
      q_map = e.w;
      if (e.x) {
        // one graph at a time: unrolled, the two graphs' summaries set the kernel's register peak
#pragma unroll 1
        for (uint32_t j = 0; j < ng; ++j) {
          const uint4* gs = reinterpret_cast<const uint4*>(gsum + j * kSumDwords);
          const uint4 ga = gs[0], gb4 = gs[1], gc = gs[2], gr = gs[3];
          const uint32_t miss = (a.x & ~ga.x) | (a.y & ~ga.y) | (a.z & ~ga.z) | (a.w & ~ga.w) |
                                (b.x & ~gb4.x) | (b.y & ~gb4.y) | (b.z & ~gb4.z) | (b.w & ~gb4.w) |
                                (c.x & ~gc.x) | (c.y & ~gc.y) | (c.z & ~gc.z) | (c.w & ~gc.w);
          const bool rings = rg.x <= gr.x && rg.y <= gr.y && rg.z <= gr.z && rg.w <= gr.w;
          if (miss == 0 && rings) gpass |= 1u << j;
        }
      }
    }
    // the prefilter's record loads are complete here on every path (a lane whose query fails
    // the prefilter never reads them): without this the first node test of every query waits on
    // the vector-memory counter, and with it on the matching-order load issued ahead of the test
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    uint64_t cand = ballot(gpass != 0);
    while (cand) {
      const uint32_t li = uint32_t(ctz64(cand));
      cand &= cand - 1;
      const uint32_t qu = qb + li;
      const uint32_t gm = __builtin_amdgcn_readlane(gpass, li);
      const uint32_t rb = __builtin_amdgcn_readlane(q_rb, li);
      const uint32_t nq = __builtin_amdgcn_readlane(q_nq, li);
      const uint32_t qmap = __builtin_amdgcn_readlane(q_map, li);
      // query adjacency by depth, loaded ahead of the node loop
      const uint64_t qa8 =
          nq <= uint32_t(kSmallNq) ? ((const BMS2_CONST uint64_t*)(p.qadj8))[qu] : 0;
      uint64_t lanes = ballot((gm & myslotbit) != 0);
      uint64_t tops = lanes & T;
      // the domain of the node matched at depth d goes to lane d of dom_lo / dom_hi
      uint32_t dom_lo = 0, dom_hi = 0;
      // the query's nodes up to the last one not automorphic to an earlier one (at least one:
      // the prefilter passes only queries with more than one node); the loop ends there or as
      // soon as no graph is left
      const uint32_t* r = p.qrec + size_t(rb) * kRecDwords;
      const uint32_t* rend = r + size_t(qmap >> 24) * kRecDwords;
      do {
        // the whole record in one batch of scalar loads, then a branch-free test
        constexpr int A = 12 + SD;  // key span, key mask, domain offset
        uint32_t w[A + 3];
#pragma unroll
        for (int x = 0; x < A + 3; ++x) w[x] = cload(r, x);
        // keep the record's loads in one batch ahead of the uniform branches below
#pragma unroll
        for (int x = 0; x < A + 3; ++x) asm volatile("" ::"s"(w[x]));
        // label equal (any label for the wildcard) and degree at least the query node's in one
        // range compare on the masked key: (key & mask) - lo <= span
        uint64_t m = ballot(((d_key & w[A + 1]) - w[0]) <= w[A]) & lanes;
        // (the prefilter leaves a key match in every live graph for labels below 16 and
        // degrees below 8, so m is almost never 0 here: the test runs unconditionally)
        {
          uint32_t acc = 0;
#pragma unroll
          for (int x = 0; x < 8; ++x) acc |= w[1 + x] & d_pnot[x];
          const uint32_t c0 = d_cyc0 - w[9], c1 = d_cyc1 - w[10];
          uint32_t fail = ((c0 & c1) & k_ring_bits) ^ k_ring_bits;  // a ring field below the query's
          // every shell level at once (an empty query level adds 0 and passes every data node)
          uint64_t sh = 0;
#pragma unroll
          for (int j = 0; j < SQ; ++j)
            sh |= (uint64_t(w[12 + 2 * j]) | (uint64_t(w[13 + 2 * j]) << 32)) + d_sh[j];
          fail |= uint32_t(sh | (sh >> 32)) & k_sh_bits;
          m &= ballot((acc | fail) == 0);
        }
        // v_writelane ignores the exec mask; its lane select comes from m0 (with the data in an
        // SGPR too, one operand must be m0).  The compiler keeps nothing in m0 in this kernel (no
        // LDS DMA, no relative register moves: every m0 reference in its code is this block's), so
        // clobbering the reserved register is safe here.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
        asm volatile(
            "s_mov_b32 m0, %4\n\ts_nop 0\n\tv_writelane_b32 %0, %2, m0\n\tv_writelane_b32 %1, %3, m0"
            : "+v"(dom_lo), "+v"(dom_hi)
            : "s"(uint32_t(m)), "s"(uint32_t(m >> 32)), "s"(w[A + 2])
            : "m0");
#pragma clang diagnostic pop
        // Graphs left with a candidate: per segment, (m | top) - first has its top bit set
        // unless every non-top bit of m is clear, in which case m's own top bit decides.
        // m has bits only in lanes of live graphs, so a dead graph's top stays clear.
        const uint64_t u = (m | T) - S;
        const uint64_t alive = (u | m) & T;
        if (alive != tops) {
          if (!alive) rend = r + kRecDwords;  // no graph left: this node is the last
          lanes = ballot(mytop < 64 && ((alive >> mytop) & 1)) & node_mask;
        }
        tops = alive;
        r += kRecDwords;
      } while (r != rend);
      if (!tops) continue;
      if (lane < ng && slot_top < 64 && ((tops >> slot_top) & 1)) vbits |= uint64_t(1) << (qu & 63);
      // GMCR pairs -> staged records {q | g << 32, header, domains in matching order}: lane w
      // writes word w of each record.  The header is the query adjacency by depth (its byte 0 is
      // always 0) and nq for small records, the pack's first adjacency row and nq for large
      // ones; lane 2 + d holds the domain of the node matched at depth d (permuted from the
      // domain registers, no vector load on this path: a wait on one would also wait for every
      // store in flight).
      // Lanes past the query's words write 0 (the record's unused domain words).
      const bool small = nq <= uint32_t(kSmallNq);
      const uint32_t rw = small ? uint32_t(kSmallWords) : RWL;
      uint64_t* const stage_q = small ? stage_s : stage_l;
      uint32_t nrec = small ? ns : nl;
      const uint32_t d = lane - 2;  // depth of a domain lane (wraps for lanes 0 and 1)
      // a depth whose node was not tested takes the domain of the automorphic node's depth
      const uint32_t src = small ? (qmap >> (3 * (d & 7))) & 7u : d;
      const uint32_t wlo = uint32_t(__builtin_amdgcn_ds_bpermute(int(src * 4), int(dom_lo)));
      const uint32_t whi = uint32_t(__builtin_amdgcn_ds_bpermute(int(src * 4), int(dom_hi)));
      uint64_t word = d < nq ? (uint64_t(whi) << 32) | wlo : 0;
      if (lane == 0) word = uint64_t(qu);
      if (lane == 1) word = small ? (qa8 | nq) : (uint64_t(rows) | (uint64_t(nq) << 32));
      const uint64_t open = lane < 2 ? ~uint64_t(0) : 0;  // header words are not masked
      uint64_t at = tops;
      while (at) {
        const int tb = ctz64(at);
        at &= at - 1;
        const uint64_t seg = ballot(mytop == uint32_t(tb));
        const uint64_t gw = uint64_t(__builtin_amdgcn_readlane(mygraph, tb)) << 32;
        const uint64_t v = (word & (seg | open)) | (lane == 0 ? gw : 0);
        if (lane < rw) stage_q[nrec * rw + lane] = v;
        if (p.domsum) {
          if (d < nq)
            atomicAdd(reinterpret_cast<unsigned long long*>(
                          p.domsum + rb + p.jorder[size_t(qu) * 32 + d]),
                      (unsigned long long)popc64(v));
          if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(p.qpairs + qu), 1ull);
        }
        ++nrec;
        ++pairs;
      }
      if (small)
        ns = nrec;
      else
        nl = nrec;
      // room for the next query's records (one per graph of the pack)
      if (ns + ng > uint32_t(kStageSmall) || nl + ng > uint32_t(kStageLarge)) flush();
    }
    // GMCR bit rows of the block, and count 0 for every pair of the block (one coalesced
    // 512-byte store per graph; the joins overwrite or add to the GMCR pairs' counts)
    if (lane < ng) p.valid[uint64_t(slot_gid) * p.VW + (qb >> 6)] = vbits;
    for (uint32_t j = 0; j < ng; ++j) {
      const uint32_t gj = __builtin_amdgcn_readlane(slot_gid, j);
      if (qq < p.Q) p.counts[uint64_t(gj) * p.Q + qq] = 0;
    }
    vbits = 0;
  }
  flush();
  if (lane == 0 && pairs) atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs), (unsigned long long)pairs);
}

// ------------------------------------------------------------------ join

// Depth-first join of the large-query records (more than kSmallNq nodes), one (query, graph)
// pair per lane (isomorphism.hpp:291-472 semantics with standard backtracking); LDS stacks per
// depth: remaining candidates R, matched node F and its neighbour row NF; domain words, query
// adjacency and data adjacency rows read from global memory per step.  Once a wave has no
// records left, a lane whose search has run kSplitAfter steps hands its remaining root
// candidates to idle lanes, one each (split runs): the pieces add their counts atomically to
// the pair's count, which the screen zeroed, so one heavy pair no longer holds the kernel's
// tail alone.
template <int NQ>
__global__ __launch_bounds__(64) void join_kernel(JoinParams p) {
  __shared__ uint64_t R[NQ][64];   // remaining candidates per depth
  __shared__ uint64_t NF[NQ][64];  // neighbour row of the node matched per depth
  __shared__ uint8_t F[NQ][64];    // node matched per depth
  const uint32_t lane = lane_id();
  // after an overflow the screen skipped whole flushes, so records below ctl[0] may be stale;
  // the host redoes the chunk in smaller pieces
  if (p.ctl[kCtlOverflow]) return;
  const uint64_t total = min(p.ctl[kCtlLarge], p.rec_cap);
  const uint32_t RW = p.rec_words;
  bool busy = false, exhausted = false;
  uint64_t rec = 0, used = 0, cnt = 0, nodes = 0;
  uint32_t q = 0, g = 0, n0 = 0, nq = 0, mult = 1;
  int depth = 0;
  bool split = false;  // the pair's search is shared with other lanes
  uint64_t st_matched = 0, st_emb = 0, st_nodes = 0;
  const uint64_t* recs = p.recs;
  constexpr uint64_t kSplitAfter = 32;
  // records claimed by this wave from the global cursor, handed to idle lanes without atomics
  uint64_t local_next = 0, local_end = 0;
  constexpr uint64_t kClaim = 256;
  constexpr int kHandout = 16;  // idle lanes that trigger a handout (or the whole wave)
  while (true) {
    const uint64_t idle = ballot(!busy);
    if (popc64(idle) >= kHandout || idle == ~uint64_t(0)) {
      if (local_next >= local_end && !exhausted) {
        uint64_t base = 0;
        if (lane == 0)
          base = atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlLargeCursor), (unsigned long long)kClaim);
        base = uniform64(readlane64(base, 0));
        local_next = base;
        local_end = base >= total ? base : min(base + kClaim, total);
        if (base + kClaim >= total) exhausted = true;
      }
      const uint64_t take = min(local_end - local_next, uint64_t(popc64(idle)));
      const uint64_t rank = uint64_t(popc64(idle & low_mask(lane)));
      if (!busy && rank < take) {
        busy = true;
        rec = local_next + rank;
        const uint64_t h0 = recs[rec * RW], h1 = recs[rec * RW + 1];
        q = uint32_t(h0);
        g = uint32_t(h0 >> 32);
        n0 = uint32_t(h1);
        nq = uint32_t(h1 >> 32);
        depth = 0;
        used = 0;
        cnt = 0;
        nodes = 0;
        R[0][lane] = recs[rec * RW + 2];
        mult = p.sym ? max(uint32_t(p.qadj[size_t(q) * 32] >> 40) & 63u, 1u) : 1u;
        split = false;
      }
      local_next += take;
      if (p.split && exhausted && local_next >= local_end) {
        // no records left: long searches give their remaining root candidates to idle lanes
        uint64_t free_l = ballot(!busy);
        const uint64_t r0 = busy && depth > 0 ? R[0][lane] : 0;
        uint64_t donors = ballot(r0 != 0 && nodes >= kSplitAfter);
        while (free_l && donors) {
          const int dl = int(ctz64(donors));
          donors &= donors - 1;
          uint64_t give = readlane64(r0, dl);
          const uint64_t d_rec = readlane64(rec, dl);
          const uint32_t d_q = __builtin_amdgcn_readlane(q, dl), d_g = __builtin_amdgcn_readlane(g, dl);
          const uint32_t d_n0 = __builtin_amdgcn_readlane(n0, dl), d_nq = __builtin_amdgcn_readlane(nq, dl);
          const uint32_t d_mult = __builtin_amdgcn_readlane(mult, dl);
          while (give && free_l) {
            const int rl = int(ctz64(free_l));
            free_l &= free_l - 1;
            const uint64_t bit = give & (~give + 1);
            give &= give - 1;
            if (lane == uint32_t(rl)) {
              busy = true;
              split = true;
              rec = d_rec;
              q = d_q;
              g = d_g;
              n0 = d_n0;
              nq = d_nq;
              mult = d_mult;
              depth = 0;
              used = 0;
              cnt = 0;
              nodes = 0;
              R[0][lane] = bit;
            }
          }
          if (lane == uint32_t(dl)) {
            R[0][lane] = give;  // the candidates not handed out
            split = true;
          }
        }
      }
    }
    if (!ballot(busy)) {
      if (exhausted && local_next >= local_end) break;
      continue;
    }
    bool finished = false;
    if (busy) {
      const uint64_t Rd = R[depth][lane];
      if (Rd == 0) {
        if (depth == 0) {
          finished = true;
        } else {
          --depth;
          used &= ~(uint64_t(1) << F[depth][lane]);
        }
      } else {
        const int v = ctz64(Rd);
        R[depth][lane] = Rd & (Rd - 1);
        F[depth][lane] = uint8_t(v);
        const uint64_t nv = p.adj[n0 + v];
        NF[depth][lane] = nv;
        ++nodes;
        const int nd = depth + 1;
        const uint64_t vbit = uint64_t(1) << v;
        uint64_t Rn = recs[rec * RW + 2 + nd] & ~(used | vbit);
        const uint64_t aw = p.qadj[size_t(q) * 32 + nd];
        const uint32_t am = uint32_t(aw);
        Rn &= ((am >> depth) & 1) ? nv : ~nv;
        for (int i = 0; i < depth; ++i) {
          const uint64_t a = NF[i][lane];
          Rn &= ((am >> i) & 1) ? a : ~a;
        }
        if (p.sym && ((aw >> 37) & 1)) {
          // symmetry breaking: only candidates above the node matched at the orbit's first depth
          const uint32_t fu = F[(aw >> 32) & 31][lane];
          Rn &= ~((uint64_t(2) << fu) - 1);
        }
        if (uint32_t(nd) == nq - 1) {
          cnt += uint64_t(popc64(Rn));
          if (p.find_first && cnt) finished = true;
        } else {
          used |= vbit;
          depth = nd;
          R[nd][lane] = Rn;
        }
      }
    }
    if (finished) {
      const uint64_t c = p.find_first ? (cnt ? 1 : 0) : cnt * mult;
      if (split) {
        // the pair counts as matched once: when the first non-zero piece lands
        const uint64_t old = atomicAdd(
            reinterpret_cast<unsigned long long*>(p.counts + uint64_t(g) * p.Q + q), (unsigned long long)c);
        st_matched += old == 0 && c != 0 ? 1 : 0;
      } else {
        p.counts[uint64_t(g) * p.Q + q] = c;
        st_matched += c ? 1 : 0;
      }
      if (p.qnodes) atomicAdd(reinterpret_cast<unsigned long long*>(p.qnodes + q), (unsigned long long)nodes);
      if (p.pair_nodes) p.pair_nodes[uint64_t(g) * p.Q + q] = uint32_t(min(nodes, uint64_t(0xffffffffu)));
      st_emb += c;
      st_nodes += nodes;
      busy = false;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    st_matched += __shfl_xor(st_matched, off);
    st_emb += __shfl_xor(st_emb, off);
    st_nodes += __shfl_xor(st_nodes, off);
  }
  if (lane == 0 && st_matched + st_emb + st_nodes) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 1), (unsigned long long)st_matched);
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 2), (unsigned long long)st_emb);
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 3), (unsigned long long)st_nodes);
  }
}

struct JoinSmallParams {
  const uint64_t* recs;     // small-query records, kSmallWords u64 each
  const uint4* segs;        // {first record, count, first adjacency row, nn} per flush
  uint64_t cap_seg;
  uint32_t Q;
  const uint64_t* adj;      // [Dn] adjacency rows over pack lanes
  uint64_t* counts;
  uint64_t* ctl;
  uint32_t find_first;
  uint64_t* qnodes;         // trace (nullable)
  uint32_t* pair_nodes;     // trace (nullable)
  uint32_t sym;             // apply the symmetry breaking the records' headers carry
};

// Depth-first join of the small-query records (up to kSmallNq nodes), one (query, graph) pair
// per lane (isomorphism.hpp:291-472 semantics with standard backtracking).  A wave claims whole
// segments (the small records one screen flush wrote, all of one pack) into S slots, stages each
// slot's pack adjacency rows in LDS and queues the segment's records in an LDS ring.  Global
// memory is touched only at handouts, which happen once kBatch lanes are idle: idle lanes start
// the record they hold in reserve (in registers, loaded at an earlier handout), the pairs
// finished since the last handout store their counts, and lanes without a reserve take ring
// entries and issue their loads, which nothing waits for until the next handout.  Segments are
// staged the same way: descriptors arrive a claim (kSegClaim segments) at a time, and the next
// segment's adjacency row is loaded one refill before a slot takes it.  A started
// record's domain words go to LDS and its query adjacency to one register, so a DFS step touches
// LDS only.  USE[s] counts the records of slot s not yet finished (reserved ones included); a
// slot takes a new segment when it reaches zero.
template <int S>
__global__ __launch_bounds__(64) void join_small_kernel(JoinSmallParams p) {
  // A query of nq nodes (2 <= nq <= NQ) takes candidates at depths 0 .. nq - 2 and descends
  // from depths 0 .. nq - 3.  The current depth's remaining candidates (cur) and candidate base
  // (xcur) live in registers; X[d] = D[d + 1] & (constraint on depth d + 1 from depths < d) &
  // ~(nodes matched at depths < d) is set when the search descends to d.  Descending from depth
  // d saves the remaining candidates of d in RS[d] (bit d of `saved` when non-empty), so
  // backtracking jumps straight to the deepest saved depth; the node matched at depth d is
  // 6-bit field d of `fpack`, its neighbour row NF[d] (read by deeper descents only).
  constexpr int NQ = kSmallNq;
  constexpr uint32_t kRing = 128 + kSegRecs;
  constexpr uint32_t kBatch = 24;  // idle lanes that trigger a handout (16: +0.6%, 32: +4%)
  constexpr uint32_t kSegClaim = 8;  // segments claimed per atomic on the global cursor
  // a refill pushes a segment only while fewer than 128 entries are queued
  static_assert(128 + kSegRecs <= int(kRing), "ring must hold what a refill may queue");
  static_assert(S <= 16 && kSegRecs <= 64, "slot and offset must fit the ring entry");
  __shared__ uint64_t RS[NQ - 2][64];  // saved remaining candidates per depth
  __shared__ uint64_t X[NQ - 2][64];   // candidate base per depth (above)
  __shared__ uint64_t D[NQ - 2][64];   // domain of depth d + 2 (matching order)
  __shared__ uint64_t NF[NQ - 3][64];  // neighbour row of the node matched per depth
  __shared__ uint64_t ADJ[S][64];      // adjacency rows of each slot's pack
  __shared__ uint16_t RING[kRing];     // queued record: offset in its slot's segment | slot << 6
  __shared__ uint32_t FIRST[S];        // first record of each slot's segment
  __shared__ uint32_t USE[S];
  const uint32_t lane = lane_id();
  if (p.ctl[kCtlOverflow]) return;  // the host redoes the chunk
  const uint64_t nseg = min(p.ctl[kCtlSmall] >> kSegShift, p.cap_seg);
  if (lane < uint32_t(S)) USE[lane] = 0;
  __syncthreads();
  uint32_t qh = 0, qt = 0;  // ring head and tail (uniform)
  // Segments claimed by this wave and not yet staged: [lseg, lseg_end); lane i holds the
  // descriptor of segment cbase + i.  The staged segment (st_ok) has its adjacency row load
  // issued one refill ahead of the slot that takes it.
  uint64_t lseg = 0, lseg_end = 0, cbase = 0;
  uint4 cdesc = make_uint4(0, 0, 0, 0);
  bool st_ok = false;
  uint32_t st_first = 0, st_count = 0;
  uint64_t st_adj = 0;
  const auto stage = [&]() {
    if (lseg >= lseg_end) {
      uint64_t base = 0;
      if (lane == 0)
        base = atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlSegCursor),
                         (unsigned long long)kSegClaim);
      base = uniform64(readlane64(base, 0));
      cbase = lseg = base;
      lseg_end = base >= nseg ? base : min(base + kSegClaim, nseg);
      if (lane < lseg_end - lseg) cdesc = p.segs[lseg + lane];
    }
    if (lseg < lseg_end) {
      const int i = int(lseg - cbase);
      st_first = __builtin_amdgcn_readlane(cdesc.x, i);
      st_count = __builtin_amdgcn_readlane(cdesc.y, i);
      const uint32_t rows = __builtin_amdgcn_readlane(cdesc.z, i);
      const uint32_t nn = __builtin_amdgcn_readlane(cdesc.w, i);
      st_adj = lane < nn ? p.adj[rows + lane] : 0;
      st_ok = true;
      ++lseg;
    }
  };
  bool exhausted = false, busy = false, pf = false, has_res = false;
  uint32_t slot = 0, pslot = 0, q = 0, g = 0, nadj1 = 0, sym_v = 0, mult = 1;
  int depth = 0, last = 0, sym_du = 0, sdepth = -1;
  uint64_t used = 0, cnt = 0, nodes = 0, qa = 0, cur = 0, xcur = 0, fpack = 0;
  uint32_t saved = 0;
  uint64_t pw[kSmallWords];  // the lane's reserved record, loaded one handout ahead
#pragma unroll
  for (int x = 0; x < kSmallWords; ++x) pw[x] = 0;
  uint64_t res_at = 0, res = 0;  // count of the lane's last finished pair, stored at the handout
  uint64_t st_matched = 0, st_emb = 0, st_nodes = 0;
  while (true) {
    const uint64_t idle = ballot(!busy);
    const uint32_t nidle = uint32_t(popc64(idle));
    if (nidle >= kBatch || nidle == 64) {
      // Global memory is touched only here: the reserved records' loads and the result stores
      // of the previous handout are waited for (long landed), the new ones are not.
      if (!busy && pf) {
        busy = true;
        pf = false;
        slot = pslot;
        q = uint32_t(pw[0]);
        g = uint32_t(pw[0] >> 32);
        last = int(pw[1] & 0xfu) - 2;  // the depth whose candidates end a match
        qa = pw[1];  // byte d >= 1: earlier depths adjacent to depth d
        // symmetry breaking (bits 4-7 of byte 0: on | first orbit depth du; bit 7 of byte d:
        // depth d's candidates must exceed the node at du); sdepth = du when depth du + 1 is
        // constrained (checked per step at du), else -1; deeper depths fold it into X
        sym_v = 0;
        sdepth = -1;
        mult = 1;
        if (p.sym && ((qa >> 7) & 1)) {
#pragma unroll
          for (int d = 1; d < NQ; ++d) sym_v |= uint32_t((qa >> (8 * d + 7)) & 1) << d;
          sym_du = int((qa >> 4) & 7);
          if ((sym_v >> (sym_du + 1)) & 1) sdepth = sym_du;
          mult = uint32_t(__builtin_popcount(sym_v)) + 1;
        }
        // bit d: depths d and d + 1 not adjacent (bit d of byte d + 1 clear)
        nadj1 = ~0u;
#pragma unroll
        for (int d = 0; d < NQ - 1; ++d) nadj1 ^= uint32_t((qa >> (9 * d + 8)) & 1) << d;
#pragma unroll
        for (int d = 0; d < NQ - 2; ++d) D[d][lane] = pw[4 + d];  // words past 2 + nq are unused
        cur = pw[2];
        xcur = pw[3];
        X[0][lane] = xcur;
        saved = 0;
        depth = 0;
        used = 0;
        cnt = 0;
        nodes = 0;
      }
      const uint64_t want = ballot(!pf);  // lanes with no record in reserve
      const uint32_t nwant = uint32_t(popc64(want));
      if (!exhausted && qt - qh < nwant) {
        // refill free slots (nothing queued, reserved or running) with the staged segment, and
        // stage the next one
        bool loaded = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          if (!exhausted && qt - qh < nwant + 64 && uniform(USE[s]) == 0) {
            if (!st_ok) stage();
            if (!st_ok) {
              exhausted = true;
            } else {
              ADJ[s][lane] = st_adj;
              for (uint32_t e = lane; e < st_count; e += 64)
                RING[(qt + e) % kRing] = uint16_t(e | (uint32_t(s) << 6));
              if (lane == 0) {
                USE[s] = st_count;
                FIRST[s] = st_first;
              }
              qt += st_count;
              loaded = true;
              st_ok = false;
              stage();
            }
          }
        }
        if (loaded) __syncthreads();  // ADJ rows, ring entries and USE visible to every lane
      }
      if (has_res) {
        p.counts[res_at] = res;
        has_res = false;
      }
      const uint32_t take = min(qt - qh, nwant);
      const uint32_t rank = uint32_t(popc64(want & low_mask(lane)));
      if (!pf && rank < take) {
        // issue the record's loads; nothing waits for them before the next handout
        pf = true;
        const uint32_t ent = RING[(qh + rank) % kRing];
        pslot = ent >> 6;
        const uint64_t* r = p.recs + uint64_t(FIRST[pslot] + (ent & 63u)) * kSmallWords;
#pragma unroll
        for (int x = 0; x < kSmallWords; ++x) pw[x] = r[x];
      }
      qh += take;
      if (!ballot(busy)) {
        if (exhausted && qh == qt && !ballot(pf)) break;
        continue;
      }
    }
    bool finished = false;
    if (busy) {
      if (cur == 0) {
        // backtrack to the deepest depth with candidates left, or end the pair
        if (saved == 0) {
          finished = true;
        } else {
          const int dp = 31 - __builtin_clz(saved);
          saved &= ~(1u << dp);
          cur = RS[dp][lane];
          xcur = X[dp][lane];
          for (int i = dp; i < depth; ++i) used &= ~(uint64_t(1) << ((fpack >> (6 * i)) & 63));
          depth = dp;
        }
      }
      if (!finished) {
        const int v = ctz64(cur);
        const uint64_t vbit = uint64_t(1) << v;
        cur ^= vbit;
        // one round of LDS reads: the candidate's neighbour row, and what a descent needs
        const uint64_t nv = ADJ[slot][v];
        uint64_t nf[NQ - 3];
#pragma unroll
        for (int i = 0; i < NQ - 3; ++i) nf[i] = NF[i][lane];
        const uint64_t dn = D[min(depth, NQ - 3)][lane];
        // keep the descent's reads in this round instead of a second one inside the branch
        asm volatile("" ::"v"(nv), "v"(dn));
#pragma unroll
        for (int i = 0; i < NQ - 3; ++i) asm volatile("" ::"v"(nf[i]));
        ++nodes;
        const int nd = depth + 1;
        // adjacent ? N(v) : ~N(v) as N(v) ^ f with f all ones or zero
        const uint32_t f1 = uint32_t(__builtin_amdgcn_sbfe(int(nadj1), depth, 1));
        // at the orbit's first depth, when the next depth is constrained: candidates above v only
        const uint64_t drop = depth == sdepth ? (vbit << 1) - 1 : vbit;
        const uint64_t Rn = xcur & (nv ^ ((uint64_t(f1) << 32) | f1)) & ~drop;
        if (depth == last) {
          cnt += uint64_t(popc64(Rn));
          if (p.find_first && cnt) finished = true;
        } else if (Rn) {
          // descend: the candidate base of depth nd from the matches at depths <= depth
          const uint32_t an = uint32_t(qa >> (8 * (nd + 1))) & 0xffu;
          const uint32_t f0 = uint32_t(__builtin_amdgcn_sbfe(int(~an), depth, 1));
          const uint64_t c0 = nv ^ ((uint64_t(f0) << 32) | f0);
          // term i = (NF[i] & on_i) ^ f_i: on_i all ones for the depths on the path (i < depth),
          // f_i all ones where the depths are not adjacent or i is off the path (term = ones)
          const uint32_t on = (1u << depth) - 1, flip = ~(an & on);
          uint32_t clo = uint32_t(c0), chi = uint32_t(c0 >> 32);
#pragma unroll
          for (int i = 0; i < NQ - 3; ++i) {
            const uint32_t m = uint32_t(__builtin_amdgcn_sbfe(int(on), i, 1));
            const uint32_t f = uint32_t(__builtin_amdgcn_sbfe(int(flip), i, 1));
            clo &= (uint32_t(nf[i]) & m) ^ f;
            chi &= (uint32_t(nf[i] >> 32) & m) ^ f;
          }
          const uint64_t c = clo | (uint64_t(chi) << 32);
          used |= vbit;
          uint64_t xn = dn & c & ~used;
          if (cur) {
            RS[depth][lane] = cur;
            saved |= 1u << depth;
          }
          if (depth < NQ - 3) NF[depth][lane] = nv;
          fpack = (fpack & ~(uint64_t(63) << (6 * depth))) | (uint64_t(v) << (6 * depth));
          if ((sym_v >> (nd + 1)) & 1 & uint32_t(nd + 1 >= sym_du + 2)) {
            // depth nd + 1 constrained by the node at du (matched: du <= depth)
            const uint32_t fu = uint32_t(fpack >> (6 * sym_du)) & 63u;
            xn &= ~((uint64_t(2) << fu) - 1);
          }
          if (nd < NQ - 2) X[nd][lane] = xn;  // backtracking returns to depths <= NQ - 3 only
          depth = nd;
          cur = Rn;
          xcur = xn;
        }
      }
    }
    if (finished) {
      const uint64_t c = p.find_first ? (cnt ? 1 : 0) : cnt * mult;
      res_at = uint64_t(g) * p.Q + q;  // a finished lane stays idle until the next handout
      res = c;
      has_res = true;
      if (p.qnodes) atomicAdd(reinterpret_cast<unsigned long long*>(p.qnodes + q), (unsigned long long)nodes);
      if (p.pair_nodes) p.pair_nodes[uint64_t(g) * p.Q + q] = uint32_t(min(nodes, uint64_t(0xffffffffu)));
      atomicSub(&USE[slot], 1u);
      st_matched += c ? 1 : 0;
      st_emb += c;
      st_nodes += nodes;
      busy = false;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    st_matched += __shfl_xor(st_matched, off);
    st_emb += __shfl_xor(st_emb, off);
    st_nodes += __shfl_xor(st_nodes, off);
  }
  if (lane == 0 && st_matched + st_emb + st_nodes) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 1), (unsigned long long)st_matched);
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 2), (unsigned long long)st_emb);
    atomicAdd(reinterpret_cast<unsigned long long*>(p.ctl + kCtlPairs + 3), (unsigned long long)st_nodes);
  }
}

__global__ void zero_kernel(uint64_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Tables of the queries that copy their isomorphism class's results (built on the host).
struct ClassCopy {
  const uint32_t* reps;   // [nr] queries whose results others copy, ascending
  const uint2* dups;      // [nd] {copying query, index of its representative in reps}, by index
  const uint32_t* dpos;   // [nd] position of the copying query's GMCR-row word in words
  const uint32_t* chunk;  // [ceil(nr / 64) + 1] ranges of dups whose representative index is in
                          // [64 c, 64 c + 64)
  uint32_t nr;
  const uint32_t* words;  // [nw] GMCR-row words holding copying queries
  uint32_t nw;
};

// Outputs of the queries that were not screened because an isomorphic query earlier in the batch
// was: each copies its class representative's count and GMCR bit in every data graph.  The screen
// left their counts 0 and their GMCR bits clear, so only pairs whose representative is in the
// GMCR are written.  One wave per data graph (grid-stride): lanes load the GMCR bits and counts of
// 64 representatives at a time, every copying query takes its representative's from that lane,
// stores the count and ORs its bit into the wave's LDS copy of the words holding copying
// queries, which then go to the GMCR row.  stats[0..2] += GMCR pairs, matched pairs, embeddings
// of the copied pairs.  A wave takes 4 graphs per step, their loads in flight together.  Dynamic
// LDS: 4 nw u64 per wave.
__global__ __launch_bounds__(256) void copy_classes_kernel(
    uint64_t* __restrict__ counts, uint64_t* __restrict__ valid, uint32_t Q, uint32_t VW,
    uint64_t D, ClassCopy cc, uint64_t* __restrict__ stats) {
  constexpr int G = 4;  // graphs per wave and step: their loads are in flight together
  extern __shared__ uint64_t wbits_all[];
  const uint32_t lane = lane_id();
  uint64_t* wbits = wbits_all + (threadIdx.x / 64) * cc.nw * G;
  for (uint32_t j = lane; j < cc.nw * G; j += 64) wbits[j] = 0;
  const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x / 64);
  uint64_t gp = 0, mp = 0, emb = 0;
  for (uint64_t g0 = (uint64_t(blockIdx.x) * (blockDim.x / 64) + threadIdx.x / 64) * G; g0 < D;
       g0 += waves * G) {
    for (uint32_t c0 = 0; c0 < cc.nr; c0 += 64) {
      const uint32_t r = c0 + lane;
      const uint32_t q = r < cc.nr ? cc.reps[r] : 0;
      uint32_t in[G];
      uint64_t cnt[G];
#pragma unroll
      for (int k = 0; k < G; ++k)
        in[k] = (r < cc.nr && g0 + k < D) ? uint32_t(valid[(g0 + k) * VW + (q >> 6)] >> (q & 63)) & 1u : 0u;
#pragma unroll
      for (int k = 0; k < G; ++k) cnt[k] = in[k] ? counts[(g0 + k) * Q + q] : 0;
      const uint32_t b = cc.chunk[c0 / 64], e = cc.chunk[c0 / 64 + 1];
      for (uint32_t i0 = b; i0 < e; i0 += 64) {  // uniform trip count: every lane permutes
        const uint32_t i = i0 + lane;
        const uint2 pr = i < e ? cc.dups[i] : make_uint2(0, c0);
        const uint32_t wp = i < e ? cc.dpos[i] : 0;
        const int src = int(pr.y - c0) * 4;
#pragma unroll
        for (int k = 0; k < G; ++k) {
          const uint32_t vin = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(in[k])));
          const uint64_t v = uint64_t(uint32_t(__builtin_amdgcn_ds_bpermute(src, int(uint32_t(cnt[k]))))) |
                             (uint64_t(uint32_t(__builtin_amdgcn_ds_bpermute(src, int(uint32_t(cnt[k] >> 32))))) << 32);
          if (i < e && vin) {
            counts[(g0 + k) * Q + pr.x] = v;
            atomicOr(reinterpret_cast<unsigned long long*>(wbits + k * cc.nw + wp),
                     (unsigned long long)(uint64_t(1) << (pr.x & 63)));
            gp += 1;
            mp += v != 0;
            emb += v;
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t j = lane; j < cc.nw * G; j += 64) {
      const uint32_t k = j / cc.nw, w = j - k * cc.nw;
      const uint64_t bits = wbits[j];
      if (bits) {
        valid[(g0 + k) * VW + cc.words[w]] |= bits;
        wbits[j] = 0;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  // one atomic per wave and statistic
  for (int o = 32; o > 0; o >>= 1) {
    gp += __shfl_xor(gp, o);
    mp += __shfl_xor(mp, o);
    emb += __shfl_xor(emb, o);
  }
  if (lane == 0 && (gp | mp | emb)) {
    atomicAdd(reinterpret_cast<unsigned long long*>(stats + 0), (unsigned long long)gp);
    atomicAdd(reinterpret_cast<unsigned long long*>(stats + 1), (unsigned long long)mp);
    atomicAdd(reinterpret_cast<unsigned long long*>(stats + 2), (unsigned long long)emb);
  }
}

// ------------------------------------------------------------------ host side

// Greedy packing of consecutive graphs: at most 64 nodes and kPackGraphs graphs per pack.  A
// graph above 64 nodes gets a pack of its own, which the screen skips (the staged path serves it).
std::vector<uint32_t> build_packs(const std::vector<uint32_t>& graph_offsets) {
  const uint32_t G = uint32_t(graph_offsets.size() - 1);
  std::vector<uint32_t> begins;
  uint32_t g = 0;
  while (g < G) {
    begins.push_back(g);
    if (graph_offsets[g + 1] - graph_offsets[g] > 64) {
      ++g;
      continue;
    }
    uint32_t nodes = 0, cnt = 0;
    while (g < G && cnt < uint32_t(kPackGraphs)) {
      const uint32_t n = graph_offsets[g + 1] - graph_offsets[g];
      if (n > 64 || (cnt > 0 && nodes + n > 64)) break;
      nodes += n;
      ++cnt;
      ++g;
    }
  }
  begins.push_back(G);
  return begins;
}

// Pack plan of a resident data batch: packs of whole graphs with at most 64 nodes and
// kPackGraphs graphs, listed as graph ids (pack_gids[begins[p] ..]) in lane order.
struct DataPacks {
  std::vector<uint32_t> begins, gids;
  // window w covers graphs [win_graph[w], win_graph[w + 1]) in packs [win_pack[w], win_pack[w + 1])
  std::vector<uint32_t> win_pack, win_graph;
};

// Best-fit decreasing inside windows of kPackWindow consecutive graphs: a window's graphs, largest
// first, each go to the open pack with the least room that still holds them (a new pack when none
// does).  Filling packs to ~98% of their lanes instead of the ~78% of packing consecutive
// molecules cuts the screen's waves and query-node tests; the windows keep every pack's graphs
// close in id (locality of the count rows) and give the plan prefixes that cover exactly the
// graphs below a window edge.  A graph above 64 nodes gets a pack of its own, which the screen
// skips.
constexpr uint32_t kPackWindow = BMS2_PACK_WINDOW;
DataPacks build_data_packs(const std::vector<uint32_t>& go) {
  const uint32_t G = uint32_t(go.size() - 1);
  DataPacks d;
  d.gids.reserve(G);
  std::vector<uint32_t> ids;
  struct Bin {
    uint32_t nodes;
    std::vector<uint32_t> g;
  };
  std::vector<Bin> bins;
  std::vector<std::vector<uint32_t>> by_room(65);  // open bins by free lanes
  for (uint32_t w0 = 0; w0 < G; w0 += kPackWindow) {
    const uint32_t w1 = std::min(G, w0 + kPackWindow);
    d.win_pack.push_back(uint32_t(d.begins.size()));
    d.win_graph.push_back(w0);
    ids.clear();
    for (uint32_t g = w0; g < w1; ++g) ids.push_back(g);
    std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) {
      return go[a + 1] - go[a] > go[b + 1] - go[b];
    });
    bins.clear();
    for (auto& v : by_room) v.clear();
    for (uint32_t g : ids) {
      const uint32_t n = go[g + 1] - go[g];
      if (n > 64) {  // alone, emitted in place
        d.begins.push_back(uint32_t(d.gids.size()));
        d.gids.push_back(g);
        continue;
      }
      uint32_t b = UINT32_MAX;
      for (uint32_t r = n; r <= 64 && b == UINT32_MAX; ++r)
        if (!by_room[r].empty()) {
          b = by_room[r].back();
          by_room[r].pop_back();
        }
      if (b == UINT32_MAX) {
        b = uint32_t(bins.size());
        bins.push_back(Bin{0, {}});
      }
      Bin& bin = bins[b];
      bin.nodes += n;
      bin.g.push_back(g);
      if (bin.g.size() < size_t(kPackGraphs) && bin.nodes < 64) by_room[64 - bin.nodes].push_back(b);
    }
    for (Bin& bin : bins) {
      std::sort(bin.g.begin(), bin.g.end());
      d.begins.push_back(uint32_t(d.gids.size()));
      d.gids.insert(d.gids.end(), bin.g.begin(), bin.g.end());
    }
  }
  d.win_pack.push_back(uint32_t(d.begins.size()));
  d.win_graph.push_back(G);
  d.begins.push_back(uint32_t(d.gids.size()));
  return d;
}

// Data graphs above 64 nodes: zero their count rows (the joins add into them).
__global__ void big_zero_kernel(const uint32_t* __restrict__ big_gid, uint32_t nbig, uint32_t Q,
                                uint32_t g_limit, uint64_t* __restrict__ counts) {
  const uint32_t b = blockIdx.x;
  if (b >= nbig) return;
  const uint32_t g = big_gid[b];
  if (g >= g_limit) return;
  for (uint32_t q = threadIdx.x; q < Q; q += blockDim.x) counts[uint64_t(g) * Q + q] = 0;
}

template <class F>
int dispatch_levels(uint32_t K, F&& f) {
  switch (K) {
    case 0: return f(std::integral_constant<int, 0>{});
    case 1: return f(std::integral_constant<int, 1>{});
    case 2: return f(std::integral_constant<int, 2>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 5: return f(std::integral_constant<int, 5>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 7: return f(std::integral_constant<int, 7>{});
    case 8: return f(std::integral_constant<int, 8>{});
  }
  return set_error(BMS2_ERR_UNSUPPORTED, "refinement steps %u > %d", K, kMaxLevels);
}

}  // namespace bms2

using namespace bms2;

struct bms2_pipeline_s {
  bms2_queue q;
  bms2_device_csr data, query;
  bms2_pipeline_options opt;
  uint32_t num_packs = 0, num_qpacks = 0, nq_cap = 0, rec_words_l = 0;
  uint32_t max_nq = 0;
  uint32_t hist_map = 0, hist_compact = 0;  // labels of compact shell histograms
  uint64_t cap_s = 0, cap_seg = 0, cap_l = 0;
  uint64_t alloc_s = 0, alloc_seg = 0, alloc_l_words = 0;  // allocated record buffer sizes
  // measured records / segments per pack of each class, sizes the chunks
  double dens_s = 0.0, dens_seg = 0.0, dens_l = 0.0;
  // launch shape (autotuned): waves per CU of each join, LDS slots per small-join wave
  int cus = 256;
  uint32_t small_waves = 8, large_waves = 12, slots = 16;
  BigJoinShape big_shape;  // joins of the graphs above 64 nodes
  uint32_t* d_packs = nullptr;      // pack p: graphs d_pack_gids[d_packs[p] .. d_packs[p + 1])
  uint32_t* d_pack_gids = nullptr;
  std::vector<uint32_t> win_pack, win_graph;  // pack windows (see build_data_packs)
  uint32_t* d_qpacks = nullptr;
  uint32_t* d_qhead = nullptr;
  uint32_t* d_qpre = nullptr;
  uint32_t* d_qrec = nullptr;
  uint32_t* d_qpos = nullptr;
  uint8_t* d_jorder = nullptr;
  uint64_t* d_qadj = nullptr;
  uint64_t* d_adj = nullptr;
  uint64_t* d_qadj8 = nullptr;
  uint64_t* d_recs_s = nullptr;
  uint4* d_segs = nullptr;
  uint64_t* d_recs_l = nullptr;
  uint64_t* d_ctl = nullptr;
  // host copy of the query batch (matching orders are derived from it), a view of it, and the
  // automorphism orbits of every query (local lowest node per orbit)
  std::vector<uint32_t> qgo, qro, qci;
  std::vector<uint8_t> qlab, orbit;
  bms2_host_csr qhv{};
  // isomorphism classes of the query batch: only each class's first query is screened and
  // joined, the others copy its outputs (copy_classes_kernel) unless dedup is off
  bool dedup = true;
  uint32_t num_dups = 0;
  uint32_t* d_qrep = nullptr;       // [Q] first query of each query's class
  uint8_t* d_qorb = nullptr;        // [Qn] lowest local node of each query node's automorphism
                                    // orbit (queries of up to kSmallNq nodes; itself otherwise)
  uint32_t* d_copy_u32 = nullptr;   // ClassCopy's u32 tables: reps, chunk, words, dpos
  uint2* d_copy_pairs = nullptr;    // ClassCopy's dups
  ClassCopy copy{};
  uint64_t* d_dup_stats = nullptr;  // [4] statistics of the copied pairs
  // data graphs above 64 nodes: their own compact batch through the staged kernels
  uint32_t num_big = 0;
  bms2_device_csr big{};
  uint32_t* d_big_gid = nullptr;   // original id of each graph of the batch
  // screen order: graphs of 65-128 nodes, then 129-256, each by original id (host: original ids)
  std::vector<uint32_t> big_narrow_gid, big_wide_gid;
  uint32_t* d_big_list = nullptr;   // their batch indices
  bms2_path_signature big_path = nullptr;
  bms2_cycle_signature big_cyc = nullptr;
  uint64_t* d_big_lev = nullptr;    // [K][big nodes] shell levels
  uint64_t* d_big_qlev = nullptr;   // [K][Qn]
  uint32_t* d_big_qrec = nullptr;   // [Qn][kBigRecDwords] screening records
  BigEntries be{};                  // the screen's GMCR entries and domain records
  uint64_t* d_big_stats = nullptr;  // [1..3] statistics of the joins
  // the big-graph path runs on queues of its own beside the fused chunk (and its 4-word class
  // beside its 2-word class), forked after the run's query records and joined before its end
  bms2_queue side_q = nullptr, side_q2 = nullptr;
  hipEvent_t big_fork = nullptr, big_done = nullptr;
  hipEvent_t big_screened = nullptr;  // the chunk's screen is done: the big-graph joins may start
  bool big_pending = false;
  uint64_t big_pairs = 0;           // GMCR pairs of the pending big-graph path
  // the single-chunk kernels captured in a hipGraph, valid for these outputs and K
  hipGraphExec_t graph = nullptr;
  // with graphs above 64 nodes the chunk is two graphs, the screen's and the joins', so the
  // big-graph path can fork between them
  hipGraphExec_t graph_joins = nullptr;
  uint64_t* graph_counts = nullptr;
  uint64_t* graph_valid = nullptr;
  uint32_t graph_packs = 0;
};

namespace {

// Destroys the captured chunk graphs (the next run captures them again).
void drop_graphs(bms2_pipeline_s* p) {
  if (p->graph) (void)hipGraphExecDestroy(p->graph);
  if (p->graph_joins) (void)hipGraphExecDestroy(p->graph_joins);
  p->graph = p->graph_joins = nullptr;
}

// Releases the big-graph screen's entry and record buffers (the control words stay).
void free_big_entries(bms2_pipeline_s* p) {
  for (void* b : {(void*)p->be.gidx, (void*)p->be.rec_off, (void*)p->be.runs, (void*)p->be.recs})
    (void)hipFree(b);
  p->be.gidx = nullptr;
  p->be.rec_off = p->be.recs = nullptr;
  p->be.runs = nullptr;
  p->be.cap_e = p->be.cap_r = p->be.cap_words = 0;
}

// Sizes the entry buffers for cap_e entries and cap_r runs per class and the records for
// cap_words words.
int ensure_big_entries(bms2_pipeline_s* p, uint64_t cap_e, uint64_t cap_r, uint64_t cap_words) {
  if (p->be.gidx && p->be.cap_e >= cap_e && p->be.cap_r >= cap_r && p->be.cap_words >= cap_words)
    return BMS2_OK;
  cap_e = std::max(cap_e, p->be.cap_e);
  cap_r = std::max(cap_r, p->be.cap_r);
  cap_words = std::max(cap_words, p->be.cap_words);
  BMS2_HIP(hipStreamSynchronize(p->side_q->stream));
  free_big_entries(p);
  const uint64_t ne = kBigClasses * cap_e;
  if (int rc = device_alloc((void**)&p->be.gidx, 4 * ne)) return rc;
  if (int rc = device_alloc((void**)&p->be.rec_off, 8 * ne)) return rc;
  if (int rc = device_alloc((void**)&p->be.runs, 16 * kBigClasses * cap_r)) return rc;
  if (int rc = device_alloc((void**)&p->be.recs, 8 * cap_words)) return rc;
  p->be.cap_e = cap_e;
  p->be.cap_r = cap_r;
  p->be.cap_words = cap_words;
  return BMS2_OK;
}

void pipeline_free(bms2_pipeline p) {
  if (!p) return;
  if (p->side_q) (void)bms2_queue_wait(p->side_q);
  if (p->side_q2) (void)bms2_queue_wait(p->side_q2);
  (void)hipFree(p->d_packs);
  (void)hipFree(p->d_pack_gids);
  (void)hipFree(p->d_qpacks);
  (void)hipFree(p->d_qhead);
  (void)hipFree(p->d_qpre);
  (void)hipFree(p->d_qrec);
  (void)hipFree(p->d_qpos);
  (void)hipFree(p->d_jorder);
  (void)hipFree(p->d_qadj);
  (void)hipFree(p->d_adj);
  (void)hipFree(p->d_qadj8);
  (void)hipFree(p->d_recs_s);
  (void)hipFree(p->d_segs);
  (void)hipFree(p->d_recs_l);
  (void)hipFree(p->d_ctl);
  (void)hipFree(p->d_qrep);
  (void)hipFree(p->d_qorb);
  (void)hipFree(p->d_copy_u32);
  (void)hipFree(p->d_copy_pairs);
  (void)hipFree(p->d_dup_stats);
  drop_graphs(p);
  (void)hipFree(p->d_big_gid);
  (void)hipFree(p->d_big_list);
  (void)hipFree(p->d_big_lev);
  (void)hipFree(p->d_big_qlev);
  (void)hipFree(p->d_big_qrec);
  free_big_entries(p);
  (void)hipFree(p->be.ctl);
  (void)hipFree(p->d_big_stats);
  if (p->big_path) bms2_path_signature_destroy(p->big_path);
  if (p->big_cyc) bms2_cycle_signature_destroy(p->big_cyc);
  if (p->big.graph_offsets) bms2_device_csr_destroy(p->q, &p->big);
  if (p->big_fork) (void)hipEventDestroy(p->big_fork);
  if (p->big_screened) (void)hipEventDestroy(p->big_screened);
  if (p->big_done) (void)hipEventDestroy(p->big_done);
  if (p->side_q2) bms2_queue_destroy(p->side_q2);
  if (p->side_q) bms2_queue_destroy(p->side_q);
  delete p;
}

// Pair-record budget: a quarter of free device memory (counting record buffers already held),
// at most 16 GiB.
uint64_t record_budget_bytes(uint64_t held = 0) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 1ull << 30;
  return std::min<uint64_t>((free_b + held) / 4, 16ull << 30);
}

// Data graphs above 64 nodes: a compact device batch of them (node ids renumbered) and their
// original graph ids; the staged objects that serve them depend on the query batch.
int build_big_batch(bms2_pipeline_s* p, const std::vector<uint32_t>& go) {
  const uint32_t G = uint32_t(go.size() - 1);
  std::vector<uint32_t> gid;
  for (uint32_t g = 0; g < G; ++g)
    if (go[g + 1] - go[g] > 64) gid.push_back(g);
  p->num_big = uint32_t(gid.size());
  if (gid.empty()) return BMS2_OK;
  const bms2_device_csr& d = p->data;
  std::vector<uint32_t> ro(d.total_nodes + 1), ci(d.total_edges);
  std::vector<uint8_t> lab(d.total_nodes);
  BMS2_HIP(hipMemcpy(ro.data(), d.row_offsets, 4ull * ro.size(), hipMemcpyDeviceToHost));
  if (!ci.empty()) BMS2_HIP(hipMemcpy(ci.data(), d.column_indices, 4ull * ci.size(), hipMemcpyDeviceToHost));
  BMS2_HIP(hipMemcpy(lab.data(), d.node_labels, lab.size(), hipMemcpyDeviceToHost));
  std::vector<uint32_t> bgo{0}, bro{0}, bci;
  std::vector<uint8_t> blab, belab;
  for (uint32_t g : gid) {
    const uint32_t s = go[g], e = go[g + 1], nb = bgo.back();
    for (uint32_t u = s; u < e; ++u) {
      for (uint32_t k = ro[u]; k < ro[u + 1]; ++k) bci.push_back(ci[k] - s + nb);
      bro.push_back(uint32_t(bci.size()));
      blab.push_back(lab[u]);
    }
    bgo.push_back(nb + (e - s));
  }
  belab.assign(bci.size(), 0);  // edge labels are never read by matching (R5)
  bms2_host_csr h{};
  h.num_graphs = p->num_big;
  h.total_nodes = bgo.back();
  h.total_edges = bci.size();
  h.graph_offsets = bgo.data();
  h.row_offsets = bro.data();
  h.column_indices = bci.data();
  h.node_labels = blab.data();
  h.edge_labels = belab.data();
  if (int rc = bms2_device_csr_create(p->q, &h, &p->big)) return rc;
  // screen order: graphs of up to 128 nodes (two 64-node slots per lane), then the larger ones
  std::vector<uint32_t> list;
  p->big_narrow_gid.clear();
  p->big_wide_gid.clear();
  for (int wide = 0; wide < 2; ++wide)
    for (uint32_t b = 0; b < p->num_big; ++b)
      if ((bgo[b + 1] - bgo[b] > 128) == bool(wide)) {
        list.push_back(b);
        (wide ? p->big_wide_gid : p->big_narrow_gid).push_back(gid[b]);
      }
  const uint64_t K = std::max<uint32_t>(p->opt.refinement_steps, 1);
  if (int rc = device_alloc((void**)&p->d_big_gid, 4ull * gid.size())) return rc;
  if (int rc = device_alloc((void**)&p->d_big_list, 4ull * list.size())) return rc;
  if (int rc = device_alloc((void**)&p->d_big_lev, 8 * K * std::max<uint64_t>(h.total_nodes, 1))) return rc;
  if (int rc = device_alloc((void**)&p->be.ctl, 8 * kBigCtlWords)) return rc;
  if (int rc = device_alloc((void**)&p->d_big_stats, 8 * 8)) return rc;
  if (int rc = bms2_queue_create(p->q->device, &p->side_q)) return rc;
  if (int rc = bms2_queue_create(p->q->device, &p->side_q2)) return rc;
  BMS2_HIP(hipEventCreateWithFlags(&p->big_fork, hipEventDisableTiming));
  BMS2_HIP(hipEventCreateWithFlags(&p->big_screened, hipEventDisableTiming));
  BMS2_HIP(hipEventCreateWithFlags(&p->big_done, hipEventDisableTiming));
  BMS2_HIP(hipMemcpy(p->d_big_gid, gid.data(), 4ull * gid.size(), hipMemcpyHostToDevice));
  BMS2_HIP(hipMemcpy(p->d_big_list, list.data(), 4ull * list.size(), hipMemcpyHostToDevice));
  return BMS2_OK;
}

// Enqueues the screen and joins of the graphs above 64 nodes with an original id below g_limit
// on the side queue: the screen after the main stream's work up to p->big_fork (this run's query
// records), the joins after p->big_screened (the first chunk's screen, so they run beside the
// chunk's joins); they write those graphs' rows of the dense outputs.  The side queue's stream is synchronised
// once, to size the screen's buffers; big_finish joins it.  The signatures are the staged
// kernels' (path signatures; ring counts, which stay zero on graphs above 64 nodes as in
// signature.hpp:614) and the shell levels of every refinement step; the screen replaces the
// filter, the refinement steps and the GMCR of the staged sequence (same domains on every GMCR
// pair), the join is the staged join's kernel over the screen's runs of entries.
int big_start(bms2_pipeline_s* p, uint64_t* d_counts, uint64_t* d_valid, uint32_t g_limit) {
  p->big_pending = false;
  if (!p->num_big) return BMS2_OK;
  bms2_queue q = p->side_q;
  hipStream_t s = q->stream;
  const bms2_device_csr* dq = &p->query;
  const bms2_device_csr* db = &p->big;
  const uint32_t K = p->opt.refinement_steps;
  const uint32_t n_narrow = uint32_t(std::lower_bound(p->big_narrow_gid.begin(), p->big_narrow_gid.end(), g_limit) - p->big_narrow_gid.begin());
  const uint32_t n_wide = uint32_t(std::lower_bound(p->big_wide_gid.begin(), p->big_wide_gid.end(), g_limit) - p->big_wide_gid.begin());
  if (!n_narrow && !n_wide) return BMS2_OK;
  BMS2_HIP(hipStreamWaitEvent(s, p->big_fork, 0));
  if (int rc = bms2_path_signature_generate(p->big_path, db, BMS2_SCOPE_DATA, nullptr)) return rc;
  if (int rc = bms2_cycle_signature_generate(p->big_cyc, db, BMS2_SCOPE_DATA, nullptr)) return rc;
  if (int rc = bms2_path_signature_generate(p->big_path, dq, BMS2_SCOPE_QUERY, nullptr)) return rc;
  if (int rc = bms2_cycle_signature_generate(p->big_cyc, dq, BMS2_SCOPE_QUERY, nullptr)) return rc;
  if (int rc = big_levels(s, db, K, p->d_big_lev)) return rc;
  if (int rc = big_levels(s, dq, K, p->d_big_qlev)) return rc;
  if (int rc = big_query_records(s, dq, path_signature_bits(p->big_path, BMS2_SCOPE_QUERY),
                                 cycle_signature_values(p->big_cyc, BMS2_SCOPE_QUERY), p->d_big_qlev,
                                 K, p->d_qpos, p->d_jorder, p->d_big_qrec))
    return rc;
  const uint32_t wide_above = big_wide_above(p->big_shape);
  const uint32_t Q = dq->num_graphs;
  uint64_t ctl[kBigCtlWords];
  hipLaunchKernelGGL(big_zero_kernel, dim3(p->num_big), dim3(256), 0, s, p->d_big_gid, p->num_big,
                     Q, g_limit, d_counts);
  BMS2_CHECK_LAUNCH("big_zero_kernel");
  // first sizes: 128 entries and 2 runs per graph, 48 record words per entry; then what a run
  // measured
  if (int rc = ensure_big_entries(p, std::max<uint64_t>(1024, 128ull * p->num_big),
                                  std::max<uint64_t>(1024, 2ull * p->num_big),
                                  std::max<uint64_t>(1024, 128ull * 48 * p->num_big)))
    return rc;
  for (int attempt = 0;; ++attempt) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<uint64_t*>(p->be.ctl),
                       uint32_t(kBigCtlWords));
    BMS2_CHECK_LAUNCH("zero_kernel");
    if (int rc = big_screen(s, db, path_signature_bits(p->big_path, BMS2_SCOPE_DATA),
                            cycle_signature_values(p->big_cyc, BMS2_SCOPE_DATA), p->d_big_lev, K,
                            p->d_big_list, n_narrow, p->d_big_list + p->big_narrow_gid.size(),
                            n_wide, wide_above, Q, p->d_qpre, kPreDwords, p->d_big_qrec, p->be,
                            p->d_big_gid, d_valid))
      return rc;
    BMS2_HIP(hipMemcpyAsync(ctl, p->be.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
    BMS2_HIP(hipStreamSynchronize(s));
    if (!ctl[kBigCtlOverflow]) break;
    if (attempt)
      return set_error(BMS2_ERR_CAPACITY, "big-graph screen overflowed its measured buffers");
    // the counters ran on past the capacities: size to them, with room for the next batch
    const uint64_t need_e = std::max(ctl[0], ctl[1]), need_w = ctl[kBigCtlRecWords];
    const uint64_t need_r = std::max(ctl[kBigClasses], ctl[kBigClasses + 1]);
    if (int rc = ensure_big_entries(p, need_e + need_e / 8 + 1024, need_r + need_r / 8 + 1024,
                                    need_w + need_w / 8 + 1024))
      return rc;
  }
  const uint64_t count[kBigClasses] = {ctl[0], ctl[1]};
  const uint64_t runs[kBigClasses] = {ctl[kBigClasses], ctl[kBigClasses + 1]};
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, p->d_big_stats, 8u);
  BMS2_HIP(hipStreamWaitEvent(s, p->big_screened, 0));  // beside the chunk's joins
  if (int rc = big_join(q, p->side_q2, dq, db, p->be, count, runs, int(p->opt.find_first),
                        p->d_jorder, d_counts, p->d_big_gid, p->d_big_stats, &p->big_shape))
    return rc;
  BMS2_HIP(hipEventRecord(p->big_done, s));
  p->big_pairs = count[0] + count[1];
  p->big_pending = true;
  return BMS2_OK;
}

// Makes the main stream wait for the big-graph path and adds its statistics to stats.
int big_finish(bms2_pipeline_s* p, uint64_t stats[8]) {
  if (!p->big_pending) return BMS2_OK;
  p->big_pending = false;
  BMS2_HIP(hipStreamWaitEvent(p->q->stream, p->big_done, 0));
  uint64_t st[8];
  BMS2_HIP(hipMemcpyAsync(st, p->d_big_stats, sizeof(st), hipMemcpyDeviceToHost, p->side_q->stream));
  BMS2_HIP(hipStreamSynchronize(p->side_q->stream));
  stats[0] += p->big_pairs;
  for (int k = 1; k < 4; ++k) stats[k] += st[k];
  return BMS2_OK;
}

// Matching order per query (identity when order is null, the reference's, isomorphism.hpp:
// 371-447) and, per depth d, the bit mask of earlier depths adjacent to it in the query.
int upload_order(bms2_pipeline_s* p, const uint8_t* order) {
  const uint32_t Q = p->query.num_graphs;
  if (!Q) return BMS2_OK;
  std::vector<uint8_t> jorder(size_t(Q) * 32, 0);
  std::vector<uint64_t> qadj(size_t(Q) * 32, 0);
  std::vector<uint64_t> qadj8(Q, 0);
  for (uint32_t g = 0; g < Q; ++g) {
    const uint32_t b = p->qgo[g], n = p->qgo[g + 1] - b;
    uint8_t* o = jorder.data() + size_t(g) * 32;
    if (order) {
      uint32_t seen = 0;
      for (uint32_t d = 0; d < n; ++d) {
        const uint32_t u = order[size_t(g) * 32 + d];
        if (u >= n || ((seen >> u) & 1))
          return set_error(BMS2_ERR_INVALID, "order of query %u is not a permutation of 0..%u", g,
                           n - 1);
        seen |= 1u << u;
        o[d] = uint8_t(u);
      }
    } else {
      for (uint32_t d = 0; d < n; ++d) o[d] = uint8_t(d);
    }
    for (uint32_t d = 0; d < n; ++d) {
      const uint32_t u = b + o[d];
      uint32_t w = 0;
      for (uint32_t i = 0; i < d; ++i) {
        const uint32_t v = b + o[i];
        for (uint32_t k = p->qro[u]; k < p->qro[u + 1]; ++k)
          if (p->qci[k] == v) {
            w |= 1u << i;
            break;
          }
      }
      qadj[size_t(g) * 32 + d] = w;
      if (n <= uint32_t(kSmallNq)) qadj8[g] |= uint64_t(w) << (8 * d);
    }
    // symmetry breaking of the joins under this order
    uint8_t du = 0xff;
    uint32_t vmask = 0;
    query_symmetry(&p->qhv, g, o, p->orbit.data() + b, &du, &vmask);
    if (du != 0xff && n > uint32_t(kSmallNq)) {
      for (uint32_t d = 0; d < n; ++d)
        if ((vmask >> d) & 1) qadj[size_t(g) * 32 + d] |= (uint64_t(1) << 37) | (uint64_t(du) << 32);
      qadj[size_t(g) * 32] |= uint64_t(__builtin_popcount(vmask) + 1) << 40;
    } else if (du != 0xff) {
      // the small records' header: bits 4-7 of byte 0 (on | du), bit 7 of byte d (d in vmask)
      qadj8[g] |= 0x80u | (uint64_t(du) << 4);
      for (uint32_t d = 1; d < n; ++d)
        if ((vmask >> d) & 1) qadj8[g] |= uint64_t(1) << (8 * d + 7);
    }
  }
  hipStream_t s = p->q->stream;
  BMS2_HIP(hipMemcpyAsync(p->d_jorder, jorder.data(), jorder.size(), hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemcpyAsync(p->d_qadj, qadj.data(), 8ull * qadj.size(), hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemcpyAsync(p->d_qadj8, qadj8.data(), 8ull * qadj8.size(), hipMemcpyHostToDevice, s));
  BMS2_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
  return BMS2_OK;
}

}  // namespace

namespace {

int check_queries(const bms2_device_csr* query) {
  if (query->max_degree > 27)
    return set_error(BMS2_ERR_UNSUPPORTED,
                     "query node degree above 27: the reference's 128-entry ring-count stack could "
                     "overflow (signature.hpp:638,697), which the fused ring counts do not model");
  if (query->max_graph_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "query with %u nodes (limit %u)",
                     query->max_graph_nodes, BMS2_MAX_QUERY_NODES);
  return BMS2_OK;
}

// Releases everything that depends on the query batch (the data plan stays).
void free_query_side(bms2_pipeline_s* p) {
  for (void* b : {(void*)p->d_qpacks, (void*)p->d_qhead, (void*)p->d_qpre, (void*)p->d_qrec,
                  (void*)p->d_qpos, (void*)p->d_jorder, (void*)p->d_qadj, (void*)p->d_qadj8,
                  (void*)p->d_big_qlev, (void*)p->d_big_qrec, (void*)p->d_qrep,
                  (void*)p->d_copy_u32, (void*)p->d_copy_pairs, (void*)p->d_qorb})
    (void)hipFree(b);
  p->d_qorb = nullptr;
  p->d_qrep = p->d_copy_u32 = nullptr;
  p->d_copy_pairs = nullptr;
  p->copy = ClassCopy{};
  p->num_dups = 0;
  p->d_qpacks = p->d_qhead = p->d_qpre = p->d_qrec = p->d_qpos = nullptr;
  p->d_qadj = nullptr;
  p->d_jorder = nullptr;
  p->d_qadj8 = nullptr;
  p->d_big_qlev = nullptr;
  p->d_big_qrec = nullptr;
  if (p->big_path) bms2_path_signature_destroy(p->big_path);
  if (p->big_cyc) bms2_cycle_signature_destroy(p->big_cyc);
  p->big_path = nullptr;
  p->big_cyc = nullptr;
  drop_graphs(p);
  p->dens_s = p->dens_seg = p->dens_l = 0.0;
}

// Grows a record buffer to hold `want` items of `item` bytes (kept when already large enough).
int ensure_buffer(void** ptr, uint64_t* have, uint64_t want, uint64_t item) {
  if (*ptr && *have >= want) return BMS2_OK;
  (void)hipFree(*ptr);
  *ptr = nullptr;
  *have = 0;
  if (int rc = device_alloc(ptr, want * item)) return rc;
  *have = want;
  return BMS2_OK;
}

// Binds a query batch: screening records, orders, record buffers, and the big graphs' staged
// objects (sized by query nodes).
int bind_queries(bms2_pipeline_s* p, const bms2_device_csr* query) {
  hipStream_t s = p->q->stream;
  int rc = BMS2_OK;
  p->query = *query;
  p->nq_cap = query->max_graph_nodes <= 16 ? 16 : 32;
  p->rec_words_l = 2 + std::max<uint32_t>(query->max_graph_nodes, 2);
  p->max_nq = query->max_graph_nodes;
  const uint32_t Q = query->num_graphs;
  const uint64_t Qn = query->total_nodes;
  std::vector<uint32_t> qgo(Q + 1), qro(Qn + 1), qci(query->total_edges);
  std::vector<uint8_t> qlab(Qn);
  if (hipMemcpy(qgo.data(), query->graph_offsets, 4ull * qgo.size(), hipMemcpyDeviceToHost) ||
      hipMemcpy(qro.data(), query->row_offsets, 4ull * qro.size(), hipMemcpyDeviceToHost) ||
      (qci.size() && hipMemcpy(qci.data(), query->column_indices, 4ull * qci.size(),
                               hipMemcpyDeviceToHost)) ||
      (qlab.size() && hipMemcpy(qlab.data(), query->node_labels, qlab.size(), hipMemcpyDeviceToHost)))
    return set_error(BMS2_ERR_HIP, "copy of the query batch failed");
  std::vector<uint32_t> label_freq(256, 0);
  for (uint8_t l : qlab) label_freq[l]++;
  // shell histograms over the labels < 16 the queries carry, when there are at most 8 of them
  {
    std::vector<uint32_t> used;
    for (uint32_t l = 0; l < 16; ++l)
      if (label_freq[l]) used.push_back(l);
    p->hist_compact = used.size() <= 8 ? 1 : 0;
    p->hist_map = 0;
    if (p->hist_compact)
      for (uint32_t i = 0; i < 8; ++i)
        p->hist_map |= (i < used.size() ? used[i] : (used.empty() ? 0 : used[0])) << (4 * i);
  }
  bms2_host_csr hv{};  // host view of the query batch
  hv.num_graphs = Q;
  hv.total_nodes = Qn;
  hv.total_edges = qci.size();
  hv.graph_offsets = qgo.data();
  hv.row_offsets = qro.data();
  hv.column_indices = qci.data();
  hv.node_labels = qlab.data();
  // automorphism orbits: a node automorphic to an earlier one has its domain in every data
  // graph, so the screen tests one node per orbit of a small query; the large-query join breaks
  // the symmetry of one orbit (upload_order)
  std::vector<uint8_t> orbit(std::max<uint64_t>(Qn, 1), 0), qorb(std::max<uint64_t>(Qn, 1), 0);
  for (uint32_t g = 0; g < Q; ++g) {
    const uint32_t b = qgo[g], n = qgo[g + 1] - qgo[g];
    query_orbits(&hv, g, orbit.data() + b);
    for (uint32_t u = 0; u < n; ++u) qorb[b + u] = n <= uint32_t(kSmallNq) ? orbit[b + u] : uint8_t(u);
  }
  std::vector<uint32_t> qhead(size_t(Q) * 4, 0), qpos(std::max<uint64_t>(Qn, 1), 0);
  for (uint32_t g = 0; g < Q; ++g) {
    const uint32_t b = qgo[g], n = qgo[g + 1] - qgo[g];
    qhead[4 * g + 0] = b;
    qhead[4 * g + 1] = n;
    qhead[4 * g + 2] = n > 1 ? 1 : 0;
    uint32_t ntest = 0;
    for (uint32_t u = 0; u < n; ++u) ntest += qorb[b + u] == u;
    qhead[4 * g + 3] = ntest;
    // screening order: one node per orbit first (the others are not tested), then rarest label
    // first, then highest degree (affects speed only)
    std::vector<uint32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t c) {
      const bool ca = qorb[b + a] != a, cc = qorb[b + c] != c;
      if (ca != cc) return cc;
      const uint32_t fa = qlab[b + a] == 255 ? ~0u : label_freq[qlab[b + a]];
      const uint32_t fc = qlab[b + c] == 255 ? ~0u : label_freq[qlab[b + c]];
      if (fa != fc) return fa < fc;
      return qro[b + a + 1] - qro[b + a] > qro[b + c + 1] - qro[b + c];
    });
    for (uint32_t t = 0; t < n; ++t) qpos[b + ord[t]] = b + t;
  }
  std::vector<uint32_t> qpacks = build_packs(qgo);
  p->num_qpacks = uint32_t(qpacks.size() - 1);
  if ((rc = device_alloc((void**)&p->d_qpacks, 4ull * qpacks.size())) ||
      (rc = device_alloc((void**)&p->d_qhead, 4ull * std::max<size_t>(qhead.size(), 4))) ||
      (rc = device_alloc((void**)&p->d_qpre, 4ull * kPreDwords * std::max<uint32_t>(Q, 1))) ||
      (rc = device_alloc((void**)&p->d_qrec, 4ull * kRecDwords * std::max<uint64_t>(Qn, 1))) ||
      (rc = device_alloc((void**)&p->d_qpos, 4ull * qpos.size())) ||
      (rc = device_alloc((void**)&p->d_jorder, std::max<size_t>(size_t(Q) * 32, 32))) ||
      (rc = device_alloc((void**)&p->d_qadj, 8ull * std::max<size_t>(size_t(Q) * 32, 32))) ||
      (rc = device_alloc((void**)&p->d_qadj8, 8ull * std::max<uint32_t>(Q, 1))) ||
      (rc = device_alloc((void**)&p->d_qorb, qorb.size())))
    return rc;
  BMS2_HIP(hipMemcpyAsync(p->d_qorb, qorb.data(), qorb.size(), hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemcpyAsync(p->d_qpacks, qpacks.data(), 4ull * qpacks.size(), hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemcpyAsync(p->d_qpos, qpos.data(), 4ull * qpos.size(), hipMemcpyHostToDevice, s));
  if (Q) BMS2_HIP(hipMemcpyAsync(p->d_qhead, qhead.data(), 4ull * qhead.size(), hipMemcpyHostToDevice, s));
  // isomorphism classes and the tables copy_classes_kernel reads (ClassCopy)
  std::vector<uint32_t> qrep(std::max<uint32_t>(Q, 1)), cu32;
  std::vector<uint2> cpairs;
  p->num_dups = query_classes(&hv, qrep.data());
  std::vector<uint32_t> reps, chunk, words, dpos;
  std::vector<uint2> dups;
  if (p->num_dups) {
    std::vector<uint32_t> rindex(Q, ~0u), wpos(Q, 0);
    for (uint32_t g = 0; g < Q; ++g)
      if (qrep[g] != g && rindex[qrep[g]] == ~0u) rindex[qrep[g]] = 0;
    for (uint32_t g = 0; g < Q; ++g)
      if (rindex[g] == 0) {
        rindex[g] = uint32_t(reps.size());
        reps.push_back(g);
      }
    for (uint32_t g = 0; g < Q; ++g) {
      if (qrep[g] == g) continue;
      dups.push_back(make_uint2(g, rindex[qrep[g]]));
      if (words.empty() || words.back() != g / 64) words.push_back(g / 64);
      wpos[g] = uint32_t(words.size() - 1);
    }
    std::stable_sort(dups.begin(), dups.end(), [](uint2 a, uint2 b) { return a.y < b.y; });
    for (const uint2& d : dups) dpos.push_back(wpos[d.x]);
    // copy_classes_kernel keeps the copied words of 4 graphs in each of 4 waves in 64 KB of LDS
    if (words.size() > 512) p->num_dups = 0;
  }
  if (p->num_dups) {
    const uint32_t nchunks = uint32_t((reps.size() + 63) / 64);
    for (uint32_t c = 0, i = 0; c <= nchunks; ++c) {
      while (i < dups.size() && dups[i].y < 64 * c) ++i;
      chunk.push_back(i);
    }
    const size_t o_chunk = reps.size(), o_words = o_chunk + chunk.size(),
                 o_dpos = o_words + words.size();
    cu32 = reps;
    cu32.insert(cu32.end(), chunk.begin(), chunk.end());
    cu32.insert(cu32.end(), words.begin(), words.end());
    cu32.insert(cu32.end(), dpos.begin(), dpos.end());
    cpairs = dups;
    if ((rc = device_alloc((void**)&p->d_qrep, 4ull * qrep.size())) ||
        (rc = device_alloc((void**)&p->d_copy_u32, 4ull * cu32.size())) ||
        (rc = device_alloc((void**)&p->d_copy_pairs, sizeof(uint2) * cpairs.size())))
      return rc;
    BMS2_HIP(hipMemcpyAsync(p->d_qrep, qrep.data(), 4ull * qrep.size(), hipMemcpyHostToDevice, s));
    BMS2_HIP(hipMemcpyAsync(p->d_copy_u32, cu32.data(), 4ull * cu32.size(), hipMemcpyHostToDevice, s));
    BMS2_HIP(hipMemcpyAsync(p->d_copy_pairs, cpairs.data(), sizeof(uint2) * cpairs.size(),
                            hipMemcpyHostToDevice, s));
    p->copy.reps = p->d_copy_u32;
    p->copy.chunk = p->d_copy_u32 + o_chunk;
    p->copy.words = p->d_copy_u32 + o_words;
    p->copy.dpos = p->d_copy_u32 + o_dpos;
    p->copy.dups = p->d_copy_pairs;
    p->copy.nr = uint32_t(reps.size());
    p->copy.nw = uint32_t(words.size());
  }
  p->qgo = std::move(qgo);
  p->qro = std::move(qro);
  p->qci = std::move(qci);
  p->qlab = std::move(qlab);
  p->orbit = std::move(orbit);
  p->qhv = bms2_host_csr{};
  p->qhv.num_graphs = Q;
  p->qhv.total_nodes = Qn;
  p->qhv.total_edges = p->qci.size();
  p->qhv.graph_offsets = p->qgo.data();
  p->qhv.row_offsets = p->qro.data();
  p->qhv.column_indices = p->qci.data();
  p->qhv.node_labels = p->qlab.data();
  if ((rc = upload_order(p, nullptr))) return rc;  // synchronises: the host vectors above are done
  // pair records of both classes and the small-record segments: worst case or a share of the
  // budget, whichever is smaller (the chunks are sized to what a run measures); buffers large
  // enough from an earlier batch are kept
  uint32_t q_small = 0;
  for (uint32_t g = 0; g < Q; ++g) q_small += (p->qgo[g + 1] - p->qgo[g]) <= uint32_t(kSmallNq);
  const uint64_t budget = record_budget_bytes(8 * kSmallWords * p->alloc_s +
                                              sizeof(uint4) * p->alloc_seg + 8ull * p->alloc_l_words);
  // large queries hold few GMCR pairs (a node of every large query needs a candidate)
  const double share_l = q_small == Q ? 0.0 : (q_small == 0 ? 0.9 : 0.15);
  const uint64_t D = p->data.num_graphs;
  const uint64_t worst_s = D * q_small, worst_l = D * (Q - q_small);
  // segment descriptors hold 32-bit record indices
  p->cap_s = std::max<uint64_t>(1, std::min<uint64_t>({worst_s, uint64_t(0.95 * (1 - share_l) * budget) / (8 * kSmallWords), (1ull << 32) - 1}));
  p->cap_seg = std::max<uint64_t>(1, std::min<uint64_t>({p->cap_s, uint64_t(0.05 * (1 - share_l) * budget) / 16,
                                                          (uint64_t(1) << (64 - kSegShift)) - 1}));
  p->cap_l = std::max<uint64_t>(1, std::min<uint64_t>(worst_l, uint64_t(share_l * budget) / (8ull * p->rec_words_l)));
  uint64_t l_words = p->alloc_l_words;
  if ((rc = ensure_buffer((void**)&p->d_recs_s, &p->alloc_s, p->cap_s, 8 * kSmallWords)) ||
      (rc = ensure_buffer((void**)&p->d_segs, &p->alloc_seg, p->cap_seg, sizeof(uint4))) ||
      (rc = ensure_buffer((void**)&p->d_recs_l, &l_words, p->cap_l * p->rec_words_l, 8)))
    return rc;
  p->alloc_l_words = l_words;
  if (p->num_big) {
    const uint64_t dn = p->big.total_nodes;
    const uint64_t K = std::max<uint32_t>(p->opt.refinement_steps, 1);
    if ((rc = bms2_path_signature_create(p->side_q, dn, Qn, &p->big_path)) ||
        (rc = bms2_cycle_signature_create(p->side_q, dn, Qn, &p->big_cyc)) ||
        (rc = device_alloc((void**)&p->d_big_qlev, 8 * K * std::max<uint64_t>(Qn, 1))) ||
        (rc = device_alloc((void**)&p->d_big_qrec, 4ull * kBigRecDwords * std::max<uint64_t>(Qn, 1))))
      return rc;
  }
  p->large_waves = p->opt.join_waves ? p->opt.join_waves : (p->nq_cap <= 16 ? 12 : 6);
  BMS2_HIP(hipStreamSynchronize(s));
  return BMS2_OK;
}

}  // namespace

extern "C" int bms2_pipeline_create(bms2_queue q, const bms2_device_csr* query,
                                    const bms2_device_csr* data, const bms2_pipeline_options* opt,
                                    bms2_pipeline* out) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !opt || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (opt->refinement_steps > uint32_t(kMaxLevels))
    return set_error(BMS2_ERR_UNSUPPORTED, "refinement steps %u > %d", opt->refinement_steps,
                     kMaxLevels);
  if (data->max_graph_nodes > BMS2_MAX_GRAPH_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "data graph of %u nodes (limit %u)",
                     data->max_graph_nodes, BMS2_MAX_GRAPH_NODES);
  if (data->max_degree > 27)
    return set_error(BMS2_ERR_UNSUPPORTED,
                     "node degree above 27: the reference's 128-entry ring-count stack could "
                     "overflow (signature.hpp:638,697), which the fused ring counts do not model");
  if (int rc = check_queries(query)) return rc;
  auto* p = new bms2_pipeline_s{};
  p->q = q;
  p->data = *data;
  p->opt = *opt;
  hipStream_t s = q->stream;
  int rc = BMS2_OK;
  auto fail = [&](int code) {
    pipeline_free(p);
    return code;
  };
  // data packs: the plan of one resident batch
  std::vector<uint32_t> go(data->num_graphs + 1);
  if (hipMemcpy(go.data(), data->graph_offsets, 4ull * go.size(), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return fail(set_error(BMS2_ERR_HIP, "copy of data graph offsets failed"));
  DataPacks dp = build_data_packs(go);
  p->num_packs = uint32_t(dp.begins.size() - 1);
  p->win_pack = std::move(dp.win_pack);
  p->win_graph = std::move(dp.win_graph);
  if ((rc = build_big_batch(p, go))) return fail(rc);
  if ((rc = device_alloc((void**)&p->d_packs, 4ull * dp.begins.size())) ||
      (rc = device_alloc((void**)&p->d_pack_gids, 4ull * std::max<size_t>(dp.gids.size(), 1))) ||
      (rc = device_alloc((void**)&p->d_adj, 8ull * 64 * std::max<uint32_t>(p->num_packs, 1))) ||
      (rc = device_alloc((void**)&p->d_ctl, 8 * kCtlWords)) ||
      (rc = device_alloc((void**)&p->d_dup_stats, 8 * 4)))
    return fail(rc);
  BMS2_HIP(hipMemcpyAsync(p->d_packs, dp.begins.data(), 4ull * dp.begins.size(),
                          hipMemcpyHostToDevice, s));
  if (!dp.gids.empty())
    BMS2_HIP(hipMemcpyAsync(p->d_pack_gids, dp.gids.data(), 4ull * dp.gids.size(),
                            hipMemcpyHostToDevice, s));
  BMS2_HIP(hipStreamSynchronize(s));  // the plan's host vectors end with this function
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  p->cus = cus;
  // defaults: as many waves as LDS allows (small join with 16 slots ~20 KB per wave: 8 per CU)
  p->small_waves = opt->block_waves ? opt->block_waves : 8;
  p->slots = 16;
  if ((rc = bind_queries(p, query))) return fail(rc);
  *out = p;
  return BMS2_OK;
}

extern "C" int bms2_pipeline_set_queries(bms2_pipeline p, const bms2_device_csr* query) {
  if (!p || !query) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = use_device(p->q)) return rc;
  if (int rc = check_queries(query)) return rc;
  BMS2_HIP(hipStreamSynchronize(p->q->stream));
  free_query_side(p);
  return bind_queries(p, query);
}

extern "C" int bms2_pipeline_destroy(bms2_pipeline p) {
  if (!p) return BMS2_OK;
  if (int rc = use_device(p->q)) return rc;
  (void)hipStreamSynchronize(p->q->stream);
  pipeline_free(p);
  return BMS2_OK;
}

namespace {

struct Trace {
  uint64_t* domsum = nullptr;
  uint64_t* qpairs = nullptr;
  uint64_t* qnodes = nullptr;
  uint32_t* pair_nodes = nullptr;
};

// Screens and joins packs [0, pack_end) in chunks that fit the record buffer; stats (host)
// accumulates {GMCR pairs, matched pairs, embeddings, DFS nodes}.
int run_packs(bms2_pipeline p, uint32_t pack_end, uint32_t g_limit, uint64_t* d_counts,
              uint64_t* d_valid, const Trace& tr, EventScope& ks, uint64_t stats[8]) {
  hipStream_t s = p->q->stream;
  const uint32_t Q = p->query.num_graphs;
  if (pack_end == 0 || Q == 0) {
    if (int rc = ks.begin()) return rc;
    return ks.end();
  }
  ScreenParams sp{};
  sp.graph_offsets = p->data.graph_offsets;
  sp.row_offsets = p->data.row_offsets;
  sp.column_indices = p->data.column_indices;
  sp.labels = p->data.node_labels;
  sp.pack_gbegin = p->d_packs;
  sp.pack_gids = p->d_pack_gids;
  sp.Q = Q;
  sp.qhead = p->d_qhead;
  sp.qpre = p->d_qpre;
  sp.qrec = p->d_qrec;
  sp.jorder = p->d_jorder;
  sp.qadj8 = p->d_qadj8;
  sp.adj_out = p->d_adj;
  sp.counts = d_counts;
  sp.valid = d_valid;
  sp.VW = (Q + 63) / 64;
  sp.recs_s = p->d_recs_s;
  sp.cap_s = p->cap_s;
  sp.segs = p->d_segs;
  sp.cap_seg = p->cap_seg;
  sp.recs_l = p->d_recs_l;
  sp.cap_l = p->cap_l;
  sp.rec_words_l = p->rec_words_l;
  sp.ctl = p->d_ctl;
  sp.domsum = tr.domsum;
  sp.qpairs = tr.qpairs;
  JoinParams jp{};
  jp.recs = p->d_recs_l;
  jp.rec_cap = p->cap_l;
  jp.rec_words = p->rec_words_l;
  jp.Q = Q;
  jp.adj = p->d_adj;
  jp.qadj = p->d_qadj;
  jp.counts = d_counts;
  jp.ctl = p->d_ctl;
  jp.find_first = p->opt.find_first ? 1 : 0;
  jp.qnodes = tr.qnodes;
  jp.pair_nodes = tr.pair_nodes;
  JoinSmallParams js{};
  js.recs = p->d_recs_s;
  js.segs = p->d_segs;
  js.cap_seg = p->cap_seg;
  js.Q = Q;
  js.adj = p->d_adj;
  js.counts = d_counts;
  js.ctl = p->d_ctl;
  js.find_first = jp.find_first;
  js.qnodes = tr.qnodes;
  js.pair_nodes = tr.pair_nodes;
  QueryPrepParams qp{p->query.graph_offsets, p->query.row_offsets, p->query.column_indices,
                     p->query.node_labels, p->d_qpacks, p->d_qpos, p->d_qrec, p->hist_map,
                     p->hist_compact, p->d_jorder};
  // a run screens one query per isomorphism class; traces screen every query (their per-query
  // accumulators are the planner's input)
  const bool dedup = p->dedup && p->num_dups && !tr.domsum && !tr.qnodes && !tr.pair_nodes;
  // and the large-query join breaks the symmetry of automorphic query nodes (traces count every
  // DFS node of the plain search)
  // (count-all mode: in first-match mode the first match can come later under the constraint)
  jp.sym = p->dedup && !p->opt.find_first && !tr.domsum && !tr.qnodes && !tr.pair_nodes ? 1u : 0u;
  js.sym = jp.sym;
  jp.split = !p->opt.find_first && !tr.domsum && !tr.qnodes && !tr.pair_nodes ? 1u : 0u;
  sp.hist_map = p->hist_map;
  const size_t stage_bytes = std::max<size_t>(
      (size_t(kStageSmall) * kSmallWords + size_t(kStageLarge) * p->rec_words_l) * 8,
      size_t(kScreenScratchBytes));
  return dispatch_levels(p->opt.refinement_steps, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    hipLaunchKernelGGL(query_prep_kernel<K>, dim3(p->num_qpacks), dim3(64), 0, s, qp);
    BMS2_CHECK_LAUNCH("query_prep_kernel");
    hipLaunchKernelGGL(query_pre_kernel, dim3((Q + 255) / 256), dim3(256), 0, s, p->d_qhead,
                       p->d_qrec, Q, dedup ? p->d_qrep : nullptr, p->d_jorder, p->d_qorb,
                       p->d_qpre);
    BMS2_CHECK_LAUNCH("query_pre_kernel");
    if (p->num_big) BMS2_HIP(hipEventRecord(p->big_fork, s));  // the big-graph path's inputs
    if (int krc = ks.begin()) return krc;
    // kernels of one chunk of packs [p0, p1): the screen, then the joins
    auto launch_screen = [&](uint32_t p0, uint32_t p1) -> int {
      hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, p->d_ctl, uint32_t(kCtlWords));
      sp.pack_begin = p0;
      if (p->hist_compact)
        hipLaunchKernelGGL((screen_kernel<K, 2>), dim3(p1 - p0), dim3(64), stage_bytes, s, sp);
      else
        hipLaunchKernelGGL((screen_kernel<K, 4>), dim3(p1 - p0), dim3(64), stage_bytes, s, sp);
      BMS2_CHECK_LAUNCH("screen_kernel");
      return BMS2_OK;
    };
    auto launch_joins = [&]() -> int {
      const dim3 gs(uint32_t(p->cus) * p->small_waves), gl(uint32_t(p->cus) * p->large_waves);
      if (p->slots == 8)
        hipLaunchKernelGGL(join_small_kernel<8>, gs, dim3(64), 0, s, js);
      else
        hipLaunchKernelGGL(join_small_kernel<16>, gs, dim3(64), 0, s, js);
      BMS2_CHECK_LAUNCH("join_small_kernel");
      if (p->max_nq > uint32_t(kSmallNq)) {
        if (p->nq_cap <= 16)
          hipLaunchKernelGGL(join_kernel<16>, gl, dim3(64), 0, s, jp);
        else
          hipLaunchKernelGGL(join_kernel<32>, gl, dim3(64), 0, s, jp);
        BMS2_CHECK_LAUNCH("join_kernel");
      }
      return BMS2_OK;
    };
    auto capture = [&](auto&& launch, hipGraphExec_t* out) -> int {
      hipGraph_t g = nullptr;
      BMS2_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int lrc = launch();
      const hipError_t erc = hipStreamEndCapture(s, &g);
      if (lrc) return lrc;
      BMS2_HIP(erc);
      const hipError_t irc = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      BMS2_HIP(irc);
      return BMS2_OK;
    };
    const bool tracing = tr.domsum || tr.qnodes || tr.pair_nodes;
    // a trace of domains alone runs the screens only
    const bool screen_only = tr.domsum && !tr.qnodes && !tr.pair_nodes;
    bool big_started = screen_only;
    uint32_t p0 = 0;
    while (p0 < pack_end) {
      // chunk of packs sized so the expected records of both classes and the segments fit
      uint64_t chunk = pack_end - p0;
      auto fit = [&](double dens, uint64_t cap) {
        if (dens > 0) chunk = std::min<uint64_t>(chunk, std::max<uint64_t>(1, uint64_t(0.9 * cap / dens)));
      };
      fit(p->dens_s, p->cap_s);
      fit(p->dens_seg, p->cap_seg);
      fit(p->dens_l, p->cap_l);
      const uint32_t p1 = p0 + uint32_t(chunk);
      // a measured batch that fits one chunk replays its captured kernels
      const bool graphable = !p->opt.no_graph && !tracing && p0 == 0 && p1 == pack_end &&
                             p->dens_s + p->dens_l > 0;
      // the graphs above 64 nodes are screened beside the first chunk's screen and joined beside
      // its joins (their rows of the outputs are disjoint from the chunk's)
      const bool fork = !big_started && p->num_big;
      if (graphable && !(p->graph && p->graph_counts == d_counts && p->graph_valid == d_valid &&
                         p->graph_packs == pack_end)) {
        drop_graphs(p);
        if (p->num_big) {
          if (int crc = capture([&] { return launch_screen(p0, p1); }, &p->graph)) return crc;
          if (int crc = capture([&] { return launch_joins(); }, &p->graph_joins)) return crc;
        } else if (int crc = capture([&] {
                     const int rc = launch_screen(p0, p1);
                     return rc ? rc : launch_joins();
                   }, &p->graph)) {
          return crc;
        }
        p->graph_counts = d_counts;
        p->graph_valid = d_valid;
        p->graph_packs = pack_end;
      }
      if (graphable) {
        BMS2_HIP(hipGraphLaunch(p->graph, s));
        if (fork) BMS2_HIP(hipEventRecord(p->big_screened, s));
        if (p->graph_joins) BMS2_HIP(hipGraphLaunch(p->graph_joins, s));
      } else {
        if (int lrc = launch_screen(p0, p1)) return lrc;
        if (fork) BMS2_HIP(hipEventRecord(p->big_screened, s));
        if (!screen_only)
          if (int lrc = launch_joins()) return lrc;
      }
      if (!big_started) {
        big_started = true;
        if (int brc = big_start(p, d_counts, d_valid, g_limit)) return brc;
      }
      uint64_t ctl[kCtlWords];
      BMS2_HIP(hipMemcpyAsync(ctl, p->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
      BMS2_HIP(hipStreamSynchronize(s));
      const double np = double(p1 - p0);
      const double ds = (ctl[kCtlSmall] & low_mask(kSegShift)) / np;
      const double dg = (ctl[kCtlSmall] >> kSegShift) / np, dl = ctl[kCtlLarge] / np;
      if (ctl[kCtlOverflow]) {
        // a buffer overflowed: measure, shrink the chunk, redo it (results are rewritten)
        if (chunk == 1)
          return set_error(BMS2_ERR_CAPACITY, "pair record buffers (%llu + %llu records) too small",
                           (unsigned long long)p->cap_s, (unsigned long long)p->cap_l);
        if (tr.domsum || tr.qnodes)
          return set_error(BMS2_ERR_CAPACITY,
                           "pair record buffer overflowed during a trace; trace fewer graphs");
        p->dens_s = std::max(p->dens_s * 2.0, ds);
        p->dens_seg = std::max(p->dens_seg * 2.0, dg);
        p->dens_l = std::max(p->dens_l * 2.0, dl);
        if (graphable) drop_graphs(p);
        continue;
      }
      p->dens_s = std::max(p->dens_s, ds);
      p->dens_seg = std::max(p->dens_seg, dg);
      p->dens_l = std::max(p->dens_l, dl);
      for (int k = 0; k < 4; ++k) stats[k] += ctl[kCtlPairs + k];
      p0 = p1;
    }
    if (!big_started) {
      if (p->num_big) BMS2_HIP(hipEventRecord(p->big_screened, s));
      if (int brc = big_start(p, d_counts, d_valid, g_limit)) return brc;
    }
    if (int brc = big_finish(p, stats)) return brc;
    if (dedup) {
      // every screen and join of the run is done on this stream: copy the classes' outputs
      const uint64_t D = std::min<uint64_t>(g_limit, p->data.num_graphs);
      hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, p->d_dup_stats, 4u);
      if (D) {
        const uint32_t blocks = uint32_t(std::min<uint64_t>((D + 15) / 16, uint64_t(p->cus) * 16));
        hipLaunchKernelGGL(copy_classes_kernel, dim3(blocks), dim3(256), 4 * 4 * 8 * p->copy.nw, s,
                           d_counts, d_valid, Q, (Q + 63) / 64, D, p->copy, p->d_dup_stats);
      }
      BMS2_CHECK_LAUNCH("copy_classes_kernel");
      uint64_t dst[4];
      BMS2_HIP(hipMemcpyAsync(dst, p->d_dup_stats, sizeof(dst), hipMemcpyDeviceToHost, s));
      BMS2_HIP(hipStreamSynchronize(s));
      for (int k = 0; k < 3; ++k) stats[k] += dst[k];
    }
    return ks.end();
  });
}

}  // namespace

extern "C" int bms2_pipeline_run(bms2_pipeline p, uint64_t* d_counts, uint64_t* d_valid,
                                 uint64_t* d_stats, bms2_event* ev, bms2_event* kernel_ev) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (int rc = use_device(p->q)) return rc;
  if (!d_counts || !d_valid) return set_error(BMS2_ERR_INVALID, "null output buffer");
  hipStream_t s = p->q->stream;
  EventScope es{p->q, ev};
  if (int rc = es.begin()) return rc;
  EventScope ks{p->q, kernel_ev};
  uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (int rc = run_packs(p, p->num_packs, p->data.num_graphs, d_counts, d_valid, Trace{}, ks, stats))
    return rc;
  if (d_stats) BMS2_HIP(hipMemcpyAsync(d_stats, stats, sizeof(stats), hipMemcpyHostToDevice, s));
  if (d_stats) BMS2_HIP(hipStreamSynchronize(s));
  return es.end();
}

extern "C" int bms2_pipeline_set_dedup(bms2_pipeline p, int on, uint32_t* num_dups) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (p->dedup != (on != 0)) {
    // the captured joins hold the symmetry-breaking switch: capture them again
    if (int rc = use_device(p->q)) return rc;
    BMS2_HIP(hipStreamSynchronize(p->q->stream));
    drop_graphs(p);
  }
  p->dedup = on != 0;
  if (num_dups) *num_dups = p->num_dups;
  return BMS2_OK;
}

extern "C" int bms2_pipeline_set_order(bms2_pipeline p, const uint8_t* order) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (int rc = use_device(p->q)) return rc;
  BMS2_HIP(hipStreamSynchronize(p->q->stream));  // a running join still reads the old order
  return upload_order(p, order);
}

extern "C" int bms2_pipeline_trace(bms2_pipeline p, uint32_t max_graphs, uint64_t* d_counts,
                                   uint64_t* d_valid, uint64_t* d_domsum, uint64_t* d_qpairs,
                                   uint64_t* d_qnodes, uint32_t* d_pair_nodes,
                                   uint32_t* graphs_out) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (int rc = use_device(p->q)) return rc;
  if (!d_counts || !d_valid || !d_domsum || !d_qpairs)
    return set_error(BMS2_ERR_INVALID, "null trace buffer");
  // leading pack windows covering at least max_graphs graphs (all of them when max_graphs is)
  size_t w = 0;
  while (w + 1 < p->win_graph.size() && p->win_graph[w] < max_graphs) ++w;
  const uint32_t pe = p->win_pack[w], ge = p->win_graph[w];
  if (graphs_out) *graphs_out = ge;
  Trace tr{d_domsum, d_qpairs, d_qnodes, d_pair_nodes};
  EventScope ks{p->q, nullptr};
  uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (int rc = run_packs(p, pe, ge, d_counts, d_valid, tr, ks, stats)) return rc;
  BMS2_HIP(hipStreamSynchronize(p->q->stream));
  return BMS2_OK;
}

extern "C" int bms2_pipeline_autotune(bms2_pipeline p, uint64_t* d_counts, uint64_t* d_valid,
                                      uint32_t runs, bms2_pipeline_tuning* out) {
  if (!p) return set_error(BMS2_ERR_INVALID, "pipeline is null");
  if (int rc = use_device(p->q)) return rc;
  if (!d_counts || !d_valid) return set_error(BMS2_ERR_INVALID, "null output buffer");
  runs = std::max<uint32_t>(runs, 1);
  struct Cand {
    uint32_t small_waves, large_waves, slots;
    BigJoinShape big;
  };
  std::vector<Cand> cands;
  const uint32_t lw0 = p->nq_cap <= 16 ? 12 : 6;
  const BigJoinShape b0 = p->big_shape;
  // LDS fits 8 small-join waves per CU with 16 slots, 10 with 8
  for (uint32_t sw : {7u, 8u, 10u})
    for (uint32_t sl : {8u, 16u}) cands.push_back({sw, lw0, sl, b0});
  for (uint32_t lw : {lw0 / 2, lw0 * 3 / 4}) cands.push_back({8, lw, 16, b0});
  // graphs above 64 nodes: waves per CU, waves per workgroup and adjacency-row words of the
  // size-class joins
  if (p->num_big)
    for (uint32_t bw : {6u, 12u, 16u})
      for (uint32_t wg : {1u, 4u})
        for (uint32_t rw : {2u, 4u}) {
          BigJoinShape b;
          b.waves = bw;
          b.block_waves = wg;
          b.row_words = rw;
          cands.push_back({8, lw0, 16, b});
        }
  auto same_join_shape = [&](const Cand& c) {
    return c.small_waves == p->small_waves && c.large_waves == p->large_waves && c.slots == p->slots;
  };
  auto apply = [&](const Cand& c) {
    if (!same_join_shape(c)) drop_graphs(p);  // grid shapes are baked into the graphs
    p->small_waves = c.small_waves;
    p->large_waves = c.large_waves;
    p->slots = c.slots;
    p->big_shape = c.big;
  };
  const Cand keep{p->small_waves, p->large_waves, p->slots, p->big_shape};
  float best = 0.f;
  Cand best_c = keep;
  bms2_event ev = nullptr;
  for (size_t i = 0; i < cands.size(); ++i) {
    apply(cands[i]);
    // one untimed run (measures the batch, captures the graph), then the timed runs
    for (uint32_t r = 0; r <= runs; ++r) {
      if (int rc = bms2_pipeline_run(p, d_counts, d_valid, nullptr, nullptr, r ? &ev : nullptr)) {
        apply(keep);
        return rc;
      }
      if (r) {
        float ms = 0.f;
        if (int rc = bms2_event_elapsed_ms(ev, &ms)) return rc;
        bms2_event_destroy(ev);
        ev = nullptr;
        if (best == 0.f || ms < best) {
          best = ms;
          best_c = cands[i];
        }
      }
    }
  }
  apply(best_c);
  if (out) {
    out->small_waves = best_c.small_waves;
    out->large_waves = best_c.large_waves;
    out->slots = best_c.slots;
    out->candidates = uint32_t(cands.size());
    out->best_ms = best;
    out->big_waves = best_c.big.waves;
    out->big_block_waves = best_c.big.block_waves;
    out->big_row_words = best_c.big.row_words;
  }
  return BMS2_OK;
}

namespace bms2 {
namespace {

// Per query: data graphs with a non-zero count and the summed counts, over a range of graphs
// per block (coalesced along the query axis of the data-graph-major counts).
__global__ void query_totals_kernel(const uint64_t* __restrict__ counts, uint64_t D, uint32_t Q,
                                    uint32_t graphs_per_block, uint64_t* __restrict__ matched,
                                    uint64_t* __restrict__ emb) {
  const uint64_t g0 = uint64_t(blockIdx.x) * graphs_per_block;
  const uint64_t g1 = min(D, g0 + graphs_per_block);
  for (uint32_t q = threadIdx.x; q < Q; q += blockDim.x) {
    uint64_t m = 0, e = 0;
    for (uint64_t g = g0; g < g1; ++g) {
      const uint64_t c = counts[g * Q + q];
      m += c ? 1 : 0;
      e += c;
    }
    if (m) {
      atomicAdd(reinterpret_cast<unsigned long long*>(matched + q), (unsigned long long)m);
      atomicAdd(reinterpret_cast<unsigned long long*>(emb + q), (unsigned long long)e);
    }
  }
}

}  // namespace
}  // namespace bms2

extern "C" int bms2_pipeline_query_totals(bms2_pipeline p, const uint64_t* d_counts,
                                          uint64_t* d_matched, uint64_t* d_embeddings) {
  if (!p || !d_counts || !d_matched || !d_embeddings)
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = use_device(p->q)) return rc;
  const uint64_t D = p->data.num_graphs;
  const uint32_t Q = p->query.num_graphs;
  hipStream_t s = p->q->stream;
  if (Q) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, s, d_matched, Q);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, s, d_embeddings, Q);
  }
  if (D && Q) {
    const uint32_t per = 2048;
    hipLaunchKernelGGL(query_totals_kernel, dim3(uint32_t((D + per - 1) / per)), dim3(256), 0, s,
                       d_counts, D, Q, per, d_matched, d_embeddings);
  }
  BMS2_CHECK_LAUNCH("query_totals_kernel");
  return BMS2_OK;
}
