// Screen of the data graphs of 65-256 nodes inside the fused pipeline.  The staged sequence
// materialises the filter (isomorphism.hpp:59-97) and K refinement steps (isomorphism.hpp:142-196)
// as a query-node x data-node bit matrix, tests every pair of nodes, and reads the matrix back to
// build the GMCR (gmcr.hpp:54-127).  Here one wave takes one data graph and screens every query
// against it: the nodes' filter operands and shell signatures of every level stay in VGPRs (up
// to four nodes per lane), a query passes a prefilter on the graph's summary first, and its nodes
// are then tested in the most-selective-first order with an exit at the first node left without a
// candidate.  The domains of the pairs that survive (the GMCR pairs) go straight to a record
// buffer the join reads; pairs are never materialised otherwise.  The domains are the staged
// ones bit for bit: the same predicate, evaluated only where it can change the GMCR.
#include <algorithm>
#include <vector>

#include "device_common.hpp"
#include "runtime.hpp"

namespace bms2 {

namespace {

constexpr uint64_t kTop16 = 0x8000800080008000ull;  // top bit of each 16-bit ring count
constexpr uint64_t kFifteens = 0x0f0f0f0f0f0f0f0full;
constexpr uint64_t kByteBit4 = 0x1010101010101010ull;

__device__ inline uint32_t cload(const uint32_t* p, uint32_t i) {
  return ((const __attribute__((address_space(4))) uint32_t*)(p))[i];
}

__device__ inline uint64_t u64of(uint32_t lo, uint32_t hi) { return uint64_t(lo) | (uint64_t(hi) << 32); }

// Shell signatures of every level in one breadth-first search per node: out[k n + u] (k < K)
// holds the saturating 4-bit counts of the labels < 16 of the nodes at distance exactly k + 1
// from u inside u's graph.  That is the value refineCSRSignatures<PowerGraph>
// (signature.hpp:347-385) leaves after its (k + 1)-th call when it is called with view sizes
// 1, 2, .., k + 1 on zeroed reachability state: each call removes the nodes reached by the
// earlier calls.  One thread per node; graphs of up to 256 nodes.
__global__ void shell_levels_kernel(const uint32_t* __restrict__ go, uint32_t num_graphs,
                                    const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                    const uint8_t* __restrict__ labels, uint64_t n, uint32_t K,
                                    uint64_t* __restrict__ out) {
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < n;
       u += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t gs = go[graph_of(go, num_graphs, uint32_t(u))];
    const uint32_t ul = uint32_t(u) - gs;
    uint64_t seen[4] = {0, 0, 0, 0}, front[4] = {0, 0, 0, 0};
    seen[ul >> 6] = front[ul >> 6] = uint64_t(1) << (ul & 63);
    for (uint32_t k = 0; k < K; ++k) {
      uint64_t next[4] = {0, 0, 0, 0};
      for (int w = 0; w < 4; ++w) {
        uint64_t f = front[w];
        while (f) {
          const uint32_t x = uint32_t(64 * w + ctz64(f));
          f &= f - 1;
          for (uint32_t e = ro[gs + x]; e < ro[gs + x + 1]; ++e) {
            const uint32_t y = ci[e] - gs;
            const uint64_t bit = uint64_t(1) << (y & 63);
            if (!(seen[y >> 6] & bit)) {
              seen[y >> 6] |= bit;
              next[y >> 6] |= bit;
            }
          }
        }
      }
      uint64_t s = 0;
      for (int w = 0; w < 4; ++w) {
        uint64_t f = next[w];
        while (f) {
          s = inc_label(s, labels[gs + 64 * w + ctz64(f)]);
          f &= f - 1;
        }
        front[w] = next[w];
      }
      out[uint64_t(k) * n + u] = s;
    }
  }
}

// Screening record of one query node (u32 words): [0] label, [1] label mask (0xff, 0 for the
// wildcard: any data label), [2] degree, [3] local id, [4..11] path signature, [12..13] ring
// counts, [14] the node's depth in the matching order (jorder, identity when null), [15] 0,
// [16 + 4 k .. 19 + 4 k] shell level k < K: the even and the odd 4-bit counts spread one per
// byte (two u64).  Records sit at the node's slot of the screening order.
__global__ void big_query_rec_kernel(const uint32_t* __restrict__ qgo, uint32_t Q,
                                     const uint32_t* __restrict__ qro,
                                     const uint8_t* __restrict__ qlab, uint64_t Qn,
                                     const uint64_t* __restrict__ qpath,
                                     const uint64_t* __restrict__ qcyc,
                                     const uint64_t* __restrict__ qlev, uint32_t K,
                                     const uint32_t* __restrict__ qpos,
                                     const uint8_t* __restrict__ jorder, uint32_t* __restrict__ out) {
  const uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (u >= Qn) return;
  uint32_t* r = out + uint64_t(qpos[u]) * kBigRecDwords;
  const uint32_t lab = qlab[u];
  const uint32_t g = graph_of(qgo, Q, uint32_t(u));
  const uint32_t local = uint32_t(u) - qgo[g], nq = qgo[g + 1] - qgo[g];
  uint32_t depth = local;
  if (jorder)
    for (uint32_t i = 0; i < nq; ++i)
      if (jorder[size_t(g) * 32 + i] == local) depth = i;
  r[0] = lab;
  r[1] = lab == kWildcard ? 0u : 0xffu;
  r[2] = qro[u + 1] - qro[u];
  r[3] = local;
  for (int w = 0; w < 4; ++w) {
    const uint64_t b = qpath[4 * u + w];
    r[4 + 2 * w] = uint32_t(b);
    r[5 + 2 * w] = uint32_t(b >> 32);
  }
  const uint64_t c = qcyc[u];
  r[12] = uint32_t(c);
  r[13] = uint32_t(c >> 32);
  r[14] = depth;
  r[15] = 0;
  for (uint32_t k = 0; k < uint32_t(kBigMaxLevels); ++k) {
    const uint64_t v = k < K ? qlev[uint64_t(k) * Qn + u] : 0;
    const uint64_t e = v & kFifteens, o = (v >> 4) & kFifteens;
    r[16 + 4 * k] = uint32_t(e);
    r[17 + 4 * k] = uint32_t(e >> 32);
    r[18 + 4 * k] = uint32_t(o);
    r[19 + 4 * k] = uint32_t(o >> 32);
  }
}

struct BigScreenParams {
  const uint32_t* dgo;
  const uint32_t* dro;
  const uint8_t* dlab;
  const uint64_t* dpath;   // [dn][4]
  const uint64_t* dcyc;    // [dn]
  const uint64_t* dlev;    // [K][dn]
  uint64_t dn;
  const uint32_t* glist;   // graphs of this launch (batch indices)
  uint32_t wide_above;     // graphs above this many nodes take 4-word domains
  uint32_t Q;
  const uint32_t* qpre;    // [Q][pre_dwords] prefilter records
  uint32_t pre_dwords;
  const uint32_t* bqrec;   // [Qn][kBigRecDwords] in screening order
  BigEntries e;
  const uint32_t* gid;     // original id of each graph (its row of valid)
  uint64_t* valid;         // [graphs][VW] GMCR bit rows
  uint32_t VW;
};

constexpr uint32_t kRunBuf = 256;  // GMCR pairs of a graph gathered in LDS per run

// One wave per data graph (NWS = 64-node slots per lane: 2 up to 128 nodes, 4 above).  A block of
// 64 queries reserves record words for every candidate that passed the prefilter, so a pair's
// domains are written as its nodes pass, and stores its bits of the graph's GMCR row at once.
// The graph's GMCR pairs gather in LDS and leave as runs of up to kRunBuf entries: one atomic
// takes the entries (class c: indices c cap_e + i), one the run {graph, count, first entry}
// (class c: c cap_r + i).  Counters run on past the capacities (the overflow flag is set), so a
// rerun can size them.
template <int K, int NWS>
__global__ __launch_bounds__(64) void big_screen_kernel(BigScreenParams p) {
  __shared__ uint32_t pair_q[kRunBuf];
  __shared__ uint64_t pair_off[kRunBuf];
  const uint32_t lane = lane_id();
  const uint32_t b = p.glist[blockIdx.x];
  const uint32_t s = p.dgo[b], n = p.dgo[b + 1] - s;
  const uint32_t cls = uint32_t(big_class(n, p.wide_above));
  const uint32_t RWc = cls ? 4u : 2u;
  constexpr int KL = K > 0 ? K : 1;
  uint32_t dl[NWS];
  int32_t dd[NWS];
  uint64_t nd[NWS][4], dch[NWS], le[NWS][KL], lo[NWS][KL];
  uint32_t pres[4] = {0, 0, 0, 0}, pu[8] = {0, 0, 0, 0, 0, 0, 0, 0}, rm[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < NWS; ++j) {
    const uint32_t x = lane + 64 * j;
    if (x < n) {
      const uint64_t u = s + x;
      dl[j] = p.dlab[u];
      dd[j] = int32_t(p.dro[u + 1] - p.dro[u]);
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint64_t pw = p.dpath[4 * u + w];
        nd[j][w] = ~pw;
        pu[2 * w] |= uint32_t(pw);
        pu[2 * w + 1] |= uint32_t(pw >> 32);
      }
      const uint64_t c = p.dcyc[u];
      dch[j] = c | kTop16;
#pragma unroll
      for (int f = 0; f < 4; ++f) rm[f] = max(rm[f], uint32_t(c >> (16 * f)) & 0xffffu);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t v = p.dlev[uint64_t(k) * p.dn + u];
        le[j][k] = kFifteens - (v & kFifteens);
        lo[j][k] = kFifteens - ((v >> 4) & kFifteens);
      }
      if (dl[j] < 16) {
        const uint32_t top = min(uint32_t(dd[j]), 7u);
        pres[dl[j] >> 2] |= ((2u << top) - 1) << ((dl[j] & 3) * 8);
      }
    } else {
      // a slot past the graph fails the degree test of every query node
      dl[j] = 0;
      dd[j] = -1;
#pragma unroll
      for (int w = 0; w < 4; ++w) nd[j][w] = 0;
      dch[j] = kTop16;
#pragma unroll
      for (int k = 0; k < K; ++k) le[j][k] = lo[j][k] = kFifteens;
    }
  }
  // graph summary for the prefilter: presence bits, union of path bits, largest ring counts
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int w = 0; w < 4; ++w) pres[w] |= __shfl_xor(pres[w], off);
#pragma unroll
    for (int w = 0; w < 8; ++w) pu[w] |= __shfl_xor(pu[w], off);
#pragma unroll
    for (int f = 0; f < 4; ++f) rm[f] = max(rm[f], uint32_t(__shfl_xor(rm[f], off)));
  }

  uint32_t np = 0;  // pairs gathered
  auto flush = [&]() {
    if (!np) return;
    __syncthreads();  // lane 0's pair list visible to every lane (one wave: no waiting)
    uint64_t eb = 0, ri = 0;
    if (lane == 0) {
      eb = atomicAdd(p.e.ctl + cls, (unsigned long long)np);
      ri = atomicAdd(p.e.ctl + kBigClasses + cls, 1ull);
    }
    eb = uniform64(readlane64(eb, 0));
    ri = uniform64(readlane64(ri, 0));
    if (eb + np <= p.e.cap_e && ri < p.e.cap_r) {
      const uint64_t first = cls * p.e.cap_e + eb;
      for (uint32_t i = lane; i < np; i += 64) {
        p.e.gidx[first + i] = pair_q[i];
        p.e.rec_off[first + i] = pair_off[i];
      }
      if (lane == 0)
        p.e.runs[cls * p.e.cap_r + ri] = make_uint4(b, np, uint32_t(first), uint32_t(first >> 32));
    } else if (lane == 0) {
      atomicOr(p.e.ctl + kBigCtlOverflow, 1ull);
    }
    __syncthreads();
    np = 0;
  };

  for (uint32_t qb = 0; qb < p.Q; qb += 64) {
    const uint32_t qq = qb + lane;
    bool pass = false;
    uint32_t q_rb = 0, q_nq = 0;
    if (qq < p.Q) {
      const uint4* pr = reinterpret_cast<const uint4*>(p.qpre + size_t(qq) * p.pre_dwords);
      const uint4 a = pr[0], pb = pr[1], pc = pr[2], rg = pr[3], e = pr[4];
      q_rb = e.y;
      q_nq = e.z;
      const uint32_t miss = (a.x & ~pres[0]) | (a.y & ~pres[1]) | (a.z & ~pres[2]) | (a.w & ~pres[3]) |
                            (pb.x & ~pu[0]) | (pb.y & ~pu[1]) | (pb.z & ~pu[2]) | (pb.w & ~pu[3]) |
                            (pc.x & ~pu[4]) | (pc.y & ~pu[5]) | (pc.z & ~pu[6]) | (pc.w & ~pu[7]);
      pass = e.x && miss == 0 && rg.x <= rm[0] && rg.y <= rm[1] && rg.z <= rm[2] && rg.w <= rm[3];
    }
    uint64_t cand = ballot(pass);
    uint64_t* const vrow = p.valid + uint64_t(p.gid[b]) * p.VW + (qb >> 6);
    if (!cand) {
      if (lane == 0) *vrow = 0;
      continue;
    }
    if (np + uint32_t(popc64(cand)) > kRunBuf) flush();
    // record words of every candidate of the block: a candidate that fails leaves a hole
    uint32_t need = pass ? q_nq * RWc : 0u;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) need += __shfl_xor(need, off);
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd(p.e.ctl + kBigCtlRecWords, (unsigned long long)need);
    base = uniform64(readlane64(base, 0));
    const bool rec_ok = base + need <= p.e.cap_words;
    uint64_t roff = base;
    uint64_t vbits = 0;
    while (cand) {
      const uint32_t li = uint32_t(ctz64(cand));
      cand &= cand - 1;
      const uint32_t qu = qb + li;
      const uint32_t rb = __builtin_amdgcn_readlane(q_rb, li);
      const uint32_t nq = __builtin_amdgcn_readlane(q_nq, li);
      const uint32_t* r = p.bqrec + size_t(rb) * kBigRecDwords;
      bool alive = true;
      for (uint32_t i = 0; i < nq; ++i, r += kBigRecDwords) {
        constexpr int A = 16 + 4 * K;
        uint32_t w[A];
#pragma unroll
        for (int x = 0; x < A; ++x) w[x] = cload(r, x);
        const uint64_t qp0 = u64of(w[4], w[5]), qp1 = u64of(w[6], w[7]);
        const uint64_t qp2 = u64of(w[8], w[9]), qp3 = u64of(w[10], w[11]);
        const uint64_t qc = u64of(w[12], w[13]);
        uint64_t m[NWS], any = 0;
#pragma unroll
        for (int j = 0; j < NWS; ++j) {
          uint64_t sh = 0;
#pragma unroll
          for (int k = 0; k < K; ++k)
            sh |= (le[j][k] + u64of(w[16 + 4 * k], w[17 + 4 * k])) |
                  (lo[j][k] + u64of(w[18 + 4 * k], w[19 + 4 * k]));
          const uint64_t miss64 = (qp0 & nd[j][0]) | (qp1 & nd[j][1]) | (qp2 & nd[j][2]) |
                                  (qp3 & nd[j][3]) | (~(dch[j] - qc) & kTop16) | (sh & kByteBit4);
          const uint32_t miss = uint32_t(miss64) | uint32_t(miss64 >> 32) | ((dl[j] ^ w[0]) & w[1]) |
                                (uint32_t(dd[j] - int32_t(w[2])) & 0x80000000u);
          m[j] = ballot(miss == 0);
          any |= m[j];
        }
        if (!any) {
          alive = false;
          break;
        }
        if (rec_ok && lane < RWc) {
          uint64_t v = 0;
#pragma unroll
          for (int j = 0; j < NWS; ++j)
            if (lane == uint32_t(j)) v = m[j];
          p.e.recs[roff + uint64_t(w[14]) * RWc + lane] = v;
        }
      }
      if (alive) {
        if (lane == 0) {
          pair_q[np] = qu;
          pair_off[np] = roff;
        }
        ++np;
        vbits |= uint64_t(1) << li;
      }
      roff += uint64_t(nq) * RWc;
    }
    if (lane == 0) *vrow = vbits;
    if (!rec_ok && lane == 0) atomicOr(p.e.ctl + kBigCtlOverflow, 1ull);
  }
  flush();
}

template <class F>
int dispatch_k(uint32_t K, F&& f) {
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
  return set_error(BMS2_ERR_UNSUPPORTED, "refinement steps %u > %d", K, kBigMaxLevels);
}

inline uint32_t blocks_for(uint64_t n, uint32_t block) {
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((n + block - 1) / block, 1u << 20)));
}

}  // namespace

int big_levels(hipStream_t st, const bms2_device_csr* g, uint32_t K, uint64_t* out) {
  if (g->total_nodes && K)
    hipLaunchKernelGGL(shell_levels_kernel, dim3(blocks_for(g->total_nodes, 256)), dim3(256), 0, st,
                       g->graph_offsets, g->num_graphs, g->row_offsets, g->column_indices,
                       g->node_labels, g->total_nodes, K, out);
  BMS2_CHECK_LAUNCH("shell_levels_kernel");
  return BMS2_OK;
}

int big_query_records(hipStream_t st, const bms2_device_csr* query, const uint64_t* qpath,
                      const uint64_t* qcyc, const uint64_t* qlev, uint32_t K, const uint32_t* qpos,
                      const uint8_t* jorder, uint32_t* out) {
  if (query->total_nodes)
    hipLaunchKernelGGL(big_query_rec_kernel, dim3(blocks_for(query->total_nodes, 256)), dim3(256), 0,
                       st, query->graph_offsets, query->num_graphs, query->row_offsets,
                       query->node_labels, query->total_nodes, qpath, qcyc, qlev, K, qpos, jorder,
                       out);
  BMS2_CHECK_LAUNCH("big_query_rec_kernel");
  return BMS2_OK;
}

int big_screen(hipStream_t st, const bms2_device_csr* data, const uint64_t* dpath,
               const uint64_t* dcyc, const uint64_t* dlev, uint32_t K, const uint32_t* narrow,
               uint32_t n_narrow, const uint32_t* wide, uint32_t n_wide, uint32_t wide_above,
               uint32_t Q, const uint32_t* qpre, uint32_t pre_dwords, const uint32_t* bqrec,
               const BigEntries& e, const uint32_t* gid, uint64_t* valid) {
  BigScreenParams sp{data->graph_offsets, data->row_offsets, data->node_labels, dpath, dcyc, dlev,
                     data->total_nodes, narrow, wide_above, Q, qpre, pre_dwords, bqrec, e,
                     gid, valid, (Q + 63) / 64};
  return dispatch_k(K, [&](auto kc) {
    constexpr int KK = decltype(kc)::value;
    if (n_narrow) hipLaunchKernelGGL((big_screen_kernel<KK, 2>), dim3(n_narrow), dim3(64), 0, st, sp);
    if (n_wide) {
      BigScreenParams wp = sp;
      wp.glist = wide;
      hipLaunchKernelGGL((big_screen_kernel<KK, 4>), dim3(n_wide), dim3(64), 0, st, wp);
    }
    BMS2_CHECK_LAUNCH("big_screen_kernel");
    return BMS2_OK;
  });
}

}  // namespace bms2
