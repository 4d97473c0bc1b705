// Per-node signature tuples computed on chip for a "pack": consecutive whole graphs with at most
// 64 nodes in total, one node per lane of a wave64.  Node ids inside a pack are lane ids, so the
// adjacency of every node is one u64 row and every shell / domain is one u64 word.
#pragma once

#include "device_common.hpp"

namespace bms2 {

constexpr int kMaxLevels = 8;  // refinement steps supported (BMS2 -i)

struct Tuple {
  uint32_t label;
  uint32_t deg;
  uint32_t path[8];
  uint64_t cycle;              // u16 {5 same, 5 diff, 6 same, 6 diff}
  uint64_t hist[kMaxLevels];   // shell k+1 label histogram (16 x 4-bit)
};

// LDS scratch of one wave for tuple computation.
struct PackScratch {
  uint64_t* adj;    // [64] adjacency rows over lanes
  uint8_t* lbl;     // [64] labels
  uint32_t* bloom;  // [64 * 8] per-lane Bloom words
};

// Computes the tuple of lane `lane` (< nn) of a pack.  `base` is per lane: the lane's node is
// global node base + lane, and every neighbour v (same graph) sits in lane v - base, so a graph
// may occupy any run of lanes (base = its first node - its first lane).  Must be called by the
// whole wave (contains barriers and ballots).  Levels >= K are left zero.
template <int K>
__device__ inline void compute_tuples(uint32_t base, uint32_t nn, const uint32_t* __restrict__ ro,
                                      const uint32_t* __restrict__ ci,
                                      const uint8_t* __restrict__ labels, PackScratch s, Tuple& t) {
  const uint32_t lane = lane_id();
  const bool active = lane < nn;
  uint64_t row = 0;
  t.label = 0;
  t.deg = 0;
  bool multi = false;
  if (active) {
    const uint32_t u = base + lane;
    t.label = labels[u];
    const uint32_t b = ro[u], e = ro[u + 1];
    t.deg = e - b;
    for (uint32_t k = b; k < e; ++k) row |= uint64_t(1) << (ci[k] - base);
    multi = uint32_t(popc64(row)) != t.deg;  // parallel edges
  }
  s.adj[lane] = row;
  s.lbl[lane] = uint8_t(t.label);
#pragma unroll
  for (int w = 0; w < 8; ++w) s.bloom[lane * 8 + w] = 0;
  const bool any_multi = ballot(multi) != 0;
  __syncthreads();

  // Path Bloom filter (signature.hpp:482-514).  Walks u-v-w with w != u and u-v-w-x with
  // x not in {u, v}; parallel edges only repeat walks, so adjacency sets give the same bits.
  if (active) {
    const uint64_t self = uint64_t(1) << lane;
    const uint32_t lu = t.label;
    uint64_t vs = row;
    while (vs) {
      const int v = ctz64(vs);
      vs &= vs - 1;
      const uint32_t lv = s.lbl[v];
      uint64_t ws = s.adj[v] & ~self;
      while (ws) {
        const int w = ctz64(ws);
        ws &= ws - 1;
        const uint32_t p2 = combine_hash(lu, combine_hash(lv, s.lbl[w]));
        {
          const uint32_t h1 = hash32(p2), h2 = hash32(h1), h3 = hash32(h2);
          atomicOr(&s.bloom[lane * 8 + ((h1 & 255u) >> 5)], 1u << (h1 & 31));
          atomicOr(&s.bloom[lane * 8 + ((h2 & 255u) >> 5)], 1u << (h2 & 31));
          atomicOr(&s.bloom[lane * 8 + ((h3 & 255u) >> 5)], 1u << (h3 & 31));
        }
        uint64_t xs = s.adj[w] & ~self & ~(uint64_t(1) << v);
        while (xs) {
          const int x = ctz64(xs);
          xs &= xs - 1;
          const uint32_t h1 = hash32(combine_hash(p2, s.lbl[x]));
          const uint32_t h2 = hash32(h1), h3 = hash32(h2);
          atomicOr(&s.bloom[lane * 8 + ((h1 & 255u) >> 5)], 1u << (h1 & 31));
          atomicOr(&s.bloom[lane * 8 + ((h2 & 255u) >> 5)], 1u << (h2 & 31));
          atomicOr(&s.bloom[lane * 8 + ((h3 & 255u) >> 5)], 1u << (h3 & 31));
        }
      }
    }
  }

  // Ring counts (signature.hpp:602-715): simple paths s-v1-..-v4 (5-rings) and s-..-v5 (6-rings)
  // closed by an edge back to s, each ring found once per direction, counters kept as u16.  The
  // enumeration order differs from the reference's stack order; the totals agree whenever its
  // 128-entry stack cannot overflow, which holds for degree <= 27 (checked on the host).
  uint32_t c5s = 0, c5d = 0, c6s = 0, c6d = 0;
  if (active) {
    const uint32_t ls = t.label;
    const uint64_t m0 = uint64_t(1) << lane;
    if (!any_multi) {
      uint64_t a1 = row;
      while (a1) {
        const int v1 = ctz64(a1);
        a1 &= a1 - 1;
        const bool s1 = s.lbl[v1] == ls;
        const uint64_t m1 = m0 | (uint64_t(1) << v1);
        uint64_t a2 = s.adj[v1] & ~m1;
        while (a2) {
          const int v2 = ctz64(a2);
          a2 &= a2 - 1;
          const bool s2 = s1 && s.lbl[v2] == ls;
          const uint64_t m2 = m1 | (uint64_t(1) << v2);
          uint64_t a3 = s.adj[v2] & ~m2;
          while (a3) {
            const int v3 = ctz64(a3);
            a3 &= a3 - 1;
            const bool s3 = s2 && s.lbl[v3] == ls;
            const uint64_t m3 = m2 | (uint64_t(1) << v3);
            uint64_t a4 = s.adj[v3] & ~m3;
            while (a4) {
              const int v4 = ctz64(a4);
              a4 &= a4 - 1;
              const bool s4 = s3 && s.lbl[v4] == ls;
              const uint64_t r4 = s.adj[v4];
              if ((r4 >> lane) & 1) {
                if (s4) ++c5s; else ++c5d;
              }
              uint64_t a5 = r4 & ~(m3 | (uint64_t(1) << v4));
              while (a5) {
                const int v5 = ctz64(a5);
                a5 &= a5 - 1;
                if ((s.adj[v5] >> lane) & 1) {
                  if (s4 && s.lbl[v5] == ls) ++c6s; else ++c6d;
                }
              }
            }
          }
        }
      }
    } else {
      // Parallel edges: enumerate CSR entries so every walk counts with its multiplicity.
      const uint32_t sg = base + lane;
      for (uint32_t k1 = ro[sg]; k1 < ro[sg + 1]; ++k1) {
        const uint32_t v1 = ci[k1] - base;
        if (v1 == lane) continue;
        const bool s1 = s.lbl[v1] == ls;
        const uint64_t m1 = m0 | (uint64_t(1) << v1);
        for (uint32_t k2 = ro[base + v1]; k2 < ro[base + v1 + 1]; ++k2) {
          const uint32_t v2 = ci[k2] - base;
          if ((m1 >> v2) & 1) continue;
          const bool s2 = s1 && s.lbl[v2] == ls;
          const uint64_t m2 = m1 | (uint64_t(1) << v2);
          for (uint32_t k3 = ro[base + v2]; k3 < ro[base + v2 + 1]; ++k3) {
            const uint32_t v3 = ci[k3] - base;
            if ((m2 >> v3) & 1) continue;
            const bool s3 = s2 && s.lbl[v3] == ls;
            const uint64_t m3 = m2 | (uint64_t(1) << v3);
            for (uint32_t k4 = ro[base + v3]; k4 < ro[base + v3 + 1]; ++k4) {
              const uint32_t v4 = ci[k4] - base;
              if ((m3 >> v4) & 1) continue;
              const bool s4 = s3 && s.lbl[v4] == ls;
              const uint64_t m4 = m3 | (uint64_t(1) << v4);
              for (uint32_t k5 = ro[base + v4]; k5 < ro[base + v4 + 1]; ++k5) {
                const uint32_t v5 = ci[k5] - base;
                if (v5 == lane) {
                  if (s4) ++c5s; else ++c5d;
                  continue;
                }
                if ((m4 >> v5) & 1) continue;
                const bool s5 = s4 && s.lbl[v5] == ls;
                for (uint32_t k6 = ro[base + v5]; k6 < ro[base + v5 + 1]; ++k6)
                  if (ci[k6] - base == lane) {
                    if (s5) ++c6s; else ++c6d;
                  }
              }
            }
          }
        }
      }
    }
  }
  t.cycle = uint64_t((c5s & 0xffffu) / 2) | (uint64_t((c5d & 0xffffu) / 2) << 16) |
            (uint64_t((c6s & 0xffffu) / 2) << 32) | (uint64_t((c6d & 0xffffu) / 2) << 48);

#pragma unroll
  for (int w = 0; w < 8; ++w) t.path[w] = active ? s.bloom[lane * 8 + w] : 0u;

  // Shell histograms (refineCSRSignatures<PowerGraph>, signature.hpp:347-385, run for view
  // sizes 1..K from zeroed state): label counts of the nodes at BFS distance exactly k.
  uint64_t labmask[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) labmask[l] = ballot(active && t.label == uint32_t(l));
#pragma unroll
  for (int k = 0; k < kMaxLevels; ++k) t.hist[k] = 0;
  if (K > 0) {
    uint64_t seen = uint64_t(1) << lane, frontier = active ? seen : 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint64_t next = 0, f = frontier;
      while (f) {
        const int b = ctz64(f);
        f &= f - 1;
        next |= s.adj[b];
      }
      next &= ~seen;
      seen |= next;
      uint64_t h = 0;
#pragma unroll
      for (int l = 0; l < 16; ++l) {
        const uint32_t c = uint32_t(popc64(next & labmask[l]));
        h |= uint64_t(c > 15 ? 15 : c) << (4 * l);
      }
      t.hist[k] = h;
      frontier = next;
    }
  }
  __syncthreads();  // scratch may be reused by the caller
}

}  // namespace bms2
