// Staged entry points with the reference's per-stage contracts (signature.hpp, candidates.hpp,
// isomorphism.hpp, gmcr.hpp): each call materialises its stage's output in HBM, so callers can
// inspect the domain bit-matrix after every refinement step.  Kernels here favour exactness for
// any graph up to 256 nodes over speed; the fused pipeline is the fast path.
#include <algorithm>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "runtime.hpp"

namespace bms2 {

// ------------------------------------------------------------------ small bitsets (<= 256 nodes)
struct Bits4 {
  uint64_t w[4];
  __device__ void clear() { w[0] = w[1] = w[2] = w[3] = 0; }
  __device__ void set(uint32_t i) { w[i >> 6] |= uint64_t(1) << (i & 63); }
  __device__ bool get(uint32_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
  __device__ bool empty() const { return !(w[0] | w[1] | w[2] | w[3]); }
};

// ------------------------------------------------------------------ Signature<PowerGraph, 4>

__global__ void label_signature_kernel(const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                       const uint8_t* __restrict__ labels, uint64_t n,
                                       uint64_t* __restrict__ sig) {
  // generateCSRSignatures (signature.hpp:135-146): increments on the existing value.
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < n;
       u += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t s = sig[u];
    for (uint32_t k = ro[u]; k < ro[u + 1]; ++k) s = inc_label(s, labels[ci[k]]);
    sig[u] = s;
  }
}

__global__ void refine_signature_kernel(const uint32_t* __restrict__ go, uint32_t num_graphs,
                                        const uint32_t* __restrict__ ro,
                                        const uint32_t* __restrict__ ci,
                                        const uint8_t* __restrict__ labels, uint64_t n,
                                        uint32_t view_size, uint64_t* __restrict__ sig,
                                        uint64_t* __restrict__ reach /* [4][n] */) {
  // refineCSRSignatures<PowerGraph> (signature.hpp:347-385) with exact 256-bit sets.
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < n;
       u += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t g = graph_of(go, num_graphs, uint32_t(u));
    const uint32_t gs = go[g], gn = go[g + 1] - gs;
    const uint32_t ul = uint32_t(u) - gs;
    Bits4 reachable, frontier;
    reachable.clear();
    frontier.clear();
    reachable.set(ul);
    frontier.set(ul);
    for (uint32_t it = 0; it < view_size && !frontier.empty(); ++it) {
      Bits4 next;
      next.clear();
      for (int w = 0; w < 4; ++w) {
        uint64_t f = frontier.w[w];
        while (f) {
          const uint32_t x = uint32_t(w * 64 + ctz64(f));
          f &= f - 1;
          for (uint32_t k = ro[gs + x]; k < ro[gs + x + 1]; ++k) {
            const uint32_t y = ci[k] - gs;
            if (!reachable.get(y)) {
              reachable.set(y);
              next.set(y);
            }
          }
        }
      }
      frontier = next;
    }
    reachable.w[ul >> 6] &= ~(uint64_t(1) << (ul & 63));
    uint64_t old[4];
    for (int w = 0; w < 4; ++w) {
      old[w] = reach[uint64_t(w) * n + u];
      reachable.w[w] &= ~old[w];
    }
    uint64_t s = 0;
    for (uint32_t x = 0; x < gn; ++x)
      if (reachable.get(x)) s = inc_label(s, labels[gs + x]);
    sig[u] = s;
    for (int w = 0; w < 4; ++w) reach[uint64_t(w) * n + u] = reachable.w[w] | old[w];
  }
}

// ------------------------------------------------------------------ PathSignature

__global__ void path_signature_kernel(const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                      const uint8_t* __restrict__ labels, uint64_t n,
                                      uint64_t* __restrict__ out /* [n][4] */) {
  // generateCSRPathSignatures (signature.hpp:482-514): OR-s into the existing bits.
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < n;
       u += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t b[4] = {out[4 * u], out[4 * u + 1], out[4 * u + 2], out[4 * u + 3]};
    const uint32_t lu = labels[u];
    auto insert = [&](uint32_t h1) {
      const uint32_t h2 = hash32(h1), h3 = hash32(h2);
      const uint32_t i1 = h1 & 255u, i2 = h2 & 255u, i3 = h3 & 255u;
      for (int w = 0; w < 4; ++w) {
        uint64_t m = 0;
        if ((i1 >> 6) == uint32_t(w)) m |= uint64_t(1) << (i1 & 63);
        if ((i2 >> 6) == uint32_t(w)) m |= uint64_t(1) << (i2 & 63);
        if ((i3 >> 6) == uint32_t(w)) m |= uint64_t(1) << (i3 & 63);
        b[w] |= m;
      }
    };
    for (uint32_t i = ro[u]; i < ro[u + 1]; ++i) {
      const uint32_t v = ci[i];
      const uint32_t lv = labels[v];
      for (uint32_t j = ro[v]; j < ro[v + 1]; ++j) {
        const uint32_t w = ci[j];
        if (w == u) continue;
        const uint32_t p2 = combine_hash(lu, combine_hash(lv, labels[w]));
        insert(hash32(p2));
        for (uint32_t k = ro[w]; k < ro[w + 1]; ++k) {
          const uint32_t x = ci[k];
          if (x == v || x == u) continue;
          insert(hash32(combine_hash(p2, labels[x])));
        }
      }
    }
    for (int w = 0; w < 4; ++w) out[4 * u + w] = b[w];
  }
}

// ------------------------------------------------------------------ CycleSignature

__global__ void cycle_signature_kernel(const uint32_t* __restrict__ go, uint32_t num_graphs,
                                       const uint32_t* __restrict__ ro,
                                       const uint32_t* __restrict__ ci,
                                       const uint8_t* __restrict__ labels, uint64_t n,
                                       uint64_t* __restrict__ out) {
  // generateCSRCycleSignatures (signature.hpp:602-715): simple paths of 4 and 5 edges from s that
  // close back to s, per CSR entry (parallel edges count), u16 counters halved.  Graphs above 64
  // nodes keep their value (:614).
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < n;
       u += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t g = graph_of(go, num_graphs, uint32_t(u));
    const uint32_t gs = go[g], gn = go[g + 1] - gs;
    if (gn > 64) continue;
    const uint32_t s = uint32_t(u) - gs;
    const uint8_t ls = labels[u];
    uint32_t c5s = 0, c5d = 0, c6s = 0, c6d = 0;
    const uint64_t m0 = uint64_t(1) << s;
    for (uint32_t k1 = ro[u]; k1 < ro[u + 1]; ++k1) {
      const uint32_t v1 = ci[k1] - gs;
      if ((m0 >> v1) & 1) continue;
      const bool s1 = labels[gs + v1] == ls;
      const uint64_t m1 = m0 | (uint64_t(1) << v1);
      for (uint32_t k2 = ro[gs + v1]; k2 < ro[gs + v1 + 1]; ++k2) {
        const uint32_t v2 = ci[k2] - gs;
        if ((m1 >> v2) & 1) continue;
        const bool s2 = s1 && labels[gs + v2] == ls;
        const uint64_t m2 = m1 | (uint64_t(1) << v2);
        for (uint32_t k3 = ro[gs + v2]; k3 < ro[gs + v2 + 1]; ++k3) {
          const uint32_t v3 = ci[k3] - gs;
          if ((m2 >> v3) & 1) continue;
          const bool s3 = s2 && labels[gs + v3] == ls;
          const uint64_t m3 = m2 | (uint64_t(1) << v3);
          for (uint32_t k4 = ro[gs + v3]; k4 < ro[gs + v3 + 1]; ++k4) {
            const uint32_t v4 = ci[k4] - gs;
            if ((m3 >> v4) & 1) continue;
            const bool s4 = s3 && labels[gs + v4] == ls;
            const uint64_t m4 = m3 | (uint64_t(1) << v4);
            for (uint32_t k5 = ro[gs + v4]; k5 < ro[gs + v4 + 1]; ++k5) {
              const uint32_t v5 = ci[k5] - gs;
              if (v5 == s) {
                if (s4) ++c5s; else ++c5d;
                continue;
              }
              if ((m4 >> v5) & 1) continue;
              const bool s5 = s4 && labels[gs + v5] == ls;
              for (uint32_t k6 = ro[gs + v5]; k6 < ro[gs + v5 + 1]; ++k6)
                if (ci[k6] - gs == s) {
                  if (s5) ++c6s; else ++c6d;
                }
            }
          }
        }
      }
    }
    out[u] = uint64_t((c5s & 0xffffu) / 2) | (uint64_t((c5d & 0xffffu) / 2) << 16) |
             (uint64_t((c6s & 0xffffu) / 2) << 32) | (uint64_t((c6d & 0xffffu) / 2) << 48);
  }
}

// ------------------------------------------------------------------ filter / refine

constexpr uint64_t kCycleTop = 0x8000800080008000ull;  // top bit of each 16-bit ring count

// A query node's filter operands in one 64-byte record (one scalar load per node in the filter).
struct alignas(64) FilterQuery {
  uint32_t label;       // node label
  uint32_t label_mask;  // 0xff, or 0 for the wildcard (any label matches)
  uint32_t degree;
  uint32_t pad0;
  uint64_t path[4];     // path signature bits the data node must contain
  uint64_t cycle;       // ring counts, 16 bits each, at most 0x7fff
  uint64_t pad1;
};

__global__ void filter_queries_kernel(const uint32_t* __restrict__ qro,
                                      const uint8_t* __restrict__ qlab, uint64_t Qn,
                                      const uint64_t* __restrict__ qpath,
                                      const uint64_t* __restrict__ qcyc,
                                      FilterQuery* __restrict__ out) {
  for (uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; u < Qn;
       u += uint64_t(gridDim.x) * blockDim.x) {
    FilterQuery f{};
    f.label = qlab[u];
    f.label_mask = f.label == kWildcard ? 0u : 0xffu;
    f.degree = qro[u + 1] - qro[u];
    for (int k = 0; k < 4; ++k) f.path[k] = qpath[4 * u + k];
    f.cycle = qcyc[u];
    out[u] = f;
  }
}

// FilterCandidatesKernel_v2 (isomorphism.hpp:59-97): query node u keeps data node v when the labels
// match (or u's is the wildcard), deg(u) <= deg(v), v's path signature contains u's and v's ring
// counts are at least u's; bits are ORed into the matrix (written only where they add bits).
// One wave per 64 data nodes (one u64 word of every row): each query node's record arrives by one
// scalar load and its row word comes from one ballot over branch-free vector tests; lane j
// collects the words of rows qb + j for a band of 64 query nodes, whose old words were loaded at
// the band's start.
__global__ __launch_bounds__(256) void filter_kernel(
    const FilterQuery* __restrict__ fq, uint64_t Qn, const uint32_t* __restrict__ dro,
    const uint8_t* __restrict__ dlab, uint64_t Dn, const uint64_t* __restrict__ dpath,
    const uint64_t* __restrict__ dcyc, uint64_t* __restrict__ cand, uint64_t W) {
  const uint32_t lane = lane_id();
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x / 64;
  for (uint64_t word = (xcd_block() * uint64_t(blockDim.x) + threadIdx.x) / 64; word < W;
       word += stride) {
    const uint64_t d = word * 64 + lane;
    const bool active = d < Dn;
    uint32_t dl = 0;
    int32_t dd = -1;  // a missing node fails the degree test
    uint64_t nd[4] = {0, 0, 0, 0}, dch = 0;
    if (active) {
      dl = dlab[d];
      dd = int32_t(dro[d + 1] - dro[d]);
      for (int w = 0; w < 4; ++w) nd[w] = ~dpath[4 * d + w];
      // ring counts are at most 0x7fff, so (dc | top bits) - qc borrows across no field and
      // leaves a field's top bit set exactly when the data count is at least the query count
      dch = dcyc[d] | kCycleTop;
    }
    // the node's values have arrived: without this wait the first use of one inside the band
    // loop waits for the band's matrix words too
    __builtin_amdgcn_s_waitcnt(0);
    // row word of data node lane for query node u: one ballot of branch-free tests (a failing
    // degree test sets the top bit of the difference, folded into the miss bits)
    auto test = [&](const FilterQuery& f) {
      const uint64_t miss = (f.path[0] & nd[0]) | (f.path[1] & nd[1]) | (f.path[2] & nd[2]) |
                            (f.path[3] & nd[3]) | (~(dch - f.cycle) & kCycleTop) |
                            ((dl ^ f.label) & f.label_mask) |
                            (uint32_t(dd - int32_t(f.degree)) & 0x80000000u);
      return ballot(miss == 0);
    };
    for (uint64_t qb = 0; qb < Qn; qb += 64) {
      const uint64_t r = qb + lane;
      const bool has = r < Qn;
      // every lane loads (rows past the end read the last row and store nothing)
      const uint64_t old = cand[std::min<uint64_t>(r, Qn - 1) * W + word];
      const FilterQuery* band = fq + qb;
      uint64_t acc = 0;
      if (Qn - qb >= 64) {
#pragma unroll 8
        for (uint32_t j = 0; j < 64; ++j) acc = writelane64(test(band[j]), int(j), acc);
      } else {
        for (uint32_t j = 0; j < uint32_t(Qn - qb); ++j)
          acc = writelane64(test(band[j]), int(j), acc);
      }
      if (has && (acc & ~old)) cand[r * W + word] = old | acc;
    }
  }
}

// Bit planes of the data signatures for refinement: plane j of word column w (planes[j W + w]) has
// bit b set when bit j of data node 64 w + b's signature is set (bits 4 l .. 4 l + 3 are its count
// of label class l).  One wave per column, one ballot per plane.
__global__ __launch_bounds__(256) void sig_planes_kernel(const uint64_t* __restrict__ dsig,
                                                         uint64_t Dn, uint64_t W,
                                                         uint64_t* __restrict__ planes) {
  const uint32_t lane = lane_id();
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x / 64;
  for (uint64_t w = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / 64; w < W; w += stride) {
    const uint64_t d = w * 64 + lane;
    const uint64_t ds = d < Dn ? dsig[d] : 0;
    uint64_t mine = 0;  // lane j keeps plane j
    for (int j = 0; j < 64; ++j) mine = writelane64(ballot((ds >> j) & 1), j, mine);
    planes[lane * W + w] = mine;
  }
}

// RefineCandidatesKernel_v3 (isomorphism.hpp:142-196): data node v stays a candidate of query node
// u when each of the 16 4-bit counts of v's signature is at least u's; nodes past the batch end are
// cleared.  One wave per 64 word columns and a run of rows: lane c holds column c's 64 signature
// planes, so a row's 64 words (one coalesced 512-byte load) are refined by a bit-sliced
// comparison, 64 data nodes per instruction: for each count of the row's query signature that is
// not zero, the borrow of (data count - query count) runs through the count's four planes, and a
// node whose count borrows is dropped.  The next rows' words load while a row is refined.
constexpr uint32_t kRefineRun = 128;  // rows per wave
constexpr int kRefineAhead = 4;       // rows in flight

__global__ __launch_bounds__(64) void refine_candidates_kernel(const uint64_t* __restrict__ qsig,
                                                               const uint64_t* __restrict__ planes,
                                                               uint64_t Dn,
                                                               uint64_t* __restrict__ cand,
                                                               uint64_t W, uint64_t rows) {
  const uint64_t groups = (W + 63) / 64, runs = (rows + kRefineRun - 1) / kRefineRun;
  const uint64_t col0 = uint64_t(blockIdx.x % groups) * 64;
  const uint64_t rbeg = uint64_t(blockIdx.x / groups) * kRefineRun;
  if (rbeg >= runs * kRefineRun) return;
  const uint64_t rend = std::min<uint64_t>(rows, rbeg + kRefineRun);
  const uint64_t col = col0 + lane_id();
  const bool has = col < W;
  uint64_t P[64];
#pragma unroll
  for (int j = 0; j < 64; ++j) P[j] = has ? planes[j * W + col] : 0;
  const uint64_t valid = has ? low_mask(uint32_t(std::min<uint64_t>(64, Dn - col * 64))) : 0;
  uint64_t ahead[kRefineAhead];
#pragma unroll
  for (int k = 0; k < kRefineAhead; ++k) {
    const uint64_t r = rbeg + k;
    ahead[k] = (has && r < rend) ? cand[r * W + col] : 0;
  }
  for (uint64_t r0 = rbeg; r0 < rend; r0 += kRefineAhead) {
    uint64_t cur[kRefineAhead];
#pragma unroll
    for (int k = 0; k < kRefineAhead; ++k) {
      cur[k] = ahead[k];
      const uint64_t r = r0 + kRefineAhead + k;
      ahead[k] = (has && r < rend) ? cand[r * W + col] : 0;
    }
#pragma unroll
    for (int k = 0; k < kRefineAhead; ++k) {
      const uint64_t r = r0 + k;
      const uint64_t blk = cur[k];
      if (r >= rend || !ballot(blk != 0)) continue;
      const uint64_t qs = qsig[r];
      uint64_t drop = ~valid;
#pragma unroll
      for (int l = 0; l < 16; ++l) {
        const uint32_t c = uint32_t(qs >> (4 * l)) & 15;
        if (!c) continue;
        uint64_t borrow = 0;
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
          // query bit 1: borrow' = ~x | borrow; query bit 0: borrow' = ~x & borrow
          const uint64_t m = ((c >> k2) & 1) ? ~0ull : 0ull;
          borrow = (~P[4 * l + k2] & (borrow | m)) | (borrow & m);
        }
        drop |= borrow;
      }
      const uint64_t nb = blk & ~drop;
      if (nb != blk) cand[r * W + col] = nb;
    }
  }
}

__global__ void row_count_kernel(const uint64_t* __restrict__ cand, uint64_t rows, uint64_t W,
                                 uint32_t* __restrict__ out) {
  const uint64_t row = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / 64;
  if (row >= rows) return;
  uint32_t c = 0;
  for (uint64_t w = lane_id(); w < W; w += 64) c += popc64(cand[row * W + w]);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane_id() == 0) out[row] = c;
}

// ------------------------------------------------------------------ GMCR

// Bits [s, e) of a candidate row as words of 64 local ids (exact masks: candidates.hpp:94,140
// shift by 32 when e % 32 == 0).
__device__ inline uint64_t row_bits(const uint64_t* row, uint32_t s, uint32_t e, uint32_t k) {
  const uint32_t lo = s + 64 * k;
  if (lo >= e) return 0;
  const uint32_t w = lo >> 6, off = lo & 63;
  uint64_t x = row[w] >> off;
  if (off && (lo - off + 64) < e) x |= row[w + 1] << (64 - off);
  return x & low_mask(e - lo);
}

__device__ inline bool pair_valid(const uint64_t* cand, uint64_t W, const uint32_t* qgo, uint32_t q,
                                  uint32_t s, uint32_t e) {
  const uint32_t qb = qgo[q], nq = qgo[q + 1] - qb;
  if (nq <= 1) return false;  // gmcr.hpp:64
  for (uint32_t i = 0; i < nq; ++i) {
    const uint64_t* row = cand + uint64_t(qb + i) * W;
    bool any = false;
    for (uint32_t k = 0; s + 64 * k < e && !any; ++k) any = row_bits(row, s, e, k) != 0;
    if (!any) return false;
  }
  return true;
}

// GMCR k1 (gmcr.hpp:54-88): one wave per data graph, lanes over queries.
__global__ __launch_bounds__(256) void gmcr_count_kernel(const uint64_t* __restrict__ cand, uint64_t W,
                                                         const uint32_t* __restrict__ qgo, uint32_t Q,
                                                         const uint32_t* __restrict__ dgo, uint32_t D,
                                                         uint64_t* __restrict__ vbits, uint64_t VW,
                                                         uint32_t* __restrict__ cnt) {
  const uint64_t g = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / 64;
  if (g >= D) return;
  const uint32_t s = dgo[g], e = dgo[g + 1];
  uint32_t total = 0;
  for (uint32_t q0 = 0; q0 < Q; q0 += 64) {
    const uint32_t q = q0 + lane_id();
    const bool v = q < Q && pair_valid(cand, W, qgo, q, s, e);
    const uint64_t m = ballot(v);
    if (lane_id() == 0) vbits[g * VW + q0 / 64] = m;
    total += popc64(m);
  }
  if (lane_id() == 0) cnt[g] = total;
}

// Exclusive scans of u32 counts into u32 offsets [n + 1] (replace the serial single_task of
// gmcr.hpp:92-94).  A 1024-thread block scans a chunk of kScanChunk counts at a time.
constexpr uint32_t kScanPer = 8, kScanChunk = 1024 * kScanPer;

// Counts [base, base + kScanChunk) into offsets from `start`; returns the chunk's sum.
__device__ inline uint32_t scan_chunk(const uint32_t* __restrict__ cnt, uint64_t n, uint64_t base,
                                      uint32_t start, uint32_t* __restrict__ offsets,
                                      uint32_t* part) {
  uint32_t local[kScanPer];
  uint32_t sum = 0;
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint64_t i = base + uint64_t(threadIdx.x) * kScanPer + k;
    local[k] = i < n ? cnt[i] : 0;
    sum += local[k];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = start + part[threadIdx.x] - sum;
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint64_t i = base + uint64_t(threadIdx.x) * kScanPer + k;
    if (i < n) offsets[i] = run;
    run += local[k];
  }
  const uint32_t total = part[1023];
  __syncthreads();
  return total;
}

// One block over all chunks in turn.
__global__ __launch_bounds__(1024) void scan_kernel(const uint32_t* __restrict__ cnt, uint64_t n,
                                                    uint32_t* __restrict__ offsets) {
  __shared__ uint32_t part[1024];
  uint32_t carry = 0;
  for (uint64_t base = 0; base < n; base += kScanChunk)
    carry += scan_chunk(cnt, n, base, carry, offsets, part);
  if (threadIdx.x == 0) offsets[n] = carry;
}

// Multi-block scan, step 1: block b sums chunk b.
__global__ __launch_bounds__(1024) void scan_sums_kernel(const uint32_t* __restrict__ cnt, uint64_t n,
                                                         uint32_t* __restrict__ sums) {
  __shared__ uint32_t part[1024];
  const uint64_t base = uint64_t(blockIdx.x) * kScanChunk;
  uint32_t sum = 0;
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint64_t i = base + uint64_t(threadIdx.x) * kScanPer + k;
    if (i < n) sum += cnt[i];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 512; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// Step 3 (after the chunk sums are scanned into chunk_off [chunks + 1]): block b scans chunk b
// from its offset; block 0 writes the total.
__global__ __launch_bounds__(1024) void scan_add_kernel(const uint32_t* __restrict__ cnt, uint64_t n,
                                                        const uint32_t* __restrict__ chunk_off,
                                                        uint64_t chunks, uint32_t* __restrict__ offsets) {
  __shared__ uint32_t part[1024];
  (void)scan_chunk(cnt, n, uint64_t(blockIdx.x) * kScanChunk, chunk_off[blockIdx.x], offsets, part);
  if (blockIdx.x == 0 && threadIdx.x == 0) offsets[n] = chunk_off[chunks];
}

// GMCR k3 (gmcr.hpp:111-126) made deterministic: ascending query ids per data graph.
__global__ __launch_bounds__(256) void gmcr_fill_kernel(const uint64_t* __restrict__ vbits, uint64_t VW,
                                                        uint32_t Q, uint32_t D,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ indices) {
  const uint64_t g = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / 64;
  if (g >= D) return;
  uint32_t pos = offsets[g];
  for (uint32_t q0 = 0; q0 < Q; q0 += 64) {
    const uint64_t m = vbits[g * VW + q0 / 64];
    const uint32_t q = q0 + lane_id();
    if ((m >> lane_id()) & 1) indices[pos + popc64(m & low_mask(lane_id()))] = q;
    pos += popc64(m);
  }
}

// ------------------------------------------------------------------ join

// ---- join for data graphs of 65-256 nodes

// GMCR entries on graphs above 64 nodes are sorted into the size classes of big_class, one join
// instance per class.
constexpr int kBigDepth = BMS2_MAX_QUERY_NODES;  // depths of a lane's stash
constexpr uint32_t kBigWavesPerCU = 12;

// Multi-word adjacency rows: words [RW x, RW x + RW) of `rows` hold local node x's neighbours
// (bit y % 64 of word y / 64 for local neighbour y); one thread per data node.
__global__ void adj_rows_wide_kernel(const uint32_t* __restrict__ go, uint32_t num_graphs,
                                     const uint32_t* __restrict__ ro,
                                     const uint32_t* __restrict__ ci, uint64_t n, uint32_t RW,
                                     uint64_t* __restrict__ rows) {
  const uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (u >= n) return;
  const uint32_t s = go[graph_of(go, num_graphs, uint32_t(u))];
  uint64_t r[4] = {0, 0, 0, 0};
  for (uint32_t k = ro[u]; k < ro[u + 1]; ++k) {
    const uint32_t y = ci[k] - s;
    r[(y >> 6) & 3] |= uint64_t(1) << (y & 63);
  }
  for (uint32_t w = 0; w < RW; ++w) rows[u * RW + w] = r[w];
}

// Runs of the GMCR entries on graphs above 64 nodes, by size class (big_class): one run per data
// graph with entries, {graph, entries, first entry, 0}, one thread per graph.  Pass 0 counts the
// runs of each class into ctl[kBigClasses + c]; pass 1 writes them after the earlier classes'
// runs, in any order (ctl[c] is the fill counter).
__global__ __launch_bounds__(256) void big_classify_kernel(
    const uint32_t* __restrict__ dgo, uint32_t D, const uint32_t* __restrict__ goff, int pass,
    uint32_t wide_above, unsigned long long* __restrict__ ctl, uint4* __restrict__ runs) {
  const uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (g >= D) return;
  const uint32_t k0 = goff[g], k1 = goff[g + 1];
  if (k0 == k1) return;
  const int c = big_class(dgo[g + 1] - dgo[g], wide_above);
  if (pass == 0) {
    atomicAdd(ctl + kBigClasses + c, 1ull);
    return;
  }
  uint64_t base = 0;
  for (int cc = 0; cc < c; ++cc) base += ctl[kBigClasses + cc];
  const uint64_t i = atomicAdd(ctl + c, 1ull);
  runs[base + i] = make_uint4(uint32_t(g), k1 - k0, k0, 0u);
}

// The staged join for data graphs of 65 to 64 NW nodes (JoinCandidates3Optimized,
// isomorphism.hpp:309-468, with standard backtracking over bitsets of the graph's local ids;
// identity matching order unless jorder gives one, DFS nodes per pair when asked).  Persistent
// waves take runs of entries (consecutive GMCR entries of one data graph) from a global cursor,
// copy the run's graph's adjacency rows (NW u64 words per node) into their LDS and hand the
// run's entries to idle lanes.  At the start of a pair a lane loads every depth's domain (loads
// of four depths in flight) into its stash in global memory.  The search keeps, for the current
// depth, the remaining candidates R and X = (domain of the next depth) minus the nodes matched
// at earlier depths, intersected with their adjacency constraints, in registers; a step reads
// one adjacency row from LDS and gives the next depth's candidates as
// X & ~{v} & (adjacent ? N(v) : ~N(v)).  A descent pushes R to the stash and builds the new X
// from the stashed domain and the LDS rows of the nodes matched so far; a backtrack pops R and
// rebuilds X the same way.
// Work sharing (count-all mode): once a run's entries are handed out and at least kShareIdle
// lanes are idle, busy lanes give work to idle ones.  A donor gives all remaining candidates of
// its shallowest depth that has any (or, when only its current depth has, the upper half of
// them); the receiver copies the donor's matched nodes above that depth and searches the given
// candidates' subtrees as a unit based at that depth.  A pair's units add their counts and DFS
// nodes; the sums equal the serial search's, since every subtree is searched exactly once.  A
// heavy pair so spreads over the wave's 64 lanes instead of holding one lane while the others
// idle.  First-match mode never shares: its DFS-node count depends on the search order.

struct BigJoinParams {
  const uint64_t* cand;         // candidate rows (entries of a GMCR), or null with recs
  uint64_t W;
  const uint32_t* qgo;
  const uint32_t* qa;
  const uint32_t* chain;        // per query: bit d = depths d and d - 1 adjacent
  const uint32_t* gidx;         // query of each entry
  const uint32_t* dgo;
  const uint64_t* rows;         // adjacency rows, RW words per node
  uint32_t RW;
  const uint4* runs;            // {graph, entries, first entry (low, high)}
  uint64_t run_stride;          // class c's runs start at c run_stride, or (0) after the
                                // earlier classes' runs
  unsigned long long* ctl;      // [kBigClasses + c] runs of class c, [2 kBigClasses + c] cursors
  uint64_t* stash;
  uint32_t SD;                  // stash depth
  int find_first;
  uint64_t* total;              // nullable
  uint64_t* per_pair;           // nullable, zeroed: count per entry (find_first: 0/1)
  const uint8_t* jorder;
  uint64_t* per_pair_nodes;     // nullable, zeroed
  // domains of the fused pipeline's big-graph screen (non-null: cand and the stash's domain
  // half are unused): NW words per depth of the matching order at recs + rec_off[entry]
  const uint64_t* rec_off;
  const uint64_t* recs;
  // outputs of the fused pipeline (nullable): counts[gid[g] Q + q] (zeroed) and the statistics
  // {matched pairs, embeddings, DFS nodes} added at stats[1..3]
  uint64_t* out_counts;
  const uint32_t* out_gid;
  uint32_t Q;
  uint64_t* stats;
};

constexpr int kShareIdle = 16;  // idle lanes of a wave that start a round of work sharing
// Runs are taken in the order i -> i x (kSpread mod T) mod T, a bijection of [0, T) for every
// T < 2^32 (kSpread is a prime above 2^32): consecutive claims then reach graphs far apart in
// the batch rather than neighbours of similar size and shape.
constexpr uint64_t kSpread = 4294967311ull;

// Position of the k-th (from 0) set bit of m; m has more than k set bits.
__device__ inline uint32_t nth_bit(uint64_t m, uint32_t k) {
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t step = 32; step; step >>= 1)
    if (uint32_t(popc64(m & low_mask(lo + step))) <= k) lo += step;
  return lo;
}

template <int NW, int BW>  // BW: waves per workgroup
__global__ __launch_bounds__(64 * BW) void join_big_kernel(BigJoinParams p, int cls) {
  constexpr int NQ = kBigDepth;
  __shared__ uint8_t F[NQ][64 * BW];         // node matched per depth, per thread
  __shared__ uint64_t rows_lds[BW][64 * NW * NW];  // the current run's graph: NW words per node
  const uint32_t lane = lane_id();
  const uint32_t tid = threadIdx.x;
  const uint32_t wbase = tid - lane;  // the wave's first thread
  uint64_t* const RL = rows_lds[tid / 64];
  unsigned long long* const cursor = p.ctl + 2 * kBigClasses + cls;
  uint64_t rbase = 0;
  if (p.run_stride)
    rbase = uint64_t(cls) * p.run_stride;
  else
    for (int c = 0; c < cls; ++c) rbase += p.ctl[kBigClasses + c];
  const uint64_t NR = p.ctl[kBigClasses + cls];
  const uint64_t spread = NR > 1 ? kSpread % NR : 1;
  const uint4* const runs = p.runs + rbase;
  // this lane's stash: the domains [SD][NW], then the pushed R [SD][NW] (SD >= the batch's
  // largest query)
  uint64_t* const DS = p.stash + (uint64_t(blockIdx.x) * 64 * BW + tid) * (2 * p.SD * NW);
  uint64_t* const RSt = DS + p.SD * NW;
  const uint64_t* dom = DS;  // domains by depth: the stash, or the pair's record
  const bool share = !p.find_first;
  bool busy = false;
  uint64_t t = 0, c = 0, nodes = 0, sum = 0, matched = 0, dfs = 0;
  uint64_t R[NW], X[NW], used[NW];
  uint32_t s = 0, e = 0, g = 0, q = 0, nq = 0, chain = 0;

  uint32_t pend = 0;      // depths below the current one whose pushed R is not empty
  int depth = 0, bd = 0;  // current depth, depth the lane's unit starts at
  constexpr int kHandout = 32;
  // X of depth d: the domain of depth d + 1 minus `used`, constrained by the nodes matched at
  // depths < d (rows from LDS), the row of depth d - 1 passed in when the caller has it
  auto build_x = [&](int d, const uint64_t* last_row, uint64_t* out) {
    const uint32_t am = p.qa[size_t(q) * 32 + d + 1];
#pragma unroll
    for (int k = 0; k < NW; ++k) out[k] = dom[(d + 1) * NW + k] & ~used[k];
    const int lim = last_row ? d - 1 : d;
    for (int i = 0; i < lim; ++i) {
      const uint32_t fi = F[i][tid];
      const bool a = (am >> i) & 1;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const uint64_t r = RL[fi * NW + k];
        out[k] &= a ? r : ~r;
      }
    }
    if (last_row) {
      const bool a = (am >> (d - 1)) & 1;
#pragma unroll
      for (int k = 0; k < NW; ++k) out[k] &= a ? last_row[k] : ~last_row[k];
    }
  };
  // every depth's domain of the lane's pair into its stash, four depths' loads in flight at a
  // time; at a pair's start also R (depth 0's domain) and X (depth 1's: nq >= 2, single-node
  // queries are never GMCR pairs)
  auto load_domains = [&](bool start) {
    chain = p.chain[q];
    if (p.recs) {
      dom = p.recs + p.rec_off[t];
      if (start) {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          R[k] = dom[k];
          X[k] = dom[NW + k];
        }
      }
      return;
    }
    dom = DS;
    const uint32_t qb = p.qgo[q];
    for (uint32_t i0 = 0; i0 < nq; i0 += 4) {
      uint64_t d4[4][NW];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t i = i0 + j;
        if (i < nq) {
          const uint32_t u = p.jorder ? p.jorder[size_t(q) * 32 + i] : i;
          if (p.recs) {
            const uint64_t* dom = p.recs + p.rec_off[t] + uint64_t(u) * NW;
#pragma unroll
            for (int k = 0; k < NW; ++k) d4[j][k] = dom[k];
          } else {
            const uint64_t* row = p.cand + uint64_t(qb + u) * p.W;
#pragma unroll
            for (int k = 0; k < NW; ++k) d4[j][k] = row_bits(row, s, e, k);
          }
        }
      }
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j)
        if (i0 + j < nq) {
#pragma unroll
          for (int k = 0; k < NW; ++k) DS[(i0 + j) * NW + k] = d4[j][k];
        }
      if (start && i0 == 0) {
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          R[k] = d4[0][k];
          X[k] = d4[1][k];
        }
      }
    }
  };
  while (true) {
    uint64_t ri = 0;
    if (lane == 0) ri = atomicAdd(cursor, 1ull);
    ri = uniform64(readlane64(ri, 0));
    if (ri >= NR) break;
    const uint4 run = runs[ri * spread % NR];
    g = run.x;
    const uint32_t cnt = run.y;
    const uint64_t first = uint64_t(run.z) | (uint64_t(run.w) << 32);
    s = p.dgo[g];
    e = p.dgo[g + 1];
    // the run's graph's rows into LDS (the lanes finished the previous run's searches)
    for (uint32_t w = lane; w < (e - s) * NW; w += 64)
      RL[w] = p.rows[uint64_t(s + w / NW) * p.RW + w % NW];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t next = 0;
    while (true) {
      const uint64_t idle = ballot(!busy);
      if (next < cnt && (popc64(idle) >= kHandout || idle == ~uint64_t(0))) {
        const uint32_t take = min(cnt - next, uint32_t(popc64(idle)));
        const uint32_t rank = uint32_t(popc64(idle & low_mask(lane)));
        if (!busy && rank < take) {
          busy = true;
          t = first + next + rank;
          q = p.gidx[t];
          nq = p.qgo[q + 1] - p.qgo[q];
          load_domains(true);
#pragma unroll
          for (int k = 0; k < NW; ++k) used[k] = 0;
          depth = bd = 0;
          pend = 0;
          c = 0;
          nodes = 0;
        }
        next += take;
      }
      if (share && next >= cnt) {
        const uint64_t free_lanes = ballot(!busy);
        if (popc64(free_lanes) >= kShareIdle && free_lanes != ~uint64_t(0)) {
          // a donor's depth: the shallowest with pushed candidates, else its current one when
          // it holds two or more
          int d0 = -1;
          if (busy) {
            if (pend) {
              d0 = __builtin_ctz(pend);
            } else {
              uint32_t pc = 0;
#pragma unroll
              for (int k = 0; k < NW; ++k) pc += uint32_t(popc64(R[k]));
              if (pc >= 2) d0 = depth;
            }
          }
          const uint64_t donors = ballot(d0 >= 0);
          const uint32_t npair = min(uint32_t(popc64(donors)), uint32_t(popc64(free_lanes)));
          if (npair) {
            uint64_t give[NW];
#pragma unroll
            for (int k = 0; k < NW; ++k) give[k] = 0;
            if (d0 >= 0 && uint32_t(popc64(donors & low_mask(lane))) < npair) {
              if (d0 < depth) {
#pragma unroll
                for (int k = 0; k < NW; ++k) {
                  give[k] = RSt[d0 * NW + k];
                  RSt[d0 * NW + k] = 0;
                }
                pend &= ~(1u << d0);
              } else {
                // keep the lower half of the current candidates, give the upper half
                uint32_t pc = 0;
#pragma unroll
                for (int k = 0; k < NW; ++k) pc += uint32_t(popc64(R[k]));
                uint32_t keep = pc - pc / 2;
#pragma unroll
                for (int k = 0; k < NW; ++k) {
                  const uint32_t n = uint32_t(popc64(R[k]));
                  if (keep >= n) {
                    keep -= n;
                  } else {
                    const uint64_t m = keep ? low_mask(nth_bit(R[k], keep)) : 0;
                    give[k] = R[k] & ~m;
                    R[k] &= m;
                    keep = 0;
                  }
                }
              }
            }
            // the k-th idle lane takes the k-th donor's gift (same run: same graph)
            const uint32_t rr = uint32_t(popc64(free_lanes & low_mask(lane)));
            const bool recv = !busy && rr < npair;
            const int src = recv ? int(nth_bit(donors, rr)) : int(lane);
            const uint64_t st = __shfl(t, src);
            const uint32_t sq = __shfl(q, src);
            const int sd = __shfl(d0, src);
            uint64_t gv[NW];
#pragma unroll
            for (int k = 0; k < NW; ++k) gv[k] = __shfl(give[k], src);
            if (recv) {
              t = st;
              q = sq;
              nq = p.qgo[q + 1] - p.qgo[q];
#pragma unroll
              for (int k = 0; k < NW; ++k) used[k] = 0;
              for (int i = 0; i < sd; ++i) {
                const uint32_t f = F[i][wbase + uint32_t(src)];
                F[i][tid] = uint8_t(f);
                used[f >> 6] |= uint64_t(1) << (f & 63);
              }
              load_domains(false);
              build_x(sd, nullptr, X);
#pragma unroll
              for (int k = 0; k < NW; ++k) R[k] = gv[k];
              depth = bd = sd;
              pend = 0;
              c = 0;
              nodes = 0;
              busy = true;
            }
          }
        }
      }
      if (!ballot(busy)) {
        if (next >= cnt) break;
        continue;
      }
      bool finished = false;
      if (busy) {
        uint64_t any = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) any |= R[k];
        if (!any) {
          if (depth == bd) {
            finished = true;
          } else {
            --depth;
#pragma unroll
            for (int k = 0; k < NW; ++k) R[k] = RSt[depth * NW + k];
            pend &= ~(1u << depth);
            const uint32_t f = F[depth][tid];
            used[f >> 6] &= ~(uint64_t(1) << (f & 63));
            build_x(depth, nullptr, X);
          }
        } else {
          // lowest remaining candidate v
          uint32_t v = 0;
          bool found = false;
#pragma unroll
          for (int k = 0; k < NW; ++k)
            if (!found && R[k]) {
              v = uint32_t(64 * k + ctz64(R[k]));
              R[k] &= R[k] - 1;
              found = true;
            }
          ++nodes;
          uint64_t nv[NW];
#pragma unroll
          for (int k = 0; k < NW; ++k) nv[k] = RL[v * NW + k];
          const int nd = depth + 1;
          const bool adj = (chain >> nd) & 1;
          uint64_t Y[NW], ny = 0;
#pragma unroll
          for (int k = 0; k < NW; ++k) {
            Y[k] = X[k] & (adj ? nv[k] : ~nv[k]);
            if (k == int(v >> 6)) Y[k] &= ~(uint64_t(1) << (v & 63));
            ny |= Y[k];
          }
          if (uint32_t(nd) == nq - 1) {
#pragma unroll
            for (int k = 0; k < NW; ++k) c += uint64_t(popc64(Y[k]));
            if (p.find_first && c) finished = true;
          } else if (ny) {
            // descend: push R, then X of the new depth (v is matched at `depth` now)
            uint64_t rest = 0;
#pragma unroll
            for (int k = 0; k < NW; ++k) {
              RSt[depth * NW + k] = R[k];
              rest |= R[k];
            }
            if (rest) pend |= 1u << depth;
            F[depth][tid] = uint8_t(v);
            used[v >> 6] |= uint64_t(1) << (v & 63);
            build_x(nd, nv, X);
#pragma unroll
            for (int k = 0; k < NW; ++k) R[k] = Y[k];
            depth = nd;
          }
        }
      }
      if (finished) {
        const uint64_t cc = p.find_first ? (c ? 1 : 0) : c;
        if (p.out_counts) {
          if (cc) {
            const unsigned long long old = atomicAdd(
                reinterpret_cast<unsigned long long*>(p.out_counts + uint64_t(p.out_gid[g]) * p.Q + q),
                (unsigned long long)cc);
            if (!old) ++matched;
          }
        } else if (p.per_pair && cc) {
          atomicAdd(reinterpret_cast<unsigned long long*>(p.per_pair + t), (unsigned long long)cc);
        }
        if (p.per_pair_nodes && nodes)
          atomicAdd(reinterpret_cast<unsigned long long*>(p.per_pair_nodes + t),
                    (unsigned long long)nodes);
        sum += cc;
        dfs += nodes;
        busy = false;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    matched += __shfl_xor(matched, o);
    dfs += __shfl_xor(dfs, o);
  }
  if (lane == 0) {
    if (p.total && sum) atomicAdd(reinterpret_cast<unsigned long long*>(p.total), (unsigned long long)sum);
    if (p.stats) {
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 1), (unsigned long long)matched);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 2), (unsigned long long)sum);
      atomicAdd(reinterpret_cast<unsigned long long*>(p.stats + 3), (unsigned long long)dfs);
    }
  }
}


// Adjacency rows of a batch whose graphs have at most 64 nodes: bit y of row x = local node y is a
// neighbour of local node x (one thread per node).
__global__ void adj_rows_kernel(const uint32_t* __restrict__ go, uint32_t num_graphs,
                                const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                uint64_t n, uint64_t* __restrict__ rows) {
  const uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (u >= n) return;
  const uint32_t s = go[graph_of(go, num_graphs, uint32_t(u))];
  uint64_t r = 0;
  for (uint32_t k = ro[u]; k < ro[u + 1]; ++k) r |= uint64_t(1) << (ci[k] - s);
  rows[u] = r;
}

// Query adjacency by matching depth for the staged join: bit j of qa[32 q + i] is set when the query
// nodes matched at depths i and j of query q are adjacent (depth i matches node order[i], identity
// without an order).  One thread per (query, depth).
__global__ void join_depth_adjacency_kernel(const uint32_t* __restrict__ qgo,
                                            const uint32_t* __restrict__ qro,
                                            const uint32_t* __restrict__ qci, uint32_t Q,
                                            const uint8_t* __restrict__ jorder,
                                            uint32_t* __restrict__ qa) {
  const uint64_t x = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (x >= uint64_t(Q) * 32) return;
  const uint32_t q = uint32_t(x / 32), i = uint32_t(x % 32);
  const uint32_t qb = qgo[q], nq = qgo[q + 1] - qb;
  if (i >= nq) return;
  const uint8_t* ord = jorder ? jorder + size_t(q) * 32 : nullptr;
  const uint32_t u = ord ? ord[i] : i;
  uint32_t a = 0;
  for (uint32_t k = qro[qb + u]; k < qro[qb + u + 1]; ++k) {
    const uint32_t w = qci[k] - qb;
    uint32_t j = w;
    if (ord)
      for (j = 0; j < nq && ord[j] != w; ++j) {
      }
    a |= 1u << j;
  }
  qa[x] = a;
}

// Per query, bit d set when the query nodes matched at depths d and d - 1 are adjacent (the
// test of every DFS step), from the depth adjacency qa; one thread per query.
__global__ void depth_chain_kernel(const uint32_t* __restrict__ qgo, uint32_t Q,
                                   const uint32_t* __restrict__ qa, uint32_t* __restrict__ chain) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Q) return;
  const uint32_t nq = qgo[q + 1] - qgo[q];
  uint32_t m = 0;
  for (uint32_t d = 1; d < nq && d < 32; ++d) m |= ((qa[size_t(q) * 32 + d] >> (d - 1)) & 1u) << d;
  chain[q] = m;
}

// The staged join for data graphs of up to 64 nodes (JoinCandidates3Optimized,
// isomorphism.hpp:309-468, standard backtracking): persistent waves take GMCR entries
// 256 at a time from a global cursor and hand them to idle lanes; a lane's search keeps its
// per-depth domains (from the candidate rows), remaining candidates and matched nodes' neighbour
// rows in LDS and reads one precomputed adjacency row from global memory per step.
template <int NQ>
__global__ __launch_bounds__(64) void join_rows_kernel(
    const uint64_t* __restrict__ cand, uint64_t W, const uint32_t* __restrict__ qgo,
    const uint32_t* __restrict__ qa, const uint32_t* __restrict__ dgo, uint32_t D,
    const uint64_t* __restrict__ adj,
    const uint32_t* __restrict__ goff, const uint32_t* __restrict__ gidx, uint64_t T,
    int find_first, uint64_t* __restrict__ total, uint64_t* __restrict__ per_pair,
    const uint8_t* __restrict__ jorder, uint64_t* __restrict__ per_pair_nodes,
    unsigned long long* __restrict__ cursor) {
  __shared__ uint64_t DOM[NQ][64];  // domain per depth (local ids of the data graph)
  __shared__ uint64_t R[NQ][64];    // remaining candidates per depth
  __shared__ uint64_t NF[NQ][64];   // neighbour row of the node matched per depth
  __shared__ uint32_t QA[NQ][64];   // bit i of QA[d]: depths i and d adjacent in the query
  __shared__ uint8_t F[NQ][64];     // node matched per depth
  const uint32_t lane = lane_id();
  bool busy = false, exhausted = false;
  uint64_t t = 0, used = 0, c = 0, nodes = 0, sum = 0;
  uint32_t s = 0, nq = 0;
  int depth = 0;
  uint64_t next = 0, end = 0;
  uint32_t gnext = 0;  // data graph of entry `next`
  constexpr uint64_t kClaim = 256;
  // a handout costs the dependent loads of one entry's setup whatever the number of lanes it
  // serves, so idle lanes wait until half the wave (or all of it) can take entries
  constexpr int kHandout = 32;
  while (true) {
    const uint64_t idle = ballot(!busy);
    if (popc64(idle) >= kHandout || idle == ~uint64_t(0)) {
      if (next >= end && !exhausted) {
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd(cursor, (unsigned long long)kClaim);
        base = uniform64(readlane64(base, 0));
        next = base;
        end = base >= T ? base : min(base + kClaim, T);
        if (base + kClaim >= T) exhausted = true;
        if (next < end) {
          // data graph of the claim's first entry (utils.hpp:110-121), searched once per claim
          uint32_t lo = 0, hi = D + 1;
          while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (next >= goff[mid]) lo = mid + 1; else hi = mid;
          }
          gnext = lo - 1;
        }
      }
      const uint64_t take = min(end - next, uint64_t(popc64(idle)));
      const uint64_t rank = uint64_t(popc64(idle & low_mask(lane)));
      if (!busy && rank < take) {
        busy = true;
        t = next + rank;
        // data graph of entry t: forward from the graph of the handout's first entry (a claim
        // spans a few graphs; t < goff[D] = T ends the walk)
        uint32_t g = gnext;
        while (t >= goff[g + 1]) ++g;
        const uint32_t q = gidx[t];
        s = dgo[g];
        const uint32_t e = dgo[g + 1], qb = qgo[q];
        nq = qgo[q + 1] - qb;
        // four depths' loads in flight at a time
        for (uint32_t i0 = 0; i0 < nq; i0 += 4) {
          uint64_t dom[4];
          uint32_t a[4];
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t i = i0 + k;
            const uint32_t u = i < nq ? (jorder ? jorder[size_t(q) * 32 + i] : i) : 0;
            dom[k] = i < nq ? row_bits(cand + uint64_t(qb + u) * W, s, e, 0) : 0;
            a[k] = i < nq ? qa[size_t(q) * 32 + i] : 0;
          }
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k)
            if (i0 + k < nq) {
              DOM[i0 + k][lane] = dom[k];
              QA[i0 + k][lane] = a[k];
            }
        }
        R[0][lane] = DOM[0][lane];
        depth = 0;
        used = 0;
        c = 0;
        nodes = 0;
      }
      next += take;
      while (next < end && next >= goff[gnext + 1]) ++gnext;
    }
    if (!ballot(busy)) {
      if (exhausted && next >= end) break;
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
        const uint64_t nv = adj[s + v];
        NF[depth][lane] = nv;
        ++nodes;
        const int nd = depth + 1;
        const uint64_t vbit = uint64_t(1) << v;
        const uint32_t am = QA[nd][lane];
        uint64_t Rn = DOM[nd][lane] & ~(used | vbit) & (((am >> depth) & 1) ? nv : ~nv);
        for (int i = 0; i < depth; ++i) {
          const uint64_t x = NF[i][lane];
          Rn &= ((am >> i) & 1) ? x : ~x;
        }
        if (uint32_t(nd) == nq - 1) {
          c += uint64_t(popc64(Rn));
          if (find_first && c) finished = true;
        } else {
          used |= vbit;
          depth = nd;
          R[nd][lane] = Rn;
        }
      }
    }
    if (finished) {
      const uint64_t cc = find_first ? (c ? 1 : 0) : c;
      if (per_pair) per_pair[t] = cc;
      if (per_pair_nodes) per_pair_nodes[t] = nodes;
      sum += cc;
      busy = false;
    }
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if (lane == 0 && sum) atomicAdd(reinterpret_cast<unsigned long long*>(total), (unsigned long long)sum);
}

// ------------------------------------------------------------------ GPU-side CSR packing

// Entry count of each graph: the local row end of its last node (0 for an empty graph).
__global__ void pack_entries_kernel(const uint32_t* __restrict__ nodes,
                                    const uint32_t* __restrict__ goff, uint32_t G,
                                    const uint32_t* __restrict__ ro_local, uint32_t* __restrict__ ent) {
  const uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (g < G) ent[g] = nodes[g] ? ro_local[goff[g] + nodes[g] - 1] : 0u;
}

// Error bits of the packed batch, set by pack_csr_kernel.
constexpr uint32_t kPackRange = 1, kPackLoop = 2, kPackAsym = 4, kPackRows = 8, kPackSize = 16;

// Graph id of every node (one thread per graph writes its range) and the largest graph.
__global__ void pack_node_graph_kernel(const uint32_t* __restrict__ goff, uint32_t G,
                                       uint32_t max_nodes, uint32_t* __restrict__ node_graph,
                                       uint32_t* __restrict__ stat) {
  const uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (g >= G) return;
  const uint32_t s = goff[g], e = goff[g + 1];
  for (uint32_t u = s; u < e; ++u) node_graph[u] = uint32_t(g);
  if (e - s > max_nodes) atomicOr(stat, kPackSize);
  atomicMax(stat + 1, e - s);
}

// createDeviceCSRGraph (graph.hpp:251-310) on the GPU, one thread per node: global row end =
// local row end + its graph's first entry, global column id = local id + its graph's first node;
// every entry is checked (endpoint inside the graph, no self-loop, v lists u as often as u lists
// v) before anything reads through it.  stat[0] error bits, stat[2] largest row.
__global__ void pack_csr_kernel(const uint32_t* __restrict__ goff, const uint32_t* __restrict__ eoff,
                                const uint32_t* __restrict__ node_graph, uint64_t N,
                                const uint32_t* __restrict__ ro_local,
                                const uint32_t* __restrict__ ci_local, uint32_t* __restrict__ ro,
                                uint32_t* __restrict__ ci, uint32_t* __restrict__ stat) {
  const uint64_t u = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (u >= N) return;
  if (u == 0) ro[0] = 0;
  const uint32_t g = node_graph[u];
  const uint32_t s = goff[g], i = uint32_t(u) - s, n = goff[g + 1] - s;
  const uint32_t e0 = eoff[g], m = eoff[g + 1] - e0;
  // local row range of node j of this graph, or false when it is not a valid range
  auto range = [&](uint32_t j, uint32_t* b, uint32_t* e) {
    *b = j ? ro_local[s + j - 1] : 0u;
    *e = ro_local[s + j];
    return *b <= *e && *e <= m;
  };
  uint32_t b, e;
  const bool ok = range(i, &b, &e);
  ro[u + 1] = ro_local[u] + e0;
  if (!ok) {
    atomicOr(stat, kPackRows);
    return;
  }
  if (e - b) atomicMax(stat + 2, e - b);
  uint32_t err = 0;
  for (uint32_t k = b; k < e; ++k) {
    const uint32_t v = ci_local[e0 + k];
    ci[e0 + k] = v + s;
    if (v >= n) {
      err |= kPackRange;
      continue;
    }
    if (v == i) {
      err |= kPackLoop;
      continue;
    }
    uint32_t vb, ve;
    if (!range(v, &vb, &ve)) continue;  // flagged by v's own thread
    uint32_t fwd = 0, back = 0;
    for (uint32_t j = b; j < e; ++j) fwd += ci_local[e0 + j] == v;
    for (uint32_t j = vb; j < ve; ++j) back += ci_local[e0 + j] == i;
    if (fwd != back) err |= kPackAsym;
  }
  if (err) atomicOr(stat, err);
}

// ------------------------------------------------------------------ CS (cluster) join

// Depth-first enumeration shared by the cluster-mode kernels: maps of depths [d0, n) of one query
// into one data graph (global nodes [s, e), NW words of local ids), standard backtracking.  The
// caller fills dom[i] (domain of depth i, local bits) for i in [d0, n), qadj[i] (bit j < i: depths
// i and j adjacent in the query) for i < n, and for a fixed prefix i < d0 its local node f[i], its
// neighbour row NR[i] and its bit in `used`.  Every complete map is passed to leaf(v) with f[n-1]
// = v, in ascending DFS order (candidates of each depth by ascending node id); leaf returns true to
// stop.  count_only counts the last depth by popcount instead.  Returns the number of maps.
template <int NW>
struct CsFrame {
  uint64_t dom[BMS2_MAX_QUERY_NODES][NW];
  uint64_t NR[BMS2_MAX_QUERY_NODES][NW];
  uint64_t R[BMS2_MAX_QUERY_NODES][NW];
  uint32_t qadj[BMS2_MAX_QUERY_NODES];
  uint32_t f[BMS2_MAX_QUERY_NODES];
  uint64_t used[NW];
};

template <int NW>
__device__ inline void cs_nrow(const uint32_t* __restrict__ dro, const uint32_t* __restrict__ dci,
                               uint32_t s, uint32_t x, uint64_t* out) {
  for (int k = 0; k < NW; ++k) out[k] = 0;
  for (uint32_t k = dro[s + x]; k < dro[s + x + 1]; ++k) {
    const uint32_t y = dci[k] - s;
    out[y >> 6] |= uint64_t(1) << (y & 63);
  }
}

template <int NW>
__device__ inline void cs_candidates(CsFrame<NW>& F, int d) {
  for (int k = 0; k < NW; ++k) F.R[d][k] = F.dom[d][k] & ~F.used[k];
  for (int i = 0; i < d; ++i) {
    const bool adj = (F.qadj[d] >> i) & 1;
    for (int k = 0; k < NW; ++k) F.R[d][k] &= adj ? F.NR[i][k] : ~F.NR[i][k];
  }
}

template <int NW, typename Leaf>
__device__ uint64_t cs_dfs(CsFrame<NW>& F, int d0, int n, const uint32_t* __restrict__ dro,
                           const uint32_t* __restrict__ dci, uint32_t s, bool count_only, Leaf leaf) {
  uint64_t c = 0;
  int d = d0;
  cs_candidates(F, d);
  while (true) {
    if (count_only && d == n - 1) {
      for (int k = 0; k < NW; ++k) {
        c += uint64_t(popc64(F.R[d][k]));
        F.R[d][k] = 0;
      }
    }
    int k0 = -1;
    for (int k = 0; k < NW; ++k)
      if (F.R[d][k]) {
        k0 = k;
        break;
      }
    if (k0 < 0) {
      if (d == d0) break;
      --d;
      F.used[F.f[d] >> 6] &= ~(uint64_t(1) << (F.f[d] & 63));
      continue;
    }
    const uint32_t v = uint32_t(k0 * 64 + ctz64(F.R[d][k0]));
    F.R[d][k0] &= F.R[d][k0] - 1;
    F.f[d] = v;
    if (d == n - 1) {
      ++c;
      if (leaf(v)) break;
      continue;
    }
    cs_nrow<NW>(dro, dci, s, v, F.NR[d]);
    F.used[v >> 6] |= uint64_t(1) << (v & 63);
    ++d;
    cs_candidates(F, d);
  }
  return c;
}

// Query 0's domains (depths 0..end-1) and adjacency inside data graph g.
template <int NW>
__device__ inline void cs_prototype(CsFrame<NW>& F, const uint64_t* __restrict__ cand, uint64_t W,
                                    const uint32_t* __restrict__ qro, const uint32_t* __restrict__ qci,
                                    uint32_t qb, uint32_t end, uint32_t s, uint32_t e) {
  for (uint32_t i = 0; i < end; ++i) {
    const uint64_t* row = cand + uint64_t(qb + i) * W;
    for (int k = 0; k < NW; ++k) F.dom[i][k] = row_bits(row, s, e, k);
    uint32_t a = 0;
    for (uint32_t k = qro[qb + i]; k < qro[qb + i + 1]; ++k) {
      const uint32_t j = qci[k] - qb;
      if (j < i) a |= 1u << j;
    }
    F.qadj[i] = a;
  }
  for (int k = 0; k < NW; ++k) F.used[k] = 0;
}

// joinPartialCandidates2 (isomorphism.hpp:1021-1178): one lane per data graph; pass 0 counts the
// embeddings of query 0's first `end` nodes (cnt[g], total), pass 1 writes them at off[g] as
// records {end global data node ids, data graph id} (records at or past cap are dropped).
template <int NW>
__global__ __launch_bounds__(64) void cs_partial_kernel(
    int pass, const uint64_t* __restrict__ cand, uint64_t W, const uint32_t* __restrict__ qgo,
    const uint32_t* __restrict__ qro, const uint32_t* __restrict__ qci,
    const uint32_t* __restrict__ dgo, uint32_t D, const uint32_t* __restrict__ dro,
    const uint32_t* __restrict__ dci, uint32_t end, uint32_t* __restrict__ cnt,
    const uint32_t* __restrict__ off, uint64_t* __restrict__ total, int32_t* __restrict__ out,
    uint64_t cap) {
  const uint64_t g = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  uint64_t c = 0;
  if (g < D) {
    const uint32_t s = dgo[g], e = dgo[g + 1];
    CsFrame<NW> F;
    cs_prototype<NW>(F, cand, W, qro, qci, qgo[0], end, s, e);
    if (pass == 0) {
      c = cs_dfs<NW>(F, 0, int(end), dro, dci, s, true, [](uint32_t) { return false; });
      cnt[g] = uint32_t(c);
    } else {
      uint64_t pos = off[g];
      cs_dfs<NW>(F, 0, int(end), dro, dci, s, false, [&](uint32_t) {
        if (pos < cap) {
          int32_t* r = out + pos * (end + 1);
          for (uint32_t i = 0; i < end; ++i) r[i] = int32_t(s + F.f[i]);
          r[end] = int32_t(g);
        }
        ++pos;
        return false;
      });
    }
  }
  if (pass == 0) {
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0 && c) atomicAdd(reinterpret_cast<unsigned long long*>(total), (unsigned long long)c);
  }
}

// joinWithPartialMatches (isomorphism.hpp:756-977): one lane per partial match; over the GMCR
// queries of its data graph, n_q == end counts the match, n_q > end counts its extensions to
// depths end..n_q-1 (q's domains, q's induced adjacency against every earlier depth, prefix
// included).  Per GMCR entry: counts added (count-all) or a 0 -> 1 flag (find_first, the
// pair_done flags of :836-866); the total is added to *total.
template <int NW>
__global__ __launch_bounds__(64) void cs_extend_kernel(
    const uint64_t* __restrict__ cand, uint64_t W, const uint32_t* __restrict__ qgo,
    const uint32_t* __restrict__ qro, const uint32_t* __restrict__ qci,
    const uint32_t* __restrict__ dgo, const uint32_t* __restrict__ dro,
    const uint32_t* __restrict__ dci, const uint32_t* __restrict__ goff,
    const uint32_t* __restrict__ gidx, uint32_t D, uint32_t end,
    const int32_t* __restrict__ partial, uint64_t num_partial, int find_first,
    uint64_t* __restrict__ per_pair, uint64_t* __restrict__ total) {
  const uint64_t m = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  uint64_t sum = 0;
  // a record naming a graph or node outside the batch (not one step 1 wrote) is skipped
  bool ok = m < num_partial;
  const int32_t* rec = partial + (ok ? m : 0) * (end + 1);
  const uint32_t g = ok ? uint32_t(rec[end]) : 0u;
  ok = ok && g < D;
  const uint32_t s = ok ? dgo[g] : 0u, e = ok ? dgo[g + 1] : 0u;
  for (uint32_t i = 0; i < end && ok; ++i) ok = uint32_t(rec[i]) >= s && uint32_t(rec[i]) < e;
  if (ok) {
    CsFrame<NW> F;
    for (int k = 0; k < NW; ++k) F.used[k] = 0;
    for (uint32_t i = 0; i < end; ++i) {
      const uint32_t x = uint32_t(rec[i]) - s;
      F.f[i] = x;
      F.used[x >> 6] |= uint64_t(1) << (x & 63);
      cs_nrow<NW>(dro, dci, s, x, F.NR[i]);
    }
    uint64_t used0[NW];
    for (int k = 0; k < NW; ++k) used0[k] = F.used[k];
    for (uint32_t t = goff[g]; t < goff[g + 1]; ++t) {
      if (find_first && __atomic_load_n(per_pair + t, __ATOMIC_RELAXED)) continue;
      const uint32_t q = gidx[t];
      const uint32_t qb = qgo[q], nq = qgo[q + 1] - qb;
      if (nq < end) continue;
      uint64_t c = 1;
      if (nq > end) {
        for (uint32_t i = 0; i < nq; ++i) {
          uint32_t a = 0;
          for (uint32_t k = qro[qb + i]; k < qro[qb + i + 1]; ++k) {
            const uint32_t j = qci[k] - qb;
            if (j < i) a |= 1u << j;
          }
          F.qadj[i] = a;
          if (i >= end) {
            const uint64_t* row = cand + uint64_t(qb + i) * W;
            for (int k = 0; k < NW; ++k) F.dom[i][k] = row_bits(row, s, e, k);
          }
        }
        for (int k = 0; k < NW; ++k) F.used[k] = used0[k];
        c = cs_dfs<NW>(F, int(end), int(nq), dro, dci, s, !find_first,
                       [&](uint32_t) { return find_first != 0; });
      }
      if (!c) continue;
      if (find_first) {
        if (atomicCAS(reinterpret_cast<unsigned long long*>(per_pair + t), 0ull, 1ull) == 0ull) ++sum;
      } else {
        atomicAdd(reinterpret_cast<unsigned long long*>(per_pair + t), (unsigned long long)c);
        sum += c;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if (lane_id() == 0 && sum) atomicAdd(reinterpret_cast<unsigned long long*>(total), (unsigned long long)sum);
}

inline uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

// Blocks of 256 threads for n threads of a wave-strided kernel: a multiple of 8 (whole rounds of
// the XCDs, for xcd_block) and at most 2^20.
inline uint32_t grid8(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return uint32_t(std::min<uint64_t>((b + 7) & ~uint64_t(7), 1u << 20));
}

}  // namespace bms2

using namespace bms2;

// ------------------------------------------------------------------ handles

struct bms2_signature_s {
  bms2_queue q;
  uint64_t n[2];
  uint64_t* sig[2] = {nullptr, nullptr};
  uint64_t* reach[2] = {nullptr, nullptr};
};

struct bms2_path_signature_s {
  bms2_queue q;
  uint64_t n[2];
  uint64_t* bits[2] = {nullptr, nullptr};
};

struct bms2_cycle_signature_s {
  bms2_queue q;
  uint64_t n[2];
  uint64_t* sig[2] = {nullptr, nullptr};
};

struct bms2_candidates_s {
  bms2_queue q;
  uint64_t rows, targets, W;
  uint64_t* bits = nullptr;
};

struct bms2_gmcr_s {
  bms2_queue q;
  uint32_t num_data = 0;
  uint64_t total = 0;
  uint32_t* offsets = nullptr;
  uint32_t* indices = nullptr;
};

namespace {

int check_scope(int scope) {
  return (scope == BMS2_SCOPE_DATA || scope == BMS2_SCOPE_QUERY)
             ? BMS2_OK
             : set_error(BMS2_ERR_INVALID, "scope must be BMS2_SCOPE_DATA or BMS2_SCOPE_QUERY");
}

int check_graph(const bms2_device_csr* g, uint64_t nodes) {
  if (!g) return set_error(BMS2_ERR_INVALID, "graph is null");
  if (g->total_nodes != nodes)
    return set_error(BMS2_ERR_INVALID, "graph has %llu nodes, the object was sized for %llu",
                     (unsigned long long)g->total_nodes, (unsigned long long)nodes);
  return BMS2_OK;
}

}  // namespace

namespace bms2 {

__global__ void fill_zero_kernel(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

// Exclusive scan of n u32 counts into offsets [n + 1] on stream st: one block for a few chunks,
// otherwise chunk sums, their scan and a rescan of every chunk from its offset (the chunk sums
// in stream-ordered memory; the single block when that allocation fails).  Errors surface
// through hipGetLastError.
void scan_counts(hipStream_t st, const uint32_t* cnt, uint64_t n, uint32_t* offsets) {
  const uint64_t chunks = (n + kScanChunk - 1) / kScanChunk;
  uint32_t* tmp = nullptr;
  if (chunks <= 4 || hipMallocAsync((void**)&tmp, 4 * (2 * chunks + 1), st) != hipSuccess) {
    (void)hipGetLastError();  // a failed allocation is not an error here
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, cnt, n, offsets);
    return;
  }
  hipLaunchKernelGGL(scan_sums_kernel, dim3(uint32_t(chunks)), dim3(1024), 0, st, cnt, n, tmp);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, tmp, chunks, tmp + chunks);
  hipLaunchKernelGGL(scan_add_kernel, dim3(uint32_t(chunks)), dim3(1024), 0, st, cnt, n,
                     tmp + chunks, chunks, offsets);
  (void)hipFreeAsync(tmp, st);
}

int staged_reset(bms2_signature sig, bms2_candidates cand) {
  hipStream_t st = sig->q->stream;
  for (int sc = 0; sc < 2; ++sc) {
    if (sig->n[sc]) {
      hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(sig->n[sc], 256, 4096)), dim3(256), 0, st,
                         sig->reach[sc], 4 * sig->n[sc]);
      hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(sig->n[sc], 256, 4096)), dim3(256), 0, st,
                         sig->sig[sc], sig->n[sc]);
    }
  }
  const uint64_t words = cand->rows * cand->W;
  if (words)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(words, 256, 1u << 16)), dim3(256), 0, st,
                       cand->bits, words);
  BMS2_CHECK_LAUNCH("fill_zero_kernel");
  return BMS2_OK;
}

namespace {

int device_cus() {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return cus;
}

// Depths of a lane's stash in the joins of graphs above 64 nodes: the batch's largest query.
uint32_t stash_depth(const bms2_device_csr* query) {
  return std::max<uint32_t>(2, std::min<uint32_t>(query->max_graph_nodes, kBigDepth));
}

// Bytes of the staged join's scratch; for graphs above 64 nodes with `waves` waves per CU in
// the size-class joins (their stashes: 2 x stash depth x row words x 8 bytes per lane).
uint64_t join_scratch_bytes(const bms2_device_csr* query, const bms2_device_csr* data, uint64_t T,
                            uint32_t waves, int cus, uint32_t rw_above = 0) {
  const uint64_t nrows = std::max<uint64_t>(1, data->total_nodes);
  const uint64_t qa = 4 * 32 * uint64_t(query->num_graphs);
  // graphs of up to 64 nodes: one adjacency row per data node, the entry cursor, 32
  // depth-adjacency words per query
  if (data->max_graph_nodes <= 64) return 8 * (nrows + 1) + qa;
  // larger graphs: rows of 2 or 4 words per node, the depth adjacency, class counters and
  // cursors (padded to 16 bytes), one run of entries per data graph (16 bytes each), the lanes'
  // stashes
  (void)T;
  const uint64_t rw = rw_above ? rw_above : data->max_graph_nodes > 128 ? 4 : 2;
  const uint64_t stash = uint64_t(cus) * waves * 64 * 2 * stash_depth(query) * rw * 8;
  const uint64_t qac = qa + 4 * uint64_t(query->num_graphs);  // and the depth chains
  const uint64_t head = (8 * rw * nrows + (qac + 7) / 8 * 8 + 8 * 3 * kBigClasses + 15) / 16 * 16;
  return head + 16 * std::max<uint64_t>(data->num_graphs, 1) + stash;
}

}  // namespace

uint64_t staged_join_scratch_bytes(const bms2_device_csr* query, const bms2_device_csr* data,
                                   uint64_t T) {
  return join_scratch_bytes(query, data, T, kBigWavesPerCU, device_cus());
}

int staged_join(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                bms2_candidates cand, bms2_gmcr gmcr, int find_first, const uint8_t* d_jorder,
                uint64_t* d_total, uint64_t* d_pair_counts, uint64_t* d_pair_nodes,
                void* scratch, uint64_t scratch_bytes, const BigJoinShape* shape) {
  const uint64_t T = gmcr->total;
  if (!T) return BMS2_OK;
  if (data->max_graph_nodes <= 64) {
    // adjacency rows, the entry cursor and the queries' depth adjacency in the caller's scratch
    // (or the queue's)
    hipStream_t st = q->stream;
    const uint64_t nrows = std::max<uint64_t>(1, data->total_nodes);
    const uint64_t nqa = uint64_t(query->num_graphs) * 32;
    uint64_t* rows = static_cast<uint64_t*>(scratch);
    if (!rows)
      if (int rc = queue_scratch(q, staged_join_scratch_bytes(query, data, T), (void**)&rows)) return rc;
    auto* cursor = reinterpret_cast<unsigned long long*>(rows + nrows);
    auto* qa = reinterpret_cast<uint32_t*>(rows + nrows + 1);
    if (nqa)
      hipLaunchKernelGGL(join_depth_adjacency_kernel, dim3(grid_for(nqa, 256)), dim3(256), 0, st,
                         query->graph_offsets, query->row_offsets, query->column_indices,
                         query->num_graphs, d_jorder, qa);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, st, rows + nrows, uint64_t(1));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && data->total_nodes)
      hipLaunchKernelGGL(adj_rows_kernel, dim3(grid_for(data->total_nodes, 256, 1u << 30)), dim3(256),
                         0, st, data->graph_offsets, data->num_graphs, data->row_offsets,
                         data->column_indices, data->total_nodes, rows);
    if (e == hipSuccess) e = hipGetLastError();
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    if (e == hipSuccess) {
      // waves per CU as LDS allows: 29 KB per wave for queries of up to 16 nodes, 58 KB for 32
      const bool q16 = query->max_graph_nodes <= 16;
      const dim3 grid(uint32_t(cus) * (q16 ? 5u : 2u));
      if (q16)
        hipLaunchKernelGGL(join_rows_kernel<16>, grid, dim3(64), 0, st, cand->bits, cand->W,
                           query->graph_offsets, qa, data->graph_offsets, data->num_graphs, rows,
                           gmcr->offsets, gmcr->indices, T, find_first, d_total, d_pair_counts,
                           d_jorder, d_pair_nodes, cursor);
      else
        hipLaunchKernelGGL(join_rows_kernel<BMS2_MAX_QUERY_NODES>, grid, dim3(64), 0, st, cand->bits,
                           cand->W, query->graph_offsets, qa, data->graph_offsets,
                           data->num_graphs, rows, gmcr->offsets, gmcr->indices, T, find_first,
                           d_total, d_pair_counts, d_jorder, d_pair_nodes, cursor);
      e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_fail(e, "join_rows_kernel", __FILE__, __LINE__);
    return BMS2_OK;
  }
  // graphs of 65-256 nodes: multi-word adjacency rows, the GMCR entries sorted into size classes
  // by row width, one persistent join per class present
  hipStream_t st = q->stream;
  const int cus = device_cus();
  const BigJoinShape sh = shape ? *shape : BigJoinShape{};
  const uint64_t nrows = std::max<uint64_t>(1, data->total_nodes);
  const uint32_t wide_above = sh.row_words == 4 ? 64 : 128;
  const uint32_t RW = data->max_graph_nodes > wide_above ? 4 : 2;
  const uint64_t nqa = uint64_t(query->num_graphs) * 32;
  const uint32_t bw = sh.block_waves == 4 ? 4 : sh.block_waves == 2 ? 2 : 1;
  uint32_t waves = std::max<uint32_t>(bw, std::min<uint32_t>(32, sh.waves ? sh.waves : kBigWavesPerCU)) / bw * bw;
  uint64_t* rows = static_cast<uint64_t*>(scratch);
  if (!rows) {
    if (int rc = queue_scratch(q, join_scratch_bytes(query, data, T, waves, cus, RW), (void**)&rows))
      return rc;
  } else {
    // a caller's scratch of scratch_bytes: as many waves per CU as their stashes fit
    while (waves > bw && join_scratch_bytes(query, data, T, waves, cus, RW) > scratch_bytes) waves -= bw;
    if (join_scratch_bytes(query, data, T, waves, cus, RW) > scratch_bytes)
      return set_error(BMS2_ERR_CAPACITY, "join scratch of %llu bytes is too small",
                       (unsigned long long)scratch_bytes);
  }
  auto* qa = reinterpret_cast<uint32_t*>(rows + RW * nrows);
  uint32_t* chain = qa + nqa;
  const uint64_t nqac = nqa + query->num_graphs;
  auto* ctl = reinterpret_cast<unsigned long long*>(rows + RW * nrows + (nqac + 1) / 2);
  const uint64_t head_words = (RW * nrows + (nqac + 1) / 2 + 3 * kBigClasses + 1) / 2 * 2;
  auto* runs = reinterpret_cast<uint4*>(rows + head_words);
  uint64_t* stash = rows + head_words + 2 * std::max<uint64_t>(data->num_graphs, 1);
  const uint32_t SD = stash_depth(query);
  if (nqa)
    hipLaunchKernelGGL(join_depth_adjacency_kernel, dim3(grid_for(nqa, 256)), dim3(256), 0, st,
                       query->graph_offsets, query->row_offsets, query->column_indices,
                       query->num_graphs, d_jorder, qa);
  if (query->num_graphs)
    hipLaunchKernelGGL(depth_chain_kernel, dim3(grid_for(query->num_graphs, 256)), dim3(256), 0, st,
                       query->graph_offsets, query->num_graphs, qa, chain);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint64_t*>(ctl),
                     uint64_t(3 * kBigClasses));
  hipLaunchKernelGGL(adj_rows_wide_kernel, dim3(grid_for(data->total_nodes, 256, 1u << 30)),
                     dim3(256), 0, st, data->graph_offsets, data->num_graphs, data->row_offsets,
                     data->column_indices, data->total_nodes, RW, rows);
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(big_classify_kernel, dim3(grid_for(data->num_graphs, 256, 1u << 30)),
                       dim3(256), 0, st, data->graph_offsets, data->num_graphs, gmcr->offsets,
                       pass, wide_above, ctl, runs);
  BMS2_CHECK_LAUNCH("big_classify_kernel");
  // units of one pair add their counts and DFS nodes
  if (d_pair_counts)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(T, 256, 4096)), dim3(256), 0, st, d_pair_counts, T);
  if (d_pair_nodes)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(T, 256, 4096)), dim3(256), 0, st, d_pair_nodes, T);
  BigJoinParams bp{cand->bits, cand->W, query->graph_offsets, qa, chain, gmcr->indices,
                   data->graph_offsets, rows, RW, runs, 0, ctl, stash, SD, find_first, d_total,
                   d_pair_counts, d_jorder, d_pair_nodes, nullptr, nullptr, nullptr, nullptr, 0,
                   nullptr};
  const dim3 grid(uint32_t(cus) * waves / bw), block(64 * bw);
  auto launch = [&](auto kern, int cls) { hipLaunchKernelGGL(kern, grid, block, 0, st, bp, cls); };
  if (bw == 4) {
    launch(join_big_kernel<2, 4>, 0);
    if (RW == 4) launch(join_big_kernel<4, 4>, 1);
  } else if (bw == 2) {
    launch(join_big_kernel<2, 2>, 0);
    if (RW == 4) launch(join_big_kernel<4, 2>, 1);
  } else {
    launch(join_big_kernel<2, 1>, 0);
    if (RW == 4) launch(join_big_kernel<4, 1>, 1);
  }
  BMS2_CHECK_LAUNCH("join_big_kernel");
  return BMS2_OK;
}

const uint64_t* path_signature_bits(bms2_path_signature s, int scope) { return s->bits[scope]; }
const uint64_t* cycle_signature_values(bms2_cycle_signature s, int scope) { return s->sig[scope]; }

namespace {

// Launch of the size-class joins under a shape: waves per workgroup and waves per CU.
void big_launch_shape(const BigJoinShape* shape, uint32_t* bw, uint32_t* waves) {
  const BigJoinShape sh = shape ? *shape : BigJoinShape{};
  *bw = sh.block_waves == 4 ? 4 : sh.block_waves == 2 ? 2 : 1;
  *waves = std::max<uint32_t>(*bw, std::min<uint32_t>(32, sh.waves ? sh.waves : kBigWavesPerCU)) / *bw * *bw;
}

}  // namespace

uint64_t big_join_scratch_bytes(const bms2_device_csr* query, const bms2_device_csr* data,
                                const BigJoinShape* shape) {
  uint32_t bw = 1, waves = 1;
  big_launch_shape(shape, &bw, &waves);
  const uint32_t wide_above = big_wide_above(shape ? *shape : BigJoinShape{});
  const uint64_t RW = data->max_graph_nodes > wide_above ? 4 : 2;
  const uint64_t nrows = std::max<uint64_t>(1, data->total_nodes);
  const uint64_t qa = (4 * 33 * uint64_t(query->num_graphs) + 7) / 8 * 8;  // with the chains
  return 8 * RW * nrows + qa + uint64_t(device_cus()) * waves * 64 * 2 * stash_depth(query) * RW * 8;
}

int big_join(bms2_queue q, bms2_queue q_wide, const bms2_device_csr* query,
             const bms2_device_csr* data, const BigEntries& e, const uint64_t count[kBigClasses],
             const uint64_t runs[kBigClasses], int find_first, const uint8_t* d_jorder,
             uint64_t* d_counts, const uint32_t* d_gid, uint64_t* d_stats,
             const BigJoinShape* shape) {
  if (!count[0] && !count[1]) return BMS2_OK;
  hipStream_t st = q->stream;
  uint32_t bw = 1, waves = 1;
  big_launch_shape(shape, &bw, &waves);
  const uint32_t wide_above = big_wide_above(shape ? *shape : BigJoinShape{});
  const uint32_t RW = data->max_graph_nodes > wide_above ? 4 : 2;
  if (count[1] && RW != 4)
    return set_error(BMS2_ERR_INVALID, "entries of 4-word graphs in a batch of 2-word graphs");
  const uint64_t nrows = std::max<uint64_t>(1, data->total_nodes);
  const uint64_t nqa = uint64_t(query->num_graphs) * 32;
  uint64_t* rows = nullptr;
  if (int rc = queue_scratch(q, big_join_scratch_bytes(query, data, shape), (void**)&rows)) return rc;
  auto* qa = reinterpret_cast<uint32_t*>(rows + RW * nrows);
  uint32_t* chain = qa + nqa;
  uint64_t* stash = rows + RW * nrows + (nqa + query->num_graphs + 1) / 2;
  if (nqa) {
    hipLaunchKernelGGL(join_depth_adjacency_kernel, dim3(grid_for(nqa, 256)), dim3(256), 0, st,
                       query->graph_offsets, query->row_offsets, query->column_indices,
                       query->num_graphs, d_jorder, qa);
    hipLaunchKernelGGL(depth_chain_kernel, dim3(grid_for(query->num_graphs, 256)), dim3(256), 0, st,
                       query->graph_offsets, query->num_graphs, qa, chain);
  }
  hipLaunchKernelGGL(adj_rows_wide_kernel, dim3(grid_for(data->total_nodes, 256, 1u << 30)),
                     dim3(256), 0, st, data->graph_offsets, data->num_graphs, data->row_offsets,
                     data->column_indices, data->total_nodes, RW, rows);
  BMS2_CHECK_LAUNCH("adj_rows_wide_kernel");
  BigJoinParams bp{nullptr, 0, query->graph_offsets, qa, chain, e.gidx, data->graph_offsets, rows, RW,
                   e.runs, e.cap_r, e.ctl, stash, stash_depth(query), find_first, nullptr, nullptr,
                   d_jorder, nullptr, e.rec_off, e.recs, d_counts, d_gid, query->num_graphs,
                   d_stats};
  const uint32_t blocks = uint32_t(device_cus()) * waves / bw;
  const dim3 block(64 * bw);
  // the 4-word class, when both classes have entries and a second queue is given, runs beside
  // the 2-word class: its few graphs' long searches overlap the other class's instead of
  // following them; one workgroup per run at most, its stashes in that queue's scratch
  const bool beside = q_wide && count[0] && count[1];
  BigJoinParams wp = bp;
  uint32_t wide_blocks = blocks;
  hipEvent_t fork = nullptr, done = nullptr;
  if (beside) {
    wide_blocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(blocks, (runs[1] + bw - 1) / bw)));
    const uint64_t bytes = uint64_t(wide_blocks) * bw * 64 * 2 * stash_depth(query) * 4 * 8;
    if (int rc = queue_scratch(q_wide, bytes, (void**)&wp.stash)) return rc;
    BMS2_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    hipError_t erc = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (erc == hipSuccess) erc = hipEventRecord(fork, st);
    if (erc == hipSuccess) erc = hipStreamWaitEvent(q_wide->stream, fork, 0);
    if (erc != hipSuccess) {
      (void)hipEventDestroy(fork);
      if (done) (void)hipEventDestroy(done);
      return hip_fail(erc, "fork of the 4-word class", __FILE__, __LINE__);
    }
  }
  auto launch = [&](auto k2, auto k4) {
    if (count[1])
      hipLaunchKernelGGL(k4, dim3(wide_blocks), block, 0, beside ? q_wide->stream : st, wp, 1);
    if (count[0]) hipLaunchKernelGGL(k2, dim3(blocks), block, 0, st, bp, 0);
  };
  if (bw == 4)
    launch(join_big_kernel<2, 4>, join_big_kernel<4, 4>);
  else if (bw == 2)
    launch(join_big_kernel<2, 2>, join_big_kernel<4, 2>);
  else
    launch(join_big_kernel<2, 1>, join_big_kernel<4, 1>);
  hipError_t erc = hipGetLastError();
  if (beside) {
    if (erc == hipSuccess) erc = hipEventRecord(done, q_wide->stream);
    if (erc == hipSuccess) erc = hipStreamWaitEvent(st, done, 0);
    (void)hipEventDestroy(fork);
    (void)hipEventDestroy(done);
  }
  if (erc != hipSuccess) return hip_fail(erc, "join_big_kernel", __FILE__, __LINE__);
  return BMS2_OK;
}

}  // namespace bms2

extern "C" {

// ---- Signature
int bms2_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                          bms2_signature* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* s = new bms2_signature_s{};
  s->q = q;
  s->n[0] = data_nodes;
  s->n[1] = query_nodes;
  for (int sc = 0; sc < 2; ++sc) {
    int rc;
    if ((rc = device_alloc((void**)&s->sig[sc], 8 * s->n[sc])) ||
        (rc = device_alloc((void**)&s->reach[sc], 32 * s->n[sc]))) {
      bms2_signature_destroy(s);
      return rc;
    }
    BMS2_HIP(hipMemsetAsync(s->sig[sc], 0, 8 * s->n[sc], q->stream));
    BMS2_HIP(hipMemsetAsync(s->reach[sc], 0, 32 * s->n[sc], q->stream));
  }
  *out = s;
  return BMS2_OK;
}

uint64_t bms2_signature_alloc_size(bms2_signature s, int scope) {
  // label histogram (8 B) + the four u64 planes of counted nodes (32 B) per node
  return (s && (scope == BMS2_SCOPE_DATA || scope == BMS2_SCOPE_QUERY)) ? 40 * s->n[scope] : 0;
}

uint64_t bms2_path_signature_alloc_size(bms2_path_signature s, int scope) {
  return (s && (scope == BMS2_SCOPE_DATA || scope == BMS2_SCOPE_QUERY)) ? 32 * s->n[scope] : 0;
}

uint64_t bms2_cycle_signature_alloc_size(bms2_cycle_signature s, int scope) {
  return (s && (scope == BMS2_SCOPE_DATA || scope == BMS2_SCOPE_QUERY)) ? 8 * s->n[scope] : 0;
}

int bms2_signature_destroy(bms2_signature s) {
  if (!s) return BMS2_OK;
  (void)hipSetDevice(s->q->device);
  (void)hipStreamSynchronize(s->q->stream);
  for (int sc = 0; sc < 2; ++sc) {
    (void)hipFree(s->sig[sc]);
    (void)hipFree(s->reach[sc]);
  }
  delete s;
  return BMS2_OK;
}

int bms2_signature_generate(bms2_signature s, const bms2_device_csr* g, int scope, bms2_event* ev) {
  if (!s) return set_error(BMS2_ERR_INVALID, "signature is null");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (int rc = check_graph(g, s->n[scope])) return rc;
  EventScope es{s->q, ev};
  if (int rc = es.begin()) return rc;
  if (g->total_nodes)
    hipLaunchKernelGGL(label_signature_kernel, dim3(grid_for(g->total_nodes, 256)), dim3(256), 0,
                       s->q->stream, g->row_offsets, g->column_indices, g->node_labels,
                       g->total_nodes, s->sig[scope]);
  BMS2_CHECK_LAUNCH("label_signature_kernel");
  return es.end();
}

int bms2_signature_refine(bms2_signature s, const bms2_device_csr* g, uint32_t view_size, int scope,
                          bms2_event* ev) {
  if (!s) return set_error(BMS2_ERR_INVALID, "signature is null");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (int rc = check_graph(g, s->n[scope])) return rc;
  if (g->max_graph_nodes > BMS2_MAX_GRAPH_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "graph of %u nodes", g->max_graph_nodes);
  EventScope es{s->q, ev};
  if (int rc = es.begin()) return rc;
  if (g->total_nodes)
    hipLaunchKernelGGL(refine_signature_kernel, dim3(grid_for(g->total_nodes, 256)), dim3(256), 0,
                       s->q->stream, g->graph_offsets, g->num_graphs, g->row_offsets,
                       g->column_indices, g->node_labels, g->total_nodes, view_size,
                       s->sig[scope], s->reach[scope]);
  BMS2_CHECK_LAUNCH("refine_signature_kernel");
  return es.end();
}

int bms2_signature_copy_to_host(bms2_signature s, int scope, uint64_t* out) {
  if (!s || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (s->n[scope]) {
    BMS2_HIP(hipMemcpyAsync(out, s->sig[scope], 8 * s->n[scope], hipMemcpyDeviceToHost, s->q->stream));
    BMS2_HIP(hipStreamSynchronize(s->q->stream));
  }
  return BMS2_OK;
}

const uint64_t* bms2_signature_device_ptr(bms2_signature s, int scope) {
  if (!s || check_scope(scope)) return nullptr;
  return s->sig[scope];
}

// ---- PathSignature
int bms2_path_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                               bms2_path_signature* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* s = new bms2_path_signature_s{};
  s->q = q;
  s->n[0] = data_nodes;
  s->n[1] = query_nodes;
  for (int sc = 0; sc < 2; ++sc) {
    if (int rc = device_alloc((void**)&s->bits[sc], 32 * s->n[sc])) {
      bms2_path_signature_destroy(s);
      return rc;
    }
    BMS2_HIP(hipMemsetAsync(s->bits[sc], 0, 32 * s->n[sc], q->stream));
  }
  *out = s;
  return BMS2_OK;
}

int bms2_path_signature_destroy(bms2_path_signature s) {
  if (!s) return BMS2_OK;
  (void)hipSetDevice(s->q->device);
  (void)hipStreamSynchronize(s->q->stream);
  for (int sc = 0; sc < 2; ++sc) (void)hipFree(s->bits[sc]);
  delete s;
  return BMS2_OK;
}

int bms2_path_signature_generate(bms2_path_signature s, const bms2_device_csr* g, int scope,
                                 bms2_event* ev) {
  if (!s) return set_error(BMS2_ERR_INVALID, "path signature is null");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (int rc = check_graph(g, s->n[scope])) return rc;
  EventScope es{s->q, ev};
  if (int rc = es.begin()) return rc;
  if (g->total_nodes)
    hipLaunchKernelGGL(path_signature_kernel, dim3(grid_for(g->total_nodes, 256)), dim3(256), 0,
                       s->q->stream, g->row_offsets, g->column_indices, g->node_labels,
                       g->total_nodes, s->bits[scope]);
  BMS2_CHECK_LAUNCH("path_signature_kernel");
  return es.end();
}

int bms2_path_signature_copy_to_host(bms2_path_signature s, int scope, uint64_t* out) {
  if (!s || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (s->n[scope]) {
    BMS2_HIP(hipMemcpyAsync(out, s->bits[scope], 32 * s->n[scope], hipMemcpyDeviceToHost, s->q->stream));
    BMS2_HIP(hipStreamSynchronize(s->q->stream));
  }
  return BMS2_OK;
}

// ---- CycleSignature
int bms2_cycle_signature_create(bms2_queue q, uint64_t data_nodes, uint64_t query_nodes,
                                bms2_cycle_signature* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* s = new bms2_cycle_signature_s{};
  s->q = q;
  s->n[0] = data_nodes;
  s->n[1] = query_nodes;
  for (int sc = 0; sc < 2; ++sc) {
    if (int rc = device_alloc((void**)&s->sig[sc], 8 * s->n[sc])) {
      bms2_cycle_signature_destroy(s);
      return rc;
    }
    BMS2_HIP(hipMemsetAsync(s->sig[sc], 0, 8 * s->n[sc], q->stream));
  }
  *out = s;
  return BMS2_OK;
}

int bms2_cycle_signature_destroy(bms2_cycle_signature s) {
  if (!s) return BMS2_OK;
  (void)hipSetDevice(s->q->device);
  (void)hipStreamSynchronize(s->q->stream);
  for (int sc = 0; sc < 2; ++sc) (void)hipFree(s->sig[sc]);
  delete s;
  return BMS2_OK;
}

int bms2_cycle_signature_generate(bms2_cycle_signature s, const bms2_device_csr* g, int scope,
                                  bms2_event* ev) {
  if (!s) return set_error(BMS2_ERR_INVALID, "cycle signature is null");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (int rc = check_graph(g, s->n[scope])) return rc;
  if (g->max_degree > 27)
    return set_error(BMS2_ERR_UNSUPPORTED,
                     "node degree %u above 27: the reference's 128-entry ring-count stack could "
                     "overflow (signature.hpp:638,697)",
                     g->max_degree);
  EventScope es{s->q, ev};
  if (int rc = es.begin()) return rc;
  if (g->total_nodes)
    hipLaunchKernelGGL(cycle_signature_kernel, dim3(grid_for(g->total_nodes, 256)), dim3(256), 0,
                       s->q->stream, g->graph_offsets, g->num_graphs, g->row_offsets,
                       g->column_indices, g->node_labels, g->total_nodes, s->sig[scope]);
  BMS2_CHECK_LAUNCH("cycle_signature_kernel");
  return es.end();
}

int bms2_cycle_signature_copy_to_host(bms2_cycle_signature s, int scope, uint64_t* out) {
  if (!s || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = check_scope(scope)) return rc;
  if (int rc = use_device(s->q)) return rc;
  if (s->n[scope]) {
    BMS2_HIP(hipMemcpyAsync(out, s->sig[scope], 8 * s->n[scope], hipMemcpyDeviceToHost, s->q->stream));
    BMS2_HIP(hipStreamSynchronize(s->q->stream));
  }
  return BMS2_OK;
}

// ---- Candidates
int bms2_candidates_create(bms2_queue q, uint64_t source_nodes, uint64_t target_nodes,
                           bms2_candidates* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* c = new bms2_candidates_s{};
  c->q = q;
  c->rows = source_nodes;
  c->targets = target_nodes;
  c->W = ceil_div(target_nodes, 64);
  if (int rc = device_alloc((void**)&c->bits, 8 * c->rows * c->W)) {
    delete c;
    return rc;
  }
  BMS2_HIP(hipMemsetAsync(c->bits, 0, 8 * c->rows * c->W, q->stream));
  *out = c;
  return BMS2_OK;
}

int bms2_candidates_destroy(bms2_candidates c) {
  if (!c) return BMS2_OK;
  (void)hipSetDevice(c->q->device);
  (void)hipStreamSynchronize(c->q->stream);
  (void)hipFree(c->bits);
  delete c;
  return BMS2_OK;
}

uint64_t bms2_candidates_words_per_row(bms2_candidates c) { return c ? c->W : 0; }

int bms2_candidates_copy_to_host(bms2_candidates c, uint64_t* out) {
  if (!c || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = use_device(c->q)) return rc;
  if (c->rows * c->W) {
    BMS2_HIP(hipMemcpyAsync(out, c->bits, 8 * c->rows * c->W, hipMemcpyDeviceToHost, c->q->stream));
    BMS2_HIP(hipStreamSynchronize(c->q->stream));
  }
  return BMS2_OK;
}

int bms2_candidates_copy_rows(bms2_candidates c, uint64_t first_row, uint64_t num_rows,
                              uint64_t* out) {
  if (!c || (num_rows && !out)) return set_error(BMS2_ERR_INVALID, "null argument");
  if (first_row > c->rows || num_rows > c->rows - first_row)
    return set_error(BMS2_ERR_INVALID, "rows [%llu, %llu) outside the %llu rows",
                     (unsigned long long)first_row, (unsigned long long)(first_row + num_rows),
                     (unsigned long long)c->rows);
  if (int rc = use_device(c->q)) return rc;
  if (num_rows * c->W) {
    BMS2_HIP(hipMemcpyAsync(out, c->bits + first_row * c->W, 8 * num_rows * c->W,
                            hipMemcpyDeviceToHost, c->q->stream));
    BMS2_HIP(hipStreamSynchronize(c->q->stream));
  }
  return BMS2_OK;
}

int bms2_candidates_row_counts(bms2_candidates c, uint32_t* out) {
  if (!c || !out) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = use_device(c->q)) return rc;
  if (!c->rows) return BMS2_OK;
  uint32_t* d = nullptr;
  if (int rc = device_alloc((void**)&d, 4 * c->rows)) return rc;
  hipLaunchKernelGGL(row_count_kernel, dim3(grid_for(c->rows * 64, 256, 1u << 30)), dim3(256), 0,
                     c->q->stream, c->bits, c->rows, c->W, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, 4 * c->rows, hipMemcpyDeviceToHost, c->q->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->q->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "row counts", __FILE__, __LINE__);
  return BMS2_OK;
}

// ---- filter / refine
int bms2_filter_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                           bms2_signature sig, bms2_path_signature path,
                           bms2_cycle_signature cycle, bms2_candidates cand, bms2_event* ev) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !sig || !path || !cycle || !cand)
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (cand->rows != query->total_nodes || cand->targets != data->total_nodes)
    return set_error(BMS2_ERR_INVALID, "candidates sized %llux%llu for a %llux%llu problem",
                     (unsigned long long)cand->rows, (unsigned long long)cand->targets,
                     (unsigned long long)query->total_nodes, (unsigned long long)data->total_nodes);
  if (path->n[0] != data->total_nodes || path->n[1] != query->total_nodes ||
      cycle->n[0] != data->total_nodes || cycle->n[1] != query->total_nodes)
    return set_error(BMS2_ERR_INVALID, "signature objects sized for another problem");
  (void)sig;  // the label histogram is not part of the initial filter (isomorphism.hpp:71)
  EventScope es{q, ev};
  if (int rc = es.begin()) return rc;
  if (cand->W && query->total_nodes) {
    // the query records live in the queue's scratch until the filter has run
    FilterQuery* fq = nullptr;
    if (int rc = queue_scratch(q, sizeof(FilterQuery) * query->total_nodes, (void**)&fq)) return rc;
    hipLaunchKernelGGL(filter_queries_kernel, dim3(grid_for(query->total_nodes, 256)), dim3(256), 0,
                       q->stream, query->row_offsets, query->node_labels, query->total_nodes,
                       path->bits[1], cycle->sig[1], fq);
    BMS2_CHECK_LAUNCH("filter_queries_kernel");
    hipLaunchKernelGGL(filter_kernel, dim3(grid8(cand->W * 64)), dim3(256), 0, q->stream, fq,
                       query->total_nodes, data->row_offsets, data->node_labels,
                       data->total_nodes, path->bits[0], cycle->sig[0], cand->bits, cand->W);
  }
  BMS2_CHECK_LAUNCH("filter_kernel");
  return es.end();
}

int bms2_refine_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                           bms2_signature sig, bms2_candidates cand, bms2_event* ev) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !sig || !cand) return set_error(BMS2_ERR_INVALID, "null argument");
  if (cand->rows != query->total_nodes || cand->targets != data->total_nodes ||
      sig->n[0] != data->total_nodes || sig->n[1] != query->total_nodes)
    return set_error(BMS2_ERR_INVALID, "objects sized for another problem");
  EventScope es{q, ev};
  if (int rc = es.begin()) return rc;
  if (cand->W && cand->rows) {
    // the signature planes live in the queue's scratch until the refinement has run
    uint64_t* planes = nullptr;
    if (int rc = queue_scratch(q, 64 * 8 * cand->W, (void**)&planes)) return rc;
    hipLaunchKernelGGL(sig_planes_kernel, dim3(grid_for(cand->W * 64, 256)), dim3(256), 0,
                       q->stream, sig->sig[0], data->total_nodes, cand->W, planes);
    BMS2_CHECK_LAUNCH("sig_planes_kernel");
    const uint64_t blocks = (cand->W + 63) / 64 * ((cand->rows + kRefineRun - 1) / kRefineRun);
    if (blocks * 64 >= (1ull << 32))
      return set_error(BMS2_ERR_UNSUPPORTED, "candidate matrix of %llu x %llu words is too large",
                       (unsigned long long)cand->rows, (unsigned long long)cand->W);
    hipLaunchKernelGGL(refine_candidates_kernel, dim3(uint32_t(blocks)), dim3(64), 0, q->stream,
                       sig->sig[1], planes, data->total_nodes, cand->bits, cand->W, cand->rows);
  }
  BMS2_CHECK_LAUNCH("refine_candidates_kernel");
  return es.end();
}

// ---- GMCR
int bms2_gmcr_create(bms2_queue q, bms2_gmcr* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* g = new bms2_gmcr_s{};
  g->q = q;
  *out = g;
  return BMS2_OK;
}

int bms2_gmcr_destroy(bms2_gmcr g) {
  if (!g) return BMS2_OK;
  (void)hipSetDevice(g->q->device);
  (void)hipStreamSynchronize(g->q->stream);
  (void)hipFree(g->offsets);
  (void)hipFree(g->indices);
  delete g;
  return BMS2_OK;
}

int bms2_gmcr_generate(bms2_gmcr gm, const bms2_device_csr* query, const bms2_device_csr* data,
                       bms2_candidates cand, bms2_event* ev) {
  if (!gm) return set_error(BMS2_ERR_INVALID, "gmcr is null");
  if (int rc = use_device(gm->q)) return rc;
  if (!query || !data || !cand) return set_error(BMS2_ERR_INVALID, "null argument");
  if (cand->rows != query->total_nodes || cand->targets != data->total_nodes)
    return set_error(BMS2_ERR_INVALID, "candidates sized for another problem");
  hipStream_t s = gm->q->stream;
  const uint32_t D = data->num_graphs, Q = query->num_graphs;
  const uint64_t VW = ceil_div(Q, 64);
  (void)hipFree(gm->offsets);
  (void)hipFree(gm->indices);
  gm->offsets = gm->indices = nullptr;
  gm->total = 0;
  gm->num_data = D;
  uint64_t* vbits = nullptr;
  uint32_t* cnt = nullptr;
  int rc;
  if ((rc = device_alloc((void**)&gm->offsets, 4ull * (D + 1))) ||
      (rc = device_alloc((void**)&vbits, 8 * std::max<uint64_t>(1, uint64_t(D) * VW))) ||
      (rc = device_alloc((void**)&cnt, 4ull * std::max<uint32_t>(D, 1)))) {
    (void)hipFree(vbits);
    (void)hipFree(cnt);
    return rc;
  }
  EventScope es{gm->q, ev};
  if ((rc = es.begin())) return rc;
  hipError_t e = hipSuccess;
  if (D) {
    if (Q)
      hipLaunchKernelGGL(gmcr_count_kernel, dim3(grid_for(uint64_t(D) * 64, 256, 1u << 30)), dim3(256),
                         0, s, cand->bits, cand->W, query->graph_offsets, Q, data->graph_offsets, D,
                         vbits, VW, cnt);
    else
      e = hipMemsetAsync(cnt, 0, 4ull * D, s);
    scan_counts(s, cnt, uint64_t(D), gm->offsets);
  } else {
    e = hipMemsetAsync(gm->offsets, 0, 4, s);
  }
  if (e == hipSuccess) e = hipGetLastError();
  uint32_t total = 0;
  // The total sizes the index array (gmcr.hpp:99 reads it back the same way).  Offsets are u32
  // as in the reference; when D x Q could exceed that, the exact 64-bit sum of the per-graph
  // counts decides whether the batch fits before anything is sized from a wrapped total.
  if (e == hipSuccess) e = hipMemcpyAsync(&total, gm->offsets + D, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess && uint64_t(D) * Q > UINT32_MAX) {
    std::vector<uint32_t> hc(D);
    e = hipMemcpy(hc.data(), cnt, 4ull * D, hipMemcpyDeviceToHost);
    uint64_t sum = 0;
    for (uint32_t v : hc) sum += v;
    if (e == hipSuccess && sum > UINT32_MAX) {
      (void)hipFree(vbits);
      (void)hipFree(cnt);
      (void)hipFree(gm->offsets);
      gm->offsets = nullptr;
      return set_error(BMS2_ERR_UNSUPPORTED,
                       "%llu GMCR pairs exceed the u32 offsets of the mapping (gmcr.hpp:54-127); "
                       "split the data batch", (unsigned long long)sum);
    }
  }
  if (e == hipSuccess) {
    gm->total = total;
    if ((rc = device_alloc((void**)&gm->indices, 4ull * std::max<uint32_t>(total, 1)))) {
      (void)hipFree(vbits);
      (void)hipFree(cnt);
      return rc;
    }
    if (D && Q && total)
      hipLaunchKernelGGL(gmcr_fill_kernel, dim3(grid_for(uint64_t(D) * 64, 256, 1u << 30)), dim3(256), 0,
                         s, vbits, VW, Q, D, gm->offsets, gm->indices);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(vbits);
  (void)hipFree(cnt);
  if (e != hipSuccess) return hip_fail(e, "generateGMCR", __FILE__, __LINE__);
  return es.end();
}

uint64_t bms2_gmcr_total(bms2_gmcr g) { return g ? g->total : 0; }

int bms2_gmcr_device_ptrs(bms2_gmcr g, const uint32_t** offsets, const uint32_t** indices) {
  if (!g) return set_error(BMS2_ERR_INVALID, "gmcr is null");
  if (offsets) *offsets = g->offsets;
  if (indices) *indices = g->indices;
  return BMS2_OK;
}

int bms2_gmcr_copy_to_host(bms2_gmcr g, uint32_t* offsets, uint32_t* indices) {
  if (!g) return set_error(BMS2_ERR_INVALID, "gmcr is null");
  if (!g->offsets) return set_error(BMS2_ERR_INVALID, "gmcr not generated");
  if (int rc = use_device(g->q)) return rc;
  hipStream_t s = g->q->stream;
  if (offsets) BMS2_HIP(hipMemcpyAsync(offsets, g->offsets, 4ull * (g->num_data + 1), hipMemcpyDeviceToHost, s));
  if (indices && g->total)
    BMS2_HIP(hipMemcpyAsync(indices, g->indices, 4ull * g->total, hipMemcpyDeviceToHost, s));
  BMS2_HIP(hipStreamSynchronize(s));
  return BMS2_OK;
}

// ---- join
int bms2_join_candidates(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
                         bms2_candidates cand, bms2_gmcr gmcr, uint64_t* d_num_matches,
                         int find_first, uint64_t* d_pair_counts, uint64_t pair_capacity,
                         bms2_event* ev) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !cand || !gmcr || !d_num_matches)
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (!gmcr->offsets) return set_error(BMS2_ERR_INVALID, "gmcr not generated");
  if (gmcr->num_data != data->num_graphs) return set_error(BMS2_ERR_INVALID, "gmcr of another batch");
  if (d_pair_counts && pair_capacity < gmcr->total)
    return set_error(BMS2_ERR_CAPACITY, "per-pair counts hold %llu entries, the GMCR has %llu",
                     (unsigned long long)pair_capacity, (unsigned long long)gmcr->total);
  if (query->max_graph_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "query with %u nodes (limit %u, isomorphism.hpp:248)",
                     query->max_graph_nodes, BMS2_MAX_QUERY_NODES);
  EventScope es{q, ev};
  if (int rc = es.begin()) return rc;
  if (int rc = staged_join(q, query, data, cand, gmcr, find_first, nullptr, d_num_matches,
                           d_pair_counts, nullptr, nullptr, 0, nullptr))
    return rc;
  return es.end();
}

}  // extern "C"

namespace {
// The join's scratch for bms2_join_candidates_noalloc: one array per GPU, held for the life of
// the process.  Host code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's own name is a host-side stand-in).
constexpr uint64_t kNoAllocScratchBytes = BMS2_JOIN_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t join_scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[256] = {};

int noalloc_scratch(int device, void** out) {
  if (device < 0 || device >= 256) return set_error(BMS2_ERR_INVALID, "device %d", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(join_scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(BMS2_ERR_HIP, "hipGetSymbolAddress of the join scratch failed: %s",
                       hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return BMS2_OK;
}
}  // namespace

extern "C" {

int bms2_join_candidates_noalloc(bms2_queue q, const bms2_device_csr* query,
                                 const bms2_device_csr* data, bms2_candidates cand, bms2_gmcr gmcr,
                                 uint64_t* d_num_matches, int find_first, uint64_t* d_pair_counts,
                                 uint64_t pair_capacity) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !cand || !gmcr || !d_num_matches)
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (!gmcr->offsets) return set_error(BMS2_ERR_INVALID, "gmcr not generated");
  if (gmcr->num_data != data->num_graphs) return set_error(BMS2_ERR_INVALID, "gmcr of another batch");
  if (d_pair_counts && pair_capacity < gmcr->total)
    return set_error(BMS2_ERR_CAPACITY, "per-pair counts hold %llu entries, the GMCR has %llu",
                     (unsigned long long)pair_capacity, (unsigned long long)gmcr->total);
  if (query->max_graph_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "query with %u nodes (limit %u, isomorphism.hpp:248)",
                     query->max_graph_nodes, BMS2_MAX_QUERY_NODES);
  const uint64_t need = join_scratch_bytes(query, data, gmcr->total, 1, device_cus());
  if (need > kNoAllocScratchBytes)
    return set_error(BMS2_ERR_CAPACITY,
                     "the join needs %llu bytes of scratch for %llu data nodes, %u queries and %llu "
                     "GMCR pairs; the allocation-free join holds %llu (8 bytes per data node + 128 "
                     "per query for graphs of up to 64 nodes; 16 or 32 bytes per data node + 128 "
                     "per query + 8 per GMCR pair above)",
                     (unsigned long long)need, (unsigned long long)data->total_nodes,
                     query->num_graphs, (unsigned long long)gmcr->total,
                     (unsigned long long)kNoAllocScratchBytes);
  void* scratch = nullptr;
  if (int rc = noalloc_scratch(q->device, &scratch)) return rc;
  return staged_join(q, query, data, cand, gmcr, find_first, nullptr, d_num_matches,
                     d_pair_counts, nullptr, scratch, kNoAllocScratchBytes, nullptr);
}

// ---- GPU-side CSR packing
int bms2_device_csr_pack(bms2_queue q, uint32_t num_graphs, const uint32_t* nodes,
                         const uint32_t* ro_local, const uint32_t* ci_local,
                         const uint8_t* node_labels, const uint8_t* edge_labels,
                         uint64_t total_nodes, uint64_t total_edges, bms2_device_csr* out) {
  if (int rc = use_device(q)) return rc;
  if (!out || (num_graphs && !nodes) || (total_nodes && (!ro_local || !node_labels)) ||
      (total_edges && !ci_local))
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (total_nodes >= (1ull << 32) || total_edges >= (1ull << 32))
    return set_error(BMS2_ERR_UNSUPPORTED, "batch exceeds 2^32-1 nodes or CSR entries (u32 ids)");
  uint64_t sum = 0;
  for (uint32_t g = 0; g < num_graphs; ++g) sum += nodes[g];
  if (sum != total_nodes)
    return set_error(BMS2_ERR_INVALID, "node counts sum to %llu, total_nodes is %llu",
                     (unsigned long long)sum, (unsigned long long)total_nodes);
  bms2_device_csr d{};
  d.num_graphs = num_graphs;
  d.total_nodes = total_nodes;
  d.total_edges = total_edges;
  uint32_t* t_nodes = nullptr;
  uint32_t* t_ro = nullptr;
  uint32_t* t_ci = nullptr;
  uint32_t* t_ent = nullptr;
  uint32_t* t_eoff = nullptr;
  uint32_t* t_stat = nullptr;
  uint32_t* t_ng = nullptr;
  hipStream_t st = q->stream;
  int rc = BMS2_OK;
  auto cleanup = [&]() {
    for (void* p : {(void*)t_nodes, (void*)t_ro, (void*)t_ci, (void*)t_ent, (void*)t_eoff,
                    (void*)t_stat, (void*)t_ng})
      (void)hipFree(p);
  };
  if ((rc = device_alloc((void**)&d.graph_offsets, 4ull * (num_graphs + 1))) ||
      (rc = device_alloc((void**)&d.row_offsets, 4ull * (total_nodes + 1))) ||
      (rc = device_alloc((void**)&d.column_indices, 4ull * total_edges)) ||
      (rc = device_alloc((void**)&d.node_labels, total_nodes)) ||
      (rc = device_alloc((void**)&d.edge_labels, total_edges)) ||
      (rc = device_alloc((void**)&t_nodes, 4ull * num_graphs)) ||
      (rc = device_alloc((void**)&t_ro, 4ull * total_nodes)) ||
      (rc = device_alloc((void**)&t_ci, 4ull * total_edges)) ||
      (rc = device_alloc((void**)&t_ent, 4ull * num_graphs)) ||
      (rc = device_alloc((void**)&t_eoff, 4ull * (num_graphs + 1))) ||
      (rc = device_alloc((void**)&t_stat, 16)) ||
      (rc = device_alloc((void**)&t_ng, 4ull * total_nodes))) {
    cleanup();
    bms2_device_csr_destroy(q, &d);
    return rc;
  }
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(st);
    cleanup();
    bms2_device_csr_destroy(q, &d);
    return code;
  };
#define PACK_HIP(x) \
  if ((x) != hipSuccess) return fail(set_error(BMS2_ERR_HIP, "%s failed", #x))
  if (num_graphs) {
    PACK_HIP(hipMemcpyAsync(t_nodes, nodes, 4ull * num_graphs, hipMemcpyHostToDevice, st));
  }
  if (total_nodes) {
    PACK_HIP(hipMemcpyAsync(t_ro, ro_local, 4ull * total_nodes, hipMemcpyHostToDevice, st));
    PACK_HIP(hipMemcpyAsync(d.node_labels, node_labels, total_nodes, hipMemcpyHostToDevice, st));
  }
  if (total_edges) {
    PACK_HIP(hipMemcpyAsync(t_ci, ci_local, 4ull * total_edges, hipMemcpyHostToDevice, st));
    if (edge_labels) {
      PACK_HIP(hipMemcpyAsync(d.edge_labels, edge_labels, total_edges, hipMemcpyHostToDevice, st));
    } else {
      PACK_HIP(hipMemsetAsync(d.edge_labels, 0, total_edges, st));
    }
  }
  hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint64_t*>(t_stat),
                     uint64_t(2));
  uint32_t E = 0;
  if (num_graphs) {
    scan_counts(st, t_nodes, uint64_t(num_graphs), d.graph_offsets);
    hipLaunchKernelGGL(pack_entries_kernel, dim3((num_graphs + 255) / 256), dim3(256), 0, st,
                       t_nodes, d.graph_offsets, num_graphs, t_ro, t_ent);
    scan_counts(st, t_ent, uint64_t(num_graphs), t_eoff);
    PACK_HIP(hipMemcpyAsync(&E, t_eoff + num_graphs, 4, hipMemcpyDeviceToHost, st));
  } else {
    PACK_HIP(hipMemsetAsync(d.graph_offsets, 0, 4, st));
  }
  PACK_HIP(hipStreamSynchronize(st));
  if (uint64_t(E) != total_edges)
    return fail(set_error(BMS2_ERR_INVALID, "row ends give %u CSR entries, total_edges is %llu", E,
                          (unsigned long long)total_edges));
  if (num_graphs) {
    hipLaunchKernelGGL(pack_node_graph_kernel, dim3((num_graphs + 255) / 256), dim3(256), 0, st,
                       d.graph_offsets, num_graphs, BMS2_MAX_GRAPH_NODES, t_ng, t_stat);
    if (total_nodes)
      hipLaunchKernelGGL(pack_csr_kernel, dim3(uint32_t((total_nodes + 255) / 256)), dim3(256), 0,
                         st, d.graph_offsets, t_eoff, t_ng, total_nodes, t_ro, t_ci, d.row_offsets,
                         d.column_indices, t_stat);
    else
      PACK_HIP(hipMemsetAsync(d.row_offsets, 0, 4, st));
  } else {
    PACK_HIP(hipMemsetAsync(d.row_offsets, 0, 4, st));
  }
  uint32_t stat[4] = {0, 0, 0, 0};
  PACK_HIP(hipMemcpyAsync(stat, t_stat, 16, hipMemcpyDeviceToHost, st));
  PACK_HIP(hipStreamSynchronize(st));
#undef PACK_HIP
  if (stat[0] & kPackSize)
    return fail(set_error(BMS2_ERR_UNSUPPORTED, "a graph has %u nodes; the limit is %u", stat[1],
                          BMS2_MAX_GRAPH_NODES));
  if (stat[0] & kPackRows) return fail(set_error(BMS2_ERR_INVALID, "local row ends not monotone"));
  if (stat[0] & kPackRange)
    return fail(set_error(BMS2_ERR_INVALID, "a CSR entry points outside its graph"));
  if (stat[0] & kPackLoop) return fail(set_error(BMS2_ERR_UNSUPPORTED, "self-loop edge"));
  if (stat[0] & kPackAsym) return fail(set_error(BMS2_ERR_INVALID, "CSR is not symmetric"));
  cleanup();
  d.max_graph_nodes = stat[1];
  d.max_degree = stat[2];
  *out = d;
  return BMS2_OK;
}

// ---- CS (cluster) join
namespace {
int check_cs(bms2_queue q, const bms2_device_csr* query, const bms2_device_csr* data,
             bms2_candidates cand, uint32_t end) {
  if (int rc = use_device(q)) return rc;
  if (!query || !data || !cand) return set_error(BMS2_ERR_INVALID, "null argument");
  if (cand->rows != query->total_nodes || cand->targets != data->total_nodes)
    return set_error(BMS2_ERR_INVALID, "candidates of another pair of batches");
  if (end == 0 || end > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_INVALID, "partial match depth %u outside 1..%u", end,
                     BMS2_MAX_QUERY_NODES);
  if (query->max_graph_nodes > BMS2_MAX_QUERY_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "query with %u nodes (limit %u, isomorphism.hpp:248)",
                     query->max_graph_nodes, BMS2_MAX_QUERY_NODES);
  if (data->max_graph_nodes > BMS2_MAX_GRAPH_NODES)
    return set_error(BMS2_ERR_UNSUPPORTED, "data graph with %u nodes (limit %u)",
                     data->max_graph_nodes, BMS2_MAX_GRAPH_NODES);
  return BMS2_OK;
}
}  // namespace

int bms2_join_partial_candidates(bms2_queue q, const bms2_device_csr* query,
                                 const bms2_device_csr* data, bms2_candidates cand, uint32_t end,
                                 int32_t* d_partial, uint64_t capacity, uint64_t* num_partial,
                                 bms2_event* ev) {
  if (int rc = check_cs(q, query, data, cand, end)) return rc;
  if (!num_partial) return set_error(BMS2_ERR_INVALID, "null argument");
  if (capacity && !d_partial) return set_error(BMS2_ERR_INVALID, "null partial-match buffer");
  *num_partial = 0;
  uint32_t qgo[2] = {0, 0};
  if (query->num_graphs) {
    BMS2_HIP(hipMemcpy(qgo, query->graph_offsets, sizeof(qgo), hipMemcpyDeviceToHost));
  }
  const uint32_t D = data->num_graphs;
  EventScope es{q, ev};
  if (int rc = es.begin()) return rc;
  // query 0 with fewer than `end` nodes yields nothing (isomorphism.hpp:1087)
  if (query->num_graphs && D && qgo[1] - qgo[0] >= end) {
    hipStream_t st = q->stream;
    uint32_t* cnt = nullptr;
    uint32_t* off = nullptr;
    uint64_t* tot = nullptr;
    int rc = BMS2_OK;
    if ((rc = device_alloc((void**)&cnt, 4ull * D)) ||
        (rc = device_alloc((void**)&off, 4ull * (D + 1))) ||
        (rc = device_alloc((void**)&tot, 8))) {
      (void)hipFree(cnt);
      (void)hipFree(off);
      return rc;
    }
    auto launch = [&](auto kern, int pass) {
      hipLaunchKernelGGL(kern, dim3((D + 63) / 64), dim3(64), 0, st, pass, cand->bits, cand->W,
                         query->graph_offsets, query->row_offsets, query->column_indices,
                         data->graph_offsets, D, data->row_offsets, data->column_indices, end, cnt,
                         off, tot, d_partial, capacity);
    };
    const bool small = data->max_graph_nodes <= 64;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(64), 0, st, tot, uint64_t(1));
    if (small) launch(cs_partial_kernel<1>, 0); else launch(cs_partial_kernel<4>, 0);
    uint64_t total = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = set_error(BMS2_ERR_HIP, "partial-match count failed");
    if (rc == BMS2_OK && total > UINT32_MAX)
      rc = set_error(BMS2_ERR_UNSUPPORTED, "%llu partial matches (limit %u)",
                     (unsigned long long)total, UINT32_MAX);
    if (rc == BMS2_OK) {
      *num_partial = total;
      if (total > capacity)
        rc = set_error(BMS2_ERR_CAPACITY, "%llu partial matches, buffer holds %llu",
                       (unsigned long long)total, (unsigned long long)capacity);
    }
    if (rc == BMS2_OK && total) {
      scan_counts(st, cnt, uint64_t(D), off);
      if (small) launch(cs_partial_kernel<1>, 1); else launch(cs_partial_kernel<4>, 1);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        rc = set_error(BMS2_ERR_HIP, "partial-match fill failed");
    }
    (void)hipFree(cnt);
    (void)hipFree(off);
    (void)hipFree(tot);
    if (rc) return rc;
  }
  return es.end();
}

int bms2_join_with_partial_matches(bms2_queue q, const bms2_device_csr* query,
                                   const bms2_device_csr* data, bms2_candidates cand,
                                   bms2_gmcr gmcr, uint32_t end, const int32_t* d_partial,
                                   uint64_t num_partial, uint64_t* d_num_matches, int find_first,
                                   uint64_t* d_pair_counts, uint64_t pair_capacity,
                                   bms2_event* ev) {
  if (int rc = check_cs(q, query, data, cand, end)) return rc;
  if (!gmcr || !d_num_matches || !d_pair_counts || (num_partial && !d_partial))
    return set_error(BMS2_ERR_INVALID, "null argument");
  if (!gmcr->offsets) return set_error(BMS2_ERR_INVALID, "gmcr not generated");
  if (gmcr->num_data != data->num_graphs) return set_error(BMS2_ERR_INVALID, "gmcr of another batch");
  if (pair_capacity < gmcr->total)
    return set_error(BMS2_ERR_CAPACITY, "per-pair counts hold %llu entries, the GMCR has %llu",
                     (unsigned long long)pair_capacity, (unsigned long long)gmcr->total);
  EventScope es{q, ev};
  if (int rc = es.begin()) return rc;
  hipStream_t st = q->stream;
  const uint64_t T = gmcr->total;
  if (T)
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(T, 256, 4096)), dim3(256), 0, st,
                       d_pair_counts, T);
  if (num_partial && T) {
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(grid_for(num_partial, 64, 1u << 30)), dim3(64), 0, st,
                         cand->bits, cand->W, query->graph_offsets, query->row_offsets,
                         query->column_indices, data->graph_offsets, data->row_offsets,
                         data->column_indices, gmcr->offsets, gmcr->indices, data->num_graphs,
                         end, d_partial, num_partial, find_first, d_pair_counts, d_num_matches);
    };
    if (data->max_graph_nodes <= 64) launch(cs_extend_kernel<1>); else launch(cs_extend_kernel<4>);
  }
  BMS2_CHECK_LAUNCH("cs_extend_kernel");
  return es.end();
}

}  // extern "C"
