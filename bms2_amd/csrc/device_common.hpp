// Device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace bms2 {

constexpr int kWave = 64;
constexpr uint8_t kWildcard = 255;  // types.hpp:18

// MurmurHash3 fmix32 (signature.hpp:393-400).
__host__ __device__ inline uint32_t hash32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// boost::hash_combine (signature.hpp:402-404).
__host__ __device__ inline uint32_t combine_hash(uint32_t h1, uint32_t h2) {
  return h1 ^ (h2 + 0x9e3779b9u + (h1 << 6) + (h1 >> 2));
}

// Bloom insert of PathSignatureDevice::insert (signature.hpp:418-430) into 8 u32 words.
__device__ inline void bloom_insert(uint32_t* bits, uint32_t h1) {
  const uint32_t h2 = hash32(h1);
  const uint32_t h3 = hash32(h2);
  const uint32_t i1 = h1 & 255u, i2 = h2 & 255u, i3 = h3 & 255u;
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    uint32_t m = 0;
    m |= ((i1 >> 5) == uint32_t(w)) ? (1u << (i1 & 31)) : 0u;
    m |= ((i2 >> 5) == uint32_t(w)) ? (1u << (i2 & 31)) : 0u;
    m |= ((i3 >> 5) == uint32_t(w)) ? (1u << (i3 & 31)) : 0u;
    bits[w] |= m;
  }
}

// SignatureDevice::incrementLabelCount (signature.hpp:43-50).
__host__ __device__ inline uint64_t inc_label(uint64_t sig, uint32_t label) {
  if (label < 16) {
    const uint64_t c = (sig >> (4 * label)) & 15u;
    if (c < 15) sig += uint64_t(1) << (4 * label);
  }
  return sig;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

template <class T>
__device__ inline T uniform(T x) {
  return static_cast<T>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}

__device__ inline uint64_t uniform64(uint64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  return (uint64_t(hi) << 32) | lo;
}

__device__ inline uint64_t readlane64(uint64_t x, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane(uint32_t(x), lane);
  uint32_t hi = __builtin_amdgcn_readlane(uint32_t(x >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

// x with lane `lane` replaced by the uniform value v.
__device__ inline uint64_t writelane64(uint64_t v, int lane, uint64_t x) {
  return int(lane_id()) == lane ? v : x;
}

// Block number in an order that keeps consecutive numbers on one XCD: the hardware deals blocks
// to the 8 XCDs round-robin, so with a grid of a multiple of 8 blocks, block b runs on XCD b % 8
// and takes number (b % 8) * (grid / 8) + b / 8.
__device__ inline uint32_t xcd_block() {
  const uint32_t n = gridDim.x, b = blockIdx.x;
  return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

__device__ inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }
__device__ inline int popc64(uint64_t x) { return __popcll(x); }

// Mask of bits [0, n) for n in [0, 64].
__host__ __device__ inline uint64_t low_mask(uint32_t n) {
  return n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
}

// Size classes of the data graphs of 65-256 nodes in their joins: adjacency rows and domains of
// 2 u64 words up to `wide_above` nodes (128, or 64 to give every graph 4 words), 4 words above.
constexpr int kBigClasses = 2;
__host__ __device__ inline int big_class(uint32_t n, uint32_t wide_above) { return n > wide_above ? 1 : 0; }

// getGraphID (graph.hpp:96-107): graph holding a global node id.
__device__ inline uint32_t graph_of(const uint32_t* graph_offsets, uint32_t num_graphs,
                                    uint32_t node) {
  uint32_t lo = 0, hi = num_graphs;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (node >= graph_offsets[mid])
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

}  // namespace bms2
