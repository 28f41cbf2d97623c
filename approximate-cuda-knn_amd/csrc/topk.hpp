// Running top-k of (key, id) pairs in registers, used by the leaf kernels.
//
// A pair is held as one unsigned 64-bit integer whose unsigned order is the
// order of key_less (device_util.hpp): the key's float bits mapped onto an
// unsigned total order in the high half, the id offset by 2^31 in the low
// half.  The two orders agree for every key that is not NaN, except that
// the map places -0 below +0; no stored key is -0 (sqrt of a sum started at
// +0, a sum of |a - b| started at +0, or 1 - c rounded to nearest, which is
// +0 when c is 1).  One 64-bit compare then replaces key_less's three
// compares and two mask operations.
//
// The list is ascending in pk[0..KMAX-1] and right-aligned: the row's k
// pairs sit in pk[KMAX-k..KMAX-1] and the slots below hold 0, which no pair
// undercuts, so an insertion never moves them.  The k-th pair, the
// threshold, is therefore always pk[KMAX-1], and every slot runs the same
// straight-line compare-and-select with no branch on k.
#pragma once

#include <cstdint>

namespace aknn {

// float -> unsigned, monotone over all non-NaN floats (-0 just below +0)
__device__ __forceinline__ uint32_t order_bits(float f) {
  const uint32_t b = __float_as_uint(f);
  return b ^ (static_cast<uint32_t>(static_cast<int32_t>(b) >> 31) | 0x80000000u);
}

__device__ __forceinline__ float order_float(uint32_t u) {
  return __uint_as_float(u ^ (static_cast<uint32_t>(static_cast<int32_t>(~u) >> 31) | 0x80000000u));
}

// id offset so that signed order becomes unsigned order
__device__ __forceinline__ uint32_t order_id(int32_t id) { return static_cast<uint32_t>(id) ^ 0x80000000u; }

// pair from a key and an id already passed through order_id
__device__ __forceinline__ uint64_t pack_pair(float key, uint32_t oid) {
  return (static_cast<uint64_t>(order_bits(key)) << 32) | oid;
}

__device__ __forceinline__ float pair_key(uint64_t v) { return order_float(static_cast<uint32_t>(v >> 32)); }
__device__ __forceinline__ int32_t pair_id(uint64_t v) {
  return static_cast<int32_t>(static_cast<uint32_t>(v) ^ 0x80000000u);
}

// True when a row of KMAX values at both addresses can move in 16-byte
// pieces: the list is full (k == KMAX, a multiple of 4) and both rows are
// 16-byte aligned (the graph arrays are; row r starts r * k values in).
template <int KMAX>
__device__ __forceinline__ bool topk_row_vector(int k, const void* vals, const void* ids) {
  static_assert(KMAX % 4 == 0, "list sizes are multiples of 4");
  return k == KMAX && ((reinterpret_cast<uintptr_t>(vals) | reinterpret_cast<uintptr_t>(ids)) & 15) == 0;
}

// Loads a graph row of k pairs (k <= KMAX) into the right-aligned list.  A
// lane without a row holds (inf, INT_MAX) pairs; its row pointers must still
// point at a readable row.  Each lane reads its own row, so a full list is
// read in 16-byte pieces: a quarter of the accesses, each to a row that
// other lanes' accesses do not share.
template <int KMAX>
__device__ __forceinline__ void topk_load(uint64_t (&pk)[KMAX], int k, bool valid, const float* __restrict__ vals,
                                          const int32_t* __restrict__ ids) {
  const uint64_t none = pack_pair(__builtin_inff(), order_id(0x7fffffff));
  if (topk_row_vector<KMAX>(k, vals, ids)) {
    // every piece is requested before any is used (the row pointer is
    // readable for every lane), then invalid lanes are masked to `none`
    float4 v[KMAX / 4];
    int4 i[KMAX / 4];
#pragma unroll
    for (int p = 0; p < KMAX / 4; ++p) {
      v[p] = reinterpret_cast<const float4*>(vals)[p];
      i[p] = reinterpret_cast<const int4*>(ids)[p];
    }
    const uint64_t keep = valid ? ~0ull : 0ull;
#pragma unroll
    for (int p = 0; p < KMAX / 4; ++p) {
      pk[4 * p + 0] = (pack_pair(v[p].x, order_id(i[p].x)) & keep) | (none & ~keep);
      pk[4 * p + 1] = (pack_pair(v[p].y, order_id(i[p].y)) & keep) | (none & ~keep);
      pk[4 * p + 2] = (pack_pair(v[p].z, order_id(i[p].z)) & keep) | (none & ~keep);
      pk[4 * p + 3] = (pack_pair(v[p].w, order_id(i[p].w)) & keep) | (none & ~keep);
    }
    return;
  }
#pragma unroll
  for (int p = 0; p < KMAX; ++p) {
    const int q = p - (KMAX - k);
    if (q < 0)
      pk[p] = 0;
    else
      pk[p] = valid ? pack_pair(vals[q], order_id(ids[q])) : none;
  }
}

template <int KMAX>
__device__ __forceinline__ void topk_store(const uint64_t (&pk)[KMAX], int k, float* __restrict__ vals,
                                           int32_t* __restrict__ ids) {
  if (topk_row_vector<KMAX>(k, vals, ids)) {
#pragma unroll
    for (int p = 0; p < KMAX; p += 4) {
      *reinterpret_cast<float4*>(vals + p) =
          make_float4(pair_key(pk[p]), pair_key(pk[p + 1]), pair_key(pk[p + 2]), pair_key(pk[p + 3]));
      *reinterpret_cast<int4*>(ids + p) = make_int4(pair_id(pk[p]), pair_id(pk[p + 1]), pair_id(pk[p + 2]),
                                                    pair_id(pk[p + 3]));
    }
    return;
  }
#pragma unroll
  for (int p = 0; p < KMAX; ++p) {
    const int q = p - (KMAX - k);
    if (q >= 0) {
      vals[q] = pair_key(pk[p]);
      ids[q] = pair_id(pk[p]);
    }
  }
}

// True when the list holds the id of v (the low halves match).
template <int KMAX>
__device__ __forceinline__ bool topk_has(const uint64_t (&pk)[KMAX], uint64_t v) {
  const uint32_t vid = static_cast<uint32_t>(v);
  bool has = false;
#pragma unroll
  for (int p = 0; p < KMAX; ++p) has |= static_cast<uint32_t>(pk[p]) == vid;
  return has;
}

// Inserts v when it undercuts the k-th pair; the caller guarantees that v's
// id is not in the list.
template <int KMAX>
__device__ __forceinline__ void topk_insert(uint64_t (&pk)[KMAX], uint64_t v) {
  if (v < pk[KMAX - 1]) {
    bool lt[KMAX];
#pragma unroll
    for (int p = 0; p < KMAX; ++p) lt[p] = v < pk[p];
#pragma unroll
    for (int p = KMAX - 1; p > 0; --p) pk[p] = lt[p - 1] ? pk[p - 1] : (lt[p] ? v : pk[p]);
    pk[0] = lt[0] ? v : pk[0];
  }
}

// Inserts v when it undercuts the k-th pair and its id is not in the list
// yet (a candidate an earlier tree already stored carries the same exact
// key, so it must not enter twice); the k-th pair drops out.
template <int KMAX>
__device__ __forceinline__ void topk_offer(uint64_t (&pk)[KMAX], uint64_t v) {
  if (v < pk[KMAX - 1] && !topk_has(pk, v)) {
    bool lt[KMAX];
#pragma unroll
    for (int p = 0; p < KMAX; ++p) lt[p] = v < pk[p];
    // lt is false below the insertion slot and true from it on: slots above
    // take their lower neighbour, the insertion slot takes v
#pragma unroll
    for (int p = KMAX - 1; p > 0; --p) pk[p] = lt[p - 1] ? pk[p - 1] : (lt[p] ? v : pk[p]);
    pk[0] = lt[0] ? v : pk[0];
  }
}

}  // namespace aknn
