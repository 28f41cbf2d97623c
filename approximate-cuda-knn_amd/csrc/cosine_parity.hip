// COSINE initial pass with the reference's exact chunk semantics.
//
// findInitialStateOfAproximatedCosineKNN (euclideandistancematrixgpu.cu:714-1031)
// stores every selected entry as 1 - key (.cu:1007, :1021) and re-enters the
// carried entries into the next chunk keyed by that stored value (.cu:806), so
// the result depends on how cand(i) is cut into chunks and cannot be
// decomposed over trees.  This kernel therefore materialises cand(i) = the
// union of i's start-node ranges over all trees, minus i, in ascending id
// order (the chunk membership fixed as the legal execution: the first
// dimIdx - k unused candidates in id order), computes every key once, and runs
// the chunk loop: pool = carried k (key = stored) + next chunk (key = distance),
// first k by (key, id), store 1 - key, re-sort by (stored, id).
//
// MI355X mapping: one workgroup per point (persistent grid), points taken in
// the first tree's leaf order with each XCD's groups on a contiguous stretch
// of it, so the groups running together on one XCD read overlapping candidate
// rows through its L2.  The union is a
// two-level bitmap in a per-workgroup global scratch (one bit per point plus
// one summary bit per 32-bit word, both L2-resident), enumerated in id order
// with a block scan and cleared as it is read.  Keys are computed by all four
// waves (one candidate per lane, row streamed from HBM, x_i in LDS); the
// sequential chunk selection runs on wave 0 with the list held one entry per
// lane.
//
// Key table (when it fits in device memory): the keys are not streamed per
// point.  Every point's start-node range of every tree is contiguous in that
// tree's leaf order, and the points of one leaf share it, so
// cosine_keys_kernel computes each tree's keys (owner, range member) with the
// exact leaf kernel's tiling (64 owners x 64 candidates, 8x8 pairs per lane,
// 32-dimension chunks staged in LDS, each pair's chain in dimension order)
// and writes them to a table at the owner's offset (the exclusive prefix sum
// of the range sizes over trees and points).  The parity kernel's union pass
// then stores each member's key at its id in a per-workgroup array while it
// sets the member's bit, and the id-ordered enumeration reads the keys back:
// every key is computed once per (point, tree) instead of once per point per
// candidate row streamed through the caches.
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"

namespace aknn {

namespace {

__device__ __forceinline__ int block_excl_scan256(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += sh[w];
  *total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return before + incl - v;
}

constexpr int kTile = 64;
constexpr int kDc = 32;  // dimensions per LDS stage (divides 256)
constexpr int kPad = 4;

// Keys of one tree for the owners of each work item against their own
// start-node ranges: key(i, perm[r]) at keys[off[i] + r - rs[i]] for r in
// [rs[i], rs[i] + rc[i]).  One wave per (item, 64-owner chunk).
template <bool BIGD>
__global__ __launch_bounds__(64) void cosine_keys_kernel(const float* __restrict__ x, int d, int ld,
                                                         const float* __restrict__ norms,
                                                         const int32_t* __restrict__ perm,
                                                         const Leaf* __restrict__ items, int nitems,
                                                         int chunks_per_item, const int32_t* __restrict__ rsp,
                                                         const int32_t* __restrict__ rcp,
                                                         const int64_t* __restrict__ off, float* __restrict__ keys) {
  __shared__ __attribute__((aligned(16))) float smem[2 * kDc * (kTile + kPad)];
  float(*As)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem);
  float(*Bs)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem + kDc * (kTile + kPad));
  float(*Dt)[kTile + 1] = reinterpret_cast<float(*)[kTile + 1]>(smem);
  __shared__ float onorm[kTile], cnorm[kTile];
  __shared__ int64_t obase[kTile];  // keys[] index of the owner's range start
  __shared__ int32_t olo[kTile], ohi[kTile];

  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t item = lb / chunks_per_item;
  if (item >= nitems) return;
  const Leaf L = items[item];
  const int o0 = static_cast<int>(lb % chunks_per_item) * kTile;
  if (o0 >= L.size) return;
  const int lane = threadIdx.x;
  const int nown = L.size - o0 < kTile ? L.size - o0 : kTile;
  const bool has_owner = lane < nown;
  const int32_t my_id = has_owner ? perm[L.start + o0 + lane] : 0;
  onorm[lane] = has_owner ? norms[my_id] : 1.0f;
  olo[lane] = has_owner ? rsp[my_id] : 0;
  ohi[lane] = has_owner ? rsp[my_id] + rcp[my_id] : 0;
  obase[lane] = has_owner ? off[my_id] - rsp[my_id] : 0;
  const int r = lane >> 3, c = lane & 7;
  const float* orow = x + static_cast<int64_t>(my_id) * ld;
  for (int cb = 0; cb < L.rcount; cb += kTile) {
    const int ncand = L.rcount - cb < kTile ? L.rcount - cb : kTile;
    __syncthreads();
    const int32_t my_c = lane < ncand ? perm[L.rstart + cb + lane] : 0;
    cnorm[lane] = lane < ncand ? norms[my_c] : 1.0f;
    float acc[8][8];
    float tot[BIGD ? 8 : 1][BIGD ? 8 : 1];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;
    if constexpr (BIGD) {
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) tot[i][j] = 0.0f;
    }
    const float* crow = x + static_cast<int64_t>(my_c) * ld;
    for (int d0 = 0; d0 < d; d0 += kDc) {
      if constexpr (BIGD) {
        // chunks of 256 dimensions start from 0 and are added in order
        if (d0 > 0 && (d0 & 255) == 0) {
#pragma unroll
          for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              tot[i][j] = tot[i][j] + acc[i][j];
              acc[i][j] = 0.0f;
            }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kDc; q += 4) {
        float4 ov = make_float4(0.f, 0.f, 0.f, 0.f), cv = ov;
        if (d0 + q < d) {
          if (has_owner) ov = *reinterpret_cast<const float4*>(orow + d0 + q);
          if (lane < ncand) cv = *reinterpret_cast<const float4*>(crow + d0 + q);
        }
        As[q + 0][lane] = ov.x;
        As[q + 1][lane] = ov.y;
        As[q + 2][lane] = ov.z;
        As[q + 3][lane] = ov.w;
        Bs[q + 0][lane] = cv.x;
        Bs[q + 1][lane] = cv.y;
        Bs[q + 2][lane] = cv.z;
        Bs[q + 3][lane] = cv.w;
      }
      __syncthreads();
      const int dn = d - d0 < kDc ? d - d0 : kDc;
      for (int q = 0; q < dn; ++q) {
        const float4 a0 = *reinterpret_cast<const float4*>(&As[q][8 * r]);
        const float4 a1 = *reinterpret_cast<const float4*>(&As[q][8 * r + 4]);
        const float4 b0 = *reinterpret_cast<const float4*>(&Bs[q][8 * c]);
        const float4 b1 = *reinterpret_cast<const float4*>(&Bs[q][8 * c + 4]);
        const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int j = 0; j < 8; ++j) acc[i][j] = chain_step<kCosine>(a[i], b[j], acc[i][j]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float t = BIGD ? tot[i][j] + acc[i][j] : 0.0f + acc[i][j];
        Dt[8 * r + i][8 * c + j] = finish_key<kCosine>(t, onorm[8 * r + i], cnorm[8 * c + j]);
      }
    __syncthreads();
    // row by row: lane j stores owner o's key for candidate position
    // L.rstart + cb + j when that position lies in o's own range
    const int32_t pos = L.rstart + cb + lane;
    for (int o = 0; o < nown; ++o) {
      if (lane < ncand && pos >= olo[o] && pos < ohi[o]) keys[obase[o] + pos] = Dt[o][lane];
    }
  }
}

// Exclusive prefix sums of int32 counts into int64 offsets, in blocks of
// 1024: per-block totals, a scan of the totals, then each block's scan.
__global__ __launch_bounds__(1024) void count_block_sums_kernel(const int32_t* __restrict__ v, int64_t n,
                                                               int64_t* __restrict__ bsum) {
  __shared__ int64_t sh[16];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x;
  int64_t t = i < n ? v[i] : 0;
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < 16; ++w) s += sh[w];
    bsum[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(1024) void scan_block_sums_kernel(int64_t* __restrict__ bsum, int64_t nb,
                                                              int64_t* __restrict__ total) {
  __shared__ int64_t sh[16];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
    const int64_t i = b0 + threadIdx.x;
    const int64_t v = i < nb ? bsum[i] : 0;
    int64_t incl = v;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    int64_t before = carry;
    for (int w = 0; w < wave; ++w) before += sh[w];
    if (i < nb) bsum[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = before + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void count_scan_kernel(const int32_t* __restrict__ v, int64_t n,
                                                         const int64_t* __restrict__ bsum,
                                                         int64_t* __restrict__ out) {
  __shared__ int64_t sh[16];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 1024 + threadIdx.x;
  const int64_t x = i < n ? v[i] : 0;
  int64_t incl = x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  int64_t before = bsum[blockIdx.x];
  for (int w = 0; w < wave; ++w) before += sh[w];
  if (i < n) out[i] = before + incl - x;
}

// First k of the pool[0, np) of packed (key, id) pairs (any order; equal
// pairs are interchangeable), by a radix select on the 64-bit values, 8 bits
// per round from the top, over all 256 threads: the threshold digit is found
// from a histogram of the entries still tied with the prefix, until the
// bucket holding the k-th entry is taken whole.  out[0, k) receives them.
__device__ void pool_select(const uint64_t* pool, int np, int k, uint64_t* out, uint32_t* hist, int* sel) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t prefix = 0, mask = 0;
  int need = k;
  bool whole = false;
  for (int shift = 56; shift >= 0 && !whole; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    for (int e = tid; e < np; e += 256) {
      const uint64_t v = pool[e];
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255], 1u);
    }
    __syncthreads();
    if (wave == 0) {
      const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2], h3 = hist[4 * lane + 3];
      const int sum = static_cast<int>(h0 + h1 + h2 + h3);
      int incl = sum;
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      const int excl = incl - sum;
      if (excl < need && need <= incl) {
        const int hb[4] = {static_cast<int>(h0), static_cast<int>(h1), static_cast<int>(h2), static_cast<int>(h3)};
        int c = excl;
        for (int b = 0; b < 4; ++b) {
          if (c + hb[b] >= need) {
            sel[0] = 4 * lane + b;
            sel[1] = c;
            sel[2] = hb[b];
            break;
          }
          c += hb[b];
        }
      }
    }
    __syncthreads();
    const int dd = sel[0], before = sel[1], cnt = sel[2];
    prefix |= static_cast<uint64_t>(dd) << shift;
    mask |= 0xFFull << shift;
    need -= before;
    whole = cnt == need;
  }
  // whole: every entry whose top bits are at most the prefix (exactly k);
  // otherwise all 64 bits are fixed and `need` of the equal entries complete it
  if (tid == 0) {
    sel[3] = 0;
    sel[4] = 0;
  }
  __syncthreads();
  const uint64_t top = prefix | ~mask;
  for (int e = tid; e < np; e += 256) {
    const uint64_t v = pool[e];
    bool take = whole ? v <= top : v < prefix;
    if (!whole && v == prefix) take = atomicAdd(&sel[4], 1) < need;
    if (take) out[atomicAdd(&sel[3], 1)] = v;
  }
  __syncthreads();
}

template <bool TABLE>
__global__ __launch_bounds__(256) void cosine_parity_kernel(const float* __restrict__ x, int d, int ld,
                                                            const float* __restrict__ norms, int64_t n, int k,
                                                            int trees, const int32_t* __restrict__ rs,
                                                            const int32_t* __restrict__ rc,
                                                            const int32_t* __restrict__ perm, int64_t cap,
                                                            CosineParityScratch sc, int32_t* __restrict__ ids,
                                                            float* __restrict__ vals,
                                                            const int32_t* __restrict__ order,
                                                            const float* __restrict__ ktab,
                                                            const int64_t* __restrict__ koff) {
  // LDS: x_i (without the key table), the summary bitmap (one bit per 32-id
  // word of the global bitmap), the selection pool and the carried list
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const int xi_words = TABLE ? 0 : (d + 3) / 4 * 4;
  const int sum_words = static_cast<int>((sc.summary_words + 3) / 4 * 4);
  float* xi = reinterpret_cast<float*>(lds);
  uint32_t* summ = reinterpret_cast<uint32_t*>(lds) + xi_words;
  uint64_t* pool = reinterpret_cast<uint64_t*>(summ + sum_words);  // [k + cap]
  __shared__ uint64_t car[64], outl[64];
  __shared__ uint32_t hist[256];
  __shared__ int sh[4], sel[5];
  uint32_t* bits = sc.bits + static_cast<int64_t>(blockIdx.x) * sc.words;
  int32_t* cand = sc.cand + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  float* ckey = sc.key + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  float* kid = TABLE ? sc.keyarr + static_cast<int64_t>(blockIdx.x) * n : nullptr;  // key by candidate id
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int w = tid; w < sum_words; w += 256) summ[w] = 0;

  const int64_t g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  // logical group g takes the stretch [g*per, (g+1)*per) of the leaf order
  for (int64_t pidx = g * per; pidx < n && pidx < (g + 1) * per; ++pidx) {
    const int64_t i = order != nullptr ? order[pidx] : pidx;
    if constexpr (!TABLE)
      for (int dd = tid; dd < d; dd += 256) xi[dd] = x[i * ld + dd];
    __syncthreads();
    // union of the start-node ranges (.cu:160-191): bits in the global
    // bitmap (L2 atomics), summary bits in LDS; with the key table each
    // member's key goes to its id's slot (a member of several ranges gets
    // the same key from each)
    for (int t = 0; t < trees; ++t) {
      const int32_t s0 = rs[static_cast<int64_t>(t) * n + i], c0 = rc[static_cast<int64_t>(t) * n + i];
      const int32_t* pt = perm + static_cast<int64_t>(t) * n + s0;
      const float* kt = TABLE ? ktab + koff[static_cast<int64_t>(t) * n + i] : nullptr;
      // four members per thread per pass, their loads issued together
      for (int p0 = 0; p0 < c0; p0 += 1024) {
        int32_t jv[4];
        float kv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = p0 + 256 * u + tid;
          jv[u] = p < c0 ? pt[p] : -1;
          if constexpr (TABLE) kv[u] = p < c0 ? kt[p] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int32_t j = jv[u];
          if (j < 0) continue;
          if constexpr (TABLE) kid[j] = kv[u];
          __hip_atomic_fetch_or(&bits[j >> 5], 1u << (j & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          atomicOr(&summ[j >> 10], 1u << ((j >> 5) & 31));
        }
      }
    }
    // every access to the bitmap is a workgroup-scope atomic and the key
    // slots are read after a workgroup barrier: coherent within the
    // workgroup, with no cache maintenance beyond the CU
    __syncthreads();
    // enumerate in ascending id order, without i itself (.cu:190-191); the
    // words of each summary word's set bits are read together and cleared
    int m = 0;
    for (int sb = 0; sb < sum_words; sb += 256) {
      const int sw = sb + tid;
      const uint32_t s = sw < sum_words ? summ[sw] : 0u;
      uint32_t bw[32];
      int cnt = 0;
      if (s != 0) {
        // only the words whose summary bit is set, all requested before use
        uint32_t* line = bits + static_cast<int64_t>(sw) * 32;
#pragma unroll
        for (int b = 0; b < 32; ++b)
          bw[b] = ((s >> b) & 1u) ? __hip_atomic_load(&line[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0u;
        const int64_t iw = i >> 5;
        if ((iw >> 5) == sw) bw[iw & 31] &= ~(1u << (i & 31));
#pragma unroll
        for (int b = 0; b < 32; ++b) cnt += __popc(bw[b]);
      }
      int tot;
      int off = m + block_excl_scan256(cnt, sh, &tot);
      if (s != 0) {
        uint32_t* wline = bits + static_cast<int64_t>(sw) * 32;
#pragma unroll
        for (int b = 0; b < 32; ++b)
          if ((s >> b) & 1u) __hip_atomic_store(&wline[b], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        summ[sw] = 0;
#pragma unroll
        for (int b = 0; b < 32; ++b)
          for (uint32_t r = bw[b]; r; r &= r - 1)
            cand[off++] = static_cast<int32_t>((static_cast<int64_t>(sw) * 32 + b) * 32 + __ffs(r) - 1);
      }
      m += tot;
    }
    if constexpr (!TABLE) {
      __syncthreads();
      const float ni = norms[i];
      for (int c = tid; c < m; c += 256) {
        const int32_t j = cand[c];
        const float t = row_chain<kCosine>(xi, x + static_cast<int64_t>(j) * ld, d);
        ckey[c] = finish_key<kCosine>(t, ni, norms[j]);
      }
    }
    __syncthreads();
    // chunk loop (.cu:224-239, :1007): the pool is the carried list (keyed by
    // its stored values) and the next cap candidates in id order (keyed by
    // their distances); its first k by (key, id) are stored as 1 - key and
    // sorted by (stored, id)
    if (tid < 64) car[tid] = pack_pair(kStubDist, order_id(kStubId));
    __syncthreads();
    for (int64_t c0 = 0; c0 < m; c0 += cap) {
      const int nc = static_cast<int>(c0 + cap < m ? cap : m - c0);
      const int np = k + nc;
      for (int e = tid; e < np; e += 256) {
        if (e < k) {
          pool[e] = car[e];
        } else {
          const int32_t j = cand[c0 + e - k];
          const float key = TABLE ? kid[j] : ckey[c0 + e - k];
          pool[e] = pack_pair(key, order_id(j));
        }
      }
      __syncthreads();
      pool_select(pool, np, k, outl, hist, sel);
      if (wave == 0) {
        float key = __builtin_inff();
        int32_t id = INT_MAX;
        if (lane < k) {
          const uint64_t v = outl[lane];
          key = static_cast<float>(1.0 - static_cast<double>(pair_key(v)));
          id = pair_id(v);
        }
        wave_sort(key, id);
        if (lane < k) car[lane] = pack_pair(key, order_id(id));
      }
      __syncthreads();
    }
    if (tid < k) {
      const uint64_t v = car[tid];
      ids[i * k + tid] = pair_id(v);
      vals[i * k + tid] = pair_key(v);
    }
    __syncthreads();
  }
}

}  // namespace

size_t cosine_parity_lds(int d, int64_t summary_words, int k, int64_t cap, bool table) {
  const size_t xi = table ? 0 : static_cast<size_t>((d + 3) / 4 * 4) * 4;
  const size_t summ = static_cast<size_t>((summary_words + 3) / 4 * 4) * 4;
  return xi + summ + static_cast<size_t>(k + cap) * 8;
}

void launch_cosine_parity_initial(const float* x, int d, int ld, const float* norms, int64_t n, int k, int trees,
                                  const int32_t* rs, const int32_t* rc, const int32_t* perm, int64_t cap,
                                  const CosineParityScratch& sc, int32_t* ids, float* vals, const float* ktab,
                                  const int64_t* koff, hipStream_t s) {
  if (k > 64) throw HipError("COSINE parity path supports k <= 64");
  if (sc.groups % 8 != 0) throw HipError("COSINE parity path: groups must be a multiple of 8");
  const bool table = ktab != nullptr;
  if (table && (koff == nullptr || sc.keyarr == nullptr)) throw HipError("COSINE parity path: key table without offsets");
  const size_t shmem = cosine_parity_lds(d, sc.summary_words, k, cap, table);
  if (shmem > 160 * 1024 - 4096) throw HipError("COSINE parity path: d, n or k too large for LDS");
  if (table)
    hipLaunchKernelGGL(cosine_parity_kernel<true>, dim3(sc.groups), dim3(256), shmem, s, x, d, ld, norms, n, k, trees,
                       rs, rc, perm, cap, sc, ids, vals, perm, ktab, koff);
  else
    hipLaunchKernelGGL(cosine_parity_kernel<false>, dim3(sc.groups), dim3(256), shmem, s, x, d, ld, norms, n, k,
                       trees, rs, rc, perm, cap, sc, ids, vals, perm, ktab, koff);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_cosine_keys(const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* items,
                        int nitems, int max_owners, const int32_t* rs, const int32_t* rc, const int64_t* off,
                        float* keys, hipStream_t s) {
  if (nitems <= 0) return;
  if (ld % 4 != 0 || ld < d) throw HipError("COSINE keys: row stride must be a multiple of 4 covering d");
  const int chunks = (max_owners + kTile - 1) / kTile;
  const int64_t grid = xcd_grid(static_cast<int64_t>(nitems) * chunks);
  if (d > 256)
    hipLaunchKernelGGL(cosine_keys_kernel<true>, dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, x, d, ld,
                       norms, perm, items, nitems, chunks, rs, rc, off, keys);
  else
    hipLaunchKernelGGL(cosine_keys_kernel<false>, dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, x, d, ld,
                       norms, perm, items, nitems, chunks, rs, rc, off, keys);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_count_scan(const int32_t* counts, int64_t n, int64_t* offsets, int64_t* block_sums, int64_t* total,
                       hipStream_t s) {
  if (n <= 0) return;
  const int64_t nb = (n + 1023) / 1024;
  hipLaunchKernelGGL(count_block_sums_kernel, dim3(static_cast<unsigned>(nb)), dim3(1024), 0, s, counts, n,
                     block_sums);
  hipLaunchKernelGGL(scan_block_sums_kernel, dim3(1), dim3(1024), 0, s, block_sums, nb, total);
  hipLaunchKernelGGL(count_scan_kernel, dim3(static_cast<unsigned>(nb)), dim3(1024), 0, s, counts, n, block_sums,
                     offsets);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
