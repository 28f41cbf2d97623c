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
// range_keys_kernel computes each tree's keys (owner, range member) with the
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
#include "wave_ops.hpp"

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
// [rs[i], rs[i] + rc[i]).  One wave per (item, 64-owner chunk).  Any metric
// (the COSINE parity pass, and the leaf merge of lists longer than 64).
template <int METRIC, bool BIGD>
__global__ __launch_bounds__(64) void range_keys_kernel(const float* __restrict__ x, int d, int ld,
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
  onorm[lane] = METRIC == kCosine && has_owner ? norms[my_id] : 1.0f;
  olo[lane] = has_owner ? rsp[my_id] : 0;
  ohi[lane] = has_owner ? rsp[my_id] + rcp[my_id] : 0;
  obase[lane] = has_owner ? off[my_id] - rsp[my_id] : 0;
  const int r = lane >> 3, c = lane & 7;
  const float* orow = x + static_cast<int64_t>(my_id) * ld;
  for (int cb = 0; cb < L.rcount; cb += kTile) {
    const int ncand = L.rcount - cb < kTile ? L.rcount - cb : kTile;
    __syncthreads();
    const int32_t my_c = lane < ncand ? perm[L.rstart + cb + lane] : 0;
    cnorm[lane] = METRIC == kCosine && lane < ncand ? norms[my_c] : 1.0f;
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
          for (int j = 0; j < 8; ++j) acc[i][j] = chain_step<METRIC>(a[i], b[j], acc[i][j]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float t = BIGD ? tot[i][j] + acc[i][j] : 0.0f + acc[i][j];
        Dt[8 * r + i][8 * c + j] = finish_key<METRIC>(t, onorm[8 * r + i], cnorm[8 * c + j]);
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

// First k of the entries a[0, na) and b[0, nb) of packed (key, id) pairs
// (any order; equal pairs are interchangeable), by a radix select on the
// 64-bit values, 8 bits per round from the top, over all 256 threads: the
// threshold digit is found from a histogram of the entries still tied with
// the prefix, until the bucket holding the k-th entry is taken whole.
// out[0, k) receives them.
__device__ void pool_select(const uint64_t* a, int na, const uint64_t* b, int nb, int k, uint64_t* out,
                            uint32_t* hist, int* sel) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int np = na + nb;
  uint64_t prefix = 0, mask = 0;
  int need = k;
  bool whole = false;
  for (int shift = 56; shift >= 0 && !whole; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    for (int e = tid; e < np; e += 256) {
      const uint64_t v = e < na ? a[e] : b[e - na];
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
        for (int q = 0; q < 4; ++q) {
          if (c + hb[q] >= need) {
            sel[0] = 4 * lane + q;
            sel[1] = c;
            sel[2] = hb[q];
            break;
          }
          c += hb[q];
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
    const uint64_t v = e < na ? a[e] : b[e - na];
    bool take = whole ? v <= top : v < prefix;
    if (!whole && v == prefix) take = atomicAdd(&sel[4], 1) < need;
    if (take) out[atomicAdd(&sel[3], 1)] = v;
  }
  __syncthreads();
}

// LDS union of a point's start-node ranges (key table form): a point whose
// ranges hold at most kUnionCap members in all (duplicates counted) takes its
// (id, key) members into registers, kUnionSlots per thread, and ranks them in
// id order through a bitmap window of kWinIds ids in LDS slid over [0, n):
// per window the members' bits are set, the set bits counted per word and
// scanned, and each member writes itself at its rank (duplicates write the
// same entry to the same slot).  The distinct members land in ascending id
// order as packed (key, id) pairs, with no global scratch.  Ranges of up to
// kUnionTrees trees are held in LDS for the members' tree lookup.
constexpr int kUnionCap = 8192;
constexpr int kUnionSlots = kUnionCap / 256;
constexpr int kWinWords = 16384;  // 64 KB of LDS
constexpr int kWinIds = kWinWords * 32;
constexpr int kWordsPerThread = kWinWords / 256;
constexpr int kGroupsPerThread = kWordsPerThread / 8;  // groups of 8 words (256 ids)
static_assert(kGroupsPerThread == 8, "group prefixes are published as one 16-byte store per thread");
constexpr int kUnionTrees = 64;
constexpr int kGatherBatch = 16;
constexpr int kRankBatch = 4;
// chunk lists (each chunk's own first k) held for the merge
constexpr int kLocalCap = 1024;

template <bool TABLE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void cosine_parity_kernel(const float* __restrict__ x, int d, int ld,
                                                            const float* __restrict__ norms, int64_t n, int k,
                                                            int trees, const int32_t* __restrict__ rs,
                                                            const int32_t* __restrict__ rc,
                                                            const int32_t* __restrict__ perm, int64_t cap,
                                                            CosineParityScratch sc, int32_t* __restrict__ ids,
                                                            float* __restrict__ vals,
                                                            const int32_t* __restrict__ order,
                                                            const float* __restrict__ ktab,
                                                            const int64_t* __restrict__ koff) {
  // LDS, two layouts over one region: the global-bitmap path keeps x_i
  // (without the key table), the summary bitmap (one bit per 32-id word of
  // the global bitmap) and the chunk; the LDS union (key table) keeps the
  // distinct members and the bitmap window
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const int xi_words = TABLE ? 0 : (d + 3) / 4 * 4;
  const int sum_words = static_cast<int>((sc.summary_words + 3) / 4 * 4);
  float* xi = reinterpret_cast<float*>(lds);
  uint32_t* summ = reinterpret_cast<uint32_t*>(lds) + xi_words;
  uint64_t* pool = reinterpret_cast<uint64_t*>(summ + sum_words);  // [cap]
  uint32_t* win = reinterpret_cast<uint32_t*>(lds);  // [kWinWords] bitmap window, then the chunk
  uint64_t* uo = TABLE ? sc.uo + static_cast<int64_t>(blockIdx.x) * kUnionCap : nullptr;  // distinct members
  __shared__ uint64_t car[64], outl[64];
  __shared__ uint32_t hist[256];
  __shared__ int sh[4], sel[5];
  __shared__ int tr_c[kUnionTrees], tr_s[kUnionTrees], tr_p[kUnionTrees + 1];
  __shared__ int64_t tr_o[kUnionTrees];
  __shared__ __attribute__((aligned(16))) uint16_t gpre[TABLE ? kWinWords / 8 : 8];  // per-group window prefixes
  __shared__ uint64_t locl[TABLE ? kLocalCap : 1];  // each chunk's own first k (LDS union path)
  uint32_t* bits = sc.bits + static_cast<int64_t>(blockIdx.x) * sc.words;
  int32_t* cand = sc.cand + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  float* ckey = sc.key + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  float* kid = TABLE ? sc.keyarr + static_cast<int64_t>(blockIdx.x) * n : nullptr;  // key by candidate id
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool lds_union = TABLE && trees <= kUnionTrees;
  // the zeroed parts of each layout: the summary bitmap (cleared as it is
  // enumerated) and the bitmap window (cleared window by window); the other
  // layout's use of the region dirties them
  bool summ_clean = false, win_clean = false;

  // chunk loop (.cu:224-239, :1007): the pool is the carried list (keyed by
  // its stored values) and the next chunk of candidates in id order (keyed
  // by their distances); its first k by (key, id) are stored as 1 - key and
  // sorted by (stored, id)
  auto select_chunk = [&](const uint64_t* chunk, int nc) {
    pool_select(car, k, chunk, nc, k, outl, hist, sel);
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
  };

  const int64_t g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  // logical group g takes the stretch [g*per, (g+1)*per) of the leaf order
  for (int64_t pidx = g * per; pidx < n && pidx < (g + 1) * per; ++pidx) {
    const int64_t i = order != nullptr ? order[pidx] : pidx;
    if (tid < 64) car[tid] = pack_pair(kStubDist, order_id(kStubId));
    int total = INT_MAX;
    if (lds_union) {
      if (tid < trees) {
        const int64_t o = static_cast<int64_t>(tid) * n + i;
        tr_c[tid] = rc[o];
        tr_s[tid] = rs[o];
        tr_o[tid] = koff[o];
      }
      __syncthreads();
      if (wave == 0) {
        const int incl = wave_incl_add(lane < trees ? tr_c[lane] : 0);
        tr_p[lane + 1] = incl;
        if (lane == 0) tr_p[0] = 0;
      }
      __syncthreads();
      total = tr_p[trees];
    }
    if (lds_union && total <= kUnionCap) {
      // members into registers: member e is position e - tr_p[t] of tree t's
      // range, t the last tree with tr_p[t] <= e; the point itself is left out
      // (.cu:190-191)
      // (in batches of kGatherBatch slots: 2 * kGatherBatch loads in flight)
      const int nslots = (total + 255) / 256;
      int32_t jv[kUnionSlots];
      float kv[kUnionSlots];
#pragma unroll
      for (int b0 = 0; b0 < kUnionSlots; b0 += kGatherBatch) {
#pragma unroll
        for (int u = b0; u < b0 + kGatherBatch; ++u) {
          jv[u] = -1;
          kv[u] = 0.0f;
          const int e = 256 * u + tid;
          if (b0 < nslots && e < total) {
            int t = 0;
#pragma unroll
            for (int st = kUnionTrees / 2; st >= 1; st >>= 1)
              if (t + st <= trees && tr_p[t + st] <= e) t += st;
            const int p = e - tr_p[t];
            jv[u] = perm[static_cast<int64_t>(t) * n + tr_s[t] + p];
            kv[u] = ktab[tr_o[t] + p];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int u = 0; u < kUnionSlots; ++u)
        if (jv[u] == i) jv[u] = -1;
      if (!win_clean)
        for (int w = tid; w < kWinWords / 4; w += 256) reinterpret_cast<uint4*>(win)[w] = make_uint4(0u, 0u, 0u, 0u);
      win_clean = true;
      summ_clean = false;
      __syncthreads();
      // ids are below 2^31: the window offset of an absent member (-1)
      // wraps far past the window
      int m = 0;
      for (int32_t w0 = 0; w0 < n; w0 += kWinIds) {
        // the members are opaque to the compiler at each window (empty asm,
        // no instruction), so nothing derived from them is kept live across
        // windows: two registers per slot
#pragma unroll
        for (int u = 0; u < kUnionSlots; ++u) asm volatile("" : "+v"(jv[u]), "+v"(kv[u]));
#pragma unroll
        for (int b0 = 0; b0 < kUnionSlots; b0 += kGatherBatch)
          if (b0 < nslots) {
#pragma unroll
            for (int u = b0; u < b0 + kGatherBatch; ++u) {
              const uint32_t r = static_cast<uint32_t>(jv[u] - w0);
              if (r < kWinIds) atomicOr(&win[r >> 5], 1u << (r & 31));
            }
          }
        __syncthreads();
        // this thread's words [kWordsPerThread * tid, ...): set bits counted
        // per group of 8 words, scanned over the block, and each group's
        // window-relative prefix published (below 2^16: at most kUnionCap
        // members)
        const uint4* wt = reinterpret_cast<const uint4*>(win) + kWordsPerThread / 4 * tid;
        int gc[kGroupsPerThread];
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < kGroupsPerThread; ++q) {
          const uint4 a4 = wt[2 * q], b4 = wt[2 * q + 1];
          gc[q] = __popc(a4.x) + __popc(a4.y) + __popc(a4.z) + __popc(a4.w) + __popc(b4.x) + __popc(b4.y) +
                  __popc(b4.z) + __popc(b4.w);
          cnt += gc[q];
        }
        const int incl = wave_incl_add(cnt);
        if (lane == 63) sh[wave] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += sh[w];
        const int wtot = sh[0] + sh[1] + sh[2] + sh[3];
        {
          int c = before + incl - cnt;
          uint32_t pk[kGroupsPerThread / 2];
#pragma unroll
          for (int q = 0; q < kGroupsPerThread / 2; ++q) {
            const int c0 = c;
            c += gc[2 * q];
            pk[q] = static_cast<uint32_t>(c0) | (static_cast<uint32_t>(c) << 16);
            c += gc[2 * q + 1];
          }
          reinterpret_cast<uint4*>(gpre)[tid] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        }
        __syncthreads();
        // ranks: the group prefix and the group's eight words of every member
        // requested before any is used (an absent member reads group 0 and
        // writes nothing); the members go to the workgroup's global array
#pragma unroll
        for (int b0 = 0; b0 < kUnionSlots; b0 += kRankBatch)
          if (b0 < nslots) {
            uint4 wa[kRankBatch], wb[kRankBatch];
            uint32_t pv[kRankBatch];
#pragma unroll
            for (int u = 0; u < kRankBatch; ++u) {
              const uint32_t r = static_cast<uint32_t>(jv[b0 + u] - w0);
              const uint32_t gi = r < kWinIds ? r >> 8 : 0u;
              wa[u] = reinterpret_cast<const uint4*>(win)[2 * gi];
              wb[u] = reinterpret_cast<const uint4*>(win)[2 * gi + 1];
              pv[u] = gpre[gi];
            }
#pragma unroll
            for (int u = 0; u < kRankBatch; ++u) {
              const uint32_t r = static_cast<uint32_t>(jv[b0 + u] - w0);
              if (r < kWinIds) {
                const uint32_t q = (r >> 5) & 7u;
                const uint32_t w8[8] = {wa[u].x, wa[u].y, wa[u].z, wa[u].w, wb[u].x, wb[u].y, wb[u].z, wb[u].w};
                int rank = m + static_cast<int>(pv[u]);
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t) {
                  const uint32_t msk = t < q ? ~0u : (t == q ? (1u << (r & 31)) - 1u : 0u);
                  rank += __popc(w8[t] & msk);
                }
                uo[rank] = pack_pair(kv[b0 + u], order_id(jv[b0 + u]));
              }
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        __syncthreads();
        {
          uint4* wc = reinterpret_cast<uint4*>(win) + kWordsPerThread / 4 * tid;
#pragma unroll
          for (int q = 0; q < kWordsPerThread / 4; ++q) wc[q] = make_uint4(0u, 0u, 0u, 0u);
        }
        m += wtot;
        __syncthreads();
      }
      // the members (global writes ordered before these reads by the
      // workgroup barriers) into LDS over the window: at most kUnionCap
      uint64_t* cp = reinterpret_cast<uint64_t*>(lds);
      for (int e0 = 0; e0 < m; e0 += 256 * 8) {
        uint64_t t8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + 256 * u + tid;
          t8[u] = e < m ? uo[e] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int e = e0 + 256 * u + tid;
          if (e < m) cp[e] = t8[u];
        }
      }
      if (m > 0) win_clean = false;
      __syncthreads();
      const int nch = static_cast<int>((m + cap - 1) / cap);
      if (nch * k <= kLocalCap && k <= 64) {
        // The chunk loop's selections, split: the first k of (carried
        // list + chunk) are the first k of (carried list + the chunk's own
        // first k), so each chunk's own first k is taken first, four chunks
        // at a time (one wave each), and wave 0 then merges them into the
        // carried list in chunk order, flipping and sorting as before.
        uint32_t* wh = reinterpret_cast<uint32_t*>(gpre) + 256 * wave;
        for (int c = wave; c < nch; c += 4) {
          const int c0 = c * static_cast<int>(cap);
          const int nc = m - c0 < cap ? m - c0 : static_cast<int>(cap);
          const int got = wave_topk(cp + c0, nc, k, locl + c * k, wh);
          wave_sync();
          // the chunk's list in ascending packed order
          const uint64_t v = wave_sort_u64(lane < got ? locl[c * k + lane] : ~0ull);
          if (lane < got) locl[c * k + lane] = v;
          wave_sync();
        }
        __syncthreads();
        if (wave == 0) {
          // the carried list stays in ascending packed order while no stored
          // value is NaN (the (stored, id) order is then the packed order);
          // the ranks then come from binary searches, else from counts
          bool packed_sorted = true;
          for (int c = 0; c < nch; ++c) {
            const int c0 = c * static_cast<int>(cap);
            const int nc = m - c0 < cap ? m - c0 : static_cast<int>(cap);
            const int nl = nc < k ? nc : k;  // the chunk's list length
            const uint64_t* lc = locl + c * k;
            // rank of each entry among (carried list + chunk list) in packed
            // order; equal carried entries (stubs) by position
            const uint64_t cv = lane < k ? car[lane] : ~0ull;
            const uint64_t lv = lane < nl ? lc[lane] : ~0ull;
            int rcar = 0, rl = 0;
            if (packed_sorted) {
              rcar = lane + count_below(lc, nl, cv);
              rl = lane + count_below(car, k, lv);
            } else {
              // lane j's entries broadcast through scalar registers
#pragma unroll 4
              for (int j = 0; j < 64; ++j) {
                if (j < k) {
                  const uint64_t u = lane_u64(cv, j);
                  rcar += (u < cv || (u == cv && j < lane)) ? 1 : 0;
                  rl += u < lv ? 1 : 0;
                }
                if (j < nl) {
                  const uint64_t u = lane_u64(lv, j);
                  rcar += u < cv ? 1 : 0;
                  rl += u < lv ? 1 : 0;
                }
              }
            }
            wave_sync();
            if (lane < k && rcar < k) outl[rcar] = cv;
            if (lane < nl && rl < k) outl[rl] = lv;
            wave_sync();
            float key = __builtin_inff();
            int32_t id = INT_MAX;
            if (lane < k) {
              const uint64_t v = outl[lane];
              key = static_cast<float>(1.0 - static_cast<double>(pair_key(v)));
              id = pair_id(v);
            }
            // 1 - key reverses the order: with the first k ascending and no
            // two stored values equal (nor NaN), lane l goes to k - 1 - l;
            // otherwise the (stored, id) sort
            const float nkey = __shfl_down(key, 1);
            const bool nan = key != key;
            const bool tie = lane + 1 < k && nkey == key;
            if (packed_sorted && __ballot(lane < k && (nan || tie)) == 0) {
              if (lane < k) car[k - 1 - lane] = pack_pair(key, order_id(id));
            } else {
              wave_sort(key, id);
              if (lane < k) car[lane] = pack_pair(key, order_id(id));
              if (__ballot(lane < k && nan) != 0) packed_sorted = false;
            }
            wave_sync();
          }
        }
        __syncthreads();
      } else {
        for (int c0 = 0; c0 < m; c0 += static_cast<int>(cap)) {
          const int nc = static_cast<int>(c0 + cap < m ? cap : m - c0);
          select_chunk(cp + c0, nc);
        }
      }
    } else {
      if (!summ_clean)
        for (int w = tid; w < sum_words; w += 256) summ[w] = 0;
      summ_clean = true;
      win_clean = false;
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
      for (int64_t c0 = 0; c0 < m; c0 += cap) {
        const int nc = static_cast<int>(c0 + cap < m ? cap : m - c0);
        for (int e = tid; e < nc; e += 256) {
          const int32_t j = cand[c0 + e];
          const float key = TABLE ? kid[j] : ckey[c0 + e];
          pool[e] = pack_pair(key, order_id(j));
        }
        __syncthreads();
        select_chunk(pool, nc);
      }
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
  (void)k;
  const size_t xi = table ? 0 : static_cast<size_t>((d + 3) / 4 * 4) * 4;
  const size_t summ = static_cast<size_t>((summary_words + 3) / 4 * 4) * 4;
  const size_t bitmap_path = xi + summ + static_cast<size_t>(cap) * 8;
  // the LDS union's layout (distinct members, bitmap window) shares the region
  const size_t union_path = table ? size_t{kWinWords} * 4 : 0;
  return bitmap_path > union_path ? bitmap_path : union_path;
}

void launch_cosine_parity_initial(const float* x, int d, int ld, const float* norms, int64_t n, int k, int trees,
                                  const int32_t* rs, const int32_t* rc, const int32_t* perm, int64_t cap,
                                  const CosineParityScratch& sc, int32_t* ids, float* vals, const float* ktab,
                                  const int64_t* koff, hipStream_t s) {
  if (k > 64) throw HipError("COSINE parity path supports k <= 64");
  if (sc.groups % 8 != 0) throw HipError("COSINE parity path: groups must be a multiple of 8");
  const bool table = ktab != nullptr;
  if (table && (koff == nullptr || sc.keyarr == nullptr || sc.uo == nullptr))
    throw HipError("COSINE parity path: key table without offsets or scratch");
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

template <int METRIC>
void launch_keys_m(const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* items,
                   int nitems, int chunks, const int32_t* rs, const int32_t* rc, const int64_t* off, float* keys,
                   hipStream_t s) {
  const int64_t grid = xcd_grid(static_cast<int64_t>(nitems) * chunks);
  if (d > 256)
    hipLaunchKernelGGL((range_keys_kernel<METRIC, true>), dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, x, d,
                       ld, norms, perm, items, nitems, chunks, rs, rc, off, keys);
  else
    hipLaunchKernelGGL((range_keys_kernel<METRIC, false>), dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, x, d,
                       ld, norms, perm, items, nitems, chunks, rs, rc, off, keys);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_range_keys(int metric, const float* x, int d, int ld, const float* norms, const int32_t* perm,
                       const Leaf* items, int nitems, int max_owners, const int32_t* rs, const int32_t* rc,
                       const int64_t* off, float* keys, hipStream_t s) {
  if (nitems <= 0) return;
  if (ld % 4 != 0 || ld < d) throw HipError("range keys: row stride must be a multiple of 4 covering d");
  if (metric == kCosine && norms == nullptr) throw HipError("range keys: COSINE needs the norms");
  const int chunks = (max_owners + kTile - 1) / kTile;
  switch (metric) {
    case kEuclidean:
      launch_keys_m<kEuclidean>(x, d, ld, norms, perm, items, nitems, chunks, rs, rc, off, keys, s);
      break;
    case kTaxicab:
      launch_keys_m<kTaxicab>(x, d, ld, norms, perm, items, nitems, chunks, rs, rc, off, keys, s);
      break;
    case kTaxicabL1:
      launch_keys_m<kTaxicabL1>(x, d, ld, norms, perm, items, nitems, chunks, rs, rc, off, keys, s);
      break;
    case kCosine:
      launch_keys_m<kCosine>(x, d, ld, norms, perm, items, nitems, chunks, rs, rc, off, keys, s);
      break;
    default:
      throw HipError("unknown distance");
  }
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
