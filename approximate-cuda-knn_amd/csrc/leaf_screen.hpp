// The half-precision MFMA-screened leaf k-NN of one work item by one wave:
// the body of the per-tree kernel (leaf_mfma.hip) and of the path of the
// all-trees launch (leaf_all.hip) that runs it.
//
// Same contract and results as leaf_knn.hip (the exact FP32 top-k by
// (key, id) of the graph row merged with the tree's start-node candidates,
// euclideandistancematrixgpu.cu:107-415), computed differently:
//
//   1. the 64 owners x 64 candidates block of dot products of the FP16 sketch
//      rows (sketch.hip) is formed on the matrix cores,
//      v_mfma_f32_32x32x16_f16 with FP32 accumulation;
//   2. every pair gets a proven lower bound of its exact FP32 key from
//      ||xh||^2 + ||yh||^2 - 2 xh.yh, the accumulation error, the sketch radii
//      and the FP32 chain error; a pair whose bound is above the owner's
//      current k-th key cannot enter and is dropped;
//   3. the surviving pairs (near the boundary: a few percent once the graph
//      rows hold an earlier tree's neighbours) are compacted into a list and
//      re-scored with the exact FP32 chain, spread over all 64 lanes, then
//      inserted by their owners exactly as in leaf_knn.hip.
//
// The stored keys are the exact chain's, so the graph is the same bit for bit
// as the plain kernel's.  Lane layout of the 32x32x16 MFMA (checked on the
// device): A[i][k] is in lane i + 32*(k/8), element k%8; B[k][j] in lane
// j + 32*(k/8), element k%8; D[i][j] in lane j + 32*((i/4)%2), register
// 4*(i/8) + i%4.
#pragma once

#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"
#include "leaf_exact.hpp"
#include "wave_ops.hpp"

namespace aknn {
namespace leafs {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 64;
// 16-byte steps of the re-score chain unrolled together: 2 * 8 row loads in
// flight per lane (the MFMA phase's registers are free by then)
constexpr int kRescoreUnroll = 8;
// staged re-score: dimensions per LDS stage, batches of 64 pairs advanced
// together, and the fewest pairs of a tile that take it
constexpr int kStageDims = 16;
constexpr int kStageBatches = 4;
constexpr int kStageMin = 24;

// Owner threshold.  EUCLIDEAN: the smallest T such that a pair whose sketch
// distance exceeds T + eps_candidate has an exact key above thk
// (T >= thk / (1 - kap) + eps_owner, rounded up).  COSINE: the largest T such
// that a pair whose exact quotient tot / den is below T has an exact key
// 1 - quotient above thk (T <= 1 - thk - 2^-20, rounded down; the 2^-20 covers
// the FP32 rounding of the quotient and of the stored 1 - quotient).
template <int METRIC>
__device__ __forceinline__ float owner_threshold(float thk, float eo, double kap) {
  if constexpr (METRIC == kEuclidean) {
    const double t = (static_cast<double>(thk) + 1e-30) / (1.0 - kap) + static_cast<double>(eo);
    float f = static_cast<float>(t);
    if (static_cast<double>(f) < t) f = __uint_as_float(__float_as_uint(f) + 1u);
    return f;
  } else {
    const double t = 1.0 - static_cast<double>(thk) - 9.5367431640625e-07;
    float f = static_cast<float>(t);
    if (static_cast<double>(f) > t) {  // next float down
      const uint32_t b = __float_as_uint(f);
      f = f > 0.0f ? __uint_as_float(b - 1u) : (f == 0.0f ? -1.401298464e-45f : __uint_as_float(b + 1u));
    }
    return f;
  }
}

// Square of the EUCLIDEAN sketch-distance threshold of an owner against a
// candidate of sketch radius ce, rounded up: (T + ce)^2 with margins for the
// FP32 rounding of the sum and the square.
__device__ __forceinline__ float e_threshold2(float t0, float ce) {
  const float t = (t0 + ce) * 1.0000002384185791f;              // (1 + 4u)
  return __builtin_fmaf(t, t, 1e-30f) * 1.0000004768371582f;    // (1 + 8u)
}

// bits [0, x) of a 64-bit lane mask, x in [0, 64]
__device__ __forceinline__ uint64_t lane_mask_lt(int x) { return x >= 64 ? ~0ull : (1ull << x) - 1; }

struct MfmaLeafArgs {
  const float* x;
  int d, ld;
  const _Float16* xh;
  int ldh;
  const float* eps;
  const float* eps_max;  // the largest sketch radius (infinite if a row has none)
  const float* nh;
  const float* norms;  // COSINE: the exact squared-norm chains
  const int32_t* perm;
  const Leaf* leaves;
  const int32_t* rsp;  // per-point start-node ranges (nullable: the item's range)
  const int32_t* rcp;
  int nleaves, chunks_per_leaf, k;
  int32_t* gids;
  float* gvals;
  double kap;    // FP32 chain error of the exact keys (sketch_bound)
  float c1m;     // 1 - (MFMA accumulation error + evaluation margin), relative to ||xh||^2 + ||yh||^2
  float kmf;     // COSINE: MFMA accumulation error relative to ||xh|| ||yh||
  float kabf;    // COSINE: kap + 4u
  unsigned long long* rescored;  // optional: exactly re-scored pairs
  unsigned long long* screened;  // optional: screened pairs (owners x candidates)
};

// COSINE pair terms: up(G) = G + kM (A+eo)(B+ec) + eo B + (A+eo) ec
// + (kap + 4u) A B bounds the exact dot chain from the sketch dot G (MFMA
// error, sketch radii, exact chain error, quotient rounding), with
// A >= ||x_o||, B >= ||x_c|| from the exact squared norms.
__device__ __forceinline__ float norm_upper(float aa, double kap, double tiny) {
  const double t = sqrt((static_cast<double>(aa) + tiny) * (1.0 + 2.0 * kap));
  float f = static_cast<float>(t);
  if (static_cast<double>(f) < t) f = __uint_as_float(__float_as_uint(f) + 1u);
  return f;
}

// LDS of one wave of the screened item.
struct ScreenLds {
  alignas(16) int32_t oid[kTile];
  int32_t cid[kTile];
  alignas(16) float o_n[kTile];  // E: ||xh||^2; C: A (norm bound)
  alignas(16) float o_t[kTile];  // owner threshold
  alignas(16) float o_e[kTile];  // C: sketch radius
  alignas(16) float o_s[kTile];  // C: sqrtf(aa), as the exact key uses
  float c_n[kTile], c_e[kTile], c_s[kTile];
  uint64_t kp[kTile];   // re-scored pairs that can enter, packed for the owners' top-k
  uint64_t o_k[kTile];  // each owner's current k-th pair
  alignas(16) int32_t o_lo[kTile];  // owner's start-node positions [lo, hi)
  alignas(16) int32_t o_hi[kTile];
  uint16_t plist[kTile * kTile];
  alignas(16) float stg[2 * kTile * kStageDims];  // staged owner and candidate rows
};

// One work item lb (owners of one leaf, or grouped small leaves) by one wave.
// (kGroupSync: LDS ordered by workgroup barriers, the wave being the whole
// workgroup; else by the wave's own order, see leafx::item_sync)
template <int METRIC, int KMAX, bool kGroupSync = true>
__device__ __forceinline__ void screen_item(const MfmaLeafArgs& a, int64_t lb, ScreenLds& sm) {
  int32_t* oid = sm.oid;
  int32_t* cid = sm.cid;
  float* o_n = sm.o_n;
  float* o_t = sm.o_t;
  float* o_e = sm.o_e;
  float* o_s = sm.o_s;
  float* c_n = sm.c_n;
  float* c_e = sm.c_e;
  float* c_s = sm.c_s;
  uint64_t* kp = sm.kp;
  uint64_t* o_k = sm.o_k;
  int32_t* o_lo = sm.o_lo;
  int32_t* o_hi = sm.o_hi;
  uint16_t* plist = sm.plist;
  float* stg = sm.stg;

  const int64_t leaf_idx = lb / a.chunks_per_leaf;
  if (leaf_idx >= a.nleaves) return;
  const Leaf L = a.leaves[leaf_idx];
  const int o0 = static_cast<int>(lb % a.chunks_per_leaf) * kTile;
  if (o0 >= L.size) return;
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5, col = lane & 31;
  const int nown = L.size - o0 < kTile ? L.size - o0 : kTile;
  const bool has_owner = lane < nown;
  const int32_t my_id = has_owner ? a.perm[L.start + o0 + lane] : -1;
  const int k = a.k;

  const int64_t row0 = static_cast<int64_t>(has_owner ? my_id : 0) * k;
  uint64_t pk[KMAX];
  topk_load<KMAX>(pk, k, has_owner, a.gvals + row0, a.gids + row0);
  const float my_eps = has_owner ? a.eps[my_id] : 0.0f;
  const double tiny = 1e-40;
  oid[lane] = my_id;
  o_lo[lane] = a.rsp != nullptr && has_owner ? a.rsp[my_id] : L.rstart;
  o_hi[lane] = a.rsp != nullptr && has_owner ? a.rsp[my_id] + a.rcp[my_id] : L.rstart + L.rcount;
  if constexpr (METRIC == kEuclidean) {
    o_n[lane] = has_owner ? a.nh[my_id] : 0.0f;
  } else {
    const float aa = has_owner ? a.norms[my_id] : 0.0f;
    o_n[lane] = norm_upper(aa, a.kap, tiny);
    o_s[lane] = sqrtf(aa);
    o_e[lane] = my_eps;
  }
  // EUCLIDEAN: with every candidate's sketch radius bounded by the largest
  // (eps_max), the pair threshold depends on the owner only, so each owner
  // holds its squared threshold; an infinite largest radius (a row the
  // sketch cannot hold) leaves the candidates' own radii in the test
  const float ce_max = *a.eps_max;
  const bool per_cand = !(ce_max < __builtin_inff());
  auto owner_t = [&](uint64_t kth) {
    const float t = owner_threshold<METRIC>(pair_key(kth), my_eps, a.kap);
    if constexpr (METRIC == kEuclidean) return per_cand ? t : e_threshold2(t, ce_max);
    return t;
  };
  o_t[lane] = owner_t(pk[KMAX - 1]);
  if constexpr (METRIC == kEuclidean) o_n[lane] *= a.c1m;
  const int32_t my_lo = o_lo[lane], my_hi = o_hi[lane];
  const int32_t my_pos = L.start + o0 + lane;  // the owner's own position: its self pair
  leafx::item_sync<kGroupSync>();
  const int32_t oa = oid[col], ob = oid[32 + col];
  const _Float16* arow0s = a.xh + static_cast<int64_t>(oa >= 0 ? oa : 0) * a.ldh;
  const _Float16* arow1s = a.xh + static_cast<int64_t>(ob >= 0 ? ob : 0) * a.ldh;
  const int steps = (a.ldh + 15) / 16;
  const bool ktail = (a.ldh % 16) != 0;  // the last step's upper half lies past the row
  unsigned long long nres = 0;

  for (int cb = 0; cb < L.rcount; cb += kTile) {
    const int ncand = L.rcount - cb < kTile ? L.rcount - cb : kTile;
    leafx::item_sync<kGroupSync>();
    const int32_t my_c = lane < ncand ? a.perm[L.rstart + cb + lane] : -1;
    cid[lane] = my_c;
    if constexpr (METRIC == kEuclidean) {
      c_n[lane] = my_c >= 0 ? a.nh[my_c] * a.c1m : 0.0f;
    } else {
      const float bb = my_c >= 0 ? a.norms[my_c] : 0.0f;
      c_n[lane] = norm_upper(bb, a.kap, tiny);
      c_s[lane] = sqrtf(bb);
    }
    c_e[lane] = my_c >= 0 ? a.eps[my_c] : 0.0f;
    leafx::item_sync<kGroupSync>();

    // 1-2. the owner tiles ot = 0, 1 in turn, each against both candidate
    //      tiles: dot products of the sketches on the matrix cores, four K
    //      steps of fragments loaded before their MFMAs (one A fragment feeds
    //      both candidate tiles), then the screen.  This lane holds candidate
    //      column j = 32*ct + col against owner rows i = 32*ot + 8*g +
    //      4*half + q (register 4*g + q).  Loads are never behind a branch:
    //      rows of absent owners or candidates read row 0 (their pairs are
    //      masked out below), and the K tail past ldh is zeroed.  One ballot
    //      per (g, q, ct) gives the 32 candidates' verdicts for owner
    //      32*ot + 8*g + q (low half) and owner ... + 4 (high half); each
    //      owner lane keeps its half as its mask word for tile ct.
    const _Float16* bsafe0 = a.xh + static_cast<int64_t>(cid[col] >= 0 ? cid[col] : 0) * a.ldh;
    const _Float16* bsafe1 = a.xh + static_cast<int64_t>(cid[32 + col] >= 0 ? cid[32 + col] : 0) * a.ldh;
    const float cn0 = c_n[col], cn1 = c_n[32 + col];
    const float ce0 = c_e[col], ce1 = c_e[32 + col];
    // this owner lane's survivor bits for candidate tiles 0 and 1; owner
    // lane 32*ot + 8*g + 4*h + q takes half h of the ballots at (ot, g, q)
    uint32_t mw0 = 0, mw1 = 0;
    const int lkey = lane & 0x3B;
    const bool hbit = (lane & 4) != 0;
#pragma unroll 1
    for (int ot = 0; ot < 2; ++ot) {
      const _Float16* asafe = ot == 0 ? arow0s : arow1s;
      f32x16 acc2[2];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        acc2[0][r] = 0.0f;
        acc2[1][r] = 0.0f;
      }
      for (int s0 = 0; s0 < steps; s0 += 4) {
        uint4 av[4], bv0[4], bv1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = 16 * (s0 + u) + 8 * half;
          const int kc = kk < a.ldh ? kk : a.ldh - 8;
          av[u] = *reinterpret_cast<const uint4*>(asafe + kc);
          bv0[u] = *reinterpret_cast<const uint4*>(bsafe0 + kc);
          bv1[u] = *reinterpret_cast<const uint4*>(bsafe1 + kc);
        }
        if (ktail && s0 + 4 >= steps) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (16 * (s0 + u) + 8 * half >= a.ldh) {
              av[u] = make_uint4(0u, 0u, 0u, 0u);
              bv0[u] = av[u];
              bv1[u] = av[u];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (s0 + u >= steps) break;
          const half8 ah = __builtin_bit_cast(half8, av[u]);
          acc2[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, __builtin_bit_cast(half8, bv0[u]), acc2[0], 0, 0, 0);
          acc2[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, __builtin_bit_cast(half8, bv1[u]), acc2[1], 0, 0, 0);
        }
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int i0 = 32 * ot + 8 * g + 4 * half;
        const float4 on4 = *reinterpret_cast<const float4*>(&o_n[i0]);
        const float4 to4 = *reinterpret_cast<const float4*>(&o_t[i0]);
        const float on[4] = {on4.x, on4.y, on4.z, on4.w};
        const float to[4] = {to4.x, to4.y, to4.z, to4.w};
        float oe[4] = {0.0f, 0.0f, 0.0f, 0.0f}, os[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (METRIC == kCosine) {
          const float4 oe4 = *reinterpret_cast<const float4*>(&o_e[i0]);
          const float4 os4 = *reinterpret_cast<const float4*>(&o_s[i0]);
          oe[0] = oe4.x; oe[1] = oe4.y; oe[2] = oe4.z; oe[3] = oe4.w;
          os[0] = os4.x; os[1] = os4.y; os[2] = os4.z; os[3] = os4.w;
        }
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          const f32x16& acc = acc2[ct];
          const float cn = ct == 0 ? cn0 : cn1, ce = ct == 0 ? ce0 : ce1;
          const float cs = METRIC == kCosine ? c_s[32 * ct + col] : 0.0f;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            bool keep;
            if constexpr (METRIC == kEuclidean) {
              // on and cn carry the factor c1m: s_lo is a lower bound of the
              // sketch squared distance, to the owner threshold's square
              const float s_lo = __builtin_fmaf(-2.0f, acc[4 * g + q], on[q] + cn);
              keep = !(s_lo > (per_cand ? e_threshold2(to[q], ce) : to[q]));
            } else {
              // reject when up(G) + rounding margin < T_o * den * (1 - 4u)
              const float G = acc[4 * g + q];
              const float A = on[q], B = cn, eo = oe[q];
              const float Ao = A + eo;
              const float P = a.kmf * Ao * (B + ce) + eo * B + Ao * ce + a.kabf * A * B;
              const float lhs = G + P + 4.76837158203125e-07f * (__builtin_fabsf(G) + P + A * B) + 1e-30f;
              const float den = os[q] * cs;
              const float rhs = to[q] * den * 0.99999976158142090f - 1e-30f;  // (1 - 4u)
              keep = !(den > 0.0f && den <= 3.4e38f && lhs < rhs);
            }
            const uint64_t b = __ballot(keep);
            const uint32_t w = hbit ? static_cast<uint32_t>(b >> 32) : static_cast<uint32_t>(b);
            if (lkey == 32 * ot + 8 * g + q) {
              if (ct == 0)
                mw0 = w;
              else
                mw1 = w;
            }
          }
        }
      }
    }

    // 3. survivors of this owner: its own start-node positions within the
    //    tile (which also drops absent candidates), minus its self pair, minus
    //    candidates already in its row (found by an earlier tree: close
    //    points share start nodes; their exact keys are already stored).
    const int32_t tbase = L.rstart + cb;
    const int lo_c = my_lo - tbase > 0 ? my_lo - tbase : 0;
    const int hi_c = my_hi - tbase < ncand ? my_hi - tbase : ncand;
    uint64_t m = 0;
    if (has_owner && hi_c > lo_c) {
      m = (static_cast<uint64_t>(mw1) << 32) | mw0;
      m &= lane_mask_lt(hi_c) & ~lane_mask_lt(lo_c);
      const int sp = my_pos - tbase;
      if (sp >= 0 && sp < kTile) m &= ~(1ull << sp);
      uint64_t mm = m;
      while (mm) {
        const int c = __ffsll(static_cast<long long>(mm)) - 1;
        mm &= mm - 1;
        if (topk_has<KMAX>(pk, order_id(cid[c]))) m &= ~(1ull << c);
      }
    }
    const int cnt = __popcll(m);
    int off = cnt;
#pragma unroll
    for (int dl = 1; dl < 64; dl <<= 1) {
      const int t = __shfl_up(off, dl);
      if (lane >= dl) off += t;
    }
    const int total = __shfl(off, 63);
    off -= cnt;
    if (total == 0) continue;
    {
      uint64_t mm = m;
      int w = off;
      while (mm) {
        const int c = __ffsll(static_cast<long long>(mm)) - 1;
        mm &= mm - 1;
        plist[w++] = static_cast<uint16_t>((lane << 6) | c);
      }
    }
    nres += total;
    o_k[lane] = has_owner ? pk[KMAX - 1] : 0ull;
    leafx::item_sync<kGroupSync>();
    // exact keys; a pair that does not undercut its owner's current k-th
    // pair cannot enter and is not handed over.  The pairs that can are
    // compacted in list order (owner-major), so each owner finds its own
    // from the ballot and its list segment [off, off + cnt).  None of them
    // is in its owner's row: the survivor list left out the row's ids, and
    // the tiles of a range hold distinct candidates.
    auto hand_over = [&](int base, uint64_t v) {
      const uint64_t ins = __ballot(v != ~0ull);
      if (v != ~0ull) kp[__popcll(ins & lane_mask_lt(lane))] = v;
      leafx::item_sync<kGroupSync>();
      if (has_owner) {
        const int lo = (off > base ? off : base) - base;
        const int hi = (off + cnt < base + kTile ? off + cnt : base + kTile) - base;
        if (lo < hi) {
          const int q0 = __popcll(ins & lane_mask_lt(lo)), q1 = __popcll(ins & lane_mask_lt(hi));
#pragma unroll 4
          for (int q = q0; q < q1; ++q) topk_insert<KMAX>(pk, kp[q]);
          o_k[lane] = pk[KMAX - 1];
        }
      }
      leafx::item_sync<kGroupSync>();
    };
    if (total >= kStageMin) {
      // The tile's 64 owner and 64 candidate rows are staged through LDS
      // kStageDims dimensions at a time with coalesced loads (four lanes per
      // row), and up to kStageBatches batches of 64 pairs advance their
      // chains from LDS: each chain is the exact chain's steps in dimension
      // order, with its 256-dimension partial sums added in the same order.
      // Lane l stages the quarter l & 3 of rows (l >> 2) + 16 j: rows below
      // 64 are owners, the rest candidates (absent ones stage zeros).
      const float* src[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = (lane >> 2) + 16 * j;
        const int32_t rid = r < kTile ? oid[r] : cid[r - kTile];
        src[j] = rid >= 0 ? a.x + static_cast<int64_t>(rid) * a.ld : nullptr;
      }
      const int q4 = 4 * (lane & 3);
      for (int bg = 0; bg < total; bg += kTile * kStageBatches) {
        float pp[kStageBatches], tt[kStageBatches];
        int ep[kStageBatches];
#pragma unroll
        for (int b = 0; b < kStageBatches; ++b) {
          pp[b] = 0.0f;
          tt[b] = 0.0f;
          const int pi = bg + kTile * b + lane;
          ep[b] = pi < total ? static_cast<int>(plist[pi]) : -1;
        }
        // with lists of up to 24 the next stage's rows are requested before
        // this stage's chains run (longer lists keep the registers for two
        // waves per SIMD and load each stage in turn)
        constexpr bool kAhead = KMAX <= 24;
        auto stage_load = [&](float4 (&v)[8], int c) {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            v[j] = src[j] != nullptr && c + q4 < a.d ? *reinterpret_cast<const float4*>(src[j] + c + q4)
                                                     : make_float4(0.f, 0.f, 0.f, 0.f);
        };
        float4 v[8];
        if constexpr (kAhead) stage_load(v, 0);
        for (int c0 = 0; c0 < a.d; c0 += kStageDims) {
          if constexpr (!kAhead) stage_load(v, c0);
          leafx::item_sync<kGroupSync>();  // the previous stage is consumed
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int r = (lane >> 2) + 16 * j;
            *reinterpret_cast<float4*>(stg + r * kStageDims + q4) = v[j];
          }
          leafx::item_sync<kGroupSync>();
          if constexpr (kAhead) stage_load(v, c0 + kStageDims);
          const int nd = a.d - c0 < kStageDims ? a.d - c0 : kStageDims;
          const bool cut = c0 > 0 && (c0 & 255) == 0;  // a new 256-dimension partial sum
#pragma unroll
          for (int b = 0; b < kStageBatches; ++b) {
            if (ep[b] >= 0) {
              if (cut) {
                tt[b] = tt[b] + pp[b];
                pp[b] = 0.0f;
              }
              const float* so = stg + (ep[b] >> 6) * kStageDims;
              const float* sc = stg + (kTile + (ep[b] & 63)) * kStageDims;
              float p = pp[b];
              if (nd == kStageDims) {
#pragma unroll
                for (int g = 0; g < kStageDims / 4; ++g) {
                  const float4 av = *reinterpret_cast<const float4*>(so + 4 * g);
                  const float4 bv = *reinterpret_cast<const float4*>(sc + 4 * g);
                  p = chain_step<METRIC>(av.x, bv.x, p);
                  p = chain_step<METRIC>(av.y, bv.y, p);
                  p = chain_step<METRIC>(av.z, bv.z, p);
                  p = chain_step<METRIC>(av.w, bv.w, p);
                }
              } else {
                for (int dd = 0; dd < nd; ++dd) p = chain_step<METRIC>(so[dd], sc[dd], p);
              }
              pp[b] = p;
            }
          }
        }
#pragma unroll
        for (int b = 0; b < kStageBatches; ++b) {
          const int base = bg + kTile * b;
          if (base >= total) break;
          uint64_t v = ~0ull;
          if (ep[b] >= 0) {
            const int e = ep[b];
            const float t = tt[b] + pp[b];
            float key;
            if constexpr (METRIC == kEuclidean)
              key = finish_key<kEuclidean>(t, 1.0f, 1.0f);
            else
              key = finish_key<kCosine>(t, a.norms[oid[e >> 6]], a.norms[cid[e & 63]]);
            v = pack_pair(key, order_id(cid[e & 63]));
            if (!(v < o_k[e >> 6])) v = ~0ull;
          }
          hand_over(base, v);
        }
      }
    } else {
      // few pairs: one lane per pair walks both rows in global memory
      for (int base = 0; base < total; base += kTile) {
        const int p = base + lane;
        uint64_t v = ~0ull;
        if (p < total) {
          const int e = plist[p];
          const float* xo = a.x + static_cast<int64_t>(oid[e >> 6]) * a.ld;
          const float* xc = a.x + static_cast<int64_t>(cid[e & 63]) * a.ld;
          float key;
          if constexpr (METRIC == kEuclidean)
            key = finish_key<kEuclidean>(row_chain_global<kEuclidean, kRescoreUnroll>(xo, xc, a.d), 1.0f, 1.0f);
          else
            key = finish_key<kCosine>(row_chain_global<kCosine, kRescoreUnroll>(xo, xc, a.d), a.norms[oid[e >> 6]],
                                      a.norms[cid[e & 63]]);
          v = pack_pair(key, order_id(cid[e & 63]));
          if (!(v < o_k[e >> 6])) v = ~0ull;
        }
        hand_over(base, v);
      }
    }
    if (has_owner && cnt > 0) o_t[lane] = owner_t(pk[KMAX - 1]);
  }
  // counters spread over slots one 128-byte line apart (summed by the host)
  const int64_t cslot = ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kCountSlots) * kCountStride;
  if (a.rescored != nullptr && lane == 0 && nres > 0) atomicAdd(a.rescored + cslot, nres);
  if (a.screened != nullptr && lane == 0) {
    unsigned long long pairs = 0;  // the owners' own ranges
    for (int o = 0; o < nown; ++o) pairs += static_cast<unsigned long long>(o_hi[o] - o_lo[o]);
    atomicAdd(a.screened + cslot, pairs);
  }
  if (has_owner) topk_store<KMAX>(pk, k, a.gvals + row0, a.gids + row0);
}

// The fields of a screened launch that do not depend on the tree: the
// points, the sketch and its bounds, the graph, the counters.  The caller
// sets perm, leaves (work items), rsp, rcp, nleaves and chunks_per_leaf.
inline MfmaLeafArgs screen_args(int metric, const float* x, int d, int ld, const _Float16* xh, int ldh,
                                const float* eps, const float* eps_max, const float* nh, const float* norms, int k,
                                int32_t* ids, float* vals, unsigned long long* rescored,
                                unsigned long long* screened) {
  (void)metric;
  MfmaLeafArgs a{};
  a.x = x;
  a.d = d;
  a.ld = ld;
  a.xh = xh;
  a.ldh = ldh;
  a.eps = eps;
  a.eps_max = eps_max;
  a.nh = nh;
  a.norms = norms;
  a.chunks_per_leaf = 1;
  a.k = k;
  a.gids = ids;
  a.gvals = vals;
  a.kap = sketch_bound(d).kap;
  // accumulation error of the MFMA sums over ldh terms (|G - xh.yh| <=
  // (ldh + 16) u sqrt(nx ny) <= (ldh + 16) u (nx + ny) / 2, taken 8 times
  // over for unspecified internal rounding), the FP32 rounding of the stored
  // sketch norms and of the test itself
  const double u = 1.0 / 16777216.0;
  const double km = 8.0 * (ldh + 16) * u;
  a.c1m = static_cast<float>(1.0 - (2.0 * km + 16.0 * u));
  a.kmf = static_cast<float>(km * (1.0 + 1e-6));
  a.kabf = static_cast<float>((a.kap + 4.0 * u) * (1.0 + 1e-6));
  a.rescored = rescored;
  a.screened = screened;
  return a;
}

}  // namespace leafs
}  // namespace aknn
