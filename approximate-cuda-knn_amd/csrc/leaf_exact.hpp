// The exact leaf k-NN of one work item by one wave (EUCLIDEAN, TAXICAB,
// COSINE with carried distances), merged into the owners' graph rows: the
// body of the per-tree kernel (leaf_knn.hip) and of the path of the
// all-trees launch (leaf_mfma.hip) that runs it.
//
// Reference: findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN
// (euclideandistancematrixgpu.cu:107-415, :417-712, :714-1031).
#pragma once

#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"
#include "wave_ops.hpp"

namespace aknn {
namespace leafx {

constexpr int kTile = 64;
constexpr int kDc = 32;   // dimensions per LDS stage (divides 256)
constexpr int kPad = 4;

constexpr int kStageFloats = 2 * kDc * (kTile + kPad);
static_assert(kStageFloats >= kTile * (kTile + 1), "the distance block reuses the staging tiles");

// LDS of one wave: the transposed staging tiles (dimension-major) and the
// 64x64 distance block are never live at the same time and share smem.
struct ExactLds {
  alignas(16) float smem[kStageFloats];
  uint32_t cox[kTile];  // candidate ids through order_id
  float onorm[kTile];
  float cnorm[kTile];
};

struct ExactArgs {
  const float* x;
  int d, ld;
  const float* norms;
  const int32_t* perm;
  const Leaf* leaves;
  int nleaves, chunks_per_leaf;
  const int32_t* rsp;
  const int32_t* rcp;
  int k;
  int32_t* gids;
  float* gvals;
};

// LDS ordering inside an item: a workgroup barrier where the wave is the
// whole workgroup, the wave's own order where several waves of a workgroup
// take items independently.
template <bool kGroupSync>
__device__ __forceinline__ void item_sync() {
  if constexpr (kGroupSync)
    __syncthreads();
  else
    wave_sync();
}

// One (leaf, 64-owner chunk) or work item lb, by one wave.
template <int METRIC, int KMAX, bool BIGD, bool kGroupSync = true>
__device__ __forceinline__ void exact_item(const ExactArgs& ea, int64_t lb, ExactLds& sm) {
  const float* __restrict__ x = ea.x;
  const int d = ea.d, ld = ea.ld;
  const float* __restrict__ norms = ea.norms;
  const int32_t* __restrict__ perm = ea.perm;
  const Leaf* __restrict__ leaves = ea.leaves;
  const int nleaves = ea.nleaves, chunks_per_leaf = ea.chunks_per_leaf;
  const int32_t* __restrict__ rsp = ea.rsp;
  const int32_t* __restrict__ rcp = ea.rcp;
  const int k = ea.k;
  int32_t* __restrict__ gids = ea.gids;
  float* __restrict__ gvals = ea.gvals;
  float* smem = sm.smem;
  float(*As)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem);
  float(*Bs)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem + kDc * (kTile + kPad));
  float(*Dt)[kTile + 1] = reinterpret_cast<float(*)[kTile + 1]>(smem);
  uint32_t* cox = sm.cox;
  float* onorm = sm.onorm;
  float* cnorm = sm.cnorm;

  const int64_t leaf_idx = lb / chunks_per_leaf;
  if (leaf_idx >= nleaves) return;
  const Leaf L = leaves[leaf_idx];
  const int o0 = static_cast<int>(lb % chunks_per_leaf) * kTile;
  if (o0 >= L.size) return;
  const int lane = threadIdx.x & 63;
  const int nown = L.size - o0 < kTile ? L.size - o0 : kTile;
  const bool has_owner = lane < nown;
  const int32_t my_id = has_owner ? perm[L.start + o0 + lane] : -1;
  // this owner's own start-node positions; a work item may group owners of
  // several leaves and scan the union of their ranges
  const int32_t r_lo = rsp != nullptr && has_owner ? rsp[my_id] : L.rstart;
  const int32_t r_hi = rsp != nullptr && has_owner ? r_lo + rcp[my_id] : L.rstart + L.rcount;
  if constexpr (METRIC == kCosine) onorm[lane] = has_owner ? norms[my_id] : 1.0f;

  // this owner's running top-k, seeded with its graph row (earlier trees)
  const int64_t row0 = static_cast<int64_t>(has_owner ? my_id : 0) * k;
  uint64_t pk[KMAX];
  topk_load<KMAX>(pk, k, has_owner, gvals + row0, gids + row0);
  const uint32_t my_ox = order_id(my_id);

  const int r = lane >> 3, c = lane & 7;
  const float* orow = x + static_cast<int64_t>(my_id < 0 ? 0 : my_id) * ld;
  for (int cb = 0; cb < L.rcount; cb += kTile) {
    const int ncand = L.rcount - cb < kTile ? L.rcount - cb : kTile;
    item_sync<kGroupSync>();
    const int32_t my_c = lane < ncand ? perm[L.rstart + cb + lane] : -1;
    cox[lane] = order_id(my_c);
    if constexpr (METRIC == kCosine) cnorm[lane] = lane < ncand ? norms[my_c] : 1.0f;

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

    const float* crow = x + static_cast<int64_t>(my_c < 0 ? 0 : my_c) * ld;
    for (int d0 = 0; d0 < d; d0 += kDc) {
      if constexpr (BIGD) {
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
      item_sync<kGroupSync>();
#pragma unroll
      for (int q = 0; q < kDc; q += 4) {
        float4 ov = make_float4(0.f, 0.f, 0.f, 0.f), cv = ov;
        if (d0 + q < d) {
          if (has_owner) ov = *reinterpret_cast<const float4*>(orow + d0 + q);
          if (my_c >= 0) cv = *reinterpret_cast<const float4*>(crow + d0 + q);
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
      item_sync<kGroupSync>();
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
    item_sync<kGroupSync>();
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float t = BIGD ? tot[i][j] + acc[i][j] : 0.0f + acc[i][j];
        float an = 1.0f, bn = 1.0f;
        if constexpr (METRIC == kCosine) {
          an = onorm[8 * r + i];
          bn = cnorm[8 * c + j];
        }
        Dt[8 * r + i][8 * c + j] = finish_key<METRIC>(t, an, bn);
      }
    item_sync<kGroupSync>();
    if (has_owner) {
      // each lane first marks the candidates of its own range that undercut
      // its threshold, then offers only those: lanes do not wait on each
      // other's insertions
      const int base = L.rstart + cb;
      const int lo = r_lo - base < 0 ? 0 : r_lo - base;
      const int hi = r_hi - base > ncand ? ncand : r_hi - base;
      const uint64_t below_hi = hi >= 64 ? ~0ull : (hi > 0 ? (1ull << hi) - 1 : 0ull);
      const uint64_t below_lo = lo >= 64 ? ~0ull : (1ull << lo) - 1;
      const uint64_t th = pk[KMAX - 1];
      uint64_t surv = 0;
      for (int j = 0; j < ncand; ++j) {
        const uint32_t cj = cox[j];
        surv |= static_cast<uint64_t>(cj != my_ox && pack_pair(Dt[lane][j], cj) < th) << j;
      }
      surv &= below_hi & ~below_lo;
      while (surv) {
        const int j = __ffsll(static_cast<long long>(surv)) - 1;
        surv &= surv - 1;
        topk_offer<KMAX>(pk, pack_pair(Dt[lane][j], cox[j]));
      }
    }
  }
  if (has_owner) topk_store<KMAX>(pk, k, gvals + row0, gids + row0);
}

}  // namespace leafx
}  // namespace aknn
