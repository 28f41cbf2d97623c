// Initial leaf k-NN for one tree (EUCLIDEAN, TAXICAB, COSINE with carried
// distances), merged into the graph rows in place.
//
// Reference: findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN
// (euclideandistancematrixgpu.cu:107-415, :417-712, :714-1031).  There, every
// point's candidates are the union over trees of its start node's points
// (.cu:160-191), scanned in chunks against the k entries carried so far and
// reduced by a histogram select.  When the stored value equals the selection
// key (EUCLIDEAN, TAXICAB, COSINE with distances carried) that loop keeps the
// first k of (stubs + all candidates) ordered by (key, id), so it decomposes
// exactly over trees: each tree's pass merges its start-node candidates into
// the running top-k (duplicates across trees carry identical keys and are
// dropped by id).
//
// MI355X mapping: one wave per (leaf, 64-owner chunk).  A leaf's points share
// one start node, so the wave streams that node's points in tiles of 64 and
// computes the 64x64 distance block as a register-blocked outer product (8x8
// pairs per lane, dimension chunks staged transposed in LDS).  Each pair's
// chain still runs over the dimensions in order, so every key is bit-exact.
// The block goes through LDS to the owner lanes, which keep their top-k in
// registers (seeded from the graph row, so later trees reject most pairs at
// the threshold test).
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"

namespace aknn {

namespace {

constexpr int kTile = 64;
constexpr int kDc = 32;   // dimensions per LDS stage (divides 256)
constexpr int kPad = 4;

constexpr int kStageFloats = 2 * kDc * (kTile + kPad);
static_assert(kStageFloats >= kTile * (kTile + 1), "the distance block reuses the staging tiles");

template <int METRIC, int KMAX, bool BIGD>
__global__ __launch_bounds__(64) void leaf_knn_kernel(const float* __restrict__ x, int d, int ld,
                                                      const float* __restrict__ norms, const int32_t* __restrict__ perm,
                                                      const Leaf* __restrict__ leaves, int nleaves, int chunks_per_leaf,
                                                      const int32_t* __restrict__ rsp, const int32_t* __restrict__ rcp,
                                                      int k, int32_t* __restrict__ gids, float* __restrict__ gvals) {
  // The transposed staging tiles (dimension-major) and the 64x64 distance
  // block are never live at the same time and share one LDS buffer.
  __shared__ __attribute__((aligned(16))) float smem[kStageFloats];
  float(*As)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem);
  float(*Bs)[kTile + kPad] = reinterpret_cast<float(*)[kTile + kPad]>(smem + kDc * (kTile + kPad));
  float(*Dt)[kTile + 1] = reinterpret_cast<float(*)[kTile + 1]>(smem);
  __shared__ uint32_t cox[kTile];  // candidate ids through order_id
  __shared__ float onorm[kTile];
  __shared__ float cnorm[kTile];

  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t leaf_idx = lb / chunks_per_leaf;
  if (leaf_idx >= nleaves) return;
  const Leaf L = leaves[leaf_idx];
  const int o0 = static_cast<int>(lb % chunks_per_leaf) * kTile;
  if (o0 >= L.size) return;
  const int lane = threadIdx.x;
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
    __syncthreads();
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
      __syncthreads();
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
        float an = 1.0f, bn = 1.0f;
        if constexpr (METRIC == kCosine) {
          an = onorm[8 * r + i];
          bn = cnorm[8 * c + j];
        }
        Dt[8 * r + i][8 * c + j] = finish_key<METRIC>(t, an, bn);
      }
    __syncthreads();
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

template <int METRIC, int KMAX, bool BIGD>
void launch_t(const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* leaves, int nleaves,
              int max_leaf, int k, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp, hipStream_t s) {
  const int chunks = (max_leaf + kTile - 1) / kTile;
  const int64_t grid = xcd_grid(static_cast<int64_t>(nleaves) * chunks);
  hipLaunchKernelGGL((leaf_knn_kernel<METRIC, KMAX, BIGD>), dim3(static_cast<unsigned>(grid)), dim3(kTile), 0, s, x,
                     d, ld, norms, perm, leaves, nleaves, chunks, rsp, rcp, k, ids, vals);
  AKNN_HIP_TRY(hipGetLastError());
}

template <int METRIC, bool BIGD>
void launch_k(int k, const float* x, int d, int ld, const float* norms, const int32_t* perm, const Leaf* leaves,
              int nleaves, int max_leaf, int32_t* ids, float* vals, const int32_t* rsp, const int32_t* rcp,
              hipStream_t s) {
  if (k <= 8)
    launch_t<METRIC, 8, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 16)
    launch_t<METRIC, 16, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 20)
    launch_t<METRIC, 20, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 24)
    launch_t<METRIC, 24, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 32)
    launch_t<METRIC, 32, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else if (k <= 64)
    launch_t<METRIC, 64, BIGD>(x, d, ld, norms, perm, leaves, nleaves, max_leaf, k, ids, vals, rsp, rcp, s);
  else
    throw HipError("leaf k-NN supports k <= 64");
}

// stubInitializationKernel (.cu:2110-2117)
__global__ __launch_bounds__(256) void stub_init_kernel(int32_t* __restrict__ ids, float* __restrict__ vals,
                                                        int64_t count) {
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < count; e += gridDim.x * 256ll) {
    ids[e] = kStubId;
    vals[e] = kStubDist;
  }
}

}  // namespace

void launch_stub_init(int32_t* ids, float* vals, int64_t count, hipStream_t s) {
  int64_t g = (count + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(stub_init_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, s, ids, vals, count);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_leaf_knn(int metric, const float* x, int d, int ld, const float* norms, const int32_t* perm,
                     const Leaf* leaves, int nleaves, int max_leaf, int k, int32_t* ids, float* vals,
                     const int32_t* rsp, const int32_t* rcp, hipStream_t s) {
  if (nleaves <= 0) return;
  const bool big = d > 256;
  switch (metric) {
    case kEuclidean:
      big ? launch_k<kEuclidean, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kEuclidean, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kTaxicab:
      big ? launch_k<kTaxicab, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kTaxicab, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kTaxicabL1:
      big ? launch_k<kTaxicabL1, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kTaxicabL1, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    case kCosine:
      big ? launch_k<kCosine, true>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s)
          : launch_k<kCosine, false>(k, x, d, ld, norms, perm, leaves, nleaves, max_leaf, ids, vals, rsp, rcp, s);
      break;
    default:
      throw HipError("unknown distance");
  }
}

}  // namespace aknn
