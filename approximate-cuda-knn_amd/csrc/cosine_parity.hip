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
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"

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

__global__ __launch_bounds__(256) void cosine_parity_kernel(const float* __restrict__ x, int d, int ld,
                                                            const float* __restrict__ norms, int64_t n, int k,
                                                            int trees, const int32_t* __restrict__ rs,
                                                            const int32_t* __restrict__ rc,
                                                            const int32_t* __restrict__ perm, int64_t cap,
                                                            CosineParityScratch sc, int32_t* __restrict__ ids,
                                                            float* __restrict__ vals,
                                                            const int32_t* __restrict__ order) {
  extern __shared__ __attribute__((aligned(16))) float xi[];
  __shared__ int sh[4];
  uint32_t* bits = sc.bits + static_cast<int64_t>(blockIdx.x) * sc.words;
  uint32_t* summ = sc.summary + static_cast<int64_t>(blockIdx.x) * sc.summary_words;
  int32_t* cand = sc.cand + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  float* ckey = sc.key + static_cast<int64_t>(blockIdx.x) * sc.max_cand;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  const int64_t g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  // logical group g takes the stretch [g*per, (g+1)*per) of the leaf order
  for (int64_t pidx = g * per; pidx < n && pidx < (g + 1) * per; ++pidx) {
    const int64_t i = order != nullptr ? order[pidx] : pidx;
    for (int dd = tid; dd < d; dd += 256) xi[dd] = x[i * ld + dd];
    // union of the start-node ranges (.cu:160-191)
    for (int t = 0; t < trees; ++t) {
      const int32_t s0 = rs[static_cast<int64_t>(t) * n + i], c0 = rc[static_cast<int64_t>(t) * n + i];
      const int32_t* pt = perm + static_cast<int64_t>(t) * n + s0;
      for (int p = tid; p < c0; p += 256) {
        const int32_t j = pt[p];
        atomicOr(&bits[j >> 5], 1u << (j & 31));
        atomicOr(&summ[j >> 10], 1u << ((j >> 5) & 31));
      }
    }
    __threadfence_block();
    __syncthreads();
    if (tid == 0) atomicAnd(&bits[i >> 5], ~(1u << (i & 31)));  // .cu:190-191
    __threadfence_block();
    __syncthreads();
    // enumerate in ascending id order, clearing as we go
    int m = 0;
    for (int64_t sb = 0; sb < sc.summary_words; sb += 256) {
      // The bitmaps were written by L2 atomics: read them with agent-scope
      // atomics too, so no stale L1 line from an earlier point is used.
      const int64_t sw = sb + tid;
      const uint32_t s = sw < sc.summary_words ? atomicExch(&summ[sw], 0u) : 0u;
      int cnt = 0;
      for (uint32_t r = s; r; r &= r - 1)
        cnt += __popc(__hip_atomic_load(&bits[sw * 32 + __ffs(r) - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      int tot;
      int off = m + block_excl_scan256(cnt, sh, &tot);
      for (uint32_t r = s; r; r &= r - 1) {
        const int64_t w = sw * 32 + __ffs(r) - 1;
        for (uint32_t b = atomicExch(&bits[w], 0u); b; b &= b - 1) cand[off++] = static_cast<int32_t>(w * 32 + __ffs(b) - 1);
      }
      m += tot;
    }
    __threadfence_block();
    __syncthreads();
    const float ni = norms[i];
    for (int c = tid; c < m; c += 256) {
      const int32_t j = cand[c];
      const float t = row_chain<kCosine>(xi, x + static_cast<int64_t>(j) * ld, d);
      ckey[c] = finish_key<kCosine>(t, ni, norms[j]);
    }
    __threadfence_block();
    __syncthreads();
    if (wave == 0) {
      float key = lane < k ? kStubDist : __builtin_inff();
      int32_t id = lane < k ? kStubId : INT_MAX;
      for (int64_t c0 = 0; c0 < m; c0 += cap) {
        const int64_t c1 = c0 + cap < m ? c0 + cap : m;
        float thk = __shfl(key, k - 1);
        int32_t thi = __shfl(id, k - 1);
        for (int64_t b = c0; b < c1; b += 64) {
          const int64_t ci = b + lane;
          const bool valid = ci < c1;
          const float dk = valid ? ckey[ci] : __builtin_inff();
          const int32_t cj = valid ? cand[ci] : INT_MAX;
          uint64_t surv = __ballot(valid && key_less(dk, cj, thk, thi));
          while (surv) {
            const int src = __ffsll(static_cast<long long>(surv)) - 1;
            surv &= surv - 1;
            const float nk = __shfl(dk, src);
            const int32_t nid = __shfl(cj, src);
            if (key_less(nk, nid, thk, thi)) {
              wave_list_insert(key, id, k, nk, nid);
              thk = __shfl(key, k - 1);
              thi = __shfl(id, k - 1);
            }
          }
        }
        if (lane < k) key = static_cast<float>(1.0 - static_cast<double>(key));
        wave_sort(key, id);
      }
      if (lane < k) {
        ids[i * k + lane] = id;
        vals[i * k + lane] = key;
      }
    }
    __syncthreads();
  }
}

}  // namespace

void launch_cosine_parity_initial(const float* x, int d, int ld, const float* norms, int64_t n, int k, int trees,
                                  const int32_t* rs, const int32_t* rc, const int32_t* perm, int64_t cap,
                                  const CosineParityScratch& sc, int32_t* ids, float* vals, hipStream_t s) {
  if (k > 64) throw HipError("COSINE parity path supports k <= 64");
  const size_t shmem = static_cast<size_t>((d + 3) / 4 * 4) * sizeof(float);
  if (shmem > 64 * 1024) throw HipError("COSINE parity path: d too large");
  if (sc.groups % 8 != 0) throw HipError("COSINE parity path: groups must be a multiple of 8");
  hipLaunchKernelGGL(cosine_parity_kernel, dim3(sc.groups), dim3(256), shmem, s, x, d, ld, norms, n, k, trees, rs, rc,
                     perm, cap, sc, ids, vals, perm);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
