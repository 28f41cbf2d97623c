// One neighbour-propagation round (Jacobi: every read uses the previous graph).
//
// Reference: propagate{Euclidean,Taxicab,Cosine}Kernel
// (euclideandistancematrixgpu.cu:1033-1326, :1328-1621, :1623-1935), host loop
// .cu:1937-1977.  For point i the new candidates are the neighbours of its
// neighbours minus i minus its own neighbours (.cu:1078-1100, an N-byte bitmap
// per block there); if there are none the row is copied (.cu:1117); otherwise
// the k carried entries (keyed by their stored values) and the candidates
// (keyed by their distance) compete, and the first k by (key, id) are stored
// (COSINE parity stores 1 - key, .cu:1911).
//
// MI355X mapping: one wave per point.  The k*k two-hop ids are gathered with
// one load per lane-slot, deduplicated in an LDS hash set that was seeded with
// i and i's own row, and compacted in LDS.  Lanes then take one candidate each:
// the candidate's row is streamed from HBM with 16-byte loads while x_i is read
// from LDS by broadcast, giving the exact FP32 chain.  The running top-k lives
// one entry per lane; candidates below the current k-th entry (few after the
// first round) are inserted with ballot/shuffle steps.
//
// With the FP16 sketch (SK), a first pass reads each candidate's half-size
// sketch row instead and keeps only candidates whose proven lower bound of the
// exact key is not above the current k-th key (sketch.hip); the survivors are
// compacted in place and take the exact pass above.  A dropped candidate's
// exact key is above the k-th key, so it could not have entered: the stored
// graph is the same bit for bit, with about half the row bytes read.
#include <climits>

#include "device_util.hpp"
#include "kernels.hpp"
#include "topk.hpp"
#include "wave_ops.hpp"

namespace aknn {

namespace {

constexpr int kWavesPerBlock = 4;
// 16-byte row loads in flight per lane in the exact COSINE parity rounds:
// their LDS (a k^2 list beside the hash set) holds them to a few waves per
// SIMD, so registers are free and a row of up to 96 floats is read in one
// round trip
constexpr int kFlipRowLoads = 24;  // waves per block when their LDS fits (fewer otherwise)
constexpr size_t kLdsBytes = 160 * 1024;



// Open-addressing set of ids in LDS (-1 = empty), linear probing.  The table
// has at least twice as many slots as ids ever inserted, so probing ends.
__device__ __forceinline__ bool hash_insert(int32_t* hs, uint32_t size, int32_t key) {
  uint32_t h = static_cast<uint32_t>(
      (static_cast<uint64_t>(static_cast<uint32_t>(key) * 2654435761u) * size) >> 32);
  for (;;) {
    const int32_t old = atomicCAS(&hs[h], -1, key);
    if (old == -1) return true;
    if (old == key) return false;
    h = h + 1 == size ? 0 : h + 1;
  }
}

// SKM: 0 exact pass only, 1 FP16 sketch screen, 2 int8 sketch screen,
// 3 principal-subspace screen (pca.hip) on FP32 projections, 4 the same on
// FP16 projections (their rounding is inside each point's radius).  PR: the
// subspace rank of SKM 3 and 4 fixed at compile time (16), so the
// projections held in registers take only its width; 0 reads it from pca_r
// (16 or 32) with registers for 32.
template <int METRIC, bool FLIP, int SKM, int PR = 0>
__global__ __launch_bounds__(256) void propagate_kernel(const float* __restrict__ x, int d, int ld,
                                                        const _Float16* __restrict__ xh, int ldh,
                                                        const float* __restrict__ eps, SketchBound sb,
                                                        const int8_t* __restrict__ q8, int ld8,
                                                        const float* __restrict__ s8, const int32_t* __restrict__ a8,
                                                        const float* __restrict__ e8,
                                                        const float* __restrict__ pca,
                                                        const _Float16* __restrict__ pca16, int pca_r,
                                                        const float* __restrict__ pca_e, double pca_sigma,
                                                        const float* __restrict__ norms,
                                                        const int32_t* __restrict__ order, int64_t n, int k,
                                                        int hs_size, int cand_cap, int xi_stride,
                                                        const int32_t* __restrict__ ids_in,
                                                        const float* __restrict__ vals_in, int32_t* ids_out,
                                                        float* __restrict__ vals_out,
                                                        unsigned long long* __restrict__ ncand,
                                                        const int32_t* ids_prev,
                                                        const uint8_t* __restrict__ changed_prev,
                                                        uint8_t* __restrict__ changed_out) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // per wave: x_i (FP32, plain form) or, with the sketch, x_i's sketch as
  // floats, x_i (FP32) and two staged candidate rows; then the hash set, the
  // candidate list and its counter.  xi_stride is a multiple of 4.
  const size_t per_wave_words = static_cast<size_t>(xi_stride) + hs_size + cand_cap + 4;
  float* xrow = reinterpret_cast<float*>(smem) + per_wave_words * wave;  // 16-byte aligned
  constexpr bool SK = SKM > 0;
  float* xf = xrow + (SKM == 1 ? ldh : (SKM == 2 ? ld8 / 4 : (SKM >= 3 ? pca_r : 0)));
  float* stg = xf + ld;
  int32_t* hs = reinterpret_cast<int32_t*>(xrow + xi_stride);
  int32_t* cand = hs + hs_size;
  int32_t* cnt = cand + cand_cap;
  // COSINE parity, once the candidates' keys are packed: a histogram and
  // two lists of k packed pairs over the candidate list (its capacity is at
  // least 256 + 4k words then)
  uint32_t* fhist = reinterpret_cast<uint32_t*>(cand);
  uint64_t* ftop = reinterpret_cast<uint64_t*>(cand + 256);
  uint64_t* fcar = ftop + k;

  // Points are taken in `order` (the first tree's leaf order): waves running at
  // the same time then own nearby points whose candidate rows overlap, and
  // each XCD walks a contiguous part of that order, so its L2 sees the reuse.
  const int64_t pidx = xcd_block(blockIdx.x, gridDim.x) * static_cast<int64_t>(blockDim.x >> 6) + wave;
  if (pidx >= n) return;
  const int64_t i = order != nullptr ? order[pidx] : pidx;

  float key = __builtin_inff();
  int32_t id = INT_MAX;
  if (lane < k) {
    key = vals_in[i * k + lane];
    id = ids_in[i * k + lane];
  }
  // Neighbours whose whole row was already a candidate set of i last round
  // need no expansion: j was in i's previous row and j's row did not change
  // in the previous round.  Every such candidate was then in last round's
  // pool of i, and a candidate that lost there has k better entries in i's
  // current row, so it cannot enter now (ids_prev is the buffer this round
  // writes; row i of it is read here before this wave overwrites it).
  bool expand = true;
  if (ids_prev != nullptr) {
    const int32_t pid = lane < k ? ids_prev[i * k + lane] : INT_MAX;
    bool was_prev = false;
    for (int l = 0; l < k; ++l) was_prev |= __shfl(pid, l) == id;
    if (lane < k && id >= 0 && was_prev && changed_prev[id] == 0) expand = false;
  }
  for (int s = lane; s < hs_size; s += 64) hs[s] = -1;
  const float* xi = SK ? xf : xrow;
  const float* xs = xrow;
  if constexpr (SKM == 1) {
    for (int dd = lane; dd < ldh; dd += 64) xrow[dd] = static_cast<float>(xh[i * ldh + dd]);
    for (int dd = lane; dd < ld; dd += 64) xf[dd] = x[i * ld + dd];
  } else if constexpr (SKM == 2) {
    for (int dd = lane; dd < ld8 / 4; dd += 64)
      reinterpret_cast<int32_t*>(xrow)[dd] = reinterpret_cast<const int32_t*>(q8 + i * ld8)[dd];
    for (int dd = lane; dd < ld; dd += 64) xf[dd] = x[i * ld + dd];
  } else if constexpr (SKM >= 3) {
    for (int dd = lane; dd < pca_r; dd += 64)
      xrow[dd] = SKM == 4 ? static_cast<float>(pca16[i * pca_r + dd]) : pca[i * pca_r + dd];
    for (int dd = lane; dd < ld; dd += 64) xf[dd] = x[i * ld + dd];
  } else {
    for (int dd = lane; dd < d; dd += 64) xrow[dd] = x[i * ld + dd];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  const int kk = k * k;
  int m = 0;     // candidates for the exact pass
  int nnew = 0;  // new candidates (.cu:1117: none -> the row is copied, no flip)
  float thk = __shfl(key, k - 1);
  int32_t thi = __shfl(id, k - 1);
  const float ni = METRIC == kCosine ? norms[i] : 1.0f;
  {
    // two-hop ids (.cu:1078-1085), deduplicated in the LDS hash set that
    // holds i and i's row
    if (lane == 0) *cnt = 0;
    const uint32_t hsz = static_cast<uint32_t>(hs_size);
    if (lane == 0) hash_insert(hs, hsz, static_cast<int32_t>(i));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < k && id >= 0) hash_insert(hs, hsz, id);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // eight slots per lane are loaded before any of them is inserted, so
    // their memory round trips overlap instead of queueing behind the LDS
    // set operations (the set and the final candidate set are the same).
    // Slot f is neighbour f / k's entry f % k: the quotient from the float
    // reciprocal, (f + 1/2) * fl(1/k) truncated, which is exact for every
    // k <= 64 and f < 2^13.  A neighbour that is not expanded reads as -1.
    // New ids take list slots by a ballot prefix in slot order, or, in the
    // COSINE parity rounds, from a counter in LDS: the lists and graphs are
    // the same either way, and each form keeps the one measured faster for
    // it (the ballot on the subspace-screened EUCLIDEAN rounds, 4%; the
    // counter on the int8-screened parity rounds, 60%).
    constexpr int kHop = 8;
    const int kk_pad = kk + (64 - kk % 64) % 64;
    const float inv_k = 1.0f / static_cast<float>(k);
    const int32_t nbx = expand ? id : -1;
    int w = 0;
    for (int f0 = 0; f0 < kk_pad; f0 += 64 * kHop) {
      int32_t cv[kHop];
#pragma unroll
      for (int u = 0; u < kHop; ++u) {
        const int f = f0 + 64 * u + lane;
        const int hi = static_cast<int>((static_cast<float>(f) + 0.5f) * inv_k);
        const int lo = f - hi * k;
        const int32_t nb = __shfl(nbx, hi < 64 ? hi : 0);
        cv[u] = (f < kk && nb >= 0) ? ids_in[static_cast<int64_t>(nb) * k + lo] : -1;
      }
#pragma unroll
      for (int u = 0; u < kHop; ++u) {
        const int32_t c = cv[u];
        if constexpr (FLIP) {
          if (c >= 0 && hash_insert(hs, hsz, c)) {
            const int pos = atomicAdd(cnt, 1);
            cand[pos] = c;
          }
        } else {
          const bool ins = c >= 0 && hash_insert(hs, hsz, c);
          const uint64_t bal = __ballot(ins);
          if (ins) cand[w + __popcll(bal & ((1ull << lane) - 1))] = c;
          w += __popcll(bal);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    m = FLIP ? *cnt : w;
    nnew = m;
    if (ncand != nullptr && lane == 0)
      atomicAdd(ncand + (blockIdx.x % kCountSlots) * kCountStride, static_cast<unsigned long long>(m));
  }
  if constexpr (SK && !FLIP) {
    if (m > 0) {
      // screen: survivors are written back to cand[] in order (slot
      // w + rank <= base + lane was already read by this batch)
      const float ei = SKM == 2 ? e8[i] : (SKM >= 3 ? pca_e[i] : eps[i]);
      // every candidate's radius is bounded by the largest (kept at index n):
      // one load per wave instead of a scattered 4-byte load per candidate.
      // An infinite maximum (a row the sketch cannot hold) would switch the
      // screen off for all; then each candidate's own radius is read.
      const float ec_max = SKM >= 3 ? pca_e[n] : (SKM == 1 ? eps[n] : 0.0f);
      const bool per_cand = !(ec_max < __builtin_inff());
      const double si = SKM == 2 ? static_cast<double>(s8[i]) : 0.0;
      const double ti = SKM == 2 ? si * si * static_cast<double>(a8[i]) : 0.0;
      int w = 0;
      if constexpr (SKM >= 3) {
        // two batches of candidates per memory round trip: their ids and
        // projections are loaded before either is tested (compaction writes
        // slots at or below ones already read)
        const double k1 = 1.0 - sb.kap;
        // The bound below is a nondecreasing function of p, evaluated in FP64
        // by monotone steps: it exceeds thk exactly when p > P, where
        //   P = (S^2 + tiny) / k1,  S = ei + ec + sigma R,
        //   R = sqrt(((thk / k1)^2 + tiny) / k1)
        // in exact arithmetic, and its rounding moves that boundary by a few
        // units of 2^-52 relative.  So p above p_hi (P raised by 2^-13, then
        // up to a float) is certainly dropped and p below p_lo certainly
        // kept, with one FP32 comparison; the band between them, and every
        // point whose quantities leave that derivation (a per-candidate
        // radius, non-finite or negative values), take the FP64 bound.
        float p_hi = __builtin_inff(), p_lo = -__builtin_inff();
        if (!per_cand && thk >= 0.0f && thk < __builtin_inff() && ei >= 0.0f && ei < __builtin_inff() &&
            ec_max >= 0.0f) {
          const double t = static_cast<double>(thk) / k1;
          const double rr = sqrt((t * t + sb.tiny) / k1);
          const double ss = static_cast<double>(ei) + static_cast<double>(ec_max) + pca_sigma * rr;
          const double pp = (ss * ss + sb.tiny) / k1;
          if (pp < 1e300) {
            p_hi = float_up(pp * (1.0 + 1.0 / 8192.0));
            p_lo = float_down(pp * (1.0 - 1.0 / 8192.0));
          }
        }
        for (int base = 0; base < m; base += 128) {
          constexpr int RG = PR > 0 ? PR : 32;  // register width
          const int rk = PR > 0 ? PR : pca_r;
          int32_t cc[2];
          float4 pv[2][RG / 4];
          uint4 ph[2][RG / 8];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int ci = base + 64 * u + lane;
            cc[u] = ci < m ? cand[ci] : -1;
            const int64_t row = static_cast<int64_t>(cc[u] >= 0 ? cc[u] : 0) * rk;
            if constexpr (SKM == 4) {
              const uint4* pc = reinterpret_cast<const uint4*>(pca16 + row);
#pragma unroll
              for (int g = 0; g < RG / 8; ++g) ph[u][g] = g < rk / 8 ? pc[g] : make_uint4(0u, 0u, 0u, 0u);
            } else {
              const float4* pc = reinterpret_cast<const float4*>(pca + row);
#pragma unroll
              for (int g = 0; g < RG / 4; ++g) pv[u][g] = g < rk / 4 ? pc[g] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
          }
          if constexpr (SKM == 4) {
            // the FP16 values as FP32 (exact), four per group as in the FP32 form
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
              for (int g = 0; g < RG / 4; ++g) {
                const uint4 h = ph[u][g >> 1];
                const uint32_t w0 = (g & 1) ? h.z : h.x, w1 = (g & 1) ? h.w : h.y;
                pv[u][g] = make_float4(half_lo(w0), half_hi(w0), half_lo(w1), half_hi(w1));
              }
          }
          const float4* pq = reinterpret_cast<const float4*>(xs);
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            float p = 0.0f;
#pragma unroll
            for (int g = 0; g < RG / 4; ++g) {
              if (g < rk / 4) {
                const float4 v = pq[g], q = pv[u][g];
                float t = v.x - q.x;
                p = __builtin_fmaf(t, t, p);
                t = v.y - q.y;
                p = __builtin_fmaf(t, t, p);
                t = v.z - q.z;
                p = __builtin_fmaf(t, t, p);
                t = v.w - q.w;
                p = __builtin_fmaf(t, t, p);
              }
            }
            bool drop = p > p_hi;
            if (!drop && !(p < p_lo)) {
              // ||x - y|| >= (||p(x) - p(y)|| - e(x) - e(y)) / sigma, with the
              // r-term chain's rounding taken off; then the exact chain's error
              const double sp = fmax(0.0, static_cast<double>(p) * k1 - sb.tiny);
              const float ec = per_cand ? pca_e[cc[u] >= 0 ? cc[u] : 0] : ec_max;
              const double r = fmax(0.0, sqrt(sp) - static_cast<double>(ei) - static_cast<double>(ec)) / pca_sigma;
              const double lb = sqrt(fmax(0.0, r * r * k1 - sb.tiny)) * k1;
              drop = lb > static_cast<double>(thk);
            }
            const bool pass = cc[u] >= 0 && !drop;
            const uint64_t bal = __ballot(pass);
            if (pass) cand[w + __popcll(bal & ((1ull << lane) - 1))] = cc[u];
            w += __popcll(bal);
          }
        }
      } else
      for (int base = 0; base < m; base += 64) {
        const int ci = base + lane;
        const int32_t c = ci < m ? cand[ci] : -1;
        bool pass = false;
        if (c >= 0) {
          double lb;
          if constexpr (SKM == 2) {
            // exact integer dot product of the int8 rows, eight 16-byte loads in flight
            const int4* qc = reinterpret_cast<const int4*>(q8 + static_cast<int64_t>(c) * ld8);
            const int4* qi = reinterpret_cast<const int4*>(xrow);
            int dot = 0;
            const int pieces = ld8 / 16;
            int g = 0;
            for (; g + 8 <= pieces; g += 8) {
              int4 v[8];
#pragma unroll
              for (int j = 0; j < 8; ++j) v[j] = qc[g + j];
#pragma unroll
              for (int j = 0; j < 8; ++j) {
                const int4 u = qi[g + j];
                dot = __builtin_amdgcn_sdot4(v[j].x, u.x, dot, false);
                dot = __builtin_amdgcn_sdot4(v[j].y, u.y, dot, false);
                dot = __builtin_amdgcn_sdot4(v[j].z, u.z, dot, false);
                dot = __builtin_amdgcn_sdot4(v[j].w, u.w, dot, false);
              }
            }
            for (; g < pieces; ++g) {
              const int4 v = qc[g], u = qi[g];
              dot = __builtin_amdgcn_sdot4(v.x, u.x, dot, false);
              dot = __builtin_amdgcn_sdot4(v.y, u.y, dot, false);
              dot = __builtin_amdgcn_sdot4(v.z, u.z, dot, false);
              dot = __builtin_amdgcn_sdot4(v.w, u.w, dot, false);
            }
            // ||s_i q_i - s_c q_c||^2 exactly from the integers, evaluated in
            // FP64 with its rounding taken off; then the radii (triangle
            // inequality) and the exact chain's error, as sketch_lower_bound
            const double sc = static_cast<double>(s8[c]);
            const double tc = sc * sc * static_cast<double>(a8[c]);
            const double sq = ti + tc - 2.0 * si * sc * static_cast<double>(dot);
            const double slo = sq - 1e-12 * (ti + tc) - sb.tiny;
            const double r = fmax(0.0, sqrt(fmax(0.0, slo)) - static_cast<double>(ei) - static_cast<double>(e8[c]));
            const double k1 = 1.0 - sb.kap;
            lb = sqrt(fmax(0.0, r * r * k1 - sb.tiny)) * k1;
          } else {
            const float p = sketch_chain<METRIC>(xs, xh + static_cast<int64_t>(c) * ldh, ldh);
            lb = sketch_lower_bound<METRIC>(p, ei, per_cand ? eps[c] : ec_max, ni, METRIC == kCosine ? norms[c] : 1.0f,
                                            sb);
          }
          pass = !(lb > static_cast<double>(thk));
        }
        const uint64_t bal = __ballot(pass);
        if (pass) cand[w + __popcll(bal & ((1ull << lane) - 1))] = c;
        w += __popcll(bal);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      m = w;
      if (ncand != nullptr && lane == 0)
        atomicAdd(ncand + (kCountSlots + blockIdx.x % kCountSlots) * kCountStride, static_cast<unsigned long long>(m));
    }
  }

  bool changed = false;
  if constexpr (FLIP) {
    // COSINE parity (carried entries keyed by their flipped stored values,
    // so nearly every candidate undercuts the k-th): the keys of all
    // candidates first (one candidate per lane streaming its row), kept as
    // packed pairs over the finished hash set; then the first k of (row +
    // candidates), which are the first k of (row + the candidates' own
    // first k): a radix select of the candidates and a merge by ranks.  A
    // NaN key (no strict order) takes the insertion path, in candidate
    // order as before.
    if (nnew > 0) {
      uint64_t* kv = reinterpret_cast<uint64_t*>(hs);  // [m] (hs has >= 2 m words)
      bool nan = lane < k && key != key;
      bool exact_all = true;
      if constexpr (SKM == 1 || SKM == 2) {
        // The screen of the candidates' own first k: each candidate's sketch
        // bounds lb <= key <= ub; T, the k-th smallest (ub, id), bounds the
        // k-th smallest exact key of the candidates from above, so a
        // candidate whose lb is above T is not among their first k and only
        // the others are keyed exactly.  Degenerate bounds, NaN carried
        // values or at most k candidates: every candidate is keyed.
        // SKM 2: the sketch is the int8 reconstruction s q (radius e8): its
        // dot s_i s_c (q_i . q_c) is an exact integer product, rounded once,
        // within the bounds' allowance for the sketch dot's rounding.
        float* lbv = reinterpret_cast<float*>(cand);  // cand[ci]'s slot once its id is in kv
        const float ei = SKM == 2 ? e8[i] : eps[i];
        const float ec_max = SKM == 2 ? __builtin_inff() : eps[n];
        const bool per_cand = !(ec_max < __builtin_inff());
        const double si = SKM == 2 ? static_cast<double>(s8[i]) : 0.0;
        bool degen = false;
        for (int base = 0; base < m; base += 64) {
          const int ci = base + lane;
          const int32_t c = ci < m ? cand[ci] : -1;
          if (c >= 0) {
            float p;
            if constexpr (SKM == 2) {
              const int4* qc = reinterpret_cast<const int4*>(q8 + static_cast<int64_t>(c) * ld8);
              const int4* qi = reinterpret_cast<const int4*>(xrow);
              int dot = 0;
              const int pieces = ld8 / 16;
              int g = 0;
              for (; g + 8 <= pieces; g += 8) {
                int4 v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = qc[g + j];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                  const int4 u = qi[g + j];
                  dot = __builtin_amdgcn_sdot4(v[j].x, u.x, dot, false);
                  dot = __builtin_amdgcn_sdot4(v[j].y, u.y, dot, false);
                  dot = __builtin_amdgcn_sdot4(v[j].z, u.z, dot, false);
                  dot = __builtin_amdgcn_sdot4(v[j].w, u.w, dot, false);
                }
              }
              for (; g < pieces; ++g) {
                const int4 v = qc[g], u = qi[g];
                dot = __builtin_amdgcn_sdot4(v.x, u.x, dot, false);
                dot = __builtin_amdgcn_sdot4(v.y, u.y, dot, false);
                dot = __builtin_amdgcn_sdot4(v.z, u.z, dot, false);
                dot = __builtin_amdgcn_sdot4(v.w, u.w, dot, false);
              }
              p = static_cast<float>(si * static_cast<double>(s8[c]) * static_cast<double>(dot));
            } else {
              p = sketch_chain<METRIC>(xs, xh + static_cast<int64_t>(c) * ldh, ldh);
            }
            const float ec = SKM == 2 ? e8[c] : (per_cand ? eps[c] : ec_max);
            const double lb = sketch_lower_bound<METRIC>(p, ei, ec, ni, norms[c], sb);
            const double ub = sketch_upper_bound_cosine(p, ei, ec, ni, norms[c], sb);
            degen |= !(lb > -__builtin_inf()) || !(ub < __builtin_inf());
            kv[ci] = pack_pair(float_up(ub), order_id(c));
            lbv[ci] = float_down(lb);
          }
        }
        wave_sync();
        if (__ballot(degen || nan) == 0 && m > k) {
          uint32_t* shist = reinterpret_cast<uint32_t*>(cand + ((m + 3) & ~3));
          uint64_t* stop = reinterpret_cast<uint64_t*>(shist + 256);
          const int got = wave_topk(kv, m, k, stop, shist);
          wave_sync();
          uint64_t tv = lane < got ? stop[lane] : 0ull;
#pragma unroll
          for (int o = 32; o >= 1; o >>= 1) {
            const uint32_t ohi = __shfl_xor(static_cast<uint32_t>(tv >> 32), o);
            const uint32_t olo = __shfl_xor(static_cast<uint32_t>(tv), o);
            const uint64_t ov = (static_cast<uint64_t>(ohi) << 32) | olo;
            tv = ov > tv ? ov : tv;
          }
          const float thr = pair_key(tv);
          // survivors keyed exactly, compacted in order into kv (slot w +
          // rank <= base + lane was read by this batch)
          int w = 0;
          for (int base = 0; base < m; base += 64) {
            const int ci = base + lane;
            const uint64_t v = ci < m ? kv[ci] : 0ull;
            const bool pass = ci < m && !(lbv[ci] > thr);
            uint64_t ev = 0ull;
            if (pass) {
              const int32_t c = pair_id(v);
              const float dk = finish_key<METRIC>(row_chain<METRIC>(xi, x + static_cast<int64_t>(c) * ld, d), ni,
                                                  norms[c]);
              nan |= dk != dk;
              ev = pack_pair(dk, order_id(c));
            }
            const uint64_t bal = __ballot(pass);
            wave_sync();
            if (pass) kv[w + __popcll(bal & ((1ull << lane) - 1))] = ev;
            w += __popcll(bal);
          }
          wave_sync();
          m = w;
          exact_all = false;
          if (ncand != nullptr && lane == 0)
            atomicAdd(ncand + (kCountSlots + blockIdx.x % kCountSlots) * kCountStride,
                      static_cast<unsigned long long>(m));
        } else {
          for (int base = 0; base < m; base += 64) {
            const int ci = base + lane;
            if (ci < m) {
              const int32_t c = pair_id(kv[ci]);
              const float dk = finish_key<METRIC>(row_chain<METRIC>(xi, x + static_cast<int64_t>(c) * ld, d), ni,
                                                  norms[c]);
              nan |= dk != dk;
              kv[ci] = pack_pair(dk, order_id(c));
            }
          }
          exact_all = false;
          if (ncand != nullptr && lane == 0)
            atomicAdd(ncand + (kCountSlots + blockIdx.x % kCountSlots) * kCountStride,
                      static_cast<unsigned long long>(m));
        }
      }
      if (exact_all) {
        for (int base = 0; base < m; base += 64) {
          const int ci = base + lane;
          const int32_t c = ci < m ? cand[ci] : -1;
          if (c >= 0) {
            // the exact form with rows of at most kFlipRowLoads 16-byte
            // pieces: the whole row in one round trip; longer rows and the
            // screened forms keep the compiler's pipelined loads (batches
            // drain between them: slower there)
            const float* xc = x + static_cast<int64_t>(c) * ld;
            const float t = SKM == 0 && d <= 4 * kFlipRowLoads ? row_chain_batched<METRIC, kFlipRowLoads>(xi, xc, d)
                                                               : row_chain<METRIC>(xi, xc, d);
            const float dk = finish_key<METRIC>(t, ni, METRIC == kCosine ? norms[c] : 1.0f);
            nan |= dk != dk;
            kv[ci] = pack_pair(dk, order_id(c));
          }
        }
      }
      wave_sync();
      if (__ballot(nan) == 0) {
        const int got = wave_topk(kv, m, k, ftop, fhist);
        wave_sync();
        const uint64_t tv = wave_sort_u64(lane < got ? ftop[lane] : ~0ull);
        const uint64_t cv = lane < k ? pack_pair(key, order_id(id)) : ~0ull;
        if (lane < got) ftop[lane] = tv;
        if (lane < k) fcar[lane] = cv;
        wave_sync();
        // ranks in (row + candidates' first k); equal row entries (stubs)
        // by position, the row being sorted
        const int rc_ = lane + count_below(ftop, got, cv);
        const int rt = lane + count_below(fcar, k, tv);
        if (lane < k && rc_ < k) kv[rc_] = cv;
        if (lane < got && rt < k) kv[rt] = tv;
        wave_sync();
        if (lane < k) {
          const uint64_t v = kv[lane];
          key = pair_key(v);
          id = pair_id(v);
        }
      } else {
        for (int base = 0; base < m; base += 64) {
          const int ci = base + lane;
          const uint64_t v = ci < m ? kv[ci] : 0ull;
          const int32_t c = ci < m ? pair_id(v) : -1;
          const float dk = ci < m ? pair_key(v) : __builtin_inff();
          uint64_t surv = __ballot(c >= 0 && key_less(dk, c, thk, thi));
          while (surv) {
            const int src = __ffsll(static_cast<long long>(surv)) - 1;
            surv &= surv - 1;
            const float nk = __shfl(dk, src);
            const int32_t nid = __shfl(c, src);
            if (key_less(nk, nid, thk, thi)) {
              wave_list_insert(key, id, k, nk, nid);
              thk = __shfl(key, k - 1);
              thi = __shfl(id, k - 1);
            }
          }
        }
      }
      if (lane < k) key = static_cast<float>(1.0 - static_cast<double>(key));
      wave_sort(key, id);
      changed = true;
    }
  } else if (nnew > 0) {
    // exact pass: in the plain form, and after the principal-subspace screen
    // (a handful of survivors: one round trip for all of them), one candidate
    // per lane streams its row; after the sketch screens the survivors' rows
    // are first staged in LDS two at a time by the whole wave, then chained
    // from LDS
    constexpr bool STAGE = SK && SKM < 3;
    const int step = STAGE ? 2 : 64;
    for (int base = 0; base < m; base += step) {
      const int ci = base + lane;
      int32_t c = -1;
      float dk = __builtin_inff();
      if constexpr (STAGE) {
        const int nrow = m - base < 2 ? m - base : 2;
        const int q4 = ld / 4;
        __builtin_amdgcn_wave_barrier();
        for (int f = lane; f < nrow * q4; f += 64) {
          const int r = f / q4, q = f - r * q4;
          const int32_t cr = cand[base + r];
          *reinterpret_cast<float4*>(stg + r * ld + 4 * q) =
              *reinterpret_cast<const float4*>(x + static_cast<int64_t>(cr) * ld + 4 * q);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < nrow) {
          c = cand[ci];
          const float t = row_chain<METRIC>(xi, stg + lane * ld, d);
          dk = finish_key<METRIC>(t, ni, METRIC == kCosine ? norms[c] : 1.0f);
        }
      } else {
        c = ci < m ? cand[ci] : -1;
        if (c >= 0) {
          // after the subspace screen (its projection registers free by
          // now) and in the plain form for rows of up to 256 floats, eight
          // 16-byte loads issued together per batch (the compiler's own
          // schedule keeps four in flight); longer rows keep its schedule
          const float* xc = x + static_cast<int64_t>(c) * ld;
          const float t = SKM >= 3 || (SKM == 0 && d <= 256) ? row_chain_batched<METRIC, 8>(xi, xc, d)
                                                             : row_chain<METRIC>(xi, xc, d);
          dk = finish_key<METRIC>(t, ni, METRIC == kCosine ? norms[c] : 1.0f);
        }
      }
      uint64_t surv = __ballot(c >= 0 && key_less(dk, c, thk, thi));
      while (surv) {
        const int src = __ffsll(static_cast<long long>(surv)) - 1;
        surv &= surv - 1;
        const float nk = __shfl(dk, src);
        const int32_t nid = __shfl(c, src);
        if (key_less(nk, nid, thk, thi)) {
          changed = true;
          wave_list_insert(key, id, k, nk, nid);
          thk = __shfl(key, k - 1);
          thi = __shfl(id, k - 1);
        }
      }
    }
  }
  if (changed_out != nullptr && lane == 0) changed_out[i] = changed ? 1 : 0;
  if (lane < k) {
    ids_out[i * k + lane] = id;
    vals_out[i * k + lane] = key;
  }
}

// LDS layout of one wave: x_i (and with a screen its screening row and two
// staged rows), the hash set, the candidate list and its counter.
struct WaveLds {
  int hs_size = 0, cand_cap = 0, xi_stride = 0, waves = 0;
  size_t bytes() const { return (static_cast<size_t>(xi_stride) + hs_size + cand_cap + 4) * 4 * waves; }
};

// Largest block (4, 2 or 1 waves) whose LDS fits; waves == 0 if none does.
// At most 1 + k + k*k ids enter the set, which has at least twice as many
// slots: a power of two when four waves fit with it (the measured layout),
// else the smallest multiple of 64.
WaveLds plan_lds(int k, int d, int ld, int ldh, int ld8, int pca_r, int skm, bool flip) {
  WaveLds w;
  const int need = 2 * (k * k + k + 1);
  // (multiples of 4 keep every wave's LDS block 16-byte aligned)
  w.cand_cap = (k * k + 3) / 4 * 4;
  // COSINE parity: the selection's histogram and lists reuse the list; with
  // the sketch screen they follow the candidates' lower bounds
  if (flip && w.cand_cap < 256 + 4 * k) w.cand_cap = (256 + 4 * k + 3) / 4 * 4;
  if (flip && (skm == 1 || skm == 2)) w.cand_cap = (k * k + 3) / 4 * 4 + 256 + 4 * k;
  w.xi_stride = skm == 1 ? ldh + 3 * ld : skm == 2 ? ld8 / 4 + 3 * ld : skm >= 3 ? pca_r + ld : (d + 3) / 4 * 4;
  int pow2 = 128;
  while (pow2 < need) pow2 *= 2;
  w.hs_size = pow2;
  w.waves = kWavesPerBlock;
  if (w.bytes() <= kLdsBytes) return w;
  w.hs_size = (need + 63) / 64 * 64;
  for (int wv = kWavesPerBlock; wv >= 1; wv /= 2) {
    w.waves = wv;
    if (w.bytes() <= kLdsBytes) return w;
  }
  w.waves = 0;
  return w;
}

template <int METRIC, bool FLIP, int SKM, int PR = 0>
void launch_w(const PropagateArgs& a, hipStream_t s) {
  const WaveLds w = plan_lds(a.k, a.d, a.ld, a.ldh, a.ld8, a.pca_r, SKM, FLIP);
  if (w.waves == 0) throw HipError("propagation: k or d too large for LDS");
  const int hs_size = w.hs_size, cand_cap = w.cand_cap, xi_stride = w.xi_stride;
  const size_t shmem = w.bytes();
  const int64_t blocks = xcd_grid((a.n + w.waves - 1) / w.waves);
  const SketchBound sb = sketch_bound(a.d);
  hipLaunchKernelGGL((propagate_kernel<METRIC, FLIP, SKM, PR>), dim3(static_cast<unsigned>(blocks)), dim3(64 * w.waves), shmem,
                     s, a.x, a.d, a.ld, a.xh, a.ldh, a.eps, sb, a.q8, a.ld8, a.s8, a.a8, a.e8, a.pca, a.pca16, a.pca_r,
                     a.pca_e,
                     a.pca_sigma, a.norms, a.order, a.n, a.k, hs_size, cand_cap,
                     xi_stride, a.ids_in, a.vals_in, a.ids_out, a.vals_out, a.ncand,
                     FLIP ? nullptr : a.hist.ids_prev, a.hist.changed_prev, a.hist.changed_out);
  AKNN_HIP_TRY(hipGetLastError());
}

// Candidate rows are streamed (width class 0) for every width: the
// register-row instantiations measured slower here (fewer waves in flight at
// 196 VGPRs outweigh the deeper per-lane prefetch).
template <int METRIC, bool FLIP>
int launch_t(const PropagateArgs& a, hipStream_t s) {
  // COSINE parity: carried entries are keyed by their flipped stored values,
  // so the k-th entry is no distance threshold; the sketch screen there
  // bounds the candidates' own first k instead
  auto fits = [&](int skm) { return plan_lds(a.k, a.d, a.ld, a.ldh, a.ld8, a.pca_r, skm, FLIP).waves > 0; };
  int skm = 0;
  if constexpr (FLIP) {
    skm = a.q8 != nullptr && fits(2) ? 2 : (a.xh != nullptr && fits(1) ? 1 : 0);
  } else {
    // a screen whose wider per-wave layout does not fit even one wave's LDS
    // gives way to the next one, and finally to the exact pass (the same
    // graph)
    if (METRIC == kEuclidean && a.pca16 != nullptr && fits(4))
      skm = 4;
    else if (METRIC == kEuclidean && a.pca != nullptr && fits(3))
      skm = 3;
    else if (METRIC == kEuclidean && a.q8 != nullptr && fits(2))
      skm = 2;
    else if (a.xh != nullptr && fits(1))
      skm = 1;
  }
  switch (skm) {
    case 1:
      launch_w<METRIC, FLIP, 1>(a, s);
      break;
    case 2:
      if constexpr (FLIP || METRIC == kEuclidean) launch_w<METRIC, FLIP, 2>(a, s);
      break;
    case 3:
      if constexpr (!FLIP && METRIC == kEuclidean) {
        if (a.pca_r == 16)
          launch_w<METRIC, FLIP, 3, 16>(a, s);
        else
          launch_w<METRIC, FLIP, 3, 0>(a, s);
      }
      break;
    case 4:
      if constexpr (!FLIP && METRIC == kEuclidean) {
        if (a.pca_r == 16)
          launch_w<METRIC, FLIP, 4, 16>(a, s);
        else
          launch_w<METRIC, FLIP, 4, 0>(a, s);
      }
      break;
    default:
      launch_w<METRIC, FLIP, 0>(a, s);
  }
  return skm;
}

}  // namespace

bool propagation_fits(int k, int d) {
  return plan_lds(k, d, (d + 3) / 4 * 4, 0, 0, 0, 0, true).waves > 0;
}

int launch_propagate(const PropagateArgs& a, hipStream_t s) {
  if (a.k > 64) throw HipError("propagation supports k <= 64");
  if (a.ld % 4 != 0 || a.ld < a.d) throw HipError("propagation: row stride must be a multiple of 4 covering d");
  if (a.q8 != nullptr && (a.s8 == nullptr || a.a8 == nullptr || a.e8 == nullptr || a.ld8 % 16 != 0 || a.ld8 < a.d))
    throw HipError("propagation: malformed int8 sketch");
  if (a.xh != nullptr && (a.eps == nullptr || a.ldh % 8 != 0 || a.ldh < a.d))
    throw HipError("propagation: malformed sketch");
  if ((a.pca != nullptr || a.pca16 != nullptr) && (a.pca_e == nullptr || (a.pca_r != 16 && a.pca_r != 32) || !(a.pca_sigma > 0.0)))
    throw HipError("propagation: malformed principal-subspace screen");
  switch (a.metric) {
    case kEuclidean:
      return launch_t<kEuclidean, false>(a, s);
    case kTaxicab:
      return launch_t<kTaxicab, false>(a, s);
    case kTaxicabL1:
      return launch_t<kTaxicabL1, false>(a, s);
    case kCosine:
      return a.flip ? launch_t<kCosine, true>(a, s) : launch_t<kCosine, false>(a, s);
    default:
      throw HipError("unknown distance");
  }
}

}  // namespace aknn
