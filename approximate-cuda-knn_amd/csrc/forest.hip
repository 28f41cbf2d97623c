// Random-projection forest, level by level.
//
// Semantics follow buildUpTheTrees / makePartitionOfLeaf0..9
// (euclideandistancematrixgpu.cu:2130-2686): a bucket with more than `leaf`
// points is split by the sign of x . (x_p2 - x_p1) (an FMA chain in dimension
// order, threshold 0), the pivot p1 is forced LEFT and p2 RIGHT, every bucket
// at a level shares one (rand(), rand()) pair for its children's pivots, and
// the loop stops at the first level with nothing to split.
//
// Layout (MI355X-first, not the reference's per-position bookkeeping): only
// buckets that still split are kept, as a compact list of records plus a
// compact array of their points (ids in bucket order).  Finished buckets are
// written once, straight into the leaf array and a leaf list.  A level is
//   side:    one lane per active point (a quad of lanes with the sketch
//            screen): projection chain -> one ballot bit, per-1024-point
//            block counts and per-wave prefixes;
//   scans:   the block counts' offsets (one workgroup), every bucket's RIGHT
//            count from popcounts (one lane per bucket), the offsets of the
//            buckets' children and of their points (one workgroup);
//   scatter: one lane per active point: stable rank from popcounts, write to
//            the next level's compact arrays or to the final leaf array; the
//            lane of a bucket's first point writes its child records
//            (pivots) and leaves.
// The partition is stable (the reference's atomic order is one legal order;
// this fixes the one where each side keeps the bucket's order).  Median-split
// trees (no shared draw stream) are built several at a time through the same
// passes: one active list and one bucket list for the batch.
#include "device_util.hpp"
#include "kernels.hpp"

namespace aknn {

namespace {

constexpr int kSideBlock = 1024;             // points per scan block
constexpr int kWordsPerBlock = kSideBlock / 64;

// Projection chain x . (x_p2 - x_p1) in dimension order.  NV > 0: the row has
// exactly 4*NV floats and the point's row is loaded into registers in one
// batch (one memory latency; the pivot rows are shared by the bucket and come
// from cache).  NV = 0: any width, streamed.  Kernels are instantiated per
// width, so each has a single load path.
template <int NV>
__device__ __forceinline__ bool side_is_right(const float* __restrict__ x, int d, int ld, int32_t id, int32_t p1,
                                              int32_t p2) {
  const float* xc = x + static_cast<int64_t>(id) * ld;
  const float* xa = x + static_cast<int64_t>(p1) * ld;
  const float* xb = x + static_cast<int64_t>(p2) * ld;
  float s = 0.0f;
  if constexpr (NV > 0) {
    float4 c[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) c[j] = *reinterpret_cast<const float4*>(xc + 4 * j);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float4 a = *reinterpret_cast<const float4*>(xa + 4 * j);
      const float4 b = *reinterpret_cast<const float4*>(xb + 4 * j);
      s = __builtin_fmaf(c[j].x, b.x - a.x, s);
      s = __builtin_fmaf(c[j].y, b.y - a.y, s);
      s = __builtin_fmaf(c[j].z, b.z - a.z, s);
      s = __builtin_fmaf(c[j].w, b.w - a.w, s);
    }
  } else {
    int dd = 0;
    for (; dd + 4 <= d; dd += 4) {
      const float4 c = *reinterpret_cast<const float4*>(xc + dd);
      const float4 a = *reinterpret_cast<const float4*>(xa + dd);
      const float4 b = *reinterpret_cast<const float4*>(xb + dd);
      s = __builtin_fmaf(c.x, b.x - a.x, s);
      s = __builtin_fmaf(c.y, b.y - a.y, s);
      s = __builtin_fmaf(c.z, b.z - a.z, s);
      s = __builtin_fmaf(c.w, b.w - a.w, s);
    }
    for (; dd < d; ++dd) s = __builtin_fmaf(xc[dd], xb[dd] - xa[dd], s);
  }
  return !(s < 0.0f);  // side = s < 0 ? LEFT : RIGHT (.cu:2157-2161)
}

__global__ __launch_bounds__(256) void forest_init_kernel(ForestBuffers fb, int32_t* __restrict__ perm,
                                                          Leaf* __restrict__ leaves, int64_t n, int32_t q1, int32_t q2,
                                                          int leaf_eff) {
  const bool root_splits = n > leaf_eff;
  for (int64_t m = blockIdx.x * 256ll + threadIdx.x; m < n; m += gridDim.x * 256ll) {
    fb.act_ids[0][m] = static_cast<int32_t>(m);
    fb.act_bkt[0][m] = 0;
    if (!root_splits) perm[m] = static_cast<int32_t>(m);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ForestCtl c{};
    c.nact[0] = root_splits ? static_cast<int32_t>(n) : 0;
    c.nbkt[0] = root_splits ? 1 : 0;
    if (root_splits) {
      Bucket b{};
      b.astart = 0;
      b.size = static_cast<int32_t>(n);
      b.fstart = 0;
      b.q1 = q1;
      b.q2 = q2;
      b.rstart = 0;
      b.rcount = static_cast<int32_t>(n);
      fb.bkt[0][0] = b;
      if (fb.piv[0] != nullptr) {  // the root's active list is the identity
        fb.piv[0][0] = q1;
        fb.piv[0][1] = q2;
      }
    } else {
      leaves[0] = Leaf{0, static_cast<int32_t>(n), 0, static_cast<int32_t>(n)};
      c.nleaves = 1;
    }
    *fb.ctl = c;
  }
}

template <int NV>
__global__ __launch_bounds__(256) void forest_side_kernel(const float* __restrict__ x, int d, int ld,
                                                          const int32_t* __restrict__ act_ids,
                                                          const int32_t* __restrict__ act_bkt,
                                                          const Bucket* __restrict__ bkt, const ForestCtl* ctl,
                                                          int parity, uint64_t* __restrict__ rmask,
                                                          int32_t* __restrict__ wpre, int32_t* __restrict__ blk_r) {
  __shared__ int wcnt[kWordsPerBlock];
  const int nact = ctl->nact[parity];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  for (int64_t blk = blockIdx.x; blk * kSideBlock < nact; blk += gridDim.x) {
    for (int sub = 0; sub < kSideBlock / 256; ++sub) {
      const int64_t m = blk * kSideBlock + sub * 256 + threadIdx.x;
      bool right = false;
      if (m < nact) {
        const Bucket& b = bkt[act_bkt[m]];
        const int32_t rel = static_cast<int32_t>(m) - b.astart;
        if (rel == b.q2) {
          right = true;  // Leaf2 writes RIGHT for p2 last (.cu:2178-2179)
        } else if (rel != b.q1) {
          right = side_is_right<NV>(x, d, ld, act_ids[m], act_ids[b.astart + b.q1], act_ids[b.astart + b.q2]);
        }
      }
      const uint64_t bal = __ballot(right);
      const int w = sub * 4 + wave;
      if (lane == 0) {
        rmask[blk * kWordsPerBlock + w] = bal;
        wcnt[w] = __popcll(bal);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int w = 0; w < kWordsPerBlock; ++w) {
        wpre[blk * kWordsPerBlock + w] = acc;
        acc += wcnt[w];
      }
      blk_r[blk] = acc;
    }
    __syncthreads();
  }
}

// side_is_right with the staged w = x_p2 - x_p1 (the same FP32 differences
// the chain forms): the same chain in dimension order, the point's row read
// eight 16-byte pieces at a time before their FMAs.
__device__ __forceinline__ bool side_is_right_w(const float* __restrict__ xc, const float* w, int d) {
  float s = 0.0f;
  int dd = 0;
  for (; dd + 32 <= d; dd += 32) {
    float4 c[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = *reinterpret_cast<const float4*>(xc + dd + 4 * u);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float4 wv = *reinterpret_cast<const float4*>(w + dd + 4 * u);
      s = __builtin_fmaf(c[u].x, wv.x, s);
      s = __builtin_fmaf(c[u].y, wv.y, s);
      s = __builtin_fmaf(c[u].z, wv.z, s);
      s = __builtin_fmaf(c[u].w, wv.w, s);
    }
  }
  for (; dd < d; ++dd) s = __builtin_fmaf(xc[dd], w[dd], s);
  return !(s < 0.0f);
}

// Side of a point decided from its FP16 sketch, or 0 when the sketch cannot
// decide.  The projection of the sketch, s' = sum xh_d * w_d with the exact
// w_d = x_p2,d - x_p1,d of the chain, differs from the exact FP32 chain s by
// at most ||w|| (eps + kap (||xh|| + eps)) (the sketch radius and the two
// chains' rounding), so |s'| above that bound fixes the sign of s: +1 RIGHT
// (s > 0), -1 LEFT (s < 0).  ||w|| is bounded from the same loop.
//
// The four lanes q of a quad share one point: lane q takes the 8-dimension
// pieces q, q+4, q+8, ... of the rows, so consecutive lanes read consecutive
// bytes of the sketch row and of both pivot rows, and the partial sums meet by
// two xor shuffles (every lane of the quad then holds the same sums and
// decides alike).  Each partial is a chain of at most ld/4 + 1 steps and the
// meeting adds two more roundings, within the d + 8 that kap covers, so the
// bound holds for this order too.
__device__ __forceinline__ int sketch_side_quad(const float* __restrict__ x, int ld, const _Float16* __restrict__ xh,
                                                int ldh, float eo, float no, int32_t id, int32_t p1, int32_t p2,
                                                const SketchBound& sb, int q) {
  const _Float16* hc = xh + static_cast<int64_t>(id) * ldh;
  const float* xa = x + static_cast<int64_t>(p1) * ld;
  const float* xb = x + static_cast<int64_t>(p2) * ld;
  float sp = 0.0f, ww = 0.0f;
  const int pieces = ld / 8;
#pragma unroll 2
  for (int pc = q; pc < pieces; pc += 4) {
    const int dd = 8 * pc;
    const uint4 h = *reinterpret_cast<const uint4*>(hc + dd);
    const float4 a0 = *reinterpret_cast<const float4*>(xa + dd);
    const float4 a1 = *reinterpret_cast<const float4*>(xa + dd + 4);
    const float4 b0 = *reinterpret_cast<const float4*>(xb + dd);
    const float4 b1 = *reinterpret_cast<const float4*>(xb + dd + 4);
    const float w[8] = {b0.x - a0.x, b0.y - a0.y, b0.z - a0.z, b0.w - a0.w,
                        b1.x - a1.x, b1.y - a1.y, b1.z - a1.z, b1.w - a1.w};
    const float hv[8] = {half_lo(h.x), half_hi(h.x), half_lo(h.y), half_hi(h.y),
                         half_lo(h.z), half_hi(h.z), half_lo(h.w), half_hi(h.w)};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      sp = __builtin_fmaf(hv[j], w[j], sp);
      ww = __builtin_fmaf(w[j], w[j], ww);
    }
  }
  if (ld % 8 != 0 && q == pieces % 4) {  // ld % 8 == 4: one group of four, taken by the next lane in turn
    const int dd = 8 * pieces;
    const uint2 h = *reinterpret_cast<const uint2*>(hc + dd);
    const float4 a0 = *reinterpret_cast<const float4*>(xa + dd);
    const float4 b0 = *reinterpret_cast<const float4*>(xb + dd);
    const float w[4] = {b0.x - a0.x, b0.y - a0.y, b0.z - a0.z, b0.w - a0.w};
    const float hv[4] = {half_lo(h.x), half_hi(h.x), half_lo(h.y), half_hi(h.y)};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      sp = __builtin_fmaf(hv[j], w[j], sp);
      ww = __builtin_fmaf(w[j], w[j], ww);
    }
  }
  sp += __shfl_xor(sp, 1);
  ww += __shfl_xor(ww, 1);
  sp += __shfl_xor(sp, 2);
  ww += __shfl_xor(ww, 2);
  const double wn = sqrt(static_cast<double>(ww)) * (1.0 + sb.kap);
  const double xn = sqrt(static_cast<double>(no)) * (1.0 + sb.kap);
  const double e = eo;
  const double bound = wn * (e + sb.kap * (xn + e)) + 2.0 * sb.tiny;
  if (static_cast<double>(sp) > bound) return 1;
  if (static_cast<double>(sp) < -bound) return -1;
  return 0;
}

// The same decision with the pivot difference w = x_p2 - x_p1 (the exact FP32
// differences) and its squared norm ww staged in LDS for the bucket: one
// point's quad then reads only its sketch row.
__device__ __forceinline__ int sketch_side_quad_w(int ld, const _Float16* __restrict__ xh, int ldh, float eo, float no,
                                                  int32_t id, const float* w, float ww, const SketchBound& sb, int q) {
  const _Float16* hc = xh + static_cast<int64_t>(id) * ldh;
  float sp = 0.0f;
  const int pieces = ld / 8;
#pragma unroll 4
  for (int pc = q; pc < pieces; pc += 4) {
    const int dd = 8 * pc;
    const uint4 h = *reinterpret_cast<const uint4*>(hc + dd);
    const float4 w0 = *reinterpret_cast<const float4*>(w + dd);
    const float4 w1 = *reinterpret_cast<const float4*>(w + dd + 4);
    const float wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
    const float hv[8] = {half_lo(h.x), half_hi(h.x), half_lo(h.y), half_hi(h.y),
                         half_lo(h.z), half_hi(h.z), half_lo(h.w), half_hi(h.w)};
#pragma unroll
    for (int j = 0; j < 8; ++j) sp = __builtin_fmaf(hv[j], wv[j], sp);
  }
  if (ld % 8 != 0 && q == pieces % 4) {  // ld % 8 == 4: one group of four, taken by the next lane in turn
    const int dd = 8 * pieces;
    const uint2 h = *reinterpret_cast<const uint2*>(hc + dd);
    const float4 w0 = *reinterpret_cast<const float4*>(w + dd);
    const float wv[4] = {w0.x, w0.y, w0.z, w0.w};
    const float hv[4] = {half_lo(h.x), half_hi(h.x), half_lo(h.y), half_hi(h.y)};
#pragma unroll
    for (int j = 0; j < 4; ++j) sp = __builtin_fmaf(hv[j], wv[j], sp);
  }
  sp += __shfl_xor(sp, 1);
  sp += __shfl_xor(sp, 2);
  const double wn = sqrt(static_cast<double>(ww)) * (1.0 + sb.kap);
  const double xn = sqrt(static_cast<double>(no)) * (1.0 + sb.kap);
  const double e = eo;
  const double bound = wn * (e + sb.kap * (xn + e)) + 2.0 * sb.tiny;
  if (static_cast<double>(sp) > bound) return 1;
  if (static_cast<double>(sp) < -bound) return -1;
  return 0;
}

// The same decision from the int8 sketch x = s q + r (||r|| <= e8): with the
// staged w, t = q . w as an FP32 chain and s' = s t, the exact chain s
// differs from s' by at most ||w|| (e8 + kap (||x|| + s ||q||)) + 2^-22 |s'|
// (the exact chain's and t's rounding, ||q|| <= 127 sqrt(d), the product's
// rounding), with ||x|| bounded by xn for every point.  One point's quad
// reads its 16-byte pieces of the int8 row.  w is staged over ld8 values,
// zero past ld.
__device__ __forceinline__ int int8_side_quad_w(int d, const int8_t* __restrict__ q8, int ld8, int32_t id, float2 se,
                                                double xn, const float* w, float ww, const SketchBound& sb, int q) {
  const int8_t* qr = q8 + static_cast<int64_t>(id) * ld8;
  float t = 0.0f;
  const int pieces = ld8 / 16;
  for (int pc = q; pc < pieces; pc += 4) {
    const int dd = 16 * pc;
    const int4 v = *reinterpret_cast<const int4*>(qr + dd);
    const int32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float4 wv = *reinterpret_cast<const float4*>(w + dd + 4 * u);
      t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>(vw[u] & 0xff)), wv.x, t);
      t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 8) & 0xff)), wv.y, t);
      t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 16) & 0xff)), wv.z, t);
      t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 24) & 0xff)), wv.w, t);
    }
  }
  t += __shfl_xor(t, 1);
  t += __shfl_xor(t, 2);
  const float sp = se.x * t;
  const double wn = sqrt(static_cast<double>(ww)) * (1.0 + sb.kap);
  const double qn = 127.0 * sqrt(static_cast<double>(d));
  const double bound = wn * (static_cast<double>(se.y) + sb.kap * (xn + static_cast<double>(se.x) * qn)) +
                       fabs(static_cast<double>(sp)) * 2.384185791015625e-07 + 2.0 * sb.tiny;
  if (static_cast<double>(sp) > bound) return 1;
  if (static_cast<double>(sp) < -bound) return -1;
  return 0;
}

// One 16-byte piece (16 int8 values) of q against the staged w: t += q . w.
__device__ __forceinline__ float int8_piece_dot(int4 v, const float* w, float t) {
  const int32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const float4 wv = *reinterpret_cast<const float4*>(w + 4 * u);
    t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>(vw[u] & 0xff)), wv.x, t);
    t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 8) & 0xff)), wv.y, t);
    t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 16) & 0xff)), wv.z, t);
    t = __builtin_fmaf(static_cast<float>(static_cast<int8_t>((vw[u] >> 24) & 0xff)), wv.w, t);
  }
  return t;
}

// int8_side_quad_w with the lane's pieces q and q + 4 already loaded (rows
// of at most 8 pieces); every lane of the quad takes part in the shuffles,
// active or not.
__device__ __forceinline__ int int8_side_pre(int d, int pieces, bool active, int4 qa, int4 qb, float2 se, double xn,
                                             const float* w, float ww, const SketchBound& sb, int q) {
  float t = 0.0f;
  if (active && q < pieces) t = int8_piece_dot(qa, w + 16 * q, t);
  if (active && q + 4 < pieces) t = int8_piece_dot(qb, w + 16 * (q + 4), t);
  t += __shfl_xor(t, 1);
  t += __shfl_xor(t, 2);
  if (!active) return 0;
  const float sp = se.x * t;
  const double wn = sqrt(static_cast<double>(ww)) * (1.0 + sb.kap);
  const double qn = 127.0 * sqrt(static_cast<double>(d));
  const double bound = wn * (static_cast<double>(se.y) + sb.kap * (xn + static_cast<double>(se.x) * qn)) +
                       fabs(static_cast<double>(sp)) * 2.384185791015625e-07 + 2.0 * sb.tiny;
  if (static_cast<double>(sp) > bound) return 1;
  if (static_cast<double>(sp) < -bound) return -1;
  return 0;
}

// Buckets of a 1024-point block whose pivot data is staged in LDS (an active
// bucket holds more than L points, so a block spans few), within a budget of
// staged values (rows wider than it stage nothing).
constexpr int kSideStage = 8;
constexpr int kSideStageFloats = 4096;  // 16 KB: up to 8 rows of 512 values, fewer of wider rows

// Side kernel with the sketch: one pass decides most points from their half
// rows, a quad of lanes per point so the rows are read coalesced; the
// decisions become the block's ballot words through LDS.  The first
// kSideStage buckets of a block have their pivot difference, its norm and
// their records staged in LDS, so a point reads only its own sketch row (the
// int8 one when given: half the bytes of the FP16 row); the points of further
// buckets read the pivot rows themselves.  The undecided
// ones of a 1024-point block are queued in LDS and take the exact chain
// together, so a straggler does not hold a whole wave.
// four waves per SIMD: a level's ~n/1024 blocks fit the chip at once
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void forest_side_sketch_kernel(
    const float* __restrict__ x, int d, int ld, const _Float16* __restrict__ xh, int ldh,
    const float* __restrict__ eps, const float* __restrict__ nh, int64_t n, SketchBound sb,
    const int32_t* __restrict__ act_ids,
    const int32_t* __restrict__ act_bkt, const Bucket* __restrict__ bkt, const ForestCtl* ctl, int parity,
    uint64_t* __restrict__ rmask, int32_t* __restrict__ wpre, int32_t* __restrict__ blk_r,
    unsigned long long* __restrict__ amb_count, const int8_t* __restrict__ q8, int ld8,
    const float2* __restrict__ se8, const int32_t* __restrict__ piv) {
  __shared__ uint64_t bits[kWordsPerBlock];
  __shared__ int wcnt[kWordsPerBlock];
  __shared__ int16_t amb[kSideBlock];
  __shared__ int8_t dec[kSideBlock];  // 0 LEFT, 1 RIGHT, 2 undecided
  __shared__ int namb;
  __shared__ __attribute__((aligned(16))) float wst[kSideStageFloats];  // row j at j * wl
  __shared__ float wwst[kSideStage];
  __shared__ int32_t bst[kSideStage][3];  // astart, q1, q2
  __shared__ int32_t pst[kSideStage][2];  // pivot ids
  __shared__ int32_t ids_s[kSideBlock], bk_s[kSideBlock];  // the block's points and their buckets
  const int nact = ctl->nact[parity];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int quad = threadIdx.x >> 2, q = threadIdx.x & 3;
  // The bound grows with the point's radius and sketch norm, so their
  // largest values (kept at index n) serve every point without a scattered
  // load per point and level; an infinite largest radius (a row the sketch
  // cannot hold) falls back to each point's own values.
  const float emax = eps[n], nmax = nh[n];
  const bool per_point = !(emax < __builtin_inff());
  // every point's ||x|| <= ||xh|| + eps (infinite when a row is out of range:
  // then the int8 bound decides nothing and the exact chain runs)
  const double xn = sqrt(static_cast<double>(nmax)) * (1.0 + sb.kap) + static_cast<double>(emax);
  const int wl = q8 != nullptr && ld8 > ld ? ld8 : ld;  // staged row width
  for (int64_t blk = blockIdx.x; blk * kSideBlock < nact; blk += gridDim.x) {
    if (threadIdx.x == 0) namb = 0;
    // the block's points and buckets, read coalesced once
    for (int u = 0; u < kSideBlock / 256; ++u) {
      const int ml = u * 256 + threadIdx.x;
      const int64_t m = blk * kSideBlock + ml;
      ids_s[ml] = m < nact ? act_ids[m] : -1;
      bk_s[ml] = m < nact ? act_bkt[m] : -1;
    }
    __syncthreads();
    // the active list runs bucket by bucket: the block's buckets are a
    // contiguous range of bucket indices
    const int cnt_blk = nact - blk * kSideBlock < kSideBlock ? static_cast<int>(nact - blk * kSideBlock) : kSideBlock;
    const int32_t b_first = bk_s[0], b_last = bk_s[cnt_blk - 1];
    const int fit = kSideStageFloats / wl < kSideStage ? kSideStageFloats / wl : kSideStage;
    const int nst = b_last - b_first + 1 < fit ? b_last - b_first + 1 : fit;
    if (threadIdx.x < nst) {
      const int32_t bi = b_first + threadIdx.x;
      const Bucket& b = bkt[bi];
      bst[threadIdx.x][0] = b.astart;
      bst[threadIdx.x][1] = b.q1;
      bst[threadIdx.x][2] = b.q2;
      pst[threadIdx.x][0] = piv != nullptr ? piv[2 * bi] : act_ids[b.astart + b.q1];
      pst[threadIdx.x][1] = piv != nullptr ? piv[2 * bi + 1] : act_ids[b.astart + b.q2];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nst * wl; e += 256) {
      const int j = e / wl, dd = e - j * wl;
      wst[j * wl + dd] = dd < ld ? x[static_cast<int64_t>(pst[j][1]) * ld + dd] - x[static_cast<int64_t>(pst[j][0]) * ld + dd]
                           : 0.0f;
    }
    __syncthreads();
    for (int j = wave; j < nst; j += 4) {
      // any order of the FP32 sum stays within the bound's rounding allowance
      float acc = 0.0f;
      for (int dd = lane; dd < ld; dd += 64) acc = __builtin_fmaf(wst[j * wl + dd], wst[j * wl + dd], acc);
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
      if (lane == 0) wwst[j] = acc;
    }
    __syncthreads();
    // one point per quad at a time (64 quads, 16 points each).  A point's
    // class: 0 LEFT without a test (past the list, or p1), 1 RIGHT (p2), 2
    // a staged bucket's test, 3 another bucket's test (pivot rows read).
    auto classify = [&](int ml, int32_t& id, int& j) {
      const int64_t m = blk * kSideBlock + ml;
      id = ids_s[ml];
      j = -1;
      if (m >= nact) return 0;
      const int32_t bi = bk_s[ml];
      const int jj = bi - b_first;
      if (jj >= 0 && jj < nst) {
        const int32_t rel = static_cast<int32_t>(m) - bst[jj][0];
        if (rel == bst[jj][2]) return 1;  // Leaf2 writes RIGHT for p2 last (.cu:2178-2179)
        if (rel == bst[jj][1]) return 0;
        j = jj;
        return 2;
      }
      const Bucket& b = bkt[bi];
      const int32_t rel = static_cast<int32_t>(m) - b.astart;
      if (rel == b.q2) return 1;
      if (rel == b.q1) return 0;
      return 3;
    };
    auto record = [&](int ml, int r) {
      if (q == 0) {
        dec[ml] = static_cast<int8_t>(r);
        if (r == 2) amb[atomicAdd(&namb, 1)] = static_cast<int16_t>(ml);
      }
    };
    auto slow = [&](int ml, int32_t id) {
      const Bucket& b = bkt[bk_s[ml]];
      return sketch_side_quad(x, ld, xh, ldh, per_point ? eps[id] : emax, per_point ? nh[id] : nmax, id,
                              act_ids[b.astart + b.q1], act_ids[b.astart + b.q2], sb, q);
    };
    if (q8 != nullptr && ld8 <= 128) {
      // int8 rows of at most 8 pieces: four points per quad in flight, their
      // pieces requested before any is used (addresses clamped to row 0 for
      // points without a staged test, whose loads are not used)
      const int pieces = ld8 / 16;
      for (int t0 = 0; t0 < kSideBlock / 64; t0 += 4) {
        int cls[4], jj[4];
        int32_t ids[4];
        int4 qa[4], qb[4];
        float2 sev[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ml = (t0 + u) * 64 + quad;
          cls[u] = classify(ml, ids[u], jj[u]);
          const int64_t row = static_cast<int64_t>(cls[u] == 2 ? ids[u] : 0) * ld8;
          qa[u] = *reinterpret_cast<const int4*>(q8 + row + 16 * (q < pieces ? q : 0));
          qb[u] = *reinterpret_cast<const int4*>(q8 + row + 16 * (q + 4 < pieces ? q + 4 : 0));
          sev[u] = se8[cls[u] == 2 ? ids[u] : 0];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int ml = (t0 + u) * 64 + quad;
          const bool act = cls[u] == 2;
          const int j = act ? jj[u] : 0;
          const int sd = int8_side_pre(d, pieces, act, qa[u], qb[u], sev[u], xn, wst + j * wl, wwst[j], sb, q);
          int r = cls[u] == 1 ? 1 : 0;
          if (act) r = sd > 0 ? 1 : (sd == 0 ? 2 : 0);
          if (cls[u] != 3) record(ml, r);
        }
        // points of unstaged buckets (rare): the pivot rows from global memory
#pragma unroll 1
        for (int u = 0; u < 4; ++u) {
          if (cls[u] != 3) continue;
          const int ml = (t0 + u) * 64 + quad;
          const int s3 = slow(ml, ids[u]);
          record(ml, s3 > 0 ? 1 : (s3 == 0 ? 2 : 0));
        }
      }
    } else {
      for (int t = 0; t < kSideBlock / 64; ++t) {
        const int ml = t * 64 + quad;
        int32_t id;
        int j;
        const int c = classify(ml, id, j);
        int r = c == 1 ? 1 : 0;
        if (c == 2) {
          const int sd = q8 != nullptr
                             ? int8_side_quad_w(d, q8, ld8, id, se8[id], xn, wst + j * wl, wwst[j], sb, q)
                             : sketch_side_quad_w(ld, xh, ldh, per_point ? eps[id] : emax, per_point ? nh[id] : nmax,
                                                  id, wst + j * wl, wwst[j], sb, q);
          r = sd > 0 ? 1 : (sd == 0 ? 2 : 0);
        } else if (c == 3) {
          const int sd = slow(ml, id);
          r = sd > 0 ? 1 : (sd == 0 ? 2 : 0);
        }
        record(ml, r);
      }
    }
    __syncthreads();
    for (int sub = 0; sub < kSideBlock / 256; ++sub) {
      const uint64_t bal = __ballot(dec[sub * 256 + threadIdx.x] == 1);
      if (lane == 0) bits[sub * 4 + wave] = bal;
    }
    __syncthreads();
    const int na = namb;
    for (int t = threadIdx.x; t < na; t += 256) {
      const int ml = amb[t];
      const int j = bk_s[ml] - b_first;
      bool right;
      if (j >= 0 && j < nst) {
        right = side_is_right_w(x + static_cast<int64_t>(ids_s[ml]) * ld, wst + j * wl, d);
      } else {
        const Bucket& b = bkt[bk_s[ml]];
        right = side_is_right<0>(x, d, ld, ids_s[ml], act_ids[b.astart + b.q1], act_ids[b.astart + b.q2]);
      }
      if (right) atomicOr(reinterpret_cast<unsigned long long*>(&bits[ml >> 6]), 1ull << (ml & 63));
    }
    __syncthreads();
    if (threadIdx.x < kWordsPerBlock) {
      rmask[blk * kWordsPerBlock + threadIdx.x] = bits[threadIdx.x];
      wcnt[threadIdx.x] = __popcll(bits[threadIdx.x]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int w = 0; w < kWordsPerBlock; ++w) {
        wpre[blk * kWordsPerBlock + w] = acc;
        acc += wcnt[w];
      }
      blk_r[blk] = acc;
      if (amb_count != nullptr && na > 0)
        atomicAdd(amb_count + (blk % kCountSlots) * kCountStride, static_cast<unsigned long long>(na));
    }
    __syncthreads();
  }
}

// Inclusive scan over a 1024-thread block; *total receives the block sum.
__device__ __forceinline__ int block_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  if (wave == 0) {
    int w = lane < nw ? sh[lane] : 0;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(w, o);
      if (lane >= o) w += t;
    }
    if (lane < nw) sh[lane] = w;
  }
  __syncthreads();
  const int res = v + (wave > 0 ? sh[wave - 1] : 0);
  *total = sh[nw - 1];
  __syncthreads();
  return res;
}

__device__ __forceinline__ int32_t rights_before(int64_t pos, int nact, int32_t total, const int32_t* blk_roff,
                                                 const int32_t* wpre, const uint64_t* rmask) {
  if (pos >= nact) return total;
  const uint64_t w = rmask[pos >> 6];
  const uint64_t below = (1ull << (pos & 63)) - 1ull;
  return blk_roff[pos >> 10] + wpre[pos >> 6] + __popcll(w & below);
}

// Where finished buckets go.  One tree: `leaves`, counted by ctl->nleaves.
// A batch of trees (median split): tree t's leaves at leaves + t * stride,
// counted by tree_nleaves[t], and tree_last[t] = the last level at which
// tree t still has a bucket to split.
struct LeafSink {
  Leaf* leaves;
  int64_t stride;
  int32_t* tree_nleaves;
  int32_t* tree_last;
};

// Exclusive scan of the side pass's per-block RIGHT counts (one workgroup:
// each thread sums a contiguous run of blocks, loads eight at a time, then
// one block-wide scan of the run sums); the level's RIGHT total goes to the
// control block.
__global__ __launch_bounds__(1024) void forest_block_scan_kernel(const int32_t* __restrict__ blk_r,
                                                                 int32_t* __restrict__ blk_roff, ForestCtl* ctl,
                                                                 int parity) {
  __shared__ int sh[16];
  const int nact = ctl->nact[parity];
  const int nblk = (nact + kSideBlock - 1) / kSideBlock;
  const int per = (nblk + 1023) / 1024;
  const int b0 = threadIdx.x * per < nblk ? threadIdx.x * per : nblk;
  const int b1 = b0 + per < nblk ? b0 + per : nblk;
  int v = 0;
  for (int i = b0; i < b1; i += 8) {
    int t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = i + u < b1 ? blk_r[i + u] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) v += t[u];
  }
  int tot;
  int run = block_scan(v, sh, &tot) - v;
  for (int i = b0; i < b1; i += 8) {
    int t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = i + u < b1 ? blk_r[i + u] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i + u < b1) blk_roff[i + u] = run;
      run += t[u];
    }
  }
  if (threadIdx.x == 0) ctl->total_right = tot;
}

// Per bucket of the level (one lane each): its RIGHT count from the ballot
// words and the block offsets, stored in the record with the RIGHT elements
// before it; the children that still split and their active elements; the
// sides that do not split written as leaves.  Leaf slots are taken with one
// atomic per wave and tree (a bucket list is in tree order, so a wave's
// buckets nearly always belong to one tree).
__global__ __launch_bounds__(256) void forest_bucket_size_kernel(Bucket* __restrict__ bcur,
                                                                 const int32_t* __restrict__ blk_roff,
                                                                 const int32_t* __restrict__ wpre,
                                                                 const uint64_t* __restrict__ rmask, ForestCtl* ctl,
                                                                 int parity, int level, int leaf_eff, int kp1,
                                                                 int32_t* __restrict__ bchild,
                                                                 int32_t* __restrict__ bact, LeafSink sink) {
  const int nb = ctl->nbkt[parity];
  const int nact = ctl->nact[parity];
  const int32_t total = ctl->total_right;
  const int lane = threadIdx.x & 63;
  for (int base = blockIdx.x * 256; base < nb; base += gridDim.x * 256) {
    const int bi = base + threadIdx.x;
    const bool valid = bi < nb;
    Bucket b{};
    int32_t sz[2] = {0, 0};
    if (valid) {
      b = bcur[bi];
      const int32_t rbef = rights_before(b.astart, nact, total, blk_roff, wpre, rmask);
      const int32_t rend = rights_before(static_cast<int64_t>(b.astart) + b.size, nact, total, blk_roff, wpre, rmask);
      sz[1] = rend - rbef;
      sz[0] = b.size - sz[1];
      bcur[bi].rbefore = rbef;
      bcur[bi].nright = sz[1];
      bchild[bi] = static_cast<int>(sz[0] > leaf_eff) + static_cast<int>(sz[1] > leaf_eff);
      bact[bi] = (sz[0] > leaf_eff ? sz[0] : 0) + (sz[1] > leaf_eff ? sz[1] : 0);
    }
    // an empty side (possible only under the bin rule of routed trees) is
    // not a leaf
    const bool leaf0 = sz[0] > 0 && sz[0] <= leaf_eff, leaf1 = sz[1] > 0 && sz[1] <= leaf_eff;
    const int nl = static_cast<int>(leaf0) + static_cast<int>(leaf1);
    const bool splits = valid && (sz[0] > leaf_eff || sz[1] > leaf_eff);
    if (sink.tree_last != nullptr) {
      // this level's children make level + 1 a level of the tree: one store
      // per run of lanes of the same tree
      const int tr = splits ? b.tree : -1;
      const int nxt = __shfl_down(tr, 1);
      if (splits && (lane == 63 || nxt != tr)) sink.tree_last[b.tree] = level + 1;
    }
    uint64_t pending = __ballot(valid && nl > 0);
    while (pending) {
      const int leader = __ffsll(static_cast<long long>(pending)) - 1;
      const int lt = __shfl(b.tree, leader);
      const bool mine = valid && nl > 0 && b.tree == lt;
      const uint64_t grp = __ballot(mine);
      int pre = mine ? nl : 0;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(pre, o);
        if (lane >= o) pre += t;
      }
      const int tot = __shfl(pre, 63);
      pre -= mine ? nl : 0;
      int slot0 = 0;
      if (lane == leader) {
        int32_t* cnt = sink.tree_nleaves != nullptr ? sink.tree_nleaves + lt : &ctl->nleaves;
        slot0 = atomicAdd(cnt, tot);
      }
      slot0 = __shfl(slot0, leader);
      if (mine) {
        int slot = slot0 + pre;
        for (int c = 0; c < 2; ++c) {
          if (!(c ? leaf1 : leaf0)) continue;
          const int32_t size = sz[c];
          const int32_t fst = b.fstart + (c ? sz[0] : 0);
          const bool own_range = size >= kp1;
          sink.leaves[b.tree * sink.stride + slot] =
              Leaf{fst, size, own_range ? fst : b.rstart, own_range ? size : b.rcount};
          ++slot;
        }
      }
      pending &= ~grp;
    }
  }
}

// Exclusive scans, in bucket order, of the children and of their active
// elements (one workgroup, contiguous runs per thread): each bucket's first
// child index and first active offset, in place.  The next level's bucket
// and element counts go to the control block.  A level with no bucket is the
// reference's last loop iteration: nothing splits, the pair is drawn anyway.
__global__ __launch_bounds__(1024) void forest_bucket_scan_kernel(int32_t* __restrict__ bchild,
                                                                  int32_t* __restrict__ bact, ForestCtl* ctl,
                                                                  int parity, int level) {
  __shared__ int sh[16];
  const int nb = ctl->nbkt[parity];
  if (nb == 0) {
    if (threadIdx.x == 0) {
      if (ctl->done == 0) {
        ctl->done = 1;
        ctl->iterations = level + 1;
      }
      ctl->nbkt[parity ^ 1] = 0;
      ctl->nact[parity ^ 1] = 0;
    }
    return;
  }
  const int per = (nb + 1023) / 1024;
  const int c0 = threadIdx.x * per < nb ? threadIdx.x * per : nb;
  const int c1 = c0 + per < nb ? c0 + per : nb;
  int vc = 0, va = 0;
  for (int i = c0; i < c1; i += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i + u < c1) {
        vc += bchild[i + u];
        va += bact[i + u];
      }
    }
  }
  int tot_c, tot_a;
  int rc = block_scan(vc, sh, &tot_c) - vc;
  int ra = block_scan(va, sh, &tot_a) - va;
  for (int i = c0; i < c1; i += 8) {
    int tc[8], ta[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      tc[u] = i + u < c1 ? bchild[i + u] : 0;
      ta[u] = i + u < c1 ? bact[i + u] : 0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i + u < c1) {
        bchild[i + u] = rc;
        bact[i + u] = ra;
      }
      rc += tc[u];
      ra += ta[u];
    }
  }
  if (threadIdx.x == 0) {
    ctl->nbkt[parity ^ 1] = tot_c;
    ctl->nact[parity ^ 1] = tot_a;
    ctl->active_sum += ctl->nact[parity];
  }
}

// The three passes above in one workgroup, for one tree (its bucket lists
// are short enough that a launch costs more than the work): block offsets,
// then buckets 1024 at a time (sizes, children, active elements and leaves,
// each counted by a block scan, leaf slots in bucket order).
__global__ __launch_bounds__(1024) void forest_level_kernel(Bucket* __restrict__ bcur,
                                                            const int32_t* __restrict__ blk_r,
                                                            int32_t* __restrict__ blk_roff,
                                                            const uint64_t* __restrict__ rmask,
                                                            const int32_t* __restrict__ wpre, ForestCtl* ctl,
                                                            int parity, int level, int leaf_eff, int kp1,
                                                            int32_t* __restrict__ bchild, int32_t* __restrict__ bact,
                                                            Leaf* __restrict__ leaves) {
  __shared__ int sh[16];
  const int nb = ctl->nbkt[parity];
  if (nb == 0) {
    if (threadIdx.x == 0) {
      if (ctl->done == 0) {
        ctl->done = 1;
        ctl->iterations = level + 1;
      }
      ctl->nbkt[parity ^ 1] = 0;
      ctl->nact[parity ^ 1] = 0;
    }
    return;
  }
  const int nact = ctl->nact[parity];
  const int nblk = (nact + kSideBlock - 1) / kSideBlock;
  int32_t total_r;
  {
    const int per = (nblk + 1023) / 1024;
    const int b0 = threadIdx.x * per < nblk ? threadIdx.x * per : nblk;
    const int b1 = b0 + per < nblk ? b0 + per : nblk;
    int v = 0;
    for (int i = b0; i < b1; ++i) v += blk_r[i];
    int tot;
    int run = block_scan(v, sh, &tot) - v;
    for (int i = b0; i < b1; ++i) {
      const int c = blk_r[i];
      blk_roff[i] = run;
      run += c;
    }
    total_r = tot;
  }
  __syncthreads();
  int run_c = 0, run_a = 0, run_l = ctl->nleaves;
  for (int base = 0; base < nb; base += 1024) {
    const int bi = base + threadIdx.x;
    Bucket b{};
    int32_t sz[2] = {0, 0};
    if (bi < nb) {
      b = bcur[bi];
      const int32_t rbef = rights_before(b.astart, nact, total_r, blk_roff, wpre, rmask);
      const int32_t rend = rights_before(static_cast<int64_t>(b.astart) + b.size, nact, total_r, blk_roff, wpre, rmask);
      sz[1] = rend - rbef;
      sz[0] = b.size - sz[1];
      bcur[bi].rbefore = rbef;
      bcur[bi].nright = sz[1];
    }
    const bool leaf0 = sz[0] > 0 && sz[0] <= leaf_eff, leaf1 = sz[1] > 0 && sz[1] <= leaf_eff;
    const int nch = static_cast<int>(sz[0] > leaf_eff) + static_cast<int>(sz[1] > leaf_eff);
    const int nl = static_cast<int>(leaf0) + static_cast<int>(leaf1);
    const int na = (sz[0] > leaf_eff ? sz[0] : 0) + (sz[1] > leaf_eff ? sz[1] : 0);
    // children and leaves (at most 2048 each per pass) share one scan
    int tot_cl, tot_a;
    const int icl = block_scan(nch | (nl << 16), sh, &tot_cl) - (nch | (nl << 16));
    const int ia = block_scan(na, sh, &tot_a) - na;
    if (bi < nb) {
      bchild[bi] = run_c + (icl & 0xffff);
      bact[bi] = run_a + ia;
      int slot = run_l + (icl >> 16);
      for (int c = 0; c < 2; ++c) {
        if (!(c ? leaf1 : leaf0)) continue;
        const int32_t fst = b.fstart + (c ? sz[0] : 0);
        const bool own_range = sz[c] >= kp1;
        leaves[slot++] = Leaf{fst, sz[c], own_range ? fst : b.rstart, own_range ? sz[c] : b.rcount};
      }
    }
    run_c += tot_cl & 0xffff;
    run_l += tot_cl >> 16;
    run_a += tot_a;
  }
  if (threadIdx.x == 0) {
    ctl->nbkt[parity ^ 1] = run_c;
    ctl->nact[parity ^ 1] = run_a;
    ctl->nleaves = run_l;
    ctl->active_sum += nact;
  }
}

// One lane per active element: its stable rank on its side from the ballot
// words, then a move to the child's place in the next level's list, or to
// the tree's leaf array when that side does not split.  The lane of each
// bucket's first element writes the bucket's children (Leaf7 pivots for the
// sign split, .cu:2317-2347).
__global__ __launch_bounds__(256) void forest_scatter_kernel(const int32_t* __restrict__ act_ids,
                                                             const int32_t* __restrict__ act_bkt,
                                                             const Bucket* __restrict__ bkt,
                                                             const int32_t* __restrict__ bchild,
                                                             const int32_t* __restrict__ bact,
                                                             const uint64_t* __restrict__ rmask,
                                                             const int32_t* __restrict__ wpre,
                                                             const int32_t* __restrict__ blk_roff, ForestCtl* ctl,
                                                             int parity, int level, int32_t v1, int32_t v2,
                                                             int leaf_eff, int kp1, int32_t* __restrict__ act_ids_n,
                                                             int32_t* __restrict__ act_bkt_n,
                                                             Bucket* __restrict__ bnext, int32_t* __restrict__ perm,
                                                             int64_t perm_stride, int32_t* __restrict__ piv_n) {
  const int nact = ctl->nact[parity];
  for (int64_t m = blockIdx.x * 256ll + threadIdx.x; m < nact; m += gridDim.x * 256ll) {
    const int32_t id = act_ids[m];
    const int32_t bi = act_bkt[m];
    const Bucket& b = bkt[bi];
    const int32_t sz[2] = {b.size - b.nright, b.nright};
    const bool act[2] = {sz[0] > leaf_eff, sz[1] > leaf_eff};
    const uint64_t w = rmask[m >> 6];
    const int bit = static_cast<int>((w >> (m & 63)) & 1ull);
    const int32_t rb = blk_roff[m >> 10] + wpre[m >> 6] + __popcll(w & ((1ull << (m & 63)) - 1ull));
    const int32_t rank = bit ? rb - b.rbefore : (static_cast<int32_t>(m) - b.astart) - (rb - b.rbefore);
    const int32_t cidx0 = bchild[bi], cast0 = bact[bi];
    if (act[bit]) {
      const int32_t off = bit && act[0];
      const int32_t cast = cast0 + (off ? sz[0] : 0);
      act_ids_n[cast + rank] = id;
      act_bkt_n[cast + rank] = cidx0 + off;
      if (piv_n != nullptr) {
        // the child's pivots (the same draws as its record's q1, q2 below)
        const int32_t size = sz[bit];
        const int32_t q1 = v1 % size;
        int32_t q2 = v2 % size;
        if (q1 == q2) q2 = (q1 + 1) % size;
        if (rank == q1) piv_n[2 * (cidx0 + off)] = id;
        if (rank == q2) piv_n[2 * (cidx0 + off) + 1] = id;
      }
    } else {
      perm[b.tree * perm_stride + b.fstart + (bit ? sz[0] : 0) + rank] = id;
    }
    if (m == b.astart) {
      int32_t cidx = cidx0, cast = cast0;
      for (int c = 0; c < 2; ++c) {
        if (!act[c]) continue;
        const int32_t size = sz[c];
        const int32_t fst = b.fstart + (c ? sz[0] : 0);
        const bool own_range = size >= kp1;
        // size >= 2 here
        int32_t q1 = v1 % size, q2 = v2 % size;
        if (q1 == q2) q2 = (q1 + 1) % size;
        Bucket nbk{};
        nbk.astart = cast;
        nbk.size = size;
        nbk.fstart = fst;
        nbk.q1 = q1;
        nbk.q2 = q2;
        nbk.rstart = own_range ? fst : b.rstart;
        nbk.rcount = own_range ? size : b.rcount;
        nbk.tree = b.tree;
        bnext[cidx] = nbk;
        ++cidx;
        cast += size;
      }
    }
  }
}

// ---- median-split rule (beside the reference's sign split) -------------------
//
// A level of a median-split tree splits each bucket by one of three rules:
//   bins:       bit `shift` of the point's router bin (routed trees' first
//               levels), RIGHT when set;
//   projection: the point's projection p = x . r_j (project_kernel) against
//               the bucket's pivot (forest_median_kernel): RIGHT when
//               (ord(p), rel) >= (ord(p*), rel*), rel = position in the bucket;
//   position:   RIGHT when rel >= size / 2 (directions exhausted).
// Both sides of a projection or position split are non-empty for size >= 2.

// Unsigned key whose order is the float order (-0 < +0, NaNs at the ends).
__device__ __forceinline__ uint32_t ord_u32(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

constexpr int kProjGroup = kProjectStep;  // directions per launch (one pass over the rows)
constexpr int kProjChunk = 32;  // dimensions staged in LDS per step
constexpr int kProjXs = kProjChunk + 4;  // LDS row stride (floats): 16-byte reads of 8 lanes hit distinct banks

// For each of G direction groups q: P_q[j * n + i] = x_i . r_qj as one FMA
// chain in dimension order (FP32, the oracle's chain), for j < nd_q <=
// kProjGroup and every point i (act_ids null, rows in order), or every point
// of a level's active list (act_ids[0 .. ctl->nact[parity]): only they, or
// their subsets at deeper levels, read these directions).  One lane per
// point, 256 points per block: each step stages the block's rows for 32
// dimensions in LDS with coalesced 16-byte loads (8 lanes per 128-byte row
// piece), and the directions' values for those dimensions ([group][dim]
// [direction], read by every lane at once); each lane then advances its
// G * kProjGroup chains from its own staged row.  Several trees' first
// directions thus share one pass over the rows.
template <int G>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                                      ProjSet ps, const int32_t* __restrict__ act_ids,
                                                      const ForestCtl* ctl, int parity) {
  __shared__ __attribute__((aligned(16))) float xs[256 * kProjXs];
  __shared__ __attribute__((aligned(16))) float dv[G][kProjChunk][kProjGroup];
  const int t = threadIdx.x;
  const int64_t i0 = blockIdx.x * 256ll;
  const int64_t rows = act_ids != nullptr ? ctl->nact[parity] : n;
  if (i0 >= rows) return;  // the grid covers n; the list is shorter at deep levels
  float acc[G][kProjGroup];
#pragma unroll
  for (int q = 0; q < G; ++q)
#pragma unroll
    for (int g = 0; g < kProjGroup; ++g) acc[q][g] = 0.0f;
  const float* xr = xs + t * kProjXs;
  for (int c0 = 0; c0 < d; c0 += kProjChunk) {
    const int cw = d - c0 < kProjChunk ? d - c0 : kProjChunk;
    __syncthreads();
    // rows i0.. (clamped into the matrix), float4 pieces past d left unread:
    // ld % 4 == 0 and ld >= d, so a piece that starts below d stays in the row
#pragma unroll
    for (int u = 0; u < 256 * kProjChunk / 4 / 256; ++u) {
      const int e = t + 256 * u;
      const int r = e / (kProjChunk / 4), q = e % (kProjChunk / 4);
      const int64_t rr = i0 + r < rows ? i0 + r : rows - 1;
      const int64_t row = act_ids != nullptr ? act_ids[rr] : rr;
      const int c = c0 + 4 * q;
      const float4 v = c < d ? *reinterpret_cast<const float4*>(x + row * ld + c) : make_float4(0.f, 0.f, 0.f, 0.f);
      *reinterpret_cast<float4*>(xs + r * kProjXs + 4 * q) = v;
    }
    // direction values of this chunk (directions past nd read the last one:
    // their chains are not stored; dimensions past d are not read)
    for (int e = t; e < G * kProjChunk * kProjGroup; e += 256) {
      const int q = e / (kProjChunk * kProjGroup), r = e % (kProjChunk * kProjGroup);
      const int g = r / kProjChunk, c = r % kProjChunk;
      const int nq = ps.nd[q];
      dv[q][c][g] = c < cw ? ps.dirs[q][static_cast<int64_t>(g < nq ? g : nq - 1) * ld + c0 + c] : 0.0f;
    }
    __syncthreads();
    for (int c = 0; c < cw; ++c) {
      const float v = xr[c];
#pragma unroll
      for (int q = 0; q < G; ++q)
#pragma unroll
        for (int g4 = 0; g4 < kProjGroup / 4; ++g4) {
          const float4 r = *reinterpret_cast<const float4*>(&dv[q][c][4 * g4]);
          acc[q][4 * g4 + 0] = __builtin_fmaf(v, r.x, acc[q][4 * g4 + 0]);
          acc[q][4 * g4 + 1] = __builtin_fmaf(v, r.y, acc[q][4 * g4 + 1]);
          acc[q][4 * g4 + 2] = __builtin_fmaf(v, r.z, acc[q][4 * g4 + 2]);
          acc[q][4 * g4 + 3] = __builtin_fmaf(v, r.w, acc[q][4 * g4 + 3]);
        }
    }
  }
  if (i0 + t < rows) {
    const int64_t i = act_ids != nullptr ? act_ids[i0 + t] : i0 + t;
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
      for (int g = 0; g < kProjGroup; ++g)
        if (g < ps.nd[q]) ps.proj[q][static_cast<int64_t>(g) * n + i] = acc[q][g];
  }
}

constexpr int kMedianSamples = 511;

// Which rule splits the buckets of batch tree tl at a level (arg: the bin
// bit, or the direction whose projections decide).
__device__ __forceinline__ int rule_of(const MedianRules& r, int tl, int level, int& arg) {
  const int bb = r.tree0 + tl < r.routed_trees ? r.bin_bits : 0;
  if (level < bb) {
    arg = bb - 1 - level;
    return 0;
  }
  if (level - bb < r.nd) {
    arg = level - bb;
    return 1;
  }
  arg = 0;
  return 2;
}

__device__ __forceinline__ const float* rule_proj(const MedianRules& r, int tl, int dir) {
  return r.proj + (static_cast<int64_t>(tl) * r.nd + dir) * r.n;
}

__device__ __forceinline__ const int32_t* rule_bins(const MedianRules& r, int tl) {
  return r.bins + static_cast<int64_t>((r.tree0 + tl) % r.nsets) * r.n;
}

// Pivot of every active bucket under the projection rule: m = min(size, 511)
// samples at relative positions q * size / m, ordered by (ord(p), rel) with a
// bitonic sort in LDS; the sample at index m / 2 is stored as (q1 = rel*,
// q2 = ord(p*)).  Buckets of trees at a bin or position level are skipped.
__global__ __launch_bounds__(256) void forest_median_kernel(MedianRules rules, int level,
                                                            const int32_t* __restrict__ act_ids,
                                                            Bucket* __restrict__ bkt, const ForestCtl* ctl,
                                                            int parity) {
  __shared__ uint64_t key[512];
  const int nb = ctl->nbkt[parity];
  for (int bi = blockIdx.x; bi < nb; bi += gridDim.x) {
    const int32_t astart = bkt[bi].astart, size = bkt[bi].size, tl = bkt[bi].tree;
    int dir;
    if (rule_of(rules, tl, level, dir) != 1) continue;
    const float* pj = rule_proj(rules, tl, dir);
    const int m = size < kMedianSamples ? size : kMedianSamples;
    for (int q = threadIdx.x; q < 512; q += 256) {
      uint64_t kv = ~0ull;
      if (q < m) {
        const int32_t rel = static_cast<int32_t>(static_cast<int64_t>(q) * size / m);
        kv = (static_cast<uint64_t>(ord_u32(pj[act_ids[astart + rel]])) << 32) | static_cast<uint32_t>(rel);
      }
      key[q] = kv;
    }
    __syncthreads();
    for (int kk = 2; kk <= 512; kk <<= 1) {
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const int t = threadIdx.x;
        const int lo = 2 * j * (t / j) + (t % j);
        const int hi = lo + j;
        const bool up = (lo & kk) == 0;
        const uint64_t a = key[lo], c = key[hi];
        if ((a > c) == up) {
          key[lo] = c;
          key[hi] = a;
        }
        __syncthreads();
      }
    }
    if (threadIdx.x == 0) {
      const uint64_t pv = key[m / 2];
      bkt[bi].q1 = static_cast<int32_t>(static_cast<uint32_t>(pv));
      bkt[bi].q2 = static_cast<int32_t>(static_cast<uint32_t>(pv >> 32));
    }
    __syncthreads();
  }
}

// Side pass of a median-split level over the active points of every tree of
// the batch, each under its tree's rule (bins, projection or position), with
// the same ballot words and prefixes as forest_side_kernel.
__global__ __launch_bounds__(256) void forest_side_rule_kernel(MedianRules rules, int level,
                                                               const int32_t* __restrict__ act_ids,
                                                               const int32_t* __restrict__ act_bkt,
                                                               const Bucket* __restrict__ bkt, const ForestCtl* ctl,
                                                               int parity, uint64_t* __restrict__ rmask,
                                                               int32_t* __restrict__ wpre, int32_t* __restrict__ blk_r) {
  __shared__ int wcnt[kWordsPerBlock];
  const int nact = ctl->nact[parity];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  for (int64_t blk = blockIdx.x; blk * kSideBlock < nact; blk += gridDim.x) {
    for (int sub = 0; sub < kSideBlock / 256; ++sub) {
      const int64_t m = blk * kSideBlock + sub * 256 + threadIdx.x;
      bool right = false;
      if (m < nact) {
        const int32_t id = act_ids[m];
        const Bucket& b = bkt[act_bkt[m]];
        int arg;
        const int rule = rule_of(rules, b.tree, level, arg);
        const int32_t rel = static_cast<int32_t>(m) - b.astart;
        if (rule == 0) {
          right = ((rule_bins(rules, b.tree)[id] >> arg) & 1) != 0;
        } else if (rule == 1) {
          const uint32_t u = ord_u32(rule_proj(rules, b.tree, arg)[id]), pu = static_cast<uint32_t>(b.q2);
          right = u > pu || (u == pu && rel >= b.q1);
        } else {
          right = rel >= b.size / 2;
        }
      }
      const uint64_t bal = __ballot(right);
      const int w = sub * 4 + wave;
      if (lane == 0) {
        rmask[blk * kWordsPerBlock + w] = bal;
        wcnt[w] = __popcll(bal);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int w = 0; w < kWordsPerBlock; ++w) {
        wpre[blk * kWordsPerBlock + w] = acc;
        acc += wcnt[w];
      }
      blk_r[blk] = acc;
    }
    __syncthreads();
  }
}

// Roots of a batch of median-split trees: tree tl's points at [tl * n, (tl +
// 1) * n) of the active list, bucket tl.  When the root does not split every
// tree is one leaf of all points in order.
__global__ __launch_bounds__(256) void forest_init_batch_kernel(ForestBuffers fb, int ntrees, int32_t* __restrict__ perm,
                                                                LeafSink sink, int64_t n, int leaf_eff) {
  const bool root_splits = n > leaf_eff;
  const int64_t tot = n * ntrees;
  for (int64_t m = blockIdx.x * 256ll + threadIdx.x; m < tot; m += gridDim.x * 256ll) {
    const int32_t tl = static_cast<int32_t>(m / n);
    const int32_t id = static_cast<int32_t>(m - tl * n);
    fb.act_ids[0][m] = id;
    fb.act_bkt[0][m] = tl;
    if (!root_splits) perm[m] = id;
  }
  if (blockIdx.x == 0) {
    for (int tl = threadIdx.x; tl < ntrees; tl += 256) {
      if (root_splits) {
        Bucket b{};
        b.astart = static_cast<int32_t>(tl * n);
        b.size = static_cast<int32_t>(n);
        b.fstart = 0;
        b.rstart = 0;
        b.rcount = static_cast<int32_t>(n);
        b.tree = tl;
        fb.bkt[0][tl] = b;
        sink.tree_nleaves[tl] = 0;
        sink.tree_last[tl] = 0;
      } else {
        sink.leaves[tl * sink.stride] = Leaf{0, static_cast<int32_t>(n), 0, static_cast<int32_t>(n)};
        sink.tree_nleaves[tl] = 1;
        sink.tree_last[tl] = -1;
      }
    }
    if (threadIdx.x == 0) {
      ForestCtl c{};
      c.nact[0] = root_splits ? static_cast<int32_t>(tot) : 0;
      c.nbkt[0] = root_splits ? ntrees : 0;
      *fb.ctl = c;
    }
  }
}

__global__ __launch_bounds__(256) void leaf_ranges_kernel(const int32_t* __restrict__ perm,
                                                          const Leaf* __restrict__ leaves, int nleaves,
                                                          int32_t* __restrict__ rs, int32_t* __restrict__ rc) {
  const int li = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (li >= nleaves) return;
  const Leaf l = leaves[li];
  for (int p = threadIdx.x & 63; p < l.size; p += 64) {
    const int32_t pt = perm[l.start + p];
    rs[pt] = l.rstart;
    rc[pt] = l.rcount;
  }
}

int clamp_grid(int64_t g, int cap) { return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g)); }

// The passes of a level after its side pass: block offsets, bucket sizes and
// leaves, bucket offsets (one workgroup for a single tree, three passes for a
// batch of trees), then the scatter that also writes the next level's
// buckets.
void level_tail(const ForestBuffers& fb, int par, int level, int32_t v1, int32_t v2, int leaf_eff, int kp1,
                int32_t* perm, LeafSink sink, hipStream_t s) {
  if (sink.tree_nleaves == nullptr) {
    hipLaunchKernelGGL(forest_level_kernel, dim3(1), dim3(1024), 0, s, fb.bkt[par], fb.blk_r, fb.blk_roff, fb.rmask,
                       fb.wpre, fb.ctl, par, level, leaf_eff, kp1, fb.bchild, fb.bact, sink.leaves);
    AKNN_HIP_TRY(hipGetLastError());
  } else {
    hipLaunchKernelGGL(forest_block_scan_kernel, dim3(1), dim3(1024), 0, s, fb.blk_r, fb.blk_roff, fb.ctl, par);
    AKNN_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(forest_bucket_size_kernel, dim3(clamp_grid((fb.max_buckets + 255) / 256, 4096)), dim3(256), 0,
                       s, fb.bkt[par], fb.blk_roff, fb.wpre, fb.rmask, fb.ctl, par, level, leaf_eff, kp1, fb.bchild,
                       fb.bact, sink);
    AKNN_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(forest_bucket_scan_kernel, dim3(1), dim3(1024), 0, s, fb.bchild, fb.bact, fb.ctl, par, level);
    AKNN_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(forest_scatter_kernel, dim3(clamp_grid((fb.n + 255) / 256, 8192)), dim3(256), 0, s,
                     fb.act_ids[par], fb.act_bkt[par], fb.bkt[par], fb.bchild, fb.bact, fb.rmask, fb.wpre, fb.blk_roff,
                     fb.ctl, par, level, v1, v2, leaf_eff, kp1, fb.act_ids[par ^ 1], fb.act_bkt[par ^ 1],
                     fb.bkt[par ^ 1], perm, sink.stride, fb.piv[par ^ 1]);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace

void launch_forest_init(const ForestBuffers& fb, int32_t* perm, Leaf* leaves, int64_t n, int32_t q1, int32_t q2,
                        int leaf_eff, int kp1, hipStream_t s) {
  (void)kp1;
  hipLaunchKernelGGL(forest_init_kernel, dim3(clamp_grid((n + 255) / 256, 4096)), dim3(256), 0, s, fb, perm, leaves,
                     n, q1, q2, leaf_eff);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_forest_level(const ForestBuffers& fb, const float* x, int d, int ld, int32_t* perm, Leaf* leaves,
                         int level, int32_t v1, int32_t v2, int leaf_eff, int kp1, hipStream_t s) {
  const int par = level & 1;
  const dim3 sg(clamp_grid((fb.n + kSideBlock - 1) / kSideBlock, 2048));
  if (fb.xh != nullptr) {
    if (fb.eps == nullptr || fb.nh == nullptr || fb.ldh % 8 != 0 || fb.ldh < ld || ld < d || ld % 4 != 0)
      throw HipError("forest: malformed sketch");
    if (fb.q8 != nullptr && (fb.se8 == nullptr || fb.ld8 % 16 != 0 || fb.ld8 < d))
      throw HipError("forest: malformed int8 sketch");
    hipLaunchKernelGGL(forest_side_sketch_kernel, sg, dim3(256), 0, s, x, d, ld, fb.xh, fb.ldh, fb.eps, fb.nh,
                       fb.n, sketch_bound(d), fb.act_ids[par], fb.act_bkt[par], fb.bkt[par], fb.ctl, par, fb.rmask,
                       fb.wpre, fb.blk_r, fb.amb_count, fb.q8, fb.ld8, fb.se8, fb.piv[par]);
  } else switch (row_width_class(d)) {
    case 8:
      hipLaunchKernelGGL(forest_side_kernel<8>, sg, dim3(256), 0, s, x, d, ld, fb.act_ids[par], fb.act_bkt[par],
                         fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
      break;
    case 16:
      hipLaunchKernelGGL(forest_side_kernel<16>, sg, dim3(256), 0, s, x, d, ld, fb.act_ids[par], fb.act_bkt[par],
                         fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
      break;
    case 24:
      hipLaunchKernelGGL(forest_side_kernel<24>, sg, dim3(256), 0, s, x, d, ld, fb.act_ids[par], fb.act_bkt[par],
                         fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
      break;
    case 32:
      hipLaunchKernelGGL(forest_side_kernel<32>, sg, dim3(256), 0, s, x, d, ld, fb.act_ids[par], fb.act_bkt[par],
                         fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
      break;
    default:
      hipLaunchKernelGGL(forest_side_kernel<0>, sg, dim3(256), 0, s, x, d, ld, fb.act_ids[par], fb.act_bkt[par],
                         fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
  }
  AKNN_HIP_TRY(hipGetLastError());
  level_tail(fb, par, level, v1, v2, leaf_eff, kp1, perm, LeafSink{leaves, 0, nullptr, nullptr}, s);
}

void launch_project(const float* x, int64_t n, int d, int ld, const float* dirs, int nd, float* proj,
                    const ForestBuffers* fb, int level, hipStream_t s) {
  if (nd <= 0 || n <= 0) return;
  if (ld % 4 != 0 || ld < d) throw HipError("project: malformed row width");
  if (nd > kProjGroup) throw HipError("project: at most 16 directions per launch");
  const int par = level & 1;
  ProjSet ps{};
  ps.dirs[0] = dirs;
  ps.proj[0] = proj;
  ps.nd[0] = nd;
  // over an active list the grid covers its capacity (a batch of trees holds
  // up to fb->n entries); the kernel stops at the list's length
  const int64_t rows = fb ? fb->n : n;
  hipLaunchKernelGGL(project_kernel<1>, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, s, x, n, d, ld,
                     ps, fb ? fb->act_ids[par] : nullptr, fb ? fb->ctl : nullptr, par);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_project_set(const float* x, int64_t n, int d, int ld, const ProjSet& ps, int groups, hipStream_t s) {
  if (groups <= 0 || n <= 0) return;
  if (ld % 4 != 0 || ld < d) throw HipError("project: malformed row width");
  for (int q = 0; q < groups; ++q)
    if (ps.nd[q] <= 0 || ps.nd[q] > kProjGroup) throw HipError("project: 1 to 16 directions per group");
  const dim3 g(static_cast<unsigned>((n + 255) / 256)), b(256);
  switch (groups) {
    case 1: hipLaunchKernelGGL(project_kernel<1>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    case 2: hipLaunchKernelGGL(project_kernel<2>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    case 3: hipLaunchKernelGGL(project_kernel<3>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    case 4: hipLaunchKernelGGL(project_kernel<4>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    case 5: hipLaunchKernelGGL(project_kernel<5>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    case 6: hipLaunchKernelGGL(project_kernel<6>, g, b, 0, s, x, n, d, ld, ps, nullptr, nullptr, 0); break;
    default: throw HipError("project: at most 6 direction groups per pass");
  }
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_forest_init_batch(const ForestBuffers& fb, int ntrees, int32_t* perm, Leaf* leaves, int32_t* tree_nleaves,
                              int32_t* tree_last, int64_t n, int leaf_eff, hipStream_t s) {
  if (ntrees < 1 || n * ntrees > fb.n || n * ntrees > 0x7fffffffLL) throw HipError("forest: batch exceeds its buffers");
  hipLaunchKernelGGL(forest_init_batch_kernel, dim3(clamp_grid((n * ntrees + 255) / 256, 8192)), dim3(256), 0, s, fb,
                     ntrees, perm, LeafSink{leaves, n, tree_nleaves, tree_last}, n, leaf_eff);
  AKNN_HIP_TRY(hipGetLastError());
}

void launch_forest_level_batch(const ForestBuffers& fb, const MedianRules& rules, bool any_median, int level,
                               int32_t* perm, Leaf* leaves, int32_t* tree_nleaves, int32_t* tree_last, int leaf_eff,
                               int kp1, hipStream_t s) {
  if (rules.nsets < 1 || (rules.routed_trees > rules.tree0 && rules.bins == nullptr))
    throw HipError("forest: bin rule without bins");
  if (rules.nd > 0 && rules.proj == nullptr) throw HipError("forest: projection rule without projections");
  const int par = level & 1;
  if (any_median) {
    hipLaunchKernelGGL(forest_median_kernel, dim3(clamp_grid(fb.max_buckets, 8192)), dim3(256), 0, s, rules, level,
                       fb.act_ids[par], fb.bkt[par], fb.ctl, par);
    AKNN_HIP_TRY(hipGetLastError());
  }
  const dim3 sg(clamp_grid((fb.n + kSideBlock - 1) / kSideBlock, 4096));
  hipLaunchKernelGGL(forest_side_rule_kernel, sg, dim3(256), 0, s, rules, level, fb.act_ids[par], fb.act_bkt[par],
                     fb.bkt[par], fb.ctl, par, fb.rmask, fb.wpre, fb.blk_r);
  AKNN_HIP_TRY(hipGetLastError());
  // the pivot draw (v1, v2) is unused by these rules
  level_tail(fb, par, level, 0, 0, leaf_eff, kp1, perm, LeafSink{leaves, rules.n, tree_nleaves, tree_last}, s);
}

void launch_leaf_ranges(const int32_t* perm, const Leaf* leaves, int nleaves, int32_t* rs, int32_t* rc,
                        hipStream_t s) {
  if (nleaves <= 0) return;
  hipLaunchKernelGGL(leaf_ranges_kernel, dim3((nleaves + 3) / 4), dim3(256), 0, s, perm, leaves, nleaves, rs, rc);
  AKNN_HIP_TRY(hipGetLastError());
}

}  // namespace aknn
