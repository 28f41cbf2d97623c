// Learned router: MLP forward/backward on the matrix cores, Adam, balanced
// graph labels, and bin assignment.  See router.hpp for the model.
//
// Every product is one NT GEMM kernel, C[M][N] = A[M][K] . B[N][K]^T with
// BF16 operands (K contiguous) on v_mfma_f32_32x32x16_bf16 and FP32
// accumulation, with a fused epilogue: bias, ReLU, ReLU-mask of the backward
// pass, FP32 and BF16 stores, and a transposed BF16 store that feeds the
// next weight-gradient product (dW = dZ^T H is NT over the batch once both
// operands are stored transposed).  Lane layout of the 32x32x16 MFMA:
// A[i][k] in lane i + 32*(k/8), element k%8; B[k][j] in lane j + 32*(k/8),
// element k%8; D[i][j] in lane j + 32*((i/4)%2), register 4*(i/8) + i%4.
#include "router.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "directions.hpp"
#include "engine.hpp"

namespace aknn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16 pair on gfx950's converter (v_cvt_pk_bf16_f32, round to
// nearest even): lo in bits 0..15, hi in bits 16..31.  Every BF16 value the
// library makes goes through it, so the fused and layer-by-layer paths round
// alike.
__device__ __forceinline__ uint32_t bf16_pack2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}
__device__ __forceinline__ uint16_t f2bf(float f) { return static_cast<uint16_t>(bf16_pack2(f, 0.0f)); }

__device__ inline float bf2f(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }

struct GemmArgs {
  const uint16_t* a;
  int lda;
  const uint16_t* b;
  int ldb;
  int m, n, k;
  const float* bias;      // [n] or null
  int relu;               // max(0, .) after the bias
  const uint16_t* mask;   // keep where mask[row][col] > 0 (BF16, ldm), or null
  int ldm;
  float* of;              // FP32 out [m][ldof] or null
  int ldof;
  uint16_t* ob;           // BF16 out [m][ldob] or null
  int ldob;
  uint16_t* obt;          // BF16 out transposed [n][ldobt] or null
  int ldobt;
};

// One 64x64 tile per workgroup, one 32x32 quadrant per wave, K in steps of 16.
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * 64 + (w >> 1) * 32;
  const int col0 = blockIdx.y * 64 + (w & 1) * 32;
  const int li = lane & 31, kh = (lane >> 5) * 8;
  const int64_t ar = row0 + li;
  const int bc = col0 + li;
  const uint16_t* ap = ar < g.m ? g.a + ar * g.lda + kh : nullptr;
  const uint16_t* bp = bc < g.n ? g.b + static_cast<int64_t>(bc) * g.ldb + kh : nullptr;
  f32x16 acc = {};
  for (int k0 = 0; k0 < g.k; k0 += 16) {
    const uint4 av = ap ? *reinterpret_cast<const uint4*>(ap + k0) : make_uint4(0u, 0u, 0u, 0u);
    const uint4 bv = bp ? *reinterpret_cast<const uint4*>(bp + k0) : make_uint4(0u, 0u, 0u, 0u);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                  acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int64_t row = row0 + i;
    const int col = col0 + li;
    if (row < g.m && col < g.n) {
      float v = acc[r];
      if (g.bias) v += g.bias[col];
      if (g.relu) v = v > 0.0f ? v : 0.0f;
      if (g.mask && !(bf2f(g.mask[row * g.ldm + col]) > 0.0f)) v = 0.0f;
      if (g.of) g.of[row * g.ldof + col] = v;
      const uint16_t hb = f2bf(v);
      if (g.ob) g.ob[row * g.ldob + col] = hb;
      if (g.obt) g.obt[static_cast<int64_t>(col) * g.ldobt + row] = hb;
    }
  }
}

// Batch staging: row r < m of the batch is point idx[base + r] (or base + r),
// standardised u = (x - mu) * inv_sigma in FP32, rounded to BF16; rows >= m
// and columns >= d are zero.  u [mpad][dp], ut [dp][mpad] (nullable), and the
// batch labels (-1 past m).
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ x, int ld, int d, int dp,
                                                     const int32_t* __restrict__ idx, int64_t base, int64_t m,
                                                     int64_t mpad, const float* __restrict__ mu, float inv_sigma,
                                                     const int32_t* __restrict__ labels_all,
                                                     uint16_t* __restrict__ u, uint16_t* __restrict__ ut,
                                                     int32_t* __restrict__ lab) {
  const int64_t e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= mpad * dp) return;
  const int64_t r = e / dp;
  const int c = static_cast<int>(e % dp);
  uint16_t b = 0;
  int64_t src = -1;
  if (r < m) {
    src = idx ? idx[base + r] : base + r;
    if (c < d) b = f2bf((x[src * ld + c] - mu[c]) * inv_sigma);
  }
  u[r * dp + c] = b;
  if (ut) ut[static_cast<int64_t>(c) * mpad + r] = b;
  if (lab && c == 0) lab[r] = (r < m && labels_all) ? labels_all[src] : -1;
}

__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Lowest column holding the row maximum (NaN never wins; an all-NaN row gives 0).
__device__ inline int wave_argmax(const float* zr, int nb, int lane) {
  float best = -__builtin_inff();
  int bi = nb;
  for (int c = lane; c < nb; c += 64) {
    const float v = zr[c];
    if (v > best) {
      best = v;
      bi = c;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float ob = __shfl_xor(best, o);
    const int oi = __shfl_xor(bi, o);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  return bi >= nb ? 0 : bi;
}

// Softmax cross-entropy of one row per wave: dz = (softmax - onehot) / m in
// FP32 and BF16 (row-major and transposed), loss and correct argmax summed.
__global__ __launch_bounds__(256) void xent_kernel(const float* __restrict__ z, int nb,
                                                   const int32_t* __restrict__ lab, int64_t mpad, int64_t m,
                                                   float inv_m, float* __restrict__ dzf, uint16_t* __restrict__ dz,
                                                   uint16_t* __restrict__ dzt, double* __restrict__ acc) {
  const int64_t row = blockIdx.x * 4ll + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= mpad) return;
  const int32_t y = row < m ? lab[row] : -1;
  const float* zr = z + row * nb;
  if (y < 0 || y >= nb) {
    for (int c = lane; c < nb; c += 64) {
      dzf[row * nb + c] = 0.0f;
      dz[row * nb + c] = 0;
      dzt[static_cast<int64_t>(c) * mpad + row] = 0;
    }
    return;
  }
  float mx = -__builtin_inff();
  for (int c = lane; c < nb; c += 64) mx = fmaxf(mx, zr[c]);
  mx = wave_max(mx);
  float s = 0.0f;
  for (int c = lane; c < nb; c += 64) s += expf(zr[c] - mx);
  s = wave_sum(s);
  const int am = wave_argmax(zr, nb, lane);
  for (int c = lane; c < nb; c += 64) {
    const float p = expf(zr[c] - mx) / s;
    const float gv = (p - (c == y ? 1.0f : 0.0f)) * inv_m;
    dzf[row * nb + c] = gv;
    const uint16_t hb = f2bf(gv);
    dz[row * nb + c] = hb;
    dzt[static_cast<int64_t>(c) * mpad + row] = hb;
  }
  if (lane == 0) {
    atomicAdd(&acc[0], static_cast<double>(logf(s) + mx - zr[y]));
    if (am == y) atomicAdd(&acc[1], 1.0);
  }
}

__global__ __launch_bounds__(256) void argmax_kernel(const float* __restrict__ z, int nb, int64_t m,
                                                     int32_t* __restrict__ bins) {
  const int64_t row = blockIdx.x * 4ll + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= m) return;
  const int am = wave_argmax(z + row * nb, nb, lane);
  if (lane == 0) bins[row] = am;
}

// Column sums (bias gradients): a block per 64 columns, its sixteen waves
// summing interleaved rows (each lane one column, coalesced rows, eight loads
// in flight per lane), then the sixteen partial sums added in wave order.
__global__ __launch_bounds__(1024) void colsum_kernel(const float* __restrict__ a, int64_t m, int n,
                                                      float* __restrict__ out) {
  __shared__ float part[16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (c < n) {
    int64_t r = wave;
    for (; r + 16 * 7 < m; r += 16 * 8) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = a[(r + 16 * u) * n + c];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; r < m; r += 16) s += a[r * n + c];
  }
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && c < n) {
    float t = part[0][lane];
    for (int w = 1; w < 16; ++w) t += part[w][lane];
    out[c] = t;
  }
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m1, float* __restrict__ m2, int64_t cnt,
                                                   float lr, float bc1, float bc2) {
  const int64_t i = blockIdx.x * 256ll + threadIdx.x;
  if (i >= cnt) return;
  const float gi = g[i];
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float b = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
}

// BF16 copy of a [rows][cols] FP32 matrix, and its transpose (nullable).
__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ src, int rows, int cols,
                                                      uint16_t* __restrict__ dst, uint16_t* __restrict__ dstt) {
  const int64_t e = blockIdx.x * 256ll + threadIdx.x;
  if (e >= static_cast<int64_t>(rows) * cols) return;
  const int r = static_cast<int>(e / cols), c = static_cast<int>(e % cols);
  const uint16_t b = f2bf(src[e]);
  dst[e] = b;
  if (dstt) dstt[static_cast<int64_t>(c) * rows + r] = b;
}

// ---- fused inference: 128 rows per workgroup through all three layers ------
//
// The rows are standardised into LDS as BF16, each layer's activations stay
// in one LDS buffer (row stride padded by 8 halves so the 32 rows of an A
// fragment fall in different banks), the weights stream from L2 as B
// fragments, and the logits never leave registers: each lane's (value,
// column) pairs reduce across its 32-lane half by shuffles, then across waves
// through LDS, to the lowest column holding the row maximum.  Wave w owns
// output column tiles [w*cpw, (w+1)*cpw) of a layer for all four 32-row
// tiles, so a B fragment feeds four MFMAs: the weights cross from L2 once per
// 128 rows (2.5 KB per row for a 128-256-256-256 router).  Per output tile the
// K steps and lane layouts are the ones of gemm_nt_kernel, so the logits, and
// the bins, equal the unfused path's.
constexpr int kFuseRows = 128;                 // rows per workgroup of the fused path
constexpr int kFuseRT = kFuseRows / 32;         // 32-row MFMA tiles per workgroup
constexpr int kFuseH = 256;       // hidden width of the fused path
constexpr int kFuseLd = 256 + 8;  // LDS row stride (halves)

// A [n][k] BF16 weight matrix in MFMA B-fragment order: fragment (column
// tile ct, K step ks) is 64 lanes x 8 halves stored contiguously, lane l
// holding row ct*32 + l%32, columns ks*16 + (l/32)*8 .. +8, so a wave reads
// one fragment as 1 KB of consecutive bytes.
__global__ __launch_bounds__(256) void pack_fragments_kernel(const uint16_t* __restrict__ w, int n, int k,
                                                             uint16_t* __restrict__ out) {
  const int64_t e = blockIdx.x * 256ll + threadIdx.x;  // one 16-byte piece
  const int ksn = k / 16;
  if (e >= static_cast<int64_t>(n / 32) * ksn * 64) return;
  const int lane = static_cast<int>(e % 64);
  const int64_t f = e / 64;
  const int ks = static_cast<int>(f % ksn), ct = static_cast<int>(f / ksn);
  const uint4 v = *reinterpret_cast<const uint4*>(w + static_cast<int64_t>(ct * 32 + (lane & 31)) * k + ks * 16 +
                                                  (lane >> 5) * 8);
  *reinterpret_cast<uint4*>(out + e * 8) = v;
}

// One layer of the fused path for a workgroup's kFuseRows rows: wave w
// computes column tiles w*CPW .. w*CPW+CPW-1 of all kFuseRT row tiles, so each
// B fragment read from L2 feeds kFuseRT MFMAs.  The B fragments stream in
// groups of kFuseGroup K steps, double-buffered in registers: a group's loads
// are issued a whole group of MFMAs ahead of their use.  The A fragments
// (LDS) of step ks+1 are loaded during the MFMAs of step ks.  Every
// accumulator still takes its K steps in order.
constexpr int kFuseGroup = 4;

template <int CPW>
__device__ __forceinline__ void fused_layer(const uint16_t* __restrict__ src, const uint16_t* __restrict__ w, int k,
                                            int n, int wave, int lane, f32x16 (&acc)[2][kFuseRT]) {
  const int li = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int rt = 0; rt < kFuseRT; ++rt) acc[c][rt] = f32x16{};
  const int nt = n / 32, ksn = k / 16;
  auto load_a = [&](uint4 (&a)[kFuseRT], int ks) {
#pragma unroll
    for (int rt = 0; rt < kFuseRT; ++rt)
      a[rt] = *reinterpret_cast<const uint4*>(src + (rt * 32 + li) * kFuseLd + ks * 16 + kh);
  };
  // w is fragment-packed: fragment (ct, ks) is 64 consecutive 16-byte pieces.
  // Steps past k and tiles past n load valid fragments that are never used,
  // so no load sits behind a branch.
  auto load_group = [&](uint4 (&b)[kFuseGroup][CPW], int g0) {
#pragma unroll
    for (int j = 0; j < kFuseGroup; ++j) {
      const int ks = g0 + j < ksn ? g0 + j : ksn - 1;
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const int ct = wave * CPW + c < nt ? wave * CPW + c : nt - 1;
        b[j][c] = *reinterpret_cast<const uint4*>(w + ((static_cast<int64_t>(ct) * ksn + ks) * 64 + lane) * 8);
      }
    }
  };
  auto mma = [&](const uint4 (&a)[kFuseRT], const uint4 (&b)[CPW]) {
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      if (wave * CPW + c >= nt) continue;
      const bf16x8 bb = __builtin_bit_cast(bf16x8, b[c]);
#pragma unroll
      for (int rt = 0; rt < kFuseRT; ++rt)
        acc[c][rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bb, __builtin_bit_cast(bf16x8, a[rt]), acc[c][rt], 0, 0, 0);
    }
  };
  auto run_group = [&](const uint4 (&b)[kFuseGroup][CPW], int g0) {
    uint4 a0[kFuseRT], a1[kFuseRT];
    load_a(a0, g0);
#pragma unroll
    for (int j = 0; j < kFuseGroup; j += 2) {
      if (g0 + j >= ksn) break;
      if (g0 + j + 1 < ksn) load_a(a1, g0 + j + 1);
      mma(a0, b[j]);
      if (g0 + j + 1 >= ksn) break;
      if (j + 2 < kFuseGroup && g0 + j + 2 < ksn) load_a(a0, g0 + j + 2);
      mma(a1, b[j + 1]);
    }
  };
  // sched_barrier keeps each group's loads ahead of the previous group's
  // MFMAs: regrouped, the loads would be waited for as soon as issued
  uint4 bA[kFuseGroup][CPW], bB[kFuseGroup][CPW];
  load_group(bA, 0);
  for (int g0 = 0; g0 < ksn; g0 += 2 * kFuseGroup) {
    if (g0 + kFuseGroup < ksn) load_group(bB, g0 + kFuseGroup);
    __builtin_amdgcn_sched_barrier(0);
    run_group(bA, g0);
    __builtin_amdgcn_sched_barrier(0);
    if (g0 + kFuseGroup < ksn) {
      if (g0 + 2 * kFuseGroup < ksn) load_group(bA, g0 + 2 * kFuseGroup);
      __builtin_amdgcn_sched_barrier(0);
      run_group(bB, g0 + kFuseGroup);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// acc[c][rt] holds the transposed tile: lane (i + 32*hh) has data row
// rt*32 + i and, in registers 4g..4g+3, the four consecutive columns
// ct*32 + 8g + 4hh + 0..3.  Bias, ReLU and BF16 rounding, then one 8-byte
// LDS store per four columns.
template <int CPW>
__device__ __forceinline__ void fused_hidden_store(const f32x16 (&acc)[2][kFuseRT], const float* __restrict__ sbias,
                                                   int n, int wave, int lane, uint16_t* __restrict__ dst) {
  const int il = lane & 31, hh = lane >> 5;
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    const int ct = wave * CPW + c;
    if (ct >= n / 32) continue;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int n0 = ct * 32 + 8 * g + 4 * hh;
      const float4 b4 = *reinterpret_cast<const float4*>(sbias + n0);
#pragma unroll
      for (int rt = 0; rt < kFuseRT; ++rt) {
        // ReLU as gemm_nt_kernel's epilogue writes it (NaN and -0 give +0)
        const float u0 = acc[c][rt][4 * g + 0] + b4.x, u1 = acc[c][rt][4 * g + 1] + b4.y;
        const float u2 = acc[c][rt][4 * g + 2] + b4.z, u3 = acc[c][rt][4 * g + 3] + b4.w;
        const float v0 = u0 > 0.0f ? u0 : 0.0f, v1 = u1 > 0.0f ? u1 : 0.0f;
        const float v2 = u2 > 0.0f ? u2 : 0.0f, v3 = u3 > 0.0f ? u3 : 0.0f;
        *reinterpret_cast<uint2*>(dst + (rt * 32 + il) * kFuseLd + n0) =
            make_uint2(bf16_pack2(v0, v1), bf16_pack2(v2, v3));
      }
    }
  }
}

// Router inference for kFuseRows rows per workgroup through all three
// layers.  The activations live in one LDS buffer that each hidden layer
// overwrites once every wave has read all of it; the logits stay in
// registers and reduce to the arg-max (lowest bin on ties).
template <int CPW3>
__global__ __launch_bounds__(256, 2) void router_fused_kernel(const float* __restrict__ x, int ld, int d, int dp,
                                                           int64_t n, const float* __restrict__ mu, float inv_sigma,
                                                           const uint16_t* __restrict__ w1, const float* __restrict__ b1,
                                                           const uint16_t* __restrict__ w2, const float* __restrict__ b2,
                                                           const uint16_t* __restrict__ w3, const float* __restrict__ b3,
                                                           int nb, int32_t* __restrict__ bins) {
  __shared__ __attribute__((aligned(16))) uint16_t buf[kFuseRows * kFuseLd];
  __shared__ __attribute__((aligned(16))) float smu[kFuseH];
  __shared__ __attribute__((aligned(16))) float sb[3][kFuseH];  // b1, b2, b3 (zero past nb)
  __shared__ float bval[4][kFuseRows];
  __shared__ int bidx[4][kFuseRows];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int il = lane & 31, hh = lane >> 5;
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kFuseRows;
  // the mean (zero past d) and the biases; dp, nb <= kFuseH = blockDim
  {
    const int t = threadIdx.x;
    if (t < dp) smu[t] = t < d ? mu[t] : 0.0f;
    sb[0][t] = b1[t];
    sb[1][t] = b2[t];
    sb[2][t] = t < nb ? b3[t] : 0.0f;
  }
  __syncthreads();
  // standardised rows into buf (BF16), zero past d and past n: in rounds of
  // 16 float4 per thread, every load of a round issued before any is converted
  {
    const int q4 = dp / 4;  // 4-float groups per row (dp % 16 == 0, ld % 4 == 0)
    const int total = kFuseRows * q4;
    for (int base = 0; base < total; base += 256 * 16) {
      float4 v[16];
      // addresses clamped into the matrix (ld % 4 == 0), so no load sits
      // behind a branch; the conversion zeroes what lies past n or d
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = base + threadIdx.x + 256 * u < total ? base + threadIdx.x + 256 * u : total - 1;
        const int r = e / q4, c = 4 * (e % q4);
        const int64_t row = row0 + r < n ? row0 + r : n - 1;
        v[u] = *reinterpret_cast<const float4*>(x + row * ld + (c < ld ? c : ld - 4));
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = base + threadIdx.x + 256 * u;
        if (e < total) {
          const int r = e / q4, c = 4 * (e % q4);
          const float4 m4 = *reinterpret_cast<const float4*>(smu + c);
          const bool live = row0 + r < n;
          const float f0 = live && c + 0 < d ? (v[u].x - m4.x) * inv_sigma : 0.0f;
          const float f1 = live && c + 1 < d ? (v[u].y - m4.y) * inv_sigma : 0.0f;
          const float f2 = live && c + 2 < d ? (v[u].z - m4.z) * inv_sigma : 0.0f;
          const float f3 = live && c + 3 < d ? (v[u].w - m4.w) * inv_sigma : 0.0f;
          *reinterpret_cast<uint2*>(buf + r * kFuseLd + c) = make_uint2(bf16_pack2(f0, f1), bf16_pack2(f2, f3));
        }
      }
    }
  }
  __syncthreads();
  f32x16 acc[2][kFuseRT];
  fused_layer<2>(buf, w1, dp, kFuseH, wave, lane, acc);
  __syncthreads();  // every wave has read the layer's input
  fused_hidden_store<2>(acc, sb[0], kFuseH, wave, lane, buf);
  __syncthreads();
  fused_layer<2>(buf, w2, kFuseH, kFuseH, wave, lane, acc);
  __syncthreads();
  fused_hidden_store<2>(acc, sb[1], kFuseH, wave, lane, buf);
  __syncthreads();
  fused_layer<CPW3>(buf, w3, kFuseH, nb, wave, lane, acc);
  // argmax.  Lane (i + 32*hh) holds, for data row rt*32 + i, the columns
  // ct*32 + 8g + 4hh + q: it visits them in increasing order (strict '>', so
  // the lowest column wins a tie and NaN never wins), then meets lane ^ 32
  // (the other four of each eight columns), then the other waves in LDS.
#pragma unroll
  for (int rt = 0; rt < kFuseRT; ++rt) {
    float best = -__builtin_inff();
    int bi = nb;
#pragma unroll
    for (int c = 0; c < CPW3; ++c) {
      const int ct = wave * CPW3 + c;
      if (ct >= nb / 32) continue;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int n0 = ct * 32 + 8 * g + 4 * hh;
        const float4 b4 = *reinterpret_cast<const float4*>(&sb[2][n0]);
        const float u[4] = {acc[c][rt][4 * g] + b4.x, acc[c][rt][4 * g + 1] + b4.y, acc[c][rt][4 * g + 2] + b4.z,
                            acc[c][rt][4 * g + 3] + b4.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (u[q] > best) {
            best = u[q];
            bi = n0 + q;
          }
      }
    }
    const float ob = __shfl_xor(best, 32);
    const int oi = __shfl_xor(bi, 32);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
    if (hh == 0) {
      bval[wave][rt * 32 + il] = best;
      bidx[wave][rt * 32 + il] = bi;
    }
  }
  __syncthreads();
  if (threadIdx.x < kFuseRows) {
    const int row = threadIdx.x;
    float best = -__builtin_inff();
    int bi = nb;
    for (int w = 0; w < 4; ++w) {
      const float v = bval[w][row];
      const int i = bidx[w][row];
      if (v > best || (v == best && i < bi)) {
        best = v;
        bi = i;
      }
    }
    if (row0 + row < n) bins[row0 + row] = bi >= nb ? 0 : bi;
  }
}

unsigned grid_of(int64_t work, int per) { return static_cast<unsigned>((work + per - 1) / per); }

void gemm(const GemmArgs& g, hipStream_t s) {
  if (g.k % 16 != 0 || g.lda % 8 != 0 || g.ldb % 8 != 0) throw HipError("router gemm: K and strides must be aligned");
  if (g.m <= 0 || g.n <= 0) return;
  hipLaunchKernelGGL(gemm_nt_kernel, dim3(grid_of(g.m, 64), grid_of(g.n, 64)), dim3(256), 0, s, g);
  AKNN_HIP_TRY(hipGetLastError());
}

GemmArgs gemm_args(const uint16_t* a, int lda, const uint16_t* b, int ldb, int m, int n, int k) {
  GemmArgs g{};
  g.a = a;
  g.lda = lda;
  g.b = b;
  g.ldb = ldb;
  g.m = m;
  g.n = n;
  g.k = k;
  return g;
}

constexpr int64_t kAssignChunk = 1 << 17;  // rows per assignment pass

}  // namespace

template <typename T>
T* Router::alloc(size_t count) {
  void* p = nullptr;
  AKNN_HIP_TRY(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  allocs_.push_back(p);
  return static_cast<T*>(p);
}

Router::Router(int device, int d, int hidden, int bin_bits, uint32_t seed)
    : device_(device), d_(d), h_(hidden), bin_bits_(bin_bits), seed_(seed) {
  if (d < 1 || d > 65536) throw std::invalid_argument("router: D must be in [1, 65536]");
  if (hidden < 32 || hidden > 4096 || hidden % 32 != 0)
    throw std::invalid_argument("router: hidden width must be a multiple of 32 in [32, 4096]");
  if (bin_bits < 5 || bin_bits > 12) throw std::invalid_argument("router: bin_bits must be in [5, 12]");
  int count = 0;
  AKNN_HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) throw HipError("device " + std::to_string(device) + " is not present");
  AKNN_HIP_TRY(hipSetDevice(device_));
  AKNN_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  AKNN_HIP_TRY(hipEventCreate(&ev_[0]));
  AKNN_HIP_TRY(hipEventCreate(&ev_[1]));
  dp_ = (d + 15) / 16 * 16;
  nb_ = 1 << bin_bits;
  off_[0] = 0;
  off_[1] = off_[0] + static_cast<size_t>(h_) * dp_;
  off_[2] = off_[1] + h_;
  off_[3] = off_[2] + static_cast<size_t>(h_) * h_;
  off_[4] = off_[3] + h_;
  off_[5] = off_[4] + static_cast<size_t>(nb_) * h_;
  off_[6] = off_[5] + nb_;
  const size_t np = off_[6];
  p_ = alloc<float>(np);
  g_ = alloc<float>(np);
  m1_ = alloc<float>(np);
  m2_ = alloc<float>(np);
  w1b_ = alloc<uint16_t>(static_cast<size_t>(h_) * dp_);
  w2b_ = alloc<uint16_t>(static_cast<size_t>(h_) * h_);
  w2t_ = alloc<uint16_t>(static_cast<size_t>(h_) * h_);
  w3b_ = alloc<uint16_t>(static_cast<size_t>(nb_) * h_);
  w3t_ = alloc<uint16_t>(static_cast<size_t>(nb_) * h_);
  w1p_ = alloc<uint16_t>(static_cast<size_t>(h_) * dp_);
  w2p_ = alloc<uint16_t>(static_cast<size_t>(h_) * h_);
  w3p_ = alloc<uint16_t>(static_cast<size_t>(nb_) * h_);
  mu_ = alloc<float>(dp_);
  acc_ = alloc<double>(2);
  // He-normal weights from the seed (splitmix64 + Box-Muller), zero biases;
  // the padding columns of W1 stay zero (their inputs are zero, so are their
  // gradients and Adam steps)
  std::vector<float> host(np, 0.0f);
  SplitMix64 gen((static_cast<uint64_t>(seed) << 32) ^ 0x5851F42D4C957F2Dull);
  auto normal = [&]() {
    const double u1 = gen.open01(), u2 = gen.open01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  for (int o = 0; o < h_; ++o)
    for (int c = 0; c < d_; ++c) host[off_[0] + static_cast<size_t>(o) * dp_ + c] = static_cast<float>(normal() * std::sqrt(2.0 / d_));
  for (size_t i = off_[2]; i < off_[3]; ++i) host[i] = static_cast<float>(normal() * std::sqrt(2.0 / h_));
  for (size_t i = off_[4]; i < off_[5]; ++i) host[i] = static_cast<float>(normal() * std::sqrt(2.0 / h_));
  mu_host_.assign(dp_, 0.0f);
  set_params(host.data(), mu_host_.data(), 1.0f);
}

Router::~Router() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : allocs_) (void)hipFree(p);
  for (auto e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Router::refresh_weights(hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_of(static_cast<int64_t>(h_) * dp_, 256)), dim3(256), 0, s, p_ + off_[0],
                     h_, dp_, w1b_, static_cast<uint16_t*>(nullptr));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_of(static_cast<int64_t>(h_) * h_, 256)), dim3(256), 0, s, p_ + off_[2],
                     h_, h_, w2b_, w2t_);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_of(static_cast<int64_t>(nb_) * h_, 256)), dim3(256), 0, s,
                     p_ + off_[4], nb_, h_, w3b_, w3t_);
  AKNN_HIP_TRY(hipGetLastError());
  if (w1p_) {
    // fragment-ordered copies for the fused inference kernel
    hipLaunchKernelGGL(pack_fragments_kernel, dim3(grid_of(static_cast<int64_t>(h_) * dp_ / 8, 256)), dim3(256), 0, s,
                       w1b_, h_, dp_, w1p_);
    hipLaunchKernelGGL(pack_fragments_kernel, dim3(grid_of(static_cast<int64_t>(h_) * h_ / 8, 256)), dim3(256), 0, s,
                       w2b_, h_, h_, w2p_);
    hipLaunchKernelGGL(pack_fragments_kernel, dim3(grid_of(static_cast<int64_t>(nb_) * h_ / 8, 256)), dim3(256), 0, s,
                       w3b_, nb_, h_, w3p_);
    AKNN_HIP_TRY(hipGetLastError());
  }
}

void Router::ensure_batch(int64_t mpad) {
  if (mpad <= cap_) return;
  const size_t m = static_cast<size_t>(mpad);
  u_ = alloc<uint16_t>(m * dp_);
  ut_ = alloc<uint16_t>(m * dp_);
  h1_ = alloc<uint16_t>(m * h_);
  h1t_ = alloc<uint16_t>(m * h_);
  h2_ = alloc<uint16_t>(m * h_);
  h2t_ = alloc<uint16_t>(m * h_);
  z_ = alloc<float>(m * nb_);
  dz3f_ = alloc<float>(m * nb_);
  dz3_ = alloc<uint16_t>(m * nb_);
  dz3t_ = alloc<uint16_t>(m * nb_);
  dz2f_ = alloc<float>(m * h_);
  dz2_ = alloc<uint16_t>(m * h_);
  dz2t_ = alloc<uint16_t>(m * h_);
  dz1f_ = alloc<float>(m * h_);
  dz1t_ = alloc<uint16_t>(m * h_);
  lab_ = alloc<int32_t>(m);
  cap_ = mpad;
}

void Router::forward(int64_t mpad, hipStream_t s, bool train) {
  const int m = static_cast<int>(mpad);
  GemmArgs g = gemm_args(u_, dp_, w1b_, dp_, m, h_, dp_);
  g.bias = p_ + off_[1];
  g.relu = 1;
  g.ob = h1_;
  g.ldob = h_;
  g.obt = train ? h1t_ : nullptr;
  g.ldobt = m;
  gemm(g, s);
  g = gemm_args(h1_, h_, w2b_, h_, m, h_, h_);
  g.bias = p_ + off_[3];
  g.relu = 1;
  g.ob = h2_;
  g.ldob = h_;
  g.obt = train ? h2t_ : nullptr;
  g.ldobt = m;
  gemm(g, s);
  g = gemm_args(h2_, h_, w3b_, h_, m, nb_, h_);
  g.bias = p_ + off_[5];
  g.of = z_;
  g.ldof = nb_;
  gemm(g, s);
}

void Router::backward(int64_t mpad, int64_t m, int64_t count, hipStream_t s) {
  const int mp = static_cast<int>(mpad);
  hipLaunchKernelGGL(xent_kernel, dim3(grid_of(mpad, 4)), dim3(256), 0, s, z_, nb_, lab_, mpad, m,
                     1.0f / static_cast<float>(std::max<int64_t>(count, 1)), dz3f_, dz3_, dz3t_, acc_);
  AKNN_HIP_TRY(hipGetLastError());
  // dW3 = dZ3^T H2, db3 = column sums of dZ3
  GemmArgs g = gemm_args(dz3t_, mp, h2t_, mp, nb_, h_, mp);
  g.of = g_ + off_[4];
  g.ldof = h_;
  gemm(g, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(grid_of(nb_, 64)), dim3(1024), 0, s, dz3f_, mpad, nb_, g_ + off_[5]);
  // dZ2 = (dZ3 W3) masked by H2 > 0
  g = gemm_args(dz3_, nb_, w3t_, nb_, mp, h_, nb_);
  g.mask = h2_;
  g.ldm = h_;
  g.of = dz2f_;
  g.ldof = h_;
  g.ob = dz2_;
  g.ldob = h_;
  g.obt = dz2t_;
  g.ldobt = mp;
  gemm(g, s);
  g = gemm_args(dz2t_, mp, h1t_, mp, h_, h_, mp);
  g.of = g_ + off_[2];
  g.ldof = h_;
  gemm(g, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(grid_of(h_, 64)), dim3(1024), 0, s, dz2f_, mpad, h_, g_ + off_[3]);
  // dZ1 = (dZ2 W2) masked by H1 > 0
  g = gemm_args(dz2_, h_, w2t_, h_, mp, h_, h_);
  g.mask = h1_;
  g.ldm = h_;
  g.of = dz1f_;
  g.ldof = h_;
  g.obt = dz1t_;
  g.ldobt = mp;
  gemm(g, s);
  g = gemm_args(dz1t_, mp, ut_, mp, h_, dp_, mp);
  g.of = g_ + off_[0];
  g.ldof = dp_;
  gemm(g, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(grid_of(h_, 64)), dim3(1024), 0, s, dz1f_, mpad, h_, g_ + off_[1]);
  AKNN_HIP_TRY(hipGetLastError());
}

void Router::graph_labels(Engine& e, int sweeps, std::vector<int32_t>& labels) {
  const int64_t n = e.n();
  const int k = e.params().k;
  if (n < nb_) throw std::invalid_argument("router: fewer training points than bins");
  std::vector<int32_t> ids(static_cast<size_t>(n) * k), perm(n), rs(n), rc(n);
  std::vector<float> vals(static_cast<size_t>(n) * k);
  e.get_graph(ids.data(), vals.data());
  // routers with different seeds start from different trees' leaf orders
  e.get_forest(static_cast<int>(seed_ % static_cast<uint32_t>(e.params().trees)), perm.data(), rs.data(), rc.data());
  labels.assign(n, 0);
  for (int64_t p = 0; p < n; ++p) labels[perm[p]] = static_cast<int32_t>(p * nb_ / n);
  // symmetrised adjacency (CSR): every graph edge in both directions
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) {
      const int32_t t = ids[i * k + j];
      if (t >= 0 && t != i) {
        ++start[i + 1];
        ++start[t + 1];
      }
    }
  for (int64_t i = 0; i < n; ++i) start[i + 1] += start[i];
  std::vector<int32_t> adj(start[n]);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) {
      const int32_t t = ids[i * k + j];
      if (t >= 0 && t != i) {
        adj[fill[i]++] = t;
        adj[fill[t]++] = static_cast<int32_t>(i);
      }
    }
  auto cut = [&]() {
    int64_t c = 0, tot = 0;
    for (int64_t i = 0; i < n; ++i)
      for (int j = 0; j < k; ++j) {
        const int32_t t = ids[i * k + j];
        if (t >= 0) {
          ++tot;
          c += labels[t] != labels[i];
        }
      }
    return tot ? static_cast<double>(c) / tot : 0.0;
  };
  stats_.cut_before = cut();
  // capacity-bounded label propagation, points in index order: a point moves
  // to the part holding strictly more of its neighbours than its own, if the
  // target is below 1.05 n/bins and its own part above 0.95 n/bins
  std::vector<int64_t> size(nb_, 0);
  for (int64_t i = 0; i < n; ++i) ++size[labels[i]];
  const int64_t hi = static_cast<int64_t>(std::ceil(1.05 * n / nb_)), lo = static_cast<int64_t>(std::floor(0.95 * n / nb_));
  std::vector<int32_t> cnt(nb_, 0), touched;
  for (int sw = 0; sw < sweeps; ++sw) {
    int64_t moved = 0;
    for (int64_t i = 0; i < n; ++i) {
      touched.clear();
      for (int64_t a = start[i]; a < start[i + 1]; ++a) {
        const int32_t l = labels[adj[a]];
        if (cnt[l]++ == 0) touched.push_back(l);
      }
      const int32_t own = labels[i];
      int32_t best = own, bc = cnt[own];
      for (int32_t l : touched)
        if (cnt[l] > bc || (cnt[l] == bc && best != own && l < best)) {
          best = l;
          bc = cnt[l];
        }
      for (int32_t l : touched) cnt[l] = 0;
      if (best != own && size[best] < hi && size[own] > lo) {
        labels[i] = best;
        --size[own];
        ++size[best];
        ++moved;
      }
    }
    if (moved == 0) break;
  }
  stats_.cut_after = cut();
}

void Router::fit(Engine& e, const std::vector<int32_t>& labels, int epochs, int batch, float lr) {
  if (e.device() != device_) throw std::invalid_argument("router and graph live on different devices");
  if (e.d() != d_) throw std::invalid_argument("router and graph have different D");
  const int64_t n = e.n();
  if (static_cast<int64_t>(labels.size()) != n) throw std::invalid_argument("one label per training point");
  for (int32_t l : labels)
    if (l < 0 || l >= nb_) throw std::invalid_argument("label out of range");
  if (epochs < 1 || batch < 16 || batch > (1 << 16) || !(lr > 0.0f)) throw std::invalid_argument("bad training settings");
  AKNN_HIP_TRY(hipSetDevice(device_));
  const auto t0 = std::chrono::steady_clock::now();
  const int ld = e.ld();
  const float* x = e.raw_points();
  // standardisation: per-dimension mean, one global scale (RMS of the centred values)
  {
    std::vector<float> host(static_cast<size_t>(n) * ld);
    AKNN_HIP_TRY(hipMemcpy(host.data(), x, sizeof(float) * host.size(), hipMemcpyDeviceToHost));
    std::vector<double> mu(d_, 0.0);
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < d_; ++c) mu[c] += host[i * ld + c];
    double ss = 0.0;
    for (int c = 0; c < d_; ++c) mu[c] /= static_cast<double>(n);
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < d_; ++c) {
        const double v = host[i * ld + c] - mu[c];
        ss += v * v;
      }
    const double sigma = std::sqrt(ss / (static_cast<double>(n) * d_));
    mu_host_.assign(dp_, 0.0f);
    for (int c = 0; c < d_; ++c) mu_host_[c] = static_cast<float>(mu[c]);
    inv_sigma_ = sigma > 0.0 ? static_cast<float>(1.0 / sigma) : 1.0f;
    AKNN_HIP_TRY(hipMemcpy(mu_, mu_host_.data(), sizeof(float) * dp_, hipMemcpyHostToDevice));
  }
  int32_t* lab_all = alloc<int32_t>(n);
  idx_ = alloc<int32_t>(n);
  AKNN_HIP_TRY(hipMemcpy(lab_all, labels.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  const int64_t mpad = (static_cast<int64_t>(batch) + 63) / 64 * 64;
  ensure_batch(mpad);
  AKNN_HIP_TRY(hipMemsetAsync(m1_, 0, sizeof(float) * param_count(), stream_));
  AKNN_HIP_TRY(hipMemsetAsync(m2_, 0, sizeof(float) * param_count(), stream_));
  step_ = 0;
  std::vector<int32_t> order(n);
  SplitMix64 shuf((static_cast<uint64_t>(seed_) << 32) ^ 0x2545F4914F6CDD1Dull);
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = 0; i < n; ++i) order[i] = static_cast<int32_t>(i);
    for (int64_t i = n - 1; i > 0; --i) std::swap(order[i], order[shuf.next() % static_cast<uint64_t>(i + 1)]);
    AKNN_HIP_TRY(hipMemcpyAsync(idx_, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream_));
    AKNN_HIP_TRY(hipMemsetAsync(acc_, 0, sizeof(double) * 2, stream_));
    for (int64_t base = 0; base < n; base += batch) {
      const int64_t m = std::min<int64_t>(batch, n - base);
      hipLaunchKernelGGL(gather_kernel, dim3(grid_of(mpad * dp_, 256)), dim3(256), 0, stream_, x, ld, d_, dp_, idx_,
                         base, m, mpad, mu_, inv_sigma_, lab_all, u_, ut_, lab_);
      AKNN_HIP_TRY(hipGetLastError());
      forward(mpad, stream_, true);
      backward(mpad, m, m, stream_);
      ++step_;
      const float bc1 = 1.0f - std::pow(0.9f, static_cast<float>(step_));
      const float bc2 = 1.0f - std::pow(0.999f, static_cast<float>(step_));
      hipLaunchKernelGGL(adam_kernel, dim3(grid_of(static_cast<int64_t>(param_count()), 256)), dim3(256), 0, stream_,
                         p_, g_, m1_, m2_, static_cast<int64_t>(param_count()), lr, bc1, bc2);
      AKNN_HIP_TRY(hipGetLastError());
      refresh_weights(stream_);
    }
    double acc[2] = {0, 0};
    AKNN_HIP_TRY(hipMemcpyAsync(acc, acc_, sizeof(acc), hipMemcpyDeviceToHost, stream_));
    AKNN_HIP_TRY(hipStreamSynchronize(stream_));
    stats_.loss = acc[0] / n;
    stats_.accuracy = acc[1] / n;
    stats_.epochs = ep + 1;
  }
  stats_.train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

double Router::gradients(const float* x_raw, int64_t m, const int32_t* labels, float* grads) {
  if (m < 1 || m > (1 << 20)) throw std::invalid_argument("router: batch must hold 1..2^20 rows");
  for (int64_t i = 0; i < m; ++i)
    if (labels[i] >= nb_) throw std::invalid_argument("label out of range");
  AKNN_HIP_TRY(hipSetDevice(device_));
  const int64_t mpad = (m + 63) / 64 * 64;
  ensure_batch(mpad);
  float* xd = nullptr;
  int32_t* ld = nullptr;
  AKNN_HIP_TRY(hipMalloc(&xd, sizeof(float) * m * d_));
  if (hipMalloc(&ld, sizeof(int32_t) * m) != hipSuccess) {
    (void)hipFree(xd);
    throw HipError("router: out of device memory");
  }
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{xd, ld};
  AKNN_HIP_TRY(hipMemcpy(xd, x_raw, sizeof(float) * m * d_, hipMemcpyHostToDevice));
  AKNN_HIP_TRY(hipMemcpy(ld, labels, sizeof(int32_t) * m, hipMemcpyHostToDevice));
  AKNN_HIP_TRY(hipMemsetAsync(acc_, 0, sizeof(double) * 2, stream_));
  hipLaunchKernelGGL(gather_kernel, dim3(grid_of(mpad * dp_, 256)), dim3(256), 0, stream_, xd, d_, d_, dp_,
                     static_cast<const int32_t*>(nullptr), 0, m, mpad, mu_, inv_sigma_, ld, u_, ut_, lab_);
  AKNN_HIP_TRY(hipGetLastError());
  forward(mpad, stream_, true);
  int64_t real = 0;
  for (int64_t i = 0; i < m; ++i) real += labels[i] >= 0;
  backward(mpad, m, real, stream_);
  double acc[2] = {0, 0};
  AKNN_HIP_TRY(hipMemcpyAsync(acc, acc_, sizeof(acc), hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipMemcpyAsync(grads, g_, sizeof(float) * param_count(), hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  return acc[0] / static_cast<double>(std::max<int64_t>(real, 1));
}

void Router::logits(const float* x_raw, int64_t m, float* z) {
  if (m < 1 || m > (1 << 20)) throw std::invalid_argument("router: batch must hold 1..2^20 rows");
  AKNN_HIP_TRY(hipSetDevice(device_));
  const int64_t mpad = (m + 63) / 64 * 64;
  ensure_batch(mpad);
  float* xd = nullptr;
  AKNN_HIP_TRY(hipMalloc(&xd, sizeof(float) * m * d_));
  struct Free {
    void* a;
    ~Free() { (void)hipFree(a); }
  } fr{xd};
  AKNN_HIP_TRY(hipMemcpy(xd, x_raw, sizeof(float) * m * d_, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(gather_kernel, dim3(grid_of(mpad * dp_, 256)), dim3(256), 0, stream_, xd, d_, d_, dp_,
                     static_cast<const int32_t*>(nullptr), 0, m, mpad, mu_, inv_sigma_,
                     static_cast<const int32_t*>(nullptr), u_, ut_, static_cast<int32_t*>(nullptr));
  AKNN_HIP_TRY(hipGetLastError());
  forward(mpad, stream_, false);
  AKNN_HIP_TRY(hipMemcpyAsync(z, z_, sizeof(float) * m * nb_, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
}

void Router::assign(const float* x_raw, int64_t n, int ld, int32_t* bins, hipStream_t s) {
  AKNN_HIP_TRY(hipSetDevice(device_));
  if (fused_ && h_ == kFuseH && dp_ <= kFuseH && (nb_ == 256 || nb_ == 128 || nb_ == 64 || nb_ == 32)) {
    AKNN_HIP_TRY(hipEventRecord(ev_[0], s));
    const unsigned grid = grid_of(n, kFuseRows);
    if (nb_ == 256)
      hipLaunchKernelGGL(router_fused_kernel<2>, dim3(grid), dim3(256), 0, s, x_raw, ld, d_, dp_, n, mu_, inv_sigma_,
                         w1p_, p_ + off_[1], w2p_, p_ + off_[3], w3p_, p_ + off_[5], nb_, bins);
    else
      hipLaunchKernelGGL(router_fused_kernel<1>, dim3(grid), dim3(256), 0, s, x_raw, ld, d_, dp_, n, mu_, inv_sigma_,
                         w1p_, p_ + off_[1], w2p_, p_ + off_[3], w3p_, p_ + off_[5], nb_, bins);
    AKNN_HIP_TRY(hipGetLastError());
    AKNN_HIP_TRY(hipEventRecord(ev_[1], s));
    assigned_ = true;
    return;
  }
  const int64_t chunk = std::min<int64_t>(kAssignChunk, (n + 63) / 64 * 64);
  ensure_batch(chunk);
  AKNN_HIP_TRY(hipEventRecord(ev_[0], s));
  for (int64_t base = 0; base < n; base += chunk) {
    const int64_t m = std::min<int64_t>(chunk, n - base);
    const int64_t mpad = (m + 63) / 64 * 64;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_of(mpad * dp_, 256)), dim3(256), 0, s, x_raw, ld, d_, dp_,
                       static_cast<const int32_t*>(nullptr), base, m, mpad, mu_, inv_sigma_,
                       static_cast<const int32_t*>(nullptr), u_, static_cast<uint16_t*>(nullptr),
                       static_cast<int32_t*>(nullptr));
    AKNN_HIP_TRY(hipGetLastError());
    forward(mpad, s, false);
    hipLaunchKernelGGL(argmax_kernel, dim3(grid_of(m, 4)), dim3(256), 0, s, z_, nb_, m, bins + base);
    AKNN_HIP_TRY(hipGetLastError());
  }
  AKNN_HIP_TRY(hipEventRecord(ev_[1], s));
  assigned_ = true;
}

void Router::bins_of(const float* x_raw, int64_t m, int32_t* bins) {
  if (m < 1 || m > (1 << 24)) throw std::invalid_argument("router: 1..2^24 rows");
  AKNN_HIP_TRY(hipSetDevice(device_));
  float* xd = nullptr;
  int32_t* bd = nullptr;
  AKNN_HIP_TRY(hipMalloc(&xd, sizeof(float) * m * d_));
  if (hipMalloc(&bd, sizeof(int32_t) * m) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(xd);
    throw HipError("router: out of device memory");
  }
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } fr{xd, bd};
  AKNN_HIP_TRY(hipMemcpy(xd, x_raw, sizeof(float) * m * d_, hipMemcpyHostToDevice));
  assign(xd, m, d_, bd, stream_);
  AKNN_HIP_TRY(hipMemcpyAsync(bins, bd, sizeof(int32_t) * m, hipMemcpyDeviceToHost, stream_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
}

double Router::last_assign_ms() const {
  if (!assigned_) return 0.0;
  float ms = 0.0f;
  if (hipEventSynchronize(ev_[1]) != hipSuccess || hipEventElapsedTime(&ms, ev_[0], ev_[1]) != hipSuccess) {
    (void)hipGetLastError();  // leave no error behind for the next launch check
    return 0.0;
  }
  return ms;
}

void Router::get_params(float* p, float* mu, float* inv_sigma) const {
  AKNN_HIP_TRY(hipSetDevice(device_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  if (p) AKNN_HIP_TRY(hipMemcpy(p, p_, sizeof(float) * param_count(), hipMemcpyDeviceToHost));
  if (mu) std::copy(mu_host_.begin(), mu_host_.begin() + d_, mu);
  if (inv_sigma) *inv_sigma = inv_sigma_;
}

void Router::set_params(const float* p, const float* mu, float inv_sigma) {
  AKNN_HIP_TRY(hipSetDevice(device_));
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
  if (p) {
    std::vector<float> host(p, p + param_count());
    for (int o = 0; o < h_; ++o)
      for (int c = d_; c < dp_; ++c) host[off_[0] + static_cast<size_t>(o) * dp_ + c] = 0.0f;
    AKNN_HIP_TRY(hipMemcpy(p_, host.data(), sizeof(float) * param_count(), hipMemcpyHostToDevice));
  }
  if (mu) {
    mu_host_.assign(dp_, 0.0f);
    std::copy(mu, mu + d_, mu_host_.begin());
    AKNN_HIP_TRY(hipMemcpy(mu_, mu_host_.data(), sizeof(float) * dp_, hipMemcpyHostToDevice));
  }
  if (!(inv_sigma > 0.0f) || !std::isfinite(inv_sigma)) throw std::invalid_argument("inv_sigma must be positive");
  inv_sigma_ = inv_sigma;
  refresh_weights(stream_);
  AKNN_HIP_TRY(hipStreamSynchronize(stream_));
}

namespace {
constexpr char kMagic[8] = {'A', 'K', 'N', 'N', 'R', 'T', 'R', '1'};
}

void Router::save(const std::string& path) const {
  std::vector<float> p(param_count());
  get_params(p.data(), nullptr, nullptr);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const int32_t hdr[3] = {d_, h_, bin_bits_};
  f.write(kMagic, 8);
  f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  f.write(reinterpret_cast<const char*>(mu_host_.data()), sizeof(float) * d_);
  f.write(reinterpret_cast<const char*>(&inv_sigma_), sizeof(float));
  f.write(reinterpret_cast<const char*>(p.data()), sizeof(float) * p.size());
  if (!f) throw std::runtime_error("write failed: " + path);
}

void Router::load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  char magic[8];
  int32_t hdr[3];
  f.read(magic, 8);
  f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
  if (!f || std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error(path + " is not a router file");
  if (hdr[0] != d_ || hdr[1] != h_ || hdr[2] != bin_bits_)
    throw std::invalid_argument("router file shape differs from this router");
  std::vector<float> mu(d_), p(param_count());
  float inv = 1.0f;
  f.read(reinterpret_cast<char*>(mu.data()), sizeof(float) * d_);
  f.read(reinterpret_cast<char*>(&inv), sizeof(float));
  f.read(reinterpret_cast<char*>(p.data()), sizeof(float) * p.size());
  if (!f) throw std::runtime_error("truncated router file " + path);
  set_params(p.data(), mu.data(), inv);
}

}  // namespace aknn
