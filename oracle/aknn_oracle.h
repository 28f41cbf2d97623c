/*
 * CPU oracle for the approximate k-NN graph pipeline of gosteq/Approximate-CUDA-kNN.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test-suite
 * (tests/), the smoke check in __graft_entry__.py and the cpu_baseline leg of
 * bench.py may load it.  The product (libaknn.so, the aknn CLI) never links,
 * loads or calls it, and must fail loudly instead of falling back to it.
 *
 * It restates, in plain C, one legal execution of the reference's CUDA code
 * (the reference is nondeterministic through atomics; every free choice is
 * fixed here as "stable order" and "ties broken by (key, lower id)"):
 *
 *   normalize            euclideandistancematrixgpu.cu:19-85
 *   glibc rand() stream  used at euclideandistancematrixgpu.cu:2483-2484, :2589
 *   forest (one tree)    euclideandistancematrixgpu.cu:2130-2405, :2465-2655
 *   start ranges         euclideandistancematrixgpu.cu:87-105, :166-176
 *   initial k-NN         euclideandistancematrixgpu.cu:107-1031
 *   propagation round    euclideandistancematrixgpu.cu:1033-1977
 *   distances            euclideandistancematrixgpu.cu:243-273, :540-570, :840-881
 *   result file          euclideandistancematrixgpu.cu:2908-2978
 *
 * Build flags matter: -ffp-contract=off and explicit fmaf() reproduce the
 * reference's contracted FMA chains bit for bit; no -ffast-math (denormals kept,
 * IEEE sqrt and division).
 */
#ifndef AKNN_ORACLE_H
#define AKNN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EUCLIDEAN 0
#define ORC_TAXICAB 1
#define ORC_COSINE 2
/* fixed mode of TAXICAB: keys and stored values are the L1 total itself (the
 * reference stores its sqrtf, .cu:568, :1477) */
#define ORC_TAXICAB_L1 3

#define ORC_STUB_ID (-1)
#define ORC_STUB_DIST 100000.0f

/* glibc random_r TYPE_3 (x_i = x_{i-3} + x_{i-31}, output >> 1). */
typedef struct {
    int32_t state[31];
    int f;
    int r;
} orc_rng;

void orc_rng_seed(orc_rng* g, uint32_t seed);
int32_t orc_rng_next(orc_rng* g);
/* int(((double)r / RAND_MAX) * INT_MAX) % count, as at .cu:2482-2484 */
int32_t orc_pick(int32_t r, int32_t count);

/* A.2: in place on a row-major [n][d] array. */
void orc_normalize(float* x, int64_t n, int d);
/* scale_all 1 (fixed mode): the max-abs and the division cover every
 * dimension instead of the first 256. */
void orc_normalize_ex(float* x, int64_t n, int d, int scale_all);
/* orc_normalize_ex with the mean taken (mean_fp64 != 0) as the library's
 * option mean_fp64 takes it: FP64 sums over blocks of 2048 points in point
 * order, block sums added in block order, the mean rounded once to FP32 (not
 * the reference's arithmetic). */
void orc_normalize_mode(float* x, int64_t n, int d, int scale_all, int mean_fp64);

/* A.7: the FP32 key of a pair (EUCLIDEAN sqrt(L2^2), TAXICAB sqrt(L1),
 * COSINE float(1 - (double)cos)). */
float orc_distance(const float* a, const float* b, int d, int metric);

/* A.4 + A.5: builds one tree, advancing the shared rand() stream g.
 * rng_order 0: Leaf7's (rand1, rand2) = (second draw, first draw) (g++ right-to-left
 * argument evaluation); 1: (first, second).
 * perm[n]: leaf array entity order.  range_start/range_count[n] (indexed by
 * point id): positions in perm of the start node of that point.
 * Returns 0, or -1 on bad arguments. */
int orc_build_tree(const float* x, int64_t n, int d, int leaf_max, int k,
                   orc_rng* g, int rng_order,
                   int32_t* perm, int32_t* range_start, int32_t* range_count,
                   int32_t* n_iterations, int64_t* n_nodes, int64_t* active_elems);

/* Median-split forest (a rule of this rebuild, not the reference's).
 * orc_directions: `count` seeded Gaussian directions of length d for tree
 * `tree` (splitmix64 + Box-Muller in double, rounded to float), level-major.
 * orc_median_depth: directions per tree, ceil(log2(n / leaf)) + 8.
 * orc_build_tree_median: bins (nullable) routes the first bin_bits levels by
 * the bits of each point's bin, most significant first; the later levels
 * split at a sampled median of x . r[j]; see aknn_oracle.c for the rule. */
void orc_directions(uint32_t seed, int tree, int count, int d, float* r);
int orc_median_depth(int64_t n, int leaf);
int orc_build_tree_median(const float* x, int64_t n, int d, int leaf_max, int k,
                          const float* dirs, int ndirs, const int32_t* bins, int bin_bits,
                          int32_t* perm, int32_t* range_start, int32_t* range_count,
                          int32_t* n_levels);

/* A.6.1.  perm/rs/rc are [trees][n].  Output rows [n][k] sorted by
 * (stored value, id).  cosine_fixed=1 carries distances instead of the
 * reference's per-store 1-key flip. */
int orc_initial_knn(const float* x, int64_t n, int d, int k, int trees, int metric,
                    const int32_t* perm, const int32_t* rs, const int32_t* rc,
                    int32_t* ids, float* vals, int cosine_fixed);

/* A.6.2: one Jacobi propagation round. */
int orc_propagate(const float* x, int64_t n, int d, int k, int metric,
                  const int32_t* ids_in, const float* vals_in,
                  int32_t* ids_out, float* vals_out, int cosine_fixed);

/* The whole calculate() bracket: normalize x in place, build the forest,
 * initial k-NN, props rounds.  stage_ms[4] (may be NULL): normalize, forest,
 * initial, propagation wall times. */
int orc_run(float* x, int64_t n, int d, int k, int trees, int leaf, int props, int metric,
            uint32_t seed, int rng_order, int cosine_fixed,
            int32_t* ids, float* vals, double* stage_ms);

/* orc_run with the forest rule: split 0 (reference sign split) or 1
 * (median split; trees t < router_trees take bins[n] as their first levels),
 * and scale_all for the normalization; metric may be ORC_TAXICAB_L1. */
int orc_run_ex(float* x, int64_t n, int d, int k, int trees, int leaf, int props, int metric,
               uint32_t seed, int rng_order, int cosine_fixed, int split,
               const int32_t* bins, int bin_bits, int router_trees, int scale_all,
               int32_t* ids, float* vals, double* stage_ms);

/* Router MLP forward (and backward when labels and grads are given): returns
 * the mean softmax cross-entropy over rows with a label in [0, nb), writes
 * grads (flat parameter layout) and logits[m][nb] (nullable).  round_bf16
 * rounds where the device rounds to BF16; 0 computes in FP64 throughout. */
double orc_mlp(const float* x, int64_t m, int d, int dp, int h, int nb,
               const float* params, const float* mu, float inv_sigma,
               const int32_t* labels, int round_bf16, float* grads, float* logits);

/* FP64 exact k-th neighbour distance (self excluded) for each query. */
void orc_exact_kth(const float* x, int64_t n, int d, int k, int metric,
                   const int64_t* queries, int64_t nq, double* kth);
/* FP64 distance used by the recall check. */
double orc_distance64(const float* a, const float* b, int d, int metric);
/* recall@k: hits counted when d64(q, returned) <= kth(q). */
double orc_recall(const float* x, int64_t n, int d, int k, int metric,
                  const int64_t* queries, int64_t nq, const double* kth,
                  const int32_t* ids);

/* A.9: the reference's binary result file. */
int orc_write_result(const char* path, int64_t n, int k, int metric,
                     const int32_t* ids, const float* vals);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
