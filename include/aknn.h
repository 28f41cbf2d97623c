/*
 * aknn: approximate k-NN graph builder for AMD Instinct MI355X (gfx950).
 *
 * C ABI of libaknn.so.  It replaces the public interface of the reference's
 * class EuclideanDistanceMatrixGPU (gosteq/Approximate-CUDA-kNN,
 * euclideandistancematrixgpu.h:50-63) one call per method, so a binding
 * (ctypes, cgo, JNI, a C++ wrapper) can drive the same lifecycle:
 *
 *   create -> set_data_file -> set_results_file -> initialize -> calculate
 *          -> save_result -> deinitialize -> destroy
 *
 * Conventions: every call returns AKNN_OK (0) or a nonzero status and never
 * throws; aknn_last_error() gives the message.  The handle owns all device
 * memory; arrays passed in or out are owned by the caller.  One handle per
 * host thread.  Arrays are plain host pointers, row-major.
 */
#ifndef AKNN_H
#define AKNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* euclideandistancematrixgpu.h:16-18 */
#define AKNN_DISTANCE_EUCLIDEAN 0
#define AKNN_DISTANCE_TAXICAB 1
#define AKNN_DISTANCE_COSINE 2

#define AKNN_OK 0
#define AKNN_ERR_ARGUMENT 1
#define AKNN_ERR_IO 2
#define AKNN_ERR_DEVICE 3
#define AKNN_ERR_STATE 4

typedef struct aknn_handle aknn_handle;

/* EuclideanDistanceMatrixGPU(bool debugMode), euclideandistancematrixgpu.cu:2696-2702.
 * debug_mode != 0 makes aknn_save_result also write <results>DEBUG. */
aknn_handle* aknn_create(int debug_mode);
/* ~EuclideanDistanceMatrixGPU(), euclideandistancematrixgpu.cu:2704-2706 */
void aknn_destroy(aknn_handle* h);

/* setDataFile, euclideandistancematrixgpu.cu:2708-2710 */
int aknn_set_data_file(aknn_handle* h, const char* path);
/* setResultsFile, euclideandistancematrixgpu.cu:2898-2900 */
int aknn_set_results_file(aknn_handle* h, const char* path);

/* initialize(N, D, k, device, typeOfDistance, klaster, trees, propagations,
 * minSize), euclideandistancematrixgpu.cu:2759-2819: selects the GPU, reads
 * the CSV data file (two header lines, then N lines of D comma-separated
 * values, plus a trailing integer label when klaster != 0, .cu:2712-2757) and
 * uploads it.  A data file named *.fbin is read as binary instead (uint32 N,
 * uint32 D, then N*D float32, row-major; no reference counterpart).  Rejects inputs the reference cannot run correctly: N < k+1,
 * trees < 1, missing or short rows. */
int aknn_initialize(aknn_handle* h, int n, int d, int k, int device, int distance, int klaster, int trees,
                    int propagations, int min_size);

/* loadData (euclideandistancematrixgpu.cu:2712-2757) on its own, without a
 * device: reads n points of d values from the CSV (or *.fbin) file at path
 * into x[n*d] (row-major) and, with klaster, the labels into labels[n].
 * threads: CSV parsing threads (0: up to 16, 1: sequential; the values are
 * the same).  On failure returns AKNN_ERR_IO or AKNN_ERR_ARGUMENT with the
 * message in error[error_len] (may be NULL). */
int aknn_read_points(const char* path, int64_t n, int d, int klaster, int threads, float* x, int32_t* labels,
                     char* error, int error_len);

/* Same as aknn_initialize with the points given in memory instead of a file
 * (x: n*d floats, row-major, point-major).  No reference counterpart: it is
 * initialize() with loadData() (.cu:2712-2757) replaced by a copy. */
int aknn_initialize_host(aknn_handle* h, const float* x, int64_t n, int d, int k, int device, int distance,
                         int trees, int propagations, int min_size);

/* calculate(), euclideandistancematrixgpu.cu:2861-2896: normalize, forest,
 * initial k-NN, propagation rounds, all on the device.  Re-runnable: every call
 * starts again from the uploaded points. */
int aknn_calculate(aknn_handle* h);

/* saveResultToResultFile(), euclideandistancematrixgpu.cu:2908-2978: writes
 * "N;k;8\n", the 8-byte marker 0x01020304, then per point k {int32 id,
 * float32 value} records sorted by value (COSINE: descending), ties by lower
 * id; <results>VALIDATION; <results>DEBUG when created with debug_mode. */
int aknn_save_result(aknn_handle* h);

/* deinitialize(), euclideandistancematrixgpu.cu:2821-2859 */
int aknn_deinitialize(aknn_handle* h);

/* Options, set before aknn_calculate (no reference counterpart; defaults
 * reproduce the reference):
 *   "seed"          rand() stream seed (default 1: the reference never seeds)
 *   "rng_order"     0: each level's Leaf7(rand(), rand()) receives (second,
 *                   first) draw, as g++ evaluates arguments right to left; 1:
 *                   (first, second)
 *   "cosine_fixed"  0: COSINE stores 1 - key on every selection and re-keys
 *                   carried entries by it (.cu:1007, :806); 1: carry distances
 *   "level_batch"   forest levels enqueued per host check (default 16)
 *   "skip_settled"  1 (default): a propagation round does not re-examine the
 *                   neighbours of a neighbour whose row it already examined
 *                   in the previous round unchanged (EUCLIDEAN, TAXICAB and
 *                   COSINE with carried distances; the results are the same
 *                   as without the skip); 0: examine all
 *   "sketch"        1: an FP16 copy of the points with proven error bounds
 *                   decides most forest side tests (the exact chain runs where
 *                   the bound leaves the sign open), and propagation screens
 *                   each candidate with a lower bound of its exact FP32 key,
 *                   computing the exact key only where the bound does not
 *                   exclude it (the results are the same); 0 (default): exact
 *                   chains everywhere
 *   "sketch_int8"   1: with "sketch", EUCLIDEAN propagation screens with an
 *                   int8 copy (per-point scale, exact integer dot products,
 *                   half the bytes) instead of the FP16 one; 0 (default): FP16
 *                   (its tighter bound re-scores fewer candidates)
 *   "flip_int8"     1 (default): COSINE parity rounds with "sketch" screen their
 *                   candidates with the int8 sketch where it is made (the sign
 *                   forest's, "forest_int8"): an exact integer dot product of
 *                   half the bytes of the FP16 row (the graph is the same); 0:
 *                   the FP16 row
 *   "forest_int8"   1 (default): with "sketch", the forest side test of the
 *                   buckets a block stages (D <= 512) reads each point's int8
 *                   row (per-point scale and radius, a proven bound; the exact
 *                   chain decides the rest) instead of its FP16 row, half the
 *                   bytes (the trees are the same); 0: the FP16 row
 *   "leaf_mfma"     1: EUCLIDEAN leaf k-NN of every tree after the first
 *                   screens its pairs with FP16 dot products on the matrix
 *                   cores and proven lower bounds, computing exact FP32 keys
 *                   only for pairs the bound does not exclude (the results are
 *                   the same); 0 (default): exact keys for every pair
 *   "split"         forest rule.  0 (default): the reference's sign of
 *                   x . (x_p2 - x_p1) with forced pivots (.cu:2138-2181).
 *                   1: median split, a rule of this library that replaces the
 *                   reference's: tree t draws seeded Gaussian directions (seed
 *                   option, tree index), each level projects its buckets on
 *                   one direction and splits each bucket at a sampled median
 *                   (balanced leaves of leaf/2..leaf points, no rand() use)
 *   "taxicab_l1"    1: TAXICAB keys and stores the L1 total instead of its
 *                   square root (the reference stores sqrtf, .cu:568, :1477)
 *   "scale_all"     1: normalization divides every dimension by the max-abs
 *                   over all dimensions (the reference: the first 256 only,
 *                   .cu:57-65, :76-84)
 *   "fixed"         sets cosine_fixed, taxicab_l1 and scale_all together
 *   "mean_fp64"     1: normalization's per-dimension mean is summed in FP64
 *                   over blocks of 2048 points in point order, the block sums
 *                   added in block order and the mean rounded once to FP32
 *                   (the reference: one sequential FP32 sum over all points,
 *                   .cu:29-40, a chain of N dependent adds); 0 (default)
 *   "pca_screen"    1: EUCLIDEAN propagation screens each candidate with the
 *                   distance of its projection on the data's top 16 or 32
 *                   principal directions (recomputed per build from a sample;
 *                   used when they hold 97% of the second moment, D in
 *                   [16, 256]), a proven lower bound of the exact key; the
 *                   exact chain runs where the bound does not exclude a
 *                   candidate (the results are the same); 0 (default): off
 *   "pca_fp16"      1 (default): the screen's projections are stored in FP16
 *                   (32 or 64 bytes read per candidate) and each point's
 *                   radius includes its projection's FP16 rounding, so the
 *                   bound stays proven; 0: FP32 projections (the results are
 *                   the same)
 *   "io_threads"    host threads parsing the CSV in aknn_initialize (0, the
 *                   default: up to 16; 1: the sequential loader); the values
 *                   are the same for any count
 *   "cosine_table"  1 (default): the COSINE parity initial pass computes each
 *                   tree's keys (point, start-range member) once into a
 *                   table, when it and its per-workgroup arrays fit in 70% of
 *                   the free device memory (4 bytes per range member), and
 *                   reads them in id order; 0: each point computes the keys
 *                   of its candidates from their rows (the results are the
 *                   same)
 *   "graph_rounds"  1 (default): once the propagation round mode is fixed (the
 *                   first calculate() probes it, or nothing is screened), the
 *                   rounds are captured into one hipGraph and replayed by later
 *                   calculate() calls (same results); 0: launch round by round
 *   "leaf_batch"    initial leaf k-NN launches (same results in every form).
 *                   1: every tree in one persistent cooperative launch (trees
 *                   in order with a grid barrier between them, each tree's
 *                   work items from per-XCD queues, the screen switched per
 *                   tree on the device); 0: one launch per tree (also taken
 *                   when the runtime refuses the cooperative grid); 2
 *                   (default): the faster of the two on this data set, timed
 *                   on its first calculate() (both run there, the second from
 *                   stub rows again) and kept until the data or an option
 *                   changes
 *   "router_trees"  with split 1: trees 0..router_trees-1 take the router bins
 *                   (aknn_set_bins or an attached router) as their first
 *                   levels, one bit per level, before the median levels */
int aknn_set_option(aknn_handle* h, const char* name, int64_t value);

/* The graph as the result file holds it before sorting: ids[n*k], vals[n*k]
 * (the device-to-host copy of .cu:2911-2915; rows sorted by (value, id)). */
int aknn_get_graph(aknn_handle* h, int32_t* ids, float* vals);

/* Points after normalization (.cu:19-85), n*d floats, row-major. */
int aknn_get_normalized(aknn_handle* h, float* x);

/* Tree `tree` of the last calculate(): perm[n] is the leaf array's point order
 * (TreeNodeLeaf::entityNumber, .cu:2391-2405); range_start[p], range_count[p]
 * give, for point p, the positions in perm of its start node
 * (findGraphTraversalStartPoint, .cu:87-105, and the leftmost-leaf descent,
 * .cu:166-176).  iterations (may be NULL) receives the number of level-loop
 * iterations (.cu:2509-2637), i.e. rand() pairs the tree drew after its root. */
int aknn_get_forest(aknn_handle* h, int tree, int32_t* perm, int32_t* range_start, int32_t* range_count,
                    int32_t* iterations);

/* Times of the last calculate() in ms: [0] normalize, [1] forest, [2] initial
 * k-NN, [3] propagation, [4] total (the reference's event bracket,
 * .cu:2867-2893).  n: capacity of ms. */
int aknn_get_stage_times(aknn_handle* h, double* ms, int n);

/* Kernel statistics of the last calculate(): name is "leaf_knn",
 * "propagate", "forest", "leaf_mfma" (the screened trees of the leaf k-NN
 * alone), "leaf_exact" (the leaf k-NN trees that ran the exact chain on every
 * pair) or "propagate_graph" (the rounds replayed from the captured graph;
 * zero launches when they were launched one by one).  launches (for leaf_mfma
 * and leaf_exact: trees, each a launch or a segment of the all-trees launch),
 * summed time (HIP events on the launch stream; a segment's from the device
 * clock at the launch's grid barriers), and algorithmic units summed over
 * them (flops for leaf_knn and leaf_exact: 3D per pair; GEMM-form flops for
 * leaf_mfma: 2 * padded D per pair; bytes for propagate and forest). */
int aknn_get_kernel_stats(aknn_handle* h, const char* name, int64_t* launches, double* total_ms, double* units);

/* The leaf k-NN launch form of the last calculate(): *form 1 when every tree
 * ran in one persistent launch, 0 one launch per tree; per_tree_ms and
 * all_trees_ms (nullable) receive the two forms' times measured by the call
 * that chose it (option leaf_batch 2; 0 before any choice).  No counterpart
 * in the reference (its initial pass is one kernel over all trees,
 * euclideandistancematrixgpu.cu:2067-2108): an accessor of this library's
 * launch-form tuning, for callers measuring it. */
int aknn_get_leaf_form(aknn_handle* h, int32_t* form, double* per_tree_ms, double* all_trees_ms);

/* Candidate counts of the last calculate().  stage "propagate": new
 * candidates examined over all rounds, and those whose exact FP32 distance was
 * computed (all of them without "sketch").  stage "leaf_knn": pairs screened
 * on the matrix cores (trees after the first, with "leaf_mfma"), and those
 * whose exact distance was computed.  stage "forest": side tests of active
 * points over all levels, and those the sketch left to the exact chain (0
 * without "sketch").  stage "pca": the principal-subspace screen's width
 * (16, 32, or 0 when unused) and the share of the sample's second moment its
 * top 16 directions hold.  Stage "cosine_table": 1 when the COSINE parity
 * initial pass read its keys from the key table (option cosine_table), else 0.
 * No reference counterpart: the work accounting of the kernel statistics
 * above. */
int aknn_get_candidate_counts(aknn_handle* h, const char* stage, double* examined, double* rescored);

/* Exact k-NN in FP64 of nq query points (indices into the data) against all
 * points, on the normalized points of the last calculate(), with the handle's
 * distance: EUCLIDEAN the squared L2 distance, TAXICAB L1, COSINE 1 - cos;
 * dimensions summed in order in FP64.  The query itself is excluded.  ids and
 * dist are [nq][kq], sorted by (distance, id); 1 <= kq <= 64.  No reference
 * counterpart: the measurement reference for recall. */
int aknn_exact_knn(aknn_handle* h, const int64_t* queries, int64_t nq, int kq, int32_t* ids, double* dist);

/* recall@k of the current graph over nq query points: the fraction of the
 * queries' graph neighbours whose FP64 distance (as aknn_exact_knn) is at most
 * the query's k-th exact distance (ties at the k-th distance count as hits).
 * No reference counterpart. */
int aknn_recall(aknn_handle* h, const int64_t* queries, int64_t nq, double* recall);

/* aknn_recall without device allocation: the same result, bit for bit, from
 * the same kernels, whose buffers (query indices, exact ids and distances
 * [nq][k], k-th distances, hit counter, and for COSINE the FP64 squared norms
 * of all n points) come from one static device array of 48 MiB per GPU that
 * the library holds until the process ends; the call makes no hipMalloc or
 * hipFree.  They need 8 + 12k + 8 bytes per query (rounded up per buffer to
 * 256 bytes), plus 8n for COSINE; a call needing more than the array returns
 * AKNN_ERR_ARGUMENT.  The array is shared by every call on a GPU and keeps
 * the last call's contents: calls that overlap on one GPU (several host
 * threads or handles on the same device) race and must be serialised.
 * No reference counterpart: for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc). */
int aknn_recall_noalloc(aknn_handle* h, const int64_t* queries, int64_t nq, double* recall);

/* Router bins of the points for the median-split forest's routed trees:
 * bins[n] with values in [0, 2^bin_bits), 1 <= bin_bits <= 20, n the number
 * of points.  The first bin_bits levels of a routed tree split a bucket by
 * the bits of its points' bins, most significant first.  No reference
 * counterpart: the reference's trees split only on pivot differences
 * (.cu:2138-2164); this feeds learned leaf assignments to the same forest. */
int aknn_set_bins(aknn_handle* h, const int32_t* bins, int64_t n, int bin_bits);

/* Message of the last failed call on h (empty if none). */
const char* aknn_last_error(aknn_handle* h);

/* ---- Learned router of the median-split forest ------------------------------
 * No reference counterpart: the reference draws split directions from random
 * pivot pairs (.cu:2138-2164, :2291-2365).  The router is an MLP
 *   u = (x - mu) * inv_sigma, h1 = relu(W1 u + b1), h2 = relu(W2 h1 + b2),
 *   logits = W3 h2 + b3 (2^bin_bits of them), bin = argmax (lowest on ties)
 * over the raw points, with BF16 weights and activations, FP32 accumulation,
 * trained (Adam, FP32 master weights) on the device by the library's own
 * forward and backward kernels.  Its labels are a balanced partition of a
 * k-NN graph the library built on a training set, so its bins keep graph
 * neighbours together.  Attached to a handle with option split = 1 and
 * router_trees > 0, it assigns each point's bin inside aknn_calculate and the
 * routed trees split on the bin bits before their median levels. */
typedef struct aknn_router aknn_router;

/* d: point dimension; hidden: width of both hidden layers (multiple of 32);
 * bin_bits in [5, 12]; seed: initial weights and batch order.  NULL on error. */
aknn_router* aknn_router_create(int device, int d, int hidden, int bin_bits, uint32_t seed);
void aknn_router_destroy(aknn_router* r);

/* Balanced graph labels of the points of `trained` (a handle after
 * aknn_calculate): equal runs of the leaf order of tree (router seed mod
 * trees), then `sweeps` sweeps of
 * capacity-bounded label propagation over the symmetrised graph (parts within
 * 0.95..1.05 n/bins).  labels[n]; cut (may be NULL) receives the fraction of
 * graph edges cut before and after the sweeps. */
int aknn_router_labels(aknn_router* r, aknn_handle* trained, int sweeps, int32_t* labels, double* cut);

/* Trains on the raw points of `trained` with labels[n] (NULL: the graph
 * labels of aknn_router_labels with `sweeps`): standardisation from those
 * points, then `epochs` of Adam (learning rate lr) over seeded shuffled batches
 * of `batch` points minimising the softmax cross-entropy. */
int aknn_router_fit(aknn_router* r, aknn_handle* trained, const int32_t* labels, int sweeps, int epochs, int batch,
                    float lr);

/* One forward and backward pass over m raw points x[m][d] with labels[m]
 * (label < 0: row ignored): gradients of the mean cross-entropy in the flat
 * parameter layout W1 [hidden][dp], b1, W2 [hidden][hidden], b2,
 * W3 [2^bin_bits][hidden], b3 (dp = d rounded up to 16, padding columns
 * zero), and the loss. */
int aknn_router_gradients(aknn_router* r, const float* x, int64_t m, const int32_t* labels, float* grads,
                          double* loss);
/* FP32 logits z[m][2^bin_bits] of m raw points. */
int aknn_router_logits(aknn_router* r, const float* x, int64_t m, float* z);
/* Bins of m raw points x[m][d], as aknn_calculate assigns them: fused 1 runs
 * the one-kernel inference (hidden 256, d <= 256, up to 256 bins; other
 * shapes fall back), 0 the layer-by-layer products; both give the same bins. */
int aknn_router_bins(aknn_router* r, const float* x, int64_t m, int fused, int32_t* bins);
int aknn_router_param_count(aknn_router* r, int64_t* count);
/* Parameters (flat layout above), mu[d] and inv_sigma; any pointer may be NULL. */
int aknn_router_get_params(aknn_router* r, float* params, float* mu, float* inv_sigma);
/* params or mu may be NULL (kept); inv_sigma must be positive. */
int aknn_router_set_params(aknn_router* r, const float* params, const float* mu, float inv_sigma);
int aknn_router_save(aknn_router* r, const char* path);
int aknn_router_load(aknn_router* r, const char* path);
/* out[0..6]: epochs trained, last epoch's loss and accuracy, training ms, graph
 * edges cut by the labels before and after refinement, ms of the last bin
 * assignment. */
int aknn_router_stats(aknn_router* r, double* out, int n);
const char* aknn_router_last_error(aknn_router* r);

/* Attaches a router to a handle (NULL detaches every router); detach before
 * destroying the router.  It must have the handle's D and device. */
int aknn_set_router(aknn_handle* h, aknn_router* r);
/* Appends another router (same bin_bits): routed tree t takes router t mod R,
 * so routers trained on different labels give routed trees different tops. */
int aknn_add_router(aknn_handle* h, aknn_router* r);

#ifdef __cplusplus
}
#endif

#endif /* AKNN_H */
