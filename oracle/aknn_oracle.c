/*
 * CPU oracle: plain-C restatement of the reference's approximate k-NN graph
 * pipeline.  TEST INFRASTRUCTURE ONLY (see aknn_oracle.h): used by tests/, the
 * smoke check and bench.py's cpu_baseline leg, never by the product.
 *
 * Every function cites the reference lines it restates (file
 * euclideandistancematrixgpu.cu, abbreviated ".cu").  Where the reference is
 * nondeterministic (atomic partition order, atomic slot order, chunk
 * membership, boundary-bucket complement) this file fixes one legal execution:
 * stable partitions, ascending-id chunks, and selections by (key, lower id).
 */
#include "aknn_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SIDE_LEFT 0  /* .cu:9 */
#define SIDE_RIGHT 2 /* .cu:11 */

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * glibc rand(): the published random_r TYPE_3 algorithm (degree 31, separation
 * 3, seeded by the Park-Miller LCG 16807 mod 2^31-1 via Schrage's method, first
 * 310 outputs discarded).  The reference never calls srand(), so its stream is
 * seed 1 (.cu:2483-2484, :2589).
 * ------------------------------------------------------------------------- */
void orc_rng_seed(orc_rng* g, uint32_t seed) {
    int32_t word;
    int i;
    if (seed == 0) seed = 1;
    g->state[0] = (int32_t)seed;
    word = (int32_t)seed;
    for (i = 1; i < 31; ++i) {
        long hi = word / 127773;
        long lo = word % 127773;
        long w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        word = (int32_t)w;
        g->state[i] = word;
    }
    g->f = 3;
    g->r = 0;
    for (i = 0; i < 310; ++i) (void)orc_rng_next(g);
}

int32_t orc_rng_next(orc_rng* g) {
    uint32_t val = (uint32_t)g->state[g->f] + (uint32_t)g->state[g->r];
    g->state[g->f] = (int32_t)val;
    if (++g->f >= 31) g->f = 0;
    if (++g->r >= 31) g->r = 0;
    return (int32_t)(val >> 1);
}

/* .cu:2482-2484 and :2322-2323: int(((double)r/(RAND_MAX))*INT_MAX) % count */
int32_t orc_pick(int32_t r, int32_t count) {
    double v = ((double)r / 2147483647.0) * 2147483647.0;
    return ((int32_t)v) % count;
}

/* ---------------------------------------------------------------------------
 * Normalization, .cu:19-42 (step 1) and .cu:44-85 (step 2).
 * Step 1: per dimension, a sequential FP32 sum over all points, divided by
 * (float)N, then subtracted.  Step 2: max |x| over dimensions < min(D, 256)
 * only, and only those dimensions are divided by it (.cu:57-65, :76-84 use
 * `if (tid < D)`, not a loop).
 * ------------------------------------------------------------------------- */
void orc_normalize(float* x, int64_t n, int d) {
    int dd;
    int dm = d < 256 ? d : 256;
    float g = 0.0f;
#pragma omp parallel for schedule(static)
    for (dd = 0; dd < d; ++dd) {
        float m = 0.0f;
        int64_t i;
        for (i = 0; i < n; ++i) m = m + x[i * d + dd];
        m = m / (float)n;
        for (i = 0; i < n; ++i) x[i * d + dd] = x[i * d + dd] - m;
    }
    for (dd = 0; dd < dm; ++dd) {
        int64_t i;
        for (i = 0; i < n; ++i) {
            float a = fabsf(x[i * d + dd]);
            g = a > g ? a : g;
        }
    }
#pragma omp parallel for schedule(static)
    for (dd = 0; dd < dm; ++dd) {
        int64_t i;
        for (i = 0; i < n; ++i) x[i * d + dd] = x[i * d + dd] / g;
    }
}

/* ---------------------------------------------------------------------------
 * Distances (.cu:243-273 EUCLIDEAN, .cu:540-570 TAXICAB, .cu:840-881 COSINE).
 * Dimensions go in chunks of 256; each chunk's partial starts at 0.0f and is
 * added into the running total (distances_device[...] += distance).  nvcc's
 * default contraction turns `p += t*t` / `ab += a*b` into FMAs.
 * ------------------------------------------------------------------------- */
float orc_distance(const float* a, const float* b, int d, int metric) {
    int c0, dd;
    if (metric == ORC_COSINE) {
        float ab = 0.0f, aa = 0.0f, bb = 0.0f, c;
        for (c0 = 0; c0 < d; c0 += 256) {
            int c1 = c0 + 256 < d ? c0 + 256 : d;
            float pab = 0.0f, paa = 0.0f, pbb = 0.0f;
            for (dd = c0; dd < c1; ++dd) {
                pab = fmaf(a[dd], b[dd], pab);
                paa = fmaf(a[dd], a[dd], paa);
                pbb = fmaf(b[dd], b[dd], pbb);
            }
            ab = ab + pab;
            aa = aa + paa;
            bb = bb + pbb;
        }
        c = ab / (sqrtf(aa) * sqrtf(bb));
        return (float)(1.0 - (double)c); /* (-1.0*distance)+1.0, .cu:879 */
    } else {
        float tot = 0.0f;
        for (c0 = 0; c0 < d; c0 += 256) {
            int c1 = c0 + 256 < d ? c0 + 256 : d;
            float p = 0.0f;
            if (metric == ORC_EUCLIDEAN) {
                for (dd = c0; dd < c1; ++dd) {
                    float t = a[dd] - b[dd];
                    p = fmaf(t, t, p);
                }
            } else {
                for (dd = c0; dd < c1; ++dd) p = p + fabsf(a[dd] - b[dd]);
            }
            tot = tot + p;
        }
        return sqrtf(tot);
    }
}

/* ---------------------------------------------------------------------------
 * One random-projection tree, restating buildUpTheTrees (.cu:2465-2655) and
 * makePartitionOfLeaf0..9 (.cu:2130-2405) with per-position arrays.  The
 * atomic scatters of Leaf4/Leaf5 become a stable partition; node ids are
 * handed out in ascending bucket-start order, left child first.
 * ------------------------------------------------------------------------- */
int orc_build_tree(const float* x, int64_t n, int d, int leaf_max, int k,
                   orc_rng* g, int rng_order,
                   int32_t* perm, int32_t* range_start, int32_t* range_count,
                   int32_t* n_iterations, int64_t* n_nodes, int64_t* active_elems) {
    int32_t *elems, *size, *bias, *q1, *q2, *par;
    int32_t *elems2, *size2, *bias2, *q12, *q22, *par2;
    uint8_t *side, *side2;
    int32_t *nparent, *nleft, *nright, *ncount;
    int32_t* base_i;
    uint8_t* base_c;
    int64_t nn = 1, p, active = 0;
    int32_t a, b, p1, p2, iters = 0;
    int ni = (int)n;

    if (n < 1 || d < 1 || leaf_max < 0 || k < 0 || n < (int64_t)k + 1 || n > 0x7fffffffLL) return -1;

    elems = (int32_t*)malloc(sizeof(int32_t) * n * 12);
    side = (uint8_t*)malloc(n * 2);
    nparent = (int32_t*)malloc(sizeof(int32_t) * 2 * n * 4);
    if (!elems || !side || !nparent) {
        free(elems); free(side); free(nparent);
        return -1;
    }
    base_i = elems;
    base_c = side;
    size = elems + n; bias = size + n; q1 = bias + n; q2 = q1 + n; par = q2 + n;
    elems2 = par + n; size2 = elems2 + n; bias2 = size2 + n; q12 = bias2 + n; q22 = q12 + n; par2 = q22 + n;
    side2 = side + n;
    nleft = nparent + 2 * n; nright = nleft + 2 * n; ncount = nright + 2 * n;

    /* root node and its two pivots (.cu:2473-2487) */
    nparent[0] = -1; nleft[0] = -1; nright[0] = -1; ncount[0] = ni;
    a = orc_rng_next(g);
    b = orc_rng_next(g);
    p1 = orc_pick(a, ni);
    p2 = orc_pick(b, ni);
    if (p1 == p2) p2 = (p1 + 1) % ni;
    for (p = 0; p < n; ++p) {
        elems[p] = (int32_t)p; size[p] = ni; bias[p] = 0;
        q1[p] = p1; q2[p] = p2; par[p] = 0; side[p] = SIDE_LEFT;
    }

    for (;;) {
        int split = 0;
        int32_t r_first, r_second, r1, r2;
        int32_t* ti;
        uint8_t* tc;
        ++iters;

        /* Leaf1 (.cu:2138-2164): side of x . (x_p2 - x_p1), sign threshold 0 */
#pragma omp parallel for schedule(static)
        for (p = 0; p < n; ++p) {
            if (size[p] > leaf_max) {
                const float* xc = x + (int64_t)elems[p] * d;
                const float* xa = x + (int64_t)elems[bias[p] + q1[p]] * d;
                const float* xb = x + (int64_t)elems[bias[p] + q2[p]] * d;
                float s = 0.0f;
                int dd;
                for (dd = 0; dd < d; ++dd) s = fmaf(xc[dd], xb[dd] - xa[dd], s);
                side[p] = s < 0 ? SIDE_LEFT : SIDE_RIGHT;
            }
        }
        for (p = 0; p < n; ++p)
            if (size[p] > leaf_max) ++active;
        /* Leaf2 (.cu:2166-2181): forced pivots, RIGHT written last */
        for (p = 0; p < n; ++p) {
            if (p == bias[p] && size[p] > leaf_max) {
                side[bias[p] + q1[p]] = SIDE_LEFT;
                side[bias[p] + q2[p]] = SIDE_RIGHT;
            }
        }
        /* Leaf3..Leaf6 (.cu:2183-2289): stable partition of every bucket */
        for (p = 0; p < n;) {
            int64_t h = p, e = p + size[p], j;
            int32_t node = par[h], nl = 0, nr, w;
            for (j = h; j < e; ++j) nl += side[j] == SIDE_LEFT;
            nr = (int32_t)(e - h) - nl;
            w = (int32_t)h;
            for (j = h; j < e; ++j)
                if (side[j] == SIDE_LEFT) {
                    elems2[w] = elems[j]; side2[w] = SIDE_LEFT; size2[w] = nl;
                    bias2[w] = (int32_t)h; par2[w] = node; ++w;
                }
            for (j = h; j < e; ++j)
                if (side[j] != SIDE_LEFT) {
                    elems2[w] = elems[j]; side2[w] = SIDE_RIGHT; size2[w] = nr;
                    bias2[w] = (int32_t)h + nl; par2[w] = node; ++w;
                }
            nleft[node] = nl; /* Leaf3 reuses the child fields as counters */
            nright[node] = nr;
            p = e;
        }
        /* Leaf7 (.cu:2291-2365): one (rand(), rand()) pair per level, shared by
         * every bucket; the pair is drawn even on the last, non-splitting level. */
        r_first = orc_rng_next(g);
        r_second = orc_rng_next(g);
        if (rng_order == 0) { r1 = r_second; r2 = r_first; }
        else { r1 = r_first; r2 = r_second; }
        for (p = 0; p < n;) {
            int64_t h = p;
            int32_t node = par2[h], nl = nleft[node], nr = nright[node];
            p = h + nl + nr;
            if (nl > 0 && nr > 0) {
                int32_t idl = (int32_t)nn++, idr = (int32_t)nn++, side_i;
                split = 1;
                nparent[idl] = node; nleft[idl] = -1; nright[idl] = -1; ncount[idl] = nl;
                nparent[idr] = node; nleft[idr] = -1; nright[idr] = -1; ncount[idr] = nr;
                nleft[node] = idl; nright[node] = idr;
                for (side_i = 0; side_i < 2; ++side_i) {
                    int64_t hh = side_i == 0 ? h : h + nl;
                    int32_t cnt = side_i == 0 ? nl : nr;
                    par2[hh] = side_i == 0 ? idl : idr;
                    if (cnt > 1) {
                        int32_t i1 = orc_pick(r1, cnt), i2 = orc_pick(r2, cnt);
                        if (i1 == i2) i2 = (i1 + 1) % cnt;
                        q12[hh] = i1; q22[hh] = i2;
                    } else {
                        q12[hh] = 0; q22[hh] = 0;
                    }
                }
            } else {
                q12[h] = 0; q22[h] = 0;
            }
        }
        /* Leaf8 (.cu:2367-2389): broadcast the head's node and pivots */
        for (p = 0; p < n; ++p) {
            int32_t hb = bias2[p];
            par2[p] = par2[hb]; q12[p] = q12[hb]; q22[p] = q22[hb];
        }
        /* double-buffer swap (.cu:2606-2632) */
        ti = elems; elems = elems2; elems2 = ti;
        ti = size; size = size2; size2 = ti;
        ti = bias; bias = bias2; bias2 = ti;
        ti = q1; q1 = q12; q12 = ti;
        ti = q2; q2 = q22; q22 = ti;
        ti = par; par = par2; par2 = ti;
        tc = side; side = side2; side2 = tc;
        if (!split) break;
    }

    /* Leaf9 (.cu:2391-2405): leaf node left = offset, right = -1 */
    for (p = 0; p < n; ++p) {
        perm[p] = elems[p];
        if (p == bias[p]) { nleft[par[p]] = (int32_t)p; nright[par[p]] = -1; }
    }
    /* findGraphTraversalStartPoint (.cu:87-105) and the leftmost-leaf descent
     * of the initial k-NN kernels (.cu:166-176) */
    for (p = 0; p < n; ++p) {
        int32_t node = par[p], cnt = ncount[node], v;
        while (cnt < k + 1) { node = nparent[node]; cnt = ncount[node]; }
        v = node;
        while (nright[v] != -1) v = nleft[v];
        range_start[elems[p]] = nleft[v];
        range_count[elems[p]] = cnt;
    }
    if (n_iterations) *n_iterations = iters;
    if (n_nodes) *n_nodes = nn;
    if (active_elems) *active_elems = active;
    free(base_i); free(base_c); free(nparent);
    return 0;
}

/* ------------------------------------------------------------------------- */
typedef struct { float key; int32_t id; } orc_ent;

static int ent_cmp(const void* pa, const void* pb) {
    const orc_ent* a = (const orc_ent*)pa;
    const orc_ent* b = (const orc_ent*)pb;
    if (a->key < b->key) return -1;
    if (a->key > b->key) return 1;
    return (a->id > b->id) - (a->id < b->id);
}

static int i32_cmp(const void* pa, const void* pb) {
    int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
    return (a > b) - (a < b);
}

/* The histogram radix select (.cu:266-410 and five copies) keeps exactly the
 * first k of the pool by (key, id) in the execution fixed here.  Stores the
 * selected entries back with the reference's store rule and sorts the row by
 * (stored, id). */
static void select_and_store(orc_ent* pool, int np, int k, int flip, int32_t* ids, float* vals) {
    int j;
    qsort(pool, np, sizeof(orc_ent), ent_cmp);
    for (j = 0; j < k; ++j)
        if (flip) pool[j].key = (float)(1.0 - (double)pool[j].key); /* .cu:1007 */
    qsort(pool, k, sizeof(orc_ent), ent_cmp);
    for (j = 0; j < k; ++j) { ids[j] = pool[j].id; vals[j] = pool[j].key; }
}

/* Initial k-NN, findInitialStateOfAproximated{Euclidean,Taxicab,Cosine}KNN
 * (.cu:107-415, :417-712, :714-1031).  cand(i) is the union of the start-node
 * ranges over all trees minus i (.cu:160-191); it is consumed in ascending-id
 * chunks of dimIdx - k candidates (.cu:207-239, dimIdx at .cu:2768), each
 * chunk competing with the k entries carried from the previous one. */
int orc_initial_knn(const float* x, int64_t n, int d, int k, int trees, int metric,
                    const int32_t* perm, const int32_t* rs, const int32_t* rc,
                    int32_t* ids, float* vals, int cosine_fixed) {
    int64_t dim_idx = (int64_t)k * k + k < n ? (int64_t)k * k + k : n;
    int64_t cap = dim_idx - k;
    int flip = metric == ORC_COSINE && !cosine_fixed;
    int64_t i;
    int err = 0;
    if (k < 1 || n < k + 1 || cap < 1) return -1;
#pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < n; ++i) {
        int64_t tot = 0, m = 0, j, c0;
        int t;
        int32_t* cand;
        orc_ent* pool;
        int32_t* lid = ids + i * k;
        float* lv = vals + i * k;
        for (t = 0; t < trees; ++t) tot += rc[(int64_t)t * n + i];
        cand = (int32_t*)malloc(sizeof(int32_t) * (tot + 1));
        pool = (orc_ent*)malloc(sizeof(orc_ent) * (k + cap));
        if (!cand || !pool) {
            free(cand); free(pool);
#pragma omp atomic write
            err = 1;
            continue;
        }
        for (t = 0; t < trees; ++t) {
            const int32_t* pt = perm + (int64_t)t * n;
            int64_t s = rs[(int64_t)t * n + i], c = rc[(int64_t)t * n + i];
            for (j = 0; j < c; ++j) cand[m++] = pt[s + j];
        }
        qsort(cand, m, sizeof(int32_t), i32_cmp);
        {
            int64_t w = 0;
            for (j = 0; j < m; ++j)
                if (cand[j] != i && (w == 0 || cand[w - 1] != cand[j])) cand[w++] = cand[j];
            m = w;
        }
        for (j = 0; j < k; ++j) { lid[j] = ORC_STUB_ID; lv[j] = ORC_STUB_DIST; } /* .cu:2110-2117 */
        for (c0 = 0; c0 < m; c0 += cap) {
            int64_t c1 = c0 + cap < m ? c0 + cap : m;
            int np = 0;
            for (j = 0; j < k; ++j) { pool[np].key = lv[j]; pool[np].id = lid[j]; ++np; }
            for (j = c0; j < c1; ++j) {
                pool[np].key = orc_distance(x + i * d, x + (int64_t)cand[j] * d, d, metric);
                pool[np].id = cand[j];
                ++np;
            }
            select_and_store(pool, np, k, flip, lid, lv);
        }
        free(cand);
        free(pool);
    }
    return err ? -1 : 0;
}

/* One propagation round, propagate{Euclidean,Taxicab,Cosine}Kernel
 * (.cu:1033-1326, :1328-1621, :1623-1935), host loop .cu:1937-1977.
 * C(i) = 2-hop ids minus i minus i's own neighbours (.cu:1078-1100); an empty
 * C(i) leaves the row untouched (the while loop never runs, .cu:1117). */
int orc_propagate(const float* x, int64_t n, int d, int k, int metric,
                  const int32_t* ids_in, const float* vals_in,
                  int32_t* ids_out, float* vals_out, int cosine_fixed) {
    int flip = metric == ORC_COSINE && !cosine_fixed;
    int64_t i;
    int err = 0;
    if (k < 1 || n < k + 1) return -1;
#pragma omp parallel for schedule(dynamic, 64)
    for (i = 0; i < n; ++i) {
        const int32_t* own = ids_in + i * k;
        int32_t* cand = (int32_t*)malloc(sizeof(int32_t) * ((int64_t)k * k + 1));
        orc_ent* pool = (orc_ent*)malloc(sizeof(orc_ent) * ((int64_t)k * k + k));
        int m = 0, j, l, w;
        if (!cand || !pool) {
            free(cand); free(pool);
#pragma omp atomic write
            err = 1;
            continue;
        }
        for (j = 0; j < k; ++j) {
            int32_t nb = own[j];
            if (nb < 0) continue;
            for (l = 0; l < k; ++l) {
                int32_t c = ids_in[(int64_t)nb * k + l];
                if (c >= 0) cand[m++] = c;
            }
        }
        qsort(cand, m, sizeof(int32_t), i32_cmp);
        w = 0;
        for (j = 0; j < m; ++j) {
            int32_t c = cand[j];
            int skip = c == i || (w > 0 && cand[w - 1] == c);
            for (l = 0; l < k && !skip; ++l) skip = own[l] == c;
            if (!skip) cand[w++] = c;
        }
        m = w;
        if (m == 0) {
            memcpy(ids_out + i * k, own, sizeof(int32_t) * k);
            memcpy(vals_out + i * k, vals_in + i * k, sizeof(float) * k);
        } else {
            int np = 0;
            for (j = 0; j < k; ++j) { pool[np].key = vals_in[i * k + j]; pool[np].id = own[j]; ++np; }
            for (j = 0; j < m; ++j) {
                pool[np].key = orc_distance(x + i * d, x + (int64_t)cand[j] * d, d, metric);
                pool[np].id = cand[j];
                ++np;
            }
            select_and_store(pool, np, k, flip, ids_out + i * k, vals_out + i * k);
        }
        free(cand);
        free(pool);
    }
    return err ? -1 : 0;
}

/* calculate() (.cu:2861-2896): normalize, forest, stub + initial, rounds. */
int orc_run(float* x, int64_t n, int d, int k, int trees, int leaf, int props, int metric,
            uint32_t seed, int rng_order, int cosine_fixed,
            int32_t* ids, float* vals, double* stage_ms) {
    orc_rng g;
    int32_t *perm, *rs, *rc, *id2;
    float* v2;
    double t0, t1, t2, t3, t4;
    int t, p, rcode = 0;
    if (n < k + 1 || k < 1 || trees < 1 || d < 1) return -1;
    perm = (int32_t*)malloc(sizeof(int32_t) * n * trees * 3);
    id2 = (int32_t*)malloc(sizeof(int32_t) * n * k);
    v2 = (float*)malloc(sizeof(float) * n * k);
    if (!perm || !id2 || !v2) { free(perm); free(id2); free(v2); return -1; }
    rs = perm + n * trees;
    rc = rs + n * trees;
    orc_rng_seed(&g, seed);
    t0 = now_ms();
    orc_normalize(x, n, d);
    t1 = now_ms();
    for (t = 0; t < trees && !rcode; ++t)
        rcode = orc_build_tree(x, n, d, leaf, k, &g, rng_order, perm + (int64_t)t * n,
                               rs + (int64_t)t * n, rc + (int64_t)t * n, NULL, NULL, NULL);
    t2 = now_ms();
    if (!rcode) rcode = orc_initial_knn(x, n, d, k, trees, metric, perm, rs, rc, ids, vals, cosine_fixed);
    t3 = now_ms();
    for (p = 0; p < props && !rcode; ++p) {
        rcode = orc_propagate(x, n, d, k, metric, ids, vals, id2, v2, cosine_fixed);
        memcpy(ids, id2, sizeof(int32_t) * n * k);
        memcpy(vals, v2, sizeof(float) * n * k);
    }
    t4 = now_ms();
    if (stage_ms) {
        stage_ms[0] = t1 - t0; stage_ms[1] = t2 - t1; stage_ms[2] = t3 - t2; stage_ms[3] = t4 - t3;
    }
    free(perm); free(id2); free(v2);
    return rcode;
}

/* --------------------------------------------------------------------------
 * FP64 exact reference for recall@k (not part of the reference repository).
 * ------------------------------------------------------------------------ */
double orc_distance64(const float* a, const float* b, int d, int metric) {
    int dd;
    if (metric == ORC_COSINE) {
        double ab = 0, aa = 0, bb = 0;
        for (dd = 0; dd < d; ++dd) {
            ab += (double)a[dd] * b[dd];
            aa += (double)a[dd] * a[dd];
            bb += (double)b[dd] * b[dd];
        }
        return 1.0 - ab / (sqrt(aa) * sqrt(bb));
    } else {
        double s = 0;
        for (dd = 0; dd < d; ++dd) {
            double t = (double)a[dd] - (double)b[dd];
            s += metric == ORC_EUCLIDEAN ? t * t : fabs(t);
        }
        return s;
    }
}

void orc_exact_kth(const float* x, int64_t n, int d, int k, int metric,
                   const int64_t* queries, int64_t nq, double* kth) {
    int64_t qi;
#pragma omp parallel for schedule(dynamic, 4)
    for (qi = 0; qi < nq; ++qi) {
        int64_t q = queries[qi], j;
        double* best = (double*)malloc(sizeof(double) * k);
        int cnt = 0, l;
        for (j = 0; j < n; ++j) {
            double v;
            if (j == q) continue;
            v = orc_distance64(x + q * d, x + j * d, d, metric);
            if (cnt < k) {
                l = cnt++;
                while (l > 0 && best[l - 1] > v) { best[l] = best[l - 1]; --l; }
                best[l] = v;
            } else if (v < best[k - 1]) {
                l = k - 1;
                while (l > 0 && best[l - 1] > v) { best[l] = best[l - 1]; --l; }
                best[l] = v;
            }
        }
        kth[qi] = cnt == k ? best[k - 1] : INFINITY;
        free(best);
    }
}

double orc_recall(const float* x, int64_t n, int d, int k, int metric,
                  const int64_t* queries, int64_t nq, const double* kth,
                  const int32_t* ids) {
    int64_t qi, hits = 0;
    (void)n;
#pragma omp parallel for reduction(+ : hits)
    for (qi = 0; qi < nq; ++qi) {
        int64_t q = queries[qi];
        int j;
        for (j = 0; j < k; ++j) {
            int32_t c = ids[q * k + j];
            if (c < 0 || c == q) continue;
            if (orc_distance64(x + q * d, x + (int64_t)c * d, d, metric) <= kth[qi]) ++hits;
        }
    }
    return nq ? (double)hits / ((double)nq * k) : 0.0;
}

/* --------------------------------------------------------------------------
 * Result file (.cu:2908-2978): header "N;k;8\n", the long 0x01020304, then per
 * point k records of {int32 id, float32 value}, sorted by value ascending
 * (COSINE: descending), ties by lower id.
 * ------------------------------------------------------------------------ */
static int cmp_desc(const void* pa, const void* pb) {
    const orc_ent* a = (const orc_ent*)pa;
    const orc_ent* b = (const orc_ent*)pb;
    if (a->key > b->key) return -1;
    if (a->key < b->key) return 1;
    return (a->id > b->id) - (a->id < b->id);
}

int orc_write_result(const char* path, int64_t n, int k, int metric,
                     const int32_t* ids, const float* vals) {
    FILE* f = fopen(path, "wb");
    int64_t i;
    long marker = 0x01020304;
    orc_ent* row;
    if (!f) return -1;
    row = (orc_ent*)malloc(sizeof(orc_ent) * k);
    fprintf(f, "%lld;%d;%d\n", (long long)n, k, (int)sizeof(long));
    fwrite(&marker, sizeof(long), 1, f);
    for (i = 0; i < n; ++i) {
        int j;
        for (j = 0; j < k; ++j) { row[j].id = ids[i * k + j]; row[j].key = vals[i * k + j]; }
        qsort(row, k, sizeof(orc_ent), metric == ORC_COSINE ? cmp_desc : ent_cmp);
        for (j = 0; j < k; ++j) {
            fwrite(&row[j].id, 4, 1, f);
            fwrite(&row[j].key, 4, 1, f);
        }
    }
    free(row);
    return fclose(f) == 0 ? 0 : -1;
}
