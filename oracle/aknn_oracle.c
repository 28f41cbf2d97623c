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
void orc_normalize(float* x, int64_t n, int d) { orc_normalize_ex(x, n, d, 0); }

/* scale_all (fixed mode): every dimension enters the max-abs and is scaled. */
void orc_normalize_ex(float* x, int64_t n, int d, int scale_all) { orc_normalize_mode(x, n, d, scale_all, 0); }

void orc_normalize_mode(float* x, int64_t n, int d, int scale_all, int mean_fp64) {
    int dd;
    int dm = scale_all ? d : (d < 256 ? d : 256);
    float g = 0.0f;
#pragma omp parallel for schedule(static)
    for (dd = 0; dd < d; ++dd) {
        float m = 0.0f;
        int64_t i;
        if (mean_fp64) {
            double tot = 0.0;
            int64_t b;
            for (b = 0; b < n; b += 2048) {
                double part = 0.0;
                int64_t e = b + 2048 < n ? b + 2048 : n;
                for (i = b; i < e; ++i) part = part + (double)x[i * d + dd];
                tot = tot + part;
            }
            m = (float)(tot / (double)n);
        } else {
            for (i = 0; i < n; ++i) m = m + x[i * d + dd];
            m = m / (float)n;
        }
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
        /* ORC_TAXICAB_L1 (fixed mode): the L1 total itself */
        return metric == ORC_TAXICAB_L1 ? tot : sqrtf(tot);
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

/* ---------------------------------------------------------------------------
 * Median-split forest (not in the reference: its split is the sign of
 * x . (x_p2 - x_p1), .cu:2151-2162; this is the balanced rule the rebuild
 * offers beside it, with its own semantics fixed here).
 *
 * Directions: tree t's level directions are seeded Gaussian vectors, drawn as
 * splitmix64 uniforms (0, 1) -> Box-Muller pairs (cos, then sin) in double,
 * rounded to float, level-major: r[j][d].
 * ------------------------------------------------------------------------- */
static uint64_t sm64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double sm64_open01(uint64_t* s) {
    return ((double)(sm64_next(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_directions(uint32_t seed, int tree, int count, int d, float* r) {
    uint64_t s = ((uint64_t)seed << 32) ^ ((uint64_t)(uint32_t)(tree + 1) * 0xD1B54A32D192ED03ull);
    int64_t total = (int64_t)count * d, i = 0;
    while (i < total) {
        double u1 = sm64_open01(&s), u2 = sm64_open01(&s);
        double rad = sqrt(-2.0 * log(u1)), th = 6.283185307179586 * u2;
        r[i++] = (float)(rad * cos(th));
        if (i < total) r[i++] = (float)(rad * sin(th));
    }
}

/* Ordered key of a float: unsigned order equals the float order (-0 < +0,
 * NaNs at the ends), so every projection has a place in a total order. */
static uint32_t ord_u32(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

typedef struct { uint32_t u; int32_t rel; } orc_mkey;

static int mkey_cmp(const void* pa, const void* pb) {
    const orc_mkey* a = (const orc_mkey*)pa;
    const orc_mkey* b = (const orc_mkey*)pb;
    if (a->u != b->u) return a->u < b->u ? -1 : 1;
    return (a->rel > b->rel) - (a->rel < b->rel);
}

typedef struct {
    const float* x;
    int64_t n;
    int d, leaf_max, k, ndirs, bin_bits;
    const float* dirs;
    const int32_t* bins;
    int32_t *perm, *rs, *rc, *tmp;
    float* proj;
    uint8_t* side;
    int32_t levels;
} orc_mtree;

/* One bucket [s, s+c) at level lv; (ps, pc) is the range of its deepest
 * ancestor with at least k+1 points.  The side of position p:
 *   lv < bin_bits (routed trees): bit (bin_bits-1-lv) of the point's bin;
 *   else j = lv - bin_bits < ndirs: proj = fmaf chain of x . r[j] in dimension
 *     order; m = min(c, 511) samples at relative positions q*c/m (q < m) are
 *     ordered by (ord_u32(proj), rel); the one at index m/2 is the pivot; a
 *     position is RIGHT when (ord_u32(proj), rel) >= the pivot's (both sides
 *     non-empty for c >= 2);
 *   else (directions exhausted): RIGHT when rel >= c/2.
 * The partition is stable; an empty child is dropped. */
static void mtree_bucket(orc_mtree* T, int64_t s, int64_t c, int lv, int32_t ps, int32_t pc) {
    int64_t p, nl = 0, w;
    int32_t rs0 = c >= T->k + 1 ? (int32_t)s : ps;
    int32_t rc0 = c >= T->k + 1 ? (int32_t)c : pc;
    if (lv + 1 > T->levels) T->levels = lv + 1;
    if (c <= T->leaf_max) {
        for (p = s; p < s + c; ++p) { T->rs[T->perm[p]] = rs0; T->rc[T->perm[p]] = rc0; }
        return;
    }
    if (T->bins && lv < T->bin_bits) {
        int sh = T->bin_bits - 1 - lv;
        for (p = s; p < s + c; ++p) T->side[p] = (uint8_t)((T->bins[T->perm[p]] >> sh) & 1);
    } else {
        int j = lv - (T->bins ? T->bin_bits : 0);
        if (j >= T->ndirs) {
            for (p = s; p < s + c; ++p) T->side[p] = (uint8_t)((p - s) >= c / 2);
        } else {
            const float* r = T->dirs + (int64_t)j * T->d;
            int64_t m = c < 511 ? c : 511, q;
            orc_mkey* smp = (orc_mkey*)malloc(sizeof(orc_mkey) * m);
            orc_mkey piv;
            for (p = s; p < s + c; ++p) {
                const float* xr = T->x + (int64_t)T->perm[p] * T->d;
                float acc = 0.0f;
                int dd;
                for (dd = 0; dd < T->d; ++dd) acc = fmaf(xr[dd], r[dd], acc);
                T->proj[p] = acc;
            }
            for (q = 0; q < m; ++q) {
                int64_t rel = q * c / m;
                smp[q].u = ord_u32(T->proj[s + rel]);
                smp[q].rel = (int32_t)rel;
            }
            qsort(smp, m, sizeof(orc_mkey), mkey_cmp);
            piv = smp[m / 2];
            free(smp);
            for (p = s; p < s + c; ++p) {
                uint32_t u = ord_u32(T->proj[p]);
                int32_t rel = (int32_t)(p - s);
                T->side[p] = (uint8_t)(u > piv.u || (u == piv.u && rel >= piv.rel));
            }
        }
    }
    for (p = s; p < s + c; ++p) nl += T->side[p] == 0;
    w = s;
    for (p = s; p < s + c; ++p) if (T->side[p] == 0) T->tmp[w++] = T->perm[p];
    for (p = s; p < s + c; ++p) if (T->side[p] != 0) T->tmp[w++] = T->perm[p];
    memcpy(T->perm + s, T->tmp + s, sizeof(int32_t) * c);
    if (nl > 0) mtree_bucket(T, s, nl, lv + 1, rs0, rc0);
    if (c - nl > 0) mtree_bucket(T, s + nl, c - nl, lv + 1, rs0, rc0);
}

int orc_build_tree_median(const float* x, int64_t n, int d, int leaf_max, int k,
                          const float* dirs, int ndirs, const int32_t* bins, int bin_bits,
                          int32_t* perm, int32_t* range_start, int32_t* range_count,
                          int32_t* n_levels) {
    orc_mtree T;
    int64_t p;
    if (n < 1 || d < 1 || leaf_max < 0 || k < 0 || n < (int64_t)k + 1 || n > 0x7fffffffLL) return -1;
    if (ndirs < 0 || bin_bits < 0 || bin_bits > 20 || (ndirs > 0 && !dirs)) return -1;
    memset(&T, 0, sizeof(T));
    T.x = x; T.n = n; T.d = d; T.leaf_max = leaf_max < 1 ? 1 : leaf_max; T.k = k;
    T.dirs = dirs; T.ndirs = ndirs; T.bins = bins; T.bin_bits = bins ? bin_bits : 0;
    T.perm = perm; T.rs = range_start; T.rc = range_count;
    T.tmp = (int32_t*)malloc(sizeof(int32_t) * n);
    T.proj = (float*)malloc(sizeof(float) * n);
    T.side = (uint8_t*)malloc(n);
    if (!T.tmp || !T.proj || !T.side) { free(T.tmp); free(T.proj); free(T.side); return -1; }
    for (p = 0; p < n; ++p) perm[p] = (int32_t)p;
    mtree_bucket(&T, 0, n, 0, 0, (int32_t)n);
    if (n_levels) *n_levels = T.levels;
    free(T.tmp); free(T.proj); free(T.side);
    return 0;
}

int orc_median_depth(int64_t n, int leaf) {
    int lv = 0;
    int64_t c = n, l = leaf < 1 ? 1 : leaf;
    while (c > l) { c = (c + 1) / 2; ++lv; }
    return lv + 8;
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

/* calculate() (.cu:2861-2896): normalize, forest, stub + initial, rounds.
 * split 0: the reference's sign-split forest on the rand() stream; split 1:
 * the median-split forest (orc_build_tree_median) with tree t's directions
 * orc_directions(seed, t, orc_median_depth(n, leaf), d), and for t <
 * router_trees the shared bins[n] (bin_bits bits) as its first levels. */
int orc_run_ex(float* x, int64_t n, int d, int k, int trees, int leaf, int props, int metric,
               uint32_t seed, int rng_order, int cosine_fixed, int split,
               const int32_t* bins, int bin_bits, int router_trees, int scale_all,
               int32_t* ids, float* vals, double* stage_ms) {
    orc_rng g;
    int32_t *perm, *rs, *rc, *id2;
    float *v2, *dirs = NULL;
    double t0, t1, t2, t3, t4;
    int t, p, rcode = 0, ndirs = 0;
    if (n < k + 1 || k < 1 || trees < 1 || d < 1 || split < 0 || split > 1) return -1;
    perm = (int32_t*)malloc(sizeof(int32_t) * n * trees * 3);
    id2 = (int32_t*)malloc(sizeof(int32_t) * n * k);
    v2 = (float*)malloc(sizeof(float) * n * k);
    if (split) {
        ndirs = orc_median_depth(n, leaf);
        dirs = (float*)malloc(sizeof(float) * (size_t)ndirs * d);
    }
    if (!perm || !id2 || !v2 || (split && !dirs)) { free(perm); free(id2); free(v2); free(dirs); return -1; }
    rs = perm + n * trees;
    rc = rs + n * trees;
    orc_rng_seed(&g, seed);
    t0 = now_ms();
    orc_normalize_ex(x, n, d, scale_all);
    t1 = now_ms();
    for (t = 0; t < trees && !rcode; ++t) {
        if (split) {
            orc_directions(seed, t, ndirs, d, dirs);
            rcode = orc_build_tree_median(x, n, d, leaf, k, dirs, ndirs, t < router_trees ? bins : NULL, bin_bits,
                                          perm + (int64_t)t * n, rs + (int64_t)t * n, rc + (int64_t)t * n, NULL);
        } else {
            rcode = orc_build_tree(x, n, d, leaf, k, &g, rng_order, perm + (int64_t)t * n,
                                   rs + (int64_t)t * n, rc + (int64_t)t * n, NULL, NULL, NULL);
        }
    }
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
    free(perm); free(id2); free(v2); free(dirs);
    return rcode;
}

int orc_run(float* x, int64_t n, int d, int k, int trees, int leaf, int props, int metric,
            uint32_t seed, int rng_order, int cosine_fixed,
            int32_t* ids, float* vals, double* stage_ms) {
    return orc_run_ex(x, n, d, k, trees, leaf, props, metric, seed, rng_order, cosine_fixed, 0, NULL, 0, 0, 0,
                      ids, vals, stage_ms);
}

/* ---------------------------------------------------------------------------
 * Router MLP (not in the reference): forward and backward in FP64.  With
 * round_bf16 the values the device rounds to BF16 are rounded here at the
 * same points (inputs, weights, hidden activations, the logit gradient and
 * the hidden gradients that feed the products) and the FP32 steps (bias add,
 * standardisation) are done in float; without it everything is FP64.
 * params: W1 [h][dp], b1 [h], W2 [h][h], b2 [h], W3 [nb][h], b3 [nb].
 * ------------------------------------------------------------------------- */
static float bf16_round(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) u |= 0x00400000u;
    else u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

double orc_mlp(const float* x, int64_t m, int d, int dp, int h, int nb,
               const float* params, const float* mu, float inv_sigma,
               const int32_t* labels, int round_bf16, float* grads, float* logits) {
    const float *W1 = params, *b1 = W1 + (int64_t)h * dp, *W2 = b1 + h, *b2 = W2 + (int64_t)h * h;
    const float *W3 = b2 + h, *b3 = W3 + (int64_t)nb * h;
    double *w1, *w2, *w3, *u, *h1, *h2, *dz3, *dz3b, *dz2, *dz2b, *dz1, *dz1b, loss = 0.0;
    double *g = NULL;
    int64_t r, count = 0, np = (int64_t)h * dp + h + (int64_t)h * h + h + (int64_t)nb * h + nb;
    int o, c, i;
#define RB(v) (round_bf16 ? (double)bf16_round((float)(v)) : (double)(v))
    w1 = (double*)malloc(sizeof(double) * ((int64_t)h * dp + (int64_t)h * h + (int64_t)nb * h));
    u = (double*)malloc(sizeof(double) * m * (dp + 2 * h + 2 * nb + 4 * h));
    if (!w1 || !u) { free(w1); free(u); return NAN; }
    w2 = w1 + (int64_t)h * dp; w3 = w2 + (int64_t)h * h;
    h1 = u + m * dp; h2 = h1 + m * h; dz3 = h2 + m * h; dz3b = dz3 + m * nb;
    dz2 = dz3b + m * nb; dz2b = dz2 + m * h; dz1 = dz2b + m * h; dz1b = dz1 + m * h;
    for (i = 0; i < h * dp; ++i) w1[i] = RB(W1[i]);
    for (i = 0; i < h * h; ++i) w2[i] = RB(W2[i]);
    for (i = 0; i < nb * h; ++i) w3[i] = RB(W3[i]);
    for (r = 0; r < m; ++r)
        for (c = 0; c < dp; ++c) {
            double v = 0.0;
            if (c < d) {
                if (round_bf16) v = RB((x[r * d + c] - mu[c]) * inv_sigma);
                else v = ((double)x[r * d + c] - mu[c]) * (double)inv_sigma;
            }
            u[r * dp + c] = v;
        }
    for (r = 0; r < m; ++r) {
        for (o = 0; o < h; ++o) {
            double s = 0.0, v;
            for (c = 0; c < dp; ++c) s += w1[(int64_t)o * dp + c] * u[r * dp + c];
            v = round_bf16 ? (double)((float)s + b1[o]) : s + b1[o];
            h1[r * h + o] = RB(v > 0 ? v : 0.0);
        }
        for (o = 0; o < h; ++o) {
            double s = 0.0, v;
            for (c = 0; c < h; ++c) s += w2[(int64_t)o * h + c] * h1[r * h + c];
            v = round_bf16 ? (double)((float)s + b2[o]) : s + b2[o];
            h2[r * h + o] = RB(v > 0 ? v : 0.0);
        }
        for (o = 0; o < nb; ++o) {
            double s = 0.0, v;
            for (c = 0; c < h; ++c) s += w3[(int64_t)o * h + c] * h2[r * h + c];
            v = round_bf16 ? (double)((float)s + b3[o]) : s + b3[o];
            dz3[r * nb + o] = v;  /* logits for now */
            if (logits) logits[r * nb + o] = (float)v;
        }
    }
    if (!labels) { free(w1); free(u); return 0.0; }
    for (r = 0; r < m; ++r) count += labels[r] >= 0 && labels[r] < nb;
    for (r = 0; r < m; ++r) {
        double* z = dz3 + r * nb;
        double mx = -INFINITY, s = 0.0;
        int y = labels[r];
        if (y < 0 || y >= nb) {
            for (o = 0; o < nb; ++o) { z[o] = 0.0; dz3b[r * nb + o] = 0.0; }
            continue;
        }
        for (o = 0; o < nb; ++o) if (z[o] > mx) mx = z[o];
        for (o = 0; o < nb; ++o) s += exp(z[o] - mx);
        loss += log(s) + mx - z[y];
        for (o = 0; o < nb; ++o) {
            double gv = (exp(z[o] - mx) / s - (o == y ? 1.0 : 0.0)) / (double)(count ? count : 1);
            if (round_bf16) gv = (double)(float)gv;
            z[o] = gv;
            dz3b[r * nb + o] = RB(gv);
        }
    }
    if (!grads) { free(w1); free(u); return loss / (double)(count ? count : 1); }
    g = (double*)calloc(np, sizeof(double));
    if (!g) { free(w1); free(u); return NAN; }
    {
        double *gW1 = g, *gb1 = gW1 + (int64_t)h * dp, *gW2 = gb1 + h, *gb2 = gW2 + (int64_t)h * h;
        double *gW3 = gb2 + h, *gb3 = gW3 + (int64_t)nb * h;
        for (r = 0; r < m; ++r) {
            for (o = 0; o < nb; ++o) {
                gb3[o] += dz3[r * nb + o];
                for (c = 0; c < h; ++c) gW3[(int64_t)o * h + c] += dz3b[r * nb + o] * h2[r * h + c];
            }
            for (c = 0; c < h; ++c) {
                double s = 0.0;
                for (o = 0; o < nb; ++o) s += dz3b[r * nb + o] * w3[(int64_t)o * h + c];
                if (!(h2[r * h + c] > 0)) s = 0.0;
                if (round_bf16) s = (double)(float)s;
                dz2[r * h + c] = s;
                dz2b[r * h + c] = RB(s);
            }
            for (o = 0; o < h; ++o) {
                gb2[o] += dz2[r * h + o];
                for (c = 0; c < h; ++c) gW2[(int64_t)o * h + c] += dz2b[r * h + o] * h1[r * h + c];
            }
            for (c = 0; c < h; ++c) {
                double s = 0.0;
                for (o = 0; o < h; ++o) s += dz2b[r * h + o] * w2[(int64_t)o * h + c];
                if (!(h1[r * h + c] > 0)) s = 0.0;
                if (round_bf16) s = (double)(float)s;
                dz1[r * h + c] = s;
                dz1b[r * h + c] = RB(s);
            }
            for (o = 0; o < h; ++o) {
                gb1[o] += dz1[r * h + o];
                for (c = 0; c < dp; ++c) gW1[(int64_t)o * dp + c] += dz1b[r * h + o] * u[r * dp + c];
            }
        }
    }
    for (r = 0; r < np; ++r) grads[r] = (float)g[r];
    free(g); free(w1); free(u);
#undef RB
    return loss / (double)(count ? count : 1);
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
