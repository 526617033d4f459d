/*
 * CPU oracle for the IVF-Flat query path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker.  Conventions and reference citations: oracle.h.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static int g_mut = 0;

void orc_set_mutation(int m) { g_mut = m; }
int orc_get_mutation(void) { return g_mut; }

/* Strict total order on (value, id): does (a, ia) come before (b, ib) in a
 * result row?  L2 ranks small distances first, IP large ones
 * (src/IVFScan.hpp:131 vs :116); equal values rank by ascending id (faiss
 * cmp2, third_party/faiss/faiss/utils/Heap.h:127-129).  NaN never ranks
 * first. */
static inline int precedes(int metric, float a, int64_t ia, float b,
                           int64_t ib) {
    if (a == b) return g_mut == 2 ? ia > ib : ia < ib;
    return metric == ORC_METRIC_L2 ? a < b : a > b;
}

static inline float sentinel(int metric) {
    /* src/heap.hpp:13-25 */
    return metric == ORC_METRIC_L2 ? FLT_MAX : -FLT_MAX;
}

static inline int64_t eff_dim(int64_t d) {
    return (g_mut == 1 && d > 0) ? d - 1 : d;
}

/* Partial L2 over dims [lo, hi), continuing the chain in acc. */
static inline float l2_range(const float* a, const float* b, int64_t lo,
                             int64_t hi, float acc) {
    for (int64_t i = lo; i < hi; i++) {
        float t = a[i] - b[i];
        acc = fmaf(t, t, acc);
    }
    return acc;
}

float orc_distance(int metric, const float* a, const float* b, int64_t d) {
    int64_t n = eff_dim(d);
    if (metric == ORC_METRIC_L2) return l2_range(a, b, 0, n, 0.0f);
    float acc = 0.0f;
    for (int64_t i = 0; i < n; i++) acc = fmaf(a[i], b[i], acc);
    return acc;
}

/* ---- k-heap: worst entry at the root (src/heap.hpp:35-71 shape, with the
 * (value, id) order above). 0-based arrays, 1-based arithmetic inside. */
typedef struct {
    int metric;
    int64_t k;
    float* v;
    int64_t* id;
} heap_t;

static void heap_init(heap_t* h, int metric, int64_t k, float* v,
                      int64_t* id) {
    h->metric = metric;
    h->k = k;
    h->v = v;
    h->id = id;
    float s = sentinel(metric);
    for (int64_t i = 0; i < k; i++) {
        v[i] = s;
        id[i] = -1;
    }
}

/* Put (val, vid) at the root of a heap of size n and sift it down. */
static void heap_sift_root(int metric, int64_t n, float* v, int64_t* id,
                           float val, int64_t vid) {
    int64_t i = 0;
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= n) break;
        /* pick the worse child: the one the other child precedes */
        if (c + 1 < n && precedes(metric, v[c], id[c], v[c + 1], id[c + 1]))
            c = c + 1;
        if (precedes(metric, v[c], id[c], val, vid)) break;
        v[i] = v[c];
        id[i] = id[c];
        i = c;
    }
    v[i] = val;
    id[i] = vid;
}

static inline void heap_push(heap_t* h, float val, int64_t vid) {
    if (h->k > 0 && precedes(h->metric, val, vid, h->v[0], h->id[0]))
        heap_sift_root(h->metric, h->k, h->v, h->id, val, vid);
}

/* In-place heap sort, best first (src/heap.hpp:73-82). */
static void heap_sort(heap_t* h) {
    for (int64_t i = h->k - 1; i > 0; i--) {
        float val = h->v[i];
        int64_t vid = h->id[i];
        h->v[i] = h->v[0];
        h->id[i] = h->id[0];
        heap_sift_root(h->metric, i, h->v, h->id, val, vid);
    }
}

void orc_topk(int metric, int64_t n, const float* val, const int64_t* ids,
              int64_t k, float* out_val, int64_t* out_ids) {
    heap_t h;
    heap_init(&h, metric, k, out_val, out_ids);
    for (int64_t i = 0; i < n; i++) heap_push(&h, val[i], ids[i]);
    heap_sort(&h);
}

void orc_merge(int metric, int64_t nq, int64_t parts, int64_t k_in,
               const float* D_parts, const int64_t* I_parts, int64_t k,
               float* D, int64_t* I) {
    int64_t used = (g_mut == 6 && parts > 0) ? parts - 1 : parts;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nq; q++) {
        heap_t h;
        heap_init(&h, metric, k, D + q * k, I + q * k);
        for (int64_t p = 0; p < used; p++) {
            const float* dv = D_parts + (p * nq + q) * k_in;
            const int64_t* iv = I_parts + (p * nq + q) * k_in;
            for (int64_t i = 0; i < k_in; i++)
                if (iv[i] >= 0) heap_push(&h, dv[i], iv[i]);
        }
        heap_sort(&h);
    }
}

static inline int64_t clamp_nprobe(int64_t nprobe, int64_t nlist) {
    return (nprobe <= 0 || nprobe > nlist) ? nlist : nprobe;
}

int64_t orc_coarse(int metric, int64_t nq, const float* q, int64_t nlist,
                   const float* cent, int64_t d, int64_t nprobe,
                   float* out_dist, int64_t* out_list) {
    int64_t P = clamp_nprobe(nprobe, nlist);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < nq; i++) {
        heap_t h;
        heap_init(&h, metric, P, out_dist + i * P, out_list + i * P);
        for (int64_t c = 0; c < nlist; c++)
            heap_push(&h, orc_distance(metric, q + i * d, cent + c * d, d), c);
        heap_sort(&h);
    }
    return P;
}

void orc_assign(int metric, int64_t n, const float* x, int64_t nlist,
                const float* cent, int64_t d, float* out_dist,
                int64_t* out_list) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        heap_t h;
        heap_init(&h, metric, 1, out_dist + i, out_list + i);
        for (int64_t c = 0; c < nlist; c++)
            heap_push(&h, orc_distance(metric, x + i * d, cent + c * d, d), c);
    }
}

typedef struct {
    int64_t list;
    float dist;
    int64_t row;
} fill_t;

static int fill_cmp_l2(const void* pa, const void* pb) {
    const fill_t* a = (const fill_t*)pa;
    const fill_t* b = (const fill_t*)pb;
    if (a->list != b->list) return a->list < b->list ? -1 : 1;
    if (a->dist != b->dist) return a->dist < b->dist ? -1 : 1;
    return a->row < b->row ? -1 : (a->row > b->row);
}

static int fill_cmp_ip(const void* pa, const void* pb) {
    const fill_t* a = (const fill_t*)pa;
    const fill_t* b = (const fill_t*)pb;
    if (a->list != b->list) return a->list < b->list ? -1 : 1;
    if (a->dist != b->dist) return a->dist > b->dist ? -1 : 1;
    return a->row < b->row ? -1 : (a->row > b->row);
}

void orc_build_lists(int metric, int64_t n, const float* x, int64_t nlist,
                     const float* cent, int64_t d, int64_t* list_off,
                     int64_t* order) {
    float* dist = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int64_t* lst = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    fill_t* f = (fill_t*)malloc(sizeof(fill_t) * (size_t)(n > 0 ? n : 1));
    orc_assign(metric, n, x, nlist, cent, d, dist, lst);
    for (int64_t i = 0; i < n; i++) {
        f[i].list = lst[i];
        f[i].dist = dist[i];
        f[i].row = i;
    }
    qsort(f, (size_t)n, sizeof(fill_t),
          metric == ORC_METRIC_L2 ? fill_cmp_l2 : fill_cmp_ip);
    for (int64_t l = 0; l <= nlist; l++) list_off[l] = 0;
    for (int64_t i = 0; i < n; i++) list_off[f[i].list + 1]++;
    for (int64_t l = 0; l < nlist; l++) list_off[l + 1] += list_off[l];
    for (int64_t i = 0; i < n; i++) order[i] = f[i].row;
    free(f);
    free(lst);
    free(dist);
}

void orc_warmup_tau(const orc_index* ix, int64_t nq, const float* q,
                    const int64_t* probes, int64_t nprobe, int64_t k,
                    int64_t L, int64_t W, float* tau) {
    int64_t Le = L < nprobe ? L : nprobe; /* src/Index.cpp:1008-1011 */
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < nq; i++) {
        float* hv = (float*)malloc(sizeof(float) * (size_t)(k > 0 ? k : 1));
        int64_t* hi = (int64_t*)malloc(sizeof(int64_t) * (size_t)(k > 0 ? k : 1));
        heap_t h;
        heap_init(&h, ix->metric, k, hv, hi);
        const float* qv = q + i * ix->d;
        for (int64_t j = 0; j < Le; j++) {
            int64_t l = probes[i * nprobe + j];
            int64_t n_l = ix->list_off[l + 1] - ix->list_off[l];
            int64_t rows = n_l < W ? n_l : W;
            if (g_mut == 5 && rows > 0) rows--;
            for (int64_t r = 0; r < rows; r++) {
                int64_t g = ix->list_off[l] + r;
                heap_push(&h,
                          orc_distance(ix->metric, qv, ix->codes + g * ix->d,
                                       ix->d),
                          ix->ids[g]);
            }
        }
        tau[i] = k > 0 ? hv[0] : sentinel(ix->metric);
        free(hv);
        free(hi);
    }
}

/* One query's scan over its probes.  dim_block <= 0: plain full distances.
 * Otherwise L2 block-pruned as described in oracle.h. */
static void scan_one(const orc_index* ix, const float* qv,
                     const int64_t* probes, int64_t P, int64_t k,
                     int64_t dim_block, float tau, int use_running_kth,
                     float* D, int64_t* I, orc_stats* st) {
    heap_t h;
    heap_init(&h, ix->metric, k, D, I);
    int64_t d = ix->d;
    int64_t de = eff_dim(d);
    int64_t Pe = (g_mut == 4 && P > 0) ? P - 1 : P;
    int64_t nb = dim_block > 0 ? (d + dim_block - 1) / dim_block : 1;
    for (int64_t j = 0; j < Pe; j++) {
        int64_t l = probes[j];
        int64_t lo = ix->list_off[l], hi = ix->list_off[l + 1];
        st->candidates += hi - lo;
        for (int64_t g = lo; g < hi; g++) {
            const float* x = ix->codes + g * d;
            if (dim_block <= 0) {
                heap_push(&h, orc_distance(ix->metric, qv, x, d), ix->ids[g]);
                continue;
            }
            /* src/node.cpp:639-656: running distance checked after each
             * block but the last; drop the row once it exceeds the bound. */
            st->blocks_total += nb;
            float acc = 0.0f;
            int pruned = 0;
            for (int64_t b = 0; b < nb; b++) {
                int64_t a0 = b * dim_block;
                int64_t a1 = a0 + dim_block < de ? a0 + dim_block : de;
                acc = l2_range(qv, x, a0, a1, acc);
                if (b == nb - 1) break;
                float bound = tau;
                if (use_running_kth && k > 0 && h.v[0] < bound)
                    bound = h.v[0];
                int cut = (g_mut == 3) ? (acc >= bound) : (acc > bound);
                if (cut) {
                    st->blocks_skipped += nb - 1 - b;
                    pruned = 1;
                    break;
                }
            }
            if (!pruned) heap_push(&h, acc, ix->ids[g]);
        }
    }
    heap_sort(&h);
}

static void scan_all(const orc_index* ix, int64_t nq, const float* q,
                     const int64_t* probes, int64_t nprobe, int64_t k,
                     int64_t dim_block, const float* tau, int use_running_kth,
                     float* D, int64_t* I, int threads, orc_stats* st) {
    int64_t cand = 0, bt = 0, bs = 0;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : cand, bt, bs) \
    num_threads(nt)
    for (int64_t i = 0; i < nq; i++) {
        orc_stats s = {0, 0, 0};
        float t = tau ? tau[i] : sentinel(ix->metric);
        scan_one(ix, q + i * ix->d, probes + i * nprobe, nprobe, k, dim_block,
                 t, use_running_kth, D + i * k, I + i * k, &s);
        cand += s.candidates;
        bt += s.blocks_total;
        bs += s.blocks_skipped;
    }
    if (st) {
        st->candidates = cand;
        st->blocks_total = bt;
        st->blocks_skipped = bs;
    }
}

void orc_scan(const orc_index* ix, int64_t nq, const float* q,
              const int64_t* probes, int64_t nprobe, int64_t k, float* D,
              int64_t* I, int threads, orc_stats* st) {
    scan_all(ix, nq, q, probes, nprobe, k, 0, NULL, 0, D, I, threads, st);
}

void orc_search_pruned(const orc_index* ix, int64_t nq, const float* q,
                       const int64_t* probes, int64_t nprobe, int64_t k,
                       int64_t dim_block, const float* tau,
                       int use_running_kth, float* D, int64_t* I, int threads,
                       orc_stats* st) {
    if (dim_block <= 0) dim_block = ix->d;
    scan_all(ix, nq, q, probes, nprobe, k, dim_block, tau, use_running_kth, D,
             I, threads, st);
}

int64_t orc_search(const orc_index* ix, int64_t nq, const float* q,
                   int64_t nprobe, int64_t k, float* D, int64_t* I,
                   int threads, orc_stats* st) {
    int64_t P = clamp_nprobe(nprobe, ix->nlist);
    float* cd = (float*)malloc(sizeof(float) * (size_t)(nq * P + 1));
    int64_t* cl = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nq * P + 1));
    orc_coarse(ix->metric, nq, q, ix->nlist, ix->centroids, ix->d, P, cd, cl);
    orc_scan(ix, nq, q, cl, P, k, D, I, threads, st);
    free(cd);
    free(cl);
    return P;
}

void orc_kmeans(int metric, int64_t n, const float* x, int64_t nlist,
                int64_t d, int niter, float* cent) {
    const float* xa = x;
    float* xn = NULL;
    if (metric == ORC_METRIC_IP) {
        /* IP assignment uses unit-norm rows (src/Clustering.cpp:125-131) */
        xn = (float*)malloc(sizeof(float) * (size_t)(n * d + 1));
        for (int64_t i = 0; i < n; i++) {
            float s = 0.0f;
            for (int64_t j = 0; j < d; j++)
                s = fmaf(x[i * d + j], x[i * d + j], s);
            float r = sqrtf(s);
            for (int64_t j = 0; j < d; j++)
                xn[i * d + j] = r > 0.0f ? x[i * d + j] / r : x[i * d + j];
        }
        xa = xn;
    }
    float* dist = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    int64_t* lst = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    double* sum = (double*)malloc(sizeof(double) * (size_t)(nlist * d + 1));
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nlist + 1));
    for (int it = 0; it < niter; it++) {
        orc_assign(metric, n, xa, nlist, cent, d, dist, lst);
        memset(sum, 0, sizeof(double) * (size_t)(nlist * d));
        memset(cnt, 0, sizeof(int64_t) * (size_t)nlist);
        for (int64_t i = 0; i < n; i++) {
            int64_t c = lst[i];
            cnt[c]++;
            for (int64_t j = 0; j < d; j++) sum[c * d + j] += (double)x[i * d + j];
        }
        for (int64_t c = 0; c < nlist; c++)
            if (cnt[c] > 0)
                for (int64_t j = 0; j < d; j++)
                    cent[c * d + j] = (float)(sum[c * d + j] / (double)cnt[c]);
    }
    if (metric == ORC_METRIC_IP) {
        /* src/Index.cpp:409-421: unit-norm centroids for IP */
        for (int64_t c = 0; c < nlist; c++) {
            float s = 0.0f;
            for (int64_t j = 0; j < d; j++)
                s = fmaf(cent[c * d + j], cent[c * d + j], s);
            if (s != 0.0f) {
                float r = sqrtf(s);
                for (int64_t j = 0; j < d; j++) cent[c * d + j] /= r;
            }
        }
    }
    free(cnt);
    free(sum);
    free(lst);
    free(dist);
    free(xn);
}

double orc_recall(const int64_t* I, const int64_t* GT, int64_t nq, int64_t k,
                  int64_t gt_k) {
    if (gt_k <= 0) gt_k = k;
    int64_t hit = 0;
    for (int64_t i = 0; i < nq; i++) {
        for (int64_t j = 0; j < k; j++) {
            int64_t v = I[i * k + j];
            if (v == -1) break;
            for (int64_t t = 0; t < k; t++)
                if (GT[i * gt_k + t] == v) {
                    hit++;
                    break;
                }
        }
    }
    return nq * k > 0 ? (double)hit / (double)(nq * k) : 0.0;
}
