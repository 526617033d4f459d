/*
 * CPU oracle for the IVF-Flat query path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing on the product path links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * It restates, in plain C, the reference's plain IVF-Flat search and the
 * dimension-block pruned scan:
 *   distances          src/utils.h:480-487 (L2), src/utils.h:512-519 (IP)
 *   top-k heap         src/heap.hpp:13-90, ordered by (value, id) as in
 *                      third_party/faiss/faiss/utils/Heap.h:111-149 (cmp2)
 *   coarse quantizer   src/Index.cpp:967-995 + src/IVFScan.hpp:108-138
 *   warm-up bound      src/Index.cpp:999-1051
 *   plain scan         src/Index.cpp:1420-1554 + src/IVFScan.hpp:174-287
 *   pruned scan        src/node.cpp:606-678 (4 checks of d/4) generalised
 *                      to blocks of `dim_block` dims
 *   list build         src/Index.cpp:567-579, 581-698
 *   k-means            src/Clustering.cpp:14-152 (Lloyd iterations)
 *
 * Canonical conventions (stricter than the reference, which orders equal
 * distances by heap internals):
 *   - a distance is one sequential FMA chain over dims 0..d-1:
 *       L2: acc = fmaf(a_i - b_i, a_i - b_i, acc);  IP: acc = fmaf(a_i, b_i, acc)
 *   - results are ordered by (distance, id): ascending distance for L2,
 *     descending for IP, ascending id among equal distances;
 *   - missing entries are (FLT_MAX, -1) for L2 and (-FLT_MAX, -1) for IP.
 *
 * Parity pins for every function are listed in tests/test_oracle.py.
 */
#ifndef HIVF_ORACLE_H
#define HIVF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_METRIC_IP = 0, ORC_METRIC_L2 = 1 }; /* src/common.h:20-21 */

/* Planted-mistake switch: tests set it to show that the pins catch a wrong
 * oracle.  0 = correct.  1 = drop the last dimension of every distance.
 * 2 = order equal distances by descending id.  3 = prune with `>=` instead
 * of `>`.  4 = skip the last probed list.  5 = warm-up bound uses one row
 * fewer per list.  6 = the merge ignores the last partition. */
void orc_set_mutation(int m);
int orc_get_mutation(void);

/* One distance, canonical FMA chain (src/utils.h:480-487, 512-519). */
float orc_distance(int metric, const float* a, const float* b, int64_t d);

/* Streaming top-k over n (value, id) pairs visited in the given order, with
 * strict admission (src/IVFScan.hpp:131) into a k-heap ordered by (value,
 * id), then sorted (src/heap.hpp:73-90).  Output k entries, padded with
 * sentinels. */
void orc_topk(int metric, int64_t n, const float* val, const int64_t* ids,
              int64_t k, float* out_val, int64_t* out_ids);

/* Coarse quantizer: for each query the nprobe best centroids, sorted
 * (src/Index.cpp:967-995).  nprobe is clamped to nlist
 * (src/Index.cpp:1646-1648); outputs are [nq x nprobe_clamped]. */
int64_t orc_coarse(int metric, int64_t nq, const float* q, int64_t nlist,
                   const float* cent, int64_t d, int64_t nprobe,
                   float* out_dist, int64_t* out_list);

/* Nearest centroid of each row, lowest id on exact ties
 * (src/Index.cpp:567-579). */
void orc_assign(int metric, int64_t n, const float* x, int64_t nlist,
                const float* cent, int64_t d, float* out_dist,
                int64_t* out_list);

/* List construction (src/Index.cpp:581-698): assign every row, then fill
 * each list in ascending (L2) / descending (IP) order of the row's distance
 * to its centroid, equal distances by ascending row id.
 * Outputs list_off[nlist+1] (CSR offsets) and order[n] (row ids in list
 * order). */
void orc_build_lists(int metric, int64_t n, const float* x, int64_t nlist,
                     const float* cent, int64_t d, int64_t* list_off,
                     int64_t* order);

/* Index searched by the scans below, in CSR form: list l holds rows
 * [list_off[l], list_off[l+1]) of ids[] and codes[] (row-major, d floats). */
typedef struct {
    int metric;
    int64_t d;
    int64_t nlist;
    const float* centroids; /* [nlist x d] */
    const int64_t* list_off; /* [nlist + 1] */
    const int64_t* ids;      /* [N] */
    const float* codes;      /* [N x d] */
} orc_index;

typedef struct {
    int64_t candidates;     /* sum over queries of scanned rows */
    int64_t blocks_total;   /* (row, dim-block) pairs considered */
    int64_t blocks_skipped; /* pairs skipped by pruning */
} orc_stats;

/* Warm-up bound (src/Index.cpp:999-1051): the k-th best distance among the
 * first min(W, n_l) rows of the first L probed lists of each query, or the
 * sentinel when fewer than k rows were seen.  probes: [nq x nprobe]. */
void orc_warmup_tau(const orc_index* ix, int64_t nq, const float* q,
                    const int64_t* probes, int64_t nprobe, int64_t k,
                    int64_t L, int64_t W, float* tau);

/* Plain search (src/Index.cpp:1420-1554, 1638-1706): coarse quantizer, scan
 * of every probed list, sorted top-k.  Query-parallel over `threads` OpenMP
 * threads (src/Index.cpp:1667-1699); 0 = all.  Returns the clamped nprobe. */
int64_t orc_search(const orc_index* ix, int64_t nq, const float* q,
                   int64_t nprobe, int64_t k, float* D, int64_t* I,
                   int threads, orc_stats* st);

/* Dimension-block pruned scan over given probes (L2 only;
 * src/node.cpp:606-678, 1056-1174): the running distance is checked after
 * every block of dim_block dims except the last and the row is dropped when
 * it exceeds the bound.  bound = tau[q] (the reference's heapTops), and when
 * use_running_kth is set, min(tau[q], current k-th distance).  tau may be
 * NULL (no external bound). */
void orc_search_pruned(const orc_index* ix, int64_t nq, const float* q,
                       const int64_t* probes, int64_t nprobe, int64_t k,
                       int64_t dim_block, const float* tau,
                       int use_running_kth, float* D, int64_t* I, int threads,
                       orc_stats* st);

/* Scan over given probes without pruning (same as orc_search after the
 * coarse step). */
void orc_scan(const orc_index* ix, int64_t nq, const float* q,
              const int64_t* probes, int64_t nprobe, int64_t k, float* D,
              int64_t* I, int threads, orc_stats* st);

/* Lloyd k-means (src/Clustering.cpp:93-152): `niter` rounds of nearest
 * centroid assignment (orc_assign) then centroid = mean of its rows,
 * accumulated in double in row order and rounded to float; empty clusters
 * keep their centroid.  cent holds the initial centroids on entry. */
void orc_kmeans(int metric, int64_t n, const float* x, int64_t nlist,
                int64_t d, int niter, float* cent);

/* Top-k merge of per-partition results (src/Index.cpp:1327-1343): for
 * each query, the entries of partitions 0..parts-1 (each k_in entries, laid
 * out [parts][nq][k_in]) are offered in partition order to a k-heap with
 * strict admission, then sorted.  Entries with a negative id (the sentinel
 * padding of a partial result) are skipped. */
void orc_merge(int metric, int64_t nq, int64_t parts, int64_t k_in,
               const float* D_parts, const int64_t* I_parts, int64_t k,
               float* D, int64_t* I);

/* recall@k as in src/utils.h:661-686: |I n GT| / (nq k), a row stops at
 * its first -1. */
double orc_recall(const int64_t* I, const int64_t* GT, int64_t nq, int64_t k,
                  int64_t gt_k);

#ifdef __cplusplus
}
#endif

#endif
