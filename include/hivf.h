/*
 * hivf -- MI355X-native IVF-Flat query path behind a C ABI.
 *
 * Drop-in boundary for the reference's harmony::Index (src/Index.h:15-93):
 * one opaque handle per index, caller-owned query/result buffers, status
 * codes instead of exceptions.  Plain pointers and sizes only.
 *
 * Conventions (src/heap.hpp:13-90, src/Index.cpp:1652):
 *   - results are row-major D[n*k] float32 and I[n*k] int64, ascending
 *     distance for L2 and descending for IP, equal distances by ascending id;
 *   - missing entries are (FLT_MAX, -1) for L2 and (-FLT_MAX, -1) for IP;
 *   - a distance is one sequential FP32 FMA chain over dims 0..d-1
 *     (L2: sum (q_i - x_i)^2, IP: sum q_i x_i), src/utils.h:480-519.
 *
 * Threading: one search in flight per handle (the reference's
 * Index::search is not reentrant either, src/Index.h:142-147); separate
 * handles may be used concurrently.  Every call returns 0 or a negative
 * HIVF_E* code; hivf_last_error() then describes the failure.
 */
#ifndef HIVF_H
#define HIVF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MetricType, src/common.h:20-21 */
#define HIVF_METRIC_IP 0
#define HIVF_METRIC_L2 1

#define HIVF_OK 0
#define HIVF_EINVAL (-1)      /* bad argument or shape */
#define HIVF_ENOMEM (-2)      /* device or host allocation failed */
#define HIVF_EHIP (-3)        /* HIP runtime error */
#define HIVF_ESTATE (-4)      /* call out of order (e.g. search before add) */
#define HIVF_EUNSUPPORTED (-5)/* combination not implemented */
#define HIVF_EIO (-6)         /* file read/write failed */

/* Scan algorithm. */
#define HIVF_ALGO_PLAIN 0   /* per-query scan, full distances (src/Index.cpp:1420-1554) */
#define HIVF_ALGO_PRUNED 1  /* per-query scan, dimension-block pruning (src/node.cpp:606-678) */
#define HIVF_ALGO_BATCHED 2 /* list-major batched scan with MFMA contractions */

/* Source of the per-query pruning bound. */
#define HIVF_TAU_NONE 0    /* running k-th only */
#define HIVF_TAU_WARMUP 1  /* warm-up bound, src/Index.cpp:999-1051 */
#define HIVF_TAU_LEARNED 2 /* learned bound from the attached model (exactness kept by fallback) */

/* Arithmetic of the batched scan's contractions.  Results are the FP32
 * path's either way (ids and distance bits); only the speed differs.
 * FP32: the batched scan picks its form -- exact int8 MFMA when every list
 * value and every query is an integer in [0, 255] and d <= 256, else the
 * BF16 shadow with a proven error margin and exact FP32 rescoring of the
 * candidates inside it.  The per-query scans always compute in FP32.
 * BF16: the batched scan uses the BF16 shadow even where the int8 form
 * applies (EUNSUPPORTED for the per-query scans). */
#define HIVF_PREC_FP32 0
#define HIVF_PREC_BF16 1

typedef struct hivf_index hivf_index;

typedef struct {
    int64_t nprobe;                  /* <= 0: the index default (hivf_index_set_nprobe) */
    const int32_t* nprobe_per_query; /* optional [n] per-query probe counts, host memory */
    int32_t algo;                    /* HIVF_ALGO_* */
    int32_t tau_source;              /* HIVF_TAU_* */
    int32_t precision;               /* HIVF_PREC_* */
    int32_t dim_block;               /* dims per pruning block; 0 = d/4 (src/node.cpp:640) */
    int32_t warmup_lists;            /* L, default 10 (src/query.cpp:845) */
    int32_t warmup_list_size;        /* W, 0 = k (src/query.cpp:847) */
    int32_t survivor_cap;            /* batched: candidates kept per query before its exact rescan; 0 = 16384;
                                        rounded up to an even count */
    int32_t batch_queries;           /* batched: queries per work item (32, 64, 128); 0 = tuned/default */
    int32_t batch_rows;              /* batched: list rows per work item (multiple of 32); 0 = tuned/default */
    int32_t batch_waves;             /* batched: resident waves per CU of the persistent launch; 0 = tuned/default */
    int32_t use_graph;               /* search_device only, no stats: capture the search in a hipGraph on the
                                        first call of a shape and replay it on later calls with the same
                                        shape and buffers */
    int32_t reserved;
    const int32_t* nprobe_per_query_dev; /* search_device only: optional [n] per-query probe counts in device
                                        memory, read by the search on its stream (no host copy, so a captured
                                        search replays with new counts); nprobe must be set and bounds them:
                                        a count above it reads as nprobe, one below 1 probes nothing */
} hivf_search_params;

typedef struct {
    double total_ms;       /* device time of the whole search */
    double coarse_ms;      /* coarse quantizer */
    double tau_ms;         /* bound (warm-up scan or model) */
    double scan_ms;        /* list scan (dominant kernel) */
    double select_ms;      /* final selection / merge */
    double fallback_ms;    /* exact rescans after a failed learned bound */
    int64_t candidates;    /* sum over queries of scanned rows */
    int64_t blocks_total;  /* (row, dim-block) pairs of the pruned scan */
    int64_t blocks_skipped;/* pairs skipped by pruning */
    int64_t survivors;     /* batched scan: rows kept for the final selection */
    int64_t fallback_queries; /* batched scan: queries its first pass could not finish (rescanned) */
    int64_t scan_launches; /* launches of the dominant kernel in this search */
    int64_t exact_rescans; /* batched scan: queries left to the per-query exact scan */
    int64_t batch_queries; /* batched scan: tile shape used (queries, rows, waves per CU) */
    int64_t batch_rows;
    int64_t batch_waves;
    double plan_ms;        /* batched scan: probe inversion and work-item planning */
    int64_t scan_kind;     /* batched scan that ran: 0 BF16 with error margin, 1 BF16 exact,
                              2 int8 (integer data) */
    int64_t overflow_queries; /* batched scan: of fallback_queries, those whose survivors
                                 overflowed survivor_cap (the rest: learned bound too low) */
} hivf_search_stats;

/* Fill params with defaults (plain scan, index nprobe). */
void hivf_search_params_default(hivf_search_params* p);

/* Index(d, nlist, nprobe, metric) ctor, src/Index.h:15-25 / src/Index.cpp:24-43.
 * Returns NULL on bad arguments or when no GPU is usable. */
hivf_index* hivf_index_create(int64_t d, int64_t nlist, int metric);
void hivf_index_destroy(hivf_index* h);
/* Last error message of this handle ("" when none); NULL handle: global. */
const char* hivf_last_error(const hivf_index* h);

/* Public nprobe field, src/Index.h:101; clamped to nlist at search time
 * (src/Index.cpp:1646-1648). */
int hivf_index_set_nprobe(hivf_index* h, int64_t nprobe);

/* Index::train, src/Index.cpp:381-425: Lloyd k-means (niter rounds) from
 * the given initial centroids over the n training rows; the assignment
 * step runs on the GPU.  init_centroids: [nlist x d]. */
int hivf_index_train(hivf_index* h, int64_t n, const float* x,
                     const float* init_centroids, int32_t niter);
/* Index::train's default recipe (src/Index.cpp:381-425 with
 * src/Clustering.cpp:14-152): above nlist * max_points_per_centroid rows
 * (<= 0: 256) the rows are subsampled, then nlist sample rows seed the
 * centroids, both picked by std::shuffle under
 * std::default_random_engine(seed) exactly as the reference does (the
 * reference passes seed 6666 and 20 rounds, 2 with `lite`); then niter
 * Lloyd rounds as hivf_index_train. */
int hivf_index_train_kmeans(hivf_index* h, int64_t n, const float* x, int32_t niter,
                            uint64_t seed, int32_t max_points_per_centroid);
/* The rows that recipe picks, without a device: sample_rows[*n_sample]
 * (row numbers of x, in sample order; *n_sample = min(n, nlist * max)) and
 * init_rows[nlist] (positions in the sample). */
int hivf_kmeans_sample(int64_t n, int64_t nlist, int32_t max_points_per_centroid, uint64_t seed,
                       int64_t* sample_rows, int64_t* n_sample, int64_t* init_rows);
/* Inject centroids trained elsewhere (e.g. by the reference). */
int hivf_index_set_centroids(hivf_index* h, const float* centroids);
int hivf_index_get_centroids(const hivf_index* h, float* centroids);

/* Index::add, src/Index.cpp:581-698: assigns every row to its nearest
 * centroid on the GPU and fills each list in ascending (L2) / descending
 * (IP) order of centroid distance, equal distances by row id; ids are row
 * numbers.  Replaces all lists (src/Index.cpp:636-639). */
int hivf_index_add(hivf_index* h, int64_t n, const float* x);

/* Load lists verbatim: list l holds rows [list_off[l], list_off[l+1]) of
 * ids[] / codes[] (row-major, d floats).  ids must lie in [0, 2^32 - 2]. */
int hivf_index_set_lists(hivf_index* h, const int64_t* list_off,
                         const int64_t* ids, const float* codes);
int64_t hivf_index_ntotal(const hivf_index* h);
/* Shape of the index: d, nlist, metric (any pointer may be NULL). */
int hivf_index_info(const hivf_index* h, int64_t* d, int64_t* nlist, int* metric);
int hivf_index_list_size(const hivf_index* h, int64_t list, int64_t* size);
/* Copy list `list` back in list order: ids[size], codes[size*d]. */
int hivf_index_get_list(const hivf_index* h, int64_t list, int64_t* ids,
                        float* codes);

/* Index::search, src/Index.h:82 / src/Index.cpp:1638-1706.  Host buffers:
 * x[n*d], D[n*k], I[n*k].  params / stats may be NULL. */
int hivf_index_search(hivf_index* h, int64_t n, const float* x, int64_t k,
                      float* D, int64_t* I, const hivf_search_params* params,
                      hivf_search_stats* stats);

/* Same with device buffers already resident in HBM, enqueued on `stream`
 * (a hipStream_t; NULL = the handle's own stream).  Returns once the work
 * is enqueued unless stats are requested, in which case it waits. */
int hivf_index_search_device(hivf_index* h, int64_t n, const float* x_dev,
                             int64_t k, float* D_dev, int64_t* I_dev,
                             const hivf_search_params* params,
                             hivf_search_stats* stats, void* stream);

/* ---- search without device allocation -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph stream capture, tight serving loops, harnesses
 * that forbid hipMalloc): the per-query search (algo PLAIN or PRUNED, FP32,
 * warm-up bound or none; coarse quantizer by the per-query kernel) whose
 * scratch -- probe lists and coarse distances, warm-up bounds, padded
 * queries when d % 4 != 0 or x_dev is not 16-byte aligned -- lives in one
 * fixed device array of the library (16.25 MiB per GPU, held until the
 * process ends).  Limits per call: n <= 65536 queries, n * P <= 2^20
 * (query, probe) pairs (P = nprobe clamped to nlist, <= 1024), and
 * n * (d rounded up to 4) <= 2^21 floats when the queries are staged; larger
 * calls return HIVF_EINVAL, other algorithms HIVF_EUNSUPPORTED.  Calls no
 * hipMalloc / hipFree.  Every call on a GPU shares that GPU's array and
 * leaves its contents there: calls that overlap on one GPU (other streams or
 * threads) race.  Per-query probe counts only from device memory
 * (nprobe_per_query_dev).  Same results as hivf_index_search_device. */
int hivf_index_search_device_noalloc(hivf_index* h, int64_t n, const float* x_dev, int64_t k,
                                     float* D_dev, int64_t* I_dev, const hivf_search_params* params,
                                     void* stream);

/* ---- top-k merge across partitions --------------------------------------
 * src/Index.cpp:1327-1343 (Index::search_divide_ivf: the master merges the
 * per-worker heaps, distancesHeapBuffer[rank][q * k + i]).  Inputs are the
 * sorted or unsorted top-k_in rows of `parts` partitions with disjoint ids,
 * laid out [parts][n][k_in]; entries with a negative id are padding.  The
 * output row is the best k of their union, sorted, sentinel padded, equal
 * distances by ascending id (-0 and +0 rank equal; a zero comes back as +0).
 * k <= 1024; ids below 2^32 - 1 (the host entry point checks; the device one
 * assumes it).  No handle: the current device.
 * Errors: hivf_last_error(NULL). */
int hivf_merge_topk(int metric, int64_t n, int64_t parts, int64_t k_in, const float* D_parts,
                    const int64_t* I_parts, int64_t k, float* D, int64_t* I);
int hivf_merge_topk_device(int metric, int64_t n, int64_t parts, int64_t k_in, const float* D_parts,
                           const int64_t* I_parts, int64_t k, float* D, int64_t* I, void* stream);

/* ---- tile-shape autotuner ----------------------------------------------
 * No counterpart in the reference.  Times the batched scan on n device
 * queries for every candidate (queries per item, rows per item, waves per
 * CU), keeps the fastest as this index's default (used when a search leaves
 * those parameters 0) and reports it. */
typedef struct {
    int32_t batch_queries;
    int32_t batch_rows;
    int32_t batch_waves;
    int32_t tried;   /* candidates timed */
    double ms;       /* mean device time of one search with the best shape */
    double ms_default; /* same with the built-in default shape */
    int32_t reserved[4];
} hivf_tune_result;

int hivf_index_autotune(hivf_index* h, int64_t n, const float* x_dev, int64_t k,
                        const hivf_search_params* base, int32_t reps, hivf_tune_result* out);
/* Set the default tile shape (e.g. from a saved tuning table); 0 = built-in. */
int hivf_index_set_tune(hivf_index* h, int32_t batch_queries, int32_t batch_rows,
                        int32_t batch_waves);

/* Coarse quantizer alone (src/Index.cpp:967-995): host buffers,
 * dist[n*P], lists[n*P] with P = min(nprobe, nlist). */
int hivf_index_coarse(hivf_index* h, int64_t n, const float* x, int64_t nprobe,
                      float* dist, int64_t* lists);

/* Warm-up bound alone (src/Index.cpp:999-1051) over caller-given probes
 * [n x nprobe]: tau[n]. */
int hivf_index_warmup_tau(hivf_index* h, int64_t n, const float* x,
                          const int64_t* probes, int64_t nprobe, int64_t k,
                          int64_t L, int64_t W, float* tau);

/* Scan alone over caller-given probes [n x nprobe] (src/Index.cpp:1432-1549,
 * src/node.cpp:606-678): algo PLAIN or PRUNED; tau (optional, [n]) is the
 * external bound of the pruned scan. */
int hivf_index_scan(hivf_index* h, int64_t n, const float* x,
                    const int64_t* probes, int64_t nprobe, int64_t k,
                    const float* tau, const hivf_search_params* params,
                    float* D, int64_t* I, hivf_search_stats* stats);

/* ---- learned pruning bound --------------------------------------------
 * No counterpart in the reference (its bound is the warm-up scan,
 * src/Index.cpp:999-1051).  An MLP maps a query's 32 nearest coarse
 * distances and its norm (33 features, see hivf_index_tau_traces) to a
 * predicted k-th neighbour distance; the batched scan uses it as the bound
 * and rescans exactly every query whose prediction was too low. */
typedef struct hivf_tau_model hivf_tau_model;

typedef struct {
    int32_t epochs;        /* passes over the training split */
    int32_t batch;         /* rows per step; 0 = 1024 */
    float lr;              /* Adam learning rate */
    float quantile;        /* pinball-loss quantile of log1p(k-th distance) */
    float miss_rate;       /* target fraction of held-out bounds below the k-th (offset calibration); 0 = 0.01 */
    int32_t holdout_every; /* every n-th trace is held out for calibration; 0 = 10 */
    uint64_t seed;
    float target_scale;    /* distance scale s of the targets y = log1p(D_k / s), stored in the model
                              (tau = s max(expm1(y + offset), 0)); <= 0: 1 */
    int32_t reserved[3];
} hivf_train_params;

typedef struct {
    double final_loss;           /* mean pinball loss of the last batch */
    double holdout_miss_rate;    /* held-out fraction with bound < k-th distance */
    double holdout_mean_abs_err; /* mean |prediction - log1p(k-th)| before the offset */
    float offset;                /* calibrated offset added to predictions */
    int64_t steps;
    int64_t holdout;
    int64_t reserved[4];
} hivf_train_stats;

/* in_dim must be 33 for use as a search bound; parameters He-initialised from seed. */
hivf_tau_model* hivf_tau_model_create(int32_t in_dim, int32_t hidden, uint64_t seed);
void hivf_tau_model_destroy(hivf_tau_model* m);
int64_t hivf_tau_model_num_params(const hivf_tau_model* m);
/* Flat parameters: W1[in*h], b1[h], W2[h*h], b2[h], w3[h], b3, mu[in], sd[in], scale, offset
 * (set_params rejects a scale that is not finite and positive). */
int hivf_tau_model_get_params(const hivf_tau_model* m, float* out);
int hivf_tau_model_set_params(hivf_tau_model* m, const float* in);
/* y[n] = network output (prediction of log1p(k-th distance / scale), no offset) for raw features x[n*in]. */
int hivf_tau_model_forward(hivf_tau_model* m, int64_t n, const float* x, float* y);
/* grad = d(sum_i dy[i] * y[i]) / d(parameters), flat layout (zeros for mu, sd, offset). */
int hivf_tau_model_grad(hivf_tau_model* m, int64_t n, const float* x, const float* dy, float* grad);
/* Train on traces (features x[n*in], targets y[n] = log1p(k-th distance / p->target_scale)) on the GPU. */
int hivf_tau_model_train(hivf_tau_model* m, int64_t n, const float* x, const float* y,
                         const hivf_train_params* p, hivf_train_stats* stats);
/* Attach a copy of the model to the index (NULL detaches). */
int hivf_index_set_tau_model(hivf_index* h, const hivf_tau_model* m);
/* Exact search of n queries, returning their bound-model features
 * features[n*33] and k-th distances kth[n] (FLT_MAX if fewer than k). */
int hivf_index_tau_traces(hivf_index* h, int64_t n, const float* x, int64_t k,
                          int64_t nprobe, float* features, float* kth);
/* The learned bound tau[n] the batched search would use for these queries:
 * the attached model on their coarse distances, tau = scale max(expm1(y + offset), 0)
 * (no counterpart in the reference; for inspecting a model). */
int hivf_index_tau_bounds(hivf_index* h, int64_t n, const float* x, int64_t k,
                          int64_t nprobe, float* tau);

/* Index::save_index / load_index file format (src/Index.cpp:1708-1770,
 * src/IVF.cpp:76-133): header (size_t d, nlist; int metric, opt level;
 * size_t sub_k, sub_nlist, sub_nprobe), float centroids[nlist*d], then per
 * non-empty list: size_t list_id and an IVF record (size_t n, d, sub_k; int
 * opt; size_t ids[n]; float codes[n*d]; float norms[n]; optional arrays of
 * the opt level, skipped on load).  Load returns NULL and sets *status on
 * failure. */
int hivf_index_save_harmony(const hivf_index* h, const char* path);
hivf_index* hivf_index_load_harmony(const char* path, int* status);

/* ---- serve front-end -----------------------------------------------------
 * The reference's TCP protocol (src/master_server.h:5-8,
 * src/master_server.cpp:105-197, src/README.md:24-44), little-endian:
 *   INSERT phase (the first connection, when insert_phase is set):
 *     [u8 1][u64 n][n*d f32] -> [u8 status]   rows are buffered
 *     [u8 2]                 -> [u8 status]   build: k-means with the
 *                                             reference's recipe when the index
 *                                             has no centroids, then add
 *     [u8 0]                                  ends the INSERT phase unbuilt
 *   QUERY phase (every later connection, one client at a time):
 *     [u64 nq][u64 k][nq*d f32] -> [nq*k f32 distances][nq*k i64 labels]
 *     [u64 0][u64 0]            -> the server stops (hivf_server_wait returns)
 *   status: 0 OK, 1 error.  A request with k outside [1, 1024] or more than
 *   max_request_rows rows closes its connection.
 * One listener thread serves the connections.  Each request is searched as
 * batches of `batch` rows (the last one padded), so the search runs as one
 * captured hipGraph replayed per batch.  Stop the server before destroying
 * its index.  No reference counterpart for the function-pointer form: it
 * runs a caller's search (for example a list-sharded one) behind the same
 * transport and batching. */
typedef struct hivf_server hivf_server;

typedef struct {
    int64_t batch;            /* rows per captured search; 0 = 4096 */
    int32_t insert_phase;     /* 1: the first connection speaks the INSERT protocol */
    int32_t train_niter;      /* k-means rounds at BUILD_DONE; 0 = 20 (src/Index.cpp:389) */
    int64_t max_request_rows; /* rows per INSERT batch or query request; 0 = 2^24 */
    hivf_search_params search;/* the search of every batch (use_graph is forced on) */
    int32_t reserved[4];
} hivf_serve_params;

typedef struct {
    int64_t connections;   /* closed connections */
    int64_t requests;      /* query requests answered */
    int64_t queries;       /* query rows answered */
    int64_t batches;       /* batch searches run */
    int64_t rows_inserted; /* rows received in the INSERT phase */
    int64_t reserved[3];
} hivf_server_stats;

typedef int (*hivf_search_fn)(void* user, int64_t n, const float* x, int64_t k, float* D, int64_t* I);
typedef int (*hivf_build_fn)(void* user, int64_t n, const float* x);

/* params: defaults of hivf_serve_params_default when NULL.  address: IPv4
 * text ("0.0.0.0", "127.0.0.1"); port 0 picks a free port
 * (hivf_server_port).  NULL + *status on failure. */
void hivf_serve_params_default(hivf_serve_params* p);
hivf_server* hivf_server_start(hivf_index* h, const char* address, int port,
                               const hivf_serve_params* params, int* status);
hivf_server* hivf_server_start_fn(int64_t d, const char* address, int port, int64_t batch,
                                  int insert_phase, hivf_search_fn search, hivf_build_fn build,
                                  void* user, int* status);
int hivf_server_port(const hivf_server* s);
/* Blocks until a client sent the shutdown header (or the server was
 * stopped); returns the last error status of the serving thread (0 = none). */
int hivf_server_wait(hivf_server* s);
int hivf_server_stats_get(hivf_server* s, hivf_server_stats* out);
/* Message of the last serving error ("" when none); valid until the next error. */
const char* hivf_server_error(hivf_server* s);
/* Closes the listener and any open connection, joins the thread, frees s. */
void hivf_server_stop(hivf_server* s);

#ifdef __cplusplus
}
#endif

#endif /* HIVF_H */
