/*
 * harmony_ivf.h — C-ABI of the MI355X IVF-flat search library (libharmony_amd.so).
 *
 * This is the drop-in boundary for Harmony's per-node IVF-flat search path. The reference
 * exposes it as the C++ class harmony::Index (src/Index.h:13-149) plus the scanner operator
 * interface IVFScanBase (src/IVFScan.hpp:15-65); each entry point below names the member it
 * replaces. Conventions:
 *   - no exceptions cross the ABI; every call returns an int status (HV_OK == 0) and the
 *     message of the last failure is available from hv_last_error();
 *   - host pointers are caller-owned, read/written synchronously inside the call;
 *   - device memory is owned by the handle; one HIP stream per handle; one search at a time
 *     per handle (the reference's Index::search is not re-entrant either, Index.h:142-147);
 *   - results are ordered by (distance, id), padded with (FLT_MAX, -1) (heap.hpp:13-25).
 *   - ids are the insertion positions 0..ntotal-1 (Index.cpp:663); ntotal < 2^32 - 1.
 */
#ifndef HARMONY_IVF_H
#define HARMONY_IVF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HV_OK 0
#define HV_ERR_INVALID_ARG (-1) /* bad size / pointer / parameter */
#define HV_ERR_STATE (-2)       /* call out of order (search before add, ...) */
#define HV_ERR_DEVICE (-3)      /* HIP runtime failure */
#define HV_ERR_IO (-4)          /* file open / read / format */
#define HV_ERR_CAPACITY (-5)    /* size beyond a documented limit */
#define HV_ERR_NO_DEVICE (-6)   /* no GPU visible */

/* MetricType (src/common.h:20-21). METRIC_IP (Index.cpp:408-418, 567-671, 1440-1551): k-means
 * with normalized assignment and normalized centroids, lists ordered by descending inner product
 * with their centroid, and the FP32 search: the nprobe largest <q, c>, then the exact top-k by
 * (inner product descending, id ascending), padded with (-FLT_MAX, -1). With OPT_TRIANGLE both at
 * add and at search, each list is scanned only over the angular band a result can lie in (the
 * reference's triangle pruning, Index.cpp:1509-1546); on unit vectors that changes no result.
 * The IP search takes the FP32 form with one nprobe and one k per call: no BF16 / FP8 forms,
 * learned predictor, external bounds, per-query nprobe or k, ground truth or no-allocation
 * variant (those return HV_ERR_INVALID_ARG). */
#define HV_METRIC_IP 0
#define HV_METRIC_L2 1

/* OptLevel bits (src/common.h:23-32). The L2 search ignores them, as the reference does
 * (its TRIANGLE / SUBNN code in IVFScan.hpp:223-428 is commented out); they are kept so the
 * opt-level subset check of Index::search (Index.cpp:1642-1645) behaves the same. */
#define HV_OPT_NONE 0
#define HV_OPT_TRIANGLE 1
#define HV_OPT_SUBNN_L2 2
#define HV_OPT_SUBNN_IP 4
#define HV_OPT_ALL 7

/* bound modes for the pruning bound */
#define HV_BOUND_WARMUP 0   /* Index::warmUpSearch (Index.cpp:999-1051) */
#define HV_BOUND_EXTERNAL 1 /* caller-provided per-query bound (--fullWarmUp, query.cpp:716-725); FP32 and
                               BF16 forms only: a search with it and HV_PRECISION_FP8 or learned = 1
                               returns HV_ERR_INVALID_ARG */

/* precision of the candidate distances */
#define HV_PRECISION_FP32 0 /* FP32 blocked partial sums on every candidate (reference arithmetic) */
#define HV_PRECISION_BF16 1 /* BF16 MFMA filter with a certified error bound, then the FP32 distance
                               of each survivor: same results as HV_PRECISION_FP32 */
#define HV_PRECISION_FP8 2  /* FP8 (e4m3, power-of-two scale per list and per query slot) MFMA
                               estimates; the k * rerank_r smallest per query are re-ranked with the
                               FP32 distance (approximate: recall below 1 is possible, the returned
                               distances are exact) */

typedef struct hv_index hv_index_t;

/* k-means parameters (ClusteringParameters, src/Clustering.h:13-18). */
typedef struct {
    int64_t niter;                   /* Lloyd iterations (reference Index::train uses 20) */
    uint64_t seed;                   /* shuffle seed for subsample + init */
    int64_t max_points_per_centroid; /* subsample cap = nlist * this (0 = no cap) */
} hv_kmeans_params;

/* Search parameters: the mutable knobs of harmony::Index (nprobe, opt_level) and of
 * Index::Param (pruning, fullWarmUp, heapTops; src/Index.h:48-68), plus the preSearch
 * warm-up sizes (src/Index.h:90) and the per-node dimension-block count (node.cpp:639). */
typedef struct {
    int64_t nprobe;           /* clamped to nlist (Index.cpp:1646-1648) */
    int32_t opt_level;        /* must be a subset of the added opt level */
    int32_t dim_blocks;       /* B >= 1; blocks of floor(d/B) dims, last takes the rest */
    int32_t pruning;          /* 1: drop a candidate once a partial sum exceeds its bound */
    int32_t bound_mode;       /* HV_BOUND_* */
    int64_t warmup_list;      /* lists used for the warm-up bound (default 10) */
    int64_t warmup_list_size; /* vectors per warm-up list (0 -> k, query.cpp:328-333) */
    int32_t precision;        /* HV_PRECISION_* */
    int32_t refine_bound;     /* 1: after each query's nearest list, lower its bound to the running
                                 k-th distance (min with the warm-up bound; results unchanged) */
    int64_t survivor_cap;     /* per-query candidate buffer (0 = automatic) */
    int32_t tile_shape;       /* list-scan tile shape of the remaining-lists phase: 0 = automatic,
                                 s > 0 = compiled shape s-1 */
    int32_t tile_shape_first; /* same, for the nearest-list phase */
    /* learned probe/bound predictor (no counterpart in the reference; needs a trained model) */
    int32_t learned;              /* 1: per-query nprobe and pruning bound from the predictor */
    float learned_nprobe_scale;   /* probes = ceil(predicted nprobe * scale), at most nprobe, and at
                                     least enough lists to hold k vectors */
    float learned_bound_scale;    /* seeded bound = predicted k-th distance * scale (<= 0: none) */
    int32_t learned_min_nprobe;   /* lower limit of the per-query probe count */
    /* FP8 form (HV_PRECISION_FP8): shortlist width factor r, k * r <= 4096 (default 4) */
    int32_t rerank_r;
    /* 1: capture the device work of a search as a HIP graph on the second call with the same
     * shape, buffers and parameters, and replay it from then on (one launch per search; phase
     * times are then reported as one figure in ms_scan). Not with HV_BOUND_EXTERNAL. */
    int32_t graph;
    /* vectors per work item of the MFMA filter scans (a multiple of 16; 0 = 256) */
    int32_t scan_vsplit;
    /* learned search, FP32 and BF16 forms: a query's probe count grows past the prediction to every
     * list of its first nprobe whose lower bound sqrt(|q - c|^2) - radius (radius = the list's
     * largest |x - c|) is within sqrt(margin * seeded bound). 0: the prediction alone; >= 1 (default
     * 1): with the search's check that the k-th result is <= the seed (else the query is redone
     * over all nprobe lists), every learned result equals the plain search's at nprobe. */
    float learned_margin;
    /* BF16 form, k <= 128: 1 (default) scans each query's nearest list in one pass that keeps the
     * query's candidates in on-chip memory and reduces them to the k-th smallest upper bound
     * before anything is written; 0 writes every pair of the nearest list out and reduces it in a
     * second kernel. Same results either way. */
    int32_t nearest_pass;
} hv_search_params;

/* Per-search report (what the reference's Stats + query.cpp timing report, stats.h:14-134). */
typedef struct {
    double ms_total;         /* host wall time of the call */
    double ms_upload;        /* host->device queries (hv_search only) */
    double ms_coarse;        /* coarse distances + nprobe select */
    double ms_warmup;        /* warm-up bound */
    double ms_group;         /* (query, list) grouping + work plan */
    double ms_scan;          /* list scan */
    double ms_select;        /* per-query top-k selection */
    double ms_download;      /* device->host results (hv_search only) */
    int64_t candidates;      /* sum over queries of the probed list sizes (C_q) */
    int64_t block_evals;     /* dimension blocks evaluated (executed work, wave granularity) */
    int64_t block_evals_max; /* candidates * B (unpruned work) */
    int64_t survivors;       /* candidates that reached the top-k selection */
    int64_t overflow_queries;/* queries whose survivors exceeded the buffer (exact fallback) */
    int64_t work_items;      /* scan work items (list x query tile) */
    double ms_main;          /* the dominant kernel: the scan of the remaining probed lists */
    int64_t main_pairs;      /* (query, vector) pairs of that kernel (unpruned) */
    int64_t fallback_queries;/* learned search: queries whose seeded bound was too low, redone exactly */
    double ms_rerank;        /* BF16 form: the exact FP32 re-rank kernel */
    int64_t rerank_pairs;    /* pairs it re-ranked */
    double ms_scan_kernels;  /* summed duration of every candidate-scan kernel launch (the FP32 list
                                scan, warm-up included, or the MFMA filter scan), eager searches only */
    int64_t scan_launches;   /* those launches */
    int64_t scan_pairs;      /* (query, vector) pairs they covered, unpruned (warm-up included);
                                METRIC_IP: the inner products computed (after the triangle cut) */
    int64_t widened_queries; /* learned search: queries whose probe count the margin test widened */
} hv_stats;

/* ---- lifetime (Index::Index, Index.h:15-25) ---- */
hv_index_t* hv_create(int64_t d, int64_t nlist, int32_t metric, int32_t opt_level);
void hv_destroy(hv_index_t* h);
const char* hv_last_error(const hv_index_t* h); /* h may be NULL: last hv_create failure */

/* Dimension order applied to every vector on entry (blocks then follow this order). perm is
 * a permutation of 0..d-1; NULL restores the identity. Must precede hv_train. */
int hv_set_dim_order(hv_index_t* h, const int32_t* perm);

/* ---- build (Index::train / Index::add, Index.h:69,80) ---- */
void hv_default_kmeans_params(hv_kmeans_params* p);
int hv_train(hv_index_t* h, int64_t n, const float* x, const hv_kmeans_params* p);
int hv_set_centroids(hv_index_t* h, const float* centroids); /* nlist*d, in caller dim order */
int hv_add(hv_index_t* h, int64_t n, const float* x);

/* ---- search (Index::search, Index.h:82; preSearch knobs Index.h:90) ---- */
void hv_default_search_params(hv_search_params* p);
int hv_set_search_params(hv_index_t* h, const hv_search_params* p);
int hv_get_search_params(const hv_index_t* h, hv_search_params* p);
/* per-query external bound for HV_BOUND_EXTERNAL (heapTops, Index.h:60) */
int hv_set_external_bounds(hv_index_t* h, int64_t n, const float* bounds);
/* per-query probe counts for the next searches of n queries (each clamped to [1, nprobe]; the
 * search still selects nprobe lists per query and scans the first nprobe_q of them); n = 0
 * clears. No counterpart in the reference (one nprobe for every query, Index.h:117); not
 * combined with the learned predictor, which sets its own counts. */
int hv_set_query_nprobe(hv_index_t* h, int64_t n, const int32_t* nprobe);
/* per-query k for the next searches of n queries (each clamped to [1, k] of the search); n = 0
 * clears. Row q of D / I then holds the query's k_q results, padded with (FLT_MAX, -1) after them:
 * the same as a search of that query alone at k = k_q. Bounds are taken at each query's own k.
 * No counterpart in the reference (one k per call, Index.h:82); not combined with the learned
 * predictor (trained for one k). */
int hv_set_query_k(hv_index_t* h, int64_t n, const int32_t* k);
/* host buffers: q [n][d] f32 row-major, D [n][k] f32, I [n][k] int64 */
int hv_search(hv_index_t* h, int64_t n, const float* q, int64_t k, float* D, int64_t* I, hv_stats* st);
/* device buffers (already in HBM): same layouts */
int hv_search_device(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D_dev, int64_t* I_dev,
                     hv_stats* st);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): hv_search_device with every
 * scratch buffer of the search carved from a fixed block of 128 MiB of device memory that the
 * library holds on each GPU it runs on (reserved when the library is loaded, kept until the process
 * ends), so the call allocates no device or pinned host memory. Same results as hv_search_device.
 * Limits: the batch's scratch must fit the block (about 90 KB per query at k = 100 with the default
 * 4096-entry survivor buffer, BF16 form; otherwise HV_ERR_CAPACITY before any work is queued:
 * search fewer queries per call or lower survivor_cap); graph mode is ignored; the FP8 form needs
 * its codes built by an earlier FP8 hv_search / hv_search_device (else HV_ERR_STATE);
 * hv_get_last_* report nothing afterwards. Calls on one GPU share its block: the library
 * serializes them (a call waits for the one in progress), and the block keeps the last call's
 * contents (a search clears what it needs cleared with its own kernels). */
int hv_search_device_noalloc(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D_dev,
                             int64_t* I_dev, hv_stats* st);

/* Exact kNN in FP64 over every vector of the index, ordered by (distance, id): the ground
 * truth recall is measured against (the reference computes it with faiss IndexIVFFlat at
 * nprobe = nlist, query.cpp:481-514). Runs the current search first to bound each query's
 * k-th distance, then a full FP64 scan keeps the candidates under the bound. Host buffers. */
int hv_ground_truth_f64(hv_index_t* h, int64_t n, const float* q, int64_t k, double* D, int64_t* I);

/* ---- persistence: the reference's index file (Index::save_index / load_index,
 * Index.cpp:1708-1770; per-list IVF record, IVF.cpp:76-134). Layout, little-endian:
 *   size_t d, size_t nlist, int32 metric, int32 added_opt_level, size_t sub_k, sub_nlist,
 *   sub_nprobe, float centroids[nlist][d], then per non-empty list: size_t list_id, size_t
 *   list_size, size_t d, size_t sub_k, int32 opt_level, size_t ids[list_size],
 *   float codes[list_size][d], float norms[list_size] (+ optional arrays per opt_level bits),
 *   read until end of file. hv_save writes per-list opt_level 0 (no optional arrays);
 *   hv_load skips optional arrays, replaces the handle's d / nlist with the file's, keeps each
 *   list's vector order, sets the search opt level to HV_OPT_NONE (Index.cpp:1743) and the added
 *   opt level to the file's. The file is parsed and size-checked completely before the handle
 *   changes: a truncated or malformed file returns HV_ERR_IO and leaves the handle as it was.
 *   Codes are stored in caller dimension order. */
int hv_save(const hv_index_t* h, const char* path);
int hv_load(hv_index_t* h, const char* path);

/* ---- introspection (what the reference keeps public in Index / IVF, Index.h:117-119) ---- */
int64_t hv_ntotal(const hv_index_t* h);
int64_t hv_dim(const hv_index_t* h);
int64_t hv_nlist(const hv_index_t* h);
int hv_get_centroids(const hv_index_t* h, float* out);       /* nlist*d, caller dim order */
int hv_get_list_offsets(const hv_index_t* h, int64_t* out);  /* nlist+1 */
int hv_get_list_ids(const hv_index_t* h, int64_t* out);      /* ntotal, list order */
int hv_get_list_codes(const hv_index_t* h, float* out);      /* ntotal*d, list order, caller dims */
/* METRIC_IP: each vector's inner product with its list's centroid, list order (the per-list
 * candidate2centroid arrays of src/IVF.h:21, filled at Index.cpp:662-664) */
int hv_get_list_c2c(const hv_index_t* h, float* out);        /* ntotal */
/* Introspection of the last search (listidqueries / heapTops, src/Index.h:142-147). Each call
 * takes the capacity of `out` in elements and returns HV_ERR_INVALID_ARG without writing when
 * it is below what the last search holds: hv_last_n(h) * hv_last_probe_stride(h) probes,
 * hv_last_n(h) bounds or counts. HV_ERR_STATE when no search has run since the last build. */
int64_t hv_last_n(const hv_index_t* h);                                     /* queries of the last search */
int hv_get_last_probes(const hv_index_t* h, int64_t* out, int64_t capacity); /* [n][stride] list ids */
int hv_get_last_bounds(const hv_index_t* h, float* out, int64_t capacity);   /* n pruning bounds */
/* per-query probe counts of the last search (nprobe each, unless the predictor or
 * hv_set_query_nprobe trimmed them), and the probe-list stride of hv_get_last_probes (nprobe, or
 * at least 64 with the predictor) */
int hv_get_last_nprobe(const hv_index_t* h, int32_t* out, int64_t capacity);
int64_t hv_last_probe_stride(const hv_index_t* h);

/* ---- learned probe/bound predictor (no counterpart in the reference; the reference always
 * probes nprobe lists and takes its bound from a warm-up scan).
 * A 65 -> 128 (ReLU) -> 2 MLP. Features of a query (HV_MLP_IN): log1p of the 64 smallest upper
 * bounds of its squared centroid distances that the coarse step forms on the matrix cores
 * (d~ + kappa (|q|^2 + |c|^2), d~ = |q|^2 + |c|^2 - 2 q.c, kappa = 1.02 (12 d + 8) 2^-24: between
 * the FP32 chain distance and that plus 2 kappa (|q|^2 + |c|^2); exact distances when nlist >
 * 4096), ascending, then log1p(|q|^2). The learned search needs nprobe <= 512 when nlist <= 4096
 * (its features then come from the same rows). Targets (raw, HV_MLP_OUT):
 * log1p(k-th neighbour distance), log(nprobe needed for the target recall). The network maps
 * standardized features to standardized targets (hv_mlp_set_norm). Parameters, flat FP32
 * (HV_MLP_PARAMS): W1 [65][128], b1 [128], W2 [128][2], b2 [2]. Training (FP32, Adam) and
 * gradients run on the GPU; prediction and serving use BF16 weights with FP32 accumulation.
 * All buffers are host memory. */
#define HV_MLP_IN 65
#define HV_MLP_HIDDEN 128
#define HV_MLP_OUT 2
#define HV_MLP_PARAMS (HV_MLP_IN * HV_MLP_HIDDEN + HV_MLP_HIDDEN + HV_MLP_HIDDEN * HV_MLP_OUT + HV_MLP_OUT)
int hv_query_features(hv_index_t* h, int64_t n, const float* q, float* feat);  /* feat [n][HV_MLP_IN] */
/* norm: feature mean [65], feature std [65], target mean [2], target std [2] */
int hv_mlp_set_norm(hv_index_t* h, const float* norm);
int hv_mlp_get_norm(const hv_index_t* h, float* norm);
int hv_mlp_init(hv_index_t* h, uint64_t seed); /* He-uniform weights, zero biases */
int hv_mlp_set_params(hv_index_t* h, const float* params);
int hv_mlp_get_params(const hv_index_t* h, float* params);
/* loss 1/(2n) sum (y - t)^2 over standardized outputs/targets, and its gradient (HV_MLP_PARAMS) */
int hv_mlp_grad(hv_index_t* h, int64_t n, const float* feat, const float* target, float* grad, float* loss);
/* Adam (lr, beta 0.9 / 0.999, eps 1e-8) over epochs of shuffled minibatches; epoch_loss
 * [epochs] (may be NULL) gets each epoch's mean minibatch loss */
int hv_mlp_train(hv_index_t* h, int64_t n, const float* feat, const float* target, int32_t epochs, int32_t batch,
                 float lr, uint64_t seed, float* epoch_loss);
int hv_mlp_predict(hv_index_t* h, int64_t n, const float* feat, float* out); /* out [n][2], raw targets */

/* ---- device memory plumbing for callers without a GPU framework ---- */
int hv_device_count(void);
int hv_device_malloc(void** ptr, int64_t bytes);
int hv_device_free(void* ptr);
int hv_memcpy_htod(void* dst, const void* src, int64_t bytes);
int hv_memcpy_dtoh(void* dst, const void* src, int64_t bytes);
int hv_memcpy_dtod(void* dst, const void* src, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* HARMONY_IVF_H */
