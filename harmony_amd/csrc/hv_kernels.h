// Host-side launch wrappers for the HIP kernels (implemented in hv_*.hip). All launches go on
// the given stream; none of them synchronise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hv {

constexpr int kMaxChunks = 96;   // staging chunks per vector (dimension blocks split by kChunkDims)
constexpr int kChunkDims = 32;   // dims staged in LDS per step
constexpr int kNumShapes = 7;    // compiled list-scan tile shapes (see hv_scan.hip)
constexpr int kNumXcd = 8;      // MI355X: 8 XCDs, workgroups dispatched round-robin over them
constexpr int kMaxK = 1024;      // largest k of the GPU top-k
constexpr int kMaxSelectRow = 8192;   // largest row (nlist) of the coarse select
constexpr int kMaxWarmup = 8192;      // largest warm-up candidate set per query

// Staging plan of the dimension loop: the B pruning blocks, each cut into chunks of at most
// kChunkDims dims. ends_block[c] = 1 when chunk c closes a block; prune_after[c] = 1 when the
// bound is checked after it (a block that is not the last).
struct ChunkPlan {
    int n;
    int lo[kMaxChunks];
    int len[kMaxChunks];
    int ends_block[kMaxChunks];
    int prune_after[kMaxChunks];
};

struct ScanItem {
    int list;
    int pair_begin;  // range of the grouped (query) pairs
    int pair_end;
    int vec_begin;   // first vector of the item's vector range
};

// --- distances / selects
void launch_pair_l2_full(const float* Q, int64_t nq, const float* C, int64_t nc, int d, float* out, int64_t ldo,
                         hipStream_t s);
void launch_select_rows(const float* dist, int64_t n, int m, int p, int* out_ids, float* out_d, hipStream_t s);
void launch_argmin_rows(const float* dist, int64_t n, int m, int* labels, float* best, hipStream_t s);

// --- search path
void launch_permute_rows(const float* src, int64_t n, int d, const int* perm, float* dst, hipStream_t s);
// kq (optional, here and below): each query's own k (<= k); the bound is taken at the query's k
void launch_kth_survivors(const unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int k,
                          float* bound_ref, float* bound, const int* kq, hipStream_t s);
void launch_compact_topk(unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int k, float* bound,
                         const int* kq, hipStream_t s);
// entries at or past each query's own k -> (FLT_MAX, -1); nothing when kq is null
void launch_truncate_rows(float* D, int64_t* I, int64_t n, int k, const int* kq, hipStream_t s);
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s);
void launch_fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
// up to kFillSegs fills of 32-bit words (pointer, word count, value) in one launch
constexpr int kFillSegs = 12;
struct FillSegs {
    int* p[kFillSegs];
    int64_t n[kFillSegs];
    int v[kFillSegs];
    int nseg;
};
void launch_fill_multi(const FillSegs& f, hipStream_t s);
// qn (optional): per-query probe counts (ranks at or past qn[q] are skipped); nranks (<= nprobe):
// the ranks looked at (every qn[q] <= nranks) and the bucket stride: buckets [nlist][nranks]
void launch_count_pairs(const int* probes, int64_t nq, int nprobe, int nranks, const int* qn, int* bucket_cnt,
                        hipStream_t s);
// exclusive scan of n ints into out (n+1 entries, out[n] = total); tmp >= scan_tmp_ints(n)
int64_t scan_tmp_ints(int64_t n);
void launch_exclusive_scan(const int* in, int64_t n, int* out, int* tmp, hipStream_t s);
void launch_scatter_pairs(const int* probes, int64_t nq, int nprobe, int nranks, const int* qn, const int* bucket_off,
                          int* fill, int* pairs, int* pair_slot, hipStream_t s);
// queue: 3 x 8 ints (per-XCD length, head, base) + nlist ints (first item of each list rank)
// vector range [vmin, vmax) of each list (vmin a multiple of the scan's vector tile)
// nbr: ranks per list in the buckets (bucket of (list l, rank j) = l * nbr + j; r1 <= nbr)
void launch_plan_items(const int* bucket_off, int nlist, int nbr, int r0, int r1, int vmin, int vmax, int qtile,
                       int vsplit, const int* list_rank_order, const int64_t* list_off, int64_t qcap, ScanItem* items,
                       int* queue, unsigned long long* stats, hipStream_t s);
// several phases planned at once (one workgroup each): ranks [r0, r1), vectors [vmin, vmax), and
// the phase's own queue, items (qcap entries) and counters
constexpr int kMaxPlanPhases = 12;
struct PlanSpec {
    int r0, r1, vmin, vmax;
    int* queue;
    ScanItem* items;
    unsigned long long* stats;
    int qtile = 0, vsplit = 0;  // queries / vectors per item of this phase (0: the batch's)
};
void launch_plan_phases(const int* bucket_off, int nlist, int nbr, int qtile, int vsplit,
                        const int* list_rank_order, const int64_t* list_off, int64_t qcap, const PlanSpec* phases,
                        int nph, hipStream_t s);
int scan_query_tile(int shape);   // queries per work item of a tile shape
int scan_vector_tile(int shape);  // vectors per tile of a tile shape
int scan_grid_size(int device, int shape, int d);
void launch_list_scan(int shape, const float* Qp, int d, const float* codes, const int64_t* code_off,
                      const int64_t* list_off, const int* list_stride, const uint32_t* ids,
                      const ChunkPlan& plan, int prune, const float* bound, const int* pairs,
                      const ScanItem* items, int* queue, int vmax, int vsplit, int accept_all,
                      unsigned long long* surv, unsigned int* surv_cnt, int cap, unsigned long long* stats, int grid,
                      hipStream_t s);
// the queries the exact path redoes -> list / list_cnt (both zeroed by the caller): overflowed
// survivor buffers (cnt[q] > cap; counts[0]) and, with check_bound, queries whose k-th result is
// missing or above bound[q] (counts[1]); qn_out (optional): each listed query's probe count is set
// to qn_full (its redo covers them all); bound_out (optional): its bound to the k-th result it has
// (FLT_MAX when overflowed or short). publish (optional, pinned host memory): the last workgroup
// of the grid copies stats[0..16) there; done counts the workgroups (cleared by the caller).
// The select launchers take it and launch k_redo_list after their last kernel.
struct RedoArgs {
    const unsigned int* cnt;
    int cap;
    const float* D;
    const int64_t* I;
    int k;
    const float* bound;
    int check_bound;
    int64_t n;
    int* list;
    int* list_cnt;
    unsigned long long* counts;
    int* qn_out;
    int qn_full;
    float* bound_out;
    const unsigned long long* stats;
    unsigned int* done;
    unsigned long long* publish;
};
void launch_select_survivors(const unsigned long long* surv, const unsigned int* surv_cnt, int64_t nq, int cap,
                             int k, float* D, int64_t* I, int* overflow_list, int* overflow_cnt,
                             unsigned long long* stats, hipStream_t s, const RedoArgs* redo = nullptr);
// exact path for the queries qlist[q0 .. q0 + min(*nql_dev - q0, nql_max)) (the count read on the
// device) over their probe ranks [0, nranks) (and below qn[q] when given; probe rows are nprobe
// wide), lists cut into exact_segments(max list length) segments: per-(query, rank, segment)
// top-k runs appended per query (runs: nql_max * nranks * nseg * k keys, then nql_max ints of
// counts, which the call clears unless counts_zeroed), merged into D/I. bound may be null (none).
constexpr int kExactGrid = 2048;
int exact_segments(int64_t max_list);
void launch_query_exact(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* Qp, int d,
                        const float* codes, const int64_t* code_off, const int64_t* list_off, const int* list_stride,
                        const uint32_t* ids, const int* probes, int nprobe, int nranks, const int* qn, int nseg,
                        const ChunkPlan& plan, int prune, const float* bound, int k, unsigned long long* runs, float* D,
                        int64_t* I, bool counts_zeroed, hipStream_t s);

// --- BF16 precision form (hv_bf16.hip)
int bf16_ksteps(int d);  // 32-dim MFMA steps per vector (dpad = 32 * ksteps)
int bf16_grid_size(int device);
// rerr[p]: norm of the BF16 rounding error of row p's residual (+ 2^-23 of its norm)
void launch_build_resid(const float* x, int d, int dpad, const int* order, const int64_t* list_off,
                        const float* centroids, const int64_t* roff, int nlist, __bf16* rcodes, float* rnorm,
                        float* rerr, hipStream_t s);
// rows in list order: by_pos[p] = rows[order[p]]
void launch_gather_rows(const float* rows, int64_t n, int d, const int* order, float* by_pos, hipStream_t s);
// BF16 (q - c) rows of dpad dims for every (query, probe rank in [r0, nranks)) slot, and the
// norm of each slot's rounding error (+ 2^-23 of its norm) in qerr
void launch_pair_resid(const float* Qp, const float* centroids, const int* probes, const int* qn, int64_t nq,
                       int nprobe, int r0, int nranks, int d, int dpad, __bf16* out, float* qerr, hipStream_t s);
// FP8 (e4m3) codes of every list from the rows in list order, scaled per list by a power of two
// (inverse in inv_scale[l]), in the BF16 codes' order, one byte per value
void launch_build_resid_fp8(const float* rows_by_pos, int d, int dpad, const int64_t* list_off,
                            const float* centroids, const int64_t* roff, int nlist, uint8_t* codes, float* inv_scale,
                            hipStream_t s);
// FP8 (q - c) rows of every slot, scaled per slot (inverse in inv_scale[slot])
void launch_pair_resid_fp8(const float* Qp, const float* centroids, const int* probes, const int* qn, int64_t nq,
                           int nprobe, int nranks, int d, int dpad, uint8_t* out, float* inv_scale, hipStream_t s);
// the MFMA filter scan: elem 0 = BF16 codes (qaux = per-slot rounding-error norms, rerr =
// per-vector ones, lscale unused), 1 = FP8 codes (qaux = per-slot inverse scales, lscale =
// per-list ones, rerr and the kappas unused)
void launch_mfma_scan(int grid, int elem, const void* qres, const float* qaux, int d, const float* pdist,
                      int nprobe, const void* rcodes, const float* lscale, const int64_t* roff, const float* rnorm,
                      const float* rerr, const int64_t* list_off, const float* bound, const int* pair_slot,
                      const ScanItem* items, const int* queue, int vmax, int vsplit, int mode, float kappa1,
                      float kappa_g, float kappa2, unsigned long long* surv, float* aux, unsigned int* surv_cnt,
                      int cap, unsigned long long* stats, hipStream_t s);
// The BF16 form's nearest-list pass: per item (list, <= kNearQueries queries whose nearest list it
// is, <= kNearSplit vectors; planned with that query tile and vector split), the filter scan with
// every pair's bounds kept in LDS, the k-th smallest upper bound (k per query with kq) and the
// pairs below it. A list held by one item leaves each query's survivors, count, bound and
// compaction count (last) as the direct-slot scan and its compaction would; parts of split lists
// append to the survivor buffers and leave `last`, for the compaction the caller runs after the
// pass. lrad: each list's radius (the largest |x - c|), which sets the bounds' fixed-point scale.
constexpr int kNearQueries = 16;
constexpr int kNearSplit = 1024;  // vectors per item: longer lists are cut into several items
constexpr int kNearMaxK = 1024;  // a query's k fits the range's bounds up to k <= kNearSplit
// The pass forms the rank-0 query residuals itself from Qp and the centroids (BF16 hi + lo parts
// for d <= 256). bound_ref (optional): the bound of every query the pass finishes is written there
// too.
void launch_mfma_nearest(int grid, const float* Qp, const float* centroids, int d, const float* pdist, int nprobe, const __bf16* rcodes, const int64_t* roff,
                         const float* rnorm, const float* rerr, const float* lrad, const int64_t* list_off,
                         float* bound, float* bound_ref, const int* pair_slot, const ScanItem* items,
                         const int* queue, float kappa1, float kappa_g, float kappa2, int k, const int* kq,
                         unsigned long long* surv, float* aux, unsigned int* surv_cnt, unsigned int* last, int cap,
                         int fused, unsigned long long* stats, hipStream_t s);
// FP8 shortlist: each query keeps its kr smallest estimates (ties at the kr-th included), in place;
// bound[q] drops to the kr-th estimate
// with kq: each query keeps min(kr, kmul * k_q)
void launch_shortlist(unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int kr, float* bound,
                      const int* kq, int kmul, hipStream_t s);
// counts after the direct-slot scan of the nearest lists (vectors v < cap)
void launch_first_counts(const int* probes, int64_t nq, int nprobe, const int64_t* list_off, int cap,
                         unsigned int* cnt, hipStream_t s);
// last (optional, nq counts; start at 0xffffffff): the count each query's previous compaction
// left; a query whose count still equals it is skipped
// with kq: the bound at each query's min(k, kmul * k_q)-th upper bound
// bound_ref (optional): each compacted query's new bound is written there too
void launch_bf16_compact(unsigned long long* surv, float* aux, unsigned int* cnt, int64_t nq, int cap, int k,
                         float* bound, unsigned int* last, const int* kq, int kmul, hipStream_t s,
                         float* bound_ref = nullptr);
// re-rank + final top-k of the MFMA forms, in three launches:
//  launch_rerank_items: work items of 64 stored pairs (items: nq * ceil(cap/64) entries, n_items
//    zeroed by the caller); overflowed queries listed when overflow_list is given; stats[2]
//    counts the pairs. Queries in index order, or, with the (list, rank)-sorted pairs and their
//    slots given (npairs on the device, at most max_pairs), in the order of their rank-0 pairs;
//  launch_rerank: each pair's exact key written to its slot of kept (nq * cap keys; padding when
//    above bound[q], no filter when bound is null);
//  launch_select_kept: the first k kept keys of every non-overflowed query into D/I (done: nq
//    flags of scratch).
constexpr int kRerankGrid = 1536;  // the re-rank's resident workgroups (76 VGPRs: 6 wavefronts per SIMD)
constexpr int kRerankTileGrid = 1024;  // the LDS-staged re-rank's (36 KB of LDS: 4 workgroups per CU)
void launch_rerank_items(const unsigned int* cnt, int64_t nq, int cap, const int* pairs, const int* pair_slot,
                         const int* npairs, int64_t max_pairs, int nprobe, int2* items, unsigned int* n_items,
                         int* overflow_list, int* overflow_cnt, unsigned long long* stats, hipStream_t s);
void launch_rerank(const float* Qp, int d, const float* rows_by_pos, const uint32_t* ids, const ChunkPlan& plan,
                   const unsigned long long* surv, int64_t nq, int cap, const float* bound, const int2* items,
                   const unsigned int* n_items, unsigned long long* kept, hipStream_t s);
void launch_select_kept(const unsigned long long* kept, const unsigned int* cnt, int64_t nq, int cap, int k, float* D,
                        int64_t* I, unsigned int* done, hipStream_t s, const RedoArgs* redo = nullptr);
void launch_redo_list(const RedoArgs& ra, hipStream_t s);
// the workgroup select alone: queries whose done flag is 0 (with done), else those with more than
// kWaveSortMax survivors; with redo, then the redo list and the counters (see RedoArgs)
void launch_select_block(const unsigned long long* surv, const unsigned int* surv_cnt, int64_t nq, int cap, int k,
                         float* D, int64_t* I, const unsigned int* done, hipStream_t s, const RedoArgs* redo = nullptr);

// --- learned probe/bound predictor (hv_learn.hip)
constexpr int kNearest = 64;                                            // nearest-centroid distances
constexpr int kMlpIn = kNearest + 1, kMlpHidden = 128, kMlpOut = 2;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpOut + kMlpOut;
constexpr int kMlpW1Serve = (kMlpIn + 3) / 4 * 4 * kMlpHidden;  // BF16 serving copy of W1 (padded)
// norm: fmean [65], fstd [65], tmean [2], tstd [2]
void launch_features(const float* pdist, int P, const float* Qp, int d, int64_t n, float* feat, hipStream_t s);
void launch_mlp_step(const float* X, const float* T, const int* idx, int B, float* P, const float* norm, float* Xn,
                     float* H, float* Y, float* dY, float* dH, float* row_loss, float* G, float* loss_acc,
                     hipStream_t s);
void launch_adam(float* P, const float* G, float* M, float* V, float lr, float beta1, float beta2, float eps, int t,
                 hipStream_t s);
void launch_mlp_to_bf16(const float* P, __bf16* w1, __bf16* w2, hipStream_t s);
void launch_mlp_predict(const float* X, int64_t n, const __bf16* w1, const __bf16* w2, const float* P,
                        const float* norm, float* out, hipStream_t s);
// features from pdist [n][pc] (ascending coarse distances) and the queries Qp [n][d], as
// launch_features forms them; bound2, bound3 (optional): more copies of the seeded bound. With
// margin > 0, lrad (list radii) and cent (centroids), each query's count widens to every list of
// its first nprobe probes whose triangle or Voronoi lower bound is within the seed (widened:
// count of such queries).
void launch_mlp_serve(const float* near, const float* pdist, const int* probes, int pc, const float* Qp, int d,
                      int64_t n,
                      const __bf16* w1, const __bf16* w2, const float* P, const float* norm, float nprobe_scale,
                      float bound_scale, int min_nprobe, int nprobe, const float* lrad, const float* cent,
                      float margin, float slack, const int64_t* list_off, int kmin, int* qn, float* bound,
                      float* pred, float* bound2, float* bound3, unsigned long long* widened, int* bucket_cnt,
                      int nranks, hipStream_t s);
// lrad[l] = max over list l of sqrt(rnorm) (the list's radius around its centroid)
void launch_list_radius(const float* rnorm, const int64_t* list_off, int nlist, float* lrad, hipStream_t s);

// --- coarse quantizer on the matrix cores with an exact boundary re-check (hv_coarse.hip)
void launch_row_norms(const float* X, int64_t n, int d, float* out, hipStream_t s);
float coarse_kappa(int d);
// also writes the squared norms of the rows (qn [nq], cn [nc]) that the select uses
void launch_coarse_mfma(const float* Q, int64_t nq, const float* C, int nc, int d, float* qn, float* cn,
                        float* approx, hipStream_t s, const FillSegs* fill = nullptr);
// the kNearest smallest distance values of each row of v [n][nc] ascending, a short row's largest
// repeated, into out [n][kNearest] (nc <= kMaxSelectRow): with qn / cn, v holds coarse estimates
// and the values are their upper bounds d~ + kappa (|q|^2 + |c|^2); without, v holds exact
// distances (negatives as 0)
void launch_near_rows(const float* v, int64_t n, int nc, const float* qn, const float* cn, float kappa, float* out,
                      hipStream_t s);
// P nearest per query, exact (distance, id) order; redo_cnt must be 0 on entry. near (optional):
// the kNearest smallest upper bounds of each query, sorted, as launch_near_rows forms them.
// bucket_cnt (optional): the (list, rank < nranks) pair counts of the grouping are added as the
// probes are written; returns whether they were (the workgroup kernel leaves them to the caller)
bool launch_coarse_select(const float* approx, const float* Q, const float* C, int nc, int d, const float* qn,
                          const float* cn, float kappa, int64_t nq, int P, int* probes, float* pdist,
                          int* redo_list, int* redo_cnt, float* near, int* bucket_cnt, int nranks, hipStream_t s);

// --- inner-product metric (hv_ip.hip; the IP list scan is launch_query_exact_ip)
void launch_ip_rows(const float* X, int64_t n, const float* C, int64_t nc, int d, float* out, int64_t ldo,
                    hipStream_t s);
void launch_argmax_rows(const float* v, int64_t n, int m, int* labels, float* best, hipStream_t s);
void launch_normalize_rows(const float* x, int64_t n, int d, float* out, hipStream_t s);
void launch_normalize_centroids(float* c, int nlist, int d, hipStream_t s);
// the P largest of each row [n][m] (m <= kMaxSelectRow), descending, ties to the lower column
void launch_ip_select_rows(const float* v, int64_t n, int m, int P, int* ids, float* vals, hipStream_t s);
void launch_ip_simi(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* D, const int64_t* I, int k,
                    float* simi, hipStream_t s);
void launch_iota_list(int* list, int64_t n, hipStream_t s);  // list[0] = n, list[1 + i] = i
// exact inner-product top-k of the listed queries over probe ranks [r0, r1): keys by (inner product
// descending, id ascending), D = inner products, padding (-FLT_MAX, -1). The kept keys of every
// list are runs (nprobe * nseg * k keys per query, then nql_max counts); append keeps the runs of
// an earlier call (ranks before r0) and D, I are the top-k of all of them. With simi, each list is
// scanned only over the angular band of candidate2centroid (c2c) a result better than simi[q]
// needs; scanned (may be null) accumulates the inner products computed.
void launch_query_exact_ip(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* Qp, int d,
                           const float* codes, const int64_t* code_off, const int64_t* list_off,
                           const int* list_stride, const uint32_t* ids, const int* probes, int nprobe, int r0,
                           int r1, int nseg, const float* c2c, const float* pval, const float* simi, float slack,
                           int k, unsigned long long* runs, bool append, float* D, int64_t* I,
                           unsigned long long* scanned, hipStream_t s);

// --- FP64 ground truth
void launch_gt_f64(const float* Qp, int64_t nq, int d, const float* codes, const int64_t* code_off,
                   const int64_t* list_off, const int* list_stride, const uint32_t* ids, int nlist,
                   const double* tau, int cap, double* sd, uint32_t* sid, unsigned int* cnt, hipStream_t s);

// --- index build
void launch_centroid_update_f64(const float* x, int d, const int* perm, const int64_t* cl_off, int nlist,
                                float* centroids, hipStream_t s);
void launch_build_layout(const float* x, int d, const int* order, const int64_t* list_off,
                         const int64_t* code_off, const int* list_stride, int nlist, float* codes, hipStream_t s);

}  // namespace hv
