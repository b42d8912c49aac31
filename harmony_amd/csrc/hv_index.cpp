// Host side of the MI355X IVF-flat library: the index object behind the C-ABI of
// include/harmony_ivf.h. It mirrors harmony::Index (src/Index.h:13-149): train (k-means),
// add (assignment + list order), search (coarse -> warm-up bound -> pruned list scan ->
// top-k), and keeps every list on the GPU in the list-major, dimension-major layout the
// kernels read. All device work is issued on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/harmony_ivf.h"
#include "hv_device.h"
#include "hv_kernels.h"

#include <atomic>
#include <functional>
#include <mutex>

// Scratch of hv_search_device_noalloc: a fixed block of device memory on every GPU the library
// runs on, reserved when its code object is loaded and held until the process ends. Host code
// never uses this name as an address: hv_create looks up each device's copy with
// hipGetSymbolAddress (with that device current) and the search passes that address on.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

thread_local std::string g_create_error;

// Bumped by every device allocation and free of the search buffers: a captured search graph
// holds their addresses and is replayed only while this has not moved.
std::atomic<uint64_t> g_alloc_gen{0};

// Bump allocator over a fixed block of device memory (the no-allocation search's scratch): the
// search buffers are carved from it in order, 256-byte aligned; a buffer that was the last one
// carved grows in place. Carving past the end fails instead of allocating.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* last = nullptr;
    size_t high = 0;  // largest `used` of the current call
};

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    Arena* arena = nullptr;  // set: ensure() carves from the arena and never allocates
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (arena) {
            const size_t b = (std::max<size_t>(bytes, 256) + 255) / 256 * 256;
            char* at = (p && p == arena->last) ? static_cast<char*>(p) : arena->base + arena->used;
            const size_t end = (size_t)(at - arena->base) + b;
            if (end > arena->cap) return hipErrorOutOfMemory;
            p = at;
            cap = b;
            arena->last = at;
            arena->used = end;
            arena->high = std::max(arena->high, end);
            return hipSuccess;
        }
        g_alloc_gen.fetch_add(1);
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t b = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) cap = b;
        return e;
    }
    void release() {
        if (arena) {  // arena memory is never freed
            p = nullptr;
            cap = 0;
            return;
        }
        if (p) (void)hipFree(p);
        g_alloc_gen.fetch_add(1);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Pinned host staging (a graph's host-to-device copy reads it at replay time)
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        g_alloc_gen.fetch_add(1);
        if (p) {
            hipError_t e = hipHostFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        hipError_t e = hipHostMalloc(&p, std::max<size_t>(bytes, 256), hipHostMallocDefault);
        if (e == hipSuccess) cap = std::max<size_t>(bytes, 256);
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// What the host does after a search's device work: read the counters, add redo batches, report.
struct SearchPost {
    int batch = 1;
    bool main_recorded = false, mfma = false, per_query = false;
    int B = 1;
    int64_t nprobe = 0, nprobe_cap = 0;
    int* redo_cnt = nullptr;
    std::function<void(int)> redo;
    std::function<void()> truncate;
};

// A search captured as a HIP graph, replayed while the signature and the buffers match
struct SearchGraph {
    std::vector<unsigned char> sig;
    uint64_t gen = 0;
    hipGraphExec_t exec = nullptr;
    SearchPost post;
    std::vector<unsigned char> pending_sig;  // an eager search waiting for a second identical call
    uint64_t pending_gen = 0;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        sig.clear();
        pending_sig.clear();
    }
};

struct HvError {
    int code;
    std::string msg;
};

#define HV_HIP(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            throw HvError{HV_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)};              \
    } while (0)

#define HV_CHECK_LAUNCH(what)                                                                             \
    do {                                                                                                  \
        hipError_t e_ = hipGetLastError();                                                                \
        if (e_ != hipSuccess) throw HvError{HV_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)}; \
    } while (0)

void require(bool ok, int code, const std::string& msg) {
    if (!ok) throw HvError{code, msg};
}

// Per-device address of `scratch` (looked up once per device) and the lock that serializes the
// no-allocation searches on that device, which share it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_setup_mu;
std::mutex g_scratch_mu[kMaxDevices];
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// With `dev` current: the device address of its scratch block, or null (error in *e).
void* scratch_setup(int dev, hipError_t* e) {
    if (dev < 0 || dev >= kMaxDevices) {
        *e = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_setup_mu);
    if (!g_scratch_done[dev]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[dev] = r;
        g_scratch_ptr[dev] = r == hipSuccess ? p : nullptr;
        g_scratch_done[dev] = true;
    }
    *e = g_scratch_err[dev];
    return g_scratch_ptr[dev];
}

// SplitMix64 + Fisher-Yates: the shuffle used for the k-means subsample and init
// (replacing std::shuffle(default_random_engine(seed)) of Clustering.cpp:58,81 with a
// portable stream).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

void shuffle_indices(std::vector<int64_t>& v, uint64_t seed) {
    SplitMix64 rng(seed);
    for (int64_t i = (int64_t)v.size() - 1; i > 0; --i) {
        int64_t j = (int64_t)(rng.next() % (uint64_t)(i + 1));
        std::swap(v[i], v[j]);
    }
}

// Subsample (Clustering.cpp:48-68) and init (72-89): returns the sample row indices and the
// sample rows that seed the centroids.
void kmeans_pick(int64_t n, int64_t nlist, int64_t max_pts, uint64_t seed, std::vector<int64_t>& sample_rows,
                 std::vector<int64_t>& init_rows) {
    sample_rows.resize(n);
    std::iota(sample_rows.begin(), sample_rows.end(), 0);
    if (max_pts > 0 && n > nlist * max_pts) {
        shuffle_indices(sample_rows, seed);
        sample_rows.resize(nlist * max_pts);
    }
    const int64_t ns = (int64_t)sample_rows.size();
    init_rows.resize(ns);
    std::iota(init_rows.begin(), init_rows.end(), 0);
    shuffle_indices(init_rows, seed);
    init_rows.resize(nlist);
}

// Longest vector range of one scan work item: long lists are split so that no workgroup holds
// a whole large list while the others run dry.
constexpr int kVecSplit = 1024;
// The BF16 scan's items: its per-item setup is small (query residuals from memory), so shorter
// ranges spread long nearest lists over more workgroups.
constexpr int kVecSplitBf16 = 256;

// The BF16 form scans the probe ranks after the nearest list in one phase up to this many probes,
// beyond it in groups of growing width with the bound tightened in between
constexpr int64_t kBf16SinglePhase = 16;

// Scratch of one batch of exact redos (per-(query, rank, segment) runs of k keys); the
// no-allocation search uses smaller batches
constexpr int64_t kRedoRunBytes = 64ll << 20;
constexpr int64_t kRedoRunBytesNoAlloc = 8ll << 20;
// Coarse distance rows are formed in slices of at most this many bytes (no-allocation: smaller)
constexpr int64_t kCoarseSliceBytes = 256ll << 20;
constexpr int64_t kCoarseSliceBytesNoAlloc = 16ll << 20;

// Dimension blocks of node.cpp:639-646 (B blocks of floor(d/B) dims, the last takes the
// rest), each cut into staging chunks of at most kChunkDims dims.
hv::ChunkPlan make_plan(int d, int B) {
    hv::ChunkPlan p;
    std::memset(&p, 0, sizeof(p));
    const int w = d / B;
    for (int b = 0; b < B; ++b) {
        const int lo = b * w;
        const int hi = (b == B - 1) ? d : lo + w;
        for (int c = lo; c < hi; c += hv::kChunkDims) {
            require(p.n < hv::kMaxChunks, HV_ERR_CAPACITY, "too many dimension chunks (d too large for B)");
            p.lo[p.n] = c;
            p.len[p.n] = std::min(hv::kChunkDims, hi - c);
            p.ends_block[p.n] = (c + hv::kChunkDims >= hi) ? 1 : 0;
            p.prune_after[p.n] = (p.ends_block[p.n] && b != B - 1) ? 1 : 0;
            p.n++;
        }
    }
    return p;
}

}  // namespace

struct hv_index {
    int64_t d = 0, nlist = 0;
    int metric = HV_METRIC_L2;
    int opt_level = HV_OPT_NONE, added_opt_level = HV_OPT_NONE;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    // begin/end pairs around each candidate-scan kernel launch of an eager search
    static constexpr int kScanEv = 16;
    hipEvent_t ev_scan[2 * kScanEv] = {};
    int nscan_ev = 0;
    int scan_grid[hv::kNumShapes] = {};

    std::vector<int> perm;  // empty = identity
    DevBuf d_perm;

    bool trained = false;
    std::vector<float> centroids;  // index dim order
    DevBuf d_centroids;

    bool added = false;
    int64_t ntotal = 0, max_list = 0;
    std::vector<int64_t> list_off, code_off, ids;  // ids in list order
    std::vector<int> list_stride, rank_order;
    DevBuf d_codes, d_list_off, d_code_off, d_list_stride, d_ids, d_rank_order;
    // BF16 residual codes per list (MFMA operand order), their FP32 norms, and the FP32 rows in list order
    DevBuf d_rcodes, d_roff, d_rnorm, d_rerr, d_rows;
    DevBuf d_lrad;  // each list's radius around its centroid (the learned search's margin test)
    // inner-product metric: each vector's inner product with its centroid, list order (descending
    // within a list when c2c_sorted), the candidate2centroid array of the reference's IVF
    DevBuf d_c2c;
    std::vector<float> c2c;
    bool c2c_sorted = false;
    int dpad = 0, bf16_grid = 0;
    // FP8 (e4m3) codes per list in the same order and their per-list inverse scales, built on the
    // first FP8-form search after an add
    DevBuf d_rcodes8, d_lscale8;
    bool fp8_built = false;
    int64_t rcode_elems = 0;  // code values of all lists (both code arrays)

    hv_search_params sp{};
    std::vector<float> ext_bounds;
    std::vector<int32_t> query_nprobe;  // per-query probe counts (empty: nprobe for all)
    PinnedBuf qn_host;                  // the clamped counts of the current search
    PinnedBuf post_host;                // the search's counters, read back after its device work
    std::vector<int32_t> qn_stage;      // the same for a no-allocation search (sized when set)
    std::vector<int32_t> query_k;       // per-query k (empty: the search's k for all)
    PinnedBuf qk_host;                  // their clamped values for the current search
    std::vector<int32_t> qk_stage;      // the same for a no-allocation search
    // no-allocation search in progress: scratch comes from the static arena, graphs are off;
    // size_only: stop after sizing the scratch (the no-allocation entry point's first pass)
    bool noalloc = false, size_only = false;
    SearchGraph graph;

    // learned probe/bound predictor: FP32 parameters and Adam moments, BF16 serving weights,
    // feature/target standardization
    bool mlp_params = false, mlp_norm = false;
    DevBuf m_params, m_grad, m_adam_m, m_adam_v, m_w1b, m_w2b, m_norm;

    // search scratch (grow-only)
    DevBuf s_qk;  // per-query k of the current search
    DevBuf s_q, s_qp, s_D, s_I, s_dist, s_probes, s_bound, s_bound_ref, s_bucket, s_fill, s_bucket_off, s_scan_tmp, s_pairs,
        s_items, s_queue, s_surv, s_surv_cnt, s_overflow, s_ritems, s_kept, s_done, s_stats, s_pdist, s_pair_slot,
        s_aux, s_qres, s_runs, s_qn, s_feat, s_pred, s_seed, s_fb, s_qnorm, s_cnorm, s_redo, s_qscale, s_last,
        s_near;
    int64_t last_n = 0, last_nprobe = 0, last_stride = 0;
    bool last_learned = false;
};

namespace {

void set_err(hv_index_t* h, const HvError& e) {
    if (h) h->err = e.msg;
}

template <typename F>
int guarded(hv_index_t* h, F&& f) {
    try {
        f();
        if (h) h->err.clear();
        return HV_OK;
    } catch (const HvError& e) {
        set_err(h, e);
        return e.code;
    } catch (const std::bad_alloc&) {
        if (h) h->err = "host allocation failed";
        return HV_ERR_CAPACITY;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        return HV_ERR_INVALID_ARG;
    }
}

void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
    HV_HIP(b.ensure(bytes));
    if (bytes) HV_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
}

// Nearest centroid of every row of xp (index dim order, on the device): labels + distances.
void assign_rows(hv_index_t* h, const float* xp, int64_t n, int* labels, float* best) {
    const int64_t per = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)(256ll << 20) / (h->nlist * 4)));
    HV_HIP(h->s_dist.ensure(sizeof(float) * per * h->nlist));
    for (int64_t r0 = 0; r0 < n; r0 += per) {
        const int64_t rows = std::min(per, n - r0);
        if (h->metric == HV_METRIC_IP) {  // the largest inner product, ties to the lower id
            hv::launch_ip_rows(xp + r0 * h->d, rows, h->d_centroids.as<float>(), h->nlist, (int)h->d,
                               h->s_dist.as<float>(), h->nlist, h->stream);
            hv::launch_argmax_rows(h->s_dist.as<float>(), rows, (int)h->nlist, labels + r0, best ? best + r0 : nullptr,
                                   h->stream);
        } else {
            hv::launch_pair_l2_full(xp + r0 * h->d, rows, h->d_centroids.as<float>(), h->nlist, (int)h->d,
                                    h->s_dist.as<float>(), h->nlist, h->stream);
            hv::launch_argmin_rows(h->s_dist.as<float>(), rows, (int)h->nlist, labels + r0, best ? best + r0 : nullptr,
                                   h->stream);
        }
        HV_CHECK_LAUNCH("assignment");
    }
}

// Upload rows in caller dim order and return a device pointer to them in index dim order.
const float* upload_rows(hv_index_t* h, const float* x, int64_t n, DevBuf& raw, DevBuf& permuted) {
    upload(raw, x, sizeof(float) * n * h->d, h->stream);
    if (h->perm.empty()) return raw.as<float>();
    HV_HIP(permuted.ensure(sizeof(float) * n * h->d));
    hv::launch_permute_rows(raw.as<float>(), n, (int)h->d, h->d_perm.as<int>(), permuted.as<float>(), h->stream);
    HV_CHECK_LAUNCH("permute");
    return permuted.as<float>();
}

// Installs the inverted lists on the device: rows of xp (index dim order, on the device) are
// gathered through order[] into the list-major, dimension-major slabs; off[nlist+1] delimits
// the lists and ids[] (list order) are the vector ids.
void install_lists(hv_index_t* h, const float* xp, const std::vector<int>& order, const std::vector<int64_t>& off,
                   const std::vector<int64_t>& ids, const std::vector<float>* c2c = nullptr) {
    const int64_t d = h->d, nlist = h->nlist, n = (int64_t)order.size();
    h->list_off = off;
    h->list_stride.assign(nlist, 0);
    h->code_off.assign(nlist + 1, 0);
    for (int64_t l = 0; l < nlist; ++l) {
        const int64_t sz = off[l + 1] - off[l];
        require(sz < (1ll << 30), HV_ERR_CAPACITY, "list longer than 2^30");
        h->list_stride[l] = (int)((sz + 15) / 16 * 16);
        h->code_off[l + 1] = h->code_off[l] + d * h->list_stride[l];
    }
    h->ids = ids;
    h->rank_order.resize(nlist);
    std::iota(h->rank_order.begin(), h->rank_order.end(), 0);
    std::stable_sort(h->rank_order.begin(), h->rank_order.end(),
                     [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
    std::vector<uint32_t> ids32(ids.begin(), ids.end());
    DevBuf order_buf;
    upload(order_buf, order.data(), sizeof(int) * n, h->stream);
    upload(h->d_list_off, off.data(), sizeof(int64_t) * (nlist + 1), h->stream);
    upload(h->d_code_off, h->code_off.data(), sizeof(int64_t) * (nlist + 1), h->stream);
    upload(h->d_list_stride, h->list_stride.data(), sizeof(int) * nlist, h->stream);
    upload(h->d_ids, ids32.data(), sizeof(uint32_t) * n, h->stream);
    upload(h->d_rank_order, h->rank_order.data(), sizeof(int) * nlist, h->stream);
    HV_HIP(h->d_codes.ensure(sizeof(float) * std::max<int64_t>(1, h->code_off[nlist])));
    if (n)
        hv::launch_build_layout(xp, (int)d, order_buf.as<int>(), h->d_list_off.as<int64_t>(),
                                h->d_code_off.as<int64_t>(), h->d_list_stride.as<int>(), (int)nlist,
                                h->d_codes.as<float>(), h->stream);
    HV_CHECK_LAUNCH("layout");
    // BF16 precision form: residual codes and norms, FP32 rows by id for the exact re-rank
    h->dpad = hv::bf16_ksteps((int)d) * 32;
    std::vector<int64_t> roff(nlist + 1, 0);
    for (int64_t l = 0; l < nlist; ++l) roff[l + 1] = roff[l] + (int64_t)h->list_stride[l] * h->dpad;
    upload(h->d_roff, roff.data(), sizeof(int64_t) * (nlist + 1), h->stream);
    HV_HIP(h->d_rcodes.ensure(sizeof(__bf16) * std::max<int64_t>(1, roff[nlist])));
    h->rcode_elems = roff[nlist];
    HV_HIP(h->d_rnorm.ensure(sizeof(float) * std::max<int64_t>(1, n)));
    HV_HIP(h->d_rerr.ensure(sizeof(float) * std::max<int64_t>(1, n)));
    HV_HIP(h->d_rows.ensure(sizeof(float) * std::max<int64_t>(1, n * d)));
    if (n) {
        hv::launch_build_resid(xp, (int)d, h->dpad, order_buf.as<int>(), h->d_list_off.as<int64_t>(),
                               h->d_centroids.as<float>(), h->d_roff.as<int64_t>(), (int)nlist,
                               h->d_rcodes.as<__bf16>(), h->d_rnorm.as<float>(), h->d_rerr.as<float>(), h->stream);
        hv::launch_gather_rows(xp, n, (int)d, order_buf.as<int>(), h->d_rows.as<float>(), h->stream);
    }
    if (h->metric == HV_METRIC_IP) {
        // candidate2centroid in list order; the triangle band needs it descending within each list
        h->c2c = c2c ? *c2c : std::vector<float>(n, 0.f);
        h->c2c_sorted = c2c != nullptr;
        for (int64_t l = 0; l < nlist && h->c2c_sorted; ++l)
            for (int64_t t = off[l] + 1; t < off[l + 1]; ++t)
                if (!(h->c2c[t] <= h->c2c[t - 1])) {
                    h->c2c_sorted = false;
                    break;
                }
        upload(h->d_c2c, h->c2c.data(), sizeof(float) * std::max<int64_t>(1, n), h->stream);
    }
    HV_HIP(h->d_lrad.ensure(sizeof(float) * nlist));
    hv::launch_list_radius(h->d_rnorm.as<float>(), h->d_list_off.as<int64_t>(), (int)nlist, h->d_lrad.as<float>(),
                           h->stream);
    HV_CHECK_LAUNCH("bf16 layout");
    HV_HIP(hipStreamSynchronize(h->stream));
    order_buf.release();
    h->ntotal = n;
    h->fp8_built = false;
    h->max_list = 0;
    for (int64_t l = 0; l < nlist; ++l) h->max_list = std::max<int64_t>(h->max_list, off[l + 1] - off[l]);
    h->added = true;
}

// Coarse quantizer: the P nearest lists of each query, ascending by the FP32 chain distance,
// ties to the lower id (Index.cpp:967-997). Up to P = 512 and nlist = 4096 on the matrix cores
// with an exact re-check of the candidates near the P-th (hv_coarse.hip); otherwise full exact
// distance rows and a select. Rows in slices of <= 256 MB.
int64_t coarse_rows_per_slice(const hv_index_t* h, int64_t n) {
    const int64_t slice = h->noalloc ? kCoarseSliceBytesNoAlloc : kCoarseSliceBytes;
    return std::max<int64_t>(1, std::min<int64_t>(n, slice / (h->nlist * 4)));
}
bool coarse_on_mfma(const hv_index_t* h, int64_t P) { return P <= 512 && h->nlist <= 4096; }
// the MFMA path's re-check counter and list (slice rows + 1 ints)
int* coarse_counter(hv_index_t* h, int64_t n) {
    HV_HIP(h->s_redo.ensure(sizeof(int) * (coarse_rows_per_slice(h, n) + 1)));
    return h->s_redo.as<int>();
}
// counter_zeroed: the caller has cleared the re-check counter (the first slice skips its fill);
// near (optional, [n][kNearest]): the kNearest smallest distance estimates of each query, sorted
// (the learned predictor's features), taken from the same rows; bucket_cnt (optional): the
// grouping's (list, rank < nranks) pair counts, added as the probes are written. Returns whether
// they were (else the caller counts them).
// fill: the search's buffers that start cleared, cleared by the coarse GEMM's first launch (or by
// a fill launch of their own before the other coarse path)
bool coarse_select(hv_index_t* h, int64_t n, const float* Qp, int64_t P, int* probes, float* pdist,
                   bool counter_zeroed = false, float* near = nullptr, int* bucket_cnt = nullptr, int nranks = 0,
                   const hv::FillSegs* fill = nullptr) {
    const int64_t d = h->d, nlist = h->nlist;
    hipStream_t s = h->stream;
    const int64_t per = coarse_rows_per_slice(h, n);
    HV_HIP(h->s_dist.ensure(sizeof(float) * per * nlist));
    if (coarse_on_mfma(h, P)) {
        HV_HIP(h->s_qnorm.ensure(sizeof(float) * n));
        HV_HIP(h->s_cnorm.ensure(sizeof(float) * nlist));
        int* redo_cnt = coarse_counter(h, n);
        const float kappa = hv::coarse_kappa((int)d);
        bool counted = bucket_cnt != nullptr;
        for (int64_t r0 = 0; r0 < n; r0 += per) {
            const int64_t rows = std::min(per, n - r0);
            hv::launch_coarse_mfma(Qp + r0 * d, rows, h->d_centroids.as<float>(), (int)nlist, (int)d,
                                   h->s_qnorm.as<float>() + r0, h->s_cnorm.as<float>(), h->s_dist.as<float>(), s,
                                   r0 == 0 ? fill : nullptr);
            if (r0 > 0 || !counter_zeroed) hv::launch_fill_i32(redo_cnt, 1, 0, s);
            counted &= hv::launch_coarse_select(h->s_dist.as<float>(), Qp + r0 * d, h->d_centroids.as<float>(),
                                                (int)nlist, (int)d, h->s_qnorm.as<float>() + r0,
                                                h->s_cnorm.as<float>(), kappa, rows, (int)P, probes + r0 * P,
                                                pdist + r0 * P, redo_cnt + 1, redo_cnt,
                                                near ? near + r0 * hv::kNearest : nullptr, bucket_cnt, nranks, s);
        }
        HV_CHECK_LAUNCH("coarse");
        return counted;
    }
    if (fill) hv::launch_fill_multi(*fill, s);
    for (int64_t r0 = 0; r0 < n; r0 += per) {
        const int64_t rows = std::min(per, n - r0);
        hv::launch_pair_l2_full(Qp + r0 * d, rows, h->d_centroids.as<float>(), nlist, (int)d, h->s_dist.as<float>(),
                                nlist, s);
        hv::launch_select_rows(h->s_dist.as<float>(), rows, (int)nlist, (int)P, probes + r0 * P, pdist + r0 * P, s);
        if (near)
            hv::launch_near_rows(h->s_dist.as<float>(), rows, (int)nlist, nullptr, nullptr, 0.f,
                                 near + r0 * hv::kNearest, s);
    }
    HV_CHECK_LAUNCH("coarse");
    return false;
}

// After a search's device work: read the counters and the redo count (one copy, the one host
// sync when nothing needs redoing), run the exact redo batches when some query needs one (the
// redo kernels are launched only then), and fill the report (phase times only when the phases
// were recorded, i.e. not from a captured graph).
void finish_search(hv_index_t* h, const SearchPost& p, int64_t n, hv_stats* st, bool timed,
                   std::chrono::steady_clock::time_point t0) {
    hipStream_t s = h->stream;
    // [15]: the redo count (low word) and the re-rank item count. The redo-list kernel, the last
    // of the device work, wrote the 16 words into the handle's pinned buffer (allocated with the
    // handle); the synchronization makes them visible here.
    HV_HIP(hipStreamSynchronize(s));
    unsigned long long hs[16];
    std::memcpy(hs, h->post_host.p, sizeof(hs));
    const int nredo = (int)(hs[15] & 0xffffffffull);
    if (nredo > 0) {
        for (int q0 = 0; q0 < nredo; q0 += p.batch) p.redo(q0);
        if (p.truncate) p.truncate();
        HV_CHECK_LAUNCH("exact redo");
        HV_HIP(hipStreamSynchronize(s));
    }
    h->last_n = n;
    h->last_nprobe = p.nprobe_cap;
    h->last_stride = p.nprobe;
    h->last_learned = p.per_query;
    if (st) {
        if (timed) {
            float ms[5];
            for (int i = 0; i < 5; ++i) HV_HIP(hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]));
            st->ms_coarse += ms[0];
            st->ms_group += ms[1];
            st->ms_warmup += ms[2];
            st->ms_scan += ms[3];
            st->ms_select += ms[4];
            if (p.main_recorded) {
                float mm = 0.f;
                HV_HIP(hipEventElapsedTime(&mm, h->ev[6], h->ev[7]));
                st->ms_main += mm;
            }
            if (p.mfma) {
                float mr = 0.f;
                HV_HIP(hipEventElapsedTime(&mr, h->ev[8], h->ev[9]));
                st->ms_rerank += mr;
            }
            for (int i = 0; i < h->nscan_ev; ++i) {
                float ms1 = 0.f;
                HV_HIP(hipEventElapsedTime(&ms1, h->ev_scan[2 * i], h->ev_scan[2 * i + 1]));
                st->ms_scan_kernels += ms1;
            }
            st->scan_launches += h->nscan_ev;
        } else {
            float mg = 0.f;  // the whole device work of the graph
            HV_HIP(hipEventElapsedTime(&mg, h->ev[0], h->ev[5]));
            st->ms_scan += mg;
        }
        const int64_t cand = (int64_t)(hs[0] + hs[8]);
        st->candidates += cand;
        st->block_evals += (int64_t)(hs[1] + hs[9]);
        st->block_evals_max += cand * p.B;
        st->survivors += (int64_t)hs[2];
        st->overflow_queries += (int64_t)hs[12];
        st->fallback_queries += (int64_t)hs[13];
        st->widened_queries += (int64_t)hs[14];
        if (p.main_recorded) st->main_pairs += (int64_t)hs[8];
        st->scan_pairs += (int64_t)(hs[0] + hs[4] + hs[8]);
        if (p.mfma) st->rerank_pairs += (int64_t)hs[2];  // the select sees exactly the re-ranked pairs
        st->ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

// Inner-product search: Index::single_thread_search's IP branch (Index.cpp:1440-1449, 1509-1551)
// on the GPU. The nprobe centroids with the largest inner product (exact FP32 chains, ties to the
// lower id), then the exact top-k of the probed lists by (inner product descending, id ascending).
// With opt level OPT_TRIANGLE a query's lists are scanned only over the angular band a result
// better than its current k-th needs: the k-th after its nearest list (a first pass), so every
// list is cut against a value at most the final k-th (the reference cuts each list against the
// running k-th of the lists before it). For unit vectors the cut drops only vectors that cannot be
// results, so the results equal the unpruned search's.
void do_search_ip(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D, int64_t* I, hv_stats* st) {
    require(h->trained && h->added, HV_ERR_STATE, "search before train/add");
    require(n >= 0, HV_ERR_INVALID_ARG, "n < 0");
    require(k >= 1 && k <= hv::kMaxK, HV_ERR_CAPACITY, "k must be in [1, 1024]");
    require((h->sp.opt_level & h->added_opt_level) == h->sp.opt_level, HV_ERR_INVALID_ARG,
            "opt_level is not subset of added_opt_level");  // Index.cpp:1642-1645
    if (n == 0) return;
    require(q_dev && D && I, HV_ERR_INVALID_ARG, "null buffer");
    require(h->sp.precision == HV_PRECISION_FP32 && !h->sp.learned && h->query_nprobe.empty() && h->query_k.empty() &&
                h->sp.bound_mode == HV_BOUND_WARMUP,
            HV_ERR_INVALID_ARG,
            "the inner-product metric runs the FP32 form with one nprobe and one k per call (no learned "
            "predictor, no external bound)");
    require(!h->noalloc, HV_ERR_INVALID_ARG, "the no-allocation search is L2 only");
    const int64_t d = h->d, nlist = h->nlist;
    const int64_t nprobe = std::min<int64_t>(h->sp.nprobe, nlist);
    require(nprobe >= 1, HV_ERR_INVALID_ARG, "nprobe must be >= 1");
    require(nlist <= hv::kMaxSelectRow, HV_ERR_CAPACITY, "nlist > 8192");
    require((int64_t)n * nprobe < (1ll << 31), HV_ERR_CAPACITY, "n * nprobe >= 2^31");
    const bool tri = (h->sp.opt_level & HV_OPT_TRIANGLE) && h->c2c_sorted;
    hipStream_t s = h->stream;
    auto t0 = std::chrono::steady_clock::now();
    HV_HIP(hipEventRecord(h->ev[0], s));
    const float* Qp = q_dev;
    if (!h->perm.empty()) {
        HV_HIP(h->s_qp.ensure(sizeof(float) * n * d));
        hv::launch_permute_rows(q_dev, n, (int)d, h->d_perm.as<int>(), h->s_qp.as<float>(), s);
        Qp = h->s_qp.as<float>();
    }
    // coarse: inner-product rows in slices, the nprobe largest of each
    HV_HIP(h->s_probes.ensure(sizeof(int) * n * nprobe));
    HV_HIP(h->s_pdist.ensure(sizeof(float) * n * nprobe));
    const int64_t per = coarse_rows_per_slice(h, n);
    HV_HIP(h->s_dist.ensure(sizeof(float) * per * nlist));
    int* probes = h->s_probes.as<int>();
    float* pval = h->s_pdist.as<float>();
    for (int64_t r0 = 0; r0 < n; r0 += per) {
        const int64_t rows = std::min(per, n - r0);
        hv::launch_ip_rows(Qp + r0 * d, rows, h->d_centroids.as<float>(), nlist, (int)d, h->s_dist.as<float>(), nlist,
                           s);
        hv::launch_ip_select_rows(h->s_dist.as<float>(), rows, (int)nlist, (int)nprobe, probes + r0 * nprobe,
                                  pval + r0 * nprobe, s);
    }
    HV_CHECK_LAUNCH("ip coarse");
    HV_HIP(hipEventRecord(h->ev[1], s));
    // exact top-k over the probed lists, query batches sized to the run scratch
    HV_HIP(h->s_redo.ensure(sizeof(int) * (n + 1)));
    int* qcount = h->s_redo.as<int>();
    hv::launch_iota_list(qcount, n, s);
    const int nseg = hv::exact_segments(h->max_list);
    const int64_t run_keys = nprobe * nseg * k;
    const int batch = (int)std::max<int64_t>(1, std::min<int64_t>(n, kRedoRunBytes / (8 * run_keys)));
    HV_HIP(h->s_runs.ensure((sizeof(unsigned long long) * run_keys + sizeof(int)) * batch));
    HV_HIP(h->s_bound.ensure(sizeof(float) * n));
    // slack of the band for FP32 roundings: the products (d u each for unit vectors) and the square
    // roots of 1 - c^2 and 1 - s^2 (an error e there moves them by at most sqrt(e))
    const double du = (double)d * std::ldexp(1.0, -24);
    const float slack = (float)(4.0 * std::sqrt(2.0 * du) + 8.0 * du);
    unsigned long long* runs = h->s_runs.as<unsigned long long>();
    HV_HIP(h->s_stats.ensure(sizeof(unsigned long long) * 2));
    unsigned long long* scanned = h->s_stats.as<unsigned long long>();  // [0] first pass, [1] second
    hv::launch_fill_u64(scanned, 2, 0, s);
    for (int64_t q0 = 0; q0 < n; q0 += batch) {
        auto pass = [&](int r0, int r1, const float* simi, bool append, unsigned long long* cnt) {
            hv::launch_query_exact_ip(qcount + 1, qcount, (int)q0, batch, Qp, (int)d, h->d_codes.as<float>(),
                                      h->d_code_off.as<int64_t>(), h->d_list_off.as<int64_t>(),
                                      h->d_list_stride.as<int>(), h->d_ids.as<uint32_t>(), probes, (int)nprobe, r0, r1,
                                      nseg, h->d_c2c.as<float>(), pval, simi, slack, (int)k, runs, append, D, I, cnt,
                                      s);
        };
        if (tri && nprobe > 1) {
            // each query's nearest list first: its k-th bounds the final k-th from below; the
            // other lists then append their runs to it, each cut to the band that bound allows
            pass(0, 1, nullptr, false, scanned);
            hv::launch_ip_simi(qcount + 1, qcount, (int)q0, batch, D, I, (int)k, h->s_bound.as<float>(), s);
            pass(1, (int)nprobe, h->s_bound.as<float>(), true, scanned + 1);
        } else {
            pass(0, (int)nprobe, nullptr, false, scanned + 1);
        }
        HV_CHECK_LAUNCH("ip scan");
    }
    HV_HIP(hipEventRecord(h->ev[5], s));
    HV_HIP(hipStreamSynchronize(s));
    h->last_n = n;
    h->last_nprobe = nprobe;
    h->last_stride = nprobe;
    h->last_learned = false;
    if (st) {
        float mc = 0.f, ms = 0.f;
        HV_HIP(hipEventElapsedTime(&mc, h->ev[0], h->ev[1]));
        HV_HIP(hipEventElapsedTime(&ms, h->ev[1], h->ev[5]));
        st->ms_coarse += mc;
        st->ms_scan += ms;
        st->ms_scan_kernels += ms;
        unsigned long long cnt[2] = {0, 0};
        HV_HIP(hipMemcpy(cnt, scanned, sizeof(cnt), hipMemcpyDeviceToHost));
        st->scan_launches += tri && nprobe > 1 ? 2 : 1;
        st->scan_pairs += (int64_t)(cnt[0] + cnt[1]);
        st->ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

void do_search(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D, int64_t* I, hv_stats* st) {
    if (h->metric == HV_METRIC_IP) {
        do_search_ip(h, n, q_dev, k, D, I, st);
        return;
    }
    require(h->trained && h->added, HV_ERR_STATE, "search before train/add");
    require(n >= 0, HV_ERR_INVALID_ARG, "n < 0");
    require(k >= 1 && k <= hv::kMaxK, HV_ERR_CAPACITY, "k must be in [1, 1024]");
    require((h->sp.opt_level & h->added_opt_level) == h->sp.opt_level, HV_ERR_INVALID_ARG,
            "opt_level is not subset of added_opt_level");  // Index.cpp:1642-1645
    if (n == 0) return;
    require(q_dev && D && I, HV_ERR_INVALID_ARG, "null buffer");
    const int64_t d = h->d, nlist = h->nlist;
    int64_t nprobe = std::min<int64_t>(h->sp.nprobe, nlist);  // Index.cpp:1646-1648
    require(nprobe >= 1, HV_ERR_INVALID_ARG, "nprobe must be >= 1");
    const int B = h->sp.pruning ? h->sp.dim_blocks : 1;
    require(h->sp.dim_blocks >= 1 && h->sp.dim_blocks <= d, HV_ERR_INVALID_ARG, "dim_blocks must be in [1, d]");
    require(nlist <= hv::kMaxSelectRow, HV_ERR_CAPACITY, "nlist > 8192");
    const hv::ChunkPlan plan = make_plan((int)d, B);
    // learned predictor: per-query probe counts (at most nprobe) and seeded bounds; its features
    // (the 64 smallest coarse distance estimates) come from the coarse rows beside the selection
    const bool learned = h->sp.learned != 0;
    const bool uneven = !h->query_nprobe.empty();
    if (uneven) {
        require((int64_t)h->query_nprobe.size() == n, HV_ERR_STATE, "per-query nprobe not set for this batch");
        require(!learned, HV_ERR_INVALID_ARG, "per-query nprobe and the learned predictor are exclusive");
    }
    if (learned) {
        require(h->mlp_params && h->mlp_norm, HV_ERR_STATE,
                "learned search needs a predictor (hv_mlp_init/train or set_params, and hv_mlp_set_norm)");
        // the features come from the coarse rows: the matrix-core estimates whenever
        // hv_query_features would take them too
        require(coarse_on_mfma(h, nprobe) == coarse_on_mfma(h, 1), HV_ERR_INVALID_ARG,
                "the learned search takes nprobe <= 512 (its features come from the coarse estimates)");
    }
    const bool kq_on = !h->query_k.empty();
    if (kq_on) {
        require((int64_t)h->query_k.size() == n, HV_ERR_STATE, "per-query k not set for this batch");
        require(!learned, HV_ERR_INVALID_ARG, "per-query k and the learned predictor are exclusive");
    }
    const bool fp8 = h->sp.precision == HV_PRECISION_FP8;
    const bool bf16 = h->sp.precision == HV_PRECISION_BF16;
    const bool mfma = bf16 || fp8;
    const int64_t nprobe_cap = nprobe;
    const int64_t rank_end = nprobe;
    int64_t wl = learned ? 0 : std::min<int64_t>(h->sp.warmup_list, nprobe);  // Index.cpp:1007-1010
    int64_t wls = h->sp.warmup_list_size == 0 ? k : h->sp.warmup_list_size;  // query.cpp:328-333
    if (h->sp.bound_mode == HV_BOUND_WARMUP && wl > 0)
        require(wl * wls <= hv::kMaxWarmup, HV_ERR_CAPACITY, "warmup_list * warmup_list_size > 8192");
    require(h->sp.bound_mode != HV_BOUND_EXTERNAL || (!learned && !fp8), HV_ERR_INVALID_ARG,
            "HV_BOUND_EXTERNAL cannot be combined with the FP8 form or the learned predictor");
    const bool external = h->sp.bound_mode == HV_BOUND_EXTERNAL;
    if (external)
        require((int64_t)h->ext_bounds.size() == n, HV_ERR_STATE, "external bounds not set for this batch");
    // FP8 form: the shortlist of the kr smallest estimates per query is re-ranked exactly
    const int64_t kr = fp8 ? k * std::max(1, h->sp.rerank_r) : k;
    int64_t cap = h->sp.survivor_cap;
    if (cap == 0) cap = std::min<int64_t>(fp8 ? 4096 : 8192, std::max<int64_t>(4096, hv_next_pow2((int)(4 * k))));
    require(cap >= k && cap <= 8192, HV_ERR_INVALID_ARG, "survivor_cap must be in [k, 8192]");
    if (fp8) {
        require(cap <= 4096, HV_ERR_INVALID_ARG, "survivor_cap must be <= 4096 in the FP8 form");
        require(kr <= cap, HV_ERR_CAPACITY, "k * rerank_r must not exceed the survivor buffer (4096)");
    }
    require((int64_t)n * nprobe < (1ll << 31), HV_ERR_CAPACITY, "n * nprobe >= 2^31");

    hipStream_t s = h->stream;
    auto t0 = std::chrono::steady_clock::now();
    // per-query probe counts, clamped, in pinned memory (a captured search copies them from there;
    // a no-allocation search uses the host vector sized when the counts were set)
    auto stage_qn = [&]() -> const int* {
        int* qh;
        if (h->noalloc) {
            require((int64_t)h->qn_stage.size() >= n, HV_ERR_STATE, "per-query nprobe staging not sized");
            qh = h->qn_stage.data();
        } else {
            HV_HIP(h->qn_host.ensure(sizeof(int) * n));
            qh = static_cast<int*>(h->qn_host.p);
        }
        for (int64_t i = 0; i < n; ++i) qh[i] = (int)std::min<int64_t>(nprobe, std::max<int64_t>(1, h->query_nprobe[i]));
        return qh;
    };
    // per-query k, clamped to [1, k], staged the same way
    auto stage_qk = [&]() -> const int* {
        int* kh;
        if (h->noalloc) {
            require((int64_t)h->qk_stage.size() >= n, HV_ERR_STATE, "per-query k staging not sized");
            kh = h->qk_stage.data();
        } else {
            HV_HIP(h->qk_host.ensure(sizeof(int) * n));
            kh = static_cast<int*>(h->qk_host.p);
        }
        for (int64_t i = 0; i < n; ++i) kh[i] = (int)std::min<int64_t>(k, std::max<int64_t>(1, h->query_k[i]));
        return kh;
    };
    // graph mode: the device work of a search is captured on the second call with the same
    // signature (the first one sizes every buffer) and replayed from then on
    const bool graph_ok = h->sp.graph != 0 && !external && !h->noalloc;
    std::vector<unsigned char> sig;
    if (graph_ok) {
        auto put = [&](const void* v, size_t b) {
            const unsigned char* c = static_cast<const unsigned char*>(v);
            sig.insert(sig.end(), c, c + b);
        };
        const int64_t shape[7] = {n, k, h->ntotal, h->max_list, (int64_t)uneven, (int64_t)h->fp8_built, (int64_t)kq_on};
        const void* ptrs[3] = {q_dev, D, I};
        put(shape, sizeof(shape));
        put(ptrs, sizeof(ptrs));
        put(&h->sp, sizeof(h->sp));
        if (h->graph.exec && h->graph.sig == sig && h->graph.gen == g_alloc_gen.load()) {
            if (uneven) stage_qn();
            if (kq_on) stage_qk();
            HV_HIP(hipEventRecord(h->ev[0], s));
            HV_HIP(hipGraphLaunch(h->graph.exec, s));
            HV_HIP(hipEventRecord(h->ev[5], s));
            finish_search(h, h->graph.post, n, st, false, t0);
            return;
        }
    }
    // Every scratch buffer of the search is sized here, before the first launch: a no-allocation
    // search carves them all from its arena (or fails before doing anything); later ensure() calls
    // of the same buffers are no-ops.
    const int64_t nb = nlist * rank_end;  // (list, rank) buckets of the searched ranks
    const bool learned_seed = learned && !fp8;  // the seed is the starting bound (FP32 and BF16)
    const bool bound_fmax = !learned_seed && !external && !(wl > 0 && !mfma && !learned);
    // the exact redo's run scratch: batches of queries, each with nprobe_cap * nseg runs of k keys,
    // then one count per query of the batch
    const int nseg = hv::exact_segments(h->max_list);
    const int64_t run_keys = nprobe_cap * nseg * k;  // run keys of one redone query
    const int64_t redo_bytes = h->noalloc ? kRedoRunBytesNoAlloc : kRedoRunBytes;
    const int batch = (int)std::max<int64_t>(1, std::min<int64_t>(n, redo_bytes / (8 * run_keys)));
    // work items of one phase: at most (query tiles) x (vector ranges); query tiles hold >= 16
    // queries, vector ranges >= kVecSplit vectors
    const int vsplit_mfma = h->sp.scan_vsplit > 0 ? h->sp.scan_vsplit : kVecSplitBf16;
    const int64_t qcap = (n * nprobe / 16 + nlist + 1) * (h->max_list / std::min(kVecSplit, vsplit_mfma) + 1);
    require(qcap < (1ll << 31), HV_ERR_CAPACITY, "too many work items");
    // survivor buffer: room for the warm-up set too
    if (!learned && h->sp.bound_mode == HV_BOUND_WARMUP && wl > 0)
        cap = std::max<int64_t>(cap, hv_next_pow2((int)(wl * wls)));
    // MFMA forms: the scan phases (probe ranks [r0, r1), vectors [vmin, vmax)) in the order they run
    struct PhaseRange {
        int64_t r0, r1;
        int vmin, vmax;
        bool main;
        bool nearest;  // the nearest-list pass: items of kNearQueries queries over whole lists
    };
    // BF16 form: each query's nearest list scanned and reduced in one pass (candidates in LDS)
    const bool near_pass = bf16 && h->sp.nearest_pass != 0 && k <= hv::kNearMaxK;
    // learned BF16 with the nearest-list pass: the pass takes ranks [0, near_end) at once (every
    // row appends and lowers its query's bound; near_end <= cap / (2k), so what the items of one
    // query keep, about k or a little more each, fits its buffer)
    const int64_t near_end =
        near_pass && learned_seed
            ? std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(rank_end, kBf16SinglePhase), cap / (2 * k)))
            : 1;
    // learned BF16: the nearest lists (or the pass's ranks), then the other searched ranks in one
    // group when there are few of them (as the plain BF16 form takes up to kBf16SinglePhase
    // probes), else in groups of growing width [1, 4), [4, 16), ..., each followed by a compaction
    // that lowers the bound
    auto learned_rank_group_end = [&](int64_t r0) -> int64_t {
        if (r0 == 0) return near_pass ? near_end : 1;
        return rank_end <= kBf16SinglePhase ? rank_end : std::min<int64_t>(rank_end, r0 * 4);
    };
    std::vector<PhaseRange> ranges;
    if (mfma) {
        if (learned_seed) {
            for (int64_t r0 = 0; r0 < rank_end;) {
                const int64_t r1 = learned_rank_group_end(r0);
                ranges.push_back({r0, r1, 0, INT32_MAX, r0 != 0, near_pass && r0 == 0});
                r0 = r1;
            }
        } else {
            const int vdirect = (int)(cap / 16 * 16);
            if (near_pass) {
                ranges.push_back({0, 1, 0, INT32_MAX, false, true});
            } else {
                if (vdirect > 0) ranges.push_back({0, 1, 0, vdirect, false, false});
                if (h->max_list > vdirect) ranges.push_back({0, 1, vdirect, INT32_MAX, false, false});
            }
            if (!fp8 && nprobe <= kBf16SinglePhase) {
                if (nprobe > 1) ranges.push_back({1, nprobe, 0, INT32_MAX, true, false});
            } else {
                const int64_t end = fp8 ? rank_end : nprobe;
                for (int64_t r0 = 1; r0 < end;) {
                    const int64_t r1 = std::min<int64_t>(end, r0 * 4);
                    ranges.push_back({r0, r1, 0, INT32_MAX, true, false});
                    r0 = r1;
                }
            }
        }
        require((int)ranges.size() <= hv::kMaxPlanPhases, HV_ERR_CAPACITY, "too many scan phases");
    }
    const int64_t qlen = 3 * hv::kNumXcd + nlist;
    const int64_t nph = std::max<int64_t>(1, (int64_t)ranges.size());
    {
        auto need = [&](DevBuf& b, int64_t bytes) { HV_HIP(b.ensure((size_t)std::max<int64_t>(bytes, 1))); };
        if (!h->perm.empty()) need(h->s_qp, 4 * n * d);
        need(h->s_bucket, 4 * nb);
        need(h->s_fill, 4 * nb);
        need(h->s_stats, 8 * 17);
        need(h->s_surv_cnt, 4 * n);
        need(h->s_bound, 4 * n);
        need(h->s_bound_ref, 4 * n);
        need(h->s_overflow, 4 * (n + 2));
        if (mfma) need(h->s_last, 4 * n);
        need(h->s_runs, (8 * run_keys + 4) * batch);
        need(h->s_probes, 4 * n * nprobe);
        need(h->s_pdist, 4 * n * nprobe);
        const int64_t per = coarse_rows_per_slice(h, n);
        need(h->s_dist, 4 * per * nlist);
        if (coarse_on_mfma(h, nprobe)) {
            need(h->s_qnorm, 4 * n);
            need(h->s_cnorm, 4 * nlist);
            need(h->s_redo, 4 * (per + 1));
        }
        if (learned) {
            need(h->s_qn, 4 * n);
            need(h->s_seed, 4 * n);
            need(h->s_pred, 8 * n);
            need(h->s_near, 4 * n * hv::kNearest);
        } else if (uneven) {
            need(h->s_qn, 4 * n);
        }
        if (kq_on) need(h->s_qk, 4 * n);
        need(h->s_bucket_off, 4 * (nb + 1));
        need(h->s_scan_tmp, 4 * hv::scan_tmp_ints(nb));
        need(h->s_pairs, 4 * n * nprobe);
        need(h->s_pair_slot, 4 * n * nprobe);
        need(h->s_items, (int64_t)sizeof(hv::ScanItem) * qcap * nph);
        need(h->s_queue, 4 * qlen * nph);
        need(h->s_surv, 8 * n * cap);
        if (mfma) {
            need(h->s_aux, 4 * n * cap);
            need(h->s_qres, (fp8 ? 1 : 2) * n * nprobe * h->dpad);
            need(h->s_qscale, 4 * n * nprobe);
            need(h->s_ritems, 8 * n * ((cap + 63) / 64));
            need(h->s_kept, 8 * n * cap);
            need(h->s_done, 4 * n);
        }
    }
    if (h->size_only) return;
    const bool capture = graph_ok && h->graph.pending_sig == sig && h->graph.pending_gen == g_alloc_gen.load();
    auto rec = [&](int i) {
        if (!capture) HV_HIP(hipEventRecord(h->ev[i], s));
    };
    h->nscan_ev = 0;
    // events around one candidate-scan launch (eager searches; launches past kScanEv go untimed)
    auto scan_ev = [&](bool end) {
        if (capture || h->nscan_ev >= hv_index::kScanEv) return;
        HV_HIP(hipEventRecord(h->ev_scan[2 * h->nscan_ev + (end ? 1 : 0)], s));
        if (end) ++h->nscan_ev;
    };
    HV_HIP(hipEventRecord(h->ev[0], s));
    if (capture) {
        h->graph.reset();
        HV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    }
    SearchPost post;
    // an error thrown while capturing ends the capture (the stream stays usable)
    struct CaptureGuard {
        hipStream_t s;
        bool active;
        ~CaptureGuard() {
            if (!active) return;
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
        }
    } guard{s, capture};

    // 1. queries in index dim order
    const float* Qp = q_dev;
    if (!h->perm.empty()) {
        HV_HIP(h->s_qp.ensure(sizeof(float) * n * d));
        hv::launch_permute_rows(q_dev, n, (int)d, h->d_perm.as<int>(), h->s_qp.as<float>(), s);
        Qp = h->s_qp.as<float>();
    }
    // counters, bucket counts and (when no warm-up, seed or caller's bound applies) the bounds
    // start cleared: one fill launch for all of them
    HV_HIP(h->s_bucket.ensure(sizeof(int) * nb));
    HV_HIP(h->s_fill.ensure(sizeof(int) * nb));
    // stats: [0..3] counters of the scan phases, [4..7] of the warm-up, [8..11] of the main phase
    // (the remaining probed lists), [12] overflowed buffers, [13] learned-mode redos, [14] learned
    // queries whose probe count the margin test widened, [15] the redo list's count (low word) and
    // the re-rank item count (high word): all published to the host by the redo-list kernel, which
    // counts its finished workgroups in [16]
    HV_HIP(h->s_stats.ensure(sizeof(unsigned long long) * 17));
    HV_HIP(h->s_surv_cnt.ensure(sizeof(unsigned int) * n));
    HV_HIP(h->s_bound.ensure(sizeof(float) * n));
    HV_HIP(h->s_bound_ref.ensure(sizeof(float) * n));
    HV_HIP(h->s_overflow.ensure(sizeof(int) * (n + 2)));
    if (mfma) HV_HIP(h->s_last.ensure(sizeof(unsigned int) * n));
    HV_HIP(h->s_runs.ensure((sizeof(unsigned long long) * run_keys + sizeof(int)) * batch));
    hv::FillSegs f{};
    {
        auto add = [&](void* p, int64_t words, int v) {
            f.p[f.nseg] = static_cast<int*>(p);
            f.n[f.nseg] = words;
            f.v[f.nseg] = v;
            ++f.nseg;
        };
        add(h->s_bucket.p, nb, 0);
        add(h->s_fill.p, nb, 0);
        add(h->s_stats.p, 34, 0);
        add(h->s_surv_cnt.p, n, 0);
        if (coarse_on_mfma(h, nprobe)) add(coarse_counter(h, n), 1, 0);
        if (mfma) add(h->s_last.p, n, -1);  // no compaction yet: counts never match
        add(h->s_runs.as<unsigned long long>() + batch * run_keys, batch, 0);  // first redo batch's counts
        if (bound_fmax) {
            int fm;
            const float fmax = FLT_MAX;
            std::memcpy(&fm, &fmax, 4);
            add(h->s_bound.p, n, fm);
            add(h->s_bound_ref.p, n, fm);
        }
        // (cleared by the coarse step's first launch)
    }
    // 2. coarse: full distance rows + nprobe select (Index.cpp:967-997)
    HV_HIP(h->s_probes.ensure(sizeof(int) * n * nprobe));
    HV_HIP(h->s_pdist.ensure(sizeof(float) * n * nprobe));
    int* probes = h->s_probes.as<int>();
    float* pdist = h->s_pdist.as<float>();
    // the grouping's pair counts: by the coarse select when every query searches ranks [0, rank_end)
    // (a learned search's predictor counts its own)
    bool counted = coarse_select(h, n, Qp, nprobe, probes, pdist, true, learned ? h->s_near.as<float>() : nullptr,
                                 learned || uneven ? nullptr : h->s_bucket.as<int>(), (int)rank_end, &f);
    const int* qn = nullptr;
    if (learned) {
        HV_HIP(h->s_qn.ensure(sizeof(int) * n));
        HV_HIP(h->s_seed.ensure(sizeof(float) * n));
        HV_HIP(h->s_pred.ensure(sizeof(float) * n * 2));
        // margin widening only where the search verifies the seed afterwards (FP32 and BF16); the
        // slack covers the FP32 roundings of the coarse distances, residual norms and square roots
        const float slack = (float)(4.0 * d * std::ldexp(1.0, -24) + std::ldexp(1.0, -20));
        hv::launch_mlp_serve(h->s_near.as<float>(), pdist, probes, (int)nprobe, Qp, (int)d, n, h->m_w1b.as<__bf16>(),
                             h->m_w2b.as<__bf16>(),
                             h->m_params.as<float>(), h->m_norm.as<float>(), h->sp.learned_nprobe_scale,
                             h->sp.learned_bound_scale, std::max(1, h->sp.learned_min_nprobe), (int)nprobe_cap,
                             h->d_lrad.as<float>(), h->d_centroids.as<float>(),
                             learned_seed ? h->sp.learned_margin : 0.f, slack, h->d_list_off.as<int64_t>(), (int)k,
                             h->s_qn.as<int>(), h->s_seed.as<float>(), h->s_pred.as<float>(),
                             learned_seed ? h->s_bound.as<float>() : nullptr,
                             learned_seed ? h->s_bound_ref.as<float>() : nullptr, h->s_stats.as<unsigned long long>() + 14,
                             h->s_bucket.as<int>(), (int)rank_end, s);
        counted = true;
        HV_CHECK_LAUNCH("predictor");
        qn = h->s_qn.as<int>();
    } else if (uneven) {
        const int* qh = stage_qn();
        HV_HIP(h->s_qn.ensure(sizeof(int) * n));
        HV_HIP(hipMemcpyAsync(h->s_qn.p, qh, sizeof(int) * n, hipMemcpyHostToDevice, s));
        qn = h->s_qn.as<int>();
    }
    const int* kq = nullptr;
    if (kq_on) {
        const int* kh = stage_qk();
        HV_HIP(hipMemcpyAsync(h->s_qk.p, kh, sizeof(int) * n, hipMemcpyHostToDevice, s));
        kq = h->s_qk.as<int>();
    }
    rec(1);
    // 3. grouping by (list, rank)
    HV_HIP(h->s_bucket_off.ensure(sizeof(int) * (nb + 1)));
    HV_HIP(h->s_scan_tmp.ensure(sizeof(int) * hv::scan_tmp_ints(nb)));
    HV_HIP(h->s_pairs.ensure(sizeof(int) * n * nprobe));
    HV_HIP(h->s_pair_slot.ensure(sizeof(int) * n * nprobe));
    HV_HIP(h->s_items.ensure(sizeof(hv::ScanItem) * qcap));
    HV_HIP(h->s_queue.ensure(sizeof(int) * (3 * hv::kNumXcd + nlist)));
    unsigned long long* stats = h->s_stats.as<unsigned long long>();
    if (!counted) hv::launch_count_pairs(probes, n, (int)nprobe, (int)rank_end, qn, h->s_bucket.as<int>(), s);
    hv::launch_exclusive_scan(h->s_bucket.as<int>(), nb, h->s_bucket_off.as<int>(), h->s_scan_tmp.as<int>(), s);
    hv::launch_scatter_pairs(probes, n, (int)nprobe, (int)rank_end, qn, h->s_bucket_off.as<int>(), h->s_fill.as<int>(),
                             h->s_pairs.as<int>(), h->s_pair_slot.as<int>(), s);
    HV_CHECK_LAUNCH("grouping");
    rec(2);

    int* queue = h->s_queue.as<int>();
    HV_HIP(h->s_surv.ensure(sizeof(unsigned long long) * n * cap));
    unsigned long long* surv = h->s_surv.as<unsigned long long>();
    unsigned int* surv_cnt = h->s_surv_cnt.as<unsigned int>();
    // one scan phase over probe ranks [r0, r1), at most vmax vectors per list
    // Tile shape: the caller's, or by the expected number of queries per list in the phase
    // (few queries per list -> narrow query tiles, long vector tiles).
    auto pick_shape = [&](int r0, int r1) {
        if (h->sp.tile_shape > 0 && r1 == (int)rank_end && r1 - r0 > 1) return h->sp.tile_shape - 1;
        if (h->sp.tile_shape_first > 0 && r0 == 0 && r1 == 1) return h->sp.tile_shape_first - 1;
        const double g = (double)n * (r1 - r0) / (double)nlist;
        return g <= 16 ? 5 : (g <= 40 ? 2 : 0);
    };
    bool main_recorded = false;  // events 6-7 bracket the main phase's scan kernel
    auto scan_phase = [&](int r0, int r1, int vmax, int prune, int accept_all, unsigned long long* st_slot) {
        const int shape = pick_shape(r0, r1);
        if (h->scan_grid[shape] == 0) h->scan_grid[shape] = hv::scan_grid_size(h->device, shape, (int)d);
        const int vt = hv::scan_vector_tile(shape);
        const int vsplit = (kVecSplit + vt - 1) / vt * vt;
        hv::launch_plan_items(h->s_bucket_off.as<int>(), (int)nlist, (int)rank_end, r0, r1, 0, vmax,
                              hv::scan_query_tile(shape), vsplit, h->d_rank_order.as<int>(),
                              h->d_list_off.as<int64_t>(), qcap, h->s_items.as<hv::ScanItem>(), queue, st_slot, s);
        const bool main_phase = st_slot == stats + 8;
        main_recorded |= main_phase;
        if (main_phase) rec(6);
        scan_ev(false);
        hv::launch_list_scan(shape, Qp, (int)d, h->d_codes.as<float>(), h->d_code_off.as<int64_t>(),
                             h->d_list_off.as<int64_t>(), h->d_list_stride.as<int>(), h->d_ids.as<uint32_t>(), plan,
                             prune, h->s_bound.as<float>(), h->s_pairs.as<int>(), h->s_items.as<hv::ScanItem>(),
                             queue, vmax, vsplit, accept_all, surv, surv_cnt, (int)cap, st_slot, h->scan_grid[shape],
                             s);
        HV_CHECK_LAUNCH("list scan");
        scan_ev(true);
        if (main_phase) rec(7);
    };
    // BF16 precision form: one phase of the MFMA filter scan over probe ranks [r0, r1)
    // MFMA filter forms: BF16 (certified error bound, exact results) and FP8 (estimates only, a
    // shortlist of k * rerank_r per query re-ranked exactly)
    float* aux = nullptr;
    unsigned int* last = mfma ? h->s_last.as<unsigned int>() : nullptr;  // compaction counts
    float kappa1 = 0.f, kappa_g = 0.f, kappa2 = 0.f;
    if (mfma) {
        HV_HIP(h->s_aux.ensure(sizeof(float) * n * cap));
        aux = h->s_aux.as<float>();
        if (h->bf16_grid == 0) h->bf16_grid = hv::bf16_grid_size(h->device);
    }
    if (bf16) {
        // error bound of the BF16 estimate (hv_bf16.hip): the measured rounding-error norms of
        // both residuals (factor 2 of the cross term), FP32 accumulation over d terms (twice
        // g_d for the MFMA's order), FP32 norms, the FP32 evaluation of the estimate and the FP32
        // chain of the exact path; 1 % on top for the FP32 arithmetic of the bound itself
        const double u = std::ldexp(1.0, -24);
        const double gd = d * u / (1.0 - d * u);
        kappa1 = (float)(1.01 * 2.0);
        kappa_g = (float)(1.01 * 2.0 * 2.0 * gd);
        kappa2 = (float)(1.01 * (4.0 * d + 2.0 * B + 32.0) * u);
        HV_HIP(h->s_qres.ensure(sizeof(__bf16) * n * nprobe * h->dpad));
        HV_HIP(h->s_qscale.ensure(sizeof(float) * n * nprobe));
        // (the nearest-list pass forms its ranks' residuals itself)
        const int r_first = near_pass ? (int)near_end : 0;
        if (rank_end > r_first) {
            hv::launch_pair_resid(Qp, h->d_centroids.as<float>(), probes, qn, n, (int)nprobe, r_first, (int)rank_end,
                                  (int)d, h->dpad, h->s_qres.as<__bf16>(), h->s_qscale.as<float>(), s);
            HV_CHECK_LAUNCH("pair residuals");
        }
    }
    if (fp8) {
        if (!h->fp8_built) {
            HV_HIP(h->d_rcodes8.ensure(std::max<int64_t>(1, h->rcode_elems)));
            HV_HIP(h->d_lscale8.ensure(sizeof(float) * nlist));
            hv::launch_build_resid_fp8(h->d_rows.as<float>(), (int)d, h->dpad, h->d_list_off.as<int64_t>(),
                                       h->d_centroids.as<float>(), h->d_roff.as<int64_t>(), (int)nlist,
                                       h->d_rcodes8.as<uint8_t>(), h->d_lscale8.as<float>(), s);
            HV_CHECK_LAUNCH("fp8 codes");
            h->fp8_built = true;
        }
        HV_HIP(h->s_qres.ensure(n * nprobe * h->dpad));
        HV_HIP(h->s_qscale.ensure(sizeof(float) * n * nprobe));
        hv::launch_pair_resid_fp8(Qp, h->d_centroids.as<float>(), probes, qn, n, (int)nprobe, (int)rank_end, (int)d,
                                  h->dpad, h->s_qres.as<uint8_t>(), h->s_qscale.as<float>(), s);
        HV_CHECK_LAUNCH("pair residuals");
    }
    // MFMA forms: every scan phase (ranks, vector range), in the order the phases run below, is
    // planned here in one pair of launches, each phase into its own queue and items
    std::vector<hv::PlanSpec> mph;
    size_t mph_next = 0;
    if (mfma) {
        for (const PhaseRange& r : ranges)
            mph.push_back(hv::PlanSpec{(int)r.r0, (int)r.r1, r.vmin, r.vmax, nullptr, nullptr, r.main ? stats + 8 : stats,
                                       r.nearest ? hv::kNearQueries : 0, r.nearest ? hv::kNearSplit : 0});
        HV_HIP(h->s_items.ensure(sizeof(hv::ScanItem) * qcap * std::max<size_t>(1, mph.size())));
        HV_HIP(h->s_queue.ensure(sizeof(int) * qlen * std::max<size_t>(1, mph.size())));
        for (size_t i = 0; i < mph.size(); ++i) {
            mph[i].queue = h->s_queue.as<int>() + i * qlen;
            mph[i].items = h->s_items.as<hv::ScanItem>() + i * qcap;
        }
        hv::launch_plan_phases(h->s_bucket_off.as<int>(), (int)nlist, (int)rank_end, 64, vsplit_mfma,
                               h->d_rank_order.as<int>(), h->d_list_off.as<int64_t>(), qcap, mph.data(),
                               (int)mph.size(), s);
        HV_CHECK_LAUNCH("plan");
    }
    auto mfma_phase = [&](int r0, int r1, int vmin, int vmax, int mode, unsigned long long* st_slot) {
        // the next planned phase must be this one
        require(mph_next < mph.size() && mph[mph_next].r0 == r0 && mph[mph_next].r1 == r1 &&
                    mph[mph_next].vmin == vmin && mph[mph_next].vmax == vmax && mph[mph_next].stats == st_slot,
                HV_ERR_STATE, "scan phase out of plan");
        const hv::PlanSpec& ph = mph[mph_next++];
        const bool main_phase = st_slot == stats + 8;
        if (main_phase && !main_recorded) rec(6);  // events 6-7 span the main phases
        main_recorded |= main_phase;
        // per slot: FP8 inverse scales or BF16 rounding-error norms (s_qscale holds either)
        scan_ev(false);
        hv::launch_mfma_scan(h->bf16_grid, fp8 ? 1 : 0, h->s_qres.p, h->s_qscale.as<float>(), (int)d, pdist,
                             (int)nprobe, fp8 ? h->d_rcodes8.p : h->d_rcodes.p,
                             fp8 ? h->d_lscale8.as<float>() : nullptr, h->d_roff.as<int64_t>(),
                             h->d_rnorm.as<float>(), fp8 ? nullptr : h->d_rerr.as<float>(),
                             h->d_list_off.as<int64_t>(), h->s_bound.as<float>(), h->s_pair_slot.as<int>(),
                             ph.items, ph.queue, vmax, vsplit_mfma, mode, kappa1, kappa_g, kappa2, surv, aux,
                             surv_cnt, (int)cap, st_slot, s);
        HV_CHECK_LAUNCH("mfma scan");
        scan_ev(true);
        if (main_phase) rec(7);
    };
    // the nearest-list pass of the BF16 form (ranks [0, near_end), planned with its own query tile)
    auto near_phase = [&](unsigned long long* st_slot) {
        require(mph_next < mph.size() && mph[mph_next].r0 == 0 && mph[mph_next].r1 == near_end &&
                    mph[mph_next].qtile == hv::kNearQueries && mph[mph_next].stats == st_slot,
                HV_ERR_STATE, "nearest-list pass out of plan");
        const hv::PlanSpec& ph = mph[mph_next++];
        // rank 0 alone: at most n / 16 + nlist + ntotal / kNearSplit items, one workgroup each,
        // the dispatcher balancing their lengths (more ranks: the workgroups loop over the rest)
        const int grid = (int)std::min<int64_t>(
            1 << 15, (n / hv::kNearQueries + nlist + h->ntotal / hv::kNearSplit + hv::kNumXcd) / hv::kNumXcd * hv::kNumXcd);
        scan_ev(false);
        hv::launch_mfma_nearest(grid, Qp, h->d_centroids.as<float>(), (int)d, pdist, (int)nprobe,
                                h->d_rcodes.as<__bf16>(), h->d_roff.as<int64_t>(), h->d_rnorm.as<float>(),
                                h->d_rerr.as<float>(), h->d_lrad.as<float>(), h->d_list_off.as<int64_t>(),
                                h->s_bound.as<float>(), learned_seed ? nullptr : h->s_bound_ref.as<float>(),
                                h->s_pair_slot.as<int>(), ph.items, ph.queue, kappa1, kappa_g, kappa2, (int)k, kq, surv,
                                aux, surv_cnt, last, (int)cap, near_end > 1 ? 1 : 0, st_slot, s);
        HV_CHECK_LAUNCH("nearest-list pass");
        scan_ev(true);
    };
    // 4. bound: the reference's warm-up bound (heapTops), or the caller's. The BF16 form needs no
    //    warm-up scan: it scans each query's nearest list unpruned and takes the bound from there.
    float* bound = h->s_bound.as<float>();
    float* bound_ref = h->s_bound_ref.as<float>();
    // learned (FP32 and BF16): the seed from the start, no warm-up. The FP8 form starts without a
    // bound and filters on estimates only.
    if (learned_seed) {
        // the predictor wrote the seed into both
    } else if (external) {
        HV_HIP(hipMemcpyAsync(bound_ref, h->ext_bounds.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
        HV_HIP(hipMemcpyAsync(bound, bound_ref, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    } else if (wl > 0 && !mfma && !learned) {
        scan_phase(0, (int)wl, (int)wls, 0, 1, stats + 4);
        hv::launch_kth_survivors(surv, surv_cnt, n, (int)cap, (int)k, bound_ref, bound, kq, s);
    }  // otherwise both start at FLT_MAX (the fill above)
    HV_CHECK_LAUNCH("bound");
    rec(3);
    // 5. list scan: each query's nearest list first, then (bound lowered to the running k-th)
    //    the remaining probed lists
    const int prune = B > 1 ? 1 : 0;
    if (mfma && learned_seed) {
        // learned, BF16: the seed filters from the nearest lists on; the ranks go in groups
        // ([0, near_end) in the nearest-list pass, or [0, 1)), then [.., 4), [4, 16), ..., each
        // followed by the compaction, which lowers the bound to the k-th smallest upper bound so
        // far (when below the seed) and drops the pairs whose lower bound exceeds it
        for (int64_t r0 = 0; r0 < rank_end;) {
            const int64_t r1 = learned_rank_group_end(r0);
            if (r0 == 0 && near_pass)
                near_phase(stats);  // reduces each item as it goes; the compaction below merges
            else
                mfma_phase((int)r0, (int)r1, 0, INT32_MAX, 0, r0 == 0 ? stats : stats + 8);
            hv::launch_bf16_compact(surv, aux, surv_cnt, n, (int)cap, (int)k, bound, last, kq, 1, s);
            r0 = r1;
        }
        HV_CHECK_LAUNCH("compact");
    } else if (mfma) {
        // nearest list: vectors below cap go to their own slot unfiltered; the k-th smallest upper
        // bound (FP8: the kr-th smallest estimate) then lowers the bound and the pairs whose lower
        // bound exceeds it are dropped (vector ranges start on 16-vector blocks)
        const int vdirect = (int)(cap / 16 * 16);
        if (near_pass) {
            // the pass and this compaction (of the queries whose nearest list was split into
            // several items; the others are skipped) leave bound_ref = bound, the nearest list's
            // bound
            near_phase(stats);
            hv::launch_bf16_compact(surv, aux, surv_cnt, n, (int)cap, (int)k, bound, last, kq, 1, s, bound_ref);
            HV_CHECK_LAUNCH("compact");
        } else {
            if (vdirect > 0) mfma_phase(0, 1, 0, vdirect, 1, stats);
            hv::launch_first_counts(probes, n, (int)nprobe, h->d_list_off.as<int64_t>(), vdirect, surv_cnt, s);
            hv::launch_bf16_compact(surv, aux, surv_cnt, n, (int)cap, (int)kr, bound, last, kq, (int)(kr / k), s);
            HV_CHECK_LAUNCH("compact");
            HV_HIP(hipMemcpyAsync(bound_ref, bound, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        }
        // the rest of the nearest lists (longer than the direct part), then the other probed lists
        if (!near_pass && h->max_list > vdirect) mfma_phase(0, 1, vdirect, INT32_MAX, 0, stats);
        if (!fp8 && nprobe <= kBf16SinglePhase) {
            if (nprobe > 1) mfma_phase(1, (int)nprobe, 0, INT32_MAX, 0, stats + 8);
        } else if (!fp8) {
            // many probes: the ranks go in groups of growing width, each followed by the
            // compaction, which lowers the bound to the k-th smallest upper bound so far and drops
            // the pairs whose lower bound exceeds it (upper bounds travel with the keys)
            for (int64_t r0 = 1; r0 < nprobe;) {
                const int64_t r1 = std::min<int64_t>(nprobe, r0 * 4);
                mfma_phase((int)r0, (int)r1, 0, INT32_MAX, 0, stats + 8);
                hv::launch_bf16_compact(surv, aux, surv_cnt, n, (int)cap, (int)k, bound, last, kq, 1, s);
                r0 = r1;
            }
            HV_CHECK_LAUNCH("compact");
        } else {
            // FP8: the estimates carry no error bound, so the filter is the kr-th smallest
            // estimate so far; the probe ranks go in groups of growing width ([1, 4), [4, 16),
            // ...), each followed by the shortlist, which keeps the kr smallest and lowers the
            // filter. The last shortlist is what the exact re-rank sees.
            for (int64_t r0 = 1; r0 < rank_end;) {
                const int64_t r1 = std::min<int64_t>(rank_end, r0 * 4);
                mfma_phase((int)r0, (int)r1, 0, INT32_MAX, 0, stats + 8);
                hv::launch_shortlist(surv, surv_cnt, n, (int)cap, (int)kr, bound, kq, (int)(kr / k), s);
                r0 = r1;
            }
            if (rank_end == 1) hv::launch_shortlist(surv, surv_cnt, n, (int)cap, (int)kr, bound, kq, (int)(kr / k), s);
            HV_CHECK_LAUNCH("shortlist");
        }
    } else {
        int r_next = 0;
        if ((h->sp.refine_bound || learned) && rank_end > 1) {
            scan_phase(0, 1, INT32_MAX, prune, 0, stats);
            hv::launch_compact_topk(surv, surv_cnt, n, (int)cap, (int)k, bound, kq, s);
            HV_CHECK_LAUNCH("compact");
            r_next = 1;
        }
        scan_phase(r_next, (int)rank_end, INT32_MAX, prune, 0, r_next == 1 ? stats + 8 : stats);
    }
    rec(4);
    // 6. top-k selection. The BF16 form re-ranks its survivors (the plain path's FP32 arithmetic)
    //    and selects in the same pass.
    int* redo_cnt = reinterpret_cast<int*>(stats + 15);  // cleared with the counters above
    unsigned int* n_items = reinterpret_cast<unsigned int*>(redo_cnt + 1);
    int* redo_list = h->s_overflow.as<int>();
    // 7. (listed in the select's last kernel) the exact path redoes the queries whose survivor
    //    buffer overflowed (rare) and, in learned mode, those whose k-th result is missing or above
    //    the seeded bound (a neighbour may have been pruned), over the same probe lists, with no
    //    bound in learned mode. The list is built on the device, which also publishes the counters
    //    to the host; the host launches the redo only when the list is not empty (finish_search).
    // (the check is against the bound the seed set, bound_ref: later compactions only tighten
    //  the working bound towards the k-th distance itself)
    // (a learned query that is redone is redone over all nprobe_cap lists: its result is then the
    //  plain search's at nprobe_cap, and its reported probe count says so)
    const hv::RedoArgs redo_args{surv_cnt, (int)cap, D, I, (int)k, bound_ref, learned && !fp8 ? 1 : 0, n, redo_list,
                                 redo_cnt, stats + 12, learned ? h->s_qn.as<int>() : nullptr, (int)nprobe_cap,
                                 learned && !fp8 ? bound : nullptr, stats,
                                 reinterpret_cast<unsigned int*>(stats + 16),
                                 static_cast<unsigned long long*>(h->post_host.p)};
    if (mfma) {
        HV_HIP(h->s_ritems.ensure(sizeof(int2) * n * ((cap + 63) / 64)));
        HV_HIP(h->s_kept.ensure(sizeof(unsigned long long) * n * cap));
        HV_HIP(h->s_done.ensure(sizeof(unsigned int) * n));
        // queries in nearest-list order (their rank-0 pairs), so the re-rank reads shared rows once
        hv::launch_rerank_items(surv_cnt, n, (int)cap, h->s_pairs.as<int>(), h->s_pair_slot.as<int>(),
                                h->s_bucket_off.as<int>() + nb, n * rank_end, (int)nprobe, h->s_ritems.as<int2>(),
                                n_items, nullptr, nullptr, stats, s);
        rec(8);  // events 8-9 bracket the re-rank kernel alone
        hv::launch_rerank(Qp, (int)d, h->d_rows.as<float>(), h->d_ids.as<uint32_t>(), plan, surv, n, (int)cap,
                          fp8 ? nullptr : bound, h->s_ritems.as<int2>(), n_items, h->s_kept.as<unsigned long long>(), s);
        rec(9);
        hv::launch_select_kept(h->s_kept.as<unsigned long long>(), surv_cnt, n, (int)cap, (int)k, D, I,
                               h->s_done.as<unsigned int>(), s, &redo_args);
        HV_CHECK_LAUNCH("rerank + select");
    } else {
        hv::launch_select_survivors(surv, surv_cnt, n, (int)cap, (int)k, D, I, nullptr, nullptr, stats, s, &redo_args);
        HV_CHECK_LAUNCH("select");
    }
    const int* redo_qn = learned ? nullptr : qn;
    // learned (FP32, BF16): each redone query's own k-th result (written above); FP8: no bound
    const float* redo_bound = fp8 ? nullptr : bound;
    post.batch = batch;
    post.main_recorded = main_recorded;
    post.mfma = mfma;
    post.per_query = learned || uneven;
    post.B = B;
    post.nprobe = nprobe;
    post.nprobe_cap = nprobe_cap;
    post.redo_cnt = redo_cnt;
    unsigned long long* runs = h->s_runs.as<unsigned long long>();
    post.redo = [=](int q0) {
        hv::launch_query_exact(redo_list, redo_cnt, q0, batch, Qp, (int)d, h->d_codes.as<float>(),
                               h->d_code_off.as<int64_t>(), h->d_list_off.as<int64_t>(), h->d_list_stride.as<int>(),
                               h->d_ids.as<uint32_t>(), probes, (int)nprobe, (int)nprobe_cap, redo_qn, nseg, plan,
                               prune, redo_bound, (int)k, runs, D, I, q0 == 0, s);  // the first batch's counts: cleared above
    };
    // per-query k: each row cut to its query's k (again after any redo batch the host adds)
    post.truncate = [=]() { hv::launch_truncate_rows(D, I, n, (int)k, kq, s); };
    post.truncate();
    HV_CHECK_LAUNCH("truncate");
    if (capture) {
        hipGraph_t g = nullptr;
        guard.active = false;
        HV_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HV_HIP(ie);
        h->graph.exec = ex;
        h->graph.sig = sig;
        h->graph.gen = g_alloc_gen.load();
        h->graph.post = post;
        HV_HIP(hipGraphLaunch(ex, s));
    }
    HV_HIP(hipEventRecord(h->ev[5], s));
    finish_search(h, capture ? h->graph.post : post, n, st, !capture, t0);
    if (graph_ok && !capture) {
        h->graph.pending_sig = sig;
        h->graph.pending_gen = g_alloc_gen.load();
    }
}

}  // namespace

extern "C" {

hv_index_t* hv_create(int64_t d, int64_t nlist, int32_t metric, int32_t opt_level) {
    g_create_error.clear();
    if (d <= 0 || nlist <= 0) {
        g_create_error = "d and nlist must be positive";
        return nullptr;
    }
    if (metric != HV_METRIC_L2 && metric != HV_METRIC_IP) {
        g_create_error = "metric must be HV_METRIC_L2 or HV_METRIC_IP";
        return nullptr;
    }
    if (opt_level < 0 || opt_level > HV_OPT_ALL) {
        g_create_error = "invalid opt_level";
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_error = "no GPU device visible";
        return nullptr;
    }
    hv_index_t* h = new hv_index_t();
    h->d = d;
    h->nlist = nlist;
    h->metric = metric;
    h->opt_level = opt_level;
    hv_default_search_params(&h->sp);
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        hv_destroy(h);
        return nullptr;
    }
    if (h->post_host.ensure(sizeof(unsigned long long) * 16) != hipSuccess) {
        g_create_error = "hipHostMalloc failed";
        hv_destroy(h);
        return nullptr;
    }
    {
        // the no-allocation search's scratch on this device (current here); a failure only
        // disables that entry point, which reports it
        hipError_t e = hipSuccess;
        (void)scratch_setup(h->device, &e);
        (void)hipGetLastError();
    }
    for (auto& e : h->ev_scan) {
        if (hipEventCreate(&e) != hipSuccess) {
            g_create_error = "hipEventCreate failed";
            hv_destroy(h);
            return nullptr;
        }
    }
    for (auto& e : h->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            g_create_error = "hipEventCreate failed";
            hv_destroy(h);
            return nullptr;
        }
    }
    return h;
}

void hv_destroy(hv_index_t* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->graph.reset();
    h->qn_host.release();
    h->post_host.release();
    h->qk_host.release();
    DevBuf* bufs[] = {&h->d_perm, &h->d_centroids, &h->d_codes, &h->d_list_off, &h->d_code_off, &h->d_list_stride,
                      &h->d_ids, &h->d_rank_order, &h->s_q, &h->s_qp, &h->s_D, &h->s_I, &h->s_dist, &h->s_probes,
                      &h->s_bound, &h->s_bound_ref, &h->s_bucket, &h->s_fill, &h->s_bucket_off, &h->s_scan_tmp, &h->s_pairs,
                      &h->s_items, &h->s_queue, &h->s_surv, &h->s_surv_cnt, &h->s_overflow, &h->s_ritems,
                      &h->s_kept, &h->s_done, &h->s_stats, &h->s_pdist, &h->s_pair_slot, &h->s_aux, &h->s_qres,
                      &h->s_runs, &h->s_qn, &h->s_feat, &h->s_pred, &h->s_seed, &h->s_fb, &h->s_qnorm, &h->s_near,
                      &h->s_cnorm, &h->s_redo, &h->d_rcodes, &h->d_roff, &h->d_rnorm, &h->d_rerr, &h->d_rows, &h->s_qscale, &h->s_last,
                      &h->d_rcodes8, &h->d_lscale8, &h->d_lrad, &h->s_qk, &h->d_c2c};
    for (DevBuf* b : bufs) b->release();
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : h->ev_scan)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* hv_last_error(const hv_index_t* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int hv_set_dim_order(hv_index_t* h, const int32_t* perm) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(!h->trained, HV_ERR_STATE, "dim order must be set before train");
        if (!perm) {
            h->perm.clear();
            return;
        }
        std::vector<int> p(perm, perm + h->d), seen(h->d, 0);
        for (int v : p) {
            require(v >= 0 && v < h->d && !seen[v], HV_ERR_INVALID_ARG, "not a permutation of 0..d-1");
            seen[v] = 1;
        }
        h->perm = p;
        upload(h->d_perm, p.data(), sizeof(int) * h->d, h->stream);
        HV_HIP(hipStreamSynchronize(h->stream));
    });
}

void hv_default_kmeans_params(hv_kmeans_params* p) {
    if (!p) return;
    p->niter = 10;
    p->seed = 0;
    p->max_points_per_centroid = 256;
}

int hv_set_centroids(hv_index_t* h, const float* c) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(c != nullptr, HV_ERR_INVALID_ARG, "null centroids");
        h->centroids.assign(h->nlist * h->d, 0.f);
        for (int64_t l = 0; l < h->nlist; ++l)
            for (int64_t t = 0; t < h->d; ++t)
                h->centroids[l * h->d + t] = c[l * h->d + (h->perm.empty() ? t : h->perm[t])];
        upload(h->d_centroids, h->centroids.data(), sizeof(float) * h->nlist * h->d, h->stream);
        HV_HIP(hipStreamSynchronize(h->stream));
        h->trained = true;
        h->added = false;
    });
}

int hv_train(hv_index_t* h, int64_t n, const float* x, const hv_kmeans_params* kp) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(x != nullptr, HV_ERR_INVALID_ARG, "null x");
        require(n >= h->nlist, HV_ERR_INVALID_ARG, "need at least nlist training vectors");
        hv_kmeans_params p;
        hv_default_kmeans_params(&p);
        if (kp) p = *kp;
        require(p.niter >= 0 && p.max_points_per_centroid >= 0, HV_ERR_INVALID_ARG, "bad k-means params");
        const int64_t d = h->d, nlist = h->nlist;
        std::vector<int64_t> rows, init;
        kmeans_pick(n, nlist, p.max_points_per_centroid, p.seed, rows, init);
        const int64_t ns = (int64_t)rows.size();
        require(ns < (1ll << 31), HV_ERR_CAPACITY, "training sample too large");
        std::vector<float> sample(ns * d);
        for (int64_t i = 0; i < ns; ++i) std::memcpy(&sample[i * d], x + rows[i] * d, sizeof(float) * d);
        DevBuf raw, perm_buf, labels, off_buf, order_buf, xn_buf;
        const float* xp = upload_rows(h, sample.data(), ns, raw, perm_buf);
        // IP: rows are assigned by <normalized row, centroid> (Clustering.cpp:122-130), averaged raw
        const float* xa = xp;
        if (h->metric == HV_METRIC_IP) {
            HV_HIP(xn_buf.ensure(sizeof(float) * ns * d));
            hv::launch_normalize_rows(xp, ns, (int)d, xn_buf.as<float>(), h->stream);
            xa = xn_buf.as<float>();
        }
        h->centroids.assign(nlist * d, 0.f);
        for (int64_t c = 0; c < nlist; ++c)
            for (int64_t t = 0; t < d; ++t)
                h->centroids[c * d + t] = sample[init[c] * d + (h->perm.empty() ? t : h->perm[t])];
        upload(h->d_centroids, h->centroids.data(), sizeof(float) * nlist * d, h->stream);
        HV_HIP(labels.ensure(sizeof(int) * ns));
        HV_HIP(off_buf.ensure(sizeof(int64_t) * (nlist + 1)));
        HV_HIP(order_buf.ensure(sizeof(int) * ns));
        std::vector<int> lab(ns), order(ns);
        std::vector<int64_t> off(nlist + 1);
        for (int64_t it = 0; it < p.niter; ++it) {
            assign_rows(h, xa, ns, labels.as<int>(), nullptr);
            HV_HIP(hipMemcpyAsync(lab.data(), labels.p, sizeof(int) * ns, hipMemcpyDeviceToHost, h->stream));
            HV_HIP(hipStreamSynchronize(h->stream));
            // stable counting sort of the sample by cluster (ascending point order inside)
            std::fill(off.begin(), off.end(), 0);
            for (int64_t i = 0; i < ns; ++i) off[lab[i] + 1]++;
            for (int64_t c = 0; c < nlist; ++c) off[c + 1] += off[c];
            std::vector<int64_t> cur(off.begin(), off.end() - 1);
            for (int64_t i = 0; i < ns; ++i) order[cur[lab[i]]++] = (int)i;
            HV_HIP(hipMemcpyAsync(off_buf.p, off.data(), sizeof(int64_t) * (nlist + 1), hipMemcpyHostToDevice,
                                  h->stream));
            HV_HIP(hipMemcpyAsync(order_buf.p, order.data(), sizeof(int) * ns, hipMemcpyHostToDevice, h->stream));
            hv::launch_centroid_update_f64(xp, (int)d, order_buf.as<int>(), off_buf.as<int64_t>(), (int)nlist,
                                           h->d_centroids.as<float>(), h->stream);
            HV_CHECK_LAUNCH("centroid update");
        }
        if (h->metric == HV_METRIC_IP) {  // unit centroids (Index.cpp:408-418)
            hv::launch_normalize_centroids(h->d_centroids.as<float>(), (int)nlist, (int)d, h->stream);
            HV_CHECK_LAUNCH("centroid normalization");
        }
        HV_HIP(hipMemcpyAsync(h->centroids.data(), h->d_centroids.p, sizeof(float) * nlist * d,
                              hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipStreamSynchronize(h->stream));
        raw.release();
        perm_buf.release();
        labels.release();
        off_buf.release();
        order_buf.release();
        xn_buf.release();
        h->trained = true;
        h->added = false;
    });
}

// Index::add (Index.cpp:581-696): every vector goes to its nearest centroid (ties to the
// lower id); inside a list, vectors are ordered by ascending distance to the centroid, ties
// by id. Like the reference (IVF::reset, Index.cpp:636-639), add rebuilds the lists from the
// given vectors, whose ids are their positions 0..n-1.
int hv_add(hv_index_t* h, int64_t n, const float* x) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->trained, HV_ERR_STATE, "add before train");
        require(x != nullptr && n > 0, HV_ERR_INVALID_ARG, "empty add");
        require(n < 0xffffffffll, HV_ERR_CAPACITY, "ntotal must be < 2^32 - 1");
        const int64_t d = h->d, nlist = h->nlist;
        DevBuf raw, perm_buf, labels, best, order_buf;
        const float* xp = upload_rows(h, x, n, raw, perm_buf);
        HV_HIP(labels.ensure(sizeof(int) * n));
        HV_HIP(best.ensure(sizeof(float) * n));
        assign_rows(h, xp, n, labels.as<int>(), best.as<float>());
        std::vector<int> lab(n);
        std::vector<float> dist(n);
        HV_HIP(hipMemcpyAsync(lab.data(), labels.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipMemcpyAsync(dist.data(), best.p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipStreamSynchronize(h->stream));
        labels.release();
        best.release();
        std::vector<int64_t> off(nlist + 1, 0);
        for (int64_t i = 0; i < n; ++i) off[lab[i] + 1]++;
        for (int64_t l = 0; l < nlist; ++l) off[l + 1] += off[l];
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        std::vector<int> order(n);
        for (int64_t i = 0; i < n; ++i) order[cur[lab[i]]++] = (int)i;
        const bool ip = h->metric == HV_METRIC_IP;  // IP: descending inner product (Index.cpp:643-651)
        for (int64_t l = 0; l < nlist; ++l)
            std::sort(order.begin() + off[l], order.begin() + off[l + 1], [&](int a, int b) {
                if (dist[a] != dist[b]) return ip ? dist[a] > dist[b] : dist[a] < dist[b];
                return a < b;
            });
        std::vector<float> c2c;
        if (ip) {
            c2c.resize(n);
            for (int64_t t = 0; t < n; ++t) c2c[t] = dist[order[t]];
        }
        install_lists(h, xp, order, off, std::vector<int64_t>(order.begin(), order.end()), ip ? &c2c : nullptr);
        raw.release();
        perm_buf.release();
        h->added_opt_level = h->opt_level;
    });
}

extern "C++" {
namespace {

template <typename T>
void read_exact(FILE* f, T* dst, size_t count, const char* what) {
    if (count && std::fread(dst, sizeof(T), count, f) != count)
        throw HvError{HV_ERR_IO, std::string("truncated index file while reading ") + what};
}

template <typename T>
void write_exact(FILE* f, const T* src, size_t count) {
    if (count && std::fwrite(src, sizeof(T), count, f) != count) throw HvError{HV_ERR_IO, "short write"};
}

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

}  // namespace
}  // extern "C++"

// Index::save_index (Index.cpp:1708-1732) + IVF::save_IVF (IVF.cpp:76-99).
int hv_save(const hv_index_t* hc, const char* path) {
    if (!hc) return HV_ERR_INVALID_ARG;
    hv_index_t* h = const_cast<hv_index_t*>(hc);
    return guarded(h, [&] {
        require(path != nullptr, HV_ERR_INVALID_ARG, "null path");
        require(h->trained && h->added, HV_ERR_STATE, "save before train/add");
        const int64_t d = h->d, nlist = h->nlist;
        std::vector<float> codes(h->ntotal * d), cent(nlist * d);
        if (h->ntotal) {
            int rc = hv_get_list_codes(h, codes.data());
            if (rc != HV_OK) throw HvError{rc, h->err};
        }
        hv_get_centroids(h, cent.data());
        FILE* f = std::fopen(path, "wb");
        require(f != nullptr, HV_ERR_IO, std::string("Cannot open file ") + path);
        FileCloser fc{f};
        const uint64_t hd[2] = {(uint64_t)d, (uint64_t)nlist};
        const int32_t en[2] = {h->metric, h->added_opt_level};
        const uint64_t sub[3] = {0, 1, 1};
        write_exact(f, hd, 2);
        write_exact(f, en, 2);
        write_exact(f, sub, 3);
        write_exact(f, cent.data(), cent.size());
        std::vector<uint64_t> ids;
        std::vector<float> norms;
        const bool ip = h->metric == HV_METRIC_IP;
        for (int64_t l = 0; l < nlist; ++l) {
            const int64_t b = h->list_off[l], n = h->list_off[l + 1] - b;
            if (n == 0) continue;
            const uint64_t rec[4] = {(uint64_t)l, (uint64_t)n, (uint64_t)d, 0};
            // IP lists carry candidate2centroid (the TRIANGLE arrays, IVF.cpp:87-90)
            const int32_t opt = ip ? HV_OPT_TRIANGLE : 0;
            write_exact(f, rec, 4);
            write_exact(f, &opt, 1);
            ids.assign(h->ids.begin() + b, h->ids.begin() + b + n);
            write_exact(f, ids.data(), ids.size());
            write_exact(f, codes.data() + b * d, (size_t)(n * d));
            norms.assign(n, 0.f);  // the reference fills candidate_norms for L2 only (Index.cpp:689-696)
            for (int64_t t = 0; t < n && !ip; ++t) {
                float s = 0.f;
                for (int64_t i = 0; i < d; ++i) s = std::fmaf(codes[(b + t) * d + i], codes[(b + t) * d + i], s);
                norms[t] = s;
            }
            write_exact(f, norms.data(), norms.size());
            if (ip) {
                write_exact(f, h->c2c.data() + b, (size_t)n);
                const std::vector<float> zeros(n, 0.f);  // sqrt_candidate2centroid: L2 only (Index.cpp:678-686)
                write_exact(f, zeros.data(), zeros.size());
            }
        }
    });
}

// Index::load_index (Index.cpp:1733-1770) + IVF::load_IVF (IVF.cpp:101-134). The whole file is
// parsed and checked (sizes against the bytes left in the file, optional arrays included) before
// the handle is touched, so a bad or truncated file leaves the handle as it was. Like the
// reference, the search opt level restarts at OPT_NONE (Index.cpp:1743) and added_opt_level is the
// file's.
int hv_load(hv_index_t* h, const char* path) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(path != nullptr, HV_ERR_INVALID_ARG, "null path");
        FILE* f = std::fopen(path, "rb");
        require(f != nullptr, HV_ERR_IO, std::string("Cannot open file ") + path);
        FileCloser fc{f};
        require(std::fseek(f, 0, SEEK_END) == 0, HV_ERR_IO, "cannot size the index file");
        const long fsize = std::ftell(f);
        require(fsize >= 0 && std::fseek(f, 0, SEEK_SET) == 0, HV_ERR_IO, "cannot size the index file");
        // bytes left after the current position must cover `need`
        auto need = [&](uint64_t bytes, const char* what) {
            const long pos = std::ftell(f);
            require(pos >= 0 && (uint64_t)(fsize - pos) >= bytes, HV_ERR_IO,
                    std::string("truncated index file while reading ") + what);
        };
        uint64_t hd[2];
        int32_t en[2];
        uint64_t sub[3];
        read_exact(f, hd, 2, "header");
        read_exact(f, en, 2, "header");
        read_exact(f, sub, 3, "header");
        const int64_t d = (int64_t)hd[0], nlist = (int64_t)hd[1];
        require(d > 0 && d <= 65536 && nlist > 0 && nlist <= (1 << 24), HV_ERR_IO, "implausible d / nlist in header");
        require(en[0] == HV_METRIC_L2 || en[0] == HV_METRIC_IP, HV_ERR_IO, "unknown metric in header");
        const bool ip = en[0] == HV_METRIC_IP;
        require(en[1] >= 0 && en[1] <= HV_OPT_ALL, HV_ERR_IO, "unknown opt level in header");
        need(sizeof(float) * (uint64_t)nlist * d, "centroids");
        std::vector<float> cent(nlist * d);
        read_exact(f, cent.data(), cent.size(), "centroids");
        std::vector<std::vector<uint64_t>> lids(nlist);
        std::vector<std::vector<float>> lcodes(nlist), lc2c(nlist);
        bool all_c2c = true;  // every stored list carries candidate2centroid
        std::vector<char> seen(nlist, 0);
        int64_t total = 0;
        for (;;) {
            uint64_t lid;
            if (std::fread(&lid, sizeof(lid), 1, f) != 1) {
                require(std::feof(f) && std::ftell(f) == fsize, HV_ERR_IO, "truncated list record");
                break;
            }
            uint64_t rec[3];
            int32_t opt;
            read_exact(f, rec, 3, "list header");
            read_exact(f, &opt, 1, "list header");
            const uint64_t n = rec[0], ld = rec[1], sk = rec[2];
            require(lid < (uint64_t)nlist && ld == (uint64_t)d && n < (1ull << 32), HV_ERR_IO, "bad list record");
            require(opt >= 0 && opt <= HV_OPT_ALL && sk < (1ull << 20), HV_ERR_IO, "bad list record");
            require(!seen[lid], HV_ERR_IO, "list stored twice");
            seen[lid] = 1;
            // ids, codes, norms, then the optional arrays of IVF.cpp:117-134
            uint64_t extra = 0;
            if ((opt & HV_OPT_TRIANGLE) || (opt & HV_OPT_SUBNN_IP)) extra += 2ull * 4 * n;
            if (opt & HV_OPT_SUBNN_L2) extra += n * sk * (8 + 4);
            if (opt & HV_OPT_SUBNN_IP) extra += 2ull * n * sk * (8 + 4);
            need(n * 8 + n * (uint64_t)d * 4 + n * 4 + extra, "list arrays");
            lids[lid].resize(n);
            lcodes[lid].resize(n * d);
            read_exact(f, lids[lid].data(), n, "ids");
            read_exact(f, lcodes[lid].data(), n * d, "codes");
            require(std::fseek(f, (long)(n * 4), SEEK_CUR) == 0, HV_ERR_IO, "seek failed");  // norms
            if ((opt & HV_OPT_TRIANGLE) || (opt & HV_OPT_SUBNN_IP)) {
                lc2c[lid].resize(n);
                read_exact(f, lc2c[lid].data(), n, "candidate2centroid");
                extra -= 4 * n;
            } else {
                all_c2c = false;
            }
            require(std::fseek(f, (long)extra, SEEK_CUR) == 0, HV_ERR_IO, "seek failed");
            total += (int64_t)n;
        }
        require(total < 0xffffffffll, HV_ERR_CAPACITY, "ntotal must be < 2^32 - 1");
        std::vector<float> rows(total * d), c2c;
        std::vector<int64_t> off(nlist + 1, 0), ids(total);
        if (ip) c2c.resize(total);
        for (int64_t l = 0; l < nlist; ++l) {
            off[l + 1] = off[l] + (int64_t)lids[l].size();
            std::copy(lcodes[l].begin(), lcodes[l].end(), rows.begin() + off[l] * d);
            for (size_t t = 0; ip && t < lids[l].size(); ++t) {
                if (all_c2c) {
                    c2c[off[l] + t] = lc2c[l][t];
                } else {  // <x, c> as the add computes it (FP32 chain in the file's dimension order)
                    float acc = 0.f;
                    for (int64_t i = 0; i < d; ++i) acc = std::fmaf(lcodes[l][t * d + i], cent[l * d + i], acc);
                    c2c[off[l] + t] = acc;
                }
            }
            for (size_t t = 0; t < lids[l].size(); ++t) {
                require(lids[l][t] < 0xffffffffull, HV_ERR_CAPACITY, "id >= 2^32 - 1");
                ids[off[l] + t] = (int64_t)lids[l][t];
            }
        }
        lcodes.clear();
        // the file is good: re-initialise the handle to its shape. A device failure from here on
        // leaves the handle untrained (searches refuse with HV_ERR_STATE) rather than half loaded.
        if (d != h->d) {
            h->perm.clear();
            h->d_perm.release();
        }
        h->d = d;
        h->nlist = nlist;
        h->metric = en[0];
        h->trained = h->added = false;
        h->opt_level = HV_OPT_NONE;
        h->sp.opt_level = HV_OPT_NONE;
        h->last_n = 0;
        h->graph.reset();
        try {
            int rc = hv_set_centroids(h, cent.data());
            if (rc != HV_OK) throw HvError{rc, h->err};
            DevBuf raw, perm_buf;
            const float* xp = total ? upload_rows(h, rows.data(), total, raw, perm_buf) : nullptr;
            std::vector<int> order(total);
            std::iota(order.begin(), order.end(), 0);
            install_lists(h, xp, order, off, ids, ip ? &c2c : nullptr);
            h->added_opt_level = en[1];
        } catch (...) {
            h->trained = h->added = false;
            throw;
        }
    });
}

void hv_default_search_params(hv_search_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nprobe = 1;
    p->opt_level = HV_OPT_NONE;
    p->dim_blocks = 4;
    p->pruning = 1;
    p->bound_mode = HV_BOUND_WARMUP;
    p->warmup_list = 10;
    p->warmup_list_size = 0;
    p->precision = HV_PRECISION_FP32;
    p->rerank_r = 4;
    p->refine_bound = 1;
    p->survivor_cap = 0;
    p->learned = 0;
    p->learned_nprobe_scale = 1.0f;
    p->learned_bound_scale = 1.0f;
    p->learned_min_nprobe = 1;
    p->learned_margin = 1.0f;
    p->nearest_pass = 1;
}

int hv_set_search_params(hv_index_t* h, const hv_search_params* p) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(p != nullptr, HV_ERR_INVALID_ARG, "null params");
        require(p->nprobe >= 1, HV_ERR_INVALID_ARG, "nprobe must be >= 1");
        require(p->dim_blocks >= 1 && p->dim_blocks <= h->d, HV_ERR_INVALID_ARG, "dim_blocks must be in [1, d]");
        require(p->opt_level >= 0 && p->opt_level <= HV_OPT_ALL, HV_ERR_INVALID_ARG, "invalid opt_level");
        require(p->bound_mode == HV_BOUND_WARMUP || p->bound_mode == HV_BOUND_EXTERNAL, HV_ERR_INVALID_ARG,
                "invalid bound_mode");
        require(p->warmup_list >= 0 && p->warmup_list_size >= 0, HV_ERR_INVALID_ARG, "negative warm-up size");
        require(p->precision == HV_PRECISION_FP32 || p->precision == HV_PRECISION_BF16 ||
                    p->precision == HV_PRECISION_FP8,
                HV_ERR_INVALID_ARG, "unsupported precision");
        require(p->rerank_r >= 1, HV_ERR_INVALID_ARG, "rerank_r must be >= 1");
        require(p->scan_vsplit >= 0 && p->scan_vsplit % 16 == 0 && p->scan_vsplit <= (1 << 20), HV_ERR_INVALID_ARG,
                "scan_vsplit must be 0 or a positive multiple of 16");
        require(p->survivor_cap >= 0 && p->survivor_cap <= 8192, HV_ERR_INVALID_ARG, "survivor_cap > 8192");
        require(p->tile_shape >= 0 && p->tile_shape <= hv::kNumShapes && p->tile_shape_first >= 0 &&
                    p->tile_shape_first <= hv::kNumShapes,
                HV_ERR_INVALID_ARG, "invalid tile_shape");
        require(p->learned == 0 || p->learned == 1, HV_ERR_INVALID_ARG, "learned must be 0 or 1");
        require(p->learned_nprobe_scale > 0.f && p->learned_min_nprobe >= 1, HV_ERR_INVALID_ARG,
                "learned_nprobe_scale must be > 0 and learned_min_nprobe >= 1");
        require(p->learned_margin >= 0.f, HV_ERR_INVALID_ARG, "learned_margin must be >= 0");
        require(p->nearest_pass == 0 || p->nearest_pass == 1, HV_ERR_INVALID_ARG, "nearest_pass must be 0 or 1");
        h->sp = *p;
    });
}

int hv_get_search_params(const hv_index_t* h, hv_search_params* p) {
    if (!h || !p) return HV_ERR_INVALID_ARG;
    *p = h->sp;
    return HV_OK;
}

int hv_set_external_bounds(hv_index_t* h, int64_t n, const float* bounds) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(n >= 0 && (bounds || n == 0), HV_ERR_INVALID_ARG, "bad bounds");
        h->ext_bounds.assign(bounds, bounds + n);
    });
}

int hv_set_query_nprobe(hv_index_t* h, int64_t n, const int32_t* nprobe) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(n >= 0 && (nprobe || n == 0), HV_ERR_INVALID_ARG, "bad per-query nprobe");
        h->query_nprobe.assign(nprobe, nprobe + n);
        h->qn_stage.assign(n, 0);
    });
}

int hv_set_query_k(hv_index_t* h, int64_t n, const int32_t* k) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(n >= 0 && (k || n == 0), HV_ERR_INVALID_ARG, "bad per-query k");
        for (int64_t i = 0; i < n; ++i) require(k[i] >= 1, HV_ERR_INVALID_ARG, "per-query k must be >= 1");
        h->query_k.assign(k, k + n);
        h->qk_stage.assign(n, 0);
    });
}

int hv_search_device(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D_dev, int64_t* I_dev,
                     hv_stats* st) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        if (st) std::memset(st, 0, sizeof(*st));
        do_search(h, n, q_dev, k, D_dev, I_dev, st);
    });
}

int hv_search_device_noalloc(hv_index_t* h, int64_t n, const float* q_dev, int64_t k, float* D_dev, int64_t* I_dev,
                             hv_stats* st) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        if (st) std::memset(st, 0, sizeof(*st));
        hipError_t e = hipSuccess;
        void* base = scratch_setup(h->device, &e);
        require(base != nullptr, HV_ERR_DEVICE,
                std::string("no-allocation scratch unavailable on this device: ") + hipGetErrorString(e));
        require(h->sp.precision != HV_PRECISION_FP8 || h->fp8_built, HV_ERR_STATE,
                "the FP8 codes are built by the first FP8 search (hv_search / hv_search_device); a no-allocation "
                "search cannot build them");
        // the search's scratch buffers, swapped for arena-backed ones for the length of the call
        DevBuf* bufs[] = {&h->s_qp, &h->s_dist, &h->s_redo, &h->s_qnorm, &h->s_cnorm, &h->s_bucket, &h->s_fill,
                          &h->s_stats, &h->s_surv_cnt, &h->s_bound, &h->s_bound_ref, &h->s_overflow, &h->s_last,
                          &h->s_runs, &h->s_probes, &h->s_pdist, &h->s_qn, &h->s_seed, &h->s_pred, &h->s_bucket_off,
                          &h->s_scan_tmp, &h->s_pairs, &h->s_pair_slot, &h->s_items, &h->s_queue, &h->s_surv,
                          &h->s_aux, &h->s_qres, &h->s_qscale, &h->s_ritems, &h->s_kept, &h->s_done, &h->s_qk,
                          &h->s_near};
        constexpr int nbufs = (int)(sizeof(bufs) / sizeof(bufs[0]));
        std::lock_guard<std::mutex> lk(g_scratch_mu[h->device]);
        struct Swap {
            hv_index_t* h;
            DevBuf** b;
            void* p[nbufs];
            size_t cap[nbufs];
            Swap(hv_index_t* h_, DevBuf** b_) : h(h_), b(b_) {
                for (int i = 0; i < nbufs; ++i) {
                    p[i] = b[i]->p;
                    cap[i] = b[i]->cap;
                }
                h->noalloc = true;
            }
            void attach(Arena* a) {
                for (int i = 0; i < nbufs; ++i) {
                    b[i]->p = nullptr;
                    b[i]->cap = 0;
                    b[i]->arena = a;
                }
            }
            ~Swap() {
                // the scratch is free for the next call only once this call's work is done
                (void)hipStreamSynchronize(h->stream);
                for (int i = 0; i < nbufs; ++i) {
                    b[i]->p = p[i];
                    b[i]->cap = cap[i];
                    b[i]->arena = nullptr;
                }
                h->noalloc = false;
                h->size_only = false;
                h->last_n = 0;  // the last search's buffers are not kept
            }
        } swap(h, bufs);
        // pass 1 sizes every buffer (no launches) against an unbounded counter; pass 2 runs the
        // search in the real block
        Arena count{static_cast<char*>(base), (size_t)1 << 62, 0, nullptr, 0};
        swap.attach(&count);
        h->size_only = true;
        do_search(h, n, q_dev, k, D_dev, I_dev, nullptr);
        h->size_only = false;
        require(count.high <= kNoAllocScratchBytes, HV_ERR_CAPACITY,
                "no-allocation search: this batch needs " + std::to_string((count.high + (1 << 20) - 1) >> 20) +
                    " MiB of scratch, the limit is " + std::to_string(kNoAllocScratchBytes >> 20) +
                    " MiB (search fewer queries per call, or lower survivor_cap)");
        Arena arena{static_cast<char*>(base), kNoAllocScratchBytes, 0, nullptr, 0};
        swap.attach(&arena);
        do_search(h, n, q_dev, k, D_dev, I_dev, st);
    });
}

int hv_search(hv_index_t* h, int64_t n, const float* q, int64_t k, float* D, int64_t* I, hv_stats* st) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        if (st) std::memset(st, 0, sizeof(*st));
        require(n >= 0, HV_ERR_INVALID_ARG, "n < 0");
        require(k >= 1 && k <= hv::kMaxK, HV_ERR_CAPACITY, "k must be in [1, 1024]");
        if (n == 0) return;
        require(q && D && I, HV_ERR_INVALID_ARG, "null buffer");
        auto t0 = std::chrono::steady_clock::now();
        upload(h->s_q, q, sizeof(float) * n * h->d, h->stream);
        HV_HIP(h->s_D.ensure(sizeof(float) * n * k));
        HV_HIP(h->s_I.ensure(sizeof(int64_t) * n * k));
        HV_HIP(hipStreamSynchronize(h->stream));
        auto t1 = std::chrono::steady_clock::now();
        do_search(h, n, h->s_q.as<float>(), k, h->s_D.as<float>(), h->s_I.as<int64_t>(), st);
        auto t2 = std::chrono::steady_clock::now();
        HV_HIP(hipMemcpyAsync(D, h->s_D.p, sizeof(float) * n * k, hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipMemcpyAsync(I, h->s_I.p, sizeof(int64_t) * n * k, hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipStreamSynchronize(h->stream));
        auto t3 = std::chrono::steady_clock::now();
        if (st) {
            st->ms_upload = std::chrono::duration<double, std::milli>(t1 - t0).count();
            st->ms_download = std::chrono::duration<double, std::milli>(t3 - t2).count();
            st->ms_total = std::chrono::duration<double, std::milli>(t3 - t0).count();
        }
    });
}

int hv_ground_truth_f64(hv_index_t* h, int64_t n, const float* q, int64_t k, double* D, int64_t* I) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->metric == HV_METRIC_L2, HV_ERR_INVALID_ARG, "the FP64 ground truth is L2 only");
        require(n >= 0, HV_ERR_INVALID_ARG, "n < 0");
        if (n == 0) return;
        require(q && D && I, HV_ERR_INVALID_ARG, "null buffer");
        require(k >= 1 && k <= hv::kMaxK, HV_ERR_CAPACITY, "k must be in [1, 1024]");
        const int64_t d = h->d;
        const int64_t cap = 4096;
        // 1. the current search gives k real candidates per query; their FP32 distances
        //    bound the k-th exact distance. The FP32 chain of d+2 rounded non-negative terms
        //    is within a relative (d+2)*2^-24 of the exact value, so 4x that margin is safe.
        upload(h->s_q, q, sizeof(float) * n * d, h->stream);
        HV_HIP(h->s_D.ensure(sizeof(float) * n * k));
        HV_HIP(h->s_I.ensure(sizeof(int64_t) * n * k));
        do_search(h, n, h->s_q.as<float>(), k, h->s_D.as<float>(), h->s_I.as<int64_t>(), nullptr);
        std::vector<float> d32(n * k);
        std::vector<int64_t> i32(n * k);
        HV_HIP(hipMemcpy(d32.data(), h->s_D.p, sizeof(float) * n * k, hipMemcpyDeviceToHost));
        HV_HIP(hipMemcpy(i32.data(), h->s_I.p, sizeof(int64_t) * n * k, hipMemcpyDeviceToHost));
        const double margin = 1.0 + 4.0 * (double)(d + 2) * std::ldexp(1.0, -24);
        std::vector<double> tau(n);
        for (int64_t i = 0; i < n; ++i)
            tau[i] = i32[i * k + k - 1] >= 0 ? (double)d32[i * k + k - 1] * margin : INFINITY;
        // 2. FP64 scan of every list, keeping distances <= tau
        const float* Qp = h->s_q.as<float>();
        if (!h->perm.empty()) Qp = h->s_qp.as<float>();  // permuted by the search above
        DevBuf b_tau, b_sd, b_sid, b_cnt;
        upload(b_tau, tau.data(), sizeof(double) * n, h->stream);
        HV_HIP(b_sd.ensure(sizeof(double) * n * cap));
        HV_HIP(b_sid.ensure(sizeof(uint32_t) * n * cap));
        HV_HIP(b_cnt.ensure(sizeof(unsigned int) * n));
        hv::launch_fill_i32(b_cnt.as<int>(), n, 0, h->stream);
        hv::launch_gt_f64(Qp, n, (int)d, h->d_codes.as<float>(), h->d_code_off.as<int64_t>(),
                          h->d_list_off.as<int64_t>(), h->d_list_stride.as<int>(), h->d_ids.as<uint32_t>(),
                          (int)h->nlist, b_tau.as<double>(), (int)cap, b_sd.as<double>(), b_sid.as<uint32_t>(),
                          b_cnt.as<unsigned int>(), h->stream);
        HV_CHECK_LAUNCH("fp64 ground truth");
        std::vector<unsigned int> cnt(n);
        HV_HIP(hipMemcpyAsync(cnt.data(), b_cnt.p, sizeof(unsigned int) * n, hipMemcpyDeviceToHost, h->stream));
        HV_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < n; ++i)
            require(cnt[i] <= cap, HV_ERR_CAPACITY, "ground truth: more than 4096 candidates under the bound");
        std::vector<double> sd(n * cap);
        std::vector<uint32_t> sid(n * cap);
        HV_HIP(hipMemcpy(sd.data(), b_sd.p, sizeof(double) * n * cap, hipMemcpyDeviceToHost));
        HV_HIP(hipMemcpy(sid.data(), b_sid.p, sizeof(uint32_t) * n * cap, hipMemcpyDeviceToHost));
        // 3. sort survivors by (distance, id)
        std::vector<std::pair<double, int64_t>> row;
        for (int64_t i = 0; i < n; ++i) {
            row.clear();
            for (unsigned int j = 0; j < cnt[i]; ++j) row.emplace_back(sd[i * cap + j], (int64_t)sid[i * cap + j]);
            const int64_t m = std::min<int64_t>(k, (int64_t)row.size());
            std::partial_sort(row.begin(), row.begin() + m, row.end());
            for (int64_t j = 0; j < k; ++j) {
                D[i * k + j] = j < m ? row[j].first : INFINITY;
                I[i * k + j] = j < m ? row[j].second : -1;
            }
        }
    });
}

int64_t hv_ntotal(const hv_index_t* h) { return h ? h->ntotal : -1; }
int64_t hv_dim(const hv_index_t* h) { return h ? h->d : -1; }
int64_t hv_nlist(const hv_index_t* h) { return h ? h->nlist : -1; }

int hv_get_centroids(const hv_index_t* h, float* out) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (!h->trained) return HV_ERR_STATE;
    for (int64_t l = 0; l < h->nlist; ++l)
        for (int64_t t = 0; t < h->d; ++t)
            out[l * h->d + (h->perm.empty() ? t : h->perm[t])] = h->centroids[l * h->d + t];
    return HV_OK;
}

int hv_get_list_offsets(const hv_index_t* h, int64_t* out) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (!h->added) return HV_ERR_STATE;
    std::memcpy(out, h->list_off.data(), sizeof(int64_t) * (h->nlist + 1));
    return HV_OK;
}

int hv_get_list_ids(const hv_index_t* h, int64_t* out) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (!h->added) return HV_ERR_STATE;
    std::memcpy(out, h->ids.data(), sizeof(int64_t) * h->ntotal);
    return HV_OK;
}

int hv_get_list_c2c(const hv_index_t* h, float* out) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (!h->added) return HV_ERR_STATE;
    if (h->metric != HV_METRIC_IP) return HV_ERR_INVALID_ARG;
    std::memcpy(out, h->c2c.data(), sizeof(float) * h->ntotal);
    return HV_OK;
}

int hv_get_list_codes(const hv_index_t* h, float* out) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (!h->added) return HV_ERR_STATE;
    hv_index_t* hm = const_cast<hv_index_t*>(h);
    return guarded(hm, [&] {
        std::vector<float> slab(h->code_off[h->nlist]);
        HV_HIP(hipMemcpy(slab.data(), h->d_codes.p, sizeof(float) * slab.size(), hipMemcpyDeviceToHost));
        for (int64_t l = 0; l < h->nlist; ++l) {
            const int64_t n = h->list_off[l + 1] - h->list_off[l];
            const int64_t ps = h->list_stride[l];
            for (int64_t t = 0; t < n; ++t)
                for (int64_t i = 0; i < h->d; ++i)
                    out[(h->list_off[l] + t) * h->d + (h->perm.empty() ? i : h->perm[i])] =
                        slab[h->code_off[l] + i * ps + t];
        }
    });
}

int hv_get_last_probes(const hv_index_t* h, int64_t* out, int64_t capacity) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (h->last_n == 0) return HV_ERR_STATE;
    if (capacity < h->last_n * h->last_stride) return HV_ERR_INVALID_ARG;
    std::vector<int> p(h->last_n * h->last_stride);
    if (hipMemcpy(p.data(), h->s_probes.p, sizeof(int) * p.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return HV_ERR_DEVICE;
    for (size_t i = 0; i < p.size(); ++i) out[i] = p[i];
    return HV_OK;
}

int hv_get_last_bounds(const hv_index_t* h, float* out, int64_t capacity) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (h->last_n == 0) return HV_ERR_STATE;
    if (capacity < h->last_n) return HV_ERR_INVALID_ARG;
    if (hipMemcpy(out, h->s_bound_ref.p, sizeof(float) * h->last_n, hipMemcpyDeviceToHost) != hipSuccess)
        return HV_ERR_DEVICE;
    return HV_OK;
}

int hv_get_last_nprobe(const hv_index_t* h, int32_t* out, int64_t capacity) {
    if (!h || !out) return HV_ERR_INVALID_ARG;
    if (h->last_n == 0) return HV_ERR_STATE;
    if (capacity < h->last_n) return HV_ERR_INVALID_ARG;
    if (h->last_learned) {
        if (hipMemcpy(out, h->s_qn.p, sizeof(int32_t) * h->last_n, hipMemcpyDeviceToHost) != hipSuccess)
            return HV_ERR_DEVICE;
    } else {
        for (int64_t i = 0; i < h->last_n; ++i) out[i] = (int32_t)h->last_nprobe;
    }
    return HV_OK;
}

int64_t hv_last_n(const hv_index_t* h) { return h ? h->last_n : 0; }

int64_t hv_last_probe_stride(const hv_index_t* h) { return h ? h->last_stride : 0; }

// ---- learned probe/bound predictor
int hv_query_features(hv_index_t* h, int64_t n, const float* q, float* feat) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->trained, HV_ERR_STATE, "features need the centroids (train first)");
        require(n >= 0 && (n == 0 || (q && feat)), HV_ERR_INVALID_ARG, "bad buffers");
        if (n == 0) return;
        const int64_t d = h->d;
        hipStream_t s = h->stream;
        upload(h->s_q, q, sizeof(float) * n * d, s);
        const float* Qp = h->s_q.as<float>();
        if (!h->perm.empty()) {
            HV_HIP(h->s_qp.ensure(sizeof(float) * n * d));
            hv::launch_permute_rows(Qp, n, (int)d, h->d_perm.as<int>(), h->s_qp.as<float>(), s);
            Qp = h->s_qp.as<float>();
        }
        // the search's coarse step (one probe) with the same feature rows
        HV_HIP(h->s_probes.ensure(sizeof(int) * n));
        HV_HIP(h->s_pdist.ensure(sizeof(float) * n));
        HV_HIP(h->s_near.ensure(sizeof(float) * n * hv::kNearest));
        coarse_select(h, n, Qp, 1, h->s_probes.as<int>(), h->s_pdist.as<float>(), false, h->s_near.as<float>());
        HV_HIP(h->s_feat.ensure(sizeof(float) * n * hv::kMlpIn));
        hv::launch_features(h->s_near.as<float>(), hv::kNearest, Qp, (int)d, n, h->s_feat.as<float>(), s);
        HV_CHECK_LAUNCH("features");
        HV_HIP(hipMemcpyAsync(feat, h->s_feat.p, sizeof(float) * n * hv::kMlpIn, hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
        h->last_n = 0;  // the probe buffers no longer describe a search
    });
}

namespace {
constexpr int kNormLen = 2 * hv::kMlpIn + 2 * hv::kMlpOut;

void mlp_refresh_serving(hv_index_t* h) {
    HV_HIP(h->m_w1b.ensure(sizeof(__bf16) * hv::kMlpW1Serve));
    HV_HIP(h->m_w2b.ensure(sizeof(__bf16) * hv::kMlpHidden * hv::kMlpOut));
    hv::launch_mlp_to_bf16(h->m_params.as<float>(), h->m_w1b.as<__bf16>(), h->m_w2b.as<__bf16>(), h->stream);
    HV_CHECK_LAUNCH("mlp weights");
    HV_HIP(hipStreamSynchronize(h->stream));
}

void mlp_upload_params(hv_index_t* h, const float* params) {
    upload(h->m_params, params, sizeof(float) * hv::kMlpParams, h->stream);
    HV_HIP(h->m_adam_m.ensure(sizeof(float) * hv::kMlpParams));
    HV_HIP(h->m_adam_v.ensure(sizeof(float) * hv::kMlpParams));
    hv::launch_fill_i32(h->m_adam_m.as<int>(), hv::kMlpParams, 0, h->stream);  // Adam moments restart
    hv::launch_fill_i32(h->m_adam_v.as<int>(), hv::kMlpParams, 0, h->stream);
    h->mlp_params = true;
    mlp_refresh_serving(h);
}
}  // namespace

int hv_mlp_set_norm(hv_index_t* h, const float* norm) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(norm != nullptr, HV_ERR_INVALID_ARG, "null norm");
        for (int i = 0; i < hv::kMlpIn; ++i)
            require(norm[hv::kMlpIn + i] > 0.f, HV_ERR_INVALID_ARG, "feature std must be > 0");
        for (int o = 0; o < hv::kMlpOut; ++o)
            require(norm[2 * hv::kMlpIn + hv::kMlpOut + o] > 0.f, HV_ERR_INVALID_ARG, "target std must be > 0");
        upload(h->m_norm, norm, sizeof(float) * kNormLen, h->stream);
        HV_HIP(hipStreamSynchronize(h->stream));
        h->mlp_norm = true;
    });
}

int hv_mlp_get_norm(const hv_index_t* h, float* norm) {
    if (!h || !norm) return HV_ERR_INVALID_ARG;
    if (!h->mlp_norm) return HV_ERR_STATE;
    return hipMemcpy(norm, h->m_norm.p, sizeof(float) * kNormLen, hipMemcpyDeviceToHost) == hipSuccess
               ? HV_OK
               : HV_ERR_DEVICE;
}

int hv_mlp_init(hv_index_t* h, uint64_t seed) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        // He-uniform: U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)) from SplitMix64; biases zero
        std::vector<float> p(hv::kMlpParams, 0.f);
        uint64_t x = seed;
        auto uni = [&]() {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        const double a1 = std::sqrt(6.0 / hv::kMlpIn), a2 = std::sqrt(6.0 / hv::kMlpHidden);
        const int w1 = hv::kMlpIn * hv::kMlpHidden, w2o = w1 + hv::kMlpHidden;
        for (int i = 0; i < w1; ++i) p[i] = (float)((2.0 * uni() - 1.0) * a1);
        for (int i = 0; i < hv::kMlpHidden * hv::kMlpOut; ++i) p[w2o + i] = (float)((2.0 * uni() - 1.0) * a2);
        mlp_upload_params(h, p.data());
    });
}

int hv_mlp_set_params(hv_index_t* h, const float* params) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(params != nullptr, HV_ERR_INVALID_ARG, "null params");
        mlp_upload_params(h, params);
    });
}

int hv_mlp_get_params(const hv_index_t* h, float* params) {
    if (!h || !params) return HV_ERR_INVALID_ARG;
    if (!h->mlp_params) return HV_ERR_STATE;
    return hipMemcpy(params, h->m_params.p, sizeof(float) * hv::kMlpParams, hipMemcpyDeviceToHost) == hipSuccess
               ? HV_OK
               : HV_ERR_DEVICE;
}

namespace {
// device training set and per-batch work buffers
struct MlpWork {
    DevBuf X, T, idx, Xn, H, Y, dY, dH, row_loss, loss;
};

void mlp_prepare(MlpWork& w, int64_t n, int B, const float* feat, const float* target, hipStream_t s) {
    upload(w.X, feat, sizeof(float) * n * hv::kMlpIn, s);
    upload(w.T, target, sizeof(float) * n * hv::kMlpOut, s);
    HV_HIP(w.idx.ensure(sizeof(int) * n));
    HV_HIP(w.Xn.ensure(sizeof(float) * B * hv::kMlpIn));
    HV_HIP(w.H.ensure(sizeof(float) * B * hv::kMlpHidden));
    HV_HIP(w.Y.ensure(sizeof(float) * B * hv::kMlpOut));
    HV_HIP(w.dY.ensure(sizeof(float) * B * hv::kMlpOut));
    HV_HIP(w.dH.ensure(sizeof(float) * B * hv::kMlpHidden));
    HV_HIP(w.row_loss.ensure(sizeof(float) * B));
    HV_HIP(w.loss.ensure(sizeof(float) * 4));
}
}  // namespace

int hv_mlp_grad(hv_index_t* h, int64_t n, const float* feat, const float* target, float* grad, float* loss) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->mlp_params && h->mlp_norm, HV_ERR_STATE, "set the parameters and the standardization first");
        require(n >= 1 && n <= (1 << 20) && feat && target && grad, HV_ERR_INVALID_ARG, "bad batch");
        hipStream_t s = h->stream;
        MlpWork w;
        mlp_prepare(w, n, (int)n, feat, target, s);
        std::vector<int> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        HV_HIP(hipMemcpyAsync(w.idx.p, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        HV_HIP(h->m_grad.ensure(sizeof(float) * hv::kMlpParams));
        hv::launch_fill_i32(w.loss.as<int>(), 1, 0, s);
        hv::launch_mlp_step(w.X.as<float>(), w.T.as<float>(), w.idx.as<int>(), (int)n, h->m_params.as<float>(),
                            h->m_norm.as<float>(), w.Xn.as<float>(), w.H.as<float>(), w.Y.as<float>(), w.dY.as<float>(),
                            w.dH.as<float>(), w.row_loss.as<float>(), h->m_grad.as<float>(), w.loss.as<float>(), s);
        HV_CHECK_LAUNCH("mlp grad");
        HV_HIP(hipMemcpyAsync(grad, h->m_grad.p, sizeof(float) * hv::kMlpParams, hipMemcpyDeviceToHost, s));
        float l = 0.f;
        HV_HIP(hipMemcpyAsync(&l, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
        if (loss) *loss = l;
    });
}

int hv_mlp_train(hv_index_t* h, int64_t n, const float* feat, const float* target, int32_t epochs, int32_t batch,
                 float lr, uint64_t seed, float* epoch_loss) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->mlp_params && h->mlp_norm, HV_ERR_STATE, "set the parameters and the standardization first");
        require(n >= 1 && feat && target, HV_ERR_INVALID_ARG, "empty training set");
        require(epochs >= 1 && batch >= 1 && batch <= 65536 && lr > 0.f, HV_ERR_INVALID_ARG, "bad schedule");
        hipStream_t s = h->stream;
        const int B = (int)std::min<int64_t>(batch, n);
        MlpWork w;
        mlp_prepare(w, n, B, feat, target, s);
        HV_HIP(h->m_grad.ensure(sizeof(float) * hv::kMlpParams));
        std::vector<int> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        uint64_t x = seed;
        auto next = [&]() {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        int step = 0;
        for (int e = 0; e < epochs; ++e) {
            for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
            HV_HIP(hipMemcpyAsync(w.idx.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
            hv::launch_fill_i32(w.loss.as<int>(), 1, 0, s);
            int batches = 0;
            for (int64_t b0 = 0; b0 + B <= n; b0 += B) {  // full batches only
                hv::launch_mlp_step(w.X.as<float>(), w.T.as<float>(), w.idx.as<int>() + b0, B, h->m_params.as<float>(),
                                    h->m_norm.as<float>(), w.Xn.as<float>(), w.H.as<float>(), w.Y.as<float>(),
                                    w.dY.as<float>(), w.dH.as<float>(), w.row_loss.as<float>(), h->m_grad.as<float>(),
                                    w.loss.as<float>(), s);
                hv::launch_adam(h->m_params.as<float>(), h->m_grad.as<float>(), h->m_adam_m.as<float>(),
                                h->m_adam_v.as<float>(), lr, 0.9f, 0.999f, 1e-8f, ++step, s);
                ++batches;
            }
            HV_CHECK_LAUNCH("mlp train");
            float l = 0.f;
            HV_HIP(hipMemcpyAsync(&l, w.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
            HV_HIP(hipStreamSynchronize(s));
            if (epoch_loss) epoch_loss[e] = l / (float)std::max(1, batches);
        }
        mlp_refresh_serving(h);
    });
}

int hv_mlp_predict(hv_index_t* h, int64_t n, const float* feat, float* out) {
    if (!h) return HV_ERR_INVALID_ARG;
    return guarded(h, [&] {
        require(h->mlp_params && h->mlp_norm, HV_ERR_STATE, "set the parameters and the standardization first");
        require(n >= 0 && (n == 0 || (feat && out)), HV_ERR_INVALID_ARG, "bad buffers");
        if (n == 0) return;
        hipStream_t s = h->stream;
        upload(h->s_feat, feat, sizeof(float) * n * hv::kMlpIn, s);
        HV_HIP(h->s_pred.ensure(sizeof(float) * n * 2));
        hv::launch_mlp_predict(h->s_feat.as<float>(), n, h->m_w1b.as<__bf16>(), h->m_w2b.as<__bf16>(),
                               h->m_params.as<float>(), h->m_norm.as<float>(), h->s_pred.as<float>(), s);
        HV_CHECK_LAUNCH("mlp predict");
        HV_HIP(hipMemcpyAsync(out, h->s_pred.p, sizeof(float) * n * 2, hipMemcpyDeviceToHost, s));
        HV_HIP(hipStreamSynchronize(s));
    });
}

int hv_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
int hv_device_malloc(void** ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return HV_ERR_INVALID_ARG;
    return hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1)) == hipSuccess ? HV_OK : HV_ERR_DEVICE;
}
int hv_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? HV_OK : HV_ERR_DEVICE; }
int hv_memcpy_htod(void* dst, const void* src, int64_t bytes) {
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? HV_OK : HV_ERR_DEVICE;
}
int hv_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? HV_OK : HV_ERR_DEVICE;
}

int hv_memcpy_dtod(void* dst, const void* src, int64_t bytes) {
    return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice) == hipSuccess ? HV_OK : HV_ERR_DEVICE;
}

}  // extern "C"
