// CPU restatement of Harmony's per-node IVF-flat search path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product library
// (harmony_amd/) never links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference checkout root). Two tie modes are offered:
//   HV_TIE_CANONICAL (0): results ordered by (distance, id) lexicographically; the top-k is
//       the k smallest (distance, id) pairs. This is the order the GPU path produces.
//   HV_TIE_VERBATIM  (1): the reference's own heap: strict '<' on distance only, unstable
//       heapsort (src/heap.hpp:35-82). Differs from canonical only among equal distances.
//
// Distance arithmetic. The reference loop is `diff = a-b; dist += diff*diff` in index order
// (src/utils.h:480-487). GCC contracts that to a fused multiply-add chain unless fast-math
// reassociates it, so the exact build (default) writes the chain explicitly with fmaf; the
// GPU kernels use the same chain, so FP32 results are bit-identical on any data. Built with
// -DORACLE_REFERENCE_FLAGS the loop is written as in the reference and compiled with the
// reference Release flags (-Ofast -march=native), which lets the compiler vectorise and
// reassociate it: that build is the timed CPU baseline, not a parity checker.
//
// Parity pins (tests/test_oracle_*.py): heap vs the reference heap.hpp compiled from source
// (oracle/_ref), top-k vs full sort, scan vs FP64 brute force at nprobe = nlist, pruning on
// == off, bound >= true k-th, k-means step vs closed-form mean, assignment vs argmin, recall
// and r2 vs hand-computed values (tests/test_metrics.py).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>
#include <omp.h>

typedef int64_t idx_t;

namespace {

enum { TIE_CANONICAL = 0, TIE_VERBATIM = 1 };

// ---------------------------------------------------------------- distance (utils.h:469-488)
#ifdef ORACLE_REFERENCE_FLAGS
inline float l2_range(const float* a, const float* b, idx_t n) {
    float distance = 0.0f;
    for (idx_t i = 0; i < n; ++i) {
        float diff = a[i] - b[i];
        distance += diff * diff;
    }
    return distance;
}
#else
inline float l2_range(const float* a, const float* b, idx_t n) {
    float acc = 0.0f;
    for (idx_t i = 0; i < n; ++i) {
        float t = a[i] - b[i];
        acc = std::fmaf(t, t, acc);
    }
    return acc;
}
#endif

// Blocked L2 as in node.cpp:639-651: B blocks of floor(d/B) dims, the last block takes the
// remainder; each block's partial sum starts from zero and is added to the running total.
// Returns +inf when the running total exceeds `bound` after a non-last block (strict >).
inline float l2_blocked(const float* q, const float* x, idx_t d, idx_t B, float bound, bool prune) {
    idx_t w = d / B;
    float dis = 0.0f;
    for (idx_t b = 0; b < B; ++b) {
        idx_t lo = b * w;
        idx_t n = (b == B - 1) ? d - lo : w;
        dis += l2_range(q + lo, x + lo, n);
        if (prune && b != B - 1 && dis > bound) return INFINITY;
    }
    return dis;
}

// Inner product (utils.h:502-520): sum += a[i] * b[i] in index order, the FMA chain a contracting
// compiler emits (the reference-flags build writes the loop as the reference does).
#ifdef ORACLE_REFERENCE_FLAGS
inline float ip_range(const float* a, const float* b, idx_t n) {
    float sum = 0.0f;
    for (idx_t i = 0; i < n; ++i) sum += a[i] * b[i];
    return sum;
}
#else
inline float ip_range(const float* a, const float* b, idx_t n) {
    float acc = 0.0f;
    for (idx_t i = 0; i < n; ++i) acc = std::fmaf(a[i], b[i], acc);
    return acc;
}
#endif

// ---------------------------------------------------------------- heap (heap.hpp:13-90)
// Verbatim restatement of heap_replace_top / heap_sort for METRIC_L2 (max-heap in place,
// 1-based child arithmetic on pointers shifted by one).
inline void heap_replace_top_verbatim(idx_t k, float* bh_val, idx_t* bh_ids, float val, idx_t id) {
    bh_val--;
    bh_ids--;
    idx_t i = 1, i1, i2;
    while (true) {
        i1 = i * 2;
        i2 = i1 + 1;
        if (i1 > k) break;
        if ((i2 == k + 1) || (bh_val[i1] > bh_val[i2])) {
            if (val > bh_val[i1]) break;
            bh_val[i] = bh_val[i1];
            bh_ids[i] = bh_ids[i1];
            i = i1;
        } else {
            if (val > bh_val[i2]) break;
            bh_val[i] = bh_val[i2];
            bh_ids[i] = bh_ids[i2];
            i = i2;
        }
    }
    bh_val[i] = val;
    bh_ids[i] = id;
}

inline void heap_sort_verbatim(idx_t k, float* bh_val, idx_t* bh_ids) {
    for (idx_t i = k - 1; i >= 0; i--) {
        float val = bh_val[i];
        idx_t id = bh_ids[i];
        bh_val[i] = bh_val[0];
        bh_ids[i] = bh_ids[0];
        heap_replace_top_verbatim(i, bh_val, bh_ids, val, id);
    }
}

// Canonical heap: same sift-down, ordered lexicographically on (value, id).
inline bool gt_pair(float a, idx_t ia, float b, idx_t ib) { return a > b || (a == b && ia > ib); }

inline void heap_replace_top_canon(idx_t k, float* bh_val, idx_t* bh_ids, float val, idx_t id) {
    bh_val--;
    bh_ids--;
    idx_t i = 1, i1, i2;
    while (true) {
        i1 = i * 2;
        i2 = i1 + 1;
        if (i1 > k) break;
        idx_t c;
        if ((i2 == k + 1) || gt_pair(bh_val[i1], bh_ids[i1], bh_val[i2], bh_ids[i2])) c = i1;
        else c = i2;
        if (gt_pair(val, id, bh_val[c], bh_ids[c])) break;
        bh_val[i] = bh_val[c];
        bh_ids[i] = bh_ids[c];
        i = c;
    }
    bh_val[i] = val;
    bh_ids[i] = id;
}

inline void heap_sort_canon(idx_t k, float* bh_val, idx_t* bh_ids) {
    for (idx_t i = k - 1; i >= 0; i--) {
        float val = bh_val[i];
        idx_t id = bh_ids[i];
        bh_val[i] = bh_val[0];
        bh_ids[i] = bh_ids[0];
        heap_replace_top_canon(i, bh_val, bh_ids, val, id);
    }
}

// heap_init for L2 (heap.hpp:13-25): FLT_MAX / -1.
inline void heap_init(idx_t k, float* v, idx_t* ids) {
    for (idx_t i = 0; i < k; ++i) {
        v[i] = FLT_MAX;
        ids[i] = -1;
    }
}

// The insert guard used by every L2 caller: `if (dis < simi[0])` (IVFScan.hpp:131,267;
// node.cpp:654). Canonical mode extends it to (dis, id) < (top, top_id).
inline void heap_push(int tie, idx_t k, float* v, idx_t* ids, float dis, idx_t id) {
    if (tie == TIE_VERBATIM) {
        if (dis < v[0]) heap_replace_top_verbatim(k, v, ids, dis, id);
    } else {
        if (gt_pair(v[0], ids[0], dis, id)) heap_replace_top_canon(k, v, ids, dis, id);
    }
}

inline void heap_finish(int tie, idx_t k, float* v, idx_t* ids) {
    if (tie == TIE_VERBATIM) heap_sort_verbatim(k, v, ids);
    else heap_sort_canon(k, v, ids);
}

// Contiguous batch split used by every OpenMP driver in the reference
// (Index.cpp:1667-1700, node.cpp:553-578): nt = min(threads, n), the first n % nt batches
// get one extra element.
template <typename F>
void for_batches(idx_t n, int nthreads, F&& f) {
    if (n <= 0) return;
    int maxt = nthreads > 0 ? nthreads : omp_get_max_threads();
    idx_t nt = std::min<idx_t>(maxt, n);
    idx_t bs = n / nt, extra = n % nt;
#pragma omp parallel for num_threads((int)nt) schedule(static, 1)
    for (idx_t t = 0; t < nt; ++t) {
        idx_t s = t < extra ? t * (bs + 1) : t * bs + extra;
        idx_t e = s + (t < extra ? bs + 1 : bs);
        for (idx_t i = s; i < e; ++i) f(i);
    }
}

// ---------------------------------------------------------------- PRNG + shuffle
// SplitMix64 stream and Fisher-Yates shuffle shared with the library's index build
// (harmony_amd/csrc/hv_index.cpp). The reference uses std::default_random_engine(seed)
// with std::shuffle (Clustering.cpp:58,81), which is library-specific; the build defines its
// own portable stream instead.
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

void shuffle_indices(std::vector<idx_t>& v, uint64_t seed) {
    SplitMix64 rng(seed);
    for (idx_t i = (idx_t)v.size() - 1; i > 0; --i) {
        idx_t j = (idx_t)(rng.next() % (uint64_t)(i + 1));
        std::swap(v[i], v[j]);
    }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- exported heap primitives
void or_heap_init(idx_t k, float* v, idx_t* ids) { heap_init(k, v, ids); }
void or_heap_push(int tie, idx_t k, float* v, idx_t* ids, float dis, idx_t id) { heap_push(tie, k, v, ids, dis, id); }
void or_heap_replace_top(int tie, idx_t k, float* v, idx_t* ids, float dis, idx_t id) {
    if (tie == TIE_VERBATIM) heap_replace_top_verbatim(k, v, ids, dis, id);
    else heap_replace_top_canon(k, v, ids, dis, id);
}
void or_heap_sort(int tie, idx_t k, float* v, idx_t* ids) { heap_finish(tie, k, v, ids); }

// ---------------------------------------------------------------- distances
float or_l2(const float* a, const float* b, idx_t d) { return l2_range(a, b, d); }
float or_l2_blocked(const float* a, const float* b, idx_t d, idx_t B) { return l2_blocked(a, b, d, B, 0.f, false); }

// ---------------------------------------------------------------- coarse (Index.cpp:967-997)
// Per query: lite_scan_codes over all centroids into an nprobe-heap (IVFScan.hpp:108-138,
// strict '<' in centroid-id order), then sort_result (heap.hpp:84-90).
void or_coarse(idx_t nq, const float* q, idx_t d, idx_t nlist, const float* centroids, idx_t nprobe,
               int tie, int nthreads, idx_t* probes, float* pdist) {
    for_batches(nq, nthreads, [&](idx_t i) {
        float* v = pdist + i * nprobe;
        idx_t* ids = probes + i * nprobe;
        heap_init(nprobe, v, ids);
        for (idx_t c = 0; c < nlist; ++c) {
            float dis = l2_range(q + i * d, centroids + c * d, d);
            heap_push(tie, nprobe, v, ids, dis, c);
        }
        heap_finish(tie, nprobe, v, ids);
    });
}

// ---------------------------------------------------------------- warm-up bound
// Index.cpp:999-1051 + 1303-1305: per query, a k-heap over the first min(|L|, wls) vectors
// of each of the first wl probed lists; the bound is the heap top (FLT_MAX while fewer than
// k were seen). The reference scans with the full-d distance (lite_scan_codes); with
// blocked != 0 the blocked sum of the scan is used instead, which is what the GPU path does
// so that every warm-up vector survives its own bound.
void or_warmup_bound(idx_t nq, const float* q, idx_t d, const idx_t* list_off, const float* codes,
                     idx_t k, idx_t nprobe, const idx_t* probes, idx_t wl, idx_t wls, idx_t B, int blocked,
                     int nthreads, float* bound) {
    if (wl > nprobe) wl = nprobe;
    for_batches(nq, nthreads, [&](idx_t i) {
        std::vector<float> v(k);
        std::vector<idx_t> ids(k);
        heap_init(k, v.data(), ids.data());
        for (idx_t j = 0; j < wl; ++j) {
            idx_t l = probes[i * nprobe + j];
            idx_t n = list_off[l + 1] - list_off[l];
            idx_t m = std::min(n, wls);
            for (idx_t t = 0; t < m; ++t) {
                const float* x = codes + (list_off[l] + t) * d;
                float dis = blocked ? l2_blocked(q + i * d, x, d, B, 0.f, false) : l2_range(q + i * d, x, d);
                heap_push(TIE_VERBATIM, k, v.data(), ids.data(), dis, list_off[l] + t);
            }
        }
        bound[i] = v[0];
    });
}

// ---------------------------------------------------------------- search
// Per-node search with B dimension blocks (node.cpp:606-678), driven per query by the master
// sequence of search_divide_ivf (Index.cpp:1279-1343): coarse, warm-up bound, per-node scan
// into a fresh k-heap, merge into the output heap, sort_result. With pruning == 0 the scan
// uses the full-d distance as in node.cpp:658-667 (the same arithmetic as the single-process
// ORIGINAL path, Index.cpp:1420-1554).
//
// list_off[nlist+1], ids[nb], codes[nb][d]: the inverted lists in list order, each list in
// its within-list order. probes_out (nq*nprobe) and bound_out (nq) are optional.
int or_search(idx_t nq, const float* q, idx_t d, idx_t nlist, const float* centroids, const idx_t* list_off,
              const idx_t* ids, const float* codes, idx_t k, idx_t nprobe, idx_t B, int pruning, idx_t wl,
              idx_t wls, int tie, int bound_blocked, int nthreads, float* D, idx_t* I, idx_t* probes_out,
              float* bound_out) {
    if (nq <= 0) return 0;
    if (d <= 0 || k <= 0 || nlist <= 0 || B <= 0 || B > d) return -1;
    if (nprobe > nlist) nprobe = nlist;  // Index.cpp:1646-1648
    if (nprobe <= 0) return -1;
    if (wl > 0 && wls == 0) wls = k;      // query.cpp:328-333
    std::vector<idx_t> probes(nq * nprobe);
    std::vector<float> pdist(nq * nprobe);
    or_coarse(nq, q, d, nlist, centroids, nprobe, tie, nthreads, probes.data(), pdist.data());
    std::vector<float> bound(nq, FLT_MAX);
    if (wl > 0 && (pruning || bound_out))
        or_warmup_bound(nq, q, d, list_off, codes, k, nprobe, probes.data(), wl, wls, B, bound_blocked, nthreads,
                        bound.data());
    for_batches(nq, nthreads, [&](idx_t i) {
        const float* qi = q + i * d;
        std::vector<float> sv(k);
        std::vector<idx_t> si(k);
        heap_init(k, sv.data(), si.data());
        for (idx_t j = 0; j < nprobe; ++j) {
            idx_t l = probes[i * nprobe + j];
            for (idx_t p = list_off[l]; p < list_off[l + 1]; ++p) {
                const float* x = codes + p * d;
                float dis = pruning ? l2_blocked(qi, x, d, B, bound[i], true) : l2_range(qi, x, d);
                heap_push(tie, k, sv.data(), si.data(), dis, ids[p]);
            }
        }
        // merge into the output row (Index.cpp:1327-1343)
        float* Di = D + i * k;
        idx_t* Ii = I + i * k;
        heap_init(k, Di, Ii);
        for (idx_t t = 0; t < k; ++t) heap_push(tie, k, Di, Ii, sv[t], si[t]);
        heap_finish(tie, k, Di, Ii);
    });
    if (probes_out) std::memcpy(probes_out, probes.data(), sizeof(idx_t) * nq * nprobe);
    if (bound_out) std::memcpy(bound_out, bound.data(), sizeof(float) * nq);
    return 0;
}

// ---------------------------------------------------------------- FP64 brute force
// Exact kNN over all nb vectors in FP64, ordered by (distance, id). This replaces the
// reference's faiss ground truth (query.cpp:481-514: IndexIVFFlat at nprobe = nlist).
void or_bruteforce_f64(idx_t nb, const float* xb, idx_t nq, const float* q, idx_t d, idx_t k, int nthreads,
                       double* D, idx_t* I) {
    for_batches(nq, nthreads, [&](idx_t i) {
        std::vector<std::pair<double, idx_t>> h;
        h.reserve(k + 1);
        auto cmp = [](const std::pair<double, idx_t>& a, const std::pair<double, idx_t>& b) { return a < b; };
        for (idx_t v = 0; v < nb; ++v) {
            double s = 0.0;
            for (idx_t t = 0; t < d; ++t) {
                double df = (double)q[i * d + t] - (double)xb[v * d + t];
                s += df * df;
            }
            std::pair<double, idx_t> e(s, v);
            if ((idx_t)h.size() < k) {
                h.push_back(e);
                std::push_heap(h.begin(), h.end(), cmp);
            } else if (e < h.front()) {
                std::pop_heap(h.begin(), h.end(), cmp);
                h.back() = e;
                std::push_heap(h.begin(), h.end(), cmp);
            }
        }
        std::sort(h.begin(), h.end());
        for (idx_t t = 0; t < k; ++t) {
            if (t < (idx_t)h.size()) {
                D[i * k + t] = h[t].first;
                I[i * k + t] = h[t].second;
            } else {
                D[i * k + t] = INFINITY;
                I[i * k + t] = -1;
            }
        }
    });
}

// ---------------------------------------------------------------- index build
// Nearest centroid per vector (Index.cpp:567-579, 596-614): full-d FP32 distance, strict '<'
// in centroid order, so ties go to the lower centroid id.
void or_assign(idx_t n, const float* x, idx_t d, idx_t nlist, const float* centroids, int nthreads, idx_t* labels,
               float* dist) {
    for_batches(n, nthreads, [&](idx_t i) {
        float best = FLT_MAX;
        idx_t arg = -1;
        for (idx_t c = 0; c < nlist; ++c) {
            float dis = l2_range(x + i * d, centroids + c * d, d);
            if (dis < best) {
                best = dis;
                arg = c;
            }
        }
        labels[i] = arg;
        dist[i] = best;
    });
}

// List order (Index.cpp:624-671): vectors grouped by list, each list ascending by distance to
// its centroid. The reference sorts with an unstable std::sort on distance only; ties here go
// to the lower id so the order is reproducible. Output: list_off[nlist+1], order[n] = ids in
// list order.
void or_build_lists(idx_t n, const idx_t* labels, const float* dist, idx_t nlist, idx_t* list_off, idx_t* order) {
    std::vector<idx_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](idx_t a, idx_t b) {
        if (labels[a] != labels[b]) return labels[a] < labels[b];
        if (dist[a] != dist[b]) return dist[a] < dist[b];
        return a < b;
    });
    std::fill(list_off, list_off + nlist + 1, 0);
    for (idx_t i = 0; i < n; ++i) list_off[labels[i] + 1]++;
    for (idx_t l = 0; l < nlist; ++l) list_off[l + 1] += list_off[l];
    std::memcpy(order, idx.data(), sizeof(idx_t) * n);
}

// k-means (Clustering.cpp:14-152, with the build's parameters): subsample to at most
// max_pts*nlist points by a seeded shuffle (48-68), initialise with the first nlist points of
// a second shuffle of the sample under the same seed (72-89), then niter Lloyd steps: nearest
// centroid (strict '<'), FP64 sums in ascending sample order, centroid = sum/count rounded to
// FP32, empty clusters keep their centroid (146-150).
void or_kmeans_init(idx_t n, const float* x, idx_t d, idx_t nlist, idx_t max_pts, uint64_t seed, float* sample_out,
                    idx_t* n_sample_out, float* centroids) {
    idx_t ns = n;
    std::vector<idx_t> perm;
    const float* src = x;
    std::vector<float> sample;
    if (max_pts > 0 && n > nlist * max_pts) {
        ns = nlist * max_pts;
        perm.resize(n);
        std::iota(perm.begin(), perm.end(), 0);
        shuffle_indices(perm, seed);
        sample.resize(ns * d);
        for (idx_t i = 0; i < ns; ++i) std::memcpy(&sample[i * d], x + perm[i] * d, sizeof(float) * d);
        src = sample.data();
    }
    if (sample_out) std::memcpy(sample_out, src, sizeof(float) * ns * d);
    *n_sample_out = ns;
    std::vector<idx_t> init(ns);
    std::iota(init.begin(), init.end(), 0);
    shuffle_indices(init, seed);
    for (idx_t c = 0; c < nlist; ++c) std::memcpy(centroids + c * d, src + init[c] * d, sizeof(float) * d);
}

void or_kmeans_step(idx_t n, const float* x, idx_t d, idx_t nlist, float* centroids, int nthreads) {
    std::vector<idx_t> lab(n);
    std::vector<float> dist(n);
    or_assign(n, x, d, nlist, centroids, nthreads, lab.data(), dist.data());
    std::vector<double> sum(nlist * d, 0.0);
    std::vector<idx_t> cnt(nlist, 0);
    for (idx_t i = 0; i < n; ++i) {
        idx_t c = lab[i];
        cnt[c]++;
        for (idx_t t = 0; t < d; ++t) sum[c * d + t] += (double)x[i * d + t];
    }
    for (idx_t c = 0; c < nlist; ++c)
        if (cnt[c] > 0)
            for (idx_t t = 0; t < d; ++t) centroids[c * d + t] = (float)(sum[c * d + t] / (double)cnt[c]);
}

void or_kmeans(idx_t n, const float* x, idx_t d, idx_t nlist, idx_t niter, idx_t max_pts, uint64_t seed,
               int nthreads, float* centroids) {
    idx_t ns = 0;
    idx_t cap = (max_pts > 0 && n > nlist * max_pts) ? nlist * max_pts : n;
    std::vector<float> sample(cap * d);
    or_kmeans_init(n, x, d, nlist, max_pts, seed, sample.data(), &ns, centroids);
    for (idx_t it = 0; it < niter; ++it) or_kmeans_step(ns, sample.data(), d, nlist, centroids, nthreads);
}

// ---------------------------------------------------------------- inner-product metric
// Row normalization (Eigen's colwise().normalized() in Clustering.cpp:122, with the norm from the
// FP32 chain): x / sqrt(<x, x>), a zero row unchanged.
void or_ip_normalize_rows(idx_t n, const float* x, idx_t d, float* out) {
    for (idx_t i = 0; i < n; ++i) {
        const float nn = std::sqrt(ip_range(x + i * d, x + i * d, d));
        for (idx_t t = 0; t < d; ++t) out[i * d + t] = nn > 0.0f ? x[i * d + t] / nn : x[i * d + t];
    }
}

// Centroid normalization after training (Index.cpp:408-418): norm = <c, c>; if non-zero every
// coordinate is divided by sqrt(norm).
void or_ip_normalize_centroids(idx_t nlist, idx_t d, float* c) {
    for (idx_t l = 0; l < nlist; ++l) {
        const float norm = ip_range(c + l * d, c + l * d, d);
        if (norm != 0.0f) {
            const float r = std::sqrt(norm);
            for (idx_t t = 0; t < d; ++t) c[l * d + t] /= r;
        }
    }
}

// Assignment by the largest inner product (the IP scanner's lite_scan_codes with a 1-heap: strict
// '>' in centroid order, so ties go to the lower id; Index.cpp:567-579 / IVFScan.hpp:116-123).
void or_ip_assign(idx_t n, const float* x, idx_t d, idx_t nlist, const float* centroids, int nthreads, idx_t* labels,
                  float* best) {
    for_batches(n, nthreads, [&](idx_t i) {
        float bv = -FLT_MAX;
        idx_t bl = -1;
        for (idx_t c = 0; c < nlist; ++c) {
            const float v = ip_range(x + i * d, centroids + c * d, d);
            if (v > bv || bl < 0) {
                bv = v;
                bl = c;
            }
        }
        labels[i] = bl;
        best[i] = bv;
    });
}

// IP k-means (Clustering.cpp:121-151 + Index.cpp:408-418): the same subsample and init as L2; per
// iteration each sample row goes to the centroid with the largest <normalized row, centroid>, the
// centroid becomes the FP64 mean of its raw rows (empty clusters keep theirs); after the last
// iteration the centroids are normalized.
void or_ip_kmeans(idx_t n, const float* x, idx_t d, idx_t nlist, idx_t niter, idx_t max_pts, uint64_t seed,
                  int nthreads, float* centroids) {
    idx_t ns = 0;
    idx_t cap = (max_pts > 0 && n > nlist * max_pts) ? nlist * max_pts : n;
    std::vector<float> sample(cap * d), xn(cap * d);
    or_kmeans_init(n, x, d, nlist, max_pts, seed, sample.data(), &ns, centroids);
    or_ip_normalize_rows(ns, sample.data(), d, xn.data());
    std::vector<idx_t> lab(ns);
    std::vector<float> best(ns);
    for (idx_t it = 0; it < niter; ++it) {
        or_ip_assign(ns, xn.data(), d, nlist, centroids, nthreads, lab.data(), best.data());
        std::vector<double> sum(nlist * d, 0.0);
        std::vector<idx_t> cnt(nlist, 0);
        for (idx_t i = 0; i < ns; ++i) {
            cnt[lab[i]]++;
            for (idx_t t = 0; t < d; ++t) sum[lab[i] * d + t] += (double)sample[i * d + t];
        }
        for (idx_t c = 0; c < nlist; ++c)
            if (cnt[c] > 0)
                for (idx_t t = 0; t < d; ++t) centroids[c * d + t] = (float)(sum[c * d + t] / (double)cnt[c]);
    }
    or_ip_normalize_centroids(nlist, d, centroids);
}

// IP lists (Index.cpp:581-671): vectors by descending inner product with their centroid (the
// candidate2centroid array), ties by id; off[nlist+1], order[n].
void or_ip_build_lists(idx_t n, const idx_t* labels, const float* best, idx_t nlist, idx_t* off, idx_t* order) {
    std::fill(off, off + nlist + 1, 0);
    for (idx_t i = 0; i < n; ++i) off[labels[i] + 1]++;
    for (idx_t l = 0; l < nlist; ++l) off[l + 1] += off[l];
    std::vector<idx_t> cur(off, off + nlist);
    for (idx_t i = 0; i < n; ++i) order[cur[labels[i]]++] = i;
    for (idx_t l = 0; l < nlist; ++l)
        std::sort(order + off[l], order + off[l + 1], [&](idx_t a, idx_t b) {
            return best[a] != best[b] ? best[a] > best[b] : a < b;
        });
}

// IP search, ORIGINAL path (Index.cpp:1420-1449, 1509-1551): the nprobe centroids with the largest
// inner product (ties to the lower id), then each probed list in that order. With triangle pruning
// (opt level bit OPT_TRIANGLE) the scanned range of a list is cut from both ends by the angular
// band the current k-th result allows (Index.cpp:1515-1546): s = the heap's worst inner product
// (-FLT_MAX until k results are held, which leaves the band undefined and the list whole),
// c = <q, centroid>; below s < c only the tail with candidate2centroid < s c - sqrt(1-s^2) sqrt(1-c^2)
// is cut, otherwise also the head with candidate2centroid > s c + sqrt(1-s^2) sqrt(1-c^2). Results
// are kept as the top k by (inner product descending, id ascending), padded with (-FLT_MAX, -1).
// skipped (may be null): vectors cut by the triangle ranges, summed over the queries.
void or_ip_search(idx_t nq, const float* q, idx_t d, idx_t nlist, const float* centroids, const idx_t* list_off,
                  const idx_t* ids, const float* codes, const float* c2c, idx_t k, idx_t nprobe, int triangle,
                  int nthreads, float* D, idx_t* I, idx_t* probes, idx_t* skipped) {
    nprobe = std::min(nprobe, nlist);
    std::vector<idx_t> skip(nq, 0);
    for_batches(nq, nthreads, [&](idx_t i) {
        const float* qi = q + i * d;
        std::vector<std::pair<float, idx_t>> cs(nlist);
        for (idx_t c = 0; c < nlist; ++c) cs[c] = {ip_range(qi, centroids + c * d, d), c};
        std::partial_sort(cs.begin(), cs.begin() + nprobe, cs.end(), [](const auto& a, const auto& b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        });
        // canonical top-k: a list kept sorted by (ip desc, id asc), at most k long
        std::vector<std::pair<float, idx_t>> top;
        auto better = [](const std::pair<float, idx_t>& a, const std::pair<float, idx_t>& b) {
            return a.first != b.first ? a.first > b.first : a.second < b.second;
        };
        for (idx_t j = 0; j < nprobe; ++j) {
            const idx_t l = cs[j].second;
            if (probes) probes[i * nprobe + j] = l;
            const float cq = cs[j].first;
            const float simi = (idx_t)top.size() < k ? -FLT_MAX : top.back().first;
            const idx_t b = list_off[l], len = list_off[l + 1] - b;
            idx_t begin = 0, end = len;
            if (triangle) {
                const float s_c = std::sqrt(1 - cq * cq);
                const float s_s = std::sqrt(1 - simi * simi);
                if (simi < cq) {
                    const float min_cut = simi * cq - s_s * s_c;
                    while (begin < end && c2c[b + end - 1] < min_cut) end--;
                } else {
                    const float max_cut = simi * cq + s_s * s_c;
                    const float min_cut = simi * cq - s_s * s_c;
                    while (begin < end && c2c[b + begin] > max_cut) begin++;
                    while (begin < end && c2c[b + end - 1] < min_cut) end--;
                }
                skip[i] += begin + (len - end);
            }
            for (idx_t t = begin; t < end; ++t) {
                const std::pair<float, idx_t> cand{ip_range(qi, codes + (b + t) * d, d), ids[b + t]};
                if ((idx_t)top.size() < k) {
                    top.insert(std::upper_bound(top.begin(), top.end(), cand, better), cand);
                } else if (better(cand, top.back())) {
                    top.pop_back();
                    top.insert(std::upper_bound(top.begin(), top.end(), cand, better), cand);
                }
            }
        }
        for (idx_t j = 0; j < k; ++j) {
            D[i * k + j] = j < (idx_t)top.size() ? top[j].first : -FLT_MAX;
            I[i * k + j] = j < (idx_t)top.size() ? top[j].second : -1;
        }
    });
    if (skipped) {
        idx_t t = 0;
        for (idx_t i = 0; i < nq; ++i) t += skip[i];
        *skipped = t;
    }
}

// ---------------------------------------------------------------- metrics (utils.h:660-756)
// calculate_recall (utils.h:660-732). Returns 0 when D[0] < 0 (662-665); gt_k = 0 means k
// (666-668); k > gt_k is an error (671-673: std::invalid_argument, here *status = -1). For each
// query the ground-truth set is the first k ids of its GT row (676); the query's ids are
// counted in order until the first -1 (678-680), each one found in the set is a hit, so a
// repeated id is counted every time it appears. The distance sanity count (687-729: a result
// is "correct" when its distance is within the row's last valid GT distance, or within a
// relative 1e-4 of it; >= for IP) goes to *correct; the reference asserts true/correct > 0.99
// in Debug builds only (730). Result: (float)hits / (nq * k) in FP32, as 731 computes it.
// One deviation: a GT row whose first k ids are all -1 makes the reference's index walk
// (691-694) run below the row (undefined); here it stops at the row's first entry.
float or_recall(idx_t nq, idx_t k, const idx_t* I, const float* D, const idx_t* GT, const float* GD, idx_t gt_k,
                int metric_l2, idx_t* correct_out, int* status) {
    if (status) *status = 0;
    if (correct_out) *correct_out = 0;
    if (D && D[0] < 0) return 0.0f;
    if (gt_k == 0) gt_k = k;
    if (k > gt_k) {
        if (status) *status = -1;
        return 0.0f;
    }
    idx_t true_correct = 0, correct = 0;
    for (idx_t i = 0; i < nq; ++i) {
        std::vector<idx_t> g(GT + i * gt_k, GT + i * gt_k + k);
        std::sort(g.begin(), g.end());
        for (idx_t j = 0; j < k; ++j) {
            const idx_t v = I[i * k + j];
            if (v == -1) break;
            if (std::binary_search(g.begin(), g.end(), v)) true_correct++;
        }
    }
    if (D && GD) {
        for (idx_t i = 0; i < nq; ++i) {
            idx_t ii = k - 1;
            while (ii > 0 && GT[i * gt_k + ii] == -1) ii--;
            const float topK = GD[i * gt_k + ii];
            for (idx_t j = 0; j < k; ++j) {
                if (I[i * k + j] == -1) break;
                const float x = D[i * k + j];
                const float rel = x == 0.0f ? std::fabs(topK) : std::fabs((x - topK) / x);  // utils.h:622-628
                const bool within = metric_l2 ? x <= topK : x >= topK;
                if (within || (double)rel < 1e-4) correct++;
            }
        }
    }
    if (correct_out) *correct_out = correct;
    return (float)true_correct / (float)(nq * k);
}

// calculate_r2 (utils.h:734-756): FP32 sums of D and of GD over each row's entries before the
// first -1 in I, r2 = sum / g_sum - 1 in FP32; k > gt_k is an error (*status = -1).
float or_r2(idx_t nq, idx_t k, const idx_t* I, const float* D, const float* GD, idx_t gt_k, int* status) {
    if (status) *status = 0;
    if (gt_k == 0) gt_k = k;
    if (k > gt_k) {
        if (status) *status = -1;
        return 0.0f;
    }
    float gs = 0, s = 0;
    for (idx_t i = 0; i < nq; ++i)
        for (idx_t j = 0; j < k; ++j) {
            if (I[i * k + j] == -1) break;
            gs += GD[i * gt_k + j];
            s += D[i * k + j];
        }
    return s / gs - 1;
}

}  // extern "C"
