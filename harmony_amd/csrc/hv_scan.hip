// The IVF list scan of Harmony's per-node search path on gfx950: warm-up bound, grouping of
// (query, probed list) pairs by list, the dimension-blocked pruned list scan, and the
// per-query top-k selection.
//
// Reference semantics (src/node.cpp:606-678, src/Index.cpp:999-1051, 1279-1343): for every
// probed list of a query, every vector's L2 distance is accumulated block by block over B
// dimension blocks; after each block but the last the candidate is dropped when its partial
// sum exceeds the query's bound (strict >); the survivors compete for the top-k. Because the
// bound is the k-th smallest distance over a subset of the candidates, every result has a
// distance <= bound, so the scan keeps exactly the non-pruned candidates with distance
// <= bound and the selection sorts them by (distance, id). The result equals the k smallest
// (distance, id) over all candidates, with pruning on or off.
#include <algorithm>

#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

// Blocked distance of one (query, vector) pair for the single-thread kernels below: q in LDS,
// the vector as column t of a list's [d][stride] slab. Returns +inf once pruned.
__device__ __forceinline__ float blocked_l2_col(const float* qs, const float* __restrict__ X, int stride, int t,
                                                const ChunkPlan& plan, bool prune, float bound) {
    float dis = 0.f, p = 0.f;
    for (int c = 0; c < plan.n; ++c) {
        const int lo = plan.lo[c], len = plan.len[c];
        // the chunk's loads first (one memory latency per chunk), then its FMA chain in order
        float xv[kChunkDims];
#pragma unroll
        for (int u = 0; u < kChunkDims; ++u) xv[u] = X[(int64_t)(lo + min(u, len - 1)) * stride + t];
#pragma unroll
        for (int u = 0; u < kChunkDims; ++u) {
            if (u < len) {
                const float t2 = qs[lo + u] - xv[u];
                p = __builtin_fmaf(t2, t2, p);
            }
        }
        if (plan.ends_block[c]) {
            dis += p;
            p = 0.f;
            if (prune && plan.prune_after[c] && dis > bound) return __builtin_inff();
        }
    }
    return dis;
}

// ---------------------------------------------------------------------------------------
// Grouping: (query, probe rank) pairs bucketed by (list, rank). Within a list, queries that
// rank the list first come first, so a query tile holds queries at similar distance from the
// list and prunes (or survives) together.
// qn (optional): per-query probe count; ranks at or past it are not searched
// one thread per (query, rank < nranks); probes and slots are [nq][nprobe], buckets
// [nlist][nranks] (bucket of (list l, rank j) = l * nranks + j)
__global__ void k_count_pairs(const int* __restrict__ probes, int64_t total, int nprobe, int nranks,
                              const int* __restrict__ qn, int* __restrict__ cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int64_t q = hv_idiv(e, nranks);
    const int j = (int)(e - q * nranks);
    if (qn && j >= qn[q]) return;
    atomicAdd(&cnt[(int64_t)probes[q * nprobe + j] * nranks + j], 1);
}

void launch_count_pairs(const int* probes, int64_t nq, int nprobe, int nranks, const int* qn, int* bucket_cnt,
                        hipStream_t s) {
    const int64_t tot = nq * nranks;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, probes, tot, nprobe, nranks,
                       qn, bucket_cnt);
}

// pairs[pos] = query, pair_slot[pos] = query * nprobe + rank (the probe slot, for the coarse
// distance of the pair)
__global__ void k_scatter_pairs(const int* __restrict__ probes, int64_t total, int nprobe, int nranks,
                                const int* __restrict__ qn, const int* __restrict__ off, int* __restrict__ fill,
                                int* __restrict__ pairs, int* __restrict__ pair_slot) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int64_t q = hv_idiv(e, nranks);
    const int j = (int)(e - q * nranks);
    if (qn && j >= qn[q]) return;
    const int64_t slot = q * nprobe + j;
    const int64_t b = (int64_t)probes[slot] * nranks + j;
    const int pos = off[b] + atomicAdd(&fill[b], 1);
    pairs[pos] = (int)q;
    pair_slot[pos] = (int)slot;
}

void launch_scatter_pairs(const int* probes, int64_t nq, int nprobe, int nranks, const int* qn, const int* bucket_off,
                          int* fill, int* pairs, int* pair_slot, hipStream_t s) {
    const int64_t tot = nq * nranks;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_scatter_pairs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, probes, tot, nprobe,
                       nranks, qn, bucket_off, fill, pairs, pair_slot);
}

// Exclusive scan of n ints (three passes: tile scan, scan of tile sums, add).
constexpr int kScanTile = 2048;  // 256 threads x 8

__device__ int block_exclusive_scan_256(int v, int* s_tmp, int* total) {
    const int t = threadIdx.x;
    s_tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int a = t >= off ? s_tmp[t - off] : 0;
        __syncthreads();
        s_tmp[t] += a;
        __syncthreads();
    }
    const int incl = s_tmp[t];
    if (total) *total = s_tmp[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_tiles(const int* __restrict__ in, int64_t n, int* __restrict__ out,
                                                    int* __restrict__ tile_sums) {
    __shared__ int s_tmp[256];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 8;
    int v[8];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        sum += v[i];
    }
    int tot = 0;
    int run = block_exclusive_scan_256(sum, s_tmp, &tot);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_sums(int* __restrict__ tile_sums, int ntiles, int* __restrict__ out_total) {
    __shared__ int s_tmp[256];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < ntiles; b0 += 256) {
        const int i = b0 + threadIdx.x;
        const int v = i < ntiles ? tile_sums[i] : 0;
        int tot = 0;
        const int ex = block_exclusive_scan_256(v, s_tmp, &tot);
        if (i < ntiles) tile_sums[i] = ex + s_carry;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *out_total = s_carry;
}

__global__ void k_scan_add(int* __restrict__ out, int64_t n, const int* __restrict__ tile_sums) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tile_sums[i / kScanTile];
}

int64_t scan_tmp_ints(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// Up to kScanOne ints in one workgroup (one launch instead of three): the ints are staged in LDS
// with coalesced loads (one pad word per 16, so the 16-int runs the threads read next fall in
// distinct banks), each thread scans its run of 16, one block scan of the 1024 run sums (a
// wavefront scan and the 16 wavefront totals), and the results leave through LDS coalesced.
constexpr int kScanOne = 16384;
__global__ __launch_bounds__(1024) void k_scan_one(const int* __restrict__ in, int n, int* __restrict__ out) {
    constexpr int R = kScanOne / 1024;  // 16: the run of one thread
    __shared__ int s_v[kScanOne + kScanOne / R];
    __shared__ int s_wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
    for (int c = 0; c < R; ++c) {
        const int i = c * 1024 + t;
        s_v[i + i / R] = i < n ? in[i] : 0;
    }
    __syncthreads();
    int v[R], sum = 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
        v[j] = s_v[t * (R + 1) + j];
        sum += v[j];
    }
    const int incl = (int)wave_incl_scan((unsigned int)sum);
    if (lane == 63) s_wsum[w] = incl;
    __syncthreads();
    int run = incl - sum, total = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int x = s_wsum[i];
        run += i < w ? x : 0;
        total += x;
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        s_v[t * (R + 1) + j] = run;
        run += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < R; ++c) {
        const int i = c * 1024 + t;
        if (i < n) out[i] = s_v[i + i / R];
    }
    if (t == 0) out[n] = total;
}

void launch_exclusive_scan(const int* in, int64_t n, int* out, int* tmp, hipStream_t s) {
    if (n <= kScanOne) {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, s, in, (int)n, out);
        return;
    }
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(256), 0, s, in, n, out, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, tmp, (int)ntiles, out + n);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, tmp);
}

// ---------------------------------------------------------------------------------------
// Work plan: a list with g probing queries (in probe ranks [r0, r1)) and n vectors (those of
// [vmin, vmax)) becomes ceil(g / qtile) x ceil(n / vsplit) items (list, query tile, vector range), so
// no item is longer than vsplit vectors. Lists, taken in decreasing size order, are dealt
// round-robin to 8 per-XCD queues, so the tiles of one list run on one XCD and share its L2,
// while every XCD gets a similar mix of list sizes. The queues lie back to back in `items`.
// queue[0..8) = lengths, [8..16) = heads, [16..24) = bases, [24 + r] = first item of the list
// of rank r. k_plan_count (one workgroup of 1024 threads, thread t owns ranks 8t..8t+7, one
// per queue) sizes the queues; k_plan_write (one workgroup per list) fills them. Several phases
// (rank and vector ranges, each with its own queues and items) are planned in one launch of
// each: blockIdx.y is the phase.
struct PlanArgs {
    const int* bucket_off;
    const int* rank_order;
    const int64_t* list_off;
    int nlist, nprobe, r0, r1, vmin, vmax, qtile, vsplit;
};
struct PlanBatch {
    const int* bucket_off;
    const int* rank_order;
    const int64_t* list_off;
    int nlist, nprobe, qtile, vsplit, nph;
    PlanSpec ph[kMaxPlanPhases];
    __device__ PlanArgs args(int i) const {
        return PlanArgs{bucket_off, rank_order, list_off, nlist, nprobe, ph[i].r0, ph[i].r1, ph[i].vmin,
                        ph[i].vmax, ph[i].qtile > 0 ? ph[i].qtile : qtile, ph[i].vsplit > 0 ? ph[i].vsplit : vsplit};
    }
};

__device__ __forceinline__ void plan_list(const PlanArgs& a, int r, int& l, int& gb, int& ge, int& nlv, int& nqt,
                                          int& nvr) {
    l = a.rank_order[r];
    gb = a.bucket_off[(int64_t)l * a.nprobe + a.r0];
    ge = a.bucket_off[(int64_t)l * a.nprobe + a.r1];
    nlv = max(0, (int)min(a.list_off[l + 1] - a.list_off[l], (long long)a.vmax) - a.vmin);
    nqt = (ge - gb + a.qtile - 1) / a.qtile;
    nvr = (nlv + a.vsplit - 1) / a.vsplit;
}

__global__ __launch_bounds__(1024) void k_plan_count(PlanBatch pb) {
    const PlanArgs a = pb.args(blockIdx.y);
    int* __restrict__ queue = pb.ph[blockIdx.y].queue;
    unsigned long long* __restrict__ stats = pb.ph[blockIdx.y].stats;
    __shared__ int s_wsum[16][kNumXcd];  // each wavefront's total per queue
    __shared__ int s_base[kNumXcd];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int cnt[kNumXcd], incl[kNumXcd];
    unsigned long long cand = 0;
#pragma unroll
    for (int x = 0; x < kNumXcd; ++x) {
        const int r = t * kNumXcd + x;
        cnt[x] = 0;
        if (r < a.nlist) {
            int l, gb, ge, nlv, nqt, nvr;
            plan_list(a, r, l, gb, ge, nlv, nqt, nvr);
            cnt[x] = nqt * nvr;
            cand += (unsigned long long)(ge - gb) * (unsigned long long)nlv;
        }
        // inclusive scan over the threads, per queue: within the wavefront by DPP moves, then
        // the totals of the wavefronts before
        incl[x] = (int)wave_incl_scan((unsigned int)cnt[x]);
        if (lane == 63) s_wsum[wv][x] = incl[x];
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < kNumXcd; ++x) {
        int before = 0;
        for (int v = 0; v < wv; ++v) before += s_wsum[v][x];
        incl[x] += before;
    }
    if (t == 0) {
        int b = 0;
        for (int x = 0; x < kNumXcd; ++x) {
            int tot = 0;
            for (int v = 0; v < 16; ++v) tot += s_wsum[v][x];
            s_base[x] = b;
            queue[x] = tot;
            queue[kNumXcd + x] = 0;
            queue[2 * kNumXcd + x] = b;
            b += tot;
        }
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < kNumXcd; ++x) {
        const int r = t * kNumXcd + x;
        if (r < a.nlist) queue[3 * kNumXcd + r] = s_base[x] + incl[x] - cnt[x];
    }
    // candidates = sum over pairs of the probed list's (capped) size
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_xor(cand, off, 64);
    if ((t & 63) == 0 && cand) atomicAdd(&stats[0], cand);
}

__global__ __launch_bounds__(64) void k_plan_write(PlanBatch pb, int64_t qcap) {
    const PlanArgs a = pb.args(blockIdx.y);
    const int* __restrict__ queue = pb.ph[blockIdx.y].queue;
    ScanItem* __restrict__ items = pb.ph[blockIdx.y].items;
    const int r = blockIdx.x;
    int l, gb, ge, nlv, nqt, nvr;
    plan_list(a, r, l, gb, ge, nlv, nqt, nvr);
    const int64_t pos0 = queue[3 * kNumXcd + r];
    const int total = nqt * nvr;
    for (int j = threadIdx.x; j < total; j += blockDim.x) {
        const int c = j / nvr, v = j % nvr;
        ScanItem it;
        it.list = l;
        it.pair_begin = gb + c * a.qtile;
        it.pair_end = min(ge, it.pair_begin + a.qtile);
        it.vec_begin = a.vmin + v * a.vsplit;
        if (pos0 + j < qcap) items[pos0 + j] = it;
    }
}

void launch_plan_phases(const int* bucket_off, int nlist, int nprobe, int qtile, int vsplit,
                        const int* list_rank_order, const int64_t* list_off, int64_t qcap, const PlanSpec* phases,
                        int nph, hipStream_t s) {
    if (nph <= 0) return;
    PlanBatch pb{bucket_off, list_rank_order, list_off, nlist, nprobe, qtile, vsplit, nph, {}};
    for (int i = 0; i < nph; ++i) pb.ph[i] = phases[i];
    hipLaunchKernelGGL(k_plan_count, dim3(1, (unsigned)nph), dim3(1024), 0, s, pb);
    hipLaunchKernelGGL(k_plan_write, dim3((unsigned)nlist, (unsigned)nph), dim3(64), 0, s, pb, qcap);
}

void launch_plan_items(const int* bucket_off, int nlist, int nprobe, int r0, int r1, int vmin, int vmax, int qtile,
                       int vsplit, const int* list_rank_order, const int64_t* list_off, int64_t qcap, ScanItem* items,
                       int* queue, unsigned long long* stats, hipStream_t s) {
    const PlanSpec p{r0, r1, vmin, vmax, queue, items, stats, 0, 0};
    launch_plan_phases(bucket_off, nlist, nprobe, qtile, vsplit, list_rank_order, list_off, qcap, &p, 1, s);
}

// ---------------------------------------------------------------------------------------
// The list scan. A persistent grid; each workgroup pulls (list, query tile, vector range)
// items from its XCD's queue (then from the others'). Per item, the range is swept in tiles of
// VT vectors; per tile, the QT queries x VT vectors distances are accumulated over the staging
// chunks of the dimension blocks. Each thread owns a QPT query x VPT vector register tile, so
// one or two 16-byte LDS reads of queries and vectors feed QPT*VPT subtract + FMA pairs.
//
// The query tile is staged once per item (all dimensions, when it fits in LDS: QRES) or per
// chunk. Work is a stream of (tile, chunk) jobs; the global loads of the predicted next job are
// issued before the current one is computed, so their latency hides behind the FMAs. At a
// pruning point the prediction is the next tile's first chunk when the previous tile died at
// that point (far lists: nearly every tile), else the current tile's next chunk.
//
// After each block but the last, pairs whose partial sum exceeds the query's bound are
// dropped; inside a block a wavefront re-checks fl(dis + partial) every 8 dimensions (partial
// sums only grow, so such a pair ends above its bound). A wavefront whose pairs are all dropped
// stops computing; the workgroup leaves the tile once every wavefront has. Surviving pairs with
// distance <= bound (or every pair, with accept_all) are appended to their query's candidate
// buffer as (distance, id) keys.
template <int QPT, int VPT, int TQ, bool QRES>
__global__ __launch_bounds__(256) void k_list_scan(const float* __restrict__ Qp, int d,
                                                   const float* __restrict__ codes,
                                                   const int64_t* __restrict__ code_off,
                                                   const int64_t* __restrict__ list_off,
                                                   const int* __restrict__ list_stride,
                                                   const uint32_t* __restrict__ ids, ChunkPlan plan, int prune,
                                                   const float* __restrict__ bound, const int* __restrict__ pairs,
                                                   const ScanItem* __restrict__ items,
                                                   int* __restrict__ queue, int vmax, int vsplit, int accept_all,
                                                   unsigned long long* __restrict__ surv,
                                                   unsigned int* __restrict__ surv_cnt, int cap,
                                                   unsigned long long* __restrict__ stats) {
    constexpr int TV = 256 / TQ, QT = TQ * QPT, VT = TV * VPT, KC = kChunkDims, QS = QT + 4;
    constexpr int QLD = QRES ? 1 : (QT * KC + 255) / 256;  // query values staged per thread per chunk
    constexpr int XLD = (KC * VT / 4 + 255) / 256;         // float4 vector loads per thread per chunk
    static_assert(QPT == 2 || QPT == 4 || QPT == 8, "QPT");
    static_assert(VPT == 4 || VPT == 8, "VPT");
    static_assert(QPT * VPT <= 32, "pair mask is 32 bits");
    extern __shared__ __align__(16) float s_q[];  // [d][QS] (QRES) or [KC][QS]
    __shared__ __align__(16) float Xs[KC][VT];
    __shared__ float s_bound[QT];
    __shared__ int s_qid[QT];
    __shared__ int s_item;
    float(*Qs)[QS] = reinterpret_cast<float(*)[QS]>(s_q);
    const int tid = threadIdx.x;
    const int tq = tid / TV, tv = tid % TV;
    const int xcd = blockIdx.x % kNumXcd;
    unsigned long long evals = 0;
    const int* __restrict__ queue_len = queue;
    int* __restrict__ queue_head = queue + kNumXcd;
    const int* __restrict__ queue_base = queue + 2 * kNumXcd;
    bool predict_die = true;

    for (;;) {
        if (tid == 0) {
            // own XCD's queue first, then the others'. A head only grows, so a plain load that
            // already shows a queue exhausted (a stale value is only ever smaller) skips the
            // atomic: the read-modify-writes on a head are device-scope and serialize across the
            // XCDs, and without the check every workgroup would issue one on every queue as the
            // work runs out.
            int it = -1;
            for (int o = 0; o < kNumXcd && it < 0; ++o) {
                const int x = (xcd + o) % kNumXcd;
                const int len = queue_len[x];
                if (__hip_atomic_load(&queue_head[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
                const int slot = atomicAdd(&queue_head[x], 1);
                if (slot < len) it = queue_base[x] + slot;
            }
            s_item = it;
        }
        __syncthreads();
        const int it = s_item;
        if (it < 0) break;
        const ScanItem item = items[it];
        const int nQ = item.pair_end - item.pair_begin;
        if (tid < QT) {
            int qq = 0;
            float b = -1.f;  // padding rows: nothing survives
            if (tid < nQ) {
                qq = pairs[item.pair_begin + tid];
                b = bound[qq];
            }
            s_qid[tid] = qq;
            s_bound[tid] = b;
        }
        const int l = item.list;
        const int64_t lbeg = list_off[l];
        const int nl = (int)min(list_off[l + 1] - lbeg, (long long)vmax);
        const int ps = list_stride[l];
        const float* __restrict__ X = codes + code_off[l];
        __syncthreads();
        if constexpr (QRES) {
            // the whole query tile, row-coalesced reads, transposed into [dim][query]
            for (int e = tid; e < QT * d; e += 256) {
                const int j = e / d, i = e - j * d;
                Qs[i][j] = j < nQ ? Qp[(int64_t)s_qid[j] * d + i] : 0.f;
            }
        }
        float myb[QPT];
#pragma unroll
        for (int i = 0; i < QPT; ++i) myb[i] = s_bound[tq * QPT + i];

        // ---- per-tile state
        int vt = item.vec_begin;
        const int vend = min(nl, item.vec_begin + vsplit);
        uint32_t valid = 0, alive = 0;
        int nvalid = 0;
        bool wave_alive = false;
        float dis[QPT][VPT], p[QPT][VPT];
        auto begin_tile = [&]() {
            valid = 0;
#pragma unroll
            for (int i = 0; i < QPT; ++i)
#pragma unroll
                for (int j = 0; j < VPT; ++j)
                    if (tq * QPT + i < nQ && vt + tv * VPT + j < nl) valid |= 1u << (i * VPT + j);
            alive = valid;
            nvalid = __builtin_popcount(valid);
            wave_alive = __any(alive != 0);
#pragma unroll
            for (int i = 0; i < QPT; ++i)
#pragma unroll
                for (int j = 0; j < VPT; ++j) dis[i][j] = p[i][j] = 0.f;
        };
        // ---- prefetch registers and the job they hold
        float qv[QLD];
        float4 xv[XLD];
        int reg_vt = -1, reg_c = -1;
        auto load_job = [&](int jvt, int jc) {
            const int lo = plan.lo[jc], len = plan.len[jc];
            if constexpr (!QRES) {
#pragma unroll
                for (int u = 0; u < QLD; ++u) {
                    const int e = tid + 256 * u;
                    const int kk = e / QT;
                    qv[u] = (e < QT * KC && kk < len) ? Qp[(int64_t)s_qid[e % QT] * d + lo + kk] : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < XLD; ++u) {
                const int e = tid + 256 * u;
                const int kk = e / (VT / 4), v4 = e % (VT / 4);
                const int vv = jvt + v4 * 4;
                xv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (e < KC * VT / 4 && kk < len && vv < ps)
                    xv[u] = *reinterpret_cast<const float4*>(X + (int64_t)(lo + kk) * ps + vv);
            }
            reg_vt = jvt;
            reg_c = jc;
        };

        int c = 0;
        begin_tile();
        load_job(vt, 0);
        while (true) {
            // the registers hold job (vt, c): move it to LDS
            if constexpr (!QRES) {
#pragma unroll
                for (int u = 0; u < QLD; ++u) {
                    const int e = tid + 256 * u;
                    if (e < QT * KC) Qs[e / QT][e % QT] = qv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < XLD; ++u) {
                const int e = tid + 256 * u;
                if (e < KC * VT / 4) *reinterpret_cast<float4*>(&Xs[e / (VT / 4)][(e % (VT / 4)) * 4]) = xv[u];
            }
            __syncthreads();
            const bool prune_point = prune && plan.prune_after[c];
            if (prune_point && predict_die) {
                if (vt + VT < vend) load_job(vt + VT, 0);
            } else if (c + 1 < plan.n) {
                load_job(vt, c + 1);
            } else if (vt + VT < vend) {
                load_job(vt + VT, 0);
            }
            const int len = plan.len[c];
            const int qrow0 = QRES ? plan.lo[c] : 0;
            if (wave_alive) {
                auto step = [&](int kk) {
                    float qf[QPT], xf[VPT];
                    if constexpr (QPT == 2) {
                        const float2 a = *reinterpret_cast<const float2*>(&Qs[qrow0 + kk][tq * QPT]);
                        qf[0] = a.x;
                        qf[1] = a.y;
                    } else {
#pragma unroll
                        for (int h = 0; h < QPT / 4; ++h) {
                            const float4 a = *reinterpret_cast<const float4*>(&Qs[qrow0 + kk][tq * QPT + 4 * h]);
                            qf[4 * h] = a.x;
                            qf[4 * h + 1] = a.y;
                            qf[4 * h + 2] = a.z;
                            qf[4 * h + 3] = a.w;
                        }
                    }
#pragma unroll
                    for (int h = 0; h < VPT / 4; ++h) {
                        const float4 b = *reinterpret_cast<const float4*>(&Xs[kk][tv * VPT + 4 * h]);
                        xf[4 * h] = b.x;
                        xf[4 * h + 1] = b.y;
                        xf[4 * h + 2] = b.z;
                        xf[4 * h + 3] = b.w;
                    }
#pragma unroll
                    for (int i = 0; i < QPT; ++i)
#pragma unroll
                        for (int j = 0; j < VPT; ++j) {
                            const float t = qf[i] - xf[j];
                            p[i][j] = __builtin_fmaf(t, t, p[i][j]);
                        }
                };
                for (int s0 = 0; s0 < len; s0 += 8) {
                    const int s1 = min(s0 + 8, len);
#pragma nounroll
                    for (int kk = s0; kk < s1; ++kk) step(kk);
                    if (prune && s1 < len) {
#pragma unroll
                        for (int i = 0; i < QPT; ++i)
#pragma unroll
                            for (int j = 0; j < VPT; ++j)
                                if (dis[i][j] + p[i][j] > myb[i]) alive &= ~(1u << (i * VPT + j));
                        if (!__any(alive != 0)) break;
                    }
                }
                wave_alive = __any(alive != 0);
            }
            __syncthreads();
            bool died = false;
            if (plan.ends_block[c]) {
#pragma unroll
                for (int i = 0; i < QPT; ++i)
#pragma unroll
                    for (int j = 0; j < VPT; ++j) {
                        dis[i][j] += p[i][j];
                        p[i][j] = 0.f;
                    }
                if (wave_alive) evals += (unsigned long long)nvalid;
                if (prune_point) {
#pragma unroll
                    for (int i = 0; i < QPT; ++i)
#pragma unroll
                        for (int j = 0; j < VPT; ++j)
                            if (dis[i][j] > myb[i]) alive &= ~(1u << (i * VPT + j));
                    wave_alive = __any(alive != 0);
                    died = !__syncthreads_or(alive != 0);
                    predict_die = died;
                }
            }
            const bool tile_done = died || c + 1 == plan.n;
            if (tile_done && !died && __any(alive != 0)) {
                // Append the survivors. The TV consecutive lanes sharing a query row reserve
                // their slots with one atomic per row: a prefix sum of the per-lane counts over
                // the group, the last lane adds the group total to the query's counter.
#pragma unroll
                for (int i = 0; i < QPT; ++i) {
                    uint32_t m = 0;
#pragma unroll
                    for (int j = 0; j < VPT; ++j)
                        if (((alive >> (i * VPT + j)) & 1u) && (accept_all || dis[i][j] <= myb[i])) m |= 1u << j;
                    const int cnt = __builtin_popcount(m);
                    int incl = cnt;
#pragma unroll
                    for (int off = 1; off < TV; off <<= 1) {
                        const int y = __shfl_up(incl, off, TV);
                        if (tv >= off) incl += y;
                    }
                    const int total = __shfl(incl, TV - 1, TV);
                    const int q = s_qid[tq * QPT + i];
                    unsigned int base = 0;
                    if (tv == TV - 1 && total > 0) base = atomicAdd(&surv_cnt[q], (unsigned int)total);
                    base = (unsigned int)__shfl((int)base, TV - 1, TV);
                    unsigned int slot = base + (unsigned int)(incl - cnt);
#pragma unroll
                    for (int j = 0; j < VPT; ++j)
                        if ((m >> j) & 1u) {
                            if (slot < (unsigned int)cap)
                                surv[(int64_t)q * cap + slot] = hv_key(dis[i][j], ids[lbeg + vt + tv * VPT + j]);
                            ++slot;
                        }
                }
            }
            const int nvt = tile_done ? vt + VT : vt;
            const int nc = tile_done ? 0 : c + 1;
            if (nvt >= vend) break;
            if (reg_vt != nvt || reg_c != nc) load_job(nvt, nc);
            if (tile_done) {
                vt = nvt;
                begin_tile();
            }
            c = nc;
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) evals += __shfl_xor(evals, off, 64);
    if ((tid & 63) == 0 && evals) atomicAdd(&stats[1], evals);
}

// Compiled tile shapes: {queries per thread, vectors per thread, threads along queries}.
struct ShapeInfo {
    int qpt, vpt, tq;
};
static const ShapeInfo kShapes[kNumShapes] = {
    {4, 8, 16},  // 0: 64 queries x 128 vectors
    {4, 4, 16},  // 1: 64 x 64
    {2, 8, 16},  // 2: 32 x 128
    {2, 4, 16},  // 3: 32 x 64
    {4, 4, 4},   // 4: 16 x 256
    {2, 4, 8},   // 5: 16 x 128
    {8, 4, 8},   // 6: 64 x 128 (8 x 4 per thread)
};

int scan_query_tile(int shape) { return kShapes[shape].qpt * kShapes[shape].tq; }
int scan_vector_tile(int shape) { return kShapes[shape].vpt * (256 / kShapes[shape].tq); }

// The query tile stays in LDS for the whole item when it takes at most this many bytes.
constexpr int kQresBytes = 40 * 1024;
static bool use_qres(int shape, int d) { return (size_t)d * (scan_query_tile(shape) + 4) * 4 <= kQresBytes; }
static size_t qlds_bytes(int shape, int d) {
    return (size_t)(use_qres(shape, d) ? d : kChunkDims) * (scan_query_tile(shape) + 4) * sizeof(float);
}

template <int QPT, int VPT, int TQ, bool QRES>
static int grid_for(int device, size_t lds) {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_list_scan<QPT, VPT, TQ, QRES>, 256, lds) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 2;
    return ncu * per_cu;
}

#define HV_SHAPES(X)      \
    X(0, 4, 8, 16)        \
    X(1, 4, 4, 16)        \
    X(2, 2, 8, 16)        \
    X(3, 2, 4, 16)        \
    X(4, 4, 4, 4)         \
    X(5, 2, 4, 8)         \
    X(6, 8, 4, 8)

int scan_grid_size(int device, int shape, int d) {
    const size_t lds = qlds_bytes(shape, d);
    const bool qres = use_qres(shape, d);
#define HV_GRID_CASE(ID, A, B, C) \
    case ID: return qres ? grid_for<A, B, C, true>(device, lds) : grid_for<A, B, C, false>(device, lds);
    switch (shape) {
        HV_SHAPES(HV_GRID_CASE)
        default: return 512;
    }
#undef HV_GRID_CASE
}

void launch_list_scan(int shape, const float* Qp, int d, const float* codes, const int64_t* code_off,
                      const int64_t* list_off, const int* list_stride, const uint32_t* ids, const ChunkPlan& plan,
                      int prune, const float* bound, const int* pairs, const ScanItem* items, int* queue, int vmax,
                      int vsplit, int accept_all, unsigned long long* surv, unsigned int* surv_cnt, int cap,
                      unsigned long long* stats, int grid, hipStream_t s) {
    const size_t lds = qlds_bytes(shape, d);
    const bool qres = use_qres(shape, d);
#define HV_SCAN_LAUNCH(A, B, C, R)                                                                             \
    hipLaunchKernelGGL((k_list_scan<A, B, C, R>), dim3((unsigned)grid), dim3(256), lds, s, Qp, d, codes, code_off, \
                       list_off, list_stride, ids, plan, prune, bound, pairs, items, queue, vmax, vsplit,           \
                       accept_all, surv, surv_cnt, cap, stats)
#define HV_SCAN_CASE(ID, A, B, C)                \
    case ID:                                     \
        if (qres) HV_SCAN_LAUNCH(A, B, C, true); \
        else HV_SCAN_LAUNCH(A, B, C, false);     \
        break;
    switch (shape) {
        HV_SHAPES(HV_SCAN_CASE)
        default: break;
    }
#undef HV_SCAN_CASE
#undef HV_SCAN_LAUNCH
}

// ---------------------------------------------------------------------------------------
// Exact path for the queries the search redoes (overflowed candidate buffers, learned-mode
// fallbacks). k_exact_lists: one workgroup per (query, probed list, segment of kExactSpan
// vectors), one candidate per thread; candidates that pass the bound are collected in LDS (one
// LDS atomic per wavefront), the k-th smallest key of the collected set is found by a block-wide
// radix select over the bytes of the 64-bit keys, which are unique, and only the keys up to it
// are kept. It appends its (at most k) best keys to the query's run buffer (an atomic count per
// query, order free). k_exact_merge: one workgroup per query reduces the appended keys to the
// final top-k the same way (kExactSeg keys collected between reductions) and sorts them.
// Same arithmetic, pruning and (distance <= bound) filter as the scan, so the result is the same.
constexpr int kExactTop = 1024, kExactSeg = 4096, kExactSpan = 256;

// Block-wide reduction of the kept keys s[0, s_top) and the collected keys s_col[0, s_cnt) to the
// k smallest (all of them when there are at most k): s_top, s_kth (the k-th key once there are
// k) and s_cnt = 0 on return.
__device__ void topk_reduce(unsigned long long* s, const unsigned long long* s_col, unsigned long long* s_out, int k,
                            int& s_top, int& s_cnt, unsigned long long& s_kth, unsigned int* s_hist,
                            unsigned long long* s_sel) {
    const int top = s_top, m = s_cnt;
    if (top + m > k) {
        // the k-th smallest of the kept and the collected keys; keep the k keys <= it
        const unsigned long long kv = block_radix_kth(s, top, s_col, m, k, s_hist, s_sel);
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        for (int i0 = 0; i0 < top + m; i0 += blockDim.x) {
            const int i = i0 + threadIdx.x;
            const unsigned long long key = i < top ? s[i] : (i < top + m ? s_col[i - top] : ~0ull);
            wave_append(i < top + m && key <= kv, key, s_out, &s_cnt);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < k; i += blockDim.x) s[i] = s_out[i];
        __syncthreads();
        if (threadIdx.x == 0) {
            s_top = k;
            s_kth = kv;
            s_cnt = 0;
        }
    } else {
        for (int i = threadIdx.x; i < m; i += blockDim.x) s[top + i] = s_col[i];
        __syncthreads();
        if (threadIdx.x == 0) {
            s_top = top + m;
            s_cnt = 0;
        }
    }
    __syncthreads();
}

// IP (inner-product metric): keys are hv_ip_key of the FP32 product chain (no bound filter,
// no dimension blocks); with simi given, a query's list j is scanned only over the vectors whose
// candidate2centroid (c2c, descending within the list) lies in the angular band that a result
// better than simi[q] needs (Index.cpp:1515-1546), widened by `slack`:
//   c = <q, centroid> (pval), lo = simi c - sqrt(1-simi^2) sqrt(1-c^2), hi = simi c + ... (no hi
//   bound when simi < c); NaN (|q| > 1, or no simi yet) leaves the list whole.
template <bool IP>
__global__ __launch_bounds__(256) void k_exact_lists(const int* __restrict__ qlist,
                                                     const int* __restrict__ nql_dev, int q0, int nql_max,
                                                     const float* __restrict__ Qp, int d,
                                                     const float* __restrict__ codes,
                                                     const int64_t* __restrict__ code_off,
                                                     const int64_t* __restrict__ list_off,
                                                     const int* __restrict__ list_stride,
                                                     const uint32_t* __restrict__ ids, const int* __restrict__ probes,
                                                     int nprobe, int r0, int nranks, const int* __restrict__ qn,
                                                     int nseg, ChunkPlan plan, int prune,
                                                     const float* __restrict__ bound, int k,
                                                     unsigned long long* __restrict__ runs, int64_t run_stride,
                                                     int* __restrict__ run_cnt, const float* __restrict__ c2c,
                                                     const float* __restrict__ pval, const float* __restrict__ simi,
                                                     float slack, unsigned long long* __restrict__ scanned) {
    constexpr unsigned long long kPad = IP ? HV_IP_KEY_PAD : HV_KEY_PAD;
    extern __shared__ __align__(16) unsigned char s_raw[];
    // kept keys [k] | keys collected from the segment [kExactSpan] | reduction output [k]: a segment
    // holds at most kExactSpan vectors, one per thread, so it is collected in one round
    unsigned long long* s = reinterpret_cast<unsigned long long*>(s_raw);
    unsigned long long* s_col = s + k;
    unsigned long long* s_out = s_col + kExactSpan;
    float* qs = reinterpret_cast<float*>(s_raw + sizeof(unsigned long long) * (2 * k + kExactSpan));
    __shared__ unsigned int s_hist[256];
    __shared__ unsigned long long s_sel;
    __shared__ int s_cnt, s_top, s_base;
    __shared__ unsigned long long s_kth;
    const int64_t per_q = (int64_t)(nranks - r0) * nseg;  // ranks [r0, nranks) of rows nprobe wide
    const int nql = min(*nql_dev - q0, nql_max);
    for (int64_t w = blockIdx.x; w < (int64_t)nql * per_q; w += gridDim.x) {
        const int64_t q = qlist[q0 + w / per_q];
        const int j = r0 + (int)(w % per_q) / nseg, seg = (int)(w % nseg);
        const float b = bound ? bound[q] : 3.402823466e38f;
        {
            // an empty segment (rank past the query's probe count, or past the list's end)
            // contributes an empty run
            const int l0 = probes[q * nprobe + j];
            const int64_t len = list_off[l0 + 1] - list_off[l0];
            if ((qn && j >= qn[q]) || (int64_t)seg * kExactSpan >= len) continue;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < d; i += blockDim.x) qs[i] = Qp[q * d + i];
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_top = 0;
            s_kth = kPad;
        }
        __syncthreads();
        const int l = probes[q * nprobe + j];
        const int64_t lb = list_off[l];
        const int len = (int)(list_off[l + 1] - lb);
        int n = min(len, (seg + 1) * kExactSpan), start = seg * kExactSpan;
        if (IP && simi) {
            const float sv = simi[q];
            if (sv > -3.402823466e38f) {
                const float c = pval[q * nprobe + j];
                const float sc = sqrtf(1.f - c * c), ss = sqrtf(1.f - sv * sv);
                const float lo = sv * c - ss * sc - slack;
                const float hi = sv < c ? __builtin_inff() : sv * c + ss * sc + slack;
                if (lo == lo && hi == hi) {  // not NaN
                    // c2c descends within the list: [begin, end) = { t : lo <= c2c[t] <= hi }
                    int a = 0, z = len;
                    while (a < z) {
                        const int m = (a + z) >> 1;
                        if (c2c[lb + m] > hi) a = m + 1;
                        else z = m;
                    }
                    const int begin = a;
                    z = len;
                    while (a < z) {
                        const int m = (a + z) >> 1;
                        if (c2c[lb + m] >= lo) a = m + 1;
                        else z = m;
                    }
                    start = max(start, begin);
                    n = min(n, a);
                }
            }
        }
        if (IP && scanned && threadIdx.x == 0 && n > start) atomicAdd(scanned, (unsigned long long)(n - start));
        const float* X = codes + code_off[l];
        const int ps = list_stride[l];
        for (int t0 = start; t0 < n; t0 += blockDim.x) {
            const int t = t0 + threadIdx.x;
            const unsigned long long kth = s_kth;
            unsigned long long key = kPad;
            bool keep = false;
            if (t < n) {
                if (IP) {
                    float acc = 0.f;
                    for (int i = 0; i < d; ++i) acc = __builtin_fmaf(qs[i], X[(int64_t)i * ps + t], acc);
                    key = hv_ip_key(acc, ids[lb + t]);
                    keep = key < kth;
                } else {
                    const float dis = blocked_l2_col(qs, X, ps, t, plan, prune != 0, b);
                    key = hv_key(dis, ids[lb + t]);
                    keep = dis <= b && key < kth;
                }
            }
            wave_append(keep, key, s_col, &s_cnt);
            __syncthreads();
            // reduce when the next round might not fit, and at the end; every thread reads the
            // count before any wavefront can append the next round's keys to it
            const bool reduce = s_cnt > kExactSpan - (int)blockDim.x || t0 + (int)blockDim.x >= n;
            __syncthreads();
            if (reduce) topk_reduce(s, s_col, s_out, k, s_top, s_cnt, s_kth, s_hist, &s_sel);
        }
        // append the kept keys (at most k) to the query's runs (run_stride keys per query)
        const int top = s_top;
        const int64_t wq = w / per_q;
        if (threadIdx.x == 0) s_base = top ? atomicAdd(&run_cnt[wq], top) : 0;
        __syncthreads();
        unsigned long long* dst = runs + wq * run_stride + s_base;
        for (int i = threadIdx.x; i < top; i += blockDim.x) dst[i] = s[i];
    }
}

template <bool IP>
__global__ __launch_bounds__(256) void k_exact_merge(const int* __restrict__ qlist, const int* __restrict__ nql_dev,
                                                     int q0, int nql_max, int64_t run_stride, int k,
                                                     const unsigned long long* __restrict__ runs,
                                                     const int* __restrict__ run_cnt, float* __restrict__ D,
                                                     int64_t* __restrict__ I) {
    constexpr unsigned long long kPad = IP ? HV_IP_KEY_PAD : HV_KEY_PAD;
    __shared__ unsigned long long s[2 * kExactTop + kExactSeg];
    unsigned long long* s_col = s + kExactTop;
    unsigned long long* s_out = s_col + kExactSeg;
    __shared__ unsigned int s_hist[256];
    __shared__ unsigned long long s_sel;
    __shared__ int s_cnt, s_top;
    __shared__ unsigned long long s_kth;
    const int nql = min(*nql_dev - q0, nql_max);
    for (int w = blockIdx.x; w < nql; w += gridDim.x) {
        const int64_t q = qlist[q0 + w];
        const unsigned long long* __restrict__ src = runs + (int64_t)w * run_stride;
        const int total = run_cnt[w];
        __syncthreads();
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_top = 0;
            s_kth = kPad;
        }
        __syncthreads();
        // kMergeKeys keys per thread per round: their loads are in flight together
        constexpr int kMergeKeys = 8;
        const int round = (int)blockDim.x * kMergeKeys;
        for (int t0 = 0; t0 < total; t0 += round) {
            const unsigned long long kth = s_kth;
            unsigned long long v[kMergeKeys];
#pragma unroll
            for (int u = 0; u < kMergeKeys; ++u) {
                const int t = t0 + u * (int)blockDim.x + (int)threadIdx.x;
                v[u] = t < total ? src[t] : kPad;
            }
#pragma unroll
            for (int u = 0; u < kMergeKeys; ++u) wave_append(v[u] < kth, v[u], s_col, &s_cnt);
            __syncthreads();
            const bool reduce = s_cnt > kExactSeg - round || t0 + round >= total;
            __syncthreads();  // the count is read by every thread before the next round appends
            if (reduce) topk_reduce(s, s_col, s_out, k, s_top, s_cnt, s_kth, s_hist, &s_sel);
        }
        const int top = s_top;
        const int p2 = hv_next_pow2(max(k, 2));
        for (int i = top + threadIdx.x; i < p2; i += blockDim.x) s[i] = kPad;
        __syncthreads();
        block_bitonic_sort(s, p2);
        for (int i = threadIdx.x; i < k; i += blockDim.x) {
            const int64_t id = hv_key_id(s[i]);
            if (IP) D[q * k + i] = id < 0 ? -3.402823466e38f : hv_ip_key_value(s[i]);  // heap_init's lowest
            else D[q * k + i] = hv_key_dist(s[i]);
            I[q * k + i] = id;
        }
    }
}

int exact_segments(int64_t max_list) { return (int)std::max<int64_t>(1, (max_list + kExactSpan - 1) / kExactSpan); }

void launch_query_exact(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* Qp, int d,
                        const float* codes, const int64_t* code_off, const int64_t* list_off, const int* list_stride,
                        const uint32_t* ids, const int* probes, int nprobe, int nranks, const int* qn, int nseg,
                        const ChunkPlan& plan, int prune, const float* bound, int k, unsigned long long* runs, float* D,
                        int64_t* I, bool counts_zeroed, hipStream_t s) {
    if (nql_max <= 0) return;
    const size_t lds = sizeof(unsigned long long) * (2 * k + kExactSpan) + sizeof(float) * ((d + 3) & ~3);
    const int64_t pairs = (int64_t)nql_max * nranks * nseg;
    const int64_t stride = (int64_t)nranks * nseg * k;
    int* run_cnt = reinterpret_cast<int*>(runs + (int64_t)nql_max * stride);
    if (!counts_zeroed) launch_fill_i32(run_cnt, nql_max, 0, s);
    hipLaunchKernelGGL(k_exact_lists<false>, dim3((unsigned)std::min<int64_t>(pairs, kExactGrid)), dim3(256), lds, s,
                       qlist, nql_dev, q0, nql_max, Qp, d, codes, code_off, list_off, list_stride, ids, probes, nprobe,
                       0, nranks, qn, nseg, plan, prune, bound, k, runs, stride, run_cnt, nullptr, nullptr, nullptr,
                       0.f, nullptr);
    hipLaunchKernelGGL(k_exact_merge<false>, dim3((unsigned)std::min(nql_max, kExactGrid)), dim3(256), 0, s, qlist,
                       nql_dev, q0, nql_max, stride, k, runs, run_cnt, D, I);
}

void launch_query_exact_ip(const int* qlist, const int* nql_dev, int q0, int nql_max, const float* Qp, int d,
                           const float* codes, const int64_t* code_off, const int64_t* list_off,
                           const int* list_stride, const uint32_t* ids, const int* probes, int nprobe, int r0,
                           int r1, int nseg, const float* c2c, const float* pval, const float* simi, float slack,
                           int k, unsigned long long* runs, bool append, float* D, int64_t* I,
                           unsigned long long* scanned, hipStream_t s) {
    if (nql_max <= 0 || r1 <= r0) return;
    const size_t lds = sizeof(unsigned long long) * (2 * k + kExactSpan) + sizeof(float) * ((d + 3) & ~3);
    const int64_t pairs = (int64_t)nql_max * (r1 - r0) * nseg;
    const int64_t stride = (int64_t)nprobe * nseg * k;  // room for every rank's runs
    int* run_cnt = reinterpret_cast<int*>(runs + (int64_t)nql_max * stride);
    if (!append) launch_fill_i32(run_cnt, nql_max, 0, s);
    ChunkPlan none{};
    hipLaunchKernelGGL(k_exact_lists<true>, dim3((unsigned)std::min<int64_t>(pairs, kExactGrid)), dim3(256), lds, s,
                       qlist, nql_dev, q0, nql_max, Qp, d, codes, code_off, list_off, list_stride, ids, probes, nprobe,
                       r0, r1, nullptr, nseg, none, 0, nullptr, k, runs, stride, run_cnt, c2c, pval, simi, slack,
                       scanned);
    hipLaunchKernelGGL(k_exact_merge<true>, dim3((unsigned)std::min(nql_max, kExactGrid)), dim3(256), 0, s, qlist,
                       nql_dev, q0, nql_max, stride, k, runs, run_cnt, D, I);
}

}  // namespace hv
