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
#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

// Blocked distance of one (query, vector) pair for the single-thread kernels below: q in LDS,
// the vector as column t of a list's [d][stride] slab. Returns +inf once pruned.
__device__ __forceinline__ float blocked_l2_col(const float* qs, const float* __restrict__ X, int stride, int t,
                                                const ChunkPlan& plan, bool prune, float bound) {
    float dis = 0.f, p = 0.f;
    for (int c = 0; c < plan.n; ++c) {
        const int lo = plan.lo[c], hi = lo + plan.len[c];
        for (int i = lo; i < hi; ++i) {
            const float t2 = qs[i] - X[(int64_t)i * stride + t];
            p = __builtin_fmaf(t2, t2, p);
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
// Warm-up bound (Index::warmUpSearch, Index.cpp:999-1051; heapTops = heap top, 1303-1305):
// the k-th smallest distance over the first min(|L|, wls) vectors of each of the first wl
// probed lists, FLT_MAX when fewer than k vectors were seen. Distances use the same blocked
// sum as the scan, so every warm-up vector passes its own query's bound. One workgroup per
// query; the distances are sorted in LDS.
__global__ __launch_bounds__(256) void k_warmup_bound(const float* __restrict__ Qp, int d,
                                                      const float* __restrict__ codes,
                                                      const int64_t* __restrict__ code_off,
                                                      const int64_t* __restrict__ list_off,
                                                      const int* __restrict__ list_stride,
                                                      const int* __restrict__ probes, int nprobe, int wl, int wls,
                                                      int k, ChunkPlan plan, float* __restrict__ bound) {
    extern __shared__ __align__(16) unsigned char s_raw[];
    float* qs = reinterpret_cast<float*>(s_raw);
    unsigned int* dbuf = reinterpret_cast<unsigned int*>(s_raw + sizeof(float) * ((d + 3) & ~3));
    const int64_t q = blockIdx.x;
    for (int i = threadIdx.x; i < d; i += blockDim.x) qs[i] = Qp[q * d + i];
    __syncthreads();
    int pos = 0;
    for (int j = 0; j < wl; ++j) {
        const int l = probes[q * nprobe + j];
        const int n = (int)(list_off[l + 1] - list_off[l]);
        const int m = min(n, wls);
        const float* X = codes + code_off[l];
        const int ps = list_stride[l];
        for (int t = threadIdx.x; t < m; t += blockDim.x)
            dbuf[pos + t] = __float_as_uint(blocked_l2_col(qs, X, ps, t, plan, false, 0.f));
        pos += m;
    }
    if (pos < k) {
        if (threadIdx.x == 0) bound[q] = 3.402823466e+38f;
        return;
    }
    const int P = hv_next_pow2(max(pos, 2));
    for (int i = pos + threadIdx.x; i < P; i += blockDim.x) dbuf[i] = 0xffffffffu;
    block_bitonic_sort(dbuf, P);
    if (threadIdx.x == 0) bound[q] = __uint_as_float(dbuf[k - 1]);
}

void launch_warmup_bound(const float* Qp, int64_t nq, int d, const float* codes, const int64_t* code_off,
                         const int64_t* list_off, const int* list_stride, const int* probes, int nprobe, int wl,
                         int wls, int k, const ChunkPlan& plan, float* bound, hipStream_t s) {
    if (nq <= 0) return;
    const int pmax = hv_next_pow2(max(wl * wls, 2));
    const size_t lds = sizeof(float) * ((d + 3) & ~3) + sizeof(unsigned int) * pmax;
    hipLaunchKernelGGL(k_warmup_bound, dim3((unsigned)nq), dim3(256), lds, s, Qp, d, codes, code_off, list_off,
                       list_stride, probes, nprobe, wl, wls, k, plan, bound);
}

// ---------------------------------------------------------------------------------------
// Grouping: (query, probe rank) pairs bucketed by (list, rank). Within a list, queries that
// rank the list first come first, so a query tile holds queries at similar distance from the
// list and prunes (or survives) together.
__global__ void k_count_pairs(const int* __restrict__ probes, int64_t total, int nprobe, int* __restrict__ cnt) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int j = (int)(e % nprobe);
    atomicAdd(&cnt[(int64_t)probes[e] * nprobe + j], 1);
}

void launch_count_pairs(const int* probes, int64_t nq, int nprobe, int* bucket_cnt, hipStream_t s) {
    const int64_t tot = nq * nprobe;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, probes, tot, nprobe,
                       bucket_cnt);
}

__global__ void k_scatter_pairs(const int* __restrict__ probes, int64_t total, int nprobe,
                                const int* __restrict__ off, int* __restrict__ fill, int* __restrict__ pairs) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int j = (int)(e % nprobe);
    const int64_t b = (int64_t)probes[e] * nprobe + j;
    const int pos = off[b] + atomicAdd(&fill[b], 1);
    pairs[pos] = (int)(e / nprobe);
}

void launch_scatter_pairs(const int* probes, int64_t nq, int nprobe, const int* bucket_off, int* fill, int* pairs,
                          hipStream_t s) {
    const int64_t tot = nq * nprobe;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_scatter_pairs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, probes, tot, nprobe,
                       bucket_off, fill, pairs);
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

void launch_exclusive_scan(const int* in, int64_t n, int* out, int* tmp, hipStream_t s) {
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    if (n > 0) hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(256), 0, s, in, n, out, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, tmp, (int)ntiles, out + n);
    if (n > 0) hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, tmp);
}

// ---------------------------------------------------------------------------------------
// Work plan: a list with g probing queries becomes ceil(g / 64) items (list, query tile).
// Lists, taken in decreasing size order, are dealt round-robin to 8 per-XCD queues, so the
// tiles of one list run on one XCD and share its L2, while every XCD gets a similar mix of
// list sizes. One workgroup of 1024 threads; thread t owns ranks 8t..8t+7 (one per queue).
__global__ __launch_bounds__(1024) void k_plan_items(const int* __restrict__ bucket_off, int nlist, int nprobe,
                                                     const int* __restrict__ rank_order,
                                                     const int64_t* __restrict__ list_off, int qcap,
                                                     ScanItem* __restrict__ items, int* __restrict__ queue_len,
                                                     int* __restrict__ queue_head,
                                                     unsigned long long* __restrict__ stats) {
    __shared__ int s_scan[kNumXcd][1024];
    const int t = threadIdx.x;
    int cnt[kNumXcd], lst[kNumXcd], gb[kNumXcd], ge[kNumXcd];
    unsigned long long cand = 0;
#pragma unroll
    for (int x = 0; x < kNumXcd; ++x) {
        const int r = t * kNumXcd + x;
        cnt[x] = 0;
        lst[x] = -1;
        gb[x] = ge[x] = 0;
        if (r < nlist) {
            const int l = rank_order[r];
            lst[x] = l;
            gb[x] = bucket_off[(int64_t)l * nprobe];
            ge[x] = bucket_off[(int64_t)(l + 1) * nprobe];
            cnt[x] = (ge[x] - gb[x] + kQueryTile - 1) / kQueryTile;
            cand += (unsigned long long)(ge[x] - gb[x]) * (unsigned long long)(list_off[l + 1] - list_off[l]);
        }
        s_scan[x][t] = cnt[x];
    }
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int a[kNumXcd];
#pragma unroll
        for (int x = 0; x < kNumXcd; ++x) a[x] = t >= off ? s_scan[x][t - off] : 0;
        __syncthreads();
#pragma unroll
        for (int x = 0; x < kNumXcd; ++x) s_scan[x][t] += a[x];
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < kNumXcd; ++x) {
        const int start = s_scan[x][t] - cnt[x];
        for (int c = 0; c < cnt[x]; ++c) {
            ScanItem it;
            it.list = lst[x];
            it.pair_begin = gb[x] + c * kQueryTile;
            it.pair_end = min(ge[x], it.pair_begin + kQueryTile);
            it.pad = 0;
            items[(int64_t)x * qcap + start + c] = it;
        }
    }
    if (t == 1023) {
#pragma unroll
        for (int x = 0; x < kNumXcd; ++x) {
            queue_len[x] = s_scan[x][1023];
            queue_head[x] = 0;
        }
    }
    // candidates = sum over pairs of the probed list's size
    for (int off = 32; off > 0; off >>= 1) cand += __shfl_xor(cand, off, 64);
    if ((t & 63) == 0 && cand) atomicAdd(&stats[0], cand);
}

void launch_plan_items(const int* bucket_off, int nlist, int nprobe, const int* list_rank_order,
                       const int64_t* list_off, int qcap, ScanItem* items, int* queue_len, int* queue_head,
                       unsigned long long* stats, hipStream_t s) {
    hipLaunchKernelGGL(k_plan_items, dim3(1), dim3(1024), 0, s, bucket_off, nlist, nprobe, list_rank_order,
                       list_off, qcap, items, queue_len, queue_head, stats);
}

// ---------------------------------------------------------------------------------------
// The list scan. A persistent grid; each workgroup pulls (list, query tile) items from its
// XCD's queue (then from the others'). Per item, the list is swept in tiles of 128 vectors;
// per tile, the 64 queries x 128 vectors distances are accumulated over the staging chunks of
// the dimension blocks. Each thread owns a 4 query x 8 vector register tile, so one 16-byte
// LDS read of queries and two of vectors feed 32 subtract + 32 FMA. After each block but the
// last, pairs whose partial sum exceeds the query's bound are dropped; a wavefront whose 16
// queries x 128 vectors are all dropped stops computing, and the workgroup leaves the tile as
// soon as all four wavefronts have. Surviving pairs with distance <= bound are appended to
// their query's candidate buffer as (distance, id) keys.
__global__ __launch_bounds__(256) void k_list_scan(const float* __restrict__ Qp, int d,
                                                   const float* __restrict__ codes,
                                                   const int64_t* __restrict__ code_off,
                                                   const int64_t* __restrict__ list_off,
                                                   const int* __restrict__ list_stride,
                                                   const uint32_t* __restrict__ ids, ChunkPlan plan, int prune,
                                                   const float* __restrict__ bound, const int* __restrict__ pairs,
                                                   const ScanItem* __restrict__ items,
                                                   const int* __restrict__ queue_len, int* __restrict__ queue_head,
                                                   int qcap, unsigned long long* __restrict__ surv,
                                                   unsigned int* __restrict__ surv_cnt, int cap,
                                                   unsigned long long* __restrict__ stats) {
    constexpr int QT = kQueryTile, VT = kVecTile, KC = kChunkDims, QS = QT + 4;
    __shared__ __align__(16) float Qs[KC][QS];
    __shared__ __align__(16) float Xs[KC][VT];
    __shared__ float s_bound[QT];
    __shared__ int s_qid[QT];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    const int tq = tid >> 4, tv = tid & 15;
    const int xcd = blockIdx.x % kNumXcd;
    unsigned long long evals = 0;

    for (;;) {
        if (tid == 0) {
            int it = -1;
            for (int o = 0; o < kNumXcd && it < 0; ++o) {
                const int x = (xcd + o) % kNumXcd;
                const int slot = atomicAdd(&queue_head[x], 1);
                if (slot < queue_len[x]) it = x * qcap + slot;
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
        const int nl = (int)(list_off[l + 1] - lbeg);
        const int ps = list_stride[l];
        const float* __restrict__ X = codes + code_off[l];
        __syncthreads();
        float myb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) myb[i] = s_bound[tq * 4 + i];

        for (int vt = 0; vt < nl; vt += VT) {
            uint32_t valid = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (tq * 4 + i < nQ && vt + tv * 8 + j < nl) valid |= 1u << (i * 8 + j);
            uint32_t alive = valid;
            const int nvalid = __builtin_popcount(valid);
            bool wave_alive = __any(alive != 0);
            float dis[4][8], p[4][8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) dis[i][j] = p[i][j] = 0.f;

            for (int c = 0; c < plan.n; ++c) {
                const int lo = plan.lo[c], len = plan.len[c];
                for (int e = tid; e < QT * KC; e += 256) {
                    const int j = e & (QT - 1), kk = e / QT;
                    Qs[kk][j] = kk < len ? Qp[(int64_t)s_qid[j] * d + lo + kk] : 0.f;
                }
                for (int e = tid; e < KC * (VT / 4); e += 256) {
                    const int kk = e / (VT / 4), v4 = e % (VT / 4);
                    const int vv = vt + v4 * 4;
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (kk < len && vv < ps) v = *reinterpret_cast<const float4*>(X + (int64_t)(lo + kk) * ps + vv);
                    *reinterpret_cast<float4*>(&Xs[kk][v4 * 4]) = v;
                }
                __syncthreads();
                if (wave_alive) {
                    for (int kk = 0; kk < len; ++kk) {
                        const float4 qa = *reinterpret_cast<const float4*>(&Qs[kk][tq * 4]);
                        const float4 x0 = *reinterpret_cast<const float4*>(&Xs[kk][tv * 8]);
                        const float4 x1 = *reinterpret_cast<const float4*>(&Xs[kk][tv * 8 + 4]);
                        const float qv[4] = {qa.x, qa.y, qa.z, qa.w};
                        const float xv[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                const float t = qv[i] - xv[j];
                                p[i][j] = __builtin_fmaf(t, t, p[i][j]);
                            }
                    }
                }
                __syncthreads();
                if (plan.ends_block[c]) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            dis[i][j] += p[i][j];
                            p[i][j] = 0.f;
                        }
                    if (wave_alive) evals += (unsigned long long)nvalid;
                    if (prune && plan.prune_after[c]) {
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 8; ++j)
                                if (dis[i][j] > myb[i]) alive &= ~(1u << (i * 8 + j));
                        wave_alive = __any(alive != 0);
                        if (!__syncthreads_or(alive != 0)) break;
                    }
                }
            }
            if (alive) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (((alive >> (i * 8 + j)) & 1u) && dis[i][j] <= myb[i]) {
                            const int q = s_qid[tq * 4 + i];
                            const unsigned int slot = atomicAdd(&surv_cnt[q], 1u);
                            if (slot < (unsigned int)cap)
                                surv[(int64_t)q * cap + slot] = hv_key(dis[i][j], ids[lbeg + vt + tv * 8 + j]);
                        }
                    }
            }
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) evals += __shfl_xor(evals, off, 64);
    if ((tid & 63) == 0 && evals) atomicAdd(&stats[1], evals);
}

int scan_grid_size(int device) {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu <= 0)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_list_scan, 256, 0) != hipSuccess || per_cu <= 0)
        per_cu = 2;
    return ncu * per_cu;
}

void launch_list_scan(const float* Qp, int d, const float* codes, const int64_t* code_off, const int64_t* list_off,
                      const int* list_stride, const uint32_t* ids, const ChunkPlan& plan, int prune,
                      const float* bound, const int* pairs, const ScanItem* items, const int* queue_len,
                      int* queue_head, int qcap, unsigned long long* surv, unsigned int* surv_cnt, int cap,
                      unsigned long long* stats, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_list_scan, dim3((unsigned)grid), dim3(256), 0, s, Qp, d, codes, code_off, list_off,
                       list_stride, ids, plan, prune, bound, pairs, items, queue_len, queue_head, qcap, surv,
                       surv_cnt, cap, stats);
}

// ---------------------------------------------------------------------------------------
// Per-query top-k of the surviving candidates: sort the (distance, id) keys in LDS, write
// the first k, pad with (FLT_MAX, -1) (sort_result, heap.hpp:84-90, on (distance, id)).
// Queries whose buffer overflowed are listed for the exact per-query kernel instead.
__global__ __launch_bounds__(256) void k_select_survivors(const unsigned long long* __restrict__ surv,
                                                          const unsigned int* __restrict__ surv_cnt, int cap, int k,
                                                          float* __restrict__ D, int64_t* __restrict__ I,
                                                          int* __restrict__ overflow_list,
                                                          int* __restrict__ overflow_cnt,
                                                          unsigned long long* __restrict__ stats) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t q = blockIdx.x;
    const unsigned int n = surv_cnt[q];
    if (n > (unsigned int)cap) {
        if (threadIdx.x == 0) {
            const int slot = atomicAdd(overflow_cnt, 1);
            overflow_list[slot] = (int)q;
        }
        return;
    }
    const int P = hv_next_pow2(max(max((int)n, k), 2));
    for (int i = threadIdx.x; i < P; i += blockDim.x) s_keys[i] = i < (int)n ? surv[q * cap + i] : HV_KEY_PAD;
    block_bitonic_sort(s_keys, P);
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        const unsigned long long key = s_keys[i];
        D[q * k + i] = hv_key_dist(key);
        I[q * k + i] = hv_key_id(key);
    }
    if (threadIdx.x == 0 && n) atomicAdd(&stats[2], (unsigned long long)n);
}

void launch_select_survivors(const unsigned long long* surv, const unsigned int* surv_cnt, int64_t nq, int cap, int k,
                             float* D, int64_t* I, int* overflow_list, int* overflow_cnt, unsigned long long* stats,
                             hipStream_t s) {
    if (nq <= 0) return;
    const size_t lds = sizeof(unsigned long long) * hv_next_pow2(max(max(cap, k), 2));
    hipLaunchKernelGGL(k_select_survivors, dim3((unsigned)nq), dim3(256), lds, s, surv, surv_cnt, cap, k, D, I,
                       overflow_list, overflow_cnt, stats);
}

// ---------------------------------------------------------------------------------------
// Exact per-query path for queries whose candidate buffer overflowed: one workgroup walks the
// query's probed lists, one candidate per thread, keeping a sorted top-k in LDS and merging a
// 1024-entry buffer of candidates that beat the current k-th into it by bitonic sort. Same
// arithmetic, pruning and (distance <= bound) filter as the scan, so the result is the same.
__global__ __launch_bounds__(256) void k_query_exact(const int* __restrict__ qlist, const float* __restrict__ Qp,
                                                     int d, const float* __restrict__ codes,
                                                     const int64_t* __restrict__ code_off,
                                                     const int64_t* __restrict__ list_off,
                                                     const int* __restrict__ list_stride,
                                                     const uint32_t* __restrict__ ids, const int* __restrict__ probes,
                                                     int nprobe, ChunkPlan plan, int prune,
                                                     const float* __restrict__ bound, int k, float* __restrict__ D,
                                                     int64_t* __restrict__ I) {
    constexpr int TOP = 1024, BUF = 1024;
    extern __shared__ __align__(16) unsigned char s_raw[];
    unsigned long long* s = reinterpret_cast<unsigned long long*>(s_raw);  // [TOP + BUF]
    float* qs = reinterpret_cast<float*>(s_raw + sizeof(unsigned long long) * (TOP + BUF));
    __shared__ int s_cnt;
    __shared__ unsigned long long s_kth;
    const int64_t q = qlist[blockIdx.x];
    const float b = bound[q];
    for (int i = threadIdx.x; i < d; i += blockDim.x) qs[i] = Qp[q * d + i];
    for (int i = threadIdx.x; i < TOP + BUF; i += blockDim.x) s[i] = HV_KEY_PAD;
    if (threadIdx.x == 0) {
        s_cnt = 0;
        s_kth = HV_KEY_PAD;
    }
    __syncthreads();
    for (int j = 0; j < nprobe; ++j) {
        const int l = probes[q * nprobe + j];
        const int64_t lb = list_off[l];
        const int n = (int)(list_off[l + 1] - lb);
        const float* X = codes + code_off[l];
        const int ps = list_stride[l];
        for (int t0 = 0; t0 < n; t0 += blockDim.x) {
            const int t = t0 + threadIdx.x;
            if (t < n) {
                const float dis = blocked_l2_col(qs, X, ps, t, plan, prune != 0, b);
                if (dis <= b) {
                    const unsigned long long key = hv_key(dis, ids[lb + t]);
                    if (key < s_kth) {
                        const int pos = atomicAdd(&s_cnt, 1);
                        s[TOP + pos] = key;
                    }
                }
            }
            __syncthreads();
            if (s_cnt > BUF - (int)blockDim.x) {
                block_bitonic_sort(s, TOP + BUF);
                for (int i = k + threadIdx.x; i < TOP + BUF; i += blockDim.x) s[i] = HV_KEY_PAD;
                __syncthreads();
                if (threadIdx.x == 0) {
                    s_cnt = 0;
                    s_kth = s[k - 1];
                }
                __syncthreads();
            }
        }
    }
    block_bitonic_sort(s, TOP + BUF);
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        D[q * k + i] = hv_key_dist(s[i]);
        I[q * k + i] = hv_key_id(s[i]);
    }
}

void launch_query_exact(const int* qlist, int nql, const float* Qp, int d, const float* codes,
                        const int64_t* code_off, const int64_t* list_off, const int* list_stride,
                        const uint32_t* ids, const int* probes, int nprobe, const ChunkPlan& plan, int prune,
                        const float* bound, int k, float* D, int64_t* I, hipStream_t s) {
    if (nql <= 0) return;
    const size_t lds = sizeof(unsigned long long) * 2048 + sizeof(float) * ((d + 3) & ~3);
    hipLaunchKernelGGL(k_query_exact, dim3((unsigned)nql), dim3(256), lds, s, qlist, Qp, d, codes, code_off,
                       list_off, list_stride, ids, probes, nprobe, plan, prune, bound, k, D, I);
}

}  // namespace hv
