// Per-row / per-query selections of the IVF-flat path: the coarse nprobe select, the warm-up
// bound (k-th distance), the top-k compaction that tightens the bound, and the final top-k.
// Every one sorts 64-bit (distance, id) keys — distance bits high, id low, so the integer
// order is the (distance, id) order — and keeps a prefix. Rows of up to 2048 keys are sorted
// by one wavefront in registers (wave_bitonic_sort); larger rows by a workgroup in LDS.
#include "hv_device.h"
#include "hv_kernels.h"

namespace hv {

constexpr float kFltMax = 3.402823466e+38f;

// ---------------------------------------------------------------------------------------
// Coarse select (Index.cpp:967-997 via lite_scan_codes into an nprobe-heap + sort_result,
// IVFScan.hpp:108-138, heap.hpp:84-90): the p smallest (distance, column) of each row, in
// ascending order; the boundary tie goes to the lower column, as the reference's strict '<' in
// column order does.
__global__ __launch_bounds__(256) void k_select_rows_wave(const float* __restrict__ dist, int64_t n, int m, int p,
                                                          int* __restrict__ out_ids, float* __restrict__ out_d) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const float* __restrict__ r = dist + row * m;
    wave_sort_emit(
        m, [&](int j) { return hv_key(r[j], (uint32_t)j); },
        [&](int i, unsigned long long key) {
            if (i < p) {
                out_ids[row * p + i] = (int)(key & 0xffffffffu);
                if (out_d) out_d[row * p + i] = hv_key_dist(key);
            }
        });
}

__global__ __launch_bounds__(256) void k_select_rows_block(const float* __restrict__ dist, int m, int p,
                                                           int* __restrict__ out_ids, float* __restrict__ out_d) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t row = blockIdx.x;
    const int P = hv_next_pow2(max(m, 2));
    for (int c = threadIdx.x; c < P; c += blockDim.x)
        s_keys[c] = c < m ? hv_key(dist[row * m + c], (uint32_t)c) : HV_KEY_PAD;
    block_bitonic_sort(s_keys, P);
    for (int i = threadIdx.x; i < p; i += blockDim.x) {
        const unsigned long long k = s_keys[i];
        out_ids[row * p + i] = (int)(k & 0xffffffffu);
        if (out_d) out_d[row * p + i] = hv_key_dist(k);
    }
}

void launch_select_rows(const float* dist, int64_t n, int m, int p, int* out_ids, float* out_d, hipStream_t s) {
    if (n <= 0) return;
    if (m <= kWaveSortMax) {
        hipLaunchKernelGGL(k_select_rows_wave, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, dist, n, m, p, out_ids,
                           out_d);
    } else {
        const size_t lds = sizeof(unsigned long long) * hv_next_pow2(m);
        hipLaunchKernelGGL(k_select_rows_block, dim3((unsigned)n), dim3(256), lds, s, dist, m, p, out_ids, out_d);
    }
}

// The k-th smallest distance of the n (<= 64 E, k <= n) keys src[0, n), by one wavefront.
template <int E>
__device__ __forceinline__ float kth_dist_wave(const unsigned long long* __restrict__ src, int n, int k,
                                               unsigned int* s_tmp) {
    const int lane = threadIdx.x & 63;
    unsigned int u[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;
        u[e] = j < n ? (unsigned int)(src[j] >> 32) : 0xffffffffu;
    }
    return __uint_as_float(wave_kth_u32<E>(u, k, s_tmp));
}

// ---------------------------------------------------------------------------------------
// Warm-up bound (Index::warmUpSearch, Index.cpp:999-1051; heapTops = heap top, 1303-1305):
// the k-th smallest of the query's warm-up distances, FLT_MAX when fewer than k were seen.
// Resets the query's candidate count for the scan that follows.
__global__ __launch_bounds__(256) void k_kth_wave(const unsigned long long* __restrict__ surv,
                                                  unsigned int* __restrict__ cnt, int64_t nq, int cap, int k_all,
                                                  float* __restrict__ bound_ref, float* __restrict__ bound,
                                                  const int* __restrict__ kq) {
    __shared__ unsigned int s_tmp[4][64];
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const int k = hv_query_k(kq, q, k_all);
    const int n = (int)min(cnt[q], (unsigned int)cap);
    if (n > kWaveSortMax) return;  // k_kth_block
    const unsigned long long* __restrict__ src = surv + q * cap;
    if (n >= k) {
        // the k-th smallest distance (of the (distance, id) keys, the k-th key's distance)
        const int w = threadIdx.x >> 6;
        const float b = n <= 256   ? kth_dist_wave<4>(src, n, k, s_tmp[w])
                        : n <= 512 ? kth_dist_wave<8>(src, n, k, s_tmp[w])
                        : n <= 1024 ? kth_dist_wave<16>(src, n, k, s_tmp[w])
                                    : kth_dist_wave<32>(src, n, k, s_tmp[w]);
        if ((threadIdx.x & 63) == 0) bound_ref[q] = bound[q] = b;
    } else if ((threadIdx.x & 63) == 0) {
        bound_ref[q] = bound[q] = kFltMax;
    }
    if ((threadIdx.x & 63) == 0) cnt[q] = 0;
}

__global__ __launch_bounds__(256) void k_kth_block(const unsigned long long* __restrict__ surv,
                                                   unsigned int* __restrict__ cnt, int cap, int k_all,
                                                   float* __restrict__ bound_ref, float* __restrict__ bound,
                                                   const int* __restrict__ kq) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t q = blockIdx.x;
    const int k = hv_query_k(kq, q, k_all);
    const int n = (int)min(cnt[q], (unsigned int)cap);
    if (n <= kWaveSortMax) return;  // k_kth_wave
    const int P = hv_next_pow2(n);
    for (int i = threadIdx.x; i < P; i += blockDim.x) s_keys[i] = i < n ? surv[q * cap + i] : HV_KEY_PAD;
    block_bitonic_sort(s_keys, P);
    if (threadIdx.x == 0) {
        bound_ref[q] = bound[q] = hv_key_dist(s_keys[k - 1]);
        cnt[q] = 0;
    }
}

void launch_kth_survivors(const unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int k,
                          float* bound_ref, float* bound, const int* kq, hipStream_t s) {
    if (nq <= 0) return;
    if (cap > kWaveSortMax) {  // before the wave kernel resets the counts
        const size_t lds = sizeof(unsigned long long) * hv_next_pow2(cap);
        hipLaunchKernelGGL(k_kth_block, dim3((unsigned)nq), dim3(256), lds, s, surv, cnt, cap, k, bound_ref, bound,
                           kq);
    }
    hipLaunchKernelGGL(k_kth_wave, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, surv, cnt, nq, cap, k, bound_ref,
                       bound, kq);
}

// ---------------------------------------------------------------------------------------
// After the scan of each query's nearest list: keep only its k best keys and lower the bound to
// their k-th distance when that is tighter. The k-th of real candidates bounds the final k-th
// distance, so a tighter bound prunes the remaining lists harder without changing the result,
// and any key outside these k has k better keys and cannot be a result. A query whose buffer
// overflowed keeps its keys (it goes to the exact path) but still gets the tighter bound of
// the keys that were stored.
__global__ __launch_bounds__(256) void k_compact_wave(unsigned long long* __restrict__ surv,
                                                      unsigned int* __restrict__ cnt, int64_t nq, int cap, int k_all,
                                                      float* __restrict__ bound, const int* __restrict__ kq) {
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const int k = hv_query_k(kq, q, k_all);
    const unsigned int c = cnt[q];
    const int n = (int)min(c, (unsigned int)cap);
    if (n == 0 || n > kWaveSortMax) return;
    unsigned long long* __restrict__ buf = surv + q * cap;
    const bool keep = c <= (unsigned int)cap;
    const int m = min(n, k);
    wave_sort_emit(
        n, [&](int j) { return buf[j]; },
        [&](int i, unsigned long long key) {
            if (keep && i < m) buf[i] = key;
            if (i == k - 1 && n >= k) bound[q] = fminf(bound[q], hv_key_dist(key));
        });
    if (keep && (threadIdx.x & 63) == 0) cnt[q] = (unsigned int)m;
}

__global__ __launch_bounds__(256) void k_compact_block(unsigned long long* __restrict__ surv,
                                                       unsigned int* __restrict__ cnt, int cap, int k_all,
                                                       float* __restrict__ bound, const int* __restrict__ kq) {
    extern __shared__ unsigned long long s_keys[];
    const int64_t q = blockIdx.x;
    const int k = hv_query_k(kq, q, k_all);
    const unsigned int c = cnt[q];
    const int n = (int)min(c, (unsigned int)cap);
    if (n <= kWaveSortMax) return;  // k_compact_wave
    const int P = hv_next_pow2(n);
    for (int i = threadIdx.x; i < P; i += blockDim.x) s_keys[i] = i < n ? surv[q * cap + i] : HV_KEY_PAD;
    block_bitonic_sort(s_keys, P);
    if (c <= (unsigned int)cap) {
        const int m = min(n, k);
        for (int i = threadIdx.x; i < m; i += blockDim.x) surv[q * cap + i] = s_keys[i];
        __syncthreads();
        if (threadIdx.x == 0) cnt[q] = (unsigned int)m;
    }
    if (threadIdx.x == 0) bound[q] = fminf(bound[q], hv_key_dist(s_keys[k - 1]));
}

void launch_compact_topk(unsigned long long* surv, unsigned int* cnt, int64_t nq, int cap, int k, float* bound,
                         const int* kq, hipStream_t s) {
    if (nq <= 0) return;
    if (cap > kWaveSortMax) {  // before the wave kernel shrinks the counts
        const size_t lds = sizeof(unsigned long long) * hv_next_pow2(cap);
        hipLaunchKernelGGL(k_compact_block, dim3((unsigned)nq), dim3(256), lds, s, surv, cnt, cap, k, bound, kq);
    }
    hipLaunchKernelGGL(k_compact_wave, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, surv, cnt, nq, cap, k, bound,
                       kq);
}

// Per-query k: entries at or past each query's own k become (FLT_MAX, -1). The top k_q by
// (distance, id) is the first k_q of the top k, so a search at k cut this way is the search at k_q.
__global__ __launch_bounds__(256) void k_truncate_rows(float* __restrict__ D, int64_t* __restrict__ I, int64_t n, int k,
                                                       const int* __restrict__ kq) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t q = hv_idiv(e, k);
    if ((int)(e - q * k) >= hv_query_k(kq, q, k)) {
        D[e] = kFltMax;
        I[e] = -1;
    }
}

void launch_truncate_rows(float* D, int64_t* I, int64_t n, int k, const int* kq, hipStream_t s) {
    if (n <= 0 || !kq) return;
    hipLaunchKernelGGL(k_truncate_rows, dim3((unsigned)((n * k + 255) / 256)), dim3(256), 0, s, D, I, n, k, kq);
}

// ---------------------------------------------------------------------------------------
// Final top-k of the surviving candidates: the first k keys in (distance, id) order, padded
// with (FLT_MAX, -1) (heap_init + sort_result, heap.hpp:13-90, merge of Index.cpp:1327-1343).
// Queries whose buffer overflowed are listed for the exact per-query kernel instead (a caller
// whose counts never exceed cap may pass null lists; null stats skip the pair count).
__global__ __launch_bounds__(256) void k_select_wave(const unsigned long long* __restrict__ surv,
                                                     const unsigned int* __restrict__ surv_cnt, int64_t nq, int cap,
                                                     int k, float* __restrict__ D, int64_t* __restrict__ I,
                                                     int* __restrict__ overflow_list, int* __restrict__ overflow_cnt,
                                                     unsigned long long* __restrict__ stats) {
    __shared__ unsigned int s_tmp[4][64];
    __shared__ unsigned long long s_buf[4][kTopkBuf];
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const unsigned int n = surv_cnt[q];
    const bool lane0 = (threadIdx.x & 63) == 0;
    if (n > (unsigned int)cap) {
        if (overflow_list && lane0) overflow_list[atomicAdd(overflow_cnt, 1)] = (int)q;
        return;
    }
    if (stats && lane0 && n) atomicAdd(&stats[2], (unsigned long long)n);
    const int m = max((int)n, k);
    if (m > kWaveSortMax) return;  // k_select_block
    const unsigned long long* __restrict__ src = surv + q * cap;
    const int w = threadIdx.x >> 6;
    wave_topk_emit(
        (int)n, k, [&](int j) { return src[j]; }, s_tmp[w], s_buf[w],
        [&](int i, unsigned long long key) {
            if (i < k) {
                D[q * k + i] = hv_key_dist(key);
                I[q * k + i] = hv_key_id(key);
            }
        });
}

// The queries the wavefront select left (more keys than it sorts, or not marked done): the
// workgroups read the counts of 256 queries at a time in parallel and list the ones to do in LDS.
// A query's k smallest keys are found by a block radix select (block_radix_kth) and only those
// k are sorted (LDS bitonic), instead of sorting all of its up to cap keys.
__global__ __launch_bounds__(256) void k_select_block(const unsigned long long* __restrict__ surv,
                                                      const unsigned int* __restrict__ surv_cnt, int64_t nq, int cap,
                                                      int k, float* __restrict__ D, int64_t* __restrict__ I,
                                                      const unsigned int* __restrict__ done) {
    extern __shared__ unsigned long long s_keys[];  // [P = pow2(max(cap, k))] | [pow2(k)]
    __shared__ int s_list[256];
    __shared__ int s_n, s_cnt;
    __shared__ unsigned int s_hist[256];
    __shared__ unsigned long long s_sel;
    const int pk = hv_next_pow2(max(k, 2));
    unsigned long long* s_top = s_keys + hv_next_pow2(max(max(cap, k), 2));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = blockIdx.x; base < nq; base += stride) {
        __syncthreads();
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const int64_t q0 = base + (int64_t)threadIdx.x * gridDim.x;
        if (q0 < nq) {
            const unsigned int n = surv_cnt[q0];
            // overflow: the exact path; otherwise unless the wavefront select did it
            const bool need = n <= (unsigned int)cap && (done ? done[q0] == 0u : max((int)n, k) > kWaveSortMax);
            if (need) s_list[atomicAdd(&s_n, 1)] = (int)q0;
        }
        __syncthreads();
        const int todo = s_n;
        for (int j = 0; j < todo; ++j) {
            const int64_t q = s_list[j];
            const int n = (int)surv_cnt[q];
            for (int i = threadIdx.x; i < n; i += blockDim.x) s_keys[i] = surv[q * cap + i];
            if (threadIdx.x == 0) s_cnt = 0;
            __syncthreads();
            if (n > k) {
                // the k keys <= the k-th (keys are unique) into s_top
                const unsigned long long kv = block_radix_kth(s_keys, n, nullptr, 0, k, s_hist, &s_sel);
                for (int i0 = 0; i0 < n; i0 += blockDim.x) {
                    const int i = i0 + threadIdx.x;
                    const unsigned long long key = i < n ? s_keys[i] : HV_KEY_PAD;
                    wave_append(i < n && key <= kv, key, s_top, &s_cnt);
                }
            } else {
                for (int i = threadIdx.x; i < n; i += blockDim.x) s_top[i] = s_keys[i];
                if (threadIdx.x == 0) s_cnt = n;
            }
            __syncthreads();
            const int m = s_cnt;
            for (int i = m + threadIdx.x; i < pk; i += blockDim.x) s_top[i] = HV_KEY_PAD;
            block_bitonic_sort(s_top, pk);
            for (int i = threadIdx.x; i < k; i += blockDim.x) {
                D[q * k + i] = hv_key_dist(s_top[i]);
                I[q * k + i] = hv_key_id(s_top[i]);
            }
            __syncthreads();  // s_keys and s_top are refilled for the next query
        }
    }
}

// The queries to redo (overflowed buffers; with check_bound, a k-th result missing or above the
// query's bound), listed on the device: thread q checks query q (q < 0: none), one list append
// per wavefront. All 64 lanes of the wavefront must call it.
__device__ __forceinline__ void redo_check(const RedoArgs& ra, int64_t q) {
    const int lane = threadIdx.x & 63;
    bool ovf = false, redo = false;
    if (q >= 0 && q < ra.n) {
        ovf = ra.cnt[q] > (unsigned int)ra.cap;
        redo = ovf || (ra.check_bound && (ra.I[q * ra.k + ra.k - 1] < 0 || ra.D[q * ra.k + ra.k - 1] > ra.bound[q]));
    }
    const unsigned long long m = __ballot(redo), mo = __ballot(ovf);
    if (!m) return;
    int base = 0;
    if (lane == 0) {
        base = atomicAdd(ra.list_cnt, (int)__builtin_popcountll(m));
        if (mo) atomicAdd(&ra.counts[0], (unsigned long long)__builtin_popcountll(mo));
        if (m != mo) atomicAdd(&ra.counts[1], (unsigned long long)__builtin_popcountll(m & ~mo));
    }
    base = __builtin_amdgcn_readfirstlane(base);
    if (redo) {
        ra.list[base + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = (int)q;
        if (ra.qn_out) ra.qn_out[q] = ra.qn_full;
        // the query's own k-th result bounds the k-th over a superset of its lists
        if (ra.bound_out)
            ra.bound_out[q] = (!ovf && ra.I[q * ra.k + ra.k - 1] >= 0) ? ra.D[q * ra.k + ra.k - 1] : 3.402823466e38f;
    }
}

// The search's counters published to the host: the last workgroup of the grid to get here (a done
// count, cleared with the counters) copies the 16 counter words, the redo count included, into
// host-visible memory (ra.publish, pinned), so the host needs no copy of its own after the device
// work. Every thread of every workgroup must call it, after its last counter update.
__device__ __forceinline__ void redo_publish(const RedoArgs& ra, unsigned int* s_last) {
    if (!ra.publish) return;
    // every counter update of this workgroup is performed (the fence waits for the atomics issued
    // before) before the workgroup is counted
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
        *s_last = __hip_atomic_fetch_add(ra.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                  gridDim.x * gridDim.y - 1;
    __syncthreads();
    if (*s_last && threadIdx.x < 16)
        ra.publish[threadIdx.x] = __hip_atomic_load(ra.stats + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_redo_list(RedoArgs ra) {
    __shared__ unsigned int s_last;
    redo_check(ra, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    redo_publish(ra, &s_last);
}

void launch_redo_list(const RedoArgs& ra, hipStream_t s) {
    if (ra.n <= 0) return;
    hipLaunchKernelGGL(k_redo_list, dim3((unsigned)((ra.n + 255) / 256)), dim3(256), 0, s, ra);
}

static size_t select_block_lds(int cap, int k) {
    return sizeof(unsigned long long) * (hv_next_pow2(std::max(std::max(cap, k), 2)) + hv_next_pow2(std::max(k, 2)));
}

void launch_select_block(const unsigned long long* surv, const unsigned int* surv_cnt, int64_t nq, int cap, int k,
                         float* D, int64_t* I, const unsigned int* done, hipStream_t s, const RedoArgs* redo) {
    if (nq <= 0) return;
    const dim3 grid((unsigned)std::min<int64_t>(nq, 1024));
    hipLaunchKernelGGL(k_select_block, grid, dim3(256), select_block_lds(cap, k), s, surv, surv_cnt, nq, cap, k, D, I,
                       done);
    // (a launch of its own: few workgroups, each with one agent-scope release for the counters)
    if (redo) launch_redo_list(*redo, s);
}

void launch_select_survivors(const unsigned long long* surv, const unsigned int* surv_cnt, int64_t nq, int cap, int k,
                             float* D, int64_t* I, int* overflow_list, int* overflow_cnt, unsigned long long* stats,
                             hipStream_t s, const RedoArgs* redo) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_select_wave, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, surv, surv_cnt, nq, cap, k, D,
                       I, overflow_list, overflow_cnt, stats);
    if (cap > kWaveSortMax || k > kWaveSortMax)
        launch_select_block(surv, surv_cnt, nq, cap, k, D, I, nullptr, s, redo);
    else if (redo)
        launch_redo_list(*redo, s);
}

}  // namespace hv
