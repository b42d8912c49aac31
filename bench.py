"""Benchmark of the IVF-flat search path on one MI355X.

A step is one hv_search_device call over the whole query batch (queries and results in HBM):
coarse quantizer -> bound -> grouped list scan -> per-query top-k. The workload is the
1,000,000 x 128 integer-mixture index (nlist 1024, k 100, 4 dimension blocks, pruning on) with
10,000 queries. Each method runs at its own operating point: the smallest nprobe of the sweep
{1, 2, 3, 4, 6, 8, 12, 16} (or, for the learned form, the calibrated policy) whose recall@100
against the FP64 ground truth (calculate_recall's rule) reaches 0.95.

Forms (each with the tile shapes, MFMA item length and graph mode its autotuner picks on the
same batch; none of these change results):
  plain-fp32    warm-up bound, FP32 blocked partial sums with pruning on every candidate
  bf16          BF16 MFMA filter with a certified error bound, exact FP32 re-rank of survivors
                (same ids and distances as plain-fp32)
  fp8           FP8 e4m3 MFMA estimates (power-of-two scales per list and per query slot), the
                k * r smallest per query re-ranked in FP32; (nprobe, r) the first of the sweep x
                {1, 2, 4} reaching the target (approximate: recall below the plain path's is possible)
  learned-bf16  the bf16 form with each query's probe count (at most 16) and bound from the
                learned predictor (65 -> 128 -> 2 MLP trained on the GPU on traces of 200,000
                training queries; serving scales calibrated on 10,000 separate validation
                queries to the target + 0.003); recall measured on the 10,000 evaluation queries
`value` is the learned-bf16 QPS (the headline form). config.forms holds every form's QPS, recall,
r2, distance error against FP64, stage and kernel times, executed block fraction and, for plain,
bf16 and learned-bf16, the host-to-host QPS of hv_search (queries and results in host memory,
PCIe included). config.learned_gain compares the learned form with the fixed-nprobe BF16 form at
equal recall for higher targets (--gain-targets).

Prints one JSON line. `roofline` is for the headline form's dominant kernel, the one with the
longest average launch:
  - the candidate scan: FP32 VALU for the plain form (3d FLOPs per unpruned pair against 157.3
    TFLOP/s); for the MFMA forms whichever bound is larger of HBM (the codes, norms and error
    norms of every touched list, once per step, plus the query residuals) and MFMA (2d FLOPs per
    pair);
  - the exact re-rank of the MFMA forms: bytes per re-ranked pair (the FP32 row, 4d, plus its
    id, 4, and the key read and written, 16) against the HBM peak;
each divided by the kernel's summed duration from HIP events on the library's stream. The MFMA
forms' scan figure sums the nearest-list pass (k_mfma_nearest) and the scans of the other ranks
(k_mfma_scan): its per-launch average is not one kernel's, so it is reported under
`config.forms.*.kernels` and only the re-rank, a single launch, is compared for `roofline`.
`cpu_baseline` is the oracle's restatement of the reference CPU scan (built with the reference's
Release flags) at the plain operating point, timed on this host: this GPU's share of the cores
(pruning on and off) and one thread.

`--workload skewed-420d` measures the 1,000,000 x 420 float mixture with per-dimension scales
exp(N(0,1)) (nlist 2048, nprobe 64, k 100, B = 8 blocks over the dimensions in descending
variance order): plain-fp32, bf16 and fp8 (r = 2), autotuned, recall and distance error against
the FP64 ground truth.

`--workload zipf-96d-10m` instead measures the uneven workload: 10,000,000 x 96 unit vectors from
a Zipf-weighted mixture (nlist 4096), 20,000 queries with per-query nprobe round(exp U[ln 8,
ln 512]) and k in {10, 100}, all in one search call per step (per-query nprobe and k), each form
autotuned on the batch (tile shapes / MFMA item length) and run as one captured HIP graph; recall
of each form against the plain path's results at the same per-query nprobe and k.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@100≥0.95, 1M×128 fp32: learned+BF16 vs plain GPU vs CPU ref"
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summaries under
    profiles/ (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        kern, val = None, None
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                line = line.strip()
                if line.startswith("traffic_kernel:"):
                    kern = line.split(":", 1)[1].strip()
                elif line.startswith("traffic_bytes_per_launch:"):
                    val = int(line.split(":", 1)[1].strip())
        if kern == kernel and val:
            return {"bytes_per_launch": val, "source": "profiles/" + name}
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bench_skewed(args):
    """The skewed-420d workload: float data, d = 420, B = 8 in variance order."""
    from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, metrics
    from harmony_amd.autotune import autotune
    from harmony_amd.index import DeviceArray, Index
    w = datagen.WORKLOADS[3]
    d, nlist, B, k, nprobe = w["d"], w["nlist"], w["dim_blocks"], w["k"], w["nprobe"][0]
    t0 = time.perf_counter()
    xb, xq = datagen.make(3, nb=args.nb, nq=args.nq)
    order = datagen.variance_order(xb)
    log("skewed data %d x %d: %.1fs" % (xb.shape[0], d, time.perf_counter() - t0))
    t0 = time.perf_counter()
    index = Index(d, nlist, nprobe=nprobe, dim_order=order)
    index.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    index.add(xb)
    log("index build %.1fs" % (time.perf_counter() - t0))
    nb = xb.shape[0]
    del xb
    nq = xq.shape[0]
    q_dev, D_dev, I_dev = DeviceArray.from_host(xq), DeviceArray(nq * k * 4), DeviceArray(nq * k * 8)
    t0 = time.perf_counter()
    index.nprobe = nprobe
    GD, GI = index.ground_truth_f64(xq, k)
    log("fp64 ground truth %.1fs" % (time.perf_counter() - t0))
    forms, ref = {}, None
    for name, prec, r in (("plain-fp32", HV_PRECISION_FP32, 4), ("bf16", HV_PRECISION_BF16, 4),
                          ("fp8", HV_PRECISION_FP8, 2)):
        index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B, precision=prec,
                        rerank_r=r, graph=False, tile_shape=0, tile_shape_first=0, scan_vsplit=0)
        tuned = None if args.no_autotune else autotune(index, xq, k, sample=nq, reps=3)[0]
        for _ in range(max(2, args.warmup)):
            index.search_device(q_dev, nq, k, D_dev, I_dev)
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            index.search_device(q_dev, nq, k, D_dev, I_dev)
            times.append(time.perf_counter() - t0)
        ms = 1000.0 * float(np.mean(times))
        I = I_dev.to_host(np.int64, (nq, k))
        D = D_dev.to_host(np.float32, (nq, k))
        if ref is None:
            ref = I
        worst = 0.0
        for i in range(0, nq, 7):
            m = dict(zip(GI[i].tolist(), GD[i].tolist()))
            for dist, j in zip(D[i], I[i]):
                g = m.get(int(j))
                if g is not None and g > 0:
                    worst = max(worst, abs(float(dist) - g) / g)
        forms[name] = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 3),
                       "recall_at_100": metrics.recall(I, GI), "recall_vs_plain": metrics.recall(I, ref),
                       "max_rel_dist_error_vs_fp64": worst, "autotuned": tuned}
        log("%s: %.3f ms, %.0f QPS, recall@100 %.5f (vs plain %.5f), max rel distance error vs fp64 %.2e"
            % (name, ms, nq / (ms / 1000.0), forms[name]["recall_at_100"], forms[name]["recall_vs_plain"], worst))
    best = max(forms, key=lambda f: forms[f]["qps"] if forms[f]["recall_at_100"] >= 0.95 else 0)
    out = {
        "metric": "QPS at recall@100>=0.95, 1M x 420 fp32 skewed mixture, nprobe 64",
        "value": forms[best]["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": forms[best]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32" if best != "fp8" else "fp8+f32",
        "data": "synthetic",
        "config": {"workload": "skewed-420d: %d x %d fp32, per-dimension scale exp(N(0,1)), nlist %d, %d queries, "
                               "nprobe %d, k %d, B=%d in variance order" % (nb, d, nlist, nq, nprobe, k, B),
                   "form": best, "forms": forms},
        "roofline": forms[best].get("roofline"),
    }
    print(json.dumps(out), flush=True)


def bench_uneven(args):
    """The zipf-96d-10m workload: per-query nprobe and k, the whole batch in one graph-captured search."""
    from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, metrics
    from harmony_amd.autotune import autotune
    from harmony_amd.index import DeviceArray, Index
    w = datagen.WORKLOADS[4]
    d, nlist, B = w["d"], w["nlist"], w["dim_blocks"]
    nb = args.nb if args.nb != 1000000 else w["nb"]
    nq = args.nq if args.nq != 10000 else w["nq"]
    t0 = time.perf_counter()
    xb, xq = datagen.make(4, nb=nb, nq=nq)
    nprobe_q, k_q = datagen.uneven_plan(4, nq)
    log("zipf data %d x %d, %d queries: %.1fs" % (nb, d, nq, time.perf_counter() - t0))
    t0 = time.perf_counter()
    index = Index(d, nlist, nprobe=512)
    index.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    index.add(xb)
    sizes = np.diff(index.list_offsets())
    log("index build %.1fs; list sizes min %d median %d max %d" % (time.perf_counter() - t0, sizes.min(),
                                                                   int(np.median(sizes)), sizes.max()))
    del xb
    kmax = int(k_q.max())
    q_dev, D_dev, I_dev = DeviceArray.from_host(xq), DeviceArray(nq * kmax * 4), DeviceArray(nq * kmax * 8)
    index.set_query_nprobe(nprobe_q)
    index.set_query_k(k_q)

    forms, ref = {}, None
    for name, prec, r in (("plain-fp32", HV_PRECISION_FP32, 4), ("bf16", HV_PRECISION_BF16, 4),
                          ("fp8", HV_PRECISION_FP8, 2)):
        index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B, precision=prec,
                        rerank_r=r, graph=True, scan_vsplit=0, tile_shape=0, tile_shape_first=0)
        tuned = None
        if not args.no_autotune:
            tuned, _ = autotune(index, xq, kmax, sample=nq, reps=3)
            index._params.graph = 1  # one captured graph per search, whatever the tuner timed
            index._push()
            log("%s autotuned: %s" % (name, tuned))
        for _ in range(max(2, args.warmup)):  # graph mode: sizing call, capture
            index.search_device(q_dev, nq, kmax, D_dev, I_dev)
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            index.search_device(q_dev, nq, kmax, D_dev, I_dev)
            times.append(time.perf_counter() - t0)
        ms = 1000.0 * float(np.mean(times))
        res = I_dev.to_host(np.int64, (nq, kmax))
        # kernel timings from eager calls with the same tuned shapes (graph replays report the
        # device work as one figure), for the roofline of the step's dominant kernel
        index._params.graph = 0
        index._push()
        ests = [index.search_device(q_dev, nq, kmax, D_dev, I_dev) for _ in range(3)]
        index._params.graph = 1
        index._push()
        avg = {kk: float(np.mean([e[kk] for e in ests])) for kk in ests[0]}
        tv, slots = touched(index, sizes)
        roof, cands = roofline_of(name, avg, d, (d + 31) // 32 * 32, tv, slots)
        if cands:
            # here the grouped rank phases are most of the step: the kernel taking the most time,
            # even where its figure sums the nearest-list pass and the filter scans' launches
            roof = dict(max(cands.values(), key=lambda c: c["kernel_ms"]))
        if name == "plain-fp32":
            ref = res
        rec = float(np.mean([metrics.recall(res[k_q == kk][:, :kk], ref[k_q == kk][:, :kk])
                             for kk in np.unique(k_q)]))
        forms[name] = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 3),
                       "recall_vs_plain": round(rec, 6), "autotuned": tuned, "calls_per_step": 1,
                       "roofline": roof, "kernels": cands, "touched_vectors": tv,
                       "stage_ms": {kk: round(avg[kk], 4) for kk in ("ms_coarse", "ms_group", "ms_scan",
                                                                     "ms_select", "ms_scan_kernels", "ms_rerank")}}
        if prec == HV_PRECISION_FP8:
            forms[name]["rerank_r"] = r
        log("%s: %.3f ms per step (%d queries, one graph launch), %.0f QPS, recall vs plain %.5f"
            % (name, ms, nq, nq / (ms / 1000.0), rec))
        if roof:
            log("%s: dominant %s %.3f ms (%.1f%% of its peak)" % (name, roof["kernel"], roof["kernel_ms"],
                                                               100.0 * roof["frac"]))
    index.set_query_nprobe(None)
    index.set_query_k(None)
    best = max(forms, key=lambda f: forms[f]["qps"] if forms[f]["recall_vs_plain"] >= 0.95 else 0)
    out = {
        "metric": "QPS, 10M x 96 Zipf mixture, per-query nprobe in [8, 512], k in {10, 100}",
        "value": forms[best]["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": forms[best]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp8+f32" if best == "fp8" else "bf16+f32",
        "data": "synthetic",
        "config": {"workload": "zipf-96d-10m: %d x %d unit vectors, nlist %d, %d queries, per-query nprobe "
                               "round(exp U[ln 8, ln 512]) (mean %.1f) and k in {10, 100} in one search call, "
                               "B=%d, graph mode" % (nb, d, nlist, nq, float(nprobe_q.mean()), B),
                   "form": best, "forms": forms},
        "roofline": forms[best].get("roofline"),
    }
    print(json.dumps(out), flush=True)


def host_info():
    """CPU of this host: what lscpu reports, and the cores this process may use."""
    info = {"nproc": os.cpu_count(),
            "affinity_cores": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
            if k.strip() == "Flags":
                info["avx512f"] = " avx512f " in " " + v + " "
    except Exception as e:  # lscpu missing: the counts above still stand
        info["lscpu_error"] = str(e)
    return info


def cpu_leg(O, data, xq, nprobe, k, dim_blocks, pruning, threads, budget_s):
    """QPS of the oracle's restatement (reference Release flags) on a growing sample of the batch,
    then the mean of three timed runs of the final sample (the growing runs are the warm-up)."""
    n = 64
    while True:
        t0 = time.perf_counter()
        O.search(data, xq[:n], k, nprobe=nprobe, dim_blocks=dim_blocks, pruning=pruning, nthreads=threads, fast=True)
        dt = time.perf_counter() - t0
        if dt > budget_s / 6 or n >= xq.shape[0]:
            break
        n = min(xq.shape[0], int(n * max(2.0, (budget_s / 6) / max(dt, 1e-3))))
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.search(data, xq[:n], k, nprobe=nprobe, dim_blocks=dim_blocks, pruning=pruning, nthreads=threads, fast=True)
        ts.append(time.perf_counter() - t0)
    return n / float(np.mean(ts)), n


def cpu_baseline(index, xq, nprobe, k, dim_blocks, budget_s):
    """The oracle's restatement of the reference per-node scan (coarse + warm-up bound + blocked
    scan + merge + sort, Release flags), single process, timed on this host: with the box's CPU
    share of threads (pruning on and off) and with one thread."""
    from oracle import oracle as O
    info = host_info()
    threads = max(1, min(info["affinity_cores"] or 1, 16))  # one GPU's share of the host
    data = O.IndexData.from_gpu_index(index)
    legs = {}
    for name, thr, pr in (("threads_pruning_on", threads, True), ("threads_pruning_off", threads, False),
                          ("1_thread_pruning_on", 1, True)):
        qps, n = cpu_leg(O, data, xq, nprobe, k, dim_blocks, pr, thr, budget_s / 3)
        legs[name] = {"qps": round(qps, 1), "threads": thr, "pruning": pr, "sample_queries": n}
        log("cpu %s: %.1f QPS (%d queries, %d threads)" % (name, qps, n, thr))
    main = legs["threads_pruning_on"]
    return {"value": main["qps"], "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "%d of %d queries at nprobe %d, k %d, B=%d pruning on, single process, %d OpenMP threads "
                      "(this GPU's share of the host); mean of 3 timed runs after the sizing runs"
                      % (main["sample_queries"], xq.shape[0], nprobe, k, dim_blocks, threads),
            "host": info, "legs": legs}


def touched(index, sizes):
    """(touched vectors, searched (query, rank) slots) of the last search: vectors of the lists that
    at least one query searched, and the probe slots searched (each query's first nprobe_q ranks)."""
    pr = index.last_probes()
    qn = index.last_nprobe()
    mask = np.arange(pr.shape[1])[None, :] < qn[:, None]
    lists = np.unique(pr[mask])
    return int(sizes[lists].sum()), int(qn.sum())


def roofline_of(name, st, d, dpad, touched_vecs, slots):
    """Roofline of the kernel that dominates the search's device time: the candidate scan (FP32
    VALU for the plain form; the MFMA filter, HBM-bound on the codes of the touched lists, for the
    others) or the exact re-rank (HBM: the FP32 row, id and key of every re-ranked pair)."""
    fp8 = name.startswith("fp8")
    plain = name == "plain-fp32"
    cands = {}
    t_scan = st["ms_scan_kernels"] / 1000.0
    if t_scan > 0:
        if plain:
            flops = 3.0 * d * st["scan_pairs"]
            ach = flops / t_scan / 1e12
            cands["scan"] = {"bound": "mfma", "pipe": "fp32 valu (direct-difference sub + fma; packed FP32 off)",
                             "kernel": "k_list_scan", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 5),
                             "algorithmic_flops_per_step": flops, "kernel_ms": round(st["ms_scan_kernels"], 4),
                             "launches": int(st["scan_launches"]), "pairs": int(st["scan_pairs"])}
        else:
            s = 1 if fp8 else 2
            # one launch is the nearest-list pass alone (every searched rank in one pass); more add
            # the filter scans of the other ranks
            scan_kernel = "k_mfma_nearest" if int(st["scan_launches"]) == 1 else "k_mfma_nearest+k_mfma_scan"
            nbytes = touched_vecs * (s * dpad + (4 if fp8 else 8)) + slots * s * dpad
            flops = 2.0 * d * st["scan_pairs"]
            peak_f = 5000.0 if fp8 else PEAK_BF16_TFLOPS
            t_hbm = nbytes / (PEAK_HBM_GBS * 1e9)
            t_mfma = flops / (peak_f * 1e12)
            if t_hbm >= t_mfma:
                ach = nbytes / t_scan / 1e9
                cands["scan"] = {"bound": "hbm", "kernel": scan_kernel, "achieved": round(ach, 1),
                                 "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 5),
                                 "algorithmic_bytes_per_step": nbytes, "touched_vectors": touched_vecs,
                                 "mfma_frac": round(flops / t_scan / 1e12 / peak_f, 6)}
            else:
                ach = flops / t_scan / 1e12
                cands["scan"] = {"bound": "mfma", "kernel": scan_kernel, "achieved": round(ach, 3),
                                 "peak": peak_f,
                                 "unit": "TFLOP/s", "frac": round(ach / peak_f, 5), "algorithmic_flops_per_step": flops}
            cands["scan"].update({"kernel_ms": round(st["ms_scan_kernels"], 4), "launches": int(st["scan_launches"]),
                                  "pairs": int(st["scan_pairs"])})
    if st["ms_rerank"] > 0:
        rbytes = (4.0 * d + 20.0) * st["rerank_pairs"]
        gbs = rbytes / (st["ms_rerank"] / 1000.0) / 1e9
        # rows staged through LDS when d and every chunk start are multiples of 4 (the dims here)
        rk = "k_rerank_tile" if d % 4 == 0 else "k_rerank_flat"
        cands["rerank"] = {"bound": "hbm", "kernel": rk, "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 5),
                           "algorithmic_bytes_per_step": rbytes, "kernel_ms": round(st["ms_rerank"], 4),
                           "launches": 1, "pairs": int(st["rerank_pairs"]),
                           "note": "every pair's row counted; queries of one list share rows, read once from HBM "
                                   "through one XCD's L2, so this rate can pass the HBM peak"}
    if not cands:
        return None, {}
    # the dominant kernel: one kernel's launches (the plain scan's warm-up and main launches are
    # one kernel; an MFMA form's scan figure of several launches mixes two kernels, so the re-rank
    # stands for it then)
    single = {c: v for c, v in cands.items() if plain or c != "scan" or v["launches"] == 1} or cands
    dom = max(single, key=lambda c: single[c]["kernel_ms"] / max(1, single[c]["launches"]))
    return dict(cands[dom]), cands


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--sweep", default="1,2,3,4,6,8,12,16")
    ap.add_argument("--target", type=float, default=0.95, help="recall@100 of the operating points")
    ap.add_argument("--learned-cap", type=int, default=16, help="largest per-query nprobe of the learned form")
    ap.add_argument("--gain-targets", default="0.99",
                    help="extra recall targets where the learned form is compared with the fixed-nprobe BF16 form")
    ap.add_argument("--nq", type=int, default=10000)
    ap.add_argument("--nb", type=int, default=1000000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--train-queries", type=int, default=200000)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--no-host-to-host", action="store_true",
                    help="skip the host-buffer (hv_search) timings (profiler runs time device work only)")
    ap.add_argument("--workload", default="int-mixture-1m", choices=["int-mixture-1m", "skewed-420d", "zipf-96d-10m"])
    ap.add_argument("--trace-markers", action="store_true",
                    help="before each form's timed steps run a device copy of (form number) MiB, so the steps "
                         "can be found in a profiler's kernel trace")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run 3 device-to-device copies of this size (a copy kernel of known traffic, "
                         "for scaling HBM counters in a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build targets a single MI355X (the path shards by "
              "inverted list across GPUs, which is a design, not a measured configuration). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "zipf-96d-10m":
        bench_uneven(args)
        return
    if args.workload == "skewed-420d":
        bench_skewed(args)
        return

    from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, learned, metrics
    from harmony_amd._lib import lib
    from harmony_amd.autotune import autotune
    from harmony_amd.index import DeviceArray, Index

    if args.copy_calibration_mb:
        nbytes = args.copy_calibration_mb << 20
        src, dst = DeviceArray(nbytes), DeviceArray(nbytes)
        for _ in range(3):
            assert lib().hv_memcpy_dtod(dst.ptr, src.ptr, nbytes) == 0
        log("copy calibration: 3 device-to-device copies of %d MiB" % args.copy_calibration_mb)
        del src, dst

    k, d, nlist, B = 100, 128, 1024, 4
    dpad = (d + 31) // 32 * 32
    t0 = time.perf_counter()
    xb, xq = datagen.make(2, nb=args.nb, nq=args.nq)
    log("data %.1fs" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    index = Index(d, nlist, nprobe=16)
    index.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    index.add(xb)
    nb = int(xb.shape[0])
    del xb
    sizes = np.diff(index.list_offsets())
    log("index build (GPU k-means + add) %.1fs" % (time.perf_counter() - t0))
    nq = xq.shape[0]
    q_dev = DeviceArray.from_host(xq)
    D_dev = DeviceArray(nq * k * 4)
    I_dev = DeviceArray(nq * k * 8)

    def params(precision, use_learned=False, rerank_r=4, nprobe=None):
        if nprobe is not None:
            index.nprobe = nprobe
        index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B, precision=precision,
                        learned=use_learned, rerank_r=rerank_r, graph=False, tile_shape=0, tile_shape_first=0,
                        scan_vsplit=0)

    def run_once():
        st = index.search_device(q_dev, nq, k, D_dev, I_dev)
        return I_dev.to_host(np.int64, (nq, k)), D_dev.to_host(np.float32, (nq, k)), st

    # 1. ground truth and the recall of every nprobe of the sweep (plain and BF16 return the same
    #    ids, checked at the operating point); operating point = smallest nprobe reaching the target
    sweep = sorted(int(x) for x in args.sweep.split(","))
    t0 = time.perf_counter()
    index.nprobe = max(sweep + [16])  # the ground truth's candidate bound comes from this search
    index.preSearch(pruning=True, dim_blocks=B)
    GD, GI = index.ground_truth_f64(xq, k)
    GD32 = GD.astype(np.float32)
    log("fp64 ground truth %.1fs" % (time.perf_counter() - t0))
    recall_at = {}
    for npb in sweep:
        params(HV_PRECISION_FP32, nprobe=npb)
        I, D, _ = run_once()
        recall_at[npb] = metrics.recall(I, GI, D, GD32)
        log("plain nprobe %d recall@100 %.5f" % (npb, recall_at[npb]))
    ok = [p for p in sweep if recall_at[p] >= args.target]
    op_plain = min(ok) if ok else max(sweep)

    def dist_error(D, I):
        """max relative |D - D_fp64| over the returned ids that the FP64 ground truth holds"""
        worst = 0.0
        for i in range(0, nq, 7):
            m = dict(zip(GI[i].tolist(), GD[i].tolist()))
            for dist, j in zip(D[i], I[i]):
                ref = m.get(int(j))
                if ref is not None and ref > 0:
                    worst = max(worst, abs(float(dist) - ref) / ref)
        return worst

    forms_seen = []
    mark_src, mark_dst = (DeviceArray(16 << 20), DeviceArray(16 << 20)) if args.trace_markers else (None, None)

    def trace_mark(i):
        assert lib().hv_memcpy_dtod(mark_dst.ptr, mark_src.ptr, i << 20) == 0
        log("trace marker: %d MiB copy before the timed steps of form %d" % (i, i))

    def measure(name, precision, use_learned=False, rerank_r=4, nprobe=None, h2h=False):
        if use_learned:
            index.preSearch(precision=precision, learned=True, graph=False, tile_shape=0, tile_shape_first=0,
                            scan_vsplit=0)
            if nprobe is not None:
                index.nprobe = nprobe
        else:
            params(precision, False, rerank_r, nprobe)
        tuned = None
        if not args.no_autotune:
            tuned, _ = autotune(index, xq, k, sample=nq, reps=3)
            log("%s autotuned: %s" % (name, tuned))
        for _ in range(args.warmup):
            index.search_device(q_dev, nq, k, D_dev, I_dev)
        if args.trace_markers:
            # a device copy of a size unique to this form marks where its timed steps start in a
            # profiler's kernel trace (1 MiB times the form's number)
            trace_mark(len(forms_seen) + 1)
        forms_seen.append(name)
        times, sts = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            st = index.search_device(q_dev, nq, k, D_dev, I_dev)
            times.append(time.perf_counter() - t0)
            sts.append(st)
        ms = 1000.0 * float(np.mean(times))
        I, D = I_dev.to_host(np.int64, (nq, k)), D_dev.to_host(np.float32, (nq, k))
        tv, slots = touched(index, sizes)
        avg = {kk: float(np.mean([st[kk] for st in sts])) for kk in sts[0]}
        if tuned and tuned.get("graph"):
            # graph replays report the device work as one figure: kernel timings from eager calls
            # with the same tuned shapes
            index._params.graph = 0  # the tuned tile shapes stay
            index._push()
            ests = [index.search_device(q_dev, nq, k, D_dev, I_dev) for _ in range(3)]
            for kk in ("ms_coarse", "ms_warmup", "ms_group", "ms_scan", "ms_select", "ms_main", "ms_rerank",
                       "ms_scan_kernels", "scan_launches"):
                avg[kk] = float(np.mean([e[kk] for e in ests]))
            index._params.graph = 1
            index._push()
        roof, cands = roofline_of(name, avg, d, dpad, tv, slots)
        rec, correct = metrics.recall(I, GI, D, GD32, return_correct=True)
        f = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 4), "nprobe": int(index.nprobe),
             "recall_at_100": rec, "r2": metrics.r2(I, D, GD32), "distance_sanity_correct": correct,
             "max_rel_dist_error_vs_fp64": dist_error(D, I), "autotuned": tuned,
             "stage_ms": {kk: round(avg[kk], 4) for kk in ("ms_coarse", "ms_warmup", "ms_group", "ms_scan",
                                                           "ms_select", "ms_scan_kernels", "ms_rerank")},
             "candidates": int(avg["candidates"]), "scan_pairs": int(avg["scan_pairs"]),
             "survivors": int(avg["survivors"]), "overflow_queries": int(avg["overflow_queries"]),
             "fallback_queries": int(avg["fallback_queries"]), "touched_vectors": tv,
             "executed_block_fraction": round(avg["block_evals"] / max(1.0, avg["block_evals_max"]), 5),
             "roofline": roof, "kernels": cands}
        if use_learned:
            f["mean_probes"] = round(float(index.last_nprobe().mean()), 4)
            f["widened_queries"] = int(avg["widened_queries"])
            f["widen_margin"] = float(index._params.learned_margin)
        if precision == HV_PRECISION_FP8:
            f["rerank_r"] = rerank_r
        if h2h and not args.no_host_to_host:
            index.search(xq, k)
            t0 = time.perf_counter()
            for _ in range(3):
                index.search(xq, k)
            f["host_to_host_qps"] = round(nq / ((time.perf_counter() - t0) / 3), 1)
        log("%s: nprobe %d, %.3f ms/step, %.0f QPS%s, recall %.5f, dominant %s %.3f ms (%.1f%% of its peak)"
            % (name, f["nprobe"], ms, f["qps"],
               (", host-to-host %.0f QPS" % f["host_to_host_qps"]) if "host_to_host_qps" in f else "", rec,
               roof["kernel"] if roof else "-", roof["kernel_ms"] if roof else 0, 100 * roof["frac"] if roof else 0))
        return f, I

    forms = {}
    f, I_plain = measure("plain-fp32", HV_PRECISION_FP32, nprobe=op_plain, h2h=True)
    forms["plain-fp32"] = f
    f, I_bf16 = measure("bf16", HV_PRECISION_BF16, nprobe=op_plain, h2h=True)
    forms["bf16"] = f
    same = bool(np.array_equal(I_plain, I_bf16))
    op = {"plain-fp32": {"nprobe": op_plain, "recall_at_100": recall_at[op_plain]},
          "bf16": {"nprobe": op_plain, "recall_at_100": forms["bf16"]["recall_at_100"], "ids_equal_plain": same}}
    if not args.no_fp8:
        found = None
        for npb in [p for p in sweep if p >= op_plain]:
            for r in (1, 2, 4):
                params(HV_PRECISION_FP8, rerank_r=r, nprobe=npb)
                I, D, _ = run_once()
                rr = metrics.recall(I, GI)
                log("fp8 nprobe %d r %d recall@100 %.5f" % (npb, r, rr))
                if rr >= args.target:
                    found = (npb, r)
                    break
            if found:
                break
        if found:
            f, _ = measure("fp8", HV_PRECISION_FP8, rerank_r=found[1], nprobe=found[0])
            forms["fp8"] = f
            op["fp8"] = {"nprobe": found[0], "rerank_r": found[1], "recall_at_100": f["recall_at_100"]}

    learned_info, gain = None, []
    best = "bf16"
    if not args.no_learned:
        t0 = time.perf_counter()
        _, xt = datagen.make(2, nb=1, nq=args.train_queries, query_set=1)
        _, xv = datagen.make(2, nb=1, nq=10000, query_set=2)
        feats, targets, _ = learned.collect_traces(index, xt, k=k)
        vf, vt, vexact = learned.collect_traces(index, xv, k=k)
        del xt
        t1 = time.perf_counter()
        losses = learned.train(index, feats, targets, epochs=20, batch=1024, lr=2e-3, seed=0)
        t2 = time.perf_counter()
        cap = args.learned_cap
        index.preSearch(precision=HV_PRECISION_BF16)
        s_np, s_b, rec_v, probes_v = learned.calibrate(index, vf, vt, xv, vexact, nprobe=cap, k=k,
                                                       target_recall=args.target, margin=0.003)
        learned_info = {"train_queries": int(feats.shape[0]), "trace_s": round(t1 - t0, 2),
                        "train_s": round(t2 - t1, 2), "epochs": 20, "loss_first": float(losses[0]),
                        "loss_last": float(losses[-1]), "nprobe_cap": cap, "nprobe_scale": s_np,
                        "bound_scale": round(s_b, 4), "widen_margin": float(index._params.learned_margin),
                        "validation_recall": round(rec_v, 5), "validation_mean_probes": round(probes_v, 3)}
        log("predictor: traces %.1fs, training %.1fs (loss %.4f -> %.4f), calibration: nprobe scale %.2f, "
            "bound scale %.3f, validation recall %.4f, mean probes %.2f"
            % (t1 - t0, t2 - t1, losses[0], losses[-1], s_np, s_b, rec_v, probes_v))
        f, _ = measure("learned-bf16", HV_PRECISION_BF16, use_learned=True, nprobe=cap, h2h=True)
        forms["learned-bf16"] = f
        op["learned-bf16"] = {"nprobe_cap": cap, "nprobe_scale": s_np, "widen_margin": learned_info["widen_margin"],
                              "mean_probes": f["mean_probes"], "recall_at_100": f["recall_at_100"]}
        best = "learned-bf16"
        # what the network gains over the method without it at equal recall: the BF16 form at the
        # smallest fixed nprobe reaching each target, against the learned form calibrated to it
        for tgt in [float(x) for x in args.gain_targets.split(",") if x]:
            okp = [p for p in sweep if recall_at[p] >= tgt]
            if not okp:
                continue
            fixed, _ = measure("bf16@%g" % tgt, HV_PRECISION_BF16, nprobe=min(okp))
            s2 = learned.calibrate(index, vf, vt, xv, vexact, nprobe=cap, k=k, target_recall=tgt, margin=0.003)
            lf, _ = measure("learned-bf16@%g" % tgt, HV_PRECISION_BF16, use_learned=True, nprobe=cap)
            gain.append({"target": tgt, "widen_margin": float(index._params.learned_margin),
                         "fixed_nprobe": min(okp), "fixed_qps": fixed["qps"],
                         "fixed_recall": fixed["recall_at_100"], "fixed_candidates": fixed["candidates"],
                         "learned_nprobe_scale": s2[0], "learned_qps": lf["qps"], "learned_recall": lf["recall_at_100"],
                         "learned_candidates": lf["candidates"], "learned_mean_probes": lf["mean_probes"],
                         "speedup": round(lf["qps"] / fixed["qps"], 3)})
        index.preSearch(learned=False)

    base = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        index.nprobe = op_plain
        base = cpu_baseline(index, xq, op_plain, k, B, args.cpu_budget)
        log("cpu baseline %.1f qps (%s) in %.1fs" % (base["value"], base["sample"], time.perf_counter() - t0))

    roof = dict(forms[best]["roofline"])
    tr = profiled_traffic(roof["kernel"])
    roof["traffic"] = tr["bytes_per_launch"] if tr else None
    if tr:
        roof["traffic_source"] = tr["source"]
    qps = forms[best]["qps"]
    h2h = {n: forms[n]["host_to_host_qps"] for n in forms if "host_to_host_qps" in forms[n]}
    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": forms[best]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32",
        "data": "synthetic",
        "config": {"workload": "int-mixture-1m: %d x 128 fp32 base, %d queries, nlist 1024, k 100, B=4 dimension "
                               "blocks, pruning on" % (nb, nq),
                   "form": best, "target_recall": args.target, "operating_points": op,
                   "recall_sweep_plain": recall_at, "forms": forms,
                   "host_to_host_qps": h2h,
                   "speedup_vs_plain_gpu": round(qps / forms["plain-fp32"]["qps"], 3),
                   "speedup_vs_plain_gpu_host_to_host":
                       round(h2h[best] / h2h["plain-fp32"], 3) if best in h2h and "plain-fp32" in h2h else None,
                   "bf16_identical_to_plain": same, "predictor": learned_info, "learned_gain": gain},
        "roofline": roof,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
