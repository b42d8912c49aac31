"""Benchmark of the IVF-flat search path on one MI355X.

A step is one hv_search_device call over the whole query batch (queries already in HBM):
coarse quantizer -> bound -> grouped list scan -> per-query top-k. The workload is the
1,000,000 x 128 integer-mixture index (nlist 1024, k 100, 4 dimension blocks, pruning on) with
10,000 queries; the operating point is the smallest nprobe of {16, 64, 256} whose recall@100
against the FP64 ground truth is >= 0.95.

Forms measured at that nprobe (each with the tile shapes, MFMA item length and graph mode its
autotuner picks on the same batch; none of these change results):
  plain-fp32    warm-up bound, FP32 blocked partial sums with pruning on every candidate
  bf16          BF16 MFMA filter with a certified error bound, exact FP32 re-rank of survivors
                (same ids and distances as plain-fp32)
  fp8           FP8 e4m3 MFMA estimates (power-of-two scales per list and per query slot), the
                k * r smallest per query re-ranked in FP32; r is the smallest of 1, 2, 4 whose
                recall@100 reaches 0.95 (approximate: recall below the plain path's is possible)
  learned-bf16  the bf16 form with each query's probe count trimmed and its bound seeded by the
                learned predictor (65 -> 128 -> 2 MLP trained on the GPU on traces of 200,000
                training queries; serving scales calibrated on 10,000 separate validation
                queries; nprobe capped at the operating point); recall measured on the 10,000
                evaluation queries against the FP64 ground truth
`value` is the learned-bf16 QPS (the headline form); every form's QPS, recall@100 and distance
error against the FP64 ground truth (over the returned ids the ground truth also holds) are
under config.forms.

Prints one JSON line. `roofline` is for the headline form's dominant kernel, the longer of
  - the scan of the remaining probed lists: algorithmic FLOPs per (query, vector) pair (2d
    multiply-adds for the BF16 dot product, 3d for the FP32 subtract + multiply-add) times the
    kernel's unpruned pair count, against the MFMA (BF16) or FP32 peak;
  - the exact re-rank of the BF16 forms: algorithmic bytes per re-ranked pair (the FP32 row,
    4d, plus its id, 4, and the key read and written, 16) against the HBM peak;
each divided by the kernel's duration measured with HIP events on the library's stream.
`cpu_baseline` is the oracle's restatement of the reference CPU scan (built with the
reference's Release flags), timed on this host on a bounded query sample.

`--workload skewed-420d` measures the 1,000,000 x 420 float mixture with per-dimension scales
exp(N(0,1)) (nlist 2048, nprobe 64, k 100, B = 8 blocks over the dimensions in descending
variance order): plain-fp32, bf16 and fp8 (r = 2), autotuned, recall and distance error against
the FP64 ground truth.

`--workload zipf-96d-10m` instead measures the uneven workload: 10,000,000 x 96 unit vectors from
a Zipf-weighted mixture (nlist 4096), 20,000 queries with per-query nprobe round(exp U[ln 8,
ln 512]) and k in {10, 100} (two calls per step, one per k), each form in graph mode (one
captured HIP graph per call) with its autotuned item length; recall of each form against the
plain path's results at the same per-query nprobe.
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


def cpu_baseline(index, xq, nprobe, k, dim_blocks, budget_s):
    """Oracle restatement of the reference per-node scan, reference Release flags, on a
    bounded, growing sample of the query batch (about budget_s seconds of CPU work)."""
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    data = O.IndexData.from_gpu_index(index)
    n = 64
    spent = 0.0
    qps = None
    while True:
        sub = xq[:n]
        t0 = time.perf_counter()
        O.search(data, sub, k, nprobe=nprobe, dim_blocks=dim_blocks, pruning=True, nthreads=threads, fast=True)
        dt = time.perf_counter() - t0
        spent += dt
        qps = n / dt
        if dt > budget_s / 3 or n >= xq.shape[0] or spent > budget_s:
            break
        n = min(xq.shape[0], int(n * max(2.0, (budget_s / 3) / max(dt, 1e-3))))
    return {"value": qps, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "%d of %d queries, nprobe %d, k %d, B=%d pruning on, single process, %d OpenMP threads"
                      % (n, xq.shape[0], nprobe, k, dim_blocks, threads)}


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
    }
    print(json.dumps(out), flush=True)


def bench_uneven(args):
    """The zipf-96d-10m workload: uneven per-query nprobe and k, graph-captured searches."""
    from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, metrics
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
    groups = []
    for kk in (10, 100):
        rows = np.where(k_q == kk)[0]
        q = np.ascontiguousarray(xq[rows])
        groups.append({"k": kk, "rows": rows, "nprobe": np.ascontiguousarray(nprobe_q[rows]),
                       "q": DeviceArray.from_host(q), "D": DeviceArray(rows.size * kk * 4),
                       "I": DeviceArray(rows.size * kk * 8), "host_q": q})

    def run_step():
        for g in groups:
            index.set_query_nprobe(g["nprobe"])
            index.search_device(g["q"], g["rows"].size, g["k"], g["D"], g["I"])

    forms, ref = {}, {}
    for name, prec, r in (("plain-fp32", HV_PRECISION_FP32, 4), ("bf16", HV_PRECISION_BF16, 4),
                          ("fp8", HV_PRECISION_FP8, 2)):
        index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B, precision=prec,
                        rerank_r=r, graph=True, scan_vsplit=0, tile_shape=0, tile_shape_first=0)
        for _ in range(max(2, args.warmup)):  # graph mode: sizing call, capture
            run_step()
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            run_step()
            times.append(time.perf_counter() - t0)
        ms = 1000.0 * float(np.mean(times))
        res = [g["I"].to_host(np.int64, (g["rows"].size, g["k"])) for g in groups]
        if name == "plain-fp32":
            ref = res
        rec = float(np.mean([metrics.recall(a, b) for a, b in zip(res, ref)]))
        forms[name] = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 3),
                       "recall_vs_plain": round(rec, 6)}
        if prec == HV_PRECISION_FP8:
            forms[name]["rerank_r"] = r
        log("%s: %.3f ms per step (%d queries, two calls), %.0f QPS, recall vs plain %.5f"
            % (name, ms, nq, nq / (ms / 1000.0), rec))
    index.set_query_nprobe(None)
    best = max(forms, key=lambda f: forms[f]["qps"] if forms[f]["recall_vs_plain"] >= 0.95 else 0)
    out = {
        "metric": "QPS, 10M x 96 Zipf mixture, per-query nprobe in [8, 512], k in {10, 100}",
        "value": forms[best]["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": forms[best]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp8+f32" if best == "fp8" else "bf16+f32",
        "data": "synthetic",
        "config": {"workload": "zipf-96d-10m: %d x %d unit vectors, nlist %d, %d queries, per-query nprobe "
                               "round(exp U[ln 8, ln 512]) (mean %.1f), k in {10, 100}, B=%d, graph mode"
                               % (nb, d, nlist, nq, float(nprobe_q.mean()), B),
                   "form": best, "forms": forms},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nprobe", type=int, default=0, help="fixed nprobe (0: smallest of the sweep reaching 0.95)")
    ap.add_argument("--sweep", default="16,64,256")
    ap.add_argument("--nq", type=int, default=10000)
    ap.add_argument("--nb", type=int, default=1000000)
    ap.add_argument("--no-recall", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--train-queries", type=int, default=200000)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--workload", default="int-mixture-1m", choices=["int-mixture-1m", "skewed-420d", "zipf-96d-10m"])
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

    from harmony_amd import datagen, learned, metrics
    from harmony_amd._lib import lib
    from harmony_amd.index import DeviceArray, Index

    if args.copy_calibration_mb:
        nbytes = args.copy_calibration_mb << 20
        src, dst = DeviceArray(nbytes), DeviceArray(nbytes)
        for _ in range(3):
            assert lib().hv_memcpy_dtod(dst.ptr, src.ptr, nbytes) == 0
        log("copy calibration: 3 device-to-device copies of %d MiB" % args.copy_calibration_mb)
        del src, dst

    k, d, nlist, B = 100, 128, 1024, 4
    t0 = time.perf_counter()
    xb, xq = datagen.make(2, nb=args.nb, nq=args.nq)
    log("data %.1fs" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    index = Index(d, nlist, nprobe=16)
    index.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    index.add(xb)
    log("index build (GPU k-means + add) %.1fs" % (time.perf_counter() - t0))
    index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B)
    nq = xq.shape[0]
    q_dev = DeviceArray.from_host(xq)
    D_dev = DeviceArray(nq * k * 4)
    I_dev = DeviceArray(nq * k * 8)

    sweep = [int(s) for s in args.sweep.split(",")]
    recall_at = {}
    if not args.no_recall:
        t0 = time.perf_counter()
        index.nprobe = max(sweep)
        GD, GI = index.ground_truth_f64(xq, k)
        log("fp64 ground truth %.1fs" % (time.perf_counter() - t0))
        for npb in sweep:
            index.nprobe = npb
            index.search_device(q_dev, nq, k, D_dev, I_dev)
            I = I_dev.to_host(np.int64, (nq, k))
            recall_at[npb] = metrics.recall(I, GI)
            log("nprobe %d recall@100 %.4f" % (npb, recall_at[npb]))
    if args.nprobe:
        nprobe = args.nprobe
    elif recall_at:
        ok = [p for p in sweep if recall_at[p] >= 0.95]
        nprobe = min(ok) if ok else max(sweep)
    else:
        nprobe = sweep[0]
    index.nprobe = nprobe

    from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32

    from harmony_amd.autotune import autotune

    def measure(precision, use_learned=False, rerank_r=4):
        index.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=B, precision=precision,
                        learned=use_learned, rerank_r=rerank_r, graph=False, tile_shape=0, tile_shape_first=0,
                        scan_vsplit=0)
        tuned = None
        if not args.no_autotune:
            tuned, _ = autotune(index, xq, k, sample=nq, reps=3)
            log("autotuned: %s" % tuned)
        for _ in range(args.warmup):
            index.search_device(q_dev, nq, k, D_dev, I_dev)
        times, sts = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            st = index.search_device(q_dev, nq, k, D_dev, I_dev)
            times.append(time.perf_counter() - t0)
            sts.append(st)
        ms = 1000.0 * float(np.mean(times))
        avg = {kk: float(np.mean([st[kk] for st in sts])) for kk in sts[0]}
        if tuned and tuned.get("graph"):
            # graph replays report the whole device work as one figure: take the kernel timings
            # from an eager call with the same tuned shapes
            index.preSearch(graph=False)
            est = index.search_device(q_dev, nq, k, D_dev, I_dev)
            for kk in ("ms_coarse", "ms_warmup", "ms_group", "ms_scan", "ms_select", "ms_main", "ms_rerank"):
                avg[kk] = est[kk]
            index.preSearch(graph=True)
        return ms, avg, I_dev.to_host(np.int64, (nq, k)), D_dev.to_host(np.float32, (nq, k)), tuned

    def dist_error(D, I):
        """max relative |D - D_fp64| over the returned ids that the FP64 ground truth holds"""
        if args.no_recall:
            return None
        worst = 0.0
        for i in range(0, nq, 7):
            m = dict(zip(GI[i].tolist(), GD[i].tolist()))
            for dist, j in zip(D[i], I[i]):
                ref = m.get(int(j))
                if ref is not None and ref > 0:
                    worst = max(worst, abs(float(dist) - ref) / ref)
        return worst

    forms = {}
    ids = {}
    runs = [("plain-fp32", HV_PRECISION_FP32, False, 3.0, PEAK_FP32_TFLOPS, "fp32-valu (direct-difference sub+fma)",
             "k_list_scan"),
            ("bf16", HV_PRECISION_BF16, False, 2.0, PEAK_BF16_TFLOPS, "bf16 mfma (v_mfma_f32_16x16x32_bf16)",
             "k_mfma_scan")]
    fp8_r = None
    if not args.no_fp8:
        fp8_r = 4
        if not args.no_recall:
            for r in (1, 2, 4):
                index.preSearch(precision=HV_PRECISION_FP8, rerank_r=r, learned=False)
                index.search_device(q_dev, nq, k, D_dev, I_dev)
                rr = metrics.recall(I_dev.to_host(np.int64, (nq, k)), GI)
                log("fp8 r %d recall@100 %.4f" % (r, rr))
                if rr >= 0.95:
                    fp8_r = r
                    break
        runs.append(("fp8", HV_PRECISION_FP8, False, 2.0, 5000.0, "fp8 mfma (v_mfma_f32_16x16x32_fp8_fp8)",
                     "k_mfma_scan"))
    learned_info = None
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
        s_np, s_b, rec_v, probes_v = learned.calibrate(index, vf, vt, xv, vexact, nprobe=nprobe, k=k)
        learned_info = {"train_queries": int(feats.shape[0]), "trace_s": round(t1 - t0, 2),
                        "train_s": round(t2 - t1, 2), "epochs": 20, "loss_first": float(losses[0]),
                        "loss_last": float(losses[-1]), "nprobe_scale": s_np, "bound_scale": round(s_b, 4),
                        "validation_recall": round(rec_v, 5), "validation_mean_probes": round(probes_v, 3)}
        log("predictor: traces %.1fs, training %.1fs (loss %.4f -> %.4f), calibration: nprobe scale %.2f, "
            "bound scale %.3f, validation recall %.4f, mean probes %.2f"
            % (t1 - t0, t2 - t1, losses[0], losses[-1], s_np, s_b, rec_v, probes_v))
        runs.append(("learned-bf16", HV_PRECISION_BF16, True, 2.0, PEAK_BF16_TFLOPS,
                     "bf16 mfma (v_mfma_f32_16x16x32_bf16)", "k_mfma_scan"))
    for name, prec, use_learned, flops_per_dim, peak, pipe, kern in runs:
        ms, st, I, Dh, tuned = measure(prec, use_learned, fp8_r if prec == HV_PRECISION_FP8 else 4)
        ids[name] = I
        flops = flops_per_dim * d * st["main_pairs"]
        ach = flops / (st["ms_main"] / 1000.0) / 1e12 if st["ms_main"] > 0 else 0.0
        roof = {"bound": "mfma", "pipe": pipe, "kernel": kern, "achieved": round(ach, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(ach / peak, 5), "algorithmic_flops_per_launch": flops,
                "kernel_ms": round(st["ms_main"], 4)}
        if st["ms_rerank"] > st["ms_main"]:
            rbytes = (4.0 * d + 20.0) * st["rerank_pairs"]
            gbs = rbytes / (st["ms_rerank"] / 1000.0) / 1e9
            roof = {"bound": "hbm", "kernel": "k_rerank_flat", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 5), "algorithmic_bytes_per_launch": rbytes,
                    "kernel_ms": round(st["ms_rerank"], 4), "pairs": int(st["rerank_pairs"])}
        rec = metrics.recall(I, GI) if not args.no_recall else None
        forms[name] = {"qps": round(nq / (ms / 1000.0), 1), "ms_per_step": round(ms, 4),
                       "recall_at_100": rec, "max_rel_dist_error_vs_fp64": dist_error(Dh, I),
                       "autotuned": tuned,
                       "stage_ms": {kk: round(st[kk], 4) for kk in ("ms_coarse", "ms_warmup", "ms_group", "ms_scan",
                                                                    "ms_select", "ms_main")},
                       "candidates": int(st["candidates"]), "main_pairs": int(st["main_pairs"]),
                       "survivors": int(st["survivors"]), "overflow_queries": int(st["overflow_queries"]),
                       "fallback_queries": int(st["fallback_queries"]),
                       "rerank_ms": round(st["ms_rerank"], 4),
                       "roofline": roof}
        if use_learned:
            forms[name]["mean_probes"] = round(float(index.last_nprobe(nq).mean()), 3)
        if prec == HV_PRECISION_FP8:
            forms[name]["rerank_r"] = fp8_r
        log("%s: %.3f ms/step, %.0f QPS, recall %s, main scan %.3f ms (%.1f TFLOP/s), re-rank %.3f ms; "
            "dominant %s %.1f%% of peak" % (name, ms, nq / (ms / 1000.0), rec, st["ms_main"], ach, st["ms_rerank"],
                                           roof["kernel"], 100 * roof["frac"]))
    same = bool(np.array_equal(ids["plain-fp32"], ids["bf16"]))
    best = "learned-bf16" if "learned-bf16" in forms else "bf16"
    # secondary figure: the host-buffer C-ABI call (queries and results crossing PCIe), headline form
    index.preSearch(learned=best == "learned-bf16", precision=HV_PRECISION_BF16, graph=False)
    index.search(xq, k)
    t0 = time.perf_counter()
    for _ in range(3):
        index.search(xq, k)
    h2h_ms = 1000.0 * (time.perf_counter() - t0) / 3
    forms[best]["host_to_host_qps"] = round(nq / (h2h_ms / 1000.0), 1)
    log("%s host-to-host (hv_search, PCIe included): %.3f ms, %.0f QPS" % (best, h2h_ms, nq / (h2h_ms / 1000.0)))
    index.preSearch(learned=False)
    ms = forms[best]["ms_per_step"]
    qps = forms[best]["qps"]

    base = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        base = cpu_baseline(index, xq, nprobe, k, B, args.cpu_budget)
        log("cpu baseline %.1f qps (%s) in %.1fs" % (base["value"], base["sample"], time.perf_counter() - t0))

    roof = dict(forms[best]["roofline"])
    tr = profiled_traffic(roof["kernel"])
    roof["traffic"] = tr["bytes_per_launch"] if tr else None
    if tr:
        roof["traffic_source"] = tr["source"]
    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32",
        "data": "synthetic",
        "config": {"workload": "int-mixture-1m: 1,000,000 x 128 fp32 base, 10,000 queries, nlist 1024, k 100, "
                               "B=4 dimension blocks, pruning on",
                   "form": best,
                   "nprobe": nprobe, "nb": int(xb.shape[0]), "nq": nq,
                   "recall_at_100": forms[best]["recall_at_100"], "recall_sweep_plain": recall_at,
                   "forms": forms,
                   "bf16_identical_to_plain": same,
                   "predictor": learned_info,
                   "speedup_vs_plain_gpu": round(qps / forms["plain-fp32"]["qps"], 3)},
        "roofline": roof,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
