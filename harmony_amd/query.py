"""Command-line harness over the library, with the reference query tool's flags and files
(src/query.cpp:48-166, 268-600, 680-777): python -m harmony_amd.query --benchmarks_path B
--dataset NAME [--k 100] [--nprobes 8 16 ...] [--mode original|base|block|brute] ...

Files, as query.cpp lays them out (query.cpp:281-300, 368-382):
  B/NAME/origin/NAME_base.<input_format>, NAME_query.<input_format>   the vectors
  B/NAME/result/groundtruth_<k>.<output_format>                        made when absent
  B/NAME/index/index_nlist_<nlist>_opt_<level>_subNprobeRatio_<r>.index  the reference's index file
  B/NAME/result/log.csv (or --csv)                                     one myToCsv row per search
Behaviour kept: nlist 0 -> sqrt(nb) (query.cpp:323-325); warmup_list_size 0 -> k when a warm-up
list is used (328-333); the last nprobe 0 -> nlist (351-353); the index file found by the opt-level
subset rule (368-382) and loaded with --cache, else built and saved; search time = mean of --loop
host-to-host searches after one untimed search when loop > 1 (630-639); recall and r2 by
calculate_recall / calculate_r2 against the ground truth; --fullWarmUp uses the ground truth's k-th
distance as every query's bound (716-725); --dataset_info writes "nb, nq, d" to --csv.
--metric ip runs the inner-product index (Index.cpp:1440-1551) in every mode as the FP32 exact scan,
with triangle pruning when the opt level has OPT_TRIANGLE; its ground truth when absent is that
scan over every list.
Modes on one GPU: original = the unpruned scan (Index.cpp:1420-1554); base = the per-node scan with
4 dimension blocks and pruning against the warm-up bound (node.cpp:606-678); block = the same with
--block dimension blocks; brute = every list (the exact FP32 kNN). group / team / worker / MPI
flags have no single-GPU meaning and are not offered. Differences: the ground truth when absent is
this library's FP64 brute force (the reference runs faiss IndexIVFFlat at nprobe = nlist; its
distances are written as FP32); --early_stop ends the nprobe sweep once recall reaches 1 (the
reference's flag is passed by value into doSearch and never takes effect, query.cpp:595, 676-678).
Extras: --precision fp32|bf16|fp8 and --rerank_r (the precision forms).
"""
import argparse
import math
import os
import sys
import time

import numpy as np

from . import harness as H
from . import metrics
from ._lib import HV_BOUND_EXTERNAL, HV_BOUND_WARMUP, HV_METRIC_IP, HV_METRIC_L2, HV_PRECISION_BF16, \
    HV_PRECISION_FP8, HV_PRECISION_FP32
from .index import Index

OPT_NAMES = {"OPT_NONE": 0, "OPT_TRIANGLE": 1, "OPT_SUBNN_L2": 2, "OPT_SUBNN_IP": 4, "OPT_TRI_SUBNN_L2": 3,
             "OPT_TRI_SUBNN_IP": 5, "OPT_SUBNN_ONLY": 6, "OPT_ALL": 7}
MODE_NAMES = {"original": "Original", "base": "Baseline", "block": "Dim     ", "brute": "Brute Force"}


def opt_level(s):
    """str2OptLevel (common.h:55-85): a number 0..7 or a level name."""
    try:
        v = int(s)
    except ValueError:
        if s not in OPT_NAMES:
            raise argparse.ArgumentTypeError("Invalid optimization level")
        return OPT_NAMES[s]
    if not 0 <= v <= 7:
        raise argparse.ArgumentTypeError("Invalid optimization level")
    return v


def parser():
    p = argparse.ArgumentParser(prog="harmony", description=__doc__.split("\n")[0])
    p.add_argument("--benchmarks_path", default="/home/xuqian/Triangle/benchmarks")
    p.add_argument("--dataset", default="msong")
    p.add_argument("--input_format", default="fvecs")
    p.add_argument("--output_format", default="bin")
    p.add_argument("--k", type=int, default=100)
    p.add_argument("--nprobes", type=int, nargs="*", default=[0])
    p.add_argument("--opt_levels", type=opt_level, nargs="*", default=[0])
    p.add_argument("--train_only", action="store_true")
    p.add_argument("--cache", action="store_true")
    p.add_argument("--sub_nprobe_ratio", type=float, default=1.0)
    p.add_argument("--metric", default="l2")
    p.add_argument("--loop", type=int, default=1)
    p.add_argument("--nlist", type=int, default=0)
    p.add_argument("--verbose", action="store_true")
    p.add_argument("--ratios", type=float, nargs="*", default=[1.0])
    p.add_argument("--csv", default="")
    p.add_argument("--dataset_info", action="store_true")
    p.add_argument("--early_stop", action="store_true")
    p.add_argument("--block", type=int, default=0)
    p.add_argument("--warmup_list_size", type=int, default=0)
    p.add_argument("--warmup_list", type=int, default=10)
    p.add_argument("--disablePruning", action="store_true")
    p.add_argument("--disableOrderOpt", action="store_true")
    p.add_argument("--fullWarmUp", action="store_true")
    p.add_argument("--mode", default="base", choices=sorted(MODE_NAMES))
    p.add_argument("--brute_force_search_ratio", type=float, default=1.0)
    p.add_argument("--kmeans_niter", type=int, default=20, help="Index::train's iterations (Index.cpp:389)")
    p.add_argument("--kmeans_seed", type=int, default=6666, help="Index::train's seed (Index.cpp:393)")
    p.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp8"])
    p.add_argument("--rerank_r", type=int, default=4)
    return p


def log(args, *a):
    if args.verbose:
        print(*a, flush=True)


def run(argv=None):
    args = parser().parse_args(argv)
    if args.early_stop and (args.ratios != [1.0]):
        raise ValueError("early_stop is only allowed when ratios is 1.0")
    if args.metric.lower() == "l2":
        metric = HV_METRIC_L2
    elif args.metric.lower() == "ip":
        metric = HV_METRIC_IP
    else:
        raise ValueError("Invalid metric type")
    k = args.k
    paths = H.harness_paths(args.benchmarks_path, args.dataset, args.input_format, args.output_format, k)
    log_path = args.csv or paths["log"]
    nb, d = H.xvecs_info(paths["base"])
    if args.dataset_info:
        nq, _ = H.xvecs_info(paths["query"])
        with open(args.csv, "w") as f:
            f.write("nb, nq, d\n%d, %d, %d\n" % (nb, nq, d))
        return 0
    nlist = args.nlist or int(math.sqrt(nb))
    wls = args.warmup_list_size
    if args.warmup_list > 0 and wls == 0:
        wls = k
    nprobes = list(args.nprobes)
    if nprobes[-1] == 0:
        nprobes[-1] = nlist
    added = 0
    for o in args.opt_levels:
        added |= o

    def index_path():
        for i in range(8):
            if (added & i) == added:
                p = H.harness_paths(args.benchmarks_path, args.dataset, nlist=nlist, opt=i,
                                    sub_nprobe_ratio=args.sub_nprobe_ratio)["index"]
                if os.path.exists(p):
                    return p
        return H.harness_paths(args.benchmarks_path, args.dataset, nlist=nlist, opt=added,
                               sub_nprobe_ratio=args.sub_nprobe_ratio)["index"]

    ipath = index_path()
    query = H.load_xvecs(paths["query"])
    nq = query.shape[0]
    print("[dim:%d, nb:%d, nq:%d, k:%d, worker:0]" % (d, nb, nq, k), flush=True)

    index = Index(d, nlist, nprobe=1, metric=metric, opt_level=added)
    train_s = add_s = 0.0
    if args.cache and os.path.exists(ipath):
        log(args, "Loading index from %s" % ipath)
        index.load_index(ipath)
    else:
        base = H.load_xvecs(paths["base"])
        t0 = time.perf_counter()
        index.train(base, niter=args.kmeans_niter, seed=args.kmeans_seed, max_points_per_centroid=256)
        train_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        index.add(base)
        add_s = time.perf_counter() - t0
        os.makedirs(os.path.dirname(ipath), exist_ok=True)
        index.save_index(ipath)
        log(args, "Index saved to %s" % ipath)
        del base
    if args.train_only:
        return 0

    # ground truth: the file, or the library's FP64 brute force (L2) / the exact inner-product
    # search over every list (IP) written in the reference's format
    if os.path.exists(paths["groundtruth"]):
        GI, GD = H.load_results(paths["groundtruth"], nq, k)
    else:
        index.nprobe = nlist
        if metric == HV_METRIC_L2:
            gD, GI = index.ground_truth_f64(query, k)
            GD = gD.astype(np.float32)
        else:
            index.opt_level = 0
            GD, GI = index.search(query, k)
        H.write_results(paths["groundtruth"], GI, GD)
        log(args, "Groundtruth file %s created" % paths["groundtruth"])

    precision = {"fp32": HV_PRECISION_FP32, "bf16": HV_PRECISION_BF16, "fp8": HV_PRECISION_FP8}[args.precision]
    mode = args.mode
    pruning = not args.disablePruning and mode in ("base", "block")
    blocks = args.block if mode == "block" and args.block > 0 else 4
    rows = []
    for nprobe in nprobes:
        stop = False
        for opt in args.opt_levels:
            for ratio in args.ratios:
                index.nprobe = nlist if mode == "brute" else nprobe
                index.opt_level = opt
                bound_mode = HV_BOUND_WARMUP
                if args.fullWarmUp and metric == HV_METRIC_L2:
                    index.set_external_bounds(GD[:, k - 1].copy())
                    bound_mode = HV_BOUND_EXTERNAL
                t0 = time.perf_counter()
                if metric == HV_METRIC_L2:
                    index.preSearch(warmup_list=args.warmup_list, warmup_list_size=wls, pruning=pruning,
                                    dim_blocks=blocks, bound_mode=bound_mode, precision=precision,
                                    rerank_r=args.rerank_r)
                presearch_s = time.perf_counter() - t0
                if args.loop > 1:
                    index.search(query, k)
                t0 = time.perf_counter()
                for _ in range(args.loop):
                    D, I = index.search(query, k)
                search_s = (time.perf_counter() - t0) / args.loop
                rec = metrics.recall(I, GI, D, GD, metric_l2=metric == HV_METRIC_L2)
                r2 = metrics.r2(I, D, GD)
                variance = 0.0
                if mode not in ("original", "brute"):
                    variance = H.probe_variance(index.last_probes()[:, :index.nprobe], nlist)
                row = {"dataset": args.dataset, "mode": MODE_NAMES[mode], "nb": nb, "nq": nq, "d": d,
                       "nlist": nlist, "nprobe": nprobe, "k": k, "divideIVF": False,
                       "orderOptimize": not args.disableOrderOpt, "UnblockSend": True, "cut": pruning,
                       "block": args.block, "worker": 0, "group": 1, "team": 1,
                       "time_speedup": H.time_speedup(0.0, search_s), "query_time": search_s,
                       "original_time": 0.0, "recall": rec, "r2": r2, "variance": variance,
                       "train_time": train_s, "add_time": add_s, "preSearch_time": presearch_s, "ratio": 0.0,
                       "teamRatio": 0.0, "brute_ratio": args.brute_force_search_ratio}
                H.append_log(log_path, row, append=True)
                rows.append(row)
                print("nprobe:%d opt_level: %d simi_ratio: %g\nquery_time: %.6f\nrecall: %g r2: %g"
                      % (nprobe, opt, ratio, search_s, rec, r2), flush=True)
                if args.early_stop and rec == 1:
                    stop = True
        if stop:
            break
    return rows


def main():
    r = run()
    sys.exit(0 if r == 0 or isinstance(r, list) else 1)


if __name__ == "__main__":
    main()
