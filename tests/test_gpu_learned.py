"""GPU tests of the learned probe/bound predictor: the library's forward, gradient, Adam and BF16
serving kernels against the CPU oracle (oracle/mlp_oracle.py), the query features against the
oracle's coarse distances, the learned search's results against the oracle at each query's
probe count, and what the trained network gains over the plain path on the 1M workload."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, learned
from harmony_amd.index import Index
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test learned]", *a, flush=True)
    sys.stdout.flush()


def rand_model(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15, M.NPARAMS).astype(np.float32)
    norm = (rng.normal(5, 1, M.IN).astype(np.float32), rng.uniform(0.5, 2.0, M.IN).astype(np.float32),
            rng.normal(3, 1, M.OUT).astype(np.float32), rng.uniform(0.5, 2.0, M.OUT).astype(np.float32))
    return p, norm


def rand_batch(seed, n):
    rng = np.random.default_rng(seed)
    return (rng.normal(5, 2, (n, M.IN)).astype(np.float32), rng.normal(3, 1.5, (n, M.OUT)).astype(np.float32))


@pytest.fixture(scope="module")
def w1():
    say("workload 1: 10000 x 128, nlist 100")
    xb, xq = datagen.make(1)
    g = Index(128, 100, nprobe=10)
    g.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    g.add(xb)
    return xb, xq, g


@pytest.mark.parametrize("seed,n", [(0, 1), (1, 64), (2, 300)])
def test_gradient_matches_oracle(w1, seed, n):
    _, _, g = w1
    p, norm = rand_model(seed)
    X, T = rand_batch(seed + 10, n)
    g.mlp_set_norm(*norm)
    g.mlp_set_params(p)
    loss, grad = g.mlp_grad(X, T)
    normd = tuple(a.astype(np.float64) for a in norm)
    oloss, ograd = M.loss_grad(p.astype(np.float64), normd, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - oloss) <= 1e-5 * max(1.0, oloss)
    scale = np.abs(ograd).max()
    np.testing.assert_allclose(grad, ograd, rtol=1e-4, atol=1e-6 * scale)


def test_predict_bf16_matches_oracle(w1):
    _, _, g = w1
    p, norm = rand_model(4)
    X, _ = rand_batch(5, 513)
    g.mlp_set_norm(*norm)
    g.mlp_set_params(p)
    out = g.mlp_predict(X)
    normd = tuple(a.astype(np.float64) for a in norm)
    ref = M.predict_raw_bf16(p, normd, X)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-4)
    # the BF16 serving weights differ from FP32 ones: the FP32 prediction is measurably different
    assert np.abs(M.predict_raw(p, normd, X) - ref).max() > 1e-4


def test_adam_step_and_training_match_oracle(w1):
    _, _, g = w1
    p, norm = rand_model(6)
    normd = tuple(a.astype(np.float64) for a in norm)
    X, T = rand_batch(7, 256)
    # one step over the whole set: the oracle's Adam step from the library's own gradient
    g.mlp_set_norm(*norm)
    g.mlp_set_params(p)
    _, grad = g.mlp_grad(X, T)
    g.mlp_train(X, T, epochs=1, batch=256, lr=1e-3, seed=3)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    gd = grad.astype(np.float64)
    m, v = (1 - b1) * gd, (1 - b2) * gd * gd
    step = p.astype(np.float64) - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    np.testing.assert_allclose(g.mlp_get_params(), step, rtol=1e-5, atol=1e-7)
    # several epochs of minibatches: the loss trajectory of the oracle's schedule
    g.mlp_set_params(p)
    losses = g.mlp_train(X, T, epochs=6, batch=32, lr=3e-3, seed=5)
    _, olosses = M.train_adam(p.astype(np.float64), normd, X.astype(np.float64), T.astype(np.float64), epochs=6,
                              batch=32, lr=3e-3, seed=5)
    np.testing.assert_allclose(losses, olosses, rtol=2e-3)
    assert losses[-1] < 0.7 * losses[0]


def test_query_features_match_oracle(w1):
    """The distance features are the coarse step's upper bounds d~ + kappa (|q|^2 + |c|^2) of the
    chain distances (d~ the MFMA estimate, |d~ - chain| <= kappa (|q|^2 + |c|^2)), sorted: each
    lies between the exact distance of its rank and that plus 2 kappa (|q|^2 + max |c|^2)
    (sorting moves no value by more than the largest shift), so its log1p is at least the
    oracle's and above it by at most that over 1 + distance; the norm feature is the oracle's to
    rounding."""
    _, xq, g = w1
    f = g.query_features(xq).astype(np.float64)
    _, pd = O.coarse(xq, g.centroids(), 64)
    ref = M.features_from_coarse(pd, xq).astype(np.float64)
    d = xq.shape[1]
    kappa = 1.02 * ((16.2 * d + 8) / 2.0 ** 24 + 3.1 * 2.0 ** -16)  # coarse_kappa (split BF16 product)
    c = g.centroids().astype(np.float64)
    err = kappa * ((xq.astype(np.float64) ** 2).sum(1) + (c ** 2).sum(1).max())
    pd64 = pd.astype(np.float64)
    eps = 4e-7 * np.abs(ref[:, :64]) + 1e-7  # FP32 log1p and chain roundings
    diff = f[:, :64] - ref[:, :64]
    assert np.all(diff >= -eps)
    assert np.all(diff <= 2 * err[:, None] / (1 + pd64) + eps)
    np.testing.assert_allclose(f[:, 64], ref[:, 64], rtol=0, atol=2e-6 * np.abs(ref).max())
    assert np.all(np.diff(f[:, :64], axis=1) >= 0)
    say("features: largest shift %.3g (bound %.3g)" % (diff.max(), (2 * err).max()))


@pytest.fixture(scope="module")
def w1_model(w1):
    xb, xq, g = w1
    say("w1: traces of 3000 training queries")
    _, xt = datagen.make(1, nq=3000, query_set=1)
    feats, targets, exact = learned.collect_traces(g, xt, k=10)
    losses = learned.train(g, feats, targets, epochs=30, batch=256, lr=3e-3, seed=0)
    say("w1 predictor loss %.4f -> %.4f" % (losses[0], losses[-1]))
    assert losses[-1] < losses[0]
    return feats, targets, exact


def test_serving_outputs_follow_predictions(w1, w1_model):
    """The search's predictor gives each query qn = clamp(ceil(exp(y1) * s_np), min, nprobe), raised
    to the shortest prefix of its probe list that holds k vectors, and seed = expm1(max(y0, 0)) *
    s_b from the same network outputs as hv_mlp_predict on the query's features."""
    xb, xq, g = w1
    g.nprobe = 10
    s_np, s_b = 1.3, 1.1
    g.preSearch(learned=True, learned_nprobe_scale=s_np, learned_bound_scale=s_b, learned_min_nprobe=2,
                precision=HV_PRECISION_BF16, dim_blocks=4, learned_margin=0.0)
    g.search(xq, 10)
    qn = g.last_nprobe(xq.shape[0])
    seed = g.last_bounds(xq.shape[0])
    g.preSearch(learned=False, precision=HV_PRECISION_FP32, learned_min_nprobe=1, learned_margin=1.0)
    probes = g.last_probes(xq.shape[0])[:, :10]
    y = g.mlp_predict(g.query_features(xq)).astype(np.float64)
    np_f = np.exp(y[:, 1]) * s_np
    ref_qn = np.where(np_f >= 10, 10, np.maximum(2, np.ceil(np_f))).clip(1, 10).astype(np.int32)
    # then at least as many lists as hold k = 10 vectors
    held = np.cumsum(np.diff(g.list_offsets())[probes], axis=1)
    cover = np.minimum(np.argmax(held >= 10, axis=1) + 1, 10)
    cover[held[:, -1] < 10] = 10
    ref_qn = np.maximum(ref_qn, cover)
    away = np.abs(np_f - np.round(np_f)) > 1e-4  # not at an integer (last-bit exp differences)
    np.testing.assert_array_equal(qn[away], ref_qn[away])
    ref_seed = np.expm1(np.maximum(y[:, 0], 0.0)) * s_b
    np.testing.assert_allclose(seed, ref_seed.astype(np.float32), rtol=2e-6)


@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
def test_learned_without_trimming_equals_plain(w1, w1_model, precision):
    """Every query keeps all nprobe lists; the seeded bound and the fallback leave the results
    exactly the plain path's."""
    xb, xq, g = w1
    g.nprobe = 10
    g.preSearch(learned=False, precision=precision, dim_blocks=4)
    D0, I0 = g.search(xq, 10)
    for bscale in (1.0, 1e-6, 0.0):  # predicted bound, a bound far too low (all fall back), none
        g.preSearch(learned=True, learned_nprobe_scale=1e6, learned_bound_scale=bscale)
        D1, I1, st = g.search(xq, 10, return_stats=True)
        np.testing.assert_array_equal(I1, I0)
        np.testing.assert_array_equal(D1, D0)
        assert np.all(g.last_nprobe(xq.shape[0]) == 10)
        if bscale == 1e-6:
            assert st["fallback_queries"] == xq.shape[0]
        if bscale == 0.0:
            assert st["fallback_queries"] == 0
    g.preSearch(learned=False, precision=HV_PRECISION_FP32)


@pytest.mark.parametrize("margin", [0.0, 1.0])
@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
def test_learned_per_query_probes_match_oracle(w1, w1_model, precision, margin):
    """Trimmed probe lists: each query's results are the oracle's at that query's probe count (the
    count reported includes the margin widening, and nprobe for a query redone exactly)."""
    xb, xq, g = w1
    o = O.IndexData.from_gpu_index(g)
    g.nprobe = 10
    g.preSearch(learned=True, learned_nprobe_scale=1.0, learned_bound_scale=1.2, precision=precision,
                dim_blocks=4, learned_margin=margin)
    D, I, st = g.search(xq, 10, return_stats=True)
    qn = g.last_nprobe(xq.shape[0])
    say("w1 learned probes: mean %.2f, min %d, max %d, fallbacks %d" % (qn.mean(), qn.min(), qn.max(),
                                                                          st["fallback_queries"]))
    assert qn.min() >= 1 and qn.max() <= 10 and qn.mean() < 10
    for p in np.unique(qn):
        sel = np.nonzero(qn == p)[0]
        oD, oI, _, _ = O.search(o, xq[sel], 10, nprobe=int(p), dim_blocks=4)
        np.testing.assert_array_equal(I[sel], oI)
        np.testing.assert_array_equal(D[sel], oD)
    g.preSearch(learned=False, precision=HV_PRECISION_FP32, learned_margin=1.0)


def list_radii(o):
    """The largest FP64 |x - c| of every list (the oracle's index arrays)."""
    r = np.zeros(o.nlist)
    for l in range(o.nlist):
        a, b = o.list_off[l], o.list_off[l + 1]
        if b > a:
            r[l] = np.sqrt(((o.codes[a:b].astype(np.float64) - o.centroids[l]) ** 2).sum(1).max())
    return r


@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
def test_learned_margin_gives_plain_results(w1, w1_model, precision):
    """With the margin test on (margin 1) every learned result is the plain search's at nprobe, even
    with the predicted counts scaled far down: the widened counts cover every list the triangle
    inequality cannot exclude within the seed, and queries whose k-th result exceeds their seed
    are redone over all nprobe lists. Each reported count is at least the FP64 count of lists
    whose lower bound (the larger of |q - c| - radius and the distance to the bisector of c and
    the query's nearest centroid) is within the seed's square root, and the search still probes
    fewer lists than nprobe on average."""
    xb, xq, g = w1
    o = O.IndexData.from_gpu_index(g)
    g.nprobe = 10
    oD, oI, oprobes, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    for s_np, s_b in ((0.25, 1.2), (0.25, 3.0), (1.0, 1.0)):
        g.preSearch(learned=True, learned_nprobe_scale=s_np, learned_bound_scale=s_b, precision=precision,
                    dim_blocks=4, learned_margin=1.0)
        D, I, st = g.search(xq, 10, return_stats=True)
        qn = g.last_nprobe(xq.shape[0])
        seed = g.last_bounds(xq.shape[0])
        say("margin 1, scales %.2f / %.1f: mean probes %.2f, widened %d, redone %d" % (
            s_np, s_b, qn.mean(), st["widened_queries"], st["fallback_queries"]))
        np.testing.assert_array_equal(I, oI)
        np.testing.assert_array_equal(D, oD)
        # FP64 restatement of the widening rule, without the FP32 slack: a lower bound for the count
        rad = list_radii(o)
        c64 = o.centroids.astype(np.float64)
        cd2 = ((xq[:, None, :].astype(np.float64) - c64[oprobes]) ** 2).sum(-1)
        cc = np.sqrt(((c64[oprobes] - c64[oprobes[:, :1]]) ** 2).sum(-1))
        with np.errstate(divide="ignore", invalid="ignore"):
            vor = np.where(cc > 0, (cd2 - cd2[:, :1]) / (2 * cc), 0.0)
        lb = np.maximum(np.sqrt(cd2) - rad[oprobes], vor)
        need = (lb <= 0) | (lb ** 2 <= seed[:, None].astype(np.float64))
        last = np.where(need.any(1), 10 - np.argmax(need[:, ::-1], axis=1), 1)
        assert np.all((qn >= last) | (qn == 10))
        assert qn.mean() < 10
    g.preSearch(learned=False, precision=HV_PRECISION_FP32)


def test_learned_fp8_wide_shortlist_matches_oracle(w1, w1_model):
    """Learned probe counts with the FP8 form: a shortlist wider than any query's candidates
    (k * r = 3000) filters nothing, so each query's results are the oracle's at its probe count."""
    xb, xq, g = w1
    o = O.IndexData.from_gpu_index(g)
    g.nprobe = 10
    g.preSearch(learned=True, learned_nprobe_scale=1.0, learned_bound_scale=1.2, precision=HV_PRECISION_FP8,
                rerank_r=300, dim_blocks=4)
    D, I = g.search(xq, 10)
    qn = g.last_nprobe(xq.shape[0])
    g.preSearch(learned=False, precision=HV_PRECISION_FP32, rerank_r=4)
    assert qn.min() >= 1 and qn.max() <= 10 and qn.mean() < 10
    for p in np.unique(qn):
        sel = np.nonzero(qn == p)[0]
        oD, oI, _, _ = O.search(o, xq[sel], 10, nprobe=int(p), dim_blocks=4)
        np.testing.assert_array_equal(I[sel], oI)
        np.testing.assert_array_equal(D[sel], oD)


@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
@pytest.mark.parametrize("bscale", [1.2, 0.5])
def test_learned_graph_replays_match_eager(w1, w1_model, precision, bscale):
    """A learned search captured into one HIP graph (second call) and replayed (later calls, with
    host work between) returns the eager results and probe counts; with the lower bound scale
    some queries fall back, so the exact redo runs inside the graph too."""
    xb, xq, g = w1
    n = xq.shape[0]
    g.nprobe = 10
    kw = dict(learned=True, learned_nprobe_scale=1.0, learned_bound_scale=bscale, precision=precision,
              dim_blocks=4)
    g.preSearch(graph=False, **kw)
    D0, I0, st0 = g.search(xq, 10, return_stats=True)
    qn0 = g.last_nprobe(n)
    g.preSearch(graph=True, **kw)
    for rep in range(4):
        D, I, st = g.search(xq, 10, return_stats=True)
        np.testing.assert_array_equal(I, I0)
        np.testing.assert_array_equal(D, D0)
        np.testing.assert_array_equal(g.last_nprobe(n), qn0)
        assert st["fallback_queries"] == st0["fallback_queries"]
        _ = np.sort(xq[: 100 * (rep + 1)], axis=0)  # the program's own work between replays
    say("learned graph, precision %d, bound scale %.1f: %d fallbacks" % (precision, bscale,
                                                                           st0["fallback_queries"]))
    g.preSearch(learned=False, precision=HV_PRECISION_FP32, graph=False)


def test_learned_gain_on_1m(tmp_path):
    """What the trained network gains over the method without it, at equal recall: on held-out
    queries of the 1M workload, the BF16 form at the smallest fixed nprobe whose recall@100
    (against the exact neighbours) reaches a target, against the learned BF16 form (cap 16,
    margin test on, calibrated on the other half of the held-out set) at the same target. The
    learned form must reach at least the fixed form's recall with fewer candidates."""
    say("workload 2: 1,000,000 x 128, nlist 1024")
    xb, xq = datagen.make(2)
    g = Index(128, 1024, nprobe=16)
    g.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    g.add(xb)
    del xb
    say("w2: traces of 50,000 training queries and the 10,000 held-out ones")
    _, xt = datagen.make(2, nb=1, nq=50000, query_set=1)
    feats, targets, _ = learned.collect_traces(g, xt, k=100)
    vf, vt, vexact = learned.collect_traces(g, xq, k=100)
    losses = learned.train(g, feats, targets, epochs=20, batch=1024, lr=2e-3, seed=0)
    say("w2 predictor loss %.4f -> %.4f" % (losses[0], losses[-1]))
    h = xq.shape[0] // 2
    for target in (0.95, 0.99):
        # the method without the network: the smallest fixed nprobe reaching the target
        fixed = None
        for npb in (1, 2, 3, 4, 6, 8, 12, 16):
            g.preSearch(learned=False, precision=HV_PRECISION_BF16)
            g.nprobe = npb
            _, I0, st0 = g.search(xq[h:], 100, return_stats=True)
            r0 = learned.recall_per_query(I0, vexact[h:]).mean()
            if r0 >= target:
                fixed = (npb, r0, st0)
                break
        assert fixed is not None
        s, bs, rec_cal, probes_cal = learned.calibrate(g, vf[:h], vt[:h], xq[:h], vexact[:h], nprobe=16, k=100,
                                                       target_recall=target, margin=0.003)
        g.preSearch(learned=True, learned_nprobe_scale=s, learned_bound_scale=bs, precision=HV_PRECISION_BF16)
        g.nprobe = 16
        _, I1, st1 = g.search(xq[h:], 100, return_stats=True)
        r1 = learned.recall_per_query(I1, vexact[h:]).mean()
        probes = g.last_nprobe().mean()
        say("w2 target %.2f: fixed nprobe %d recall %.4f candidates %d | learned (scale %.2f, margin %.1f) recall "
            "%.4f candidates %d probes %.3f widened %d redone %d"
            % (target, fixed[0], fixed[1], fixed[2]["candidates"], s, g._params.learned_margin, r1,
               st1["candidates"], probes, st1["widened_queries"], st1["fallback_queries"]))
        assert r1 >= target
        if fixed[0] > 1:  # at nprobe 1 there is nothing to trim
            assert r1 >= fixed[1]
            assert st1["candidates"] < 0.8 * fixed[2]["candidates"]
    g.preSearch(learned=False)
