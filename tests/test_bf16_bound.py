"""CPU check of the BF16 form's certified error bound against real rounding.

The MFMA filter estimates |q - x|^2 = |rq|^2 + |rx|^2 - 2 rq.rx from BF16 residuals and drops a
pair when est - err exceeds the query's bound, so the form is exact only if
    est - err <= chain(q, x) <= est + err
for every pair, where chain is the FP32 direct-difference chain the plain path computes. The
bound (hv_bf16.hip) is
    err = kappa1 (|eq| |rx| + (|rq| + |eq|) |ex|) + kappa_g (|rq| + |eq|) (|rx| + |ex|)
          + kappa2 (|rq|^2 + |rx|^2)
with |eq|, |ex| the measured norms of the BF16 rounding errors of the residuals (+ 2^-23 |r|).
Here BF16 round-to-nearest-even, exact BF16 products with FP32 accumulation and the FP32 chain
are reproduced in numpy, and the inequality is checked on random data and on adversarial pairs
where one residual is parallel to the other's rounding error (Cauchy-Schwarz with equality: the
cross term of the bound is reached), which a bound short by its factor 2 fails."""
import numpy as np
import pytest

U = 2.0 ** -24


def bf16(a):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def chain(a, b):
    """FP32 chain sum_k fma(a_k - b_k, a_k - b_k, acc) over the last axis (one block)."""
    acc = np.zeros(a.shape[:-1], np.float32)
    for k in range(a.shape[-1]):
        t = (a[..., k] - b[..., k]).astype(np.float32)
        acc = (acc.astype(np.float64) + t.astype(np.float64) ** 2).astype(np.float32)
    return acc


def err_norm(r):
    """|r - bf16(r)| + 2^-23 |r| in FP32, as the kernels compute it."""
    e = (r - bf16(r)).astype(np.float32)
    return (np.sqrt(chain(e, np.zeros_like(e))) + np.float32(2.0 ** -23) * np.sqrt(chain(r, np.zeros_like(r)))
            ).astype(np.float32)


def kappas(d, B, factor=2.0):
    gd = d * U / (1.0 - d * U)
    return (np.float32(1.01 * factor), np.float32(1.01 * 2.0 * 2.0 * gd), np.float32(1.01 * (4.0 * d + 2.0 * B + 32.0) * U))


def split_err_norm(r):
    """|r - hi - lo| + 2^-23 |r| in FP32 for the split r ~ hi + lo (hi = bf16(r), lo = bf16(r - hi))."""
    hi = bf16(r)
    lo = bf16((r - hi).astype(np.float32))
    e = ((r - hi).astype(np.float32) - lo).astype(np.float32)
    return (np.sqrt(chain(e, np.zeros_like(e))) + np.float32(2.0 ** -23) * np.sqrt(chain(r, np.zeros_like(r)))
            ).astype(np.float32)


def bounds(q, x, c, B=4, factor=2.0, split=False):
    """(lb, ub, exact chain distance) of the pairs (q[i], x[i]) in the list with centroid c. With
    split, the query residual is carried as two BF16 parts hi + lo (the nearest-list pass for
    d <= 256): both are multiplied into the codes, 2d exact products accumulated in FP32, and the
    accumulation term of the bound is taken 2.1 times."""
    d = q.shape[1]
    rq = (q - c).astype(np.float32)
    rx = (x - c).astype(np.float32)
    nrq = chain(q, np.broadcast_to(c, q.shape))
    nx = chain(x, np.broadcast_to(c, x.shape))
    bq, bx = bf16(rq), bf16(rx)
    lo = bf16((rq - bq).astype(np.float32)) if split else np.zeros_like(rq)
    dot = np.zeros(q.shape[0], np.float32)
    for k in range(d):  # exact BF16 products, FP32 accumulation
        dot = (dot + bq[:, k] * bx[:, k]).astype(np.float32)
        if split:
            dot = (dot + lo[:, k] * bx[:, k]).astype(np.float32)
    eq, ex = (split_err_norm(rq) if split else err_norm(rq)), err_norm(rx)
    k1, kg, k2 = kappas(d, B, factor)
    if split:
        kg = np.float32(2.1) * kg
    sq, sx = np.sqrt(nrq), np.sqrt(nx)
    sqp = (sq + eq).astype(np.float32)
    err = (k1 * (eq * sx + sqp * ex) + kg * sqp * (sx + ex) + k2 * (nrq + nx)).astype(np.float32)
    est = (nrq + nx - np.float32(2.0) * dot).astype(np.float32)
    return np.maximum(est - err, 0), est + err, chain(q, x), err


def adversarial(n, d, seed, scale):
    """Residuals whose BF16 rounding errors are near half an ulp, paired with residuals exactly
    representable in BF16 and parallel to those errors (c = 0, so residuals are the rows)."""
    rng = np.random.default_rng(seed)
    base = bf16(rng.uniform(0.5, 2.0, size=(n, d)).astype(np.float32) * scale)
    ulp = base.view(np.uint32).astype(np.int64)
    # a value just short of the midpoint to the next BF16 value: rounding error ~ half an ulp
    q = (ulp + 0x7FFF).astype(np.uint32).view(np.float32)
    q = q * rng.choice(np.array([-1.0, 1.0], np.float32), size=(n, d))
    e = q - bf16(q)
    x = bf16(e * np.float32(2.0 ** 8))  # parallel to the rounding error, ex = 0
    c = np.zeros(d, np.float32)
    return q.astype(np.float32), x.astype(np.float32), c


@pytest.mark.parametrize("d", [8, 33, 128])
@pytest.mark.parametrize("scale", [1e-3, 1.0, 3e4])
def test_bound_holds_on_random_data(d, scale):
    rng = np.random.default_rng(d)
    c = (rng.standard_normal(d) * scale).astype(np.float32)
    q = (c + rng.standard_normal((400, d)) * scale).astype(np.float32)
    x = (c + rng.standard_normal((400, d)) * scale * rng.uniform(0.01, 3, size=(400, 1))).astype(np.float32)
    lb, ub, exact, err = bounds(q, x, c)
    assert np.all(lb <= exact) and np.all(exact <= ub)
    # measured rounding errors: well inside the worst-case bound 4 * 2^-8 |rq| |rx|
    rq, rx = np.linalg.norm(q - c, axis=1), np.linalg.norm(x - c, axis=1)
    assert np.median(err / (4 * 2.0 ** -8 * rq * rx)) < 0.6


@pytest.mark.parametrize("d", [16, 128])
@pytest.mark.parametrize("swap", [False, True])
def test_bound_holds_on_aligned_rounding(d, swap):
    q, x, c = adversarial(300, d, seed=d, scale=1.0)
    if swap:
        q, x = x, q
    lb, ub, exact, _ = bounds(q, x, c)
    assert np.all(lb <= exact) and np.all(exact <= ub)
    # the cross term is reached: a bound without its factor 2 is violated
    lb1, ub1, exact1, _ = bounds(q, x, c, factor=1.0)
    assert np.any(lb1 > exact1) or np.any(exact1 > ub1)


@pytest.mark.parametrize("d", [8, 33, 128, 256])
@pytest.mark.parametrize("scale", [1e-3, 1.0, 3e4])
def test_split_query_bound_holds_and_halves(d, scale):
    """The query residual as hi + lo: the bound still holds, and with the query's rounding term
    gone it is about half the single-part bound."""
    rng = np.random.default_rng(d + 7)
    c = (rng.standard_normal(d) * scale).astype(np.float32)
    q = (c + rng.standard_normal((400, d)) * scale).astype(np.float32)
    x = (c + rng.standard_normal((400, d)) * scale * rng.uniform(0.01, 3, size=(400, 1))).astype(np.float32)
    lb, ub, exact, err_s = bounds(q, x, c, split=True)
    assert np.all(lb <= exact) and np.all(exact <= ub)
    _, _, _, err = bounds(q, x, c)
    assert np.median(err_s / err) < 0.65


@pytest.mark.parametrize("d", [16, 128])
def test_split_query_bound_holds_on_aligned_rounding(d):
    """Codes whose rounding error is near half an ulp, queries parallel to it: the code's term
    (|rq| + |eq|) |ex| is reached and the split bound must still cover it."""
    q, x, c = adversarial(300, d, seed=d + 1, scale=1.0)
    q, x = x, q
    lb, ub, exact, _ = bounds(q, x, c, split=True)
    assert np.all(lb <= exact) and np.all(exact <= ub)
