"""CPU check of the coarse step's error bound (hv_coarse.hip, coarse_kappa): the estimate
d~ = |q|^2 + |c|^2 - 2 q.c with q.c taken as the split BF16 product qh.ch + qh.cl + ql.ch (v = hi +
lo + e, hi and lo BF16 roundings) accumulated in FP32 in any order, against the FP32 chain
fma(q - c, q - c, acc) over the dimensions in order, must satisfy
|d~ - chain| <= kappa (|q|^2 + |c|^2), kappa = 1.02 ((16.2 d + 8) u + 3.1 2^-16).
The emulation rounds to BF16 the way the GPU's converter does (to nearest, ties to even), forms
the products exactly in float64 and accumulates them in float32 in a random order. The
split-product share of the bound, (6.1 d u + 3.1 2^-16) 2 |q||c|, is checked on its own too, on
data whose left-out parts all have one sign (the share's 2^-16 term is needed there: without it the
check fails at small d)."""
import numpy as np
import pytest

U = 2.0 ** -24


def kappa(d):
    return 1.02 * ((16.2 * d + 8.0) * U + 3.1 * 2.0 ** -16)


def bf16(x):
    """Round float32 values to BF16 (to nearest, ties to even), returned as float32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def split(v):
    hi = bf16(v)
    lo = bf16((v - hi).astype(np.float32))  # v - hi is exact in float32
    return hi, lo


def f32_sum(terms, rng):
    """float32 accumulation of the float64 terms (each rounded to float32 first: the BF16
    products are exact in float32) in a random order."""
    acc = np.float32(0.0)
    for t in terms[rng.permutation(len(terms))]:
        acc = np.float32(acc + np.float32(t))
    return acc


def f32_norm(v):
    acc = np.float32(0.0)
    for x in v:
        acc = np.float32(np.float64(x) * np.float64(x) + np.float64(acc))  # one rounding per step
    return acc


def chain(q, c):
    acc = 0.0
    for a, b in zip(q, c):
        u = np.float64(np.float32(a - b))
        acc = float(np.float32(u * u + acc))  # fma: one rounding
    return acc


def split_dot_terms(q, c):
    qh, ql = split(q)
    ch, cl = split(c)
    q64 = [x.astype(np.float64) for x in (qh, ql)]
    c64 = [x.astype(np.float64) for x in (ch, cl)]
    return np.concatenate([q64[0] * c64[0], q64[0] * c64[1], q64[1] * c64[0]])


def estimate(q, c, rng):
    dot = f32_sum(split_dot_terms(q, c), rng)
    return float(np.float32(np.float32(f32_norm(q) + f32_norm(c)) - np.float32(2.0) * dot))


def cases(rng, d):
    yield "gaussian", rng.normal(0, 1, d), rng.normal(0, 1, d)
    yield "integer", rng.integers(0, 256, d).astype(np.float64), rng.integers(0, 256, d).astype(np.float64)
    x = rng.normal(5, 1, d)
    yield "near", x, x + rng.normal(0, 1e-3, d)
    scale = np.exp(rng.normal(0, 2, d))
    yield "wide range", rng.normal(0, 1, d) * scale, rng.normal(0, 1, d) * scale
    # every component's rounding errors of one sign: values just above 1 + k 2^-7 + 2^-9
    k = rng.integers(0, 128, d)
    yield "aligned", 1.0 + k / 128.0 + 2.0 ** -9 + 2.0 ** -17, 1.0 + k / 128.0 + 2.0 ** -9 + 2.0 ** -17


@pytest.mark.parametrize("d", [8, 32, 128, 420])
def test_coarse_bound_holds(d):
    rng = np.random.default_rng(d)
    for name, q, c in cases(rng, d):
        q = q.astype(np.float32)
        c = c.astype(np.float32)
        for _ in range(3):
            e = estimate(q, c, rng)
            ch = chain(q, c)
            nn = float((q.astype(np.float64) ** 2).sum() + (c.astype(np.float64) ** 2).sum())
            assert abs(e - ch) <= kappa(d) * nn, (name, d, e, ch, kappa(d) * nn)


@pytest.mark.parametrize("d", [8, 16, 128])
def test_split_product_share(d):
    """|split dot - q.c| <= (6.1 d u + 3.1 2^-16) |q||c| (the bound's share for the product, before
    its factor 2); on the aligned data the left-out parts alone exceed the accumulation term's
    6.1 d u at small d, so dropping the 2^-16 term would fail here."""
    rng = np.random.default_rng(7 + d)
    k = rng.integers(0, 128, d)
    v = (1.0 + k / 128.0 + 2.0 ** -9 + 2.0 ** -17).astype(np.float32)
    q, c = v, v.copy()
    exact = float((q.astype(np.float64) * c.astype(np.float64)).sum())
    qc = float(np.sqrt((q.astype(np.float64) ** 2).sum() * (c.astype(np.float64) ** 2).sum()))
    worst = 0.0
    for _ in range(5):
        err = abs(float(f32_sum(split_dot_terms(q, c), rng)) - exact)
        worst = max(worst, err)
        assert err <= (6.1 * d * U + 3.1 * 2.0 ** -16) * qc
    if d <= 16:
        assert worst > 6.1 * d * U * qc  # the 2^-16 term is needed
