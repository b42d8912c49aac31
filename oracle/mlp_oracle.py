"""CPU oracle of the learned probe/bound predictor (numpy, float64).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU predictor's forward pass, gradients,
Adam training and BF16 serving. The network has no counterpart in the reference (the reference
probes a fixed nprobe and takes its bound from a warm-up scan); what this restates is the
library's own definition of it:

  features f [65] -> xn = (f - fmean) / fstd
  h = relu(xn W1 + b1)         W1 [65][128], b1 [128]
  y = h W2 + b2                W2 [128][2],  b2 [2]     (standardized outputs)
  raw output = y * tstd + tmean
  loss over a batch of n rows with targets t: L = 1/(2n) sum_rows sum_o (y - (t - tmean)/tstd)^2
  parameters flat: W1 (input-major), b1, W2, b2

Pinned in tests/test_mlp_oracle.py against torch autograd (an independent implementation of
the same network and loss) and against central finite differences.
"""
import numpy as np

IN, HID, OUT = 65, 128, 2
NPARAMS = IN * HID + HID + HID * OUT + OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID)
    o += IN * HID
    b1 = p[o:o + HID]
    o += HID
    W2 = p[o:o + HID * OUT].reshape(HID, OUT)
    o += HID * OUT
    b2 = p[o:o + OUT]
    return W1, b1, W2, b2


def pack(W1, b1, W2, b2):
    return np.concatenate([W1.ravel(), b1.ravel(), W2.ravel(), b2.ravel()])


def forward(p, norm, X):
    """(standardized outputs y [n][2], hidden h [n][128], standardized inputs xn)."""
    fmean, fstd, tmean, tstd = norm
    W1, b1, W2, b2 = unpack(p)
    xn = (np.asarray(X, np.float64) - fmean) / fstd
    h = np.maximum(xn @ W1 + b1, 0.0)
    return h @ W2 + b2, h, xn


def loss_grad(p, norm, X, T):
    """(loss, flat gradient) of L = 1/(2n) sum (y - tn)^2."""
    fmean, fstd, tmean, tstd = norm
    n = X.shape[0]
    y, h, xn = forward(p, norm, X)
    tn = (np.asarray(T, np.float64) - tmean) / tstd
    e = y - tn
    loss = 0.5 * np.sum(e * e) / n
    dy = e / n
    W1, b1, W2, b2 = unpack(p)
    gW2 = h.T @ dy
    gb2 = dy.sum(axis=0)
    dh = (dy @ W2.T) * (h > 0)
    gW1 = xn.T @ dh
    gb1 = dh.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2)


def predict_raw(p, norm, X):
    """Raw-target outputs (log1p(k-th distance), log(nprobe))."""
    y, _, _ = forward(p, norm, X)
    return y * norm[3] + norm[2]


def bf16_round(a):
    """float32 -> bfloat16 -> float64, round to nearest even (finite inputs)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def predict_raw_bf16(p, norm, X):
    """Serving form: W1 and W2 rounded to BF16, biases and arithmetic in full precision."""
    W1, b1, W2, b2 = unpack(p)
    q = pack(bf16_round(W1), b1, bf16_round(W2), b2)
    return predict_raw(q, norm, X)


def _splitmix(seed):
    state = [seed & 0xFFFFFFFFFFFFFFFF]
    M = 0xFFFFFFFFFFFFFFFF

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & M
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    return nxt


def train_adam(p, norm, X, T, epochs, batch, lr, seed, beta1=0.9, beta2=0.999, eps=1e-8):
    """The library's schedule: per epoch a Fisher-Yates shuffle from SplitMix64(seed) carried
    across epochs (i from n-1 down to 1, j = next() mod (i+1)), then Adam over the full batches
    of the permutation. Returns (params, per-epoch mean batch loss)."""
    p = np.array(p, dtype=np.float64)
    n = X.shape[0]
    B = min(batch, n)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    perm = list(range(n))
    nxt = _splitmix(seed)
    t = 0
    losses = []
    for _ in range(epochs):
        for i in range(n - 1, 0, -1):
            j = nxt() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        tot, nb = 0.0, 0
        for b0 in range(0, n - B + 1, B):
            idx = perm[b0:b0 + B]
            loss, g = loss_grad(p, norm, X[idx], T[idx])
            t += 1
            m = beta1 * m + (1 - beta1) * g
            v = beta2 * v + (1 - beta2) * g * g
            mh = m / (1 - beta1 ** t)
            vh = v / (1 - beta2 ** t)
            p = p - lr * mh / (np.sqrt(vh) + eps)
            tot += loss
            nb += 1
        losses.append(tot / max(1, nb))
    return p, np.array(losses)


def features_from_coarse(pdist_sorted, q):
    """Features from ascending squared coarse distances [n][P] and the queries [n][d]."""
    pd = np.asarray(pdist_sorted, np.float64)
    n, P = pd.shape
    cols = np.minimum(np.arange(64), P - 1)
    f = np.empty((n, IN), np.float64)
    f[:, :64] = np.log1p(np.maximum(pd[:, cols], 0.0))
    qq = np.asarray(q, np.float64)
    f[:, 64] = np.log1p(np.sum(qq * qq, axis=1))
    return f
