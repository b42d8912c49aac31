"""Pins the CPU oracle of the learned predictor: forward and gradient against torch autograd
(an independent implementation of the same network and loss), gradient against central finite
differences, BF16 rounding against torch's float32 -> bfloat16 conversion, and the Adam
schedule against torch.optim.Adam on the same batches."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def rand_case(seed, n=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, M.NPARAMS)
    norm = (rng.normal(0, 1, M.IN), rng.uniform(0.5, 2.0, M.IN), rng.normal(0, 1, M.OUT), rng.uniform(0.5, 2.0, M.OUT))
    X = rng.normal(0, 1.5, (n, M.IN))
    T = rng.normal(0, 1.0, (n, M.OUT))
    return p, norm, X, T


def torch_loss(p, norm, X, T):
    fmean, fstd, tmean, tstd = [torch.tensor(a, dtype=torch.float64) for a in norm]
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = pt[:M.IN * M.HID].view(M.IN, M.HID)
    b1 = pt[M.IN * M.HID:M.IN * M.HID + M.HID]
    o = M.IN * M.HID + M.HID
    W2 = pt[o:o + M.HID * M.OUT].view(M.HID, M.OUT)
    b2 = pt[o + M.HID * M.OUT:]
    x = (torch.tensor(X, dtype=torch.float64) - fmean) / fstd
    y = torch.nn.functional.linear(torch.relu(torch.nn.functional.linear(x, W1.T, b1)), W2.T, b2)
    tn = (torch.tensor(T, dtype=torch.float64) - tmean) / tstd
    loss = 0.5 * ((y - tn) ** 2).sum() / X.shape[0]
    loss.backward()
    return loss.item(), pt.grad.numpy(), y.detach().numpy()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_forward_and_gradient_match_torch_autograd(seed):
    p, norm, X, T = rand_case(seed)
    loss, g = M.loss_grad(p, norm, X, T)
    tl, tg, ty = torch_loss(p, norm, X, T)
    y, _, _ = M.forward(p, norm, X)
    np.testing.assert_allclose(y, ty, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) <= 1e-12 * max(1.0, abs(tl))
    np.testing.assert_allclose(g, tg, rtol=1e-10, atol=1e-12)


def test_gradient_matches_finite_differences():
    p, norm, X, T = rand_case(5, n=11)
    _, g = M.loss_grad(p, norm, X, T)
    rng = np.random.default_rng(9)
    h = 1e-6
    for i in rng.choice(M.NPARAMS, 60, replace=False):
        e = np.zeros_like(p)
        e[i] = h
        num = (M.loss_grad(p + e, norm, X, T)[0] - M.loss_grad(p - e, norm, X, T)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, num, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 5000), [0.0, -0.0, 1.0, 3.0e38]]).astype(np.float32)
    # exact halfway cases: ties to even
    a = np.concatenate([a, (np.arange(1, 50, dtype=np.uint32) << 15 | 0x3F800000).view(np.float32)])
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16_round(a), ref)


def test_adam_schedule_matches_torch_optim():
    p, norm, X, T = rand_case(7, n=64)
    ours, losses = M.train_adam(p, norm, X, T, epochs=3, batch=16, lr=1e-2, seed=11)
    # the same batches through torch.optim.Adam
    nxt = M._splitmix(11)
    perm = list(range(64))
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(3):
        for i in range(63, 0, -1):
            j = nxt() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for b0 in range(0, 64, 16):
            idx = perm[b0:b0 + 16]
            opt.zero_grad()
            _, g = M.loss_grad(pt.detach().numpy(), norm, X[idx], T[idx])
            pt.grad = torch.tensor(g)
            opt.step()
    np.testing.assert_allclose(ours, pt.detach().numpy(), rtol=1e-9, atol=1e-12)
    assert losses[-1] < losses[0]


def test_features_from_coarse():
    rng = np.random.default_rng(1)
    pd = np.sort(rng.uniform(0, 1e5, (4, 70)), axis=1)
    q = rng.normal(0, 10, (4, 8))
    f = M.features_from_coarse(pd, q)
    np.testing.assert_allclose(f[:, :64], np.log1p(pd[:, :64]))
    np.testing.assert_allclose(f[:, 64], np.log1p((q * q).sum(1)))
    short = M.features_from_coarse(pd[:, :10], q)  # fewer lists than 64: the last one repeats
    np.testing.assert_allclose(short[:, 10:64], np.repeat(np.log1p(pd[:, 9:10]), 54, axis=1))
