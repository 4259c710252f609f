"""Pins of the radiance-cache oracle (oracle/radiance_cache.py) on the CPU.

  encoding        sin/cos(2^f pi p) from the double-angle recurrence agree with the direct
                  float64 values within the recurrence's error growth; pass-through and zero
                  padding of the remaining inputs
  BF16 rounding   round-to-nearest-even on hand-picked ties and halfway cases
  network         the float64 network (rounding off) equals torch.autograd on the same weights:
                  outputs, relative-L2 loss with a stop-gradient denominator, every gradient
  finite diff.    central differences of the loss with the denominator held fixed
  Adam            five steps equal torch.optim.Adam with the same betas and epsilon
"""
import numpy as np
import pytest
import torch

from oracle import radiance_cache as rc


def random_records(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(n, 16)).astype(np.float32)


def test_layout_sizes():
    assert rc.NUM_WEIGHTS == 23552 and rc.NUM_BIASES == 336 and rc.NUM_PARAMS == 23888
    live = 96 * 64 + 4 * 64 * 64 + 4 * 64 + 5 * 64 + 4
    assert live == 23108


def test_encoding_closed_form():
    raw = random_records(2000, 1)
    x = rc.encode(raw, round_bf16=False)
    p = raw[:, :3].astype(np.float64)
    for a in range(3):
        for f in range(12):
            ref_s = np.sin(2.0 ** f * np.pi * p[:, a])
            ref_c = np.cos(2.0 ** f * np.pi * p[:, a])
            tol = 4e-7 * 2.0 ** f + 1e-6
            assert np.abs(x[:, 24 * a + 2 * f] - ref_s).max() < tol * 8, (a, f)
            assert np.abs(x[:, 24 * a + 2 * f + 1] - ref_c).max() < tol * 8, (a, f)
    np.testing.assert_array_equal(x[:, 72:85], raw[:, 3:16].astype(np.float64))
    assert np.all(x[:, 85:] == 0)
    # p = 1/2: sin = 1, cos = 0 at f = 0, then sin(pi) = 0, cos(pi) = -1, then 0, 1
    h = rc.encode(np.full((1, 16), 0.5, np.float32), round_bf16=False)[0]
    np.testing.assert_allclose(h[0:6], [1, 0, 0, -1, 0, 1], atol=1e-6)


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    cases = {one + ulp / 2: one,                  # tie -> even (1.0)
             one + 3 * ulp / 2: one + 2 * ulp,    # tie -> even (1 + 2 ulp)
             one + ulp / 2 + 2.0 ** -20: one + ulp,
             -(one + ulp / 2): -one,
             0.0: 0.0}
    for x, want in cases.items():
        assert rc.bf16(np.array([x]))[0] == want, x


def torch_loss(params, raw, targets, num_out):
    Ws, bs = rc.unpack(params, round_bf16=False)
    Wt = [torch.tensor(W, dtype=torch.float64, requires_grad=True) for W in Ws]
    bt = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for b in bs]
    x = torch.tensor(rc.encode(raw, round_bf16=False), dtype=torch.float64)
    for l in range(5):
        x = torch.relu(x @ Wt[l].T + bt[l])
    y = (x @ Wt[5].T + bt[5])[:, :num_out]
    t = torch.tensor(targets[:, :num_out], dtype=torch.float64)
    loss = ((y - t) ** 2 / (y.detach() ** 2 + 0.01)).mean()
    loss.backward()
    g = np.zeros(rc.NUM_PARAMS)
    for l in range(6):
        r, c, o = rc.layer_rows(l), rc.layer_cols(l), rc.weight_offset(l)
        g[o:o + r * c] = Wt[l].grad.numpy().reshape(-1)
        g[rc.NUM_WEIGHTS + 64 * l:rc.NUM_WEIGHTS + 64 * l + r] = bt[l].grad.numpy()
    return float(loss.detach()), g, y.detach().numpy()


@pytest.mark.parametrize("num_out", [3, 4])
def test_network_matches_torch_autograd(num_out):
    params = rc.glorot_params(5)
    params[rc.NUM_WEIGHTS:rc.NUM_WEIGHTS + 320] = np.random.default_rng(6).normal(0, 0.1, 320)
    raw = random_records(300, 7)
    targets = np.random.default_rng(8).uniform(0, 2, size=(300, 4))
    y, _ = rc.forward(params, raw, round_bf16=False)
    loss, grad = rc.loss_and_grad(params, raw, targets, num_out, round_bf16=False)
    tl, tg, ty = torch_loss(params, raw, targets, num_out)
    np.testing.assert_allclose(y[:, :num_out], ty, rtol=1e-12, atol=1e-12)
    assert abs(loss - tl) < 1e-12 * max(1.0, abs(tl))
    np.testing.assert_allclose(grad, tg, rtol=1e-9, atol=1e-15)
    # padding rows of the output layer never receive a gradient
    o = rc.weight_offset(5)
    assert np.all(grad[o + 4 * 64:o + 16 * 64] == 0)


def test_gradient_finite_differences():
    params = rc.glorot_params(9).astype(np.float64)
    raw = random_records(64, 10)
    targets = np.random.default_rng(11).uniform(0, 2, size=(64, 4))
    y0, _ = rc.forward(params, raw, round_bf16=False)
    den = y0[:, :4] ** 2 + 0.01

    def fixed_loss(p):
        y, _ = rc.forward(p, raw, round_bf16=False)
        return float(((y[:, :4] - targets[:, :4]) ** 2 / den).mean())

    _, grad = rc.loss_and_grad(params, raw, targets, 4, round_bf16=False)
    rng = np.random.default_rng(12)
    picks = list(rng.choice(rc.NUM_WEIGHTS - 16 * 64, 12, replace=False)) + \
        [rc.weight_offset(5) + 5, rc.NUM_WEIGHTS + 3, rc.NUM_WEIGHTS + 320 + 2]
    for i in picks:
        h = 1e-6
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (fixed_loss(p1) - fixed_loss(p2)) / (2 * h)
        assert abs(fd - grad[i]) < 1e-4 * max(1e-3, abs(grad[i])) + 1e-9, (i, fd, grad[i])


def test_adam_matches_torch():
    rng = np.random.default_rng(13)
    p0 = rng.normal(size=1000).astype(np.float32)
    grads = [rng.normal(scale=10.0 ** -k, size=1000).astype(np.float32) for k in (1, 3, 5, 2, 6)]
    tp = torch.nn.Parameter(torch.tensor(p0))
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.99), eps=1e-15)
    master, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for step, g in enumerate(grads, start=1):
        tp.grad = torch.tensor(g)
        opt.step()
        master, m, v = rc.adam(master, m, v, g, step)
        np.testing.assert_allclose(master, tp.detach().numpy(), rtol=2e-6, atol=2e-7)
