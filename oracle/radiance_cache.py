"""CPU restatement of the radiance-cache network (test infrastructure only).

This is the checker for the GPU's radiance-cache kernels (vlr_amd/csrc/radiance_cache.hip):
the same network, input encoding, loss, backward pass and Adam update, in numpy float64 with
BF16 rounding (round to nearest even) at the points where the GPU stores BF16: encoded inputs,
weights, hidden activations and the gradients w.r.t. pre-activations.  With `bf16=False` the
rounding is skipped and the result is the exact float64 network, which the tests pin against
torch.autograd and torch.optim.Adam on the CPU.  The cache has no counterpart in the
reference; its definition is the one documented in include/vlr/vlr_amd.h.

Parameter layout (kNumParams floats): per layer l = 0..5 the row-major weights
[rows_l][cols_l] (rows 64, the output layer padded to 16 rows of which 4 are live; cols 96 for
l = 0, else 64), then the biases (5 x 64, then 16).
"""
import numpy as np

K_IN, HIDDEN, OUT, OUT_PAD, LAYERS, FREQS = 96, 64, 4, 16, 6, 12


def layer_rows(l):
    return OUT_PAD if l == LAYERS - 1 else HIDDEN


def layer_cols(l):
    return K_IN if l == 0 else HIDDEN


def weight_offset(l):
    return 0 if l == 0 else HIDDEN * K_IN + (l - 1) * HIDDEN * HIDDEN


NUM_WEIGHTS = HIDDEN * K_IN + 4 * HIDDEN * HIDDEN + OUT_PAD * HIDDEN
NUM_BIASES = 5 * HIDDEN + OUT_PAD
NUM_PARAMS = NUM_WEIGHTS + NUM_BIASES


def bias_offset(l):
    return NUM_WEIGHTS + l * HIDDEN


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(raw, round_bf16=True):
    """(n, 16) raw records -> (n, 96) network inputs.

    inputs 24a + 2f and 24a + 2f + 1: sin and cos of 2^f * pi * raw[a] (a < 3, f < 12), by the
    double-angle recurrence in float32 from one sin/cos per axis; inputs 72..84 = raw[3..15];
    85..95 = 0."""
    raw = np.asarray(raw, np.float32)
    n = raw.shape[0]
    out = np.zeros((n, K_IN), np.float32)
    two, one = np.float32(2.0), np.float32(1.0)
    for a in range(3):
        arg = np.float32(np.pi) * raw[:, a]
        s, c = np.sin(arg).astype(np.float32), np.cos(arg).astype(np.float32)
        for f in range(FREQS):
            out[:, 24 * a + 2 * f] = s
            out[:, 24 * a + 2 * f + 1] = c
            s, c = (two * s * c).astype(np.float32), (one - two * s * s).astype(np.float32)
    out[:, 72:85] = raw[:, 3:16]
    return bf16(out) if round_bf16 else out.astype(np.float64)


def unpack(params, round_bf16=True):
    p = np.asarray(params, np.float64)
    Ws, bs = [], []
    for l in range(LAYERS):
        r, c, o = layer_rows(l), layer_cols(l), weight_offset(l)
        W = p[o:o + r * c].reshape(r, c)
        Ws.append(bf16(W) if round_bf16 else W.copy())
    for l in range(LAYERS):
        n = layer_rows(l)
        bs.append(p[NUM_WEIGHTS + l * HIDDEN:NUM_WEIGHTS + l * HIDDEN + n].copy())
    return Ws, bs


def forward(params, raw, round_bf16=True):
    """Returns y (n, 4) and the stored activations x_0..x_5 ((n, cols) each)."""
    Ws, bs = unpack(params, round_bf16)
    x = encode(raw, round_bf16)
    xs = [x]
    for l in range(LAYERS - 1):
        z = x @ Ws[l].T + bs[l]
        a = np.maximum(z, 0.0)
        x = bf16(a) if round_bf16 else a
        xs.append(x)
    y = x @ Ws[-1].T + bs[-1]
    return y[:, :OUT], xs


def loss_and_grad(params, raw, targets, num_out, round_bf16=True):
    """Relative L2 loss sum((y - t)^2 / (sg(y)^2 + 0.01)) / (n * num_out) and its gradient."""
    Ws, bs = unpack(params, round_bf16)
    y, xs = forward(params, raw, round_bf16)
    n = len(raw)
    t = np.asarray(targets, np.float64)[:, :num_out]
    yy = y[:, :num_out]
    den = yy * yy + 0.01
    norm = 1.0 / (n * num_out)
    loss = float(((yy - t) ** 2 / den).sum() * norm)
    d = np.zeros((n, OUT_PAD))
    d[:, :num_out] = 2.0 * (yy - t) / den * norm
    if round_bf16:
        d = bf16(d)
    grad = np.zeros(NUM_PARAMS)
    for l in range(LAYERS - 1, -1, -1):
        r, c, o = layer_rows(l), layer_cols(l), weight_offset(l)
        grad[o:o + r * c] = (d.T @ xs[l]).reshape(-1)
        grad[NUM_WEIGHTS + l * HIDDEN:NUM_WEIGHTS + l * HIDDEN + r] = d.sum(axis=0)
        if l > 0:
            back = (d @ Ws[l]) * (xs[l] > 0)
            d = bf16(back) if round_bf16 else back
    return loss, grad


def adam(master, m, v, grad, step, lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15):
    """Adam (Kingma and Ba 2015) in float32, one step; returns (master, m, v)."""
    f = np.float32
    g = np.asarray(grad, np.float32)
    m = (f(beta1) * np.asarray(m, np.float32) + f(1 - beta1) * g).astype(np.float32)
    v = (f(beta2) * np.asarray(v, np.float32) + f(1 - beta2) * g * g).astype(np.float32)
    m_hat = m / f(1 - beta1 ** step)
    v_hat = v / f(1 - beta2 ** step)
    master = (np.asarray(master, np.float32) - f(lr) * m_hat / (np.sqrt(v_hat) + f(eps))).astype(np.float32)
    return master, m, v


def glorot_params(seed):
    """Glorot-uniform weights and zero biases (numpy's generator; for CPU tests)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NUM_PARAMS, np.float32)
    for l in range(LAYERS):
        r, c, o = layer_rows(l), layer_cols(l), weight_offset(l)
        live = OUT if l == LAYERS - 1 else r
        a = np.sqrt(6.0 / (c + live))
        W = np.zeros((r, c), np.float32)
        W[:live] = rng.uniform(-a, a, size=(live, c))
        p[o:o + r * c] = W.reshape(-1)
    return p
