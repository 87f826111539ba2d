"""CPU reference of the learned-stop MLP (float64 numpy).

TEST INFRASTRUCTURE ONLY: the checker for the HIP training kernels'
gradients.  The network has no counterpart in the reference; this is the
textbook forward/backward of a 16 -> 64 -> 64 -> 1 ReLU MLP with a
sigmoid/BCE head, with the weights rounded to BF16 (round to nearest even)
in the forward pass and the gradient passed straight through the rounding.
Its own correctness is pinned by finite differences in
tests/test_learned_cpu.py.
"""
import numpy as np

NF, NH = 16, 64
P = NH * NF + NH + NH * NH + NH + NH + 1


def bf16_round(a):
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    w1 = p[o:o + NH * NF].reshape(NH, NF); o += NH * NF
    b1 = p[o:o + NH]; o += NH
    w2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return w1, b1, w2, b2, w3, b3


def forward(p, x, round_bf16=True):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    if round_bf16:
        w1 = bf16_round(w1).astype(np.float64)
        w2 = bf16_round(w2).astype(np.float64)
        w3 = bf16_round(w3).astype(np.float64)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ w1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w2.T + b2
    h2 = np.maximum(a2, 0)
    out = h2 @ w3 + b3
    return out, (x, h1, h2, w1, w2, w3)


def loss_and_grad(p, x, y, wt, round_bf16=True):
    """Sum of wt * BCE(sigmoid(logit), y) and its gradient (flat [P])."""
    out, (x, h1, h2, w1, w2, w3) = forward(p, x, round_bf16)
    y = np.asarray(y, dtype=np.float64)
    wt = np.asarray(wt, dtype=np.float64)
    loss = (wt * (np.maximum(out, 0) - out * y + np.log1p(np.exp(-np.abs(out))))).sum()
    prob = 1.0 / (1.0 + np.exp(-out))
    dout = wt * (prob - y)
    gw3 = h2.T @ dout
    gb3 = dout.sum()
    dh2 = np.outer(dout, w3) * (h2 > 0)
    gw2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    gw1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3, [gb3]])
    return loss, g, out
