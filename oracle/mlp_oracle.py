"""FP64 reference of the proposal scorer (TEST INFRASTRUCTURE ONLY).

32-128-1 MLP, ReLU hidden layer, mean-squared-error loss; the GPU forward
uses bf16-rounded weights with fp32 accumulation, so the reference takes the
weights rounded to bf16 (round to nearest even) and computes in float64.
Parameter layout: W1[128][32], b1[128], W2[128], b2.
"""
import numpy as np

F, H = 32, 128
NP = H * F + H + H + 1


def bf16_round(w):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    a = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    lsb = (a >> 16) & 1
    r = (a + 0x7FFF + lsb) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    W1 = w[:H * F].reshape(H, F)
    b1 = w[H * F:H * F + H]
    W2 = w[H * F + H:H * F + 2 * H]
    b2 = w[-1]
    return W1, b1, W2, b2


def forward(w, X):
    W1, b1, W2, b2 = unpack(w)
    pre = X @ W1.T + b1
    return np.maximum(pre, 0.0) @ W2 + b2, pre


def loss_and_grad(w, X, Y):
    """MSE loss and its gradient (float64) at parameters w."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    W1, b1, W2, b2 = unpack(w)
    out, pre = forward(w, X)
    e = out - Y
    B = X.shape[0]
    loss = float(np.mean(e * e))
    g = 2.0 * e / B
    act = np.maximum(pre, 0.0)
    gW2 = act.T @ g
    gb2 = g.sum()
    gh = np.outer(g, W2) * (pre > 0)
    gW1 = gh.T @ X
    gb1 = gh.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2, [gb2]])


BETA1, BETA2, EPS = 0.9, 0.999, 1e-8


def adam_step(w, m, v, g, t, lr):
    """One Adam update (Kingma and Ba 2015, Algorithm 1) in float64: moments
    m, v from gradient g, bias correction by 1 - beta^t, then
    w - lr * m_hat / (sqrt(v_hat) + eps).  Returns (w, m, v)."""
    g = np.asarray(g, dtype=np.float64)
    m = BETA1 * np.asarray(m, dtype=np.float64) + (1 - BETA1) * g
    v = BETA2 * np.asarray(v, dtype=np.float64) + (1 - BETA2) * g * g
    mh = m / (1 - BETA1 ** t)
    vh = v / (1 - BETA2 ** t)
    return np.asarray(w, dtype=np.float64) - lr * mh / (np.sqrt(vh) + EPS), m, v
