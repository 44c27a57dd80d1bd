"""FP64 numpy restatement of the learned macro-electron proposal.

TEST INFRASTRUCTURE ONLY: the checker of the library's proposal kernels
(srw_amd/csrc/learn.hip).  Only tests/ and bench.py's CPU leg use it.

Network: 6 -> 64 (tanh) -> 64 (tanh) -> 104 (8 logits, 8x6 means, 8x6 log
sigmas clamped to [-3, 1]).  Proposal q = alpha p + (1 - alpha) GMM with
p = N(0, I6).  Loss over a batch: -sum_j wh_j log q(g_j).  Gradients by
reverse-mode differentiation written out by hand; finite differences of the
same loss pin them in the tests.
"""
import numpy as np

IN, H, K, D = 6, 64, 8, 6
OUT = K + 2 * K * D
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + OUT * H
NPARAM = OFF_B3 + OUT
L2P = np.log(2 * np.pi)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OFF_W1:OFF_B1].reshape(H, IN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(OUT, H), P[OFF_B3:NPARAM])


def forward(P, c):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    h1 = np.tanh(W1 @ c + b1)
    h2 = np.tanh(W2 @ h1 + b2)
    o = W3 @ h2 + b3
    return o, (np.asarray(c, np.float64), h1, h2)


def loss_and_dout(o, g, wh, alpha):
    """Loss -sum wh log q and its gradient with respect to the network output."""
    g = np.asarray(g, np.float64).reshape(-1, D)
    wh = np.asarray(wh, np.float64)
    logit = o[:K]
    mu = o[K:K + K * D].reshape(K, D)
    raw = o[K + K * D:].reshape(K, D)
    ls = np.clip(raw, -3., 1.)
    msk = ((raw > -3.) & (raw < 1.)).astype(np.float64)
    logZ = np.log(np.sum(np.exp(logit - logit.max()))) + logit.max()
    lpi = logit - logZ
    z = (g[:, None, :] - mu[None]) * np.exp(-ls)[None]
    lN = lpi[None, :] + np.sum(-0.5 * z * z - ls[None] - 0.5 * L2P, axis=2)       # (n, K)
    lm = np.logaddexp.reduce(lN, axis=1)
    lp = -0.5 * np.sum(g * g, axis=1) - 0.5 * D * L2P
    lq = np.logaddexp(np.log(alpha) + lp, np.log(1 - alpha) + lm)
    loss = -np.sum(wh * lq)
    f = -wh * (1 - alpha) * np.exp(lm - lq)                                         # (n,)
    r = np.exp(lN - lm[:, None])                                                   # responsibilities
    pi = np.exp(lpi)
    dlogit = np.sum(f[:, None] * (r - pi[None]), axis=0)
    dz = g[:, None, :] - mu[None]
    is2 = np.exp(-2 * ls)[None]
    dmu = np.sum((f[:, None] * r)[:, :, None] * dz * is2, axis=0)
    dls = np.sum((f[:, None] * r)[:, :, None] * (dz * dz * is2 - 1.), axis=0) * msk
    return loss, np.concatenate([dlogit, dmu.ravel(), dls.ravel()])


def backward(P, cache, dout):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    c, h1, h2 = cache
    G = np.zeros(NPARAM)
    G[OFF_W3:OFF_B3] = np.outer(dout, h2).ravel()
    G[OFF_B3:NPARAM] = dout
    d2 = (W3.T @ dout) * (1 - h2 * h2)
    G[OFF_W2:OFF_B2] = np.outer(d2, h1).ravel()
    G[OFF_B2:OFF_W3] = d2
    d1 = (W2.T @ d2) * (1 - h1 * h1)
    G[OFF_W1:OFF_B1] = np.outer(d1, c).ravel()
    G[OFF_B1:OFF_W2] = d1
    return G


def loss_grad(P, c, g, wh, alpha):
    o, cache = forward(P, c)
    loss, dout = loss_and_dout(o, g, wh, alpha)
    return loss, backward(P, cache, dout), o
