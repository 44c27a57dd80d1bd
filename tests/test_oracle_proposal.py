"""Pins the numpy proposal oracle: its hand-written gradient against central finite differences."""
import numpy as np

from oracle import proposal_oracle as PO


def _case(seed=3, n=40):
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(PO.NPARAM) * 0.3
    c = rng.standard_normal(6)
    g = rng.standard_normal((n, 6)) * 1.3
    wh = rng.random(n)
    wh /= wh.sum()
    return P, c, g, wh


def test_gradient_matches_finite_differences():
    P, c, g, wh = _case()
    loss, G, _ = PO.loss_grad(P, c, g, wh, 0.25)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(PO.NPARAM, 40, replace=False), [PO.OFF_B3 + 1, PO.OFF_B3 + 20, PO.OFF_B3 + 70]])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (PO.loss_grad(Pp, c, g, wh, 0.25)[0] - PO.loss_grad(Pm, c, g, wh, 0.25)[0]) / (2 * h)
        assert abs(fd - G[i]) <= 1e-6 * max(1., abs(fd)), (i, fd, G[i])


def test_defensive_mixture_bounds_weights():
    # with alpha = 0.25 the weight p/q never exceeds 1/alpha
    P, c, g, wh = _case(5, 200)
    o, _ = PO.forward(P, c)
    loss, _ = PO.loss_and_dout(o, g, np.ones(len(g)), 0.25)
    # -sum log q >= -sum log(p / 0.25) elementwise bound
    lp = -0.5 * np.sum(g * g, axis=1) - 3 * np.log(2 * np.pi)
    assert loss <= -np.sum(lp + np.log(0.25)) + 1e-9
