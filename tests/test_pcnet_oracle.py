"""Pins the numpy network oracle (oracle/pcnet.py): its backward pass against
central finite differences of its own forward + loss, its convolution
against a brute-force direct sum, and the loss's scale invariance."""
import numpy as np
import pytest

from oracle import pcnet


def _setup(n=32, seed=0):
    rng = np.random.default_rng(seed)
    P = pcnet.init_params(seed)
    for k in P:
        if k.endswith(".b"):
            P[k] = rng.normal(0, 0.05, P[k].shape)
    fields = rng.normal(size=(3, n, n, n))
    target = rng.normal(size=(n, n, n))
    return P, pcnet.make_input(fields), pcnet.pool4_target(target)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 3, 5, 8))
    W = rng.normal(size=(2, 27, 8))
    b = rng.normal(size=2)
    y = pcnet.conv(x, W, b)
    for (i, j, k) in [(0, 0, 0), (3, 2, 4), (1, 1, 2)]:
        for o in range(2):
            s = b[o]
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        ii, jj, kk = i + dx, j + dy, k + dz
                        if 0 <= ii < 4 and 0 <= jj < 3 and 0 <= kk < 5:
                            s += W[o, (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] @ x[ii, jj, kk]
            assert y[i, j, k, o] == pytest.approx(s, rel=1e-12)


def test_backward_matches_finite_differences():
    P, x0, t = _setup()
    y, C = pcnet.forward(P, x0)
    L, gy = pcnet.loss(y, t)
    G = pcnet.backward(P, C, gy)
    rng = np.random.default_rng(3)
    for name in ("e1.w", "e4.w", "d3.w", "d1.w", "out.w", "e2.b", "d2.b"):
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in P[name].shape)
            h = 1e-6
            old = P[name][idx]
            P[name][idx] = old + h
            Lp = pcnet.loss(pcnet.forward(P, x0)[0], t)[0]
            P[name][idx] = old - h
            Lm = pcnet.loss(pcnet.forward(P, x0)[0], t)[0]
            P[name][idx] = old
            fd = (Lp - Lm) / (2 * h)
            assert G[name][idx] == pytest.approx(fd, rel=2e-4, abs=1e-9), (name, idx)


def test_loss_is_scale_invariant():
    P, x0, t = _setup(seed=2)
    y, _ = pcnet.forward(P, x0)
    L1, _ = pcnet.loss(y, t)
    L2, _ = pcnet.loss(3.7 * y, -0.2 * t)
    assert L1 == pytest.approx(L2, rel=1e-12)
    assert pcnet.loss(t, t)[0] == pytest.approx(0.0, abs=1e-14)
