"""CPU tests pinning the float64 pruner-MLP oracle by central differences."""

import numpy
import pytest

from oracle import mlp_oracle as mo


def _params(rng, n_emb=5):
    p = {
        "W1": rng.normal(0, 0.3, (128, 48)), "b1": rng.normal(0, 0.1, 128),
        "W2": rng.normal(0, 0.1, (128, 128)), "b2": rng.normal(0, 0.1, 128),
        "W3": rng.normal(0, 0.1, (128, 128)), "b3": rng.normal(0, 0.1, 128),
        "w4": rng.normal(0, 0.2, 128), "b4": rng.normal(0, 0.1, 1),
        "emb": rng.normal(0, 0.5, (n_emb, 8)),
    }
    p["W1"][:, 40:] = 0.0
    return p


def test_oracle_gradients_match_central_differences():
    rng = numpy.random.default_rng(0)
    p = _params(rng)
    n = 13
    base = rng.normal(0, 1, (n, 48))
    base[:, 40:] = 0.0
    r1 = rng.integers(0, 5, n)
    r2 = rng.integers(0, 5, n)
    y = (rng.random(n) < 0.4).astype(float)
    loss, g, _ = mo.loss_and_grads(p, base, r1, r2, y, pos_weight=0.7)
    h = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _, _ = mo.loss_and_grads(p, base, r1, r2, y, pos_weight=0.7)
            flat[idx] = old - h
            lm, _, _ = mo.loss_and_grads(p, base, r1, r2, y, pos_weight=0.7)
            flat[idx] = old
            fd = (lp - lm) / (2 * h) / n
            got = g[name].reshape(-1)[idx]
            assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (name, idx, got, fd)


def test_oracle_loss_is_logistic():
    """A network whose output is a constant logit c gives the closed form
    loss sum_y [pos_w y log(1 + e^-c) + (1 - y) log(1 + e^c)]."""
    rng = numpy.random.default_rng(1)
    p = _params(rng)
    p["w4"][:] = 0.0
    p["b4"][:] = 0.3
    base = rng.normal(0, 1, (7, 48))
    y = numpy.array([1, 0, 1, 1, 0, 0, 1], dtype=float)
    loss, _, logit = mo.loss_and_grads(p, base, numpy.zeros(7, int), numpy.zeros(7, int), y, 2.0)
    c = 0.3
    ref = (2.0 * y * numpy.log1p(numpy.exp(-c)) + (1 - y) * numpy.log1p(numpy.exp(c))).sum()
    assert numpy.allclose(logit, c)
    assert abs(loss - ref) < 1e-12
