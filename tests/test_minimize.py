"""CPU tests of the segmented L-BFGS minimizer (tmol_amd/minimize.py)."""

import json
import os

import pytest
import torch

from tmol_amd.minimize import SegmentedLBFGS

HERE = os.path.dirname(os.path.abspath(__file__))


def _rosenbrock(x, seg, scale):
    nseg = int(seg.max()) + 1
    same = seg[1:] == seg[:-1]
    t1 = (1 - x) ** 2
    t2 = scale[seg[:-1]] * (x[1:] - x[:-1] ** 2) ** 2 * same
    E = torch.zeros(nseg, dtype=x.dtype).index_add(0, seg, t1)
    return E.index_add(0, seg[:-1], t2)


def test_matches_reference_lbfgs_evaluations():
    """Same problem, same settings as the reference LBFGS_Armijo run that wrote
    the fixture: the minimizer makes the same sequence of energy evaluations
    (every line-search trial of every segment) and ends at the same point."""
    gold = json.load(open(os.path.join(HERE, "golden", "lbfgs_rosenbrock.json")))
    seg = torch.cat([torch.full((n,), i, dtype=torch.int64) for i, n in enumerate(gold["dims"])])
    scale = torch.tensor(gold["scale"], dtype=torch.float64)
    x0 = torch.tensor(gold["x0"], dtype=torch.float64)
    evals = []

    def evaluate(x):
        xx = x.detach().clone().requires_grad_(True)
        E = _rosenbrock(xx, seg, scale)
        E.sum().backward()
        evals.append(E.detach())
        return E.detach(), xx.grad.detach()

    opt = SegmentedLBFGS(seg, max_iter=100, gradtol=1e-6)
    x, trace = opt.run(x0, evaluate)
    ref = torch.tensor(gold["evaluations"], dtype=torch.float64)
    got = torch.stack(evals)
    # identical decisions: same number of line-search trials and iterations
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert trace.iterations == gold["n_iter"]
    # identical values until rounding differences (the Gram matrices are kept
    # incrementally here, recomputed there) grow along Rosenbrock's valley
    torch.testing.assert_close(got[:70], ref[:70], rtol=1e-8, atol=1e-12)
    torch.testing.assert_close(got, ref, rtol=1e-2, atol=1e-7)
    torch.testing.assert_close(x, torch.tensor(gold["x_final"], dtype=torch.float64),
                               rtol=1e-3, atol=1e-3)


def _dense_bfgs_inverse(pairs, n):
    """H from the textbook BFGS inverse update, H0 = I."""
    H = torch.eye(n, dtype=torch.float64)
    for s, y in pairs:
        rho = 1.0 / float(s @ y)
        V = torch.eye(n, dtype=torch.float64) - rho * torch.outer(y, s)
        H = V.T @ H @ V + rho * torch.outer(s, s)
    return H


@pytest.mark.parametrize("history", [3, 5])
def test_compact_product_equals_dense_bfgs(history):
    """H g from the incrementally kept compact form equals the dense BFGS
    inverse Hessian applied to g, including after the history wraps."""
    torch.manual_seed(0)
    n, P = 7, 2
    seg = torch.arange(P).repeat_interleave(n)
    opt = SegmentedLBFGS(seg, history=history)
    opt._reset_history(torch.device("cpu"))
    A = [torch.randn(n, n, dtype=torch.float64) for _ in range(P)]
    A = [a @ a.T + n * torch.eye(n, dtype=torch.float64) for a in A]  # SPD: s.y > 0
    pairs = [[] for _ in range(P)]
    for k in range(history + 3):
        s = torch.randn(P, n, dtype=torch.float64)
        y = torch.stack([A[p] @ s[p] for p in range(P)])
        opt._push(s, y)
        for p in range(P):
            pairs[p].append((s[p], y[p]))
        g = torch.randn(P, n, dtype=torch.float64)
        hg = opt._hg(g)
        for p in range(P):
            H = _dense_bfgs_inverse(pairs[p][-history:], n)
            torch.testing.assert_close(hg[p], H @ g[p], rtol=1e-9, atol=1e-9)


def test_absent_history_entries_do_not_contribute():
    """A segment whose rows of an entry are zero behaves as if the entry were
    absent (the reference's per-segment history validity)."""
    torch.manual_seed(1)
    n = 5
    seg = torch.arange(2).repeat_interleave(n)
    opt = SegmentedLBFGS(seg, history=4)
    opt._reset_history(torch.device("cpu"))
    A = torch.randn(n, n, dtype=torch.float64)
    A = A @ A.T + n * torch.eye(n, dtype=torch.float64)
    s1 = torch.randn(2, n, dtype=torch.float64)
    s2 = torch.randn(2, n, dtype=torch.float64)
    s2[1] = 0.0  # segment 1 did not keep the second pair
    opt._push(s1, s1 @ A)
    opt._push(s2, s2 @ A)
    g = torch.randn(2, n, dtype=torch.float64)
    H1 = _dense_bfgs_inverse([(s1[1], A @ s1[1])], n)
    torch.testing.assert_close(opt._hg(g)[1], H1 @ g[1], rtol=1e-9, atol=1e-9)


def test_quadratic_segments_converge():
    torch.manual_seed(2)
    dims = [3, 6, 4]
    seg = torch.cat([torch.full((d,), i) for i, d in enumerate(dims)])
    A = [torch.randn(d, d, dtype=torch.float64) for d in dims]
    A = [a @ a.T + torch.eye(a.shape[0], dtype=torch.float64) for a in A]
    xs = [torch.randn(d, dtype=torch.float64) for d in dims]

    def evaluate(x):
        E, G, o = [], [], 0
        for a, c, d in zip(A, xs, dims):
            r = x[o:o + d] - c
            E.append(0.5 * r @ a @ r)
            G.append(a @ r)
            o += d
        return torch.stack(E), torch.cat(G)

    x, trace = SegmentedLBFGS(seg, gradtol=1e-8, rtol=0, atol=0).run(
        torch.zeros(seg.numel(), dtype=torch.float64), evaluate)
    torch.testing.assert_close(x, torch.cat(xs), rtol=1e-6, atol=1e-6)
    assert bool(trace.converged.all())
