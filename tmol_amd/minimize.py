"""Per-pose (segmented) L-BFGS with an Armijo line search, driving LJ+LK
minimization.

Mirrors the reference minimizer's contract (tmol/optimization/
lbfgs_armijo.py:196-744, Rosetta-derived constants):

* each pose is an independent segment: its own step length, history
  validity, restart and convergence;
* history of 128 (s, y) pairs; a pair enters a segment's history only when
  s.y > 1e-6 for that segment; the search direction is -H g with H_0 = I,
  computed in float64;
* directional-derivative repair: flip the components with g_i d_i > 0, then
  fall back to steepest descent (lbfgs_armijo.py:557-575);
* Armijo search per segment (lbfgs_armijo.py:68-188): start at
  min(2 t_prev, 1); sufficient decrease sigma = 0.1; a step that looks linear
  (sigma = 0.8) is tried once at twice the length; backtracking by 0.25;
  below minstep a downhill step is still taken, else the search fails;
* a failed search restarts the segment once from steepest descent with
  step min(1 / sqrt(|g.d|), 1); a second failure retires it;
* convergence: max |g| <= gradtol freezes a segment; the run stops when every
  segment is frozen, retired or changed its energy by <= atol or relative
  <= rtol (defaults sqrt(eps) of the dtype).

The inverse-Hessian product uses the compact form of Byrd, Nocedal and
Schnabel (Math. Prog. 63, 1994): H g = g + S p1 + Y p2 with
p2 = -R^-1 S g and p1 = R^-T ((D + Y Y^T) R^-1 S g - Y g), R the upper
triangle of S Y^T.  Here the Gram matrices S Y^T and Y Y^T are kept up to date
incrementally (one new row and column per accepted pair, O(m n)) instead of
being recomputed from the history every iteration (O(m^2 n)).
"""

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

import torch

Evaluate = Callable[[torch.Tensor], Tuple[torch.Tensor, torch.Tensor]]

_LBFGS_LIB = None


def _lbfgs_lib():
    """The HIP library's history-product entry points (tmol_lbfgs_*)."""
    global _LBFGS_LIB
    if _LBFGS_LIB is None:
        import ctypes

        from . import ops

        L = ops.lib()
        I32, I64, VP = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
        L.tmol_lbfgs_dots.argtypes = [I32, I32, I64, I32, VP, VP, I32, VP, VP]
        L.tmol_lbfgs_combine.argtypes = [I32, I32, I64, I32, VP, VP, VP, VP, VP, VP, VP]
        _LBFGS_LIB = L
    return _LBFGS_LIB


def _vp(t: torch.Tensor):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


def _stream_of(t: torch.Tensor):
    from . import ops

    return ops._stream(t.device)


def _check_lbfgs(status, what):
    from . import ops

    ops._check(status, what)


@dataclass
class Trace:
    energy: List[torch.Tensor] = field(default_factory=list)    # [P] per iteration (f64, cpu)
    grad_norm: List[torch.Tensor] = field(default_factory=list)  # [P] L2 norm of dE/dx
    grad_max: List[torch.Tensor] = field(default_factory=list)   # [P] max |dE/dx_i|
    evals: List[int] = field(default_factory=list)               # evaluations per iteration
    iterations: int = 0
    converged: Optional[torch.Tensor] = None
    stalled: Optional[torch.Tensor] = None

    def energies(self) -> torch.Tensor:
        return torch.stack(self.energy)

    def grad_norms(self) -> torch.Tensor:
        return torch.stack(self.grad_norm)


class SegmentedLBFGS:
    """Minimize sum_p E_p(x) with every segment p independent.

    ``segment_ids[i]`` is the segment of parameter element i.  ``evaluate(x)``
    returns ``(E [P] float64, dE/dx [N])`` at the flat parameter vector x."""

    def __init__(self, segment_ids: torch.Tensor, n_segments: Optional[int] = None,
                 history: int = 128, max_iter: int = 200, gradtol: float = 1.0,
                 rtol: Optional[float] = None, atol: Optional[float] = None, lr: float = 1.0,
                 minstep: float = 1e-12, verbose: bool = False):
        seg = segment_ids.reshape(-1).to(torch.int64)
        self.device = seg.device
        self.P = int(n_segments) if n_segments is not None else int(seg.max()) + 1
        counts = torch.bincount(seg, minlength=self.P)
        if bool((counts == 0).any()):
            raise ValueError("every segment needs at least one parameter")
        self.seg = seg
        self.S = int(counts.max())
        # position of each element inside its segment's padded row
        order = torch.argsort(seg, stable=True)
        start = torch.cumsum(counts, 0) - counts
        rank = torch.empty_like(seg)
        rank[order] = torch.arange(seg.numel(), device=seg.device) - start[seg[order]]
        self.flat_index = seg * self.S + rank
        self.m = int(history)
        self.max_iter = int(max_iter)
        self.gradtol = float(gradtol)
        self.rtol, self.atol = rtol, atol
        self.lr = float(lr)
        self.minstep = float(minstep)
        self.verbose = verbose

    # -- padded <-> flat -------------------------------------------------------
    def _padded(self, v: torch.Tensor, dtype=torch.float64) -> torch.Tensor:
        out = torch.zeros(self.P * self.S, dtype=dtype, device=v.device)
        out[self.flat_index] = v.to(dtype)
        return out.view(self.P, self.S)

    def _flat(self, padded: torch.Tensor) -> torch.Tensor:
        return padded.reshape(-1)[self.flat_index]

    def _seg_sum(self, v: torch.Tensor) -> torch.Tensor:
        # a row reduction of the padded layout: no atomics, the same order on
        # every run
        return self._padded(v).sum(1)

    def _seg_max(self, v: torch.Tensor) -> torch.Tensor:
        pad = torch.full((self.P * self.S,), -math.inf, dtype=torch.float64, device=v.device)
        pad[self.flat_index] = v.to(torch.float64)
        return pad.view(self.P, self.S).amax(1)

    def _per_elem(self, per_seg: torch.Tensor) -> torch.Tensor:
        return per_seg[self.seg]

    # -- products with the history (HIP kernels on the device, torch on the host)
    def _dots(self, H: torch.Tensor, n_phys: int, vecs) -> torch.Tensor:
        """[K, P, n_phys]: H[p, i, :] . v_k[p, :] for the used rows i."""
        if H.is_cuda:
            V = torch.stack(vecs).contiguous()
            out = torch.empty((len(vecs), self.P, n_phys), dtype=torch.float64, device=H.device)
            _check_lbfgs(_lbfgs_lib().tmol_lbfgs_dots(
                self.P, self.m, self.S, n_phys, _vp(H), _vp(V), len(vecs), _vp(out),
                _stream_of(H)), "lbfgs dots")
            return out
        Hv = H[:, :n_phys]
        return torch.stack([torch.bmm(Hv, v.unsqueeze(-1)).squeeze(-1) for v in vecs])

    def _combine(self, g: torch.Tensor, n_phys: int, c1: torch.Tensor, c2: torch.Tensor):
        """g + sum_i c1[:, i] Sh[:, i] + c2[:, i] Yh[:, i] for the used rows."""
        if g.is_cuda:
            out = torch.empty_like(g)
            _check_lbfgs(_lbfgs_lib().tmol_lbfgs_combine(
                self.P, self.m, self.S, n_phys, _vp(g.contiguous()), _vp(self.Sh),
                _vp(c1.contiguous()), _vp(self.Yh), _vp(c2.contiguous()), _vp(out),
                _stream_of(g)), "lbfgs combine")
            return out
        Sv, Yv = self.Sh[:, :n_phys], self.Yh[:, :n_phys]
        return g + torch.bmm(c1.unsqueeze(1), Sv).squeeze(1) + torch.bmm(c2.unsqueeze(1), Yv).squeeze(1)

    # -- history ------------------------------------------------------------------
    # (s, y) pairs live in a ring of m physical slots; ``order`` lists the
    # slots from oldest to newest.  The small Gram matrices are kept in that
    # chronological order, so only they move when the oldest entry is dropped.
    def _reset_history(self, dev):
        m, P, S = self.m, self.P, self.S
        # [P, m, S]: each segment's history rows contiguous, so the products
        # with the history are batched matrix-vector products over segments
        self.Sh = torch.zeros((P, m, S), dtype=torch.float64, device=dev)
        self.Yh = torch.zeros((P, m, S), dtype=torch.float64, device=dev)
        self.SY = torch.zeros((P, m, m), dtype=torch.float64, device=dev)  # s_i . y_j
        self.YY = torch.zeros((P, m, m), dtype=torch.float64, device=dev)  # y_i . y_j
        self.order = []
        self.n_hist = 0
        self._idx = None

    def _used(self):
        """Physical slots in use (a prefix until the ring wraps) and the
        chronological order as indices into them."""
        n_phys = self.m if self.n_hist == self.m else self.n_hist
        if self._idx is None:
            self._idx = torch.tensor(self.order, dtype=torch.int64, device=self.Sh.device)
        return n_phys, self._idx

    def _push(self, s: torch.Tensor, y: torch.Tensor):
        """Append one (s, y) pair [P, S] (rows of segments that did not keep
        it are zero) as the newest history entry."""
        if self.n_hist == self.m:
            phys = self.order.pop(0)
            self.SY = torch.roll(self.SY, (-1, -1), (1, 2))
            self.YY = torch.roll(self.YY, (-1, -1), (1, 2))
            k = self.m - 1
        else:
            phys = self.n_hist
            k = self.n_hist
            self.n_hist += 1
        self.order.append(phys)
        self._idx = None
        self.Sh[:, phys] = s
        self.Yh[:, phys] = y
        n_phys, idx = self._used()
        s_y = self._dots(self.Sh, n_phys, [y])[0][:, idx]  # s_i . y_new, chronological
        y_s, y_y = self._dots(self.Yh, n_phys, [s, y])[:, :, idx]  # y_i . s_new, y_i . y_new
        n = self.n_hist
        self.SY[:, :n, k] = s_y
        self.SY[:, k, :n] = y_s
        self.YY[:, :n, k] = y_y
        self.YY[:, k, :n] = y_y

    def _drop_segments(self, mask: torch.Tensor):
        """Forget the history of the segments in ``mask`` (device mask; no
        host synchronisation)."""
        keep = (~mask).to(torch.float64)
        self.Sh.mul_(keep[:, None, None])
        self.Yh.mul_(keep[:, None, None])
        self.SY.mul_(keep[:, None, None])
        self.YY.mul_(keep[:, None, None])

    def _hg(self, g: torch.Tensor) -> torch.Tensor:
        """H g for padded g [P, S] (float64), compact representation."""
        n = self.n_hist
        if n == 0:
            return g.clone()
        n_phys, idx = self._used()
        a = self._dots(self.Sh, n_phys, [g])[0][:, idx]
        b = self._dots(self.Yh, n_phys, [g])[0][:, idx]
        SY = self.SY[:, :n, :n]
        D = torch.diagonal(SY, dim1=1, dim2=2)
        # absent entries (zero rows) get a unit diagonal: they then contribute 0
        R = torch.triu(SY) + torch.diag_embed((D == 0).to(SY.dtype))
        u = torch.linalg.solve_triangular(R, a.unsqueeze(-1), upper=True).squeeze(-1)
        v = torch.bmm(self.YY[:, :n, :n], u.unsqueeze(-1)).squeeze(-1) + D * u - b
        p1 = torch.linalg.solve_triangular(R.transpose(1, 2), v.unsqueeze(-1),
                                           upper=False).squeeze(-1)
        # back to physical slot order for the products with the history
        coef = torch.zeros((2, self.P, n_phys), dtype=g.dtype, device=g.device)
        coef[0][:, idx] = p1
        coef[1][:, idx] = -u
        return self._combine(g, n_phys, coef[0], coef[1])

    # -- line search ------------------------------------------------------------------
    def _armijo(self, phi_at, gtd, phi0, alpha0, searching):
        """Per-segment Armijo search; returns (accepted step, energies at it,
        evaluations, failed mask, whether the last evaluation was at the
        accepted steps).  The one host synchronisation per trial is the
        decision whether to evaluate again."""
        zero = torch.zeros_like(alpha0)
        alpha = torch.where(searching, alpha0, zero)
        phi = phi_at(alpha)
        evals = 1
        linear = phi <= phi0 + alpha * 0.8 * gtd
        enough = phi <= phi0 + alpha * 0.1 * gtd
        grow = searching & linear
        shrink = searching & ~linear & ~enough
        ok = searching & (linear | enough)
        step = torch.where(ok, alpha, zero)
        phi_step = torch.where(ok, phi, phi0)
        failed = torch.zeros_like(searching)
        while True:
            more, at_step = torch.stack([(grow | shrink).any(), (alpha == step).all()]).tolist()
            if not more:
                break
            trial = torch.where(grow, alpha * 2.0, step)
            trial = torch.where(shrink, alpha * 0.25, trial)
            phi_t = phi_at(trial)
            evals += 1
            longer = grow & (phi_t < phi)
            step = torch.where(longer, trial, step)
            phi_step = torch.where(longer, phi_t, phi_step)
            grow = torch.zeros_like(grow)
            armijo = shrink & (phi_t <= phi0 + trial * 0.1 * gtd)
            step = torch.where(armijo, trial, step)
            phi_step = torch.where(armijo, phi_t, phi_step)
            low = shrink & ~armijo & (trial < self.minstep)
            downhill = low & (phi_t < phi0)
            step = torch.where(downhill, trial, step)
            phi_step = torch.where(downhill, phi_t, phi_step)
            fail = low & ~downhill
            step = torch.where(fail, zero, step)
            phi_step = torch.where(fail, phi0, phi_step)
            failed = failed | fail
            shrink = shrink & ~armijo & ~low
            alpha, phi = trial, phi_t
        return step, phi_step, evals, failed, bool(at_step)

    # -- driver ------------------------------------------------------------------------
    def run(self, x0: torch.Tensor, evaluate: Evaluate) -> Tuple[torch.Tensor, Trace]:
        """Host synchronisations: one per line-search trial (whether to
        evaluate again) and one per iteration (convergence, and whether the
        next iteration adds a history pair or resets a segment); the trace
        stays on the device until the run ends.  The vectors live in the
        padded per-segment layout [P, S] (pads hold 0 in x, g and d), so
        segment sums are row reductions and per-segment scalars broadcast
        along rows; ``evaluate`` still sees flat vectors."""
        f64 = torch.float64
        x0f = x0.detach().reshape(-1)
        x = self._padded(x0f, dtype=x0f.dtype)
        dev = x.device
        eps = torch.finfo(x.dtype).eps
        rtol = self.rtol if self.rtol is not None else math.sqrt(eps)
        atol = self.atol if self.atol is not None else math.sqrt(eps)
        self._reset_history(dev)
        P = self.P
        flags = dict(dtype=torch.bool, device=dev)
        converged = torch.zeros(P, **flags)
        stalled = torch.zeros(P, **flags)
        needs_reset = torch.zeros(P, **flags)
        trace = Trace()
        rec_E, rec_gn, rec_gm = [], [], []

        def ev(xp):
            E_, gf = evaluate(self._flat(xp))
            gf = gf.reshape(-1)
            return E_.to(f64), self._padded(gf, dtype=gf.dtype)

        def ssum(v):  # per-segment sum, float64
            return v.to(f64).sum(1)

        def smax_abs(v):  # per-segment max |v| (pads are 0)
            return v.abs().to(f64).amax(1)

        def col(v):  # a per-segment value broadcast over the segment's elements
            return v[:, None]

        E, g = ev(x)

        def record(E, g, evals):
            rec_E.append(E.detach())
            rec_gn.append(ssum(g * g).sqrt())
            rec_gm.append(smax_abs(g))
            trace.evals.append(evals)

        record(E, g, 1)
        t = torch.full((P,), self.lr, dtype=f64, device=dev)
        g_prev = None
        x_ref = x.clone()
        keep = None
        keep_any = reset_any = False
        it = 0
        while it < self.max_iter:
            it += 1
            # ---- search direction
            if g_prev is None:
                d = -g
            else:
                if keep_any:
                    ke = col(keep)
                    y = g - g_prev
                    s = x - x_ref
                    self._push(torch.where(ke, s, torch.zeros_like(s)).to(f64),
                               torch.where(ke, y, torch.zeros_like(y)).to(f64))
                    x_ref = torch.where(ke, x, x_ref)
                d = -self._hg(g.to(f64)).to(x.dtype)
            was_reset = needs_reset.clone()
            if reset_any:
                self._drop_segments(needs_reset)
                re = col(needs_reset)
                d = torch.where(re, -g, d)
                x_ref = torch.where(re, x, x_ref)
                needs_reset = torch.zeros_like(needs_reset)
            inactive = converged | stalled
            d = d * (~col(inactive)).to(d.dtype)
            g_prev = g.clone()
            E_prev = E.clone()
            # ---- directional derivative, with repair of uphill directions
            # (flip the uphill components, then steepest descent), masked per
            # segment: a repair leaves segments without an uphill direction
            # untouched
            gtd = ssum(g * d)
            for repair in (1, 2):
                bad = (gtd > -1e-5) & ~inactive
                be = col(bad)
                fixed = d * -torch.sign(g * d) if repair == 1 else -g
                d = torch.where(be, fixed, d)
                gtd = torch.where(bad, ssum(g * d), gtd)
            # ---- line search
            searching = ~inactive & (gtd < 0)
            t0 = torch.clamp(t / 0.5, max=1.0)
            x_base = x.clone()
            last = {}

            def phi_at(alpha):
                xt = x_base + d * col(alpha).to(x.dtype)
                Et, gt = ev(xt)
                last["x"], last["E"], last["g"] = xt, Et, gt
                return Et

            step, E_step, evals, failed, at_step = self._armijo(phi_at, gtd, E, t0, searching)
            x = x_base + d * col(step).to(x.dtype)
            if at_step:
                E, g = last["E"], last["g"]
            else:
                E, g = ev(x)
                evals += 1
            retry = torch.clamp((-gtd).clamp(min=self.minstep).rsqrt(), max=1.0)
            t = torch.where(failed, retry, step)
            t = torch.where(searching, t, t0)
            stalled = stalled | (failed & was_reset)
            needs_reset = needs_reset | (failed & ~was_reset)
            record(E, g, evals)
            # ---- convergence; and, for the next iteration, which segments
            # keep the new (s, y) pair (s.y > 1e-6), fetched in the same
            # synchronisation
            converged = converged | (smax_abs(g) <= self.gradtol)
            dE = (E - E_prev).abs()
            rdiff = 2 * dE / (E.abs() + E_prev.abs() + 1e-10)
            done = converged | stalled | (dE <= atol) | (rdiff <= rtol)
            keep = ssum((g - g_prev) * (x - x_ref)) > 1e-6
            all_done, keep_any, reset_any = torch.stack(
                [done.all(), keep.any(), needs_reset.any()]).tolist()
            if self.verbose:
                print("iter %4d  E=%.6f  evals=%d  done=%d/%d" % (it, float(E.sum()), evals,
                                                                  int(done.sum()), P))
            if all_done:
                break
        trace.iterations = it
        trace.converged = converged.cpu()
        trace.stalled = stalled.cpu()
        n = len(rec_E)
        for name, rec in (("energy", rec_E), ("grad_norm", rec_gn), ("grad_max", rec_gm)):
            getattr(trace, name).extend(torch.stack(rec).cpu().unbind(0) if n else [])
        return self._flat(x), trace


# ---------------------------------------------------------------------------
# evaluators


def ragged_evaluator(scorer, weights=(1.0, 0.55, 1.0), graph: bool = True) -> Evaluate:
    """Energy + weighted gradient of a ``LJLKRaggedScorer`` (fused kernel).
    With ``graph`` the neighbor pass and the scoring launch are captured once
    in a hipGraph and each evaluation copies the coordinates in and replays
    it (the scorer's default work-list capacity cannot overflow, whatever the
    trial coordinates)."""
    w = torch.tensor(weights, dtype=torch.float64)
    wd = w.to(scorer.device)
    wk = scorer._weights(w)
    state = {"captured": False}

    def evaluate(x):
        coords = x.view(-1, 3)
        if graph and coords.is_cuda:
            if not state["captured"]:
                scorer.capture(coords, wk)
                state["captured"] = True
            E, g = scorer.replay(coords)
        else:
            E, g = scorer.energy_and_grad(coords, wk)
        return (wd[:, None] * E.double()).sum(0), g.reshape(-1).clone()

    return evaluate


def op_evaluator(op_args: list, weights=(1.0, 0.55, 1.0), op=None) -> Evaluate:
    """Energy + gradient through an ``ljlk_pose_scores``-schema op and
    autograd (the drop-in path: forward kernel, saved dV/dx, backward
    kernel).  ``op_args`` are the op's 27 arguments; argument 0 is replaced by
    the coordinates being evaluated."""
    if op is None:
        from .ops import ljlk_pose_scores as op
    args = list(op_args)
    w = torch.tensor(weights, dtype=torch.float64)

    def evaluate(x):
        coords = x.detach().view(-1, 3).clone().requires_grad_(True)
        args[0] = coords
        score = op(*args)[0]
        Ew = (w.to(score.device, score.dtype)[:, None] * score).sum(0)
        Ew.sum().backward()
        return Ew.detach().double(), coords.grad.detach().reshape(-1)

    return evaluate


def main(argv=None):
    """Minimize a workload's poses with LJ+LK on the GPU; print a JSON summary.

        python -m tmol_amd.minimize --workload diff_32x400 --iters 200 \\
            [--dtype f32|f64] [--path fused|autograd] [--trace out.json]
    """
    import argparse
    import json
    import time

    from .generator import workload_poses
    from .pose import RaggedPoseStack
    from .score.ljlk_energy_term import LJLKEnergyTerm
    from .scorer import LJLKRaggedScorer, ragged_pose_scores

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="diff_32x400")
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--path", default="autograd", choices=["fused", "autograd"])
    ap.add_argument("--trace", default=None)
    args = ap.parse_args(argv)

    dev = torch.device("cuda:0")
    dtype = torch.float32 if args.dtype == "f32" else torch.float64
    poses = workload_poses(args.workload)
    stack = RaggedPoseStack(poses, dtype=dtype)
    scorer = LJLKRaggedScorer(stack, LJLKEnergyTerm(device=dev), device=dev, dtype=dtype)
    x0 = stack.coords.to(dev).reshape(-1)
    seg = stack.atom_pose.to(dev).repeat_interleave(3).long()
    w = torch.tensor([1.0, 0.55, 1.0], dtype=torch.float64, device=dev)
    if args.path == "fused":
        evaluate = ragged_evaluator(scorer)
    else:
        def evaluate(x):
            c = x.detach().view(-1, 3).clone().requires_grad_(True)
            E = ragged_pose_scores(scorer, c)
            Ew = (w.to(E.dtype)[:, None] * E).sum(0)
            Ew.sum().backward()
            return Ew.detach().double(), c.grad.reshape(-1)
    opt = SegmentedLBFGS(seg, max_iter=args.iters)
    evaluate(x0)  # warm up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, tr = opt.run(x0, evaluate)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    E = tr.energies()
    n_res = torch.tensor([p.n_blocks for p in poses], dtype=torch.float64)
    out = {
        "workload": args.workload, "dtype": args.dtype, "path": args.path,
        "poses": len(poses), "atoms": stack.n_atoms_total,
        "iterations": tr.iterations, "evaluations": int(sum(tr.evals)),
        "seconds": secs, "ms_per_iteration": 1e3 * secs / max(tr.iterations, 1),
        "energy_start": float(E[0].sum()), "energy_end": float(E[-1].sum()),
        "energy_per_residue_end": (E[-1] / n_res).tolist(),
        "grad_norm_start": float(tr.grad_norms()[0].norm()),
        "grad_norm_end": float(tr.grad_norms()[-1].norm()),
        "converged": int(tr.converged.sum()), "stalled": int(tr.stalled.sum()),
    }
    if args.trace:
        with open(args.trace, "w") as f:
            json.dump({**out, "energy": [e.tolist() for e in tr.energy],
                       "grad_norm": [g.tolist() for g in tr.grad_norm], "evals": tr.evals}, f)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
