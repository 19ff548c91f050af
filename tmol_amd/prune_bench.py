"""Train, calibrate and measure the learned block-pair pruner.

    python -m tmol_amd.prune_bench [--epochs 6] [--steps 20]

1. Training set: poses of the 512-pose batch's distribution (own seed); the
   exact kernel labels every neighbor block pair (|E| < 1e-3 kcal/mol);
   1e6 pairs are sampled and the MLP is trained by the library's kernels.
2. Calibration: on 128 other poses of the held-out batch's size, the
   smallest threshold (most pruning) whose worst per-pose total-energy error
   stays within a quarter of the bound is chosen.
3. Held-out batch: 1024 poses of 300 residues.  Exact and pruned
   energy + gradient steps are each captured in a hipGraph and timed; the
   exact kernel judges the pruned per-pose energies; a batch whose error
   exceeds 0.05 kcal/mol/residue falls back to the exact step.
"""

import argparse
import json
import sys
import time

import torch

from . import prune
from .generator import workload_poses
from .pose import RaggedPoseStack
from .score.ljlk_energy_term import LJLKEnergyTerm
from .scorer import LJLKRaggedScorer

BOUND = 0.05  # kcal/mol per residue, total energy
MARGIN = 0.25  # calibration: worst calibration pose within a quarter of the bound
W = torch.tensor(prune.WEIGHTS, dtype=torch.float64)


def _scorer(poses, dev):
    st = RaggedPoseStack(poses)
    return st, LJLKRaggedScorer(st, LJLKEnergyTerm(device=dev), device=dev)


def _weighted(E):
    return (W.to(E.device)[:, None] * E.double()).sum(0)


def train_pruner(dev, n_pairs=1_000_000, epochs=6, seed=0, log=print, n_poses=None):
    poses = workload_poses("pruner_train", n_poses)
    st, sc = _scorer(poses, dev)
    x = st.coords.to(dev)
    pairs, labels, Ew = prune.build_training_set(sc, x)
    inputs = prune.PruneInputs(sc)
    inputs.update(x)
    n_all = pairs.shape[0]
    g = torch.Generator().manual_seed(seed)
    sel = torch.randperm(n_all, generator=g)[:min(n_pairs, n_all)].to(dev)
    pairs, labels, Ew = pairs[sel].contiguous(), labels[sel].contiguous(), Ew[sel].contiguous()
    sw = prune.energy_sample_weights(Ew)
    log("training pairs %d of %d (negligible %.3f)" % (pairs.shape[0], n_all, float(labels.mean())))
    params = prune.PrunerParams(dev, seed=seed)
    t0 = time.perf_counter()
    hist = prune.train(params, inputs, pairs, labels, epochs=epochs, log=log, sample_weights=sw)
    torch.cuda.synchronize(dev)
    t_train = time.perf_counter() - t0
    _, acc = prune.evaluate_classifier(params, inputs, pairs, labels)
    return params, {"pairs": int(pairs.shape[0]), "negligible_fraction": float(labels.mean()),
                    "epochs": epochs, "loss": hist, "train_accuracy": acc, "train_seconds": t_train}


def pose_errors(pruner, sc, x, n_res):
    """Per-pose |E_pruned - E_exact| / residues (weighted total energy) and
    the fraction of pairs the pruner removed."""
    e_exact = _weighted(sc.energy_and_grad(x)[0].clone())
    sc.neighbors(x)
    n_before = int(sc.counters()[0])
    pruner.prune(x)
    cnt, _ = sc.ws.rows(sc.cstack, sc.bt.c)
    n_after = int(cnt.to(torch.int64).clamp(max=sc.capacity).sum())
    w = sc._weights(None)
    sc.score(x, 4, weights=w, energies=sc.energies, grad=sc.grad)
    e_pruned = _weighted(sc.energies.clone())
    err = (e_pruned - e_exact).abs().cpu() / n_res
    return err, 1.0 - n_after / max(n_before, 1)


def calibrate(params, dev, candidates=(0.5, 0.7, 0.8, 0.9, 0.95, 0.98, 0.99, 0.995, 0.999),
              min_gap=prune.MIN_GAP,
              log=print, n_poses=None):
    poses = workload_poses("pruner_calib", n_poses)
    st, sc = _scorer(poses, dev)
    x = st.coords.to(dev)
    n_res = torch.tensor([p.n_blocks for p in poses], dtype=torch.float64)
    pr = prune.Pruner(sc, params, min_gap=min_gap)
    table = []
    best = 1.0
    for tau in sorted(candidates, reverse=True):
        pr.set_threshold(tau)
        err, frac = pose_errors(pr, sc, x, n_res)
        table.append({"threshold": tau, "max_err_per_res": float(err.max()),
                      "mean_err_per_res": float(err.mean()), "pruned_fraction": frac})
        log("calibration tau %.2f: max err/res %.4f, pruned %.3f" % (tau, float(err.max()), frac))
        if float(err.max()) > MARGIN * BOUND:
            break  # lower thresholds prune more: stop at the first failure
        best = tau
    return best, table


def _graph(fn):
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    return g


def _time(g, steps, warmup):
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def measure(params, threshold, dev, steps=20, warmup=3, log=print, min_gap=prune.MIN_GAP,
            n_poses=None):
    poses = workload_poses("pruner_1024x300", n_poses)
    st, sc = _scorer(poses, dev)
    x = st.coords.to(dev).contiguous()
    n_res = torch.tensor([p.n_blocks for p in poses], dtype=torch.float64)
    w = sc._weights(None)
    pr = prune.Pruner(sc, params, threshold=threshold, min_gap=min_gap)
    sc.energy_and_grad(x)
    sc.check_capacity()

    def exact():
        sc.neighbors(x)
        sc.score(x, 4, weights=w, energies=sc.energies, grad=sc.grad)

    def pruned():
        sc.neighbors(x)
        pr.prune(x)
        sc.score(x, 4, weights=w, energies=sc.energies, grad=sc.grad)

    g_exact = _graph(exact)
    t_exact = _time(g_exact, steps, warmup)
    g_exact.replay()
    torch.cuda.synchronize()
    e_exact = _weighted(sc.energies.clone())
    g_pruned = _graph(pruned)
    t_pruned = _time(g_pruned, steps, warmup)
    g_pruned.replay()
    torch.cuda.synchronize()
    e_pruned = _weighted(sc.energies.clone())
    err = (e_pruned - e_exact).abs().cpu() / n_res
    sc.neighbors(x)
    n_before = int(sc.counters()[0])
    pr.prune(x)
    cnt, _ = sc.ws.rows(sc.cstack, sc.bt.c)
    n_after = int(cnt.to(torch.int64).clamp(max=sc.capacity).sum())
    within = bool(float(err.max()) <= BOUND)
    P = len(poses)
    # kernel-level times of the scoring kernel on the exact and pruned lists
    w_ = w

    def kernel_ms(prep):
        sc.neighbors(x)
        prep()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            sc.score(x, 4, weights=w_, energies=sc.energies, grad=sc.grad)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    k_exact = kernel_ms(lambda: None)
    k_pruned = kernel_ms(lambda: pr.prune(x))
    # the bound on what any classifier can gain: remove exactly the pairs the
    # exact kernel finds negligible (|E_pair| < 1e-3 kcal/mol, weighted),
    # through the same flattening and compaction kernels
    Ew, labels = prune.pair_labels(sc, x)
    pairs_all = prune.work_list_pairs(sc)
    perfect_keep = ((labels == 0) | (pairs_all[:, 0] == pairs_all[:, 1])).to(torch.uint8)
    perfect_keep = torch.cat([perfect_keep, torch.ones(pr.keep.numel() - perfect_keep.numel(),
                                                       dtype=torch.uint8, device=dev)])

    def perfect():
        sc.neighbors(x)
        pr.compact_with(x, perfect_keep)
        sc.score(x, 4, weights=w, energies=sc.energies, grad=sc.grad)

    g_perfect = _graph(perfect)
    t_perfect = _time(g_perfect, steps, warmup)
    g_perfect.replay()
    torch.cuda.synchronize()
    e_perfect = _weighted(sc.energies.clone())
    err_perfect = (e_perfect - e_exact).abs().cpu() / n_res
    k_perfect = kernel_ms(lambda: pr.compact_with(x, perfect_keep))
    n_perfect = int(perfect_keep[:pairs_all.shape[0]].to(torch.int64).sum())
    # inference alone (flatten + network + compaction)
    sc.neighbors(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        sc.neighbors(x)
        pr.prune(x)
    b.record()
    torch.cuda.synchronize()
    t_nbr_prune = a.elapsed_time(b) / steps
    a.record()
    for _ in range(steps):
        sc.neighbors(x)
    b.record()
    torch.cuda.synchronize()
    t_nbr = a.elapsed_time(b) / steps
    out = {
        "workload": "pruner_1024x300", "poses": P, "atoms": st.n_atoms_total,
        "threshold": threshold, "min_gap": min_gap,
        "exact_poses_per_s": P / t_exact, "pruned_poses_per_s": P / t_pruned,
        "speedup": t_exact / t_pruned,
        "ms_exact": 1e3 * t_exact, "ms_pruned": 1e3 * t_pruned,
        "pairs_exact": n_before, "pairs_pruned": n_after,
        "pruned_fraction": 1.0 - n_after / max(n_before, 1),
        "max_err_per_res": float(err.max()), "mean_err_per_res": float(err.mean()),
        "within_bound": within, "fallback": not within,
        # a batch over the bound is rescored exactly, and a pruned step slower
        # than the exact one is not used (the cost check): either way the
        # throughput is the exact one
        "pays_off": bool(t_pruned < t_exact),
        "effective_poses_per_s": P / t_pruned if (within and t_pruned < t_exact) else P / t_exact,
        "kernel_ms_exact": k_exact, "kernel_ms_pruned": k_pruned,
        "prune_ms": t_nbr_prune - t_nbr,
        "perfect_labels": {
            "ms": 1e3 * t_perfect, "speedup": t_exact / t_perfect,
            "kernel_ms": k_perfect, "pairs": n_perfect,
            "pruned_fraction": 1.0 - n_perfect / max(n_before, 1),
            # share of the scoring kernel's time the negligible pairs cost
            "negligible_cost_share": 1.0 - k_perfect / k_exact,
            "max_err_per_res": float(err_perfect.max()),
        },
    }
    log(json.dumps(out))
    return out


def run(dev=None, epochs=6, steps=20, warmup=3, log=print):
    dev = dev or torch.device("cuda:0")
    params, train_report = train_pruner(dev, epochs=epochs, log=log)
    tau, table = calibrate(params, dev, log=log)
    res = measure(params, tau, dev, steps=steps, warmup=warmup, log=log)
    res["training"] = train_report
    res["calibration"] = table
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--steps", type=int, default=20)
    args = ap.parse_args(argv)
    res = run(epochs=args.epochs, steps=args.steps,
              log=lambda m: print("[prune] " + m, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
