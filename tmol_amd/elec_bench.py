"""Elec benchmark: energy + gradient of the elec term on the 512-pose ragged
batch (the same poses as the LJ+LK headline), on the same scheduler.

One step = elec's neighbor pass + the elec row kernel in weighted-gradient
mode (per-pose energies and sum_p w[p] dE_p/dx), inputs resident in HBM.
Reported: poses/s of the step, the kernel alone (HIP events on its stream),
its roofline from the atom-pair census x elec FLOP table
(tmol_amd/flops.py), and, when ``cpu`` is given (bench.py passes the
reference's own CPU elec op, one thread, forward + backward), a CPU baseline
on a bounded sample of the poses.  The ``fused``
entry times LJ+LK and elec scored in one pass against the two passes run
one after the other.

    python -m tmol_amd.elec_bench [--steps 20]      (GPU legs only)
"""

import argparse
import json
import os
import sys
import time

FP32_PEAK_TFLOPS = 157.3


def run(device, poses=None, steps=20, warmup=3, kernel_reps=10, cpu=None, log=print):
    import numpy
    import torch

    from .flops import elec_algorithmic_flops, elec_far_pairs
    from .generator import workload_poses
    from .pose import RaggedPoseStack
    from .score.elec_energy_term import ElecEnergyTerm
    from .scorer import ElecRaggedScorer

    if poses is None:
        poses = workload_poses("ragged_512x50_1200")
    stack = RaggedPoseStack(poses)
    term = ElecEnergyTerm()
    scorer = ElecRaggedScorer(stack, term, device=device)
    coords = stack.coords.to(device)
    w = scorer._weights(1.0)
    scorer.neighbors(coords)
    n_items = scorer.check_capacity()
    census = scorer.census(coords)
    flops = elec_algorithmic_flops(census)
    stream = torch.cuda.current_stream(device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    scorer.neighbors(coords)
    for _ in range(3):
        scorer.score(coords, 4, weights=w, energies=scorer.energies, grad=scorer.grad)
    torch.cuda.synchronize(device)
    ev0.record(stream)
    for _ in range(kernel_reps):
        scorer.score(coords, 4, weights=w, energies=scorer.energies, grad=scorer.grad)
    ev1.record(stream)
    torch.cuda.synchronize(device)
    score_ms = ev0.elapsed_time(ev1) / kernel_reps  # + 2 fill and 1 finalize kernels
    ev0.record(stream)
    for _ in range(kernel_reps):
        scorer.neighbors(coords)
    ev1.record(stream)
    torch.cuda.synchronize(device)
    nbr_ms = ev0.elapsed_time(ev1) / kernel_reps

    for _ in range(warmup):
        scorer.energy_and_grad(coords, w)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        scorer.energy_and_grad(coords, w)
    torch.cuda.synchronize(device)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    e = scorer.energies.double().cpu()
    achieved = flops / (score_ms * 1e-3) / 1e12
    traffic = None
    prof = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                        "elec_kernel_ragged512_traffic.json")
    if os.path.exists(prof):
        try:
            traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    res = {
        "workload": "ragged_512x50_1200",
        "poses": stack.n_poses,
        "value": stack.n_poses / (ms * 1e-3),
        "unit": "poses/s",
        "ms_per_step": ms,
        "neighbors_ms": nbr_ms,
        "score_ms": score_ms,
        "block_pairs": n_items,
        "atom_pairs": census["pairs"],
        "charged_pairs_within_max_dis": census["flat"] + census["low"] + census["coulomb"]
        + census["high"],
        "charged_pairs_beyond": elec_far_pairs(census),
        "energies_finite": bool(torch.isfinite(e).all()),
        "dtype": "f32",
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "elec_rows_kernel<float, WGRAD> (FP32 vector ALU)",
            "algorithmic_flops_per_launch": flops,
        },
        "census": census,
    }
    res["fused"] = fused_leg(stack, coords, device, steps, warmup, kernel_reps, log)
    if cpu is not None:
        log("elec cpu baseline ...")
        res["cpu_baseline"] = cpu(poses)
    return res


def fused_leg(stack, coords, device, steps=20, warmup=3, kernel_reps=10, log=print):
    """LJ+LK and elec energy + gradient (beta2016 weights) as one neighbor pass
    and one scoring launch, against the two terms as separate passes; every
    step captured in a hipGraph."""
    import torch

    from .score.elec_energy_term import ElecEnergyTerm
    from .score.ljlk_energy_term import LJLKEnergyTerm
    from .scorer import ElecRaggedScorer, LJLKElecRaggedScorer, LJLKRaggedScorer

    lj_term = LJLKEnergyTerm(device=device)
    el_term = ElecEnergyTerm()
    fused = LJLKElecRaggedScorer(stack, lj_term, el_term, device=device)
    lj = LJLKRaggedScorer(stack, lj_term, device=device)
    el = ElecRaggedScorer(stack, el_term, device=device)
    w4 = fused._weights(None)
    w3 = lj._weights(None)
    we = el._weights(1.0)

    def both():
        lj.neighbors(coords)
        lj.score(coords, 4, weights=w3, energies=lj.energies, grad=lj.grad)
        el.neighbors(coords)
        el.score(coords, 4, weights=we, energies=el.energies, grad=el.grad)

    def one():
        fused.neighbors(coords)
        fused.score(coords, 4, weights=w4, energies=fused.energies, grad=fused.grad)

    def timed(fn):
        fn()
        torch.cuda.synchronize(device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device)
        s.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                fn()
        torch.cuda.current_stream(device).wait_stream(s)
        for _ in range(warmup):
            g.replay()
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize(device)
        return (time.perf_counter() - t0) * 1000.0 / steps

    ms_sep = timed(both)
    ms_fused = timed(one)
    stream = torch.cuda.current_stream(device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fused.neighbors(coords)
    fused.score(coords, 4, weights=w4, energies=fused.energies, grad=fused.grad)
    torch.cuda.synchronize(device)
    ev0.record(stream)
    for _ in range(kernel_reps):
        fused.score(coords, 4, weights=w4, energies=fused.energies, grad=fused.grad)
    ev1.record(stream)
    torch.cuda.synchronize(device)
    score_ms = ev0.elapsed_time(ev1) / kernel_reps
    e = fused.energies.double()
    log("fused LJ+LK+elec %.3f ms, separate %.3f ms" % (ms_fused, ms_sep))
    return {
        "workload": "ragged_512x50_1200",
        "terms": "fa_ljatr, fa_ljrep, fa_lk, fa_elec (weights 1, 0.55, 1, 1)",
        "value": stack.n_poses / (ms_fused * 1e-3),
        "unit": "poses/s",
        "ms_per_step": ms_fused,
        "score_ms": score_ms,
        "separate_ms_per_step": ms_sep,
        "separate_poses_per_s": stack.n_poses / (ms_sep * 1e-3),
        "speedup_vs_separate": ms_sep / ms_fused,
        "energies_finite": bool(torch.isfinite(e).all()),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="accepted; the CPU baseline runs from bench.py")
    args = ap.parse_args()
    import torch

    res = run(torch.device("cuda:0"), steps=args.steps, cpu=None,
              log=lambda m: print("[elec] " + m, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
