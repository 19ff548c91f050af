"""Rotamer-pair scoring benchmark (packing's energy table).

Workload: 16 poses x 300 residues (the repack_16x300 generator workload),
every residue with repacking rotamers (tmol_amd.rotamers, seed 1): 66.5k
rotamers, 25.1M rotamer pairs.  One step = the whole forward of
``ljlk_rotamer_scores``: rotamer and block spheres, neighbor rows, the host
read of the totals, the dispatch list and the scoring kernel, inputs
resident in HBM.

Reported: rotamer pairs/s of the step; the scoring kernel alone (HIP events
on its stream around repeated launches over a fixed dispatch list) with its
roofline from the branch census x energy-only FLOP table; and, when ``cpu``
is given (bench.py passes the reference's own CPU op, one thread, on the
first pose), the CPU baseline.
The ``elec`` entry does the same for ``elec_rotamer_scores`` on the same
rotamers (plus its dTdV-weighted backward kernel).

    python -m tmol_amd.rotamer_bench [--steps 10]      (GPU legs only)
"""

import argparse
import json
import os
import sys
import time

FP32_PEAK_TFLOPS = 157.3


def _args(ps, rs, dt, device):
    import torch

    from .rotamers import rotamer_op_args
    from .score.ljlk_energy_term import LJLKEnergyTerm

    a = rotamer_op_args(ps, rs, LJLKEnergyTerm())
    if dt == torch.float64:
        a[23], a[24] = a[23].double(), a[24].double()
    return [x.to(device) if isinstance(x, torch.Tensor) else x for x in a]


def run(device, steps=10, warmup=2, kernel_reps=10, cpu=None, elec_cpu=None, log=print):
    import numpy
    import torch

    from .flops import CENSUS_FIELDS, algorithmic_flops, pair_evaluations
    from .ops import RotamerCall, ljlk_rotamer_scores
    from .rotamers import repack_workload

    t0 = time.time()
    ps, rs = repack_workload("repack_16x300", seed=1)
    log("rotamers: %d poses, %d rotamers, %d atoms (%.1fs)"
        % (ps.n_poses, rs.n_rotamers_total, rs.coords.shape[0], time.time() - t0))
    a = _args(ps, rs, torch.float32, device)
    stream = torch.cuda.current_stream(device)

    # the whole forward, as the op runs it
    for _ in range(warmup):
        s, idx = ljlk_rotamer_scores(*a)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        s, idx = ljlk_rotamer_scores(*a)
    torch.cuda.synchronize(device)
    op_ms = (time.perf_counter() - t0) * 1000.0 / steps
    n_pairs = int(idx.shape[1])
    finite = bool(torch.isfinite(s).all())

    # the scoring kernel alone over a fixed dispatch list
    call = RotamerCall(a, torch.float32)
    call.dispatch()
    out = torch.empty((3, call.n_pairs), dtype=torch.float32, device=device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(2):
        call.energies(out)
    torch.cuda.synchronize(device)
    ev0.record(stream)
    for _ in range(kernel_reps):
        call.energies(out)
    ev1.record(stream)
    torch.cuda.synchronize(device)
    score_ms = ev0.elapsed_time(ev1) / kernel_reps  # + one 2 KB cursor fill per launch
    ev0.record(stream)
    for _ in range(kernel_reps):
        call.dispatch()
    ev1.record(stream)
    torch.cuda.synchronize(device)
    dispatch_ms = ev0.elapsed_time(ev1) / kernel_reps  # includes the host read of the totals
    census = dict(zip(CENSUS_FIELDS, call.census()))
    flops = algorithmic_flops(census, energy_only=True)
    evals = pair_evaluations(census)
    achieved = flops / (score_ms * 1e-3) / 1e12
    traffic = None
    prof = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                        "rotamer_kernel_repack16x300_traffic.json")
    if os.path.exists(prof):
        try:
            traffic = json.load(open(prof)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    res = {
        "workload": "repack_16x300",
        "poses": ps.n_poses,
        "rotamers": rs.n_rotamers_total,
        "rotamer_pairs": n_pairs,
        "value": n_pairs / (op_ms * 1e-3),
        "unit": "rotamer pairs/s",
        "ms_per_step": op_ms,
        "dispatch_ms": dispatch_ms,
        "score_ms": score_ms,
        "atom_pair_evals": evals,
        "atom_pair_evals_per_s": evals / (score_ms * 1e-3),
        "energies_finite": finite,
        "dtype": "f32",
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "rotamer_energy_kernel<float> (FP32 vector ALU)",
            "algorithmic_flops_per_launch": flops,
        },
        "census": census,
    }
    if cpu is not None:
        log("rotamer cpu baseline ...")
        res["cpu_baseline"] = cpu()
    log("elec rotamer pairs ...")
    res["elec"] = elec_leg(device, ps, rs, steps, warmup, kernel_reps, elec_cpu, log)
    return res


def _elec_args(ps, rs, device):
    import torch

    from .rotamers import rotamer_op_args
    from .score.elec_energy_term import ElecEnergyTerm

    a = rotamer_op_args(ps, rs, ElecEnergyTerm())
    return [x.to(device) if isinstance(x, torch.Tensor) else x for x in a]


def elec_leg(device, ps, rs, steps=10, warmup=2, kernel_reps=10, cpu=None, log=print):
    """``elec_rotamer_scores`` on the same rotamers: the whole forward (its
    own dispatch at elec's 5.5 A reach), the scoring kernel alone, the
    dTdV-weighted backward kernel, and the reference's CPU elec rotamer op
    (one thread) on pose 0."""
    import torch

    from .ops import ElecRotamerCall, elec_rotamer_scores

    a = _elec_args(ps, rs, device)
    for _ in range(warmup):
        s, idx = elec_rotamer_scores(*a)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        s, idx = elec_rotamer_scores(*a)
    torch.cuda.synchronize(device)
    op_ms = (time.perf_counter() - t0) * 1000.0 / steps
    n_pairs = int(idx.shape[1])
    finite = bool(torch.isfinite(s).all())
    call = ElecRotamerCall(a, torch.float32)
    call.dispatch()
    out = torch.empty((1, call.n_pairs), dtype=torch.float32, device=device)
    w = torch.ones((1, call.n_pairs), dtype=torch.float32, device=device)
    grad = torch.empty_like(call.coords)
    stream = torch.cuda.current_stream(device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize(device)
        ev0.record(stream)
        for _ in range(kernel_reps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize(device)
        return ev0.elapsed_time(ev1) / kernel_reps

    score_ms = timed(lambda: call.energies(out))
    bwd_ms = timed(lambda: call.backward(w, out=grad))
    res = {
        "workload": "repack_16x300",
        "rotamer_pairs": n_pairs,
        "value": n_pairs / (op_ms * 1e-3),
        "unit": "rotamer pairs/s",
        "ms_per_step": op_ms,
        "score_ms": score_ms,
        "backward_ms": bwd_ms,
        "energies_finite": finite,
        "dtype": "f32",
    }
    if cpu is not None:
        log("elec rotamer cpu baseline ...")
        res["cpu_baseline"] = cpu()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="accepted; the CPU baseline runs from bench.py")
    args = ap.parse_args()
    import torch

    res = run(torch.device("cuda:0"), args.steps, args.warmup, args.kernel_reps,
              cpu=None, log=lambda m: print("[rotamer] " + m, file=sys.stderr, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
