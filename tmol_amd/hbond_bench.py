"""hbond benchmark: energy + gradient of the hbond term on the 512-pose
ragged batch (the same poses as the LJ+LK headline), on the same scheduler.

One step = hbond's neighbor pass (5.5 A) + the derived base coordinates +
the row kernel in weighted-gradient mode (energies [P], gradient [NA, 3]),
inputs resident in HBM.  Reported: poses/s of the step, each kernel's time
(HIP events on its stream), the scoring kernel's roofline from the pair
census x the hbond FLOP table (tmol_amd/flops.py), and, when ``cpu`` is
given (bench.py passes the reference's own CPU ops, one thread, forward +
backward), a CPU baseline on a bounded sample of the poses.

    python -m tmol_amd.hbond_bench [--steps 20]      (GPU legs only)
"""

import argparse
import json
import os
import sys
import time

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                       "hb_score_kernel_ragged512_counters.json")


def _counters():
    """The committed rocprofv3 counter summary of the scoring kernel on this
    batch (HBM bytes per launch from FETCH_SIZE / WRITE_SIZE), or None."""
    try:
        with open(PROFILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def run(device, poses=None, steps=20, warmup=3, kernel_reps=10, cpu=None, log=print):
    import torch

    from .flops import hb_algorithmic_flops
    from .generator import workload_poses
    from .hbond import HBondRaggedScorer
    from .ops import WGRAD
    from .pose import RaggedPoseStack
    from .score.hbond_energy_term import HBondEnergyTerm

    if poses is None:
        poses = workload_poses("ragged_512x50_1200")
    stack = RaggedPoseStack(poses)
    term = HBondEnergyTerm()
    sc = HBondRaggedScorer(stack, term, device=device)
    X = stack.coords.to(device)
    w = sc._weights(None)
    sc.neighbors(X)
    n_items = sc.check_capacity()
    sc.bases(X)
    census = sc.census(X)
    flops = hb_algorithmic_flops(census)
    stream = torch.cuda.current_stream(device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def kernel_ms(fn):
        fn()
        torch.cuda.synchronize(device)
        ev[0].record(stream)
        for _ in range(kernel_reps):
            fn()
        ev[1].record(stream)
        torch.cuda.synchronize(device)
        return ev[0].elapsed_time(ev[1]) / kernel_reps

    nbr_ms = kernel_ms(lambda: sc.neighbors(X))
    bases_ms = kernel_ms(lambda: sc.bases(X))
    score_ms = kernel_ms(lambda: sc.score(X, WGRAD, weights=w, energies=sc.energies,
                                          dxyz=sc.dxyz))
    energy_ms = kernel_ms(lambda: sc.score(X, 0, energies=sc.energies))

    for _ in range(warmup):
        sc.energy_and_grad(X, w)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        sc.energy_and_grad(X, w)
    torch.cuda.synchronize(device)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    e = sc.energies.double().cpu()
    g = sc.dxyz
    achieved = flops / (score_ms * 1e-3) / 1e12
    # compulsory bytes of the scoring kernel: the work list and block types
    # read, coordinates and derived coordinates of the polar atoms read, the
    # gradient of the hbond atoms written (bounded above by all atoms)
    na = stack.n_atoms_total
    alg_bytes = 4 * n_items + na * (3 + 9) * 4 + na * 3 * 4
    prof = _counters()
    traffic = prof.get("hbm_bytes_per_launch") if prof else None
    return {
        "workload": "ragged_512x50_1200",
        "poses": stack.n_poses,
        "value": stack.n_poses / (ms * 1e-3),
        "unit": "poses/s",
        "ms_per_step": ms,
        "neighbors_ms": nbr_ms,
        "bases_ms": bases_ms,
        "score_ms": score_ms,
        "energy_only_score_ms": energy_ms,
        "block_pairs": n_items,
        "census": census,
        "energy_sum": float(e.sum()),
        "energies_finite": bool(torch.isfinite(e).all()) and bool(torch.isfinite(g).all()),
        "dtype": "f32",
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_profile": ({k: prof.get(k) for k in (
                "commit", "hbm_read_bytes_per_launch", "hbm_write_bytes_per_launch",
                "copy_calibration")} if prof else None),
            "counters": prof.get("derived") if prof else None,
            "kernel": "hb_score_kernel<float, WGRAD> (vector ALU; polynomials in FP64)",
            "algorithmic_flops_per_launch": flops,
            "algorithmic_bytes_per_launch": alg_bytes,
            "hbm_frac_algorithmic": alg_bytes / (score_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu(poses) if cpu is not None else None,
    }


def main(argv=None):
    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    a = ap.parse_args(argv)
    res = run(torch.device("cuda", 0), steps=a.steps, warmup=a.warmup,
              log=lambda m: print(m, file=sys.stderr))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
