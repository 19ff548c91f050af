"""lk_ball benchmark: energy + gradient of the four lk_ball score types on
the 512-pose ragged batch (the same poses as the LJ+LK headline), on the
same scheduler.

One step = lk_ball's neighbor pass + water generation + the row kernel in
weighted-gradient mode (energies [4, P], gradients for coordinates and
waters) + the waters' backward, inputs resident in HBM.  Reported: poses/s
of the step, each kernel's time (HIP events on its stream), the scoring
kernel's roofline from the pair census x the lk_ball FLOP table
(tmol_amd/flops.py), and, when ``cpu`` is given (bench.py passes the
reference's own CPU ops, one thread, forward + backward), a CPU baseline on
a bounded sample of the poses.

    python -m tmol_amd.lk_ball_bench [--steps 20]      (GPU legs only)
"""

import argparse
import json
import os
import sys
import time

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                       "lkb_gather_kernel_ragged512_counters.json")


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

    from .flops import lkb_algorithmic_flops
    from .generator import workload_poses
    from .lk_ball import LKBallRaggedScorer
    from .ops import WGRAD
    from .pose import RaggedPoseStack
    from .score.lk_ball_energy_term import LKBallEnergyTerm

    if poses is None:
        poses = workload_poses("ragged_512x50_1200")
    stack = RaggedPoseStack(poses)
    term = LKBallEnergyTerm()
    sc = LKBallRaggedScorer(stack, term, device=device)
    X = stack.coords.to(device)
    w = sc._weights(None)
    sc.neighbors(X)
    n_items = sc.check_capacity()
    sc.gen_waters(X)
    census = sc.census(X)
    flops = lkb_algorithmic_flops(census)
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
    wat_ms = kernel_ms(lambda: sc.gen_waters(X))
    # + 3 fill kernels and 1 finalize around the row kernel
    score_ms = kernel_ms(lambda: sc.score(X, WGRAD, weights=w, energies=sc.energies,
                                          dxyz=sc.dxyz, dwat=sc.dwat))
    back_ms = kernel_ms(lambda: sc.waters_backward(X, sc.dwat, sc.dxyz))
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
    # compulsory bytes of the scoring kernel: coordinates + waters read, the
    # two gradient planes written, the work list read (float32)
    na = stack.n_atoms_total
    alg_bytes = na * (3 + 12) * 4 + na * (3 + 12) * 4 + 4 * n_items
    prof = _counters()
    traffic = prof.get("hbm_bytes_per_launch") if prof else None
    return {
        "workload": "ragged_512x50_1200",
        "poses": stack.n_poses,
        "value": stack.n_poses / (ms * 1e-3),
        "unit": "poses/s",
        "ms_per_step": ms,
        "neighbors_ms": nbr_ms,
        "waters_ms": wat_ms,
        "score_ms": score_ms,
        "waters_backward_ms": back_ms,
        "energy_only_score_ms": energy_ms,
        "block_pairs": n_items,
        "census": census,
        "energies": [float(x) for x in e.sum(1)],
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
            "kernel": "lkb_gather_kernel<float, WGRAD> (FP32 vector ALU)",
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
