"""Profiling driver: build a workload once, then launch the neighbor pass and
the scoring kernel a fixed number of times (no autotuning, no CPU work), so
that ``rocprofv3`` sees a short, repeatable kernel sequence.

    python -m tmol_amd.profile_step --workload ragged_512x50_1200 --reps 5 \
        [--mode wgrad|dvdx|energy] [--general] [--wave-groups N] [--term ljlk|elec]

``--term elec`` runs the elec term's neighbor pass and row kernel instead
(weights 1 per pose in wgrad mode).

``--copy-bytes N`` instead copies an N-byte device buffer ``reps`` times with
torch's copy kernel: the reference point that scales the HBM counters
(FETCH_SIZE / WRITE_SIZE) of a known byte count.
"""

import argparse
import json
import sys
import time

import torch


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ragged_512x50_1200")
    ap.add_argument("--n-poses", type=int, default=None)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--mode", default="wgrad", choices=["wgrad", "dvdx", "energy"])
    ap.add_argument("--general", action="store_true")
    ap.add_argument("--wave-groups", type=int, default=0)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--term", default="ljlk", choices=["ljlk", "elec"])
    args = ap.parse_args(argv)

    if args.copy_bytes > 0:
        dev = torch.device("cuda:0")
        src = torch.ones(args.copy_bytes, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        times = []
        for _ in range(args.reps):
            t0 = time.perf_counter()
            dst.copy_(src)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        print(json.dumps({"copy_bytes": args.copy_bytes, "step_ms": [1e3 * t for t in times]}),
              flush=True)
        return 0

    from . import ops
    from .generator import workload_poses
    from .pose import RaggedPoseStack
    from .score.ljlk_energy_term import LJLKEnergyTerm
    from .scorer import LJLKRaggedScorer

    dev = torch.device("cuda:0")
    dtype = torch.float32 if args.dtype == "f32" else torch.float64
    poses = workload_poses(args.workload, args.n_poses)
    stack = RaggedPoseStack(poses, dtype=dtype)
    if args.term == "elec":
        from .score.elec_energy_term import ElecEnergyTerm
        from .scorer import ElecRaggedScorer

        sc = ElecRaggedScorer(stack, ElecEnergyTerm(), device=dev, dtype=dtype,
                              wave_groups=args.wave_groups)
        w = sc._weights(1.0)
        n_planes = 1
    else:
        sc = LJLKRaggedScorer(stack, LJLKEnergyTerm(device=dev), device=dev, dtype=dtype,
                              wave_groups=args.wave_groups)
        w = sc._weights(torch.tensor([1.0, 0.55, 1.0]))
        n_planes = 3
    x = stack.coords.to(dev)
    flags = {"wgrad": ops.WGRAD, "dvdx": ops.DVDX, "energy": 0}[args.mode]
    if args.general:
        flags |= ops.GENERAL_KERNEL
    dvdx = (torch.empty((n_planes, sc.NA, 3), dtype=dtype, device=dev)
            if args.mode == "dvdx" else None)
    sc.neighbors(x)
    n_items = sc.check_capacity()
    torch.cuda.synchronize()
    times = []
    for _ in range(args.reps):
        t0 = time.perf_counter()
        sc.neighbors(x)
        sc.score(x, flags, weights=w, energies=sc.energies, dvdx=dvdx, grad=sc.grad)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    print(json.dumps({"workload": args.workload, "term": args.term, "mode": args.mode,
                      "general": args.general,
                      "block_pairs": n_items, "step_ms": [1e3 * t for t in times]}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
