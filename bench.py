"""Benchmark: exact LJ+LK energy + gradient throughput on one MI355X.

A step = one neighbor pass + one fused scoring pass (per-pose, per-term
energies and the weighted gradient, beta2016 weights 1.0 / 0.55 / 1.0) over
the whole batch, inputs resident in HBM, replayed as a captured hipGraph.

Default workload: 512 poses, residue counts log-uniform in [50, 1200],
ragged stack, float32 (seeded synthetic poses, ideal geometry).

Prints one JSON line: value = poses/s for the whole job; roofline of the scoring kernel from
algorithmic FLOPs (branch census x FLOP table in tmol_amd/flops.py) over its
measured duration; cpu_baseline = the reference's own CPU op (oracle/_ref) or,
if absent, the C oracle, on a bounded sample of the same poses.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "poses/s LJ+LK energy+grad; exact vs MLP-pruned at ≤0.05 kcal/mol/res error"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


# the compute unit the scoring kernel's arithmetic runs on (no MFMA in it)
SCORE_KERNEL_UNIT = "valu"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _load_profile(name):
    """A committed rocprofv3 summary under profiles/ (JSON), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def f64_parity(stack, coords, weights, e32, g32, dev):
    """The timed batch once more through the FP64 build of the same kernels
    (same input coordinates, exactly representable in f64): the largest
    relative per-pose, per-term energy error of the f32 run and the largest
    per-pose relative gradient error (norm over the pose's atoms)."""
    import torch

    from tmol_amd.scorer import LJLKRaggedScorer
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm

    sc = LJLKRaggedScorer(stack, LJLKEnergyTerm(device=dev), device=dev, dtype=torch.float64)
    e64, g64 = sc.energy_and_grad(coords.double(), weights.double())
    e64, g64 = e64.double(), g64.double()
    rel_e = ((e32 - e64).abs() / e64.abs().clamp(min=1.0)).max(dim=1).values
    pose = sc.keep["atom_pose"].long()
    P = stack.n_poses
    d2 = torch.zeros(P, dtype=torch.float64, device=dev).index_add_(
        0, pose, ((g32 - g64) ** 2).sum(1))
    n2 = torch.zeros(P, dtype=torch.float64, device=dev).index_add_(0, pose, (g64 ** 2).sum(1))
    rel_g = (d2.sqrt() / n2.sqrt().clamp(min=1e-30)).max()
    del sc
    torch.cuda.empty_cache()
    return {"max_rel_f32_vs_f64": float(rel_e.max()),
            "max_rel_f32_vs_f64_per_term": [float(v) for v in rel_e.tolist()],
            "max_rel_grad_f32_vs_f64_per_pose": float(rel_g),
            "bound": 1e-4, "within_bound": bool(float(rel_e.max()) < 1e-4 and float(rel_g) < 1e-4),
            "reference_run": "the same kernels in float64 on the timed batch's coordinates"}


def padded_dropin(dev, steps=10, warmup=3):
    """The drop-in op at its own configuration: torch.ops.tmol_ljlk.
    ljlk_pose_scores, forward + backward of the beta2016-weighted total, on
    the 64-pose padded stack (residues uniform in [150, 600], float32), each
    call as a caller makes it (block-type tables, workspace and work list set
    up per call).  Reports the synchronised wall time per call, the host time
    until the call returns (table set-up and launches), and the stream time
    between events around the call."""
    import torch

    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    from tmol_amd.generator import workload_poses
    from tmol_amd.pose import PoseStack
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

    poses = workload_poses("padded_64x150_600")
    ps = PoseStack(poses, device=dev)
    args = pose_op_args(ps, LJLKEnergyTerm(device=dev))
    op = torch.ops.tmol_ljlk.ljlk_pose_scores
    w = torch.tensor([1.0, 0.55, 1.0], device=dev)[:, None]
    stream = torch.cuda.current_stream(dev)

    def call():
        c = args[0].detach().requires_grad_(True)
        s, _ = op(c, *args[1:])
        (w * s).sum().backward()
        return c.grad

    for _ in range(warmup):
        call()
    torch.cuda.synchronize(dev)
    wall, host, gpu = [], [], []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record(stream)
        call()
        b.record(stream)
        t1 = time.perf_counter()
        torch.cuda.synchronize(dev)
        t2 = time.perf_counter()
        wall.append(1e3 * (t2 - t0))
        host.append(1e3 * (t1 - t0))
        gpu.append(a.elapsed_time(b))
    import numpy

    med = lambda v: float(numpy.median(v))  # noqa: E731
    P = ps.n_poses
    return {"workload": "padded_64x150_600", "poses": P, "max_blocks": ps.max_n_blocks,
            "op": "torch.ops.tmol_ljlk.ljlk_pose_scores (forward + backward)",
            "ms_per_call_wall": med(wall), "ms_per_call_host": med(host),
            "ms_per_call_stream": med(gpu), "poses_per_s": P / (med(wall) * 1e-3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="ragged_512x50_1200")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=4)
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (waves per SIMD, persistent grid, rows per queue "
                         "grab) and use the library's default launch, e.g. so that a kernel "
                         "trace of this command holds only the measured launches")
    ap.add_argument("--wave-groups", type=int, default=0)
    ap.add_argument("--no-pruner", action="store_true",
                    help="skip the learned-pruner leg (1024 x 300-residue batch)")
    ap.add_argument("--no-rotamer", action="store_true",
                    help="skip the rotamer-pair scoring leg (16 x 300-residue repacking)")
    ap.add_argument("--no-elec", action="store_true",
                    help="skip the elec leg (same 512-pose batch, elec energy + gradient)")
    ap.add_argument("--no-leaf", action="store_true",
                    help="skip the missing-hydrogen leg (same 512-pose batch, build + backward)")
    ap.add_argument("--no-hbond", action="store_true",
                    help="skip the hbond leg (same 512-pose batch, bases + energy + gradient)")
    ap.add_argument("--no-lkball", action="store_true",
                    help="skip the lk_ball leg (same 512-pose batch, waters + energy + gradient)")
    ap.add_argument("--no-padded", action="store_true",
                    help="skip the padded drop-in op leg (64 poses, per-call set-up included)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported. Poses are "
              "independent: a multi-GPU run would shard the pose list by predicted cost with "
              "no exchange beyond gathering the [3, P] energies." % args.gpus, flush=True)
        sys.exit(2)

    import numpy
    import torch

    import bench_cpu
    from tmol_amd.flops import algorithmic_bytes, algorithmic_flops, pair_evaluations
    from tmol_amd.generator import workload_poses
    from tmol_amd.pose import RaggedPoseStack
    from tmol_amd.scorer import LJLKRaggedScorer
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm

    dev = torch.device("cuda:0")
    t0 = time.time()
    poses = workload_poses(args.workload)
    log("generated %d poses in %.1fs" % (len(poses), time.time() - t0))
    stack = RaggedPoseStack(poses)
    term = LJLKEnergyTerm(device=dev)
    scorer = LJLKRaggedScorer(stack, term, device=dev)
    coords = stack.coords.to(dev)
    weights = torch.tensor([1.0, 0.55, 1.0])
    log("stack: %d poses, %d blocks, %d atoms" % (stack.n_poses, stack.n_blocks_total, stack.n_atoms_total))

    tune = None
    if not args.no_autotune:
        # the project's autotuner picks the scoring kernel's waves per SIMD,
        # the persistent grid and the rows per queue grab by timing the
        # energy + gradient step on this batch
        tune = scorer.autotune(coords, weights, cache=False)
        log("autotune: %s" % json.dumps(tune["choice"]))
    else:
        scorer.set_launch(wave_groups=args.wave_groups)  # 0: one workgroup per resident slot
    scorer.neighbors(coords)
    n_items = scorer.check_capacity()
    census = scorer.census(coords)
    log("census: %s" % json.dumps(census))
    flops = algorithmic_flops(census)
    pair_evals = pair_evaluations(census)

    # kernel-level timing of the scoring kernel alone (work list fixed)
    scorer.neighbors(coords)
    w = scorer._weights(weights)
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        scorer.score(coords, 4, weights=w, energies=scorer.energies, grad=scorer.grad)
    torch.cuda.synchronize(dev)
    ev0.record(stream)
    for _ in range(args.kernel_reps):
        scorer.score(coords, 4, weights=w, energies=scorer.energies, grad=scorer.grad)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    score_ms = ev0.elapsed_time(ev1) / args.kernel_reps  # includes 2 tiny fill + 1 finalize kernels
    ev0.record(stream)
    for _ in range(args.kernel_reps):
        scorer.neighbors(coords)
    ev1.record(stream)
    torch.cuda.synchronize(dev)
    nbr_ms = ev0.elapsed_time(ev1) / args.kernel_reps

    # the timed step: captured graph replays
    use_graph = not args.no_graph
    if use_graph:
        scorer.capture(coords, weights)
        step = lambda: scorer.replay()  # noqa: E731
    else:
        step = lambda: scorer.energy_and_grad(coords, weights)  # noqa: E731
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t_start = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t_start
    lat = numpy.array([a.elapsed_time(b) for a, b in evs])
    ms_per_step = wall * 1000.0 / args.steps
    e = scorer.energies.double().cpu()
    finite = bool(torch.isfinite(e).all())

    e32 = scorer.energies.double().clone()
    g32 = scorer.grad.double().clone()
    parity = f64_parity(stack, coords, weights, e32, g32, dev)
    log("parity: %s" % json.dumps(parity))

    achieved_tflops = flops / (score_ms * 1e-3) / 1e12
    abytes = algorithmic_bytes(stack.n_atoms_total, stack.n_blocks_total, stack.n_poses, n_items)
    # the committed rocprofv3 summary of the kernel this run measured
    kname = scorer.kernel_name()
    counters = _load_profile(kname + "_ragged512_counters.json")
    traffic = counters
    hbm_bytes = counters.get("hbm_bytes_per_launch") if counters else None

    result = {
        "metric": METRIC,
        "value": stack.n_poses / (ms_per_step * 1e-3),
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, ideal geometry, 70/30 helix/strand, rotameric chis)",
        "config": {
            "workload": args.workload,
            "poses": stack.n_poses,
            "residues": stack.n_blocks_total,
            "atoms": stack.n_atoms_total,
            "block_pairs": n_items,
            "atom_pair_evals_per_step": pair_evals,
            "atom_pair_evals_per_s": pair_evals / (ms_per_step * 1e-3),
            "graph": use_graph,
            "launch": {"kernel": kname, "waves_per_simd": scorer.waves_per_simd,
                       "wave_groups": scorer.wave_groups,
                       "row_grab": scorer.row_grab},
            "autotune": ({"choice": tune["choice"],
                          "timings_ms": {k: round(v * 1e3, 4) for k, v in tune["timings_s"].items()}}
                         if tune else None),
            "latency_ms_p50": float(numpy.percentile(lat, 50)),
            "latency_ms_p99": float(numpy.percentile(lat, 99)),
            "neighbors_ms": nbr_ms,
            "score_ms": score_ms,
            "energies_finite": finite,
        },
        "roofline": {
            # the scoring kernel's arithmetic is FP32 on the vector ALU (no
            # matrix instructions): it is priced against the FP32 vector peak
            "bound": SCORE_KERNEL_UNIT,
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": hbm_bytes,
            "kernel": kname + "<float, WGRAD> (FP32 vector ALU; peak = MI355X FP32 vector)",
            "algorithmic_flops_per_launch": flops,
            "algorithmic_bytes_per_launch": abytes["score"],
            "hbm_frac": (hbm_bytes / (score_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
                         if hbm_bytes else None),
            "hbm_frac_algorithmic": abytes["score"] / (score_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
            "traffic_profile": ({k: traffic.get(k) for k in (
                "commit", "hbm_read_bytes_per_launch", "hbm_write_bytes_per_launch",
                "copy_calibration")} if traffic else None),
            "counters": counters.get("derived") if counters else None,
            "counters_commit": counters.get("commit") if counters else None,
        },
        "parity": parity,
    }
    if not args.no_pruner:
        from tmol_amd import prune_bench

        log("pruner: train, calibrate, measure ...")
        pr = prune_bench.run(dev, epochs=6, steps=10, warmup=3, log=log)
        result["pruned"] = {k: pr[k] for k in (
            "workload", "poses", "threshold", "exact_poses_per_s", "pruned_poses_per_s", "speedup",
            "pruned_fraction", "max_err_per_res", "mean_err_per_res", "within_bound", "fallback",
            "pays_off", "effective_poses_per_s", "kernel_ms_exact", "kernel_ms_pruned", "prune_ms",
            "perfect_labels")}
        result["pruned"]["train_pairs"] = pr["training"]["pairs"]
        result["pruned"]["train_seconds"] = pr["training"]["train_seconds"]
        result["pruned"]["train_accuracy"] = pr["training"]["train_accuracy"]
    if not args.no_rotamer:
        from tmol_amd import rotamer_bench

        log("rotamer pairs ...")
        result["rotamer"] = rotamer_bench.run(
            dev, steps=10, warmup=2, cpu=None if args.no_cpu else bench_cpu.ljlk_rotamers,
            elec_cpu=None if args.no_cpu else bench_cpu.elec_rotamers, log=log)
        result["rotamer"].pop("census", None)
    if not args.no_elec:
        from tmol_amd import elec_bench

        log("elec ...")
        result["elec"] = elec_bench.run(dev, poses=poses,
                                        cpu=None if args.no_cpu else bench_cpu.elec_poses, log=log)
        result["elec"].pop("census", None)
    if not args.no_lkball:
        from tmol_amd import lk_ball_bench

        log("lk_ball ...")
        result["lk_ball"] = lk_ball_bench.run(
            dev, poses=poses, cpu=None if args.no_cpu else bench_cpu.lk_ball_poses, log=log)
    if not args.no_hbond:
        from tmol_amd import hbond_bench

        log("hbond ...")
        result["hbond"] = hbond_bench.run(
            dev, poses=poses, cpu=None if args.no_cpu else bench_cpu.hbond_poses, log=log)
    if not args.no_leaf:
        from tmol_amd import leaf_bench

        log("leaf atoms ...")
        result["leaf_atoms"] = leaf_bench.run(
            dev, poses=poses, cpu=None if args.no_cpu else bench_cpu.leaf_atoms, log=log)
    if not args.no_padded:
        log("padded drop-in op ...")
        result["padded"] = padded_dropin(dev)
    from tmol_amd import topology_bench

    log("topology ...")
    result["topology"] = topology_bench.run(dev, cpu=None if args.no_cpu else bench_cpu.topology,
                                            log=log)
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = bench_cpu.ljlk_poses(args.workload, poses, args.cpu_stride)
        result["cpu_baseline"]["all_cores"] = bench_cpu.ljlk_poses_all_cores(
            args.workload, poses, args.cpu_stride)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
