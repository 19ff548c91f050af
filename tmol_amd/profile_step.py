"""Profiling driver: build a workload once, then launch the neighbor pass and
the scoring kernel a fixed number of times (no autotuning, no CPU work), so
that ``rocprofv3`` sees a short, repeatable kernel sequence.

    python -m tmol_amd.profile_step --workload ragged_512x50_1200 --reps 5 \
        [--mode wgrad|dvdx|energy] [--general] [--wave-groups N]
        [--term ljlk|elec|lk_ball|hbond]

``--term elec`` runs the elec term's neighbor pass and row kernel instead
(weights 1 per pose in wgrad mode); ``--term lk_ball`` lk_ball's neighbor
pass, waters and gather kernel; ``--term hbond`` hbond's neighbor pass,
bases and row kernel (both ``--mode wgrad|energy``).

``--copy-bytes N`` instead copies an N-byte device buffer ``reps`` times with
torch's copy kernel: the reference point that scales the HBM counters
(FETCH_SIZE / WRITE_SIZE) of a known byte count.

``--summarize OUT.json --kernel NAME --pmc-dir DIR [DIR ...] [--copy-dir DIR
...] [--commit C] [--copy-bytes N]`` reads rocprofv3 ``--pmc`` CSV output
(pmc_counter_collection.csv under each DIR), averages each counter over the
dispatches of the kernel whose name contains NAME, and writes the counters
with derived fractions: VALU busy = SQ_INSTS_VALU * 4 / (SIMDs * cycles)
(a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles), LDS
conflict share, HBM bytes scaled by the copy kernel's counts.
"""

import argparse
import json
import sys
import time

import torch


MI355X_SIMDS = 1024   # 256 CUs x 4
MI355X_XCDS = 8       # GRBM_GUI_ACTIVE is summed over the XCDs


def _pmc_means(dirs, kernel):
    import csv
    import os
    from collections import defaultdict

    tot, disp = defaultdict(float), defaultdict(set)
    for d in dirs:
        with open(os.path.join(d, "pmc_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                if kernel in r["Kernel_Name"]:
                    tot[r["Counter_Name"]] += float(r["Counter_Value"])
                    disp[r["Counter_Name"]].add((d, r["Dispatch_Id"]))
    return {k: tot[k] / len(disp[k]) for k in tot}


def summarize(out, kernel, pmc_dirs, copy_dirs=(), commit=None, copy_bytes=1 << 30,
              command=None):
    c = _pmc_means(pmc_dirs, kernel)
    d = {}
    if "GRBM_GUI_ACTIVE" in c:
        cycles = c["GRBM_GUI_ACTIVE"] / MI355X_XCDS
        d["kernel_cycles"] = cycles
        if "SQ_INSTS_VALU" in c:
            d["valu_busy"] = c["SQ_INSTS_VALU"] * 4 / (MI355X_SIMDS * cycles)
        if "SQ_INSTS_LDS" in c:
            d["lds_insts_per_simd_cycle"] = c["SQ_INSTS_LDS"] / (MI355X_SIMDS * cycles)
    if "SQ_WAVE_CYCLES" in c:
        for k in ("SQ_WAIT_INST_LDS", "SQ_WAIT_ANY"):
            if k in c:
                d[k.lower().replace("sq_", "") + "_frac_of_wave_cycles"] = c[k] / c["SQ_WAVE_CYCLES"]
    if "SQ_BUSY_CU_CYCLES" in c and "SQ_LDS_IDX_ACTIVE" in c:
        d["lds_idx_active_frac_of_busy_cu"] = c["SQ_LDS_IDX_ACTIVE"] / c["SQ_BUSY_CU_CYCLES"]
    if "SQ_ACTIVE_INST_LDS" in c:
        for k in ("SQ_LDS_BANK_CONFLICT", "SQ_LDS_ADDR_CONFLICT"):
            if k in c:
                d[k.lower().replace("sq_", "") + "_per_lds_active"] = c[k] / c["SQ_ACTIVE_INST_LDS"]
    if "SQ_INSTS_VALU" in c and "SQ_INSTS_SALU" in c:
        d["salu_per_valu"] = c["SQ_INSTS_SALU"] / c["SQ_INSTS_VALU"]
    res = {"kernel": kernel, "commit": commit, "command": command, "per_dispatch": c, "derived": d}
    if copy_dirs:
        cc = _pmc_means(copy_dirs, "copyBuffer")
        f = {k: cc[k] * 1024 / copy_bytes for k in ("FETCH_SIZE", "WRITE_SIZE") if k in cc}
        res["copy_calibration"] = {"bytes_copied": copy_bytes, "fetch_factor": f.get("FETCH_SIZE"),
                                   "write_factor": f.get("WRITE_SIZE")}
        rd = c["FETCH_SIZE"] * 1024 / f["FETCH_SIZE"] if "FETCH_SIZE" in c and f.get("FETCH_SIZE") else None
        wr = c["WRITE_SIZE"] * 1024 / f["WRITE_SIZE"] if "WRITE_SIZE" in c and f.get("WRITE_SIZE") else None
        res["hbm_read_bytes_per_launch"] = rd
        res["hbm_write_bytes_per_launch"] = wr
        res["hbm_bytes_per_launch"] = (rd or 0) + (wr or 0) if (rd or wr) else None
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--summarize", default=None)
    ap.add_argument("--kernel", default="score_rows_kernel<float, 4")
    ap.add_argument("--pmc-dir", nargs="*", default=[])
    ap.add_argument("--copy-dir", nargs="*", default=[])
    ap.add_argument("--commit", default=None)
    ap.add_argument("--command", default=None)
    ap.add_argument("--workload", default="ragged_512x50_1200")
    ap.add_argument("--n-poses", type=int, default=None)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--mode", default="wgrad", choices=["wgrad", "dvdx", "energy"])
    ap.add_argument("--general", action="store_true")
    ap.add_argument("--wave-groups", type=int, default=0)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--term", default="ljlk", choices=["ljlk", "elec", "lk_ball", "hbond"])
    args = ap.parse_args(argv)

    if args.summarize:
        summarize(args.summarize, args.kernel, args.pmc_dir, args.copy_dir, args.commit,
                  args.copy_bytes or (1 << 30), args.command)
        return 0

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
    if args.term == "lk_ball":
        from .lk_ball import LKBallRaggedScorer
        from .score.lk_ball_energy_term import LKBallEnergyTerm

        sc = LKBallRaggedScorer(stack, LKBallEnergyTerm(), device=dev, dtype=dtype)
        x = stack.coords.to(dev)
        w = sc._weights(None)
        flags = {"wgrad": ops.WGRAD, "energy": 0}[args.mode]
        sc.neighbors(x)
        n_items = sc.check_capacity()
        sc.gen_waters(x)
        torch.cuda.synchronize()
        times = []
        for _ in range(args.reps):
            t0 = time.perf_counter()
            sc.neighbors(x)
            sc.gen_waters(x)
            sc.score(x, flags, weights=w, energies=sc.energies, dxyz=sc.dxyz, dwat=sc.dwat)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        print(json.dumps({"workload": args.workload, "term": args.term, "mode": args.mode,
                          "block_pairs": n_items, "step_ms": [1e3 * t for t in times]}),
              flush=True)
        return 0
    if args.term == "hbond":
        from .hbond import HBondRaggedScorer
        from .score.hbond_energy_term import HBondEnergyTerm

        sc = HBondRaggedScorer(stack, HBondEnergyTerm(), device=dev, dtype=dtype)
        x = stack.coords.to(dev)
        w = sc._weights(None)
        flags = {"wgrad": ops.WGRAD, "energy": 0}[args.mode]
        sc.neighbors(x)
        n_items = sc.check_capacity()
        sc.bases(x)
        torch.cuda.synchronize()
        times = []
        for _ in range(args.reps):
            t0 = time.perf_counter()
            sc.neighbors(x)
            sc.bases(x)
            sc.score(x, flags, weights=w, energies=sc.energies, dxyz=sc.dxyz)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        print(json.dumps({"workload": args.workload, "term": args.term, "mode": args.mode,
                          "block_pairs": n_items, "step_ms": [1e3 * t for t in times]}),
              flush=True)
        return 0
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
