"""Missing-leaf-atom benchmark: every leaf hydrogen of the 512-pose batch
(the same poses as the LJ+LK headline) rebuilt from the heavy atoms.

One step = tmol_leaf_atoms_build + tmol_leaf_atoms_backward (coordinates
out, then the gradient back onto the given atoms), inputs resident in HBM.
Reported: poses/s of the step, each kernel's time (HIP events on its
stream), the drop-in op's time per call with its host-side table checks,
the forward kernel's HBM roofline (algorithmic bytes: 1 B missing flag per
atom slot, 12 B written per real atom, 12 B read per given atom, 36 B
of ancestors and 3 B of their flags per built atom), and, when ``cpu`` is
given, the reference's CPU op on the same batch.

    python -m tmol_amd.leaf_bench [--steps 20]
"""

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles",
                       "leaf_atoms_ragged512_counters.json")


def _counters():
    """The committed rocprofv3 counter summary of the leaf kernels on this
    batch (HBM bytes per launch from FETCH_SIZE / WRITE_SIZE), or None."""
    try:
        with open(PROFILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def strip_hydrogens(poses, db=None):
    """The leaf-op arguments of a padded stack of ``poses`` with every leaf
    hydrogen flagged missing, and the stack."""
    from .chemical import default_chemical_database
    from .ingest import LeafTables, PaddedLeafLayout, leaf_op_args

    db = db or default_chemical_database()
    ps = PaddedLeafLayout(poses)
    is_h = db.atom_type_set.is_hydrogen
    tab = LeafTables(ps.packed_block_types, lambda rt: is_h[rt.atom_types].astype(bool))
    lays = []
    for p in poses:
        m = []
        for r in p.residue_types:
            lf = tab.is_leaf[ps.packed_block_types.index[r.name]][: r.n_atoms]
            m.append(lf & is_h[r.atom_types].astype(bool))
        lays.append(dict(missing=m))
    args, _ = leaf_op_args(ps, db, lays, tab)
    return ps, args


def run(device, poses=None, steps=20, warmup=3, kernel_reps=10, cpu=None, log=print):
    import torch

    from .generator import workload_poses
    from .ingest import _LeafCall, gen_pose_leaf_atoms

    if poses is None:
        poses = workload_poses("ragged_512x50_1200")
    ps, args = strip_hydrogens(poses)
    dargs = [a.to(device) if torch.is_tensor(a) else a for a in args]
    call = _LeafCall(dargs)
    n_built = int(args[2].sum())
    n_real = sum(p.n_atoms for p in poses)
    n_slots = int(args[1].numel())
    dnew = torch.randn(dargs[0].shape, device=device, generator=torch.Generator(
        device=device).manual_seed(1))
    stream = torch.cuda.current_stream(device)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def kernel_ms(fn, reps=kernel_reps):
        fn()
        torch.cuda.synchronize(device)
        ev[0].record(stream)
        for _ in range(reps):
            fn()
        ev[1].record(stream)
        torch.cuda.synchronize(device)
        return ev[0].elapsed_time(ev[1]) / reps

    fwd_ms = kernel_ms(call.build)
    bwd_ms = kernel_ms(lambda: call.backward(dnew))
    op_ms = kernel_ms(lambda: gen_pose_leaf_atoms(*dargs), reps=3)
    for _ in range(warmup):
        call.build()
        call.backward(dnew)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        out = call.build()
        g = call.backward(dnew)
    torch.cuda.synchronize(device)
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    alg_bytes = n_slots * 1 + n_real * 12 + (n_real - n_built) * 12 + n_built * (36 + 3)
    achieved = alg_bytes / (fwd_ms * 1e-3) / 1e9
    prof = _counters()
    kern = (prof or {}).get("kernels", {})
    fwd_prof = kern.get("leaf_atoms_kernel<float>", {})
    return {
        "workload": "ragged_512x50_1200 hydrogens rebuilt",
        "poses": ps.n_poses,
        "atoms": n_real,
        "built_atoms": n_built,
        "value": ps.n_poses / (ms * 1e-3),
        "unit": "poses/s",
        "ms_per_step": ms,
        "forward_ms": fwd_ms,
        "backward_ms": bwd_ms,
        "dropin_op_ms": op_ms,
        "finite": bool(torch.isfinite(out).all()) and bool(torch.isfinite(g).all()),
        "dtype": "f32",
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": fwd_prof.get("hbm_bytes_per_launch"),
            "kernel": "leaf_atoms_kernel<float> (one-wave workgroups, 16 blocks each)",
            "algorithmic_bytes_per_launch": alg_bytes,
            "backward_traffic": {k: v.get("hbm_bytes_per_launch") for k, v in kern.items()
                                 if "backward" in k} or None,
            "traffic_profile": ({k: prof.get(k) for k in ("commit", "copy_calibration")}
                                if prof else None),
        },
        "cpu_baseline": cpu(args) if cpu is not None else None,
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
