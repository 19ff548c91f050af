"""Write tests/golden/min_diff_32x400_oracle.json: the oracle-driven
minimization of the 32 x 400-residue workload (seeded generator), 200
iterations of the segmented L-BFGS in float64 with energies and gradients from
the C oracle (the restatement of the reference CPU op).

The fixture holds, per iteration, every pose's weighted LJ+LK energy
(1.0 atr + 0.55 rep + 1.0 lk) and gradient norm, the evaluations made, and a
checksum of the starting coordinates so that a GPU run can confirm it starts
from the same poses.  Runs on the host CPU (threads over poses).
"""

import json
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import pose_oracle_evaluator  # noqa: E402
from tmol_amd.generator import workload_poses  # noqa: E402
from tmol_amd.minimize import SegmentedLBFGS  # noqa: E402
from tmol_amd.pose import RaggedPoseStack  # noqa: E402

WORKLOAD = "diff_32x400"
ITERS = 200


def main():
    threads = int(os.environ.get("THREADS", os.cpu_count() or 8))
    poses = workload_poses(WORKLOAD)
    st = RaggedPoseStack(poses, dtype=torch.float64)
    x0 = st.coords.reshape(-1).double()
    seg = st.atom_pose.repeat_interleave(3).long()
    ev = pose_oracle_evaluator(poses, threads=threads)
    t0 = time.time()
    x, tr = SegmentedLBFGS(seg, max_iter=ITERS).run(x0, ev)
    out = {
        "workload": WORKLOAD,
        "dtype": "float64",
        "weights": [1.0, 0.55, 1.0],
        "iterations": tr.iterations,
        "x0_checksum": [float(x0.sum()), float((x0 * x0).sum())],
        "n_atoms": [p.n_atoms for p in poses],
        "energy": [[float("%.12g" % v) for v in e.tolist()] for e in tr.energy],
        "grad_norm": [[float("%.8g" % v) for v in g.tolist()] for g in tr.grad_norm],
        "evals": tr.evals,
        "converged": tr.converged.tolist(),
        "stalled": tr.stalled.tolist(),
        "seconds": time.time() - t0,
        "threads": threads,
    }
    with open(os.path.join(HERE, "min_diff_32x400_oracle.json"), "w") as f:
        json.dump(out, f)
    print("iterations", tr.iterations, "evals", sum(tr.evals), "E0", sum(out["energy"][0]),
          "E_end", sum(out["energy"][-1]), "%.1f s" % out["seconds"])


if __name__ == "__main__":
    sys.exit(main())
