"""The reference's own CPU LJ+LK ops, run in a child interpreter -- TEST
INFRASTRUCTURE.

``tmol_ljlk::ljlk_pose_scores`` and ``tmol_ljlk::ljlk_rotamer_scores``
(tmol/score/ljlk/potentials/compiled.ops.cpp:533-670) are compiled from the
reference sources by build_ref.sh into oracle/_ref/tmol_ljlk_ref.so.
tmol_amd registers its own ``torch.ops.tmol_ljlk`` ops of the same names
(the drop-in), and one process can hold only one library that defines an op
name, so the reference's library is loaded in a separate python process.
Arguments and results travel as ``torch.save`` files written by this module
and read back with ``weights_only=True``.

Only tests/, tests/golden/ scripts and bench.py's cpu_baseline legs use it,
as the checker or the timed CPU baseline.

    python -m oracle.refop IN.pt OUT.pt      (the child side)
"""

import os
import subprocess
import sys
import tempfile
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_LIB_PATH = os.path.join(HERE, "_ref", "tmol_ljlk_ref.so")


def available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def _cpu(args):
    return [a.detach().cpu().contiguous() if isinstance(a, torch.Tensor) else a for a in args]


def _run(job: str, payload: dict, timeout: float = 900.0):
    if not available():
        raise FileNotFoundError(REF_LIB_PATH)
    with tempfile.TemporaryDirectory(prefix="tmol_refop_") as d:
        fin = os.path.join(d, "in.pt")
        fout = os.path.join(d, "out.pt")
        torch.save(dict(payload, job=job), fin)
        env = dict(os.environ)
        env["PYTHONPATH"] = ROOT + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        # the child is a CPU process: keep it off any GPU
        env["HIP_VISIBLE_DEVICES"] = ""
        env["CUDA_VISIBLE_DEVICES"] = ""
        r = subprocess.run([sys.executable, "-m", "oracle.refop", fin, fout], cwd=ROOT, env=env,
                           capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("reference op child failed (%d):\n%s" % (r.returncode, r.stderr[-4000:]))
        return torch.load(fout, weights_only=True)


def pose_scores(args, grad_weights=()):
    """The reference ``ljlk_pose_scores`` on ``args`` (its 27 arguments, CPU
    tensors).  Returns (score, block_neighbors, grads): for each weight w
    (shaped like score) in ``grad_weights``, the gradient of (score * w).sum()
    with respect to rot_coords, through the reference's own backward."""
    out = _run("ljlk_pose", {"args": _cpu(args), "weights": _cpu(list(grad_weights))})
    return out["score"], out["aux"], out["grads"]


def rotamer_scores(args, grad_weights=()):
    """The reference ``ljlk_rotamer_scores``: (score [3, n], dispatch indices
    [3, n], grads) as for ``pose_scores``."""
    out = _run("ljlk_rotamer", {"args": _cpu(args), "weights": _cpu(list(grad_weights))})
    return out["score"], out["aux"], out["grads"]


def time_pose_workload(workload: str, stride: int, weights, budget_s: float, threads: int = 1):
    """Times the reference op, forward + backward of (score * weights).sum(),
    on every ``stride``-th pose of the named generator workload
    (tmol_amd.generator.workload_poses), each pose scored alone, until
    ``budget_s`` seconds of op time are used.  The child builds the poses and
    their op arguments itself (tmol_amd's host code, which registers no op).
    Returns {"n_done", "n_residues", "seconds"}."""
    return _run("time_pose", {"workload": workload, "stride": int(stride),
                              "weights": weights.detach().cpu(), "budget": float(budget_s),
                              "threads": int(threads)})


def time_rotamer_scores(args, threads: int = 1):
    """Times one forward of the reference rotamer op.  Returns
    {"n_pairs", "seconds"}."""
    return _run("time_rotamer", {"args": _cpu(args), "threads": int(threads)})


def schemas():
    """The reference library's registered schemas, {op name: schema string}."""
    return _run("schemas", {})


# ------------------------------------------------------------------ child side


def _op_with_grads(op, args, weights):
    score, aux = op(*args)
    grads = []
    for w in weights:
        c = args[0].clone().requires_grad_(True)
        s, _ = op(c, *args[1:])
        (s * w).sum().backward()
        grads.append(c.grad.detach().clone())
    return {"score": score.detach(), "aux": aux.detach(), "grads": grads}


def _child(fin: str, fout: str):
    req = torch.load(fin, weights_only=True)
    torch.set_num_threads(int(req.get("threads", 1)))
    torch.ops.load_library(REF_LIB_PATH)
    if "tmol_amd.ops" in sys.modules:
        raise RuntimeError("the reference op child must not import tmol_amd.ops")
    ops = torch.ops.tmol_ljlk
    job = req["job"]
    if job == "ljlk_pose":
        out = _op_with_grads(ops.ljlk_pose_scores, req["args"], req["weights"])
    elif job == "ljlk_rotamer":
        out = _op_with_grads(ops.ljlk_rotamer_scores, req["args"], req["weights"])
    elif job == "time_pose":
        from tmol_amd.generator import workload_poses
        from tmol_amd.pose import PoseStack
        from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

        w = req["weights"]
        term = LJLKEnergyTerm()
        t_total, n_done, n_res = 0.0, 0, 0
        for p in workload_poses(req["workload"])[:: req["stride"]]:
            a = pose_op_args(PoseStack([p]), term)
            a[0] = a[0].clone().requires_grad_(True)
            t0 = time.perf_counter()
            s, _ = ops.ljlk_pose_scores(*a)
            (s * w).sum().backward()
            t_total += time.perf_counter() - t0
            n_done += 1
            n_res += p.n_blocks
            if t_total > req["budget"]:
                break
        out = {"n_done": n_done, "n_residues": n_res, "seconds": t_total}
    elif job == "time_rotamer":
        t0 = time.perf_counter()
        _, idx = ops.ljlk_rotamer_scores(*req["args"])
        out = {"n_pairs": int(idx.shape[1]), "seconds": time.perf_counter() - t0}
    elif job == "schemas":
        out = {n: str(getattr(ops, n).default._schema)
               for n in ("ljlk_pose_scores", "ljlk_rotamer_scores")}
    else:
        raise ValueError("unknown job %r" % job)
    torch.save(out, fout)


if __name__ == "__main__":
    _child(sys.argv[1], sys.argv[2])
