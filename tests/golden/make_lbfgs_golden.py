"""Write tests/golden/lbfgs_rosenbrock.json: input/output vectors of the
reference minimizer (tmol/optimization/lbfgs_armijo.py, LBFGS_Armijo) on a
segmented Rosenbrock problem.

Run in the development container only (it reads the reference checkout);
the committed JSON is the fixture.  The reference module imports only torch,
so it is loaded by file path without importing the tmol package.

The fixture records every energy evaluation the minimizer makes (the
per-segment loss vector of each closure call, in order), the iteration count
and the final parameters; a minimizer with the same semantics must make the
same sequence of evaluations.
"""

import importlib.util
import json
import os
import sys

import torch

REF = os.environ.get("TMOL_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

# segments of different dimension and start; Rosenbrock in each segment
DIMS = [2, 3, 5, 8, 4, 6]
SEED = 20261202


def problem():
    g = torch.Generator().manual_seed(SEED)
    seg = torch.cat([torch.full((n,), i, dtype=torch.int64) for i, n in enumerate(DIMS)])
    x0 = (torch.rand(seg.numel(), generator=g, dtype=torch.float64) * 2.4 - 1.2)
    scale = torch.tensor([1.0, 10.0, 100.0, 100.0, 30.0, 100.0], dtype=torch.float64)
    return seg, x0, scale


def rosenbrock(x, seg, scale):
    """E_p = sum_i (1 - x_i)^2 + c_p (x_{i+1} - x_i^2)^2 within segment p."""
    nseg = int(seg.max()) + 1
    same = seg[1:] == seg[:-1]
    t1 = (1 - x) ** 2
    t2 = scale[seg[:-1]] * (x[1:] - x[:-1] ** 2) ** 2 * same
    E = torch.zeros(nseg, dtype=x.dtype).index_add(0, seg, t1)
    E = E.index_add(0, seg[:-1], t2)
    return E


def main():
    path = os.path.join(REF, "tmol", "optimization", "lbfgs_armijo.py")
    spec = importlib.util.spec_from_file_location("ref_lbfgs_armijo", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)

    seg, x0, scale = problem()
    x = x0.clone().requires_grad_(True)
    evals = []

    def closure():
        if x.grad is not None:
            x.grad.zero_()
        E = rosenbrock(x, seg, scale)
        E.sum().backward()
        evals.append(E.detach().tolist())
        return E

    opt = mod.LBFGS_Armijo([x], max_iter=100, gradtol=1e-6, segment_ids=seg)
    opt.step(closure)
    out = {
        "source": "tmol/optimization/lbfgs_armijo.py LBFGS_Armijo(max_iter=100, gradtol=1e-6, "
                  "history 128, lr 1, float64), segmented Rosenbrock",
        "dims": DIMS,
        "seed": SEED,
        "scale": scale.tolist(),
        "x0": x0.tolist(),
        "evaluations": evals,
        "n_iter": int(opt.state[x]["n_iter"]),
        "x_final": x.detach().tolist(),
    }
    with open(os.path.join(HERE, "lbfgs_rosenbrock.json"), "w") as f:
        json.dump(out, f)
    print("evaluations", len(evals), "iterations", out["n_iter"], "final E",
          rosenbrock(x.detach(), seg, scale).tolist())


if __name__ == "__main__":
    sys.exit(main())
