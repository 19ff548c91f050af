"""Topology measurement: inter_block_bondsep + min_block_bondsep of the
64-pose padded workload (residue counts uniform in 150-600, B = 600,
[64, 600, 600, 2, 2] int32 = 92 M entries) on the device; bench.py passes
the CPU baseline (the reference's own all-pairs-shortest-paths op, one
thread, on the same connection-point graphs; the reference's graph assembly
in torch is not counted).

    python -m tmol_amd.topology_bench
"""

import json
import sys
import time


def run(device, reps=10, cpu=None, log=print):
    import numpy
    import torch

    from .generator import workload_poses
    from .pose import PoseStack
    from .topology import inter_block_bondsep

    ps = PoseStack(workload_poses("padded_64x150_600"))
    args = (ps.packed_block_types, ps.block_type_ind.to(device), ps.inter_residue_connections.to(device))
    for _ in range(2):
        inter_block_bondsep(*args, device=device, check_overflow=False)
    torch.cuda.synchronize(device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        inter_block_bondsep(*args, device=device, check_overflow=False)
    e1.record()
    torch.cuda.synchronize(device)
    ms = e0.elapsed_time(e1) / reps
    P, B = ps.n_poses, ps.max_n_blocks
    C = ps.packed_block_types.max_n_conn
    res = {"workload": "padded_64x150_600", "poses": P, "max_blocks": B,
           "entries": P * B * B * C * C, "ms": ms,
           "bytes_written": 4 * (P * B * B * C * C + P * B * B),
           "write_gb_per_s": 4 * (P * B * B * C * C + P * B * B) / (ms * 1e-3) / 1e9}
    if cpu is not None:
        res["cpu_baseline"] = cpu(ps)
    return res


def main():
    import torch

    res = run(torch.device("cuda:0"), log=lambda m: print(m, file=sys.stderr))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
