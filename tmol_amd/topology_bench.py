"""Topology measurement: inter_block_bondsep + min_block_bondsep of the
64-pose padded workload (residue counts uniform in 150-600, B = 600,
[64, 600, 600, 2, 2] int32 = 92 M entries) on the device, against the
reference's own all-pairs-shortest-paths op (oracle/_ref, one thread) on the
same connection-point graphs (the reference's graph assembly in torch is
not counted).

    python -m tmol_amd.topology_bench
"""

import json
import sys
import time


def run(device, reps=10, cpu=True, log=print):
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
    if cpu:
        res["cpu_baseline"] = cpu_baseline(ps)
    return res


def cpu_baseline(ps):
    """The reference's apsp op (cutoff 6) on each pose's connection-point
    graph, one thread."""
    import os

    import numpy
    import torch

    from oracle import oracle as orc

    if not os.path.exists(os.path.join(orc.HERE, "_ref", "tmol_apsp_ref.so")):
        return None
    torch.set_num_threads(1)
    apsp = orc.load_reference_apsp_op()
    pbt = ps.packed_block_types
    bti = ps.block_type_ind.numpy()
    nconn = pbt.n_conn.numpy()
    conn = pbt.conn_atom.numpy()
    pd = pbt.ljlk_bond_separation.numpy()
    irc = ps.inter_residue_connections.numpy()
    graphs = []
    for p in range(ps.n_poses):
        points = [(b, c) for b in range(ps.max_n_blocks) if bti[p, b] >= 0 for c in range(nconn[bti[p, b]])]
        index = {bc: i for i, bc in enumerate(points)}
        n = len(points)
        W = numpy.full((n, n), 6, dtype=numpy.int32)
        for i, (b, c) in enumerate(points):
            t = bti[p, b]
            for c2 in range(nconn[t]):
                W[i, index[(b, c2)]] = pd[t, conn[t, c], conn[t, c2]]
            b2, cc2 = irc[p, b, c]
            if b2 >= 0:
                W[i, index[(b2, cc2)]] = 1
        graphs.append(torch.from_numpy(W)[None])
    t0 = time.perf_counter()
    for W in graphs:
        apsp(W, 6)
    dt = time.perf_counter() - t0
    return {"value_ms": dt * 1000.0, "cores": 1, "kind": "reference",
            "sample": "all %d poses: the reference apsp op (cutoff 6) on each connection-point "
                      "graph, graph assembly not counted" % ps.n_poses}


def main():
    import torch

    res = run(torch.device("cuda:0"), log=lambda m: print(m, file=sys.stderr))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
