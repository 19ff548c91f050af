"""CPU tests of the topology oracle (inter_block_bondsep and min_bondsep from
the connection graph; the reference's pose_stack_builder.py:459-599,
1230-1349 and bond_dependent_term.py:91-100).

Pinned by: the chain-position separations of the pose stacks used in every
scoring test (linear chains); brute-force breadth-first search over the full
atom bond graph of cyclic and multi-chain poses; and the reference's own
all-pairs-shortest-paths op (oracle/_ref/tmol_apsp_ref.so) run on the
connection-point graph.
"""

import numpy
import pytest
import torch

from oracle import oracle as orc
from tmol_amd.pose import Pose, PoseStack
from tmol_amd.topology import add_connections


def cyclic_and_multichain_stack():
    """Two poses without termini patches: a 14-residue head-to-tail cycle,
    and two 7-residue chains in one pose joined head of the first to tail of
    the second."""
    from tmol_amd.chemical import default_chemical_database
    from tmol_amd.generator import build_chains_fast, workload_specs

    db = default_chemical_database()
    specs = workload_specs("padded_64x150_600", 2)
    chains = []
    for k, n in ((0, 14), (1, 14)):
        rts, phis, psis, omegas, chis = specs[k]
        rr = [db.residue_type(rt.base_name) for rt in rts[:n]]
        chains.append((rr, phis[:n], psis[:n], omegas[:n], chis[:n]))
    coords = build_chains_fast(chains)
    p0 = Pose(chains[0][0], coords[0])
    p1 = Pose(chains[1][0], coords[1], connected_to_prev=[False] + [True] * 6 + [False] + [True] * 6)
    ps = PoseStack([p0, p1])
    r0, r1 = p0.residue_types, p1.residue_types
    bonds = [
        (0, 0, r0[0].down_connection_ind, 13, r0[13].up_connection_ind),
        (1, 0, r1[0].down_connection_ind, 13, r1[13].up_connection_ind),
    ]
    irc = add_connections(ps.inter_residue_connections, bonds)
    return ps, irc


def atom_graph_bondsep(ps, irc):
    """Breadth-first search over the atom bond graph: bonds inside a block
    are its path-distance-1 pairs; a connection bonds the two connection
    atoms.  Returns inter_block_bondsep [P, B, B, C, C] capped at 6."""
    pbt = ps.packed_block_types
    P, B = ps.n_poses, ps.max_n_blocks
    C = pbt.max_n_conn
    pd = pbt.ljlk_bond_separation.numpy()
    conn = pbt.conn_atom.numpy()
    nconn = pbt.n_conn.numpy()
    bti = ps.block_type_ind.numpy()
    off = ps.block_coord_offset.numpy()
    out = numpy.full((P, B, B, C, C), 6, dtype=numpy.int32)
    for p in range(P):
        adj = {}

        def bond(a, b):
            adj.setdefault(a, set()).add(b)
            adj.setdefault(b, set()).add(a)

        for b in range(B):
            t = bti[p, b]
            if t < 0:
                continue
            n = int(pbt.n_atoms[t])
            for i in range(n):
                for j in range(i + 1, n):
                    if pd[t, i, j] == 1:
                        bond(off[p, b] + i, off[p, b] + j)
            for c in range(nconn[t]):
                b2, c2 = irc[p, b, c].tolist()
                if b2 >= 0:
                    bond(off[p, b] + conn[t, c], off[p, b2] + conn[bti[p, b2], c2])
        for b1 in range(B):
            t1 = bti[p, b1]
            if t1 < 0:
                continue
            for c1 in range(nconn[t1]):
                src = off[p, b1] + conn[t1, c1]
                dist = {src: 0}
                frontier = [src]
                for d in range(1, 6):
                    nxt = []
                    for a in frontier:
                        for x in adj.get(a, ()):
                            if x not in dist:
                                dist[x] = d
                                nxt.append(x)
                    frontier = nxt
                for b2 in range(B):
                    t2 = bti[p, b2]
                    if t2 < 0:
                        continue
                    for c2 in range(nconn[t2]):
                        out[p, b1, b2, c1, c2] = dist.get(off[p, b2] + conn[t2, c2], 6)
    return out


def test_linear_chains_match_chain_positions():
    from tmol_amd.generator import workload_poses

    ps = PoseStack(workload_poses("cpu_single_100") + workload_poses("padded_64x150_600", 2))
    ibs, mins = orc.oracle_inter_block_bondsep(ps.packed_block_types, ps.block_type_ind,
                                               ps.inter_residue_connections)
    numpy.testing.assert_array_equal(ibs, ps.inter_block_bondsep.numpy())
    numpy.testing.assert_array_equal(mins, ps.min_block_bondsep.numpy())


def test_cyclic_and_multichain_match_atom_graph():
    ps, irc = cyclic_and_multichain_stack()
    ibs, mins = orc.oracle_inter_block_bondsep(ps.packed_block_types, ps.block_type_ind, irc)
    want = atom_graph_bondsep(ps, irc)
    numpy.testing.assert_array_equal(ibs, want)
    numpy.testing.assert_array_equal(mins, want.min(axis=(3, 4)))
    # the cycle closure is visible: first and last residue are 1 bond apart
    assert mins[0, 0, 13] == 1 and mins[0, 13, 0] == 1
    # and differs from the open chain
    assert not numpy.array_equal(ibs, ps.inter_block_bondsep.numpy())


@pytest.mark.skipif(not __import__("os").path.exists(
    __import__("os").path.join(orc.HERE, "_ref", "tmol_apsp_ref.so")), reason="oracle/_ref not built")
def test_shortest_paths_match_reference_apsp_op():
    """The connection-point graph of each pose, solved by the reference's
    own apsp op with cutoff 6, scattered to blocks, equals the oracle."""
    ps, irc = cyclic_and_multichain_stack()
    pbt = ps.packed_block_types
    apsp = orc.load_reference_apsp_op()
    ibs, _ = orc.oracle_inter_block_bondsep(pbt, ps.block_type_ind, irc)
    bti = ps.block_type_ind.numpy()
    nconn = pbt.n_conn.numpy()
    conn = pbt.conn_atom.numpy()
    pd = pbt.ljlk_bond_separation.numpy()
    for p in range(ps.n_poses):
        points = [(b, c) for b in range(ps.max_n_blocks) if bti[p, b] >= 0 for c in range(nconn[bti[p, b]])]
        index = {bc: i for i, bc in enumerate(points)}
        n = len(points)
        W = torch.full((1, n, n), 6, dtype=torch.int32)
        for i, (b, c) in enumerate(points):
            t = bti[p, b]
            for c2 in range(nconn[t]):
                W[0, i, index[(b, c2)]] = int(pd[t, conn[t, c], conn[t, c2]])
            b2, cc2 = irc[p, b, c].tolist()
            if b2 >= 0:
                W[0, i, index[(b2, cc2)]] = 1
        apsp(W, 6)
        for i, (b1, c1) in enumerate(points):
            for j, (b2, c2) in enumerate(points):
                assert ibs[p, b1, b2, c1, c2] == min(int(W[0, i, j]), 6)
