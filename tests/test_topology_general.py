"""CPU tests of count-pair connectivity for any bond topology: cycles,
disulfides and inter-chain cross-links (tests/topology_cases.py).

* the padded stack's inter_block_bondsep (Dijkstra over connection points,
  tmol_amd.pose.connection_separations) equals the topology oracle, which is
  pinned by breadth-first search over the full atom bond graph and by the
  reference's own APSP op (test_topology_oracle.py);
* the ragged stack's sparse forward bonded table holds exactly the block
  pairs within 5 bonds, with the same connection-point separations.
"""

import numpy
import torch

from oracle import oracle as orc
from test_topology_oracle import atom_graph_bondsep
from tmol_amd.pose import (BONDED_MAX_SEP, BONDED_WIDTH, PoseStack, RaggedPoseStack,
                           conn_graph_bondsep)
from topology_cases import general_topology_poses


def test_padded_bondsep_matches_oracle_and_atom_graph():
    poses = general_topology_poses()
    ps = PoseStack(poses)
    irc = ps.inter_residue_connections
    ibs, mins = orc.oracle_inter_block_bondsep(ps.packed_block_types, ps.block_type_ind, irc)
    numpy.testing.assert_array_equal(ps.inter_block_bondsep.numpy(), ibs)
    numpy.testing.assert_array_equal(ps.min_block_bondsep.numpy(), mins)
    numpy.testing.assert_array_equal(ibs, atom_graph_bondsep(ps, irc))
    for p, pose in enumerate(poses):
        nb = pose.n_blocks
        C = ps.packed_block_types.max_n_conn
        fw = conn_graph_bondsep(pose)
        numpy.testing.assert_array_equal(ibs[p, :nb, :nb, : fw.shape[2], : fw.shape[3]], fw)
        assert (ibs[p, :nb, :nb, fw.shape[2]:C] == 6).all()
    # the extra bonds are visible: the cycle closes, the disulfide bridges
    # residues 5 and 9, the chains are cross-linked
    assert mins[0, 0, 3] == 1
    # down(4)..SG(5): 1 + 3 bonds, then SG(5)-SG(9)
    assert mins[1, 5, 9] == 1 and mins[1, 4, 9] == 5
    assert mins[2, 3, 7] == 1 and mins[2, 5, 6] == 6


def _decode(stack, p):
    """Dense forward separations {(b1, b2): [4, 4]} of pose p from the
    ragged bonded table."""
    start = stack.bonded_start.numpy()
    rows = stack.bonded.numpy().view(numpy.uint32).reshape(-1, BONDED_WIDTH)
    g0 = int(stack.pose_block_start[p])
    g1e = int(stack.pose_block_start[p + 1])
    out = {}
    for g in range(g0, g1e):
        prev = -1
        for k in range(start[g], start[g + 1]):
            r = rows[k]
            g2 = int(r[0])
            assert g < g2 < g1e and g2 > prev
            prev = g2
            sep = numpy.array([[(int(r[4 + c1]) >> (8 * c2)) & 255 for c2 in range(4)]
                               for c1 in range(4)])
            assert int(r[1]) == sep.min()
            out[(g - g0, g2 - g0)] = sep
    return out


def test_ragged_bonded_table_matches_padded():
    poses = general_topology_poses()
    ps = PoseStack(poses)
    rs = RaggedPoseStack(poses)
    ibs = ps.inter_block_bondsep.numpy()
    C = ps.packed_block_types.max_n_conn
    assert rs.bonded.shape[1] == BONDED_WIDTH
    for p, pose in enumerate(poses):
        got = _decode(rs, p)
        nb = pose.n_blocks
        want = {(b1, b2) for b1 in range(nb) for b2 in range(b1 + 1, nb)
                if ibs[p, b1, b2].min() <= BONDED_MAX_SEP}
        assert set(got) == want
        for (b1, b2), sep in got.items():
            numpy.testing.assert_array_equal(sep[:C, :C], ibs[p, b1, b2])
            assert (sep[C:, :] == 6).all() and (sep[:, C:] == 6).all()


def test_ragged_bonded_table_linear_workload():
    """Linear chains take the vectorised path: the table equals the padded
    stack's inter_block_bondsep on the 64-pose workload's first poses."""
    from tmol_amd.generator import workload_poses

    poses = workload_poses("padded_64x150_600", 3)
    ps = PoseStack(poses)
    rs = RaggedPoseStack(poses, dtype=torch.float64)
    ibs = ps.inter_block_bondsep.numpy()
    for p, pose in enumerate(poses):
        got = _decode(rs, p)
        nb = pose.n_blocks
        want = {(b1, b2) for b1 in range(nb) for b2 in range(b1 + 1, nb)
                if ibs[p, b1, b2].min() <= BONDED_MAX_SEP}
        assert set(got) == want and len(want) == 2 * nb - 3
        for (b1, b2), sep in got.items():
            numpy.testing.assert_array_equal(sep[:2, :2], ibs[p, b1, b2])
