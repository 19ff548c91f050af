"""GPU tests of the topology kernels (tmol_ljlk_inter_block_bondsep):
inter_block_bondsep [P, B, B, C, C] and min_block_bondsep [P, B, B] equal to
the oracle (cyclic and multi-chain poses) and, at the 64-pose padded
workload's full size (B = 600), to the chain-position separations the
scoring tests use (exact for linear chains; the oracle agrees with them)."""

import sys

import numpy
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def test_cyclic_and_multichain_match_oracle():
    from test_topology_oracle import cyclic_and_multichain_stack
    from tmol_amd.topology import inter_block_bondsep

    _say("topology cyclic/multichain")
    ps, irc = cyclic_and_multichain_stack()
    ibs, mins = inter_block_bondsep(ps.packed_block_types, ps.block_type_ind, irc, DEV)
    oi, om = orc.oracle_inter_block_bondsep(ps.packed_block_types, ps.block_type_ind, irc)
    numpy.testing.assert_array_equal(ibs.cpu().numpy(), oi)
    numpy.testing.assert_array_equal(mins.cpu().numpy(), om)


def test_padded_workload_full_size():
    from tmol_amd.generator import workload_poses
    from tmol_amd.pose import PoseStack
    from tmol_amd.topology import inter_block_bondsep

    _say("topology padded 64 x 150-600")
    ps = PoseStack(workload_poses("padded_64x150_600"))
    ibs, mins = inter_block_bondsep(ps.packed_block_types, ps.block_type_ind,
                                    ps.inter_residue_connections, DEV)
    assert torch.equal(ibs.cpu(), ps.inter_block_bondsep)
    assert torch.equal(mins.cpu(), ps.min_block_bondsep)
    # and a slice against the oracle itself
    sub = slice(0, 2)
    oi, om = orc.oracle_inter_block_bondsep(ps.packed_block_types, ps.block_type_ind[sub],
                                            ps.inter_residue_connections[sub])
    numpy.testing.assert_array_equal(ibs[sub].cpu().numpy(), oi)
