"""Host side of the lk_ball HIP path (CPU): the conversion of the
reference ops' tile tables into occluder lists and water recipes equals the
term's own tables, and the C ABI rejects bad lk_ball arguments before any
device work."""

import ctypes

import numpy
import pytest
import torch

from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack
from tmol_amd.score.lk_ball_energy_term import (
    LKBallEnergyTerm, score_op_args, watergen_op_args)


@pytest.fixture(scope="module")
def term():
    return LKBallEnergyTerm()


def test_op_table_conversion_matches_term(ubq_pdb, term):
    from tmol_amd.generator import generate_poses
    from tmol_amd.lk_ball import _occluders_from_op, _recipes_from_op

    ps = PoseStack([pose_from_pdb(ubq_pdb)] + generate_poses(4, (20, 60), seed=3))
    wa = watergen_op_args(ps, term)
    sa = score_op_args(ps, term, torch.zeros(1))
    rec = _recipes_from_op(wa[14], wa[18], wa[19], wa[20], wa[21], wa[22], wa[23], wa[24],
                           wa[25], wa[26], wa[27], wa[28], (2, 2, 1))
    n_occ, n_pol, occ, prm = _occluders_from_op(sa[16], sa[19], sa[20], sa[21], sa[22])
    tp = term.lkb_type_params.numpy()
    for i, rt in enumerate(ps.packed_block_types.active_block_types):
        b = term.block_type(rt)
        got = sorted(map(tuple, rec[i].tolist()))
        want = sorted(map(tuple, b.recipes.tolist()))
        assert got == want, rt.name
        assert n_occ[i] == len(b.occ_list) and n_pol[i] == b.n_polar
        assert occ[i, : n_occ[i]].tolist() == b.occ_list
        numpy.testing.assert_array_equal(prm[i, : n_occ[i]], tp[rt.atom_types[b.occ_list]])


def test_cabi_rejects_bad_lk_ball_arguments():
    from tmol_amd import ops
    from tmol_amd.lk_ball import CLKBallTables, _lkb_lib

    L = _lkb_lib()
    s, b, k = ops.CStack(), ops.CBlockTypes(), CLKBallTables()
    # null structs
    assert L.tmol_lkball_waters(None, None, None, None, 0, None, None) == ops.EINVAL
    # an empty stack and empty tables
    assert L.tmol_lkball_waters(ctypes.byref(s), ctypes.byref(b), ctypes.byref(k), None, 0,
                                None, None) == ops.EINVAL
    b.n_block_types, b.max_n_atoms, b.max_n_conn = 1, 4, 2
    k.max_n_occ = 33  # more occluders than the kernel stages
    s.n_poses, s.n_blocks_total = 1, 1
    assert L.tmol_lkball_score(ctypes.byref(s), ctypes.byref(b), ctypes.byref(k), None, None,
                               0, None, 0, 1, 0, None, None, None, None, 0, None) in (
        ops.EINVAL, ops.ENOTSUP)
    b.max_n_conn = 5  # more than four connections
    assert L.tmol_lkball_waters(ctypes.byref(s), ctypes.byref(b), ctypes.byref(k), None, 0,
                                None, None) == ops.ENOTSUP
    # bad dtype
    assert L.tmol_lkball_neighbors(ctypes.byref(s), ctypes.byref(b), None, 7, 6.0, None, 0,
                                   None, 0, 1, None, None) == ops.EINVAL
