"""Host side of the hbond HIP path (CPU): the conversion of the reference
ops' tile tables into donor / acceptor lists equals the term's own lists,
the derived-slot table names the oracle's base atoms, and the C ABI rejects
bad hbond arguments before any device work."""

import ctypes

import numpy
import pytest
import torch

from oracle import oracle as orc
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack
from tmol_amd.score.hbond_energy_term import HBondEnergyTerm, bases_op_args, score_op_args


@pytest.fixture(scope="module")
def term():
    return HBondEnergyTerm()


def test_op_table_conversion_matches_term(ubq_pdb, term):
    from tmol_amd.generator import generate_poses
    from tmol_amd.hbond import _lists_from_op

    ps = PoseStack([pose_from_pdb(ubq_pdb)] + generate_poses(4, (20, 60), seed=3))
    ba = bases_op_args(ps, term)
    D, I = orc.oracle_hb_bases(*ba)
    sa = score_op_args(ps, term, torch.tensor(D), torch.tensor(I))
    dons, accs = _lists_from_op(sa[16], sa[20], sa[21], sa[23], sa[24], sa[25], sa[26],
                                sa[29], sa[30], sa[27], sa[28])
    for i, rt in enumerate(ps.packed_block_types.active_block_types):
        b = term.block_type(rt)
        numpy.testing.assert_array_equal(dons[i], b.don, err_msg=rt.name)
        numpy.testing.assert_array_equal(accs[i], b.acc, err_msg=rt.name)


def test_base_slots_match_oracle(ubq_pdb, term):
    """The derived-slot source table of the bases kernel reproduces the
    oracle's derived atom indices on 1ubq (padding atoms included)."""
    from tmol_amd.hbond import _base_src

    ps = PoseStack([pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))])
    ba = bases_op_args(ps, term)
    _, I = orc.oracle_hb_bases(*ba)
    pbt = ps.packed_block_types
    bts = [term.block_type(rt) for rt in pbt.active_block_types]
    src = _base_src(len(bts), pbt.max_n_atoms, [b.don for b in bts], [b.acc for b in bts])
    want = numpy.full_like(I, -1)
    for p in range(ps.n_poses):
        for b in range(ps.max_n_blocks):
            bt = int(ps.block_type_ind[p, b])
            if bt < 0:
                continue
            off = int(ps.rot_coord_offset.view(ps.n_poses, -1)[p, b])
            for a in range(int(pbt.n_atoms[bt])):
                for k in range(3):
                    x = src[bt, a, k]
                    want[off + a, k] = off + x if x >= 0 else -1
    numpy.testing.assert_array_equal(I, want)


def test_cabi_rejects_bad_hbond_arguments():
    from tmol_amd import ops
    from tmol_amd.hbond import CHBondTables, _hb_lib

    L = _hb_lib()
    s, b, h = ops.CStack(), ops.CBlockTypes(), CHBondTables()
    assert L.tmol_hbond_bases(None, None, None, None, 0, None, None, None) == ops.EINVAL
    assert L.tmol_hbond_bases(ctypes.byref(s), ctypes.byref(b), ctypes.byref(h), None, 0, None,
                              None, None) == ops.EINVAL
    b.n_block_types, b.max_n_atoms, b.max_n_conn = 1, 40, 2
    h.n_don_types = h.n_acc_types = 1
    s.n_poses, s.n_blocks_total = 1, 1
    # block types of more than 32 atoms
    assert L.tmol_hbond_bases(ctypes.byref(s), ctypes.byref(b), ctypes.byref(h), None, 0, None,
                              None, None) == ops.ENOTSUP
    b.max_n_atoms = 20
    dummy = (ctypes.c_int32 * 4)()  # checked for null only, never read
    b.n_atoms = b.n_conn = b.conn_atoms = ctypes.cast(dummy, ctypes.c_void_p).value
    h.max_n_acc = 33  # more acceptors than a list holds
    assert L.tmol_hbond_score(ctypes.byref(s), ctypes.byref(b), ctypes.byref(h), None, None,
                              None, 0, None, 0, 1, 0, None, None, None, 0, None) == ops.ENOTSUP
    # bad dtype and reach
    assert L.tmol_hbond_neighbors(ctypes.byref(s), ctypes.byref(b), None, 7, 5.5, None, 0, None,
                                  0, 1, None, None) == ops.EINVAL
    assert L.tmol_hbond_neighbors(ctypes.byref(s), ctypes.byref(b), None, 0, -1.0, None, 0,
                                  None, 0, 1, None, None) == ops.EINVAL
