"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and its host-only helpers behave (no device calls)."""

import ctypes
import os
import re

import numpy
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tmol_ljlk.h")


def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tmol_(?:ljlk|elec|lbfgs|lkball|hbond|leaf_atoms)_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from tmol_amd import ops

    L = ops.lib()
    syms = _declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s


def test_neighbor_tiles_host():
    from tmol_amd import ops

    L = ops.lib()
    pbs = numpy.array([0, 10, 150, 150, 300], dtype=numpy.int32)  # includes an empty pose
    n = L.tmol_ljlk_count_neighbor_tiles(pbs.ctypes.data_as(ops._P32), 4)
    t = [(nb + 63) // 64 for nb in numpy.diff(pbs)]
    assert n == sum(k * (k + 1) // 2 for k in t)
    tiles = numpy.zeros((n, 4), dtype=numpy.int32)
    L.tmol_ljlk_fill_neighbor_tiles(pbs.ctypes.data_as(ops._P32), 4, tiles.ctypes.data_as(ops._P32))
    # every (b1 <= b2) pair of every pose is covered exactly once
    for p in range(4):
        s, e = pbs[p], pbs[p + 1]
        cover = numpy.zeros((e - s, e - s), dtype=int)
        for r0, c0, end, pp in tiles[tiles[:, 3] == p]:
            assert end == e
            for b1 in range(r0, min(r0 + 64, e)):
                for b2 in range(c0, min(c0 + 64, e)):
                    if b1 <= b2:
                        cover[b1 - s, b2 - s] += 1
        iu = numpy.triu_indices(e - s)
        assert (cover[iu] == 1).all()
        assert cover.sum() == len(iu[0])


def test_workspace_bytes_and_validation():
    from tmol_amd import ops

    L = ops.lib()
    s = ops.CStack()
    s.layout = ops.PADDED
    s.n_poses = 2
    s.max_n_blocks = 10
    s.n_blocks_total = 20
    s.n_atoms_total = 300
    b = ops.CBlockTypes()
    b.n_block_types = 3
    b.max_n_atoms = 24
    b.max_n_tiles = 1
    b.max_n_conn = 2
    b.n_atom_types = 105
    b.dtype = ops.F32
    n32 = L.tmol_ljlk_workspace_bytes(ctypes.byref(s), ctypes.byref(b), 1000, ops.F32)
    n64 = L.tmol_ljlk_workspace_bytes(ctypes.byref(s), ctypes.byref(b), 1000, ops.F64)
    assert n32 > 8 * 1000 and n64 > n32
    # argument validation happens before any device work
    st = L.tmol_ljlk_score(ctypes.byref(s), ctypes.byref(b), None, ops.F32, None, 0, 0, 0, None,
                           None, None, None, 0, None)
    assert st == ops.EINVAL


def test_cpu_tensor_is_rejected(ubq_pdb):
    from tmol_amd.io_pdb import pose_from_pdb
    from tmol_amd.ops import LJLKLibraryError, ljlk_pose_scores
    from tmol_amd.pose import PoseStack
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

    ps = PoseStack([pose_from_pdb(ubq_pdb, (0, 4))])
    with pytest.raises(LJLKLibraryError):
        ljlk_pose_scores(*pose_op_args(ps, LJLKEnergyTerm()))


# The signature of ljlk_pose_scores_op and ljlk_rotamer_scores_op,
# tmol/score/ljlk/potentials/compiled.ops.cpp:533-561 and 597-625, written out
# here as (type, name) pairs in their order; TORCH_LIBRARY(tmol_ljlk, m)
# (compiled.ops.cpp:667-670) returns Tensor[].
REFERENCE_SIGNATURE = [
    ("Tensor", "rot_coords"), ("Tensor", "rot_coord_offset"), ("Tensor", "pose_ind_for_atom"),
    ("Tensor", "first_rot_for_block"), ("Tensor", "first_rot_block_type"),
    ("Tensor", "block_ind_for_rot"), ("Tensor", "pose_ind_for_rot"),
    ("Tensor", "block_type_ind_for_rot"), ("Tensor", "n_rots_for_pose"),
    ("Tensor", "rot_offset_for_pose"), ("Tensor", "n_rots_for_block"),
    ("Tensor", "rot_offset_for_block"), ("int", "max_n_rots_per_pose"),
    ("Tensor", "pose_stack_min_bond_separation"), ("Tensor", "pose_stack_inter_block_bondsep"),
    ("Tensor", "block_type_n_atoms"), ("Tensor", "block_type_n_heavy_atoms_in_tile"),
    ("Tensor", "block_type_heavy_atoms_in_tile"), ("Tensor", "block_type_atom_types"),
    ("Tensor", "block_type_n_interblock_bonds"),
    ("Tensor", "block_type_atoms_forming_chemical_bonds"),
    ("Tensor", "block_type_path_distance"), ("Tensor", "block_type_is_ligand_fragment"),
    ("Tensor", "ljlk_type_params"), ("Tensor", "global_params"), ("float", "max_dis"),
    ("bool", "output_block_pair_energies"),
]


def _signature(schema):
    return [(str(a.type), a.name) for a in schema.arguments], [str(r.type) for r in schema.returns]


@pytest.mark.parametrize("name", ["ljlk_pose_scores", "ljlk_rotamer_scores"])
def test_tmol_ljlk_op_schema_matches_reference(name):
    """torch.ops.tmol_ljlk.<name> exists with the reference's argument names,
    types and order and a Tensor[] return."""
    import tmol_amd.ops  # noqa: F401

    schema = getattr(torch.ops.tmol_ljlk, name).default._schema
    args, rets = _signature(schema)
    assert schema.name == "tmol_ljlk::" + name
    assert args == REFERENCE_SIGNATURE
    assert rets == ["List[Tensor]"]


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "tmol_ljlk_ref.so")),
                    reason="oracle/_ref not built (needs the reference tree)")
def test_tmol_ljlk_op_types_match_reference_build():
    """The argument and return types equal those of the reference's own
    registration, read from the library built from its sources (the
    reference's schema is inferred from the C++ function, so its argument
    names are positional)."""
    import tmol_amd.ops  # noqa: F401
    from oracle import refop

    ref = refop.schemas()
    for name in ("ljlk_pose_scores", "ljlk_rotamer_scores"):
        ours = str(getattr(torch.ops.tmol_ljlk, name).default._schema)
        strip = lambda sch: [a.strip().rsplit(" ", 1)[0] for a in sch[sch.index("(") + 1: sch.index(")")].split(",")]  # noqa: E731
        assert strip(ours) == strip(ref[name])
        assert ours.split("->")[1].split()[0] == ref[name].split("->")[1].split()[0] == "Tensor[]"


def test_rotamer_and_topology_validation():
    """Argument checks of the rotamer-pair and topology entry points run on
    the host, before any device work."""
    from tmol_amd import ops

    L = ops._rot_lib()
    b = ops.CBlockTypes()
    b.n_block_types = 3
    b.max_n_atoms = 24
    b.max_n_tiles = 1
    b.max_n_conn = 2
    b.n_atom_types = 105
    b.dtype = ops.F32
    r = ops.CRotamerSet()
    r.n_poses, r.max_n_blocks, r.n_rots = 2, 10, 40
    r.n_atoms_total = 600
    # missing tables: the size query refuses
    assert L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F32) == -1
    dummy = ctypes.c_void_p(16)
    for f in ("rot_coord_offset", "block_ind_for_rot", "block_type_ind_for_rot", "n_rots_for_block",
              "rot_offset_for_block", "first_rot_block_type", "min_bondsep", "inter_block_bondsep"):
        setattr(r, f, dummy)
    n32 = L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F32)
    # the dtype must match the block-type tables'
    assert L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F64) == -1
    b.dtype = ops.F64
    n64 = L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F64)
    b.dtype = ops.F32
    assert n32 > 0 and n64 > n32
    # dtype mismatch with the block-type tables, and block types over 64 atoms
    assert L.tmol_ljlk_rotamer_neighbors(ctypes.byref(r), ctypes.byref(b), dummy, ops.F64, 6.0,
                                         dummy, n64, None) == ops.EINVAL
    b.max_n_atoms, b.max_n_tiles = 70, 3
    assert L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F32) == -1
    assert L.tmol_ljlk_rotamer_score(ctypes.byref(r), ctypes.byref(b), dummy, ops.F32, dummy, n32,
                                     dummy, 1, 1, 0, None, dummy, None, 0, None) == ops.ENOTSUP
    b.max_n_atoms, b.max_n_tiles = 24, 1
    # unsupported score flags; more pairs than int32 dispatch indices hold
    assert L.tmol_ljlk_rotamer_score(ctypes.byref(r), ctypes.byref(b), dummy, ops.F32, dummy, n32,
                                     dummy, 1, 1, ops.DVDX, None, dummy, None, 0, None) == ops.ENOTSUP
    assert L.tmol_ljlk_rotamer_dispatch(ctypes.byref(r), ctypes.byref(b), ops.F32, 6.0, dummy, n32,
                                        dummy, 1 << 31, dummy, 1, None) == ops.EOVERFLOW
    # topology: null pointers and too many connections
    assert ops.lib().tmol_ljlk_inter_block_bondsep(ctypes.byref(b), 2, 10, None, None, None, None,
                                                   None, None) == ops.EINVAL
    b.max_n_conn = 5
    assert ops.lib().tmol_ljlk_inter_block_bondsep(ctypes.byref(b), 2, 10, dummy, dummy, dummy, None,
                                                   None, None) == ops.ENOTSUP


def test_elec_validation():
    """Argument checks of the elec entry points run on the host before any
    device work: missing tables, block types over 32 atoms, unordered
    potential regions, non-finite parameters, bad flags."""
    from tmol_amd import ops

    L = ops._elec_lib()
    dummy = ctypes.c_void_p(16)
    s = ops.CStack()
    s.layout = ops.PADDED
    s.n_poses, s.max_n_blocks, s.n_blocks_total, s.n_atoms_total = 2, 10, 20, 300
    s.block_atom_offset = s.block_type = s.min_bondsep = s.inter_block_bondsep = dummy
    b = ops.CBlockTypes()
    b.n_block_types, b.max_n_atoms, b.max_n_tiles, b.max_n_conn = 3, 24, 1, 2
    b.dtype = ops.F32
    b.n_atoms = b.n_conn = b.conn_atoms = b.is_ligand = dummy
    e = ops.CElecTables()
    for i, v in enumerate((79.931, 6.648, 0.441546, 1.6, 5.5)):
        e.global_params[i] = v
    nbytes = L.tmol_ljlk_workspace_bytes(ctypes.byref(s), ctypes.byref(b), 10, ops.F32)

    def score(flags=0, energies=dummy):
        return L.tmol_elec_score(ctypes.byref(s), ctypes.byref(b), ctypes.byref(e), dummy,
                                 ops.F32, dummy, nbytes, 10, flags, dummy, energies, dummy, dummy,
                                 0, None)

    assert score() == ops.EINVAL  # tables missing
    e.partial_charge = dummy
    e.inter_repr_path_distance = ctypes.cast(dummy, ops._P32)
    e.intra_repr_path_distance = ctypes.cast(dummy, ops._P32)
    assert score(flags=ops.DVDX | ops.WGRAD) == ops.EINVAL
    assert score(flags=ops.BLOCK_PAIR | ops.DVDX) == ops.ENOTSUP
    assert score(flags=ops.GENERAL_KERNEL) == ops.EINVAL
    assert score(energies=None) == ops.EINVAL
    b.max_n_atoms, b.max_n_tiles = 40, 2
    assert score() == ops.ENOTSUP
    b.max_n_atoms, b.max_n_tiles = 24, 1
    e.global_params[3] = 5.0  # min_dis + 0.25 beyond max_dis - 1
    assert score() == ops.EINVAL
    e.global_params[3] = float("nan")
    assert score() == ops.EINVAL
    e.global_params[3] = 1.6
    assert L.tmol_elec_neighbors(ctypes.byref(s), ctypes.byref(b), ctypes.byref(e), dummy, ops.F32,
                                 float("inf"), dummy, 1, dummy, nbytes, 10, None, None) == ops.EINVAL
    assert L.tmol_elec_neighbors(ctypes.byref(s), ctypes.byref(b), ctypes.byref(e), dummy, ops.F32,
                                 5.5, dummy, 1, dummy, nbytes - 1, 10, None, None) == ops.EINVAL


def test_elec_rotamer_validation():
    """Argument checks of the elec rotamer entry points (host side, before
    any device work): missing elec tables, block types over 32 atoms, a
    short workspace, unsupported flags, missing outputs."""
    from tmol_amd import ops

    L = ops._elec_lib()
    ops._rot_lib()
    dummy = ctypes.c_void_p(16)
    r = ops.CRotamerSet()
    r.n_poses, r.max_n_blocks, r.n_rots, r.reserved, r.n_atoms_total = 2, 10, 50, 0, 900
    for k in ("rot_coord_offset", "block_ind_for_rot", "block_type_ind_for_rot",
              "n_rots_for_block", "rot_offset_for_block", "first_rot_block_type",
              "min_bondsep", "inter_block_bondsep"):
        setattr(r, k, dummy)
    b = ops.CBlockTypes()
    b.n_block_types, b.max_n_atoms, b.max_n_tiles, b.max_n_conn = 3, 24, 1, 2
    b.dtype = ops.F32
    b.n_atoms = b.n_conn = b.conn_atoms = b.is_ligand = dummy
    e = ops.CElecTables()
    for i, v in enumerate((79.931, 6.648, 0.441546, 1.6, 5.5)):
        e.global_params[i] = v
    nbytes = L.tmol_ljlk_rotamer_workspace_bytes(ctypes.byref(r), ctypes.byref(b), ops.F32)
    assert nbytes > 0

    def nbr(max_dis=5.5, ws=nbytes):
        return L.tmol_elec_rotamer_neighbors(ctypes.byref(r), ctypes.byref(b), ctypes.byref(e),
                                             dummy, ops.F32, max_dis, dummy, ws, None)

    def score(flags=0, energies=dummy, grad=dummy, dTdV=dummy):
        return L.tmol_elec_rotamer_score(ctypes.byref(r), ctypes.byref(b), ctypes.byref(e), dummy,
                                         ops.F32, dummy, nbytes, dummy, 4, 10, flags, dTdV,
                                         energies, grad, 0, None)

    assert nbr() == ops.EINVAL and score() == ops.EINVAL  # elec tables missing
    e.partial_charge = dummy
    e.inter_repr_path_distance = ctypes.cast(dummy, ops._P32)
    e.intra_repr_path_distance = ctypes.cast(dummy, ops._P32)
    assert nbr(ws=nbytes - 1) == ops.EINVAL
    assert nbr(max_dis=float("nan")) == ops.EINVAL
    assert score(flags=ops.WGRAD) == ops.ENOTSUP
    assert score(flags=ops.DVDX) == ops.ENOTSUP
    assert score(energies=None) == ops.EINVAL
    assert score(flags=ops.WGRAD | ops.NO_ENERGY, grad=None) == ops.EINVAL
    assert score(flags=ops.WGRAD | ops.NO_ENERGY, dTdV=None) == ops.EINVAL
    b.max_n_atoms, b.max_n_tiles = 40, 2
    assert nbr() == ops.ENOTSUP and score() == ops.ENOTSUP


def test_noalloc_validation():
    """tmol_ljlk_score_noalloc rejects, before any device work, a row capacity
    below the longest pose and a stack whose workspace exceeds its scratch."""
    from tmol_amd import ops

    L = ops.lib()
    assert L.tmol_ljlk_noalloc_scratch_bytes() == 64 << 20
    dummy = ctypes.c_void_p(16)
    s = ops.CStack()
    s.layout = ops.PADDED
    s.n_poses = 2
    s.max_n_blocks = 10
    s.n_blocks_total = 20
    s.n_atoms_total = 300
    for k in ("block_atom_offset", "block_type", "min_bondsep", "inter_block_bondsep"):
        setattr(s, k, dummy)
    b = ops.CBlockTypes()
    b.n_block_types = 3
    b.max_n_atoms = 24
    b.max_n_tiles = 1
    b.max_n_conn = 2
    b.n_atom_types = 105
    b.dtype = ops.F32

    def call(cap):
        return L.tmol_ljlk_score_noalloc(ctypes.byref(s), ctypes.byref(b), dummy, ops.F32, 6.0,
                                         dummy, 1, cap, 0, None, dummy, None, None, None, None)

    assert call(9) == ops.EINVAL  # below the longest pose
    s.n_poses = 20000
    s.max_n_blocks = 1000
    s.n_blocks_total = 20000 * 1000
    assert L.tmol_ljlk_workspace_bytes(ctypes.byref(s), ctypes.byref(b), 1000, ops.F32) > 64 << 20
    assert call(1000) == ops.ESCRATCH


def test_leaf_atoms_validation():
    """tmol_leaf_atoms_build / _backward reject missing tables, a bad dtype
    and oversized poses before touching the device."""
    from tmol_amd import ops
    from tmol_amd.ingest import CLeafArgs, _leaf_lib

    L = _leaf_lib()
    dummy = ctypes.c_double(0.0)
    p = ctypes.cast(ctypes.pointer(dummy), ctypes.c_void_p).value
    a = CLeafArgs()
    a.n_poses, a.max_n_pose_atoms, a.max_n_blocks, a.max_n_block_atoms, a.max_n_conn = 2, 100, 8, 28, 3
    for name, _ in CLeafArgs._fields_[5:]:
        setattr(a, name, p)
    assert L.tmol_leaf_atoms_build(None, p, ops.F32, p, None) == ops.EINVAL
    assert L.tmol_leaf_atoms_build(ctypes.byref(a), p, 7, p, None) == ops.EINVAL
    assert L.tmol_leaf_atoms_build(ctypes.byref(a), None, ops.F32, p, None) == ops.EINVAL
    assert L.tmol_leaf_atoms_backward(ctypes.byref(a), p, ops.F64, p, None, None) == ops.EINVAL
    a.bt_atom_ancestors_backup = None
    assert L.tmol_leaf_atoms_build(ctypes.byref(a), p, ops.F32, p, None) == ops.EINVAL
    a.bt_atom_ancestors_backup = p
    a.n_poses, a.max_n_pose_atoms = 1 << 16, 1 << 14
    assert L.tmol_leaf_atoms_build(ctypes.byref(a), p, ops.F32, p, None) == ops.ENOTSUP
