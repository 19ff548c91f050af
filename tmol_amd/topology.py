"""Pose topology on the device: ``inter_block_bondsep`` [P, B, B, C, C] and
``min_block_bondsep`` [P, B, B] computed by the HIP library from the block
types and the inter-residue connections, in place of the reference's
connection-point graph and all-pairs shortest paths
(tmol/pose/pose_stack_builder.py:459-599, 1230-1349; tmol/pose/compiled/apsp.*;
tmol/score/bond_dependent_term.py:91-100).

``inter_residue_connections`` [P, B, C, 2] holds, for each connection point,
the (block, connection) it is bonded to, -1 for none, in both directions.
``add_connections`` adds extra bonds (cyclisation, cross-links) to such a
tensor.
"""

import ctypes
from typing import Iterable, Tuple

import torch

from . import ops


def _bt_struct(pbt, device):
    """CBlockTypes with the topology fields only (n_conn, conn_atoms,
    path_distance); the tensors are kept on the returned holder."""
    keep = dict(
        n_atoms=ops._i32(pbt.n_atoms, device),
        n_conn=ops._i32(pbt.n_conn, device),
        conn_atoms=ops._i32(pbt.conn_atom, device),
        path_distance=ops._i32(pbt.ljlk_bond_separation, device),
    )
    c = ops.CBlockTypes()
    T, Amax = keep["path_distance"].shape[:2]
    c.n_block_types = T
    c.max_n_atoms = Amax
    c.max_n_tiles = (Amax - 1) // 32 + 1
    c.max_n_conn = keep["conn_atoms"].shape[1]
    c.dtype = ops.F32
    for k, v in keep.items():
        setattr(c, k, ops._p(v))
    return c, keep


def add_connections(irc: torch.Tensor, bonds: Iterable[Tuple[int, int, int, int, int]]) -> torch.Tensor:
    """A copy of ``irc`` [P, B, C, 2] with extra bonds (pose, block1, conn1,
    block2, conn2) stored in both directions."""
    out = irc.clone()
    for p, b1, c1, b2, c2 in bonds:
        out[p, b1, c1, 0], out[p, b1, c1, 1] = b2, c2
        out[p, b2, c2, 0], out[p, b2, c2, 1] = b1, c1
    return out


def inter_block_bondsep(pbt, block_type: torch.Tensor, irc: torch.Tensor, device=None,
                        check_overflow: bool = True):
    """(inter_block_bondsep [P, B, B, C, C], min_block_bondsep [P, B, B]) on
    ``device`` (a ROCm device; raises otherwise)."""
    device = torch.device(device) if device is not None else block_type.device
    if device.type != "cuda":
        raise ops.LJLKLibraryError("topology runs on ROCm devices only; got %s" % device)
    L = ops.lib()
    if not getattr(L, "_topo_bound", False):
        L.tmol_ljlk_inter_block_bondsep.argtypes = [
            ctypes.POINTER(ops.CBlockTypes), ctypes.c_int32, ctypes.c_int32, ops._VP, ops._VP,
            ops._VP, ops._VP, ops._VP, ops._VP]
        L._topo_bound = True
    bts, keep = _bt_struct(pbt, device)
    P, B = block_type.shape
    C = bts.max_n_conn
    if tuple(irc.shape) != (P, B, C, 2):
        raise ops.LJLKLibraryError("inter_residue_connections must be [P, B, C, 2]")
    bt_d = ops._i32(block_type, device)
    irc_d = ops._i32(irc, device)
    ibs = torch.empty((P, B, B, C, C), dtype=torch.int32, device=device)
    mins = torch.empty((P, B, B), dtype=torch.int32, device=device)
    over = torch.zeros(1, dtype=torch.int32, device=device)
    ops._check(L.tmol_ljlk_inter_block_bondsep(
        ctypes.byref(bts), P, B, ops._p(bt_d), ops._p(irc_d), ops._p(ibs), ops._p(mins),
        ops._p(over), ops._stream(device)), "inter_block_bondsep")
    if check_overflow and int(over.item()) != 0:
        raise ops.LJLKLibraryError("a connection point reaches more than 48 points within 5 bonds")
    return ibs, mins
