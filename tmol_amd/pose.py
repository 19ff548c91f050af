"""Pose stacks: the tensors the LJ+LK op consumes.

Two layouts of the same systems:

* ``PoseStack`` -- the reference's padded layout (tmol/pose/pose_stack.py:
  15-168): coords ``[P, A, 3]``, ``block_coord_offset [P, B]``,
  ``block_type_ind [P, B]`` (-1 = padding), ``inter_block_bondsep
  [P, B, B, C, C]`` and the derived "rotamer" index tensors of
  ``PoseStack.__attrs_post_init__``.  This is what the drop-in op takes.
* ``RaggedPoseStack`` -- CSR over poses: every real block once, atoms packed
  back to back, no padding.  Count-pair connectivity is a sparse forward
  bonded table: for each block, the later blocks of its pose within 4 bonds
  of it (through connection atoms) with their connection-point separations,
  so no ``[P, B, B, ...]`` table is ever materialised, for any bond topology
  (chains, cycles, disulfides, cross-links).

Inter-block bond separation: the reference runs an all-pairs shortest path
over the graph of inter-residue connection points (tmol/pose/
pose_stack_builder.py:436-600, 1230-1350), capped at 6.  For linear polymer
chains that graph is a path: ``... up(i-1) -1- down(i) -d(i)- up(i) -1-
down(i+1) ...`` with d(i) the intra-residue N..C path distance, so the
separation of two connection points is the difference of their positions
along the path, capped at 6, and 6 across chain breaks.  Poses with other
bonds (``Pose.extra_bonds``: cyclisation, disulfides) use a Dijkstra search
over the connection points (``connection_separations``).
``conn_graph_bondsep`` is a Floyd-Warshall used to check both.
"""

from dataclasses import dataclass, field
import heapq
from typing import Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from .chemical import MAX_SIG_BOND_SEPARATION, ResidueType, default_chemical_database

TILE_SIZE = 32


class PackedBlockTypes:
    """Per-block-type tensors, padded over the set of block types in use
    (tmol/pose/packed_block_types.py plus the annotations of
    atom_type_dependent_term.py:84-181, bond_dependent_term.py:31-53 and
    ljlk_energy_term.py:59-111)."""

    def __init__(self, block_types: Sequence[ResidueType], device="cpu"):
        self.active_block_types = list(block_types)
        self.device = torch.device(device)
        self.index = {bt.name: i for i, bt in enumerate(self.active_block_types)}
        T = len(self.active_block_types)
        self.n_types = T
        self.max_n_atoms = max(bt.n_atoms for bt in block_types)
        self.max_n_conn = max(len(bt.connections) for bt in block_types)
        self.max_n_tiles = (self.max_n_atoms - 1) // TILE_SIZE + 1
        Amax, C, nt = self.max_n_atoms, self.max_n_conn, self.max_n_tiles

        n_atoms = numpy.array([bt.n_atoms for bt in block_types], dtype=numpy.int32)
        atom_types = numpy.full((T, Amax), -1, dtype=numpy.int32)
        n_conn = numpy.array([len(bt.connections) for bt in block_types], dtype=numpy.int32)
        conn_atom = numpy.full((T, C), -1, dtype=numpy.int32)
        bond_sep = numpy.full((T, Amax, Amax), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
        heavy_in_tile = numpy.full((T, TILE_SIZE * nt), -1, dtype=numpy.int32)
        n_heavy_in_tile = numpy.zeros((T, nt), dtype=numpy.int32)
        is_lig = numpy.zeros(T, dtype=numpy.int32)
        for i, bt in enumerate(block_types):
            atom_types[i, : bt.n_atoms] = bt.atom_types
            conn_atom[i, : len(bt.connections)] = bt.conn_atom
            bond_sep[i, : bt.n_atoms, : bt.n_atoms] = bt.path_distance
            tiles, counts = tile_subset_indices(bt.heavy_atom_inds, TILE_SIZE)
            heavy_in_tile[i, : tiles.shape[0]] = tiles
            n_heavy_in_tile[i, : counts.shape[0]] = counts
            if not bt.is_polymer:
                # ligand crossover: path distance 3/4 counts as 5
                # (ljlk_energy_term.py:92-104)
                n = bt.n_atoms
                slab = bond_sep[i, :n, :n]
                slab[(slab == 3) | (slab == 4)] = 5
            is_lig[i] = int(bt.is_ligand_fragment)

        def t(a):
            return torch.tensor(a, dtype=torch.int32, device=self.device)

        self.n_atoms = t(n_atoms)
        self.atom_types = t(atom_types)
        self.n_conn = t(n_conn)
        self.conn_atom = t(conn_atom)
        self.bond_separation = t(bond_sep)
        self.ljlk_bond_separation = self.bond_separation
        self.ljlk_heavy_atoms_in_tile = t(heavy_in_tile)
        self.ljlk_n_heavy_atoms_in_tile = t(n_heavy_in_tile)
        self.ljlk_is_ligand_fragment = t(is_lig)

    def to(self, device):
        return PackedBlockTypes(self.active_block_types, device)


def tile_subset_indices(indices: numpy.ndarray, tile_size: int):
    """Group subset indices by ``i // tile_size``, left-justified within each
    tile and expressed as offsets inside the tile; returns the tiled array
    (-1 padded) and the per-tile counts.  Same contract as
    tmol/score/common/stack_condense.py:303-345 (documented example there:
    [0,3,4,7,10,12,14], tile 8 -> [0,3,4,7,-1..,2,4,6,-1..], [4,3])."""
    indices = numpy.asarray(indices, dtype=numpy.int64)
    if indices.size == 0:
        return numpy.full(tile_size, -1, dtype=numpy.int32), numpy.zeros(1, dtype=numpy.int32)
    n_tiles = int(indices.max()) // tile_size + 1
    out = numpy.full(n_tiles * tile_size, -1, dtype=numpy.int32)
    counts = numpy.zeros(n_tiles, dtype=numpy.int32)
    for ind in indices:
        tile = int(ind) // tile_size
        out[tile * tile_size + counts[tile]] = int(ind) - tile * tile_size
        counts[tile] += 1
    return out, counts


@dataclass
class Pose:
    """One system: residue types, per-residue coords, chain bonding.

    ``connected_to_prev[i]`` says residue i's "down" connection is bonded to
    residue i-1's "up" connection."""

    residue_types: List[ResidueType]
    coords: List[numpy.ndarray]  # per residue [n_atoms, 3]
    connected_to_prev: List[bool] = field(default_factory=list)
    # bonds beyond the polymer chain (cyclisation, disulfides, cross-links):
    # (block1, connection1, block2, connection2), connection indices of the
    # blocks' residue types
    extra_bonds: List[Tuple[int, int, int, int]] = field(default_factory=list)

    def __post_init__(self):
        if not self.connected_to_prev:
            self.connected_to_prev = [False] + [True] * (len(self.residue_types) - 1)
        assert len(self.connected_to_prev) == len(self.residue_types)
        for i, ok in enumerate(self.connected_to_prev):
            if ok:
                assert self.residue_types[i].down_connection_ind >= 0
                assert self.residue_types[i - 1].up_connection_ind >= 0
        used = set()
        for a, b in self.bonds():
            for bc in (a, b):
                blk, c = bc
                assert 0 <= blk < self.n_blocks and 0 <= c < len(self.residue_types[blk].connections)
                assert bc not in used, "connection %s bonded twice" % (bc,)
                used.add(bc)

    def bonds(self) -> List[Tuple[Tuple[int, int], Tuple[int, int]]]:
        """Every inter-block bond as ((block, connection), (block, connection))."""
        out = []
        for i in range(1, self.n_blocks):
            if self.connected_to_prev[i]:
                out.append(((i - 1, self.residue_types[i - 1].up_connection_ind),
                            (i, self.residue_types[i].down_connection_ind)))
        for b1, c1, b2, c2 in self.extra_bonds:
            out.append(((b1, c1), (b2, c2)))
        return out

    @property
    def is_linear(self) -> bool:
        """Only chain bonds (down/up connections of consecutive blocks)."""
        return not self.extra_bonds

    @property
    def n_blocks(self):
        return len(self.residue_types)

    @property
    def n_atoms(self):
        return sum(rt.n_atoms for rt in self.residue_types)

    def conn_positions(self):
        """Position of every connection point along its bonded segment, and
        the segment id of each block.  Returns (pos [nb, C_block], seg [nb])."""
        nb = self.n_blocks
        max_c = max(len(rt.connections) for rt in self.residue_types)
        pos = numpy.full((nb, max(max_c, 1)), -1, dtype=numpy.int64)
        seg = numpy.zeros(nb, dtype=numpy.int64)
        cur_seg = -1
        cursor = 0
        if not self.is_linear:
            raise ValueError("chain positions describe linear poses only")
        for i, rt in enumerate(self.residue_types):
            di, ui = rt.down_connection_ind, rt.up_connection_ind
            # connections other than down/up are unbonded here: position -1
            if not self.connected_to_prev[i]:
                cur_seg += 1
                # leave a gap bigger than the cap between segments
                cursor += 4 * MAX_SIG_BOND_SEPARATION
            else:
                cursor += 1  # the peptide bond up(i-1) -> down(i)
            seg[i] = cur_seg
            down_pos = cursor
            if di >= 0:
                pos[i, di] = down_pos
            if ui >= 0:
                d_du = (
                    int(rt.path_distance[rt.conn_atom[di], rt.conn_atom[ui]])
                    if di >= 0
                    else 0
                )
                pos[i, ui] = down_pos + d_du
                cursor = down_pos + d_du
            else:
                cursor = down_pos
        return pos, seg


def chain_bondsep(pos_a, seg_a, pos_b, seg_b):
    """Separation between connection points given segment positions."""
    valid = (pos_a >= 0) & (pos_b >= 0) & (seg_a == seg_b)
    d = numpy.minimum(numpy.abs(pos_a - pos_b), MAX_SIG_BOND_SEPARATION)
    return numpy.where(valid, d, MAX_SIG_BOND_SEPARATION)


def conn_graph_bondsep(pose: Pose) -> numpy.ndarray:
    """Generic inter-block bond separation [nb, nb, C, C] by Floyd-Warshall
    over connection points (small poses; used to check the chain form)."""
    rts = pose.residue_types
    nb = len(rts)
    C = max(len(rt.connections) for rt in rts)
    offs = numpy.cumsum([0] + [len(rt.connections) for rt in rts])
    n = int(offs[-1])
    INF = 10**6
    D = numpy.full((n, n), INF, dtype=numpy.int64)
    for i, rt in enumerate(rts):
        for c1 in range(len(rt.connections)):
            for c2 in range(len(rt.connections)):
                D[offs[i] + c1, offs[i] + c2] = rt.path_distance[rt.conn_atom[c1], rt.conn_atom[c2]]
    for (b1, c1), (b2, c2) in pose.bonds():
        a, b = offs[b1] + c1, offs[b2] + c2
        D[a, b] = D[b, a] = 1
    for k in range(n):
        D = numpy.minimum(D, D[:, k : k + 1] + D[k : k + 1, :])
    D = numpy.minimum(D, MAX_SIG_BOND_SEPARATION)
    out = numpy.full((nb, nb, C, C), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
    for i in range(nb):
        for j in range(nb):
            ni, nj = len(rts[i].connections), len(rts[j].connections)
            out[i, j, :ni, :nj] = D[offs[i] : offs[i] + ni, offs[j] : offs[j] + nj]
    return out


def connection_separations(pose: Pose, cutoff: int = MAX_SIG_BOND_SEPARATION
                           ) -> Dict[Tuple[int, int], Dict[Tuple[int, int], int]]:
    """Shortest bond paths between the connection points of a pose, for any
    bond topology: {(b1, c1): {(b2, c2): d}} for every d < cutoff (absent
    pairs are ``cutoff`` or more).  Dijkstra over the connection-point graph:
    inside a block, its path distance between the two connection atoms;
    across an inter-block bond, 1 (the reference's graph and all-pairs
    shortest paths, pose_stack_builder.py:459-599, 1230-1349)."""
    rts = pose.residue_types
    partner = {}
    for a, b in pose.bonds():
        partner[a] = b
        partner[b] = a
    out = {}
    for blk, rt in enumerate(rts):
        for c in range(len(rt.connections)):
            src = (blk, c)
            dist = {src: 0}
            heap = [(0, src)]
            while heap:
                d, u = heapq.heappop(heap)
                if d > dist[u]:
                    continue
                ub, uc = u
                urt = rts[ub]
                steps = [((ub, c2), int(urt.path_distance[urt.conn_atom[uc], urt.conn_atom[c2]]))
                         for c2 in range(len(urt.connections)) if c2 != uc]
                if u in partner:
                    steps.append((partner[u], 1))
                for v, w in steps:
                    nd = d + w
                    if nd < cutoff and nd < dist.get(v, cutoff):
                        dist[v] = nd
                        heapq.heappush(heap, (nd, v))
            out[src] = dist
    return out


def inter_block_bondsep_of(pose: Pose, C: int) -> numpy.ndarray:
    """inter_block_bondsep [nb, nb, C, C] of one pose (capped at 6)."""
    nb = pose.n_blocks
    out = numpy.full((nb, nb, C, C), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
    if pose.is_linear:
        pos, seg = pose.conn_positions()
        cb = pos.shape[1]
        out[:, :, :cb, :cb] = chain_bondsep(
            pos[:, None, :, None], seg[:, None, None, None],
            pos[None, :, None, :], seg[None, :, None, None],
        )
        return out
    for (b1, c1), dists in connection_separations(pose).items():
        for (b2, c2), d in dists.items():
            out[b1, b2, c1, c2] = d
    return out


# Forward bonded table of a ragged stack: for block g, the blocks g2 > g of
# its pose whose minimum connection-point separation is at most 5 (pairs at 4
# or less need the separations for their count pair; at 5 the reference
# reports the minimum separation itself), sorted by g2.  One entry is
# int32[BONDED_WIDTH]: g2, the minimum separation, 0, 0, then for c1 = 0..3
# the separations to c2 = 0..3 as bytes (byte c2 of word 4 + c1; 6 where no
# path is shorter).
BONDED_WIDTH = 8
BONDED_MAX_SEP = 5


def forward_bonded_entries(pose: Pose, C: int):
    """Block pairs b1 < b2 of the pose whose minimum connection separation is
    <= 5, ordered by (b1, b2): arrays b1 [n], b2 [n], seps [n, C, C]."""
    nb = pose.n_blocks
    if pose.is_linear:
        pos, seg = pose.conn_positions()
        cb = pos.shape[1]
        b1s, b2s, seps = [], [], []
        # chain positions grow by at least 1 per block: a pair more than 5
        # blocks apart is more than 5 bonds apart
        for k in range(1, min(BONDED_MAX_SEP + 1, nb)):
            sep = numpy.full((nb - k, C, C), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
            sep[:, :cb, :cb] = chain_bondsep(pos[:-k, :, None], seg[:-k, None, None],
                                             pos[k:, None, :], seg[k:, None, None])
            keep = numpy.nonzero(sep.reshape(nb - k, -1).min(1) <= BONDED_MAX_SEP)[0]
            b1s.append(keep)
            b2s.append(keep + k)
            seps.append(sep[keep])
        if not b1s:
            return (numpy.zeros(0, numpy.int64), numpy.zeros(0, numpy.int64),
                    numpy.zeros((0, C, C), numpy.int32))
        b1, b2, sep = numpy.concatenate(b1s), numpy.concatenate(b2s), numpy.concatenate(seps)
        order = numpy.lexsort((b2, b1))
        return b1[order], b2[order], sep[order]
    table = {}
    for (b1, c1), dists in connection_separations(pose).items():
        for (b2, c2), d in dists.items():
            if b2 > b1:
                m = table.setdefault((b1, b2), numpy.full((C, C), MAX_SIG_BOND_SEPARATION,
                                                         dtype=numpy.int32))
                m[c1, c2] = d
    keys = [k for k in sorted(table) if table[k].min() <= BONDED_MAX_SEP]
    return (numpy.array([k[0] for k in keys], dtype=numpy.int64),
            numpy.array([k[1] for k in keys], dtype=numpy.int64),
            numpy.array([table[k] for k in keys], dtype=numpy.int32).reshape(-1, C, C))


def bonded_table(poses: Sequence[Pose], pose_block_start: numpy.ndarray, C: int):
    """(bonded_start [G + 1], bonded [n, BONDED_WIDTH]) int32 of a ragged
    stack: the forward bonded table in the layout the kernels read."""
    G = int(pose_block_start[-1])
    g1s, rows = [], []
    for p, pose in enumerate(poses):
        b1, b2, sep = forward_bonded_entries(pose, C)
        if b1.size == 0:
            continue
        g0 = int(pose_block_start[p])
        n = b1.size
        sp = numpy.full((n, 4, 4), MAX_SIG_BOND_SEPARATION, dtype=numpy.int64)
        sp[:, :C, :C] = sep
        row = numpy.zeros((n, BONDED_WIDTH), dtype=numpy.int64)
        row[:, 0] = g0 + b2
        row[:, 1] = sep.reshape(n, -1).min(1)
        row[:, 4:8] = (sp << (8 * numpy.arange(4))[None, None, :]).sum(2)
        g1s.append(g0 + b1)
        rows.append(row)
    if rows:
        g1 = numpy.concatenate(g1s)
        bonded = numpy.concatenate(rows)
    else:
        g1 = numpy.zeros(0, dtype=numpy.int64)
        bonded = numpy.zeros((1, BONDED_WIDTH), dtype=numpy.int64)
    counts = numpy.bincount(g1, minlength=G)
    start = numpy.concatenate([[0], numpy.cumsum(counts)]).astype(numpy.int32)
    # words 4..7 hold four bytes each: keep their bit patterns in int32
    return start, bonded.astype(numpy.uint32).view(numpy.int32)


class PoseStack:
    """Padded stack with the reference's tensor names and conventions."""

    def __init__(self, poses: Sequence[Pose], device="cpu", packed_block_types=None,
                 dtype=torch.float32):
        device = torch.device(device)
        self.device = device
        self.poses = list(poses)
        if packed_block_types is None:
            seen = {}
            for p in poses:
                for rt in p.residue_types:
                    seen.setdefault(rt.name, rt)
            packed_block_types = PackedBlockTypes(list(seen.values()), device)
        pbt = packed_block_types
        self.packed_block_types = pbt
        P = len(poses)
        B = max(p.n_blocks for p in poses)
        A = max(p.n_atoms for p in poses)
        C = pbt.max_n_conn
        self.n_poses, self.max_n_blocks, self.max_n_atoms = P, B, A

        coords = numpy.zeros((P, A, 3), dtype=numpy.float64)
        bco = numpy.zeros((P, B), dtype=numpy.int32)
        bti = numpy.full((P, B), -1, dtype=numpy.int32)
        irc = numpy.full((P, B, C, 2), -1, dtype=numpy.int32)
        ibs = numpy.full((P, B, B, C, C), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
        for p, pose in enumerate(poses):
            off = 0
            for b, rt in enumerate(pose.residue_types):
                bco[p, b] = off
                bti[p, b] = pbt.index[rt.name]
                coords[p, off : off + rt.n_atoms] = pose.coords[b]
                off += rt.n_atoms
            for (b1, c1), (b2, c2) in pose.bonds():
                irc[p, b1, c1] = (b2, c2)
                irc[p, b2, c2] = (b1, c1)
            nb = pose.n_blocks
            ibs[p, :nb, :nb] = inter_block_bondsep_of(pose, C)

        self.coords = torch.tensor(coords, dtype=dtype, device=device)
        self.block_coord_offset = torch.tensor(bco, device=device)
        self.block_coord_offset64 = self.block_coord_offset.to(torch.int64)
        self.block_type_ind = torch.tensor(bti, device=device)
        self.block_type_ind64 = self.block_type_ind.to(torch.int64)
        self.inter_residue_connections = torch.tensor(irc, device=device)
        self.inter_block_bondsep = torch.tensor(ibs, device=device)
        self._post_init()

    def _post_init(self):
        # tmol/pose/pose_stack.py:116-168
        P, B = self.block_coord_offset.shape
        dev = self.device
        block_inds = torch.zeros_like(self.block_coord_offset)
        block_inds[:, :] = torch.arange(0, B, dtype=torch.int32, device=dev)
        self.block_ind_for_rot = block_inds.flatten()
        self.pose_ind_for_rot = (
            torch.arange(0, P, dtype=torch.int32, device=dev).unsqueeze(1).expand(P, B).flatten()
        )
        self.block_type_ind_for_rot = self.block_type_ind.flatten()
        self.rot_offset_for_block = torch.arange(0, P * B, dtype=torch.int32, device=dev).view(P, B)
        self.first_rot_for_block = self.rot_offset_for_block
        self.first_rot_block_type = self.block_type_ind
        self.n_rots_for_pose = torch.full((P,), B, dtype=torch.int32, device=dev)
        self.rot_offset_for_pose = B * torch.arange(0, P, dtype=torch.int32, device=dev)
        coord_offset_for_pose = self.coords.size(1) * torch.arange(
            0, P, dtype=torch.int32, device=dev
        )
        self.n_rots_for_block = torch.full_like(self.block_coord_offset, 1)
        self.rot_coord_offset = (
            self.block_coord_offset.flatten() + torch.repeat_interleave(coord_offset_for_pose, B)
        ).to(torch.int32)
        self.max_n_rots_per_pose = B
        A = self.coords.size(1)
        self.pose_ind_for_atom = (
            torch.arange(P * A, device=dev, dtype=torch.int32) // A
        ).to(torch.int32)
        # BondDependentTerm.setup_poses (bond_dependent_term.py:91-100)
        mbs, _ = torch.min(self.inter_block_bondsep, dim=4)
        mbs, _ = torch.min(mbs, dim=3)
        self.min_block_bondsep = mbs.contiguous()

    def __len__(self):
        return self.n_poses

    @property
    def max_n_pose_atoms(self):
        return self.coords.shape[1]


class RaggedPoseStack:
    """CSR layout over poses; blocks and atoms packed without padding."""

    def __init__(self, poses: Sequence[Pose], device="cpu", packed_block_types=None,
                 dtype=torch.float32):
        device = torch.device(device)
        self.device = device
        self.poses = list(poses)
        if packed_block_types is None:
            seen = {}
            for p in poses:
                for rt in p.residue_types:
                    seen.setdefault(rt.name, rt)
            packed_block_types = PackedBlockTypes(list(seen.values()), device)
        pbt = packed_block_types
        self.packed_block_types = pbt
        C = pbt.max_n_conn
        P = len(poses)
        nb = numpy.array([p.n_blocks for p in poses], dtype=numpy.int64)
        na = numpy.array([p.n_atoms for p in poses], dtype=numpy.int64)
        G = int(nb.sum())
        NA = int(na.sum())
        pose_block_start = numpy.concatenate([[0], numpy.cumsum(nb)]).astype(numpy.int32)
        pose_atom_start = numpy.concatenate([[0], numpy.cumsum(na)]).astype(numpy.int64)
        block_atom_offset = numpy.zeros(G, dtype=numpy.int32)
        block_type = numpy.zeros(G, dtype=numpy.int32)
        block_pose = numpy.zeros(G, dtype=numpy.int32)
        coords = numpy.zeros((NA, 3), dtype=numpy.float64)
        atom_pose = numpy.zeros(NA, dtype=numpy.int32)
        g = 0
        for p, pose in enumerate(poses):
            off = int(pose_atom_start[p])
            atom_pose[off : off + pose.n_atoms] = p
            for b, rt in enumerate(pose.residue_types):
                block_atom_offset[g] = off
                block_type[g] = pbt.index[rt.name]
                block_pose[g] = p
                coords[off : off + rt.n_atoms] = pose.coords[b]
                off += rt.n_atoms
                g += 1
        bonded_start, bonded = bonded_table(poses, pose_block_start, C)
        if NA >= 2**31:
            raise ValueError("ragged stack exceeds 2^31 atoms")

        def t(a, dt=torch.int32):
            return torch.tensor(a, dtype=dt, device=device)

        self.n_poses = P
        self.n_blocks_total = G
        self.n_atoms_total = NA
        self.max_n_blocks = int(nb.max())
        self.coords = torch.tensor(coords, dtype=dtype, device=device)
        self.pose_block_start = t(pose_block_start)
        self.pose_atom_start = t(pose_atom_start, torch.int64)
        self.block_atom_offset = t(block_atom_offset)
        self.block_type = t(block_type)
        self.block_pose = t(block_pose)
        # forward bonded table (count-pair connectivity, any bond topology)
        self.bonded_start = t(bonded_start)
        self.bonded = t(numpy.ascontiguousarray(bonded))
        self.atom_pose = t(atom_pose)
        self.n_blocks_per_pose = t(nb.astype(numpy.int32))
        self.n_atoms_per_pose = t(na.astype(numpy.int32))

    def __len__(self):
        return self.n_poses


def pose_from_sequence_coords(names3: Sequence[str], coords: Sequence[numpy.ndarray],
                              nterm=True, cterm=True, db=None) -> Pose:
    db = db or default_chemical_database()
    n = len(names3)
    rts = [
        db.residue_type(nm, nterm=(nterm and i == 0), cterm=(cterm and i == n - 1))
        for i, nm in enumerate(names3)
    ]
    return Pose(rts, list(coords))
