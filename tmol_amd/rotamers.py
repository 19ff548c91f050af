"""Rotamer sets for rotamer-pair scoring (packing).

``RotamerSet`` carries the reference's rotamer-set tensors
(tmol/pack/rotamer/rotamer_set.py:9-60): rotamers are numbered globally,
grouped by pose and by block, each with its block type and the offset of its
first atom in a flat coordinate array.

``build_repack_rotamers`` makes a synthetic repacking set for poses built by
the generator: every residue keeps its current conformation as rotamer 0
(the IncludeCurrentSampler role) and gets chi-grid rotamers on its own
backbone: each sampled chi at a centre of {60, 180, -60} degrees plus
N(0, 10) noise, all combinations of the first three chis, a fourth chi at
180.  Alanine, glycine and proline keep only the current conformation.
Coordinates come from the generator's ideal-geometry chain builder, so a
rotamer's backbone atoms coincide with the pose's.

``rotamer_op_args`` is the 27-argument list of ``ljlk_rotamer_scores``, as
TermRotamerScoringModule builds it (tmol/score/common/scoring_module.py:
133-172) followed by the term's attributes of the pose stack.
"""

import itertools
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy
import torch

from .generator import build_chains_fast, chi_definitions

CHI_CENTRES = (60.0, 180.0, -60.0)


@dataclass
class RotamerSet:
    n_rots_for_pose: torch.Tensor        # [P] int32
    rot_offset_for_pose: torch.Tensor    # [P] int32
    n_rots_for_block: torch.Tensor       # [P, B] int32
    rot_offset_for_block: torch.Tensor   # [P, B] int32
    pose_for_rot: torch.Tensor           # [R] int32
    block_type_ind_for_rot: torch.Tensor  # [R] int32
    block_ind_for_rot: torch.Tensor      # [R] int32
    coord_offset_for_rot: torch.Tensor   # [R] int32
    coords: torch.Tensor                 # [NA, 3]
    first_rot_block_type: torch.Tensor = field(default=None)
    max_n_rots_per_pose: int = field(default=None)
    pose_ind_for_atom: torch.Tensor = field(default=None)

    def __post_init__(self):
        # defaults as rotamer_set.py:21-51
        if self.first_rot_block_type is None:
            fr = torch.full_like(self.rot_offset_for_block, -1)
            has = self.n_rots_for_block != 0
            fr[has] = self.block_type_ind_for_rot[self.rot_offset_for_block[has].long()]
            self.first_rot_block_type = fr
        if self.max_n_rots_per_pose is None:
            self.max_n_rots_per_pose = int(self.n_rots_for_pose.max())
        if self.pose_ind_for_atom is None:
            # atoms [offset_r, offset_{r+1}) belong to rotamer r's pose
            n = self.coords.shape[0]
            off = torch.cat([self.coord_offset_for_rot.long(), torch.tensor([n])])
            self.pose_ind_for_atom = torch.repeat_interleave(
                self.pose_for_rot.to(torch.int32), off[1:] - off[:-1])

    @property
    def n_rotamers_total(self):
        return int(self.block_ind_for_rot.shape[0])

    def to(self, device):
        kw = {}
        for k in ("n_rots_for_pose", "rot_offset_for_pose", "n_rots_for_block",
                  "rot_offset_for_block", "pose_for_rot", "block_type_ind_for_rot",
                  "block_ind_for_rot", "coord_offset_for_rot", "coords",
                  "first_rot_block_type", "pose_ind_for_atom"):
            kw[k] = getattr(self, k).to(device)
        return RotamerSet(max_n_rots_per_pose=self.max_n_rots_per_pose, **kw)


def chi_grid(rt, current: dict, rng) -> List[dict]:
    """Chi sets of one residue: the current one first, then the grid."""
    out = [dict(current)]
    if rt.base_name in ("ALA", "GLY", "PRO"):
        return out
    names = [c for c, _ in chi_definitions(rt)]
    if not names:
        return out
    sampled, fixed = names[:3], names[3:]
    for combo in itertools.product(CHI_CENTRES, repeat=len(sampled)):
        d = {}
        for nm, c in zip(sampled, combo):
            d[nm] = float(c + rng.normal(0.0, 10.0))
        for nm in fixed:
            d[nm] = float(180.0 + rng.normal(0.0, 10.0))
        out.append(d)
    return out


def build_repack_rotamers(specs: Sequence, pose_stack, seed: int = 0) -> RotamerSet:
    """Repacking rotamers for the poses built from ``specs`` (generator
    torsion specs: rts, phis, psis, omegas, chis), numbered pose by pose,
    block by block; block types are ``pose_stack``'s packed block types."""
    pbt = pose_stack.packed_block_types
    P, B = pose_stack.n_poses, pose_stack.max_n_blocks
    rng = numpy.random.default_rng(seed)
    per_pose = []
    chains = []
    chain_index = []  # (pose, copy) of each chain
    for p, (rts, phis, psis, omegas, chis) in enumerate(specs):
        grid = [chi_grid(rt, chis[i], rng) for i, rt in enumerate(rts)]
        K = max(len(g) for g in grid)
        per_pose.append(grid)
        for k in range(K):
            ck = [g[k] if k < len(g) else g[0] for g in grid]
            chains.append((rts, phis, psis, omegas, ck))
            chain_index.append((p, k))
    built = build_chains_fast(chains)
    copy_of = {pk: built[i] for i, pk in enumerate(chain_index)}

    n_rots_for_block = numpy.zeros((P, B), dtype=numpy.int32)
    rot_offset_for_block = numpy.zeros((P, B), dtype=numpy.int32)
    n_rots_for_pose = numpy.zeros(P, dtype=numpy.int32)
    rot_offset_for_pose = numpy.zeros(P, dtype=numpy.int32)
    pose_for_rot, bt_for_rot, block_for_rot, coord_off = [], [], [], []
    xyz = []
    r = 0
    a = 0
    for p, grid in enumerate(per_pose):
        rts = specs[p][0]
        rot_offset_for_pose[p] = r
        for b in range(B):
            rot_offset_for_block[p, b] = r
            if b >= len(rts):
                continue
            n_rots_for_block[p, b] = len(grid[b])
            for k in range(len(grid[b])):
                pose_for_rot.append(p)
                bt_for_rot.append(pbt.index[rts[b].name])
                block_for_rot.append(b)
                coord_off.append(a)
                c = copy_of[(p, k)][b]
                xyz.append(c)
                a += c.shape[0]
                r += 1
        n_rots_for_pose[p] = r - rot_offset_for_pose[p]
    t32 = lambda x: torch.tensor(numpy.asarray(x, dtype=numpy.int32))
    return RotamerSet(
        n_rots_for_pose=t32(n_rots_for_pose),
        rot_offset_for_pose=t32(rot_offset_for_pose),
        n_rots_for_block=t32(n_rots_for_block),
        rot_offset_for_block=t32(rot_offset_for_block),
        pose_for_rot=t32(pose_for_rot),
        block_type_ind_for_rot=t32(bt_for_rot),
        block_ind_for_rot=t32(block_for_rot),
        coord_offset_for_rot=t32(coord_off),
        coords=torch.tensor(numpy.concatenate(xyz, 0), dtype=pose_stack.coords.dtype),
    )


def current_rotamers(pose_stack) -> RotamerSet:
    """One rotamer per real block: the pose's own conformation.  Rotamer-pair
    energies of this set equal the stack's block-pair energies."""
    P, B = pose_stack.n_poses, pose_stack.max_n_blocks
    bti = pose_stack.block_type_ind.cpu()
    real = bti >= 0
    n_rots_for_block = real.to(torch.int32)
    flat = n_rots_for_block.flatten()
    csum = torch.cumsum(flat, 0) - flat
    rot_offset_for_block = csum.view(P, B).to(torch.int32)
    n_rots_for_pose = n_rots_for_block.sum(1).to(torch.int32)
    rot_offset_for_pose = (torch.cumsum(n_rots_for_pose, 0) - n_rots_for_pose).to(torch.int32)
    pp, bb = real.nonzero(as_tuple=True)
    A = pose_stack.max_n_atoms
    off_in_pose = pose_stack.block_coord_offset.cpu()[pp, bb]
    src = pose_stack.coords.cpu()
    xyz, coord_off, a = [], [], 0
    nat = pose_stack.packed_block_types.n_atoms.cpu()
    for p, b, o in zip(pp.tolist(), bb.tolist(), off_in_pose.tolist()):
        n = int(nat[bti[p, b]])
        xyz.append(src[p, o : o + n])
        coord_off.append(a)
        a += n
    return RotamerSet(
        n_rots_for_pose=n_rots_for_pose,
        rot_offset_for_pose=rot_offset_for_pose,
        n_rots_for_block=n_rots_for_block,
        rot_offset_for_block=rot_offset_for_block,
        pose_for_rot=pp.to(torch.int32),
        block_type_ind_for_rot=bti[pp, bb].to(torch.int32),
        block_ind_for_rot=bb.to(torch.int32),
        coord_offset_for_rot=torch.tensor(coord_off, dtype=torch.int32),
        coords=torch.cat(xyz, 0).contiguous(),
    )


def rotamer_op_args(pose_stack, rotamer_set: RotamerSet, term, coords=None):
    """The 27 arguments of ``ljlk_rotamer_scores`` in TermRotamerScoringModule
    order (scoring_module.py:146-164): rot_offset_for_block doubles as
    first_rot_for_block; block-pair output is always on."""
    rs = rotamer_set
    if coords is None:
        coords = rs.coords
    i32 = lambda t: t.to(torch.int32)
    return [
        coords.flatten(0, -2),
        i32(rs.coord_offset_for_rot),
        i32(rs.pose_ind_for_atom),
        i32(rs.rot_offset_for_block),
        i32(rs.first_rot_block_type),
        i32(rs.block_ind_for_rot),
        i32(rs.pose_for_rot),
        i32(rs.block_type_ind_for_rot),
        i32(rs.n_rots_for_pose),
        i32(rs.rot_offset_for_pose),
        i32(rs.n_rots_for_block),
        i32(rs.rot_offset_for_block),
        int(rs.max_n_rots_per_pose),
    ] + term.get_score_term_attributes(pose_stack) + [True]


def repack_workload(name: str = "repack_16x300", n_poses=None, seed: int = 1,
                    dtype=torch.float32):
    """(PoseStack, RotamerSet) of a generator workload: its poses and their
    repacking rotamers (rotamer noise from ``seed``)."""
    from .generator import workload_specs
    from .pose import Pose, PoseStack

    specs = workload_specs(name, n_poses)
    coords = build_chains_fast(specs)
    poses = [Pose(sp[0], c) for sp, c in zip(specs, coords)]
    ps = PoseStack(poses, dtype=dtype)
    return ps, build_repack_rotamers(specs, ps, seed=seed)


def dispatch_checksum(indices) -> int:
    """Order-sensitive checksum of a [3, n] dispatch list: sum over entries of
    (i + 1) * (pose * 2^42 + rot1 * 2^21 + rot2), wrapping modulo 2^64."""
    idx = numpy.asarray(indices, dtype=numpy.int64).astype(numpy.uint64)
    n = idx.shape[1]
    key = (idx[0] << numpy.uint64(42)) + (idx[1] << numpy.uint64(21)) + idx[2]
    w = numpy.arange(1, n + 1, dtype=numpy.uint64)
    with numpy.errstate(over="ignore"):
        return int(numpy.sum(w * key, dtype=numpy.uint64))


def subset_rotamers(rs: RotamerSet, keep, n_atoms_of_type) -> RotamerSet:
    """The rotamers with keep[r] true, renumbered in the same order (blocks
    may end up without rotamers); ``n_atoms_of_type`` [T] gives each block
    type's atom count."""
    keep = numpy.asarray(keep, dtype=bool)
    P, B = rs.n_rots_for_block.shape
    blk = rs.block_ind_for_rot.numpy()
    pose = rs.pose_for_rot.numpy()
    bt = rs.block_type_ind_for_rot.numpy()
    off = rs.coord_offset_for_rot.numpy()
    nat = numpy.asarray(n_atoms_of_type)[bt]
    sel = numpy.nonzero(keep)[0]
    counts = numpy.zeros((P, B), dtype=numpy.int64)
    numpy.add.at(counts, (pose[sel], blk[sel]), 1)
    flat = counts.reshape(-1)
    rob = (numpy.cumsum(flat) - flat).reshape(P, B)
    nrp = counts.sum(1)
    rop = numpy.cumsum(nrp) - nrp
    x = rs.coords
    parts = [x[off[r]: off[r] + nat[r]] for r in sel]
    new_off = numpy.cumsum([0] + [nat[r] for r in sel])[:-1]
    t32 = lambda a: torch.tensor(numpy.asarray(a), dtype=torch.int32)
    return RotamerSet(
        n_rots_for_pose=t32(nrp),
        rot_offset_for_pose=t32(rop),
        n_rots_for_block=t32(counts),
        rot_offset_for_block=t32(rob),
        pose_for_rot=t32(pose[sel]),
        block_type_ind_for_rot=t32(bt[sel]),
        block_ind_for_rot=t32(blk[sel]),
        coord_offset_for_rot=t32(new_off),
        coords=torch.cat(parts, 0).contiguous(),
    )
