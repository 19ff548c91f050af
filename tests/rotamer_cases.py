"""Rotamer sets with many rotamers per block, for the rotamer-pair tests.

The repacking sets of ``tmol_amd.rotamers.build_repack_rotamers`` hold at
most 1 + 3^3 = 28 rotamers per block, while packing sets routinely carry
hundreds.  ``densify_rotamers`` grows chosen blocks of a set to a given
rotamer count by cycling through the block's rotamers and jittering every
copy after the first round (Gaussian noise on the coordinates), so the
rotamer kernels' windows past the first 64 rotamers of a block, and the
work units that cross them, are exercised.
"""

import numpy
import torch

from tmol_amd.rotamers import RotamerSet


def densify_rotamers(rs: RotamerSet, targets: dict, n_atoms_of_type, jitter: float = 0.15,
                     seed: int = 0) -> RotamerSet:
    """``rs`` with block (pose, block) -> n rotamers for every key of
    ``targets`` (other blocks unchanged), numbered pose by pose and block by
    block as the reference's rotamer sets are."""
    rng = numpy.random.default_rng(seed)
    P, B = rs.n_rots_for_block.shape
    nrb = rs.n_rots_for_block.numpy()
    rob = rs.rot_offset_for_block.numpy()
    bt = rs.block_type_ind_for_rot.numpy()
    off = rs.coord_offset_for_rot.numpy()
    nat = numpy.asarray(n_atoms_of_type)
    x = rs.coords.numpy()
    n_rots_for_block = numpy.zeros((P, B), dtype=numpy.int32)
    rot_offset_for_block = numpy.zeros((P, B), dtype=numpy.int32)
    n_rots_for_pose = numpy.zeros(P, dtype=numpy.int32)
    rot_offset_for_pose = numpy.zeros(P, dtype=numpy.int32)
    pose_for_rot, bt_for_rot, block_for_rot, coord_off, parts = [], [], [], [], []
    r = a = 0
    for p in range(P):
        rot_offset_for_pose[p] = r
        for b in range(B):
            rot_offset_for_block[p, b] = r
            n0 = int(nrb[p, b])
            n = int(targets.get((p, b), n0)) if n0 > 0 else 0
            n_rots_for_block[p, b] = n
            for k in range(n):
                src = int(rob[p, b]) + k % n0
                na = int(nat[bt[src]])
                c = x[off[src]: off[src] + na].copy()
                if k >= n0:
                    c = c + rng.normal(0.0, jitter, size=c.shape).astype(c.dtype)
                pose_for_rot.append(p)
                bt_for_rot.append(int(bt[src]))
                block_for_rot.append(b)
                coord_off.append(a)
                parts.append(c)
                a += na
                r += 1
        n_rots_for_pose[p] = r - rot_offset_for_pose[p]
    t32 = lambda v: torch.tensor(numpy.asarray(v, dtype=numpy.int32))  # noqa: E731
    return RotamerSet(
        n_rots_for_pose=t32(n_rots_for_pose),
        rot_offset_for_pose=t32(rot_offset_for_pose),
        n_rots_for_block=t32(n_rots_for_block),
        rot_offset_for_block=t32(rot_offset_for_block),
        pose_for_rot=t32(pose_for_rot),
        block_type_ind_for_rot=t32(bt_for_rot),
        block_ind_for_rot=t32(block_for_rot),
        coord_offset_for_rot=t32(coord_off),
        coords=torch.tensor(numpy.concatenate(parts, 0), dtype=rs.coords.dtype),
    )


def dense_case(n_res: int = 14, seed: int = 5):
    """Two short poses with repacking rotamers, grown so that several
    neighbouring blocks hold 65 to 200 rotamers: {dtype: (PoseStack,
    RotamerSet)}, the f32 set the f64 set's coordinates rounded."""
    from tmol_amd.chemical import default_chemical_database
    from tmol_amd.generator import build_chains_fast, workload_specs
    from tmol_amd.pose import Pose, PoseStack
    from tmol_amd.rotamers import build_repack_rotamers

    db = default_chemical_database()
    specs = workload_specs("padded_64x150_600", 2)
    sp = []
    for k, n in enumerate((n_res, n_res - 4)):
        rts, phis, psis, omegas, chis = specs[k]
        r = list(rts[:n])
        r[-1] = db.residue_type(r[-1].base_name, cterm=True)
        sp.append((r, phis[:n], psis[:n], omegas[:n], chis[:n]))
    coords = build_chains_fast(sp)
    poses = [Pose(s[0], c) for s, c in zip(sp, coords)]
    ps64 = PoseStack(poses, dtype=torch.float64)
    rs = build_repack_rotamers(sp, ps64, seed=seed)
    nat = ps64.packed_block_types.n_atoms.numpy()
    targets = {(0, 2): 65, (0, 3): 130, (0, 4): 200, (0, 7): 97, (1, 1): 129, (1, 2): 64,
               (1, 5): 150}
    dense = densify_rotamers(rs, targets, nat, seed=seed)
    ps32 = PoseStack(poses, dtype=torch.float32)
    d32 = RotamerSet(**{k: getattr(dense, k) for k in (
        "n_rots_for_pose", "rot_offset_for_pose", "n_rots_for_block", "rot_offset_for_block",
        "pose_for_rot", "block_type_ind_for_rot", "block_ind_for_rot", "coord_offset_for_rot")},
        coords=dense.coords.to(torch.float32))
    return {torch.float64: (ps64, dense), torch.float32: (ps32, d32)}
