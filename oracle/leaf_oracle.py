"""CPU restatement of the reference's missing-leaf-atom builder --
TEST INFRASTRUCTURE.

Only tests/ (and bench.py's cpu_baseline leg) use this module, as the
checker.  ``oracle_leaf_atoms`` takes the 12 arguments of
``tmol_io::gen_pose_leaf_atoms`` and follows
tmol/io/details/compiled/gen_pose_leaf_atoms.impl.hh:25-200 (one item per
pose / block / atom slot: given atoms copied, missing ones built from their
ancestors), tmol/io/details/compiled/leaf_atoms.hh (select_leaf_ancestors:
the primary ancestor set if all three resolve to present atoms, else the
backup set, else no coordinate -- NaN), tmol/score/common/uaid_util.hh
(an ancestor named by (connection, bonds from it) is the atom that many
bonds into the block across that connection) and
tmol/kinematics/compiled/kernel_operations / gen_coord.hh (build_coordinate:
the NeRF step from three ancestors and (phi, theta, d)).

The index work is plain Python; the coordinates are built with torch on the
CPU so that autograd gives the backward independently of any hand-written
derivative.  Pinned by: the reference C++ op built into oracle/_ref
(forward and backward), the reference test's geometric check on 1ubq
(rebuilt leaf atoms within 0.1 A of the deposited ones) and gradcheck.
"""

import os

import numpy
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def _resolve(u, p, b, boff, btype, irc, down):
    # uaid_util.hh: resolve_atom_from_uaid
    a, c, k = (int(x) for x in u)
    if a != -1:
        return int(boff[p, b]) + a
    if c == -1:
        return -1
    ob, oc = (int(x) for x in irc[p, b, c])
    if ob == -1:
        return -1
    idx = int(down[btype[p, ob], oc, k])
    if idx < 0:
        return -1
    return int(boff[p, ob]) + idx


def leaf_plan(args):
    """(dst, src_anc [n, 3], geom [n, 3], copy_dst, unbuilt_dst): for each
    built atom, its pose-wide slot, its three ancestors' slots and geometry
    (leaf_atoms.hh select_leaf_ancestors); the given atoms' slots; the
    missing atoms with no complete ancestor set.  Slots are pose * NA + i."""
    (orig, bmiss, pmiss, boff, btype, irc, bt_n, down, anc, geom, anc_b, geom_b) = args
    bmiss = numpy.asarray(bmiss.cpu(), dtype=bool)
    pmiss = numpy.asarray(pmiss.cpu(), dtype=bool)
    boff, btype, irc = (numpy.asarray(t.cpu()) for t in (boff, btype, irc))
    bt_n, down, anc, anc_b = (numpy.asarray(t.cpu()) for t in (bt_n, down, anc, anc_b))
    geom = numpy.asarray(geom.cpu(), dtype=numpy.float64)
    geom_b = numpy.asarray(geom_b.cpu(), dtype=numpy.float64)
    P, NA = orig.shape[0], orig.shape[1]
    B, Amax = bmiss.shape[1], bmiss.shape[2]
    dst, src, g, copy, unbuilt = [], [], [], [], []
    for p in range(P):
        for b in range(B):
            bt = int(btype[p, b])
            if bt < 0:
                continue
            for at in range(int(bt_n[bt])):
                slot = p * NA + int(boff[p, b]) + at
                if not bmiss[p, b, at]:
                    copy.append(slot)
                    continue
                found = None
                for A, G in ((anc, geom), (anc_b, geom_b)):
                    r = [_resolve(A[bt, at, j], p, b, boff, btype, irc, down) for j in range(3)]
                    if all(x != -1 and not pmiss[p, x] for x in r):
                        found = (r, G[bt, at])
                        break
                if found is None:
                    unbuilt.append(slot)
                    continue
                dst.append(slot)
                src.append([p * NA + x for x in found[0]])
                g.append(found[1])
    return (numpy.array(dst, dtype=numpy.int64), numpy.array(src, dtype=numpy.int64).reshape(-1, 3),
            numpy.array(g, dtype=numpy.float64).reshape(-1, 3),
            numpy.array(copy, dtype=numpy.int64), numpy.array(unbuilt, dtype=numpy.int64))


def build_coordinate(A, B, C, phi, theta, d):
    """NeRF step: the atom at distance d from A, bond angle theta (the
    supplement, as stored) to B and dihedral phi to C about A-B."""
    e0 = A - B
    e0 = e0 / torch.linalg.norm(e0, dim=-1, keepdim=True)
    e2 = torch.cross(e0, C - B, dim=-1)
    e2 = e2 / torch.linalg.norm(e2, dim=-1, keepdim=True)
    e1 = torch.cross(e2, e0, dim=-1)
    k0 = d * torch.cos(torch.pi - theta)
    s = d * torch.sin(torch.pi - theta)
    return (A + k0[:, None] * e0 + (s * torch.cos(phi))[:, None] * e1
            + (s * torch.sin(phi))[:, None] * e2)


def oracle_leaf_atoms(*args, plan=None):
    """The op's output [P, NA, 3] in the dtype of orig_coords; slots that are
    neither given nor built are 0, missing atoms with no ancestors NaN.
    Differentiable with respect to orig_coords (args[0])."""
    orig = args[0]
    P, NA = orig.shape[0], orig.shape[1]
    dst, src, g, copy, unbuilt = plan if plan is not None else leaf_plan(args)
    flat = orig.reshape(P * NA, 3)
    out = torch.zeros_like(flat)
    out = out.index_copy(0, torch.as_tensor(copy), flat[torch.as_tensor(copy)]) \
        if len(copy) else out
    if len(dst):
        s = torch.as_tensor(src)
        gg = torch.as_tensor(g, dtype=orig.dtype)
        W = build_coordinate(flat[s[:, 0]], flat[s[:, 1]], flat[s[:, 2]],
                             gg[:, 0], gg[:, 1], gg[:, 2])
        out = out.index_copy(0, torch.as_tensor(dst), W)
    if len(unbuilt):
        nan = torch.full((len(unbuilt), 3), float("nan"), dtype=orig.dtype)
        out = out.index_copy(0, torch.as_tensor(unbuilt), nan)
    return out.reshape(P, NA, 3)


def load_reference_io_ops():
    """The reference's own CPU ingest ops built by build_ref.sh
    (``torch.ops.tmol_io``: gen_pose_leaf_atoms, resolve_his_taut)."""
    path = os.path.join(HERE, "_ref", "tmol_io_ref.so")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    if not hasattr(torch.ops, "tmol_io") or not hasattr(torch.ops.tmol_io,
                                                         "gen_pose_leaf_atoms"):
        torch.ops.load_library(path)
    return torch.ops.tmol_io
