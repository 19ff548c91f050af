"""Canonical-form ingest: residues with named atoms -> PoseStack, missing
leaf atoms (hydrogens, carbonyl / carboxyl oxygens) built on the device.

Mirrors, for the canonical amino acids of this package's chemical database:

* tmol/io/canonical_ordering.py:77-276 -- the canonical ordering: one entry
  per residue name3 (CYD folded into CYS), its atoms the ordered union over
  the residue's variants (termini, disulfide), PDB aliases resolved;
* canonical_ordering.py:464-614 -- ``canonical_form_from_pdb``: chain ids,
  residue types and coordinates [P, L, max atoms, 3] with NaN for absent
  atoms;
* tmol/io/pose_stack_construction.py:15-387 -- ``pose_stack_from_canonical_form``:
  termini by chain (``res_not_connected`` honoured), disulfides from SG-SG
  distances (tmol/io/details/disulfide_search.py: closest pairs within
  2.5 A), block types, and the missing leaf atoms;
* tmol/io/details/build_missing_leaf_atoms.py:225-430 -- which atoms are
  leaves (parents of no internal coordinate and not the heavy fourth atom
  of a named torsion) and each atom's three ancestors and ideal (phi,
  theta, d), with the reference's walk past hydrogen and leaf ancestors and
  its backup ancestor set; the device op ``gen_pose_leaf_atoms`` builds
  them (``tmol_leaf_atoms_build``).

Not covered: the HIS_D tautomer (this database holds the HE2 tautomer
only; a HIS given with HD1 and without HE2 is refused), non-protein
residues, and the geometric completion of regenerated ligand hydrogens
(which applies to no canonical residue).
"""

import ctypes
import math
import threading
from typing import List, Optional

import numpy
import torch

from . import ops
from .chemical import default_chemical_database
from .io_pdb import parse_pdb_residues
from .pose import Pose, PoseStack

DISULFIDE_CUTOFF = 2.5  # disulfide_search.py


# ---------------------------------------------------------------- leaf tables


def _icoor_map(rt):
    return {ic["name"]: ic for ic in rt.icoors}


def leaf_atoms_of(rt, is_h) -> numpy.ndarray:
    """is_leaf [n_atoms] (build_missing_leaf_atoms.py:225-262)."""
    n = rt.n_atoms
    is_parent = numpy.zeros(n, dtype=bool)
    for ic in rt.icoors:
        par = ic["parent"]
        if par in rt.atom_index:
            is_parent[rt.atom_index[par]] = True
    for tor in rt.torsions or []:
        d = tor["d"]
        if "atom" in d and d["atom"] in rt.atom_index:
            k = rt.atom_index[d["atom"]]
            if not is_h[k]:
                is_parent[k] = True
    is_leaf = ~is_parent
    if n == 3:
        is_leaf[:] = False
    return is_leaf


def leaf_icoor_tables(rt, is_h, is_leaf):
    """(anc [n, 3, 3], geom [n, 3], anc_backup, geom_backup) of a residue
    type (build_missing_leaf_atoms.py:341-430): ancestors as unresolved
    atom ids (atom, connection, bonds from it), geometry (phi, pi - theta,
    d) in radians; the great-grandparent walks past leaf ancestors of a
    hydrogen-ancestor chain, and a backup set replaces a great-grandparent
    across a connection or a leaf one."""
    n = rt.n_atoms
    anc = numpy.full((n, 3, 3), -1, dtype=numpy.int32)
    geom = numpy.zeros((n, 3), dtype=numpy.float32)
    anc_b = numpy.full((n, 3, 3), -1, dtype=numpy.int32)
    geom_b = numpy.zeros((n, 3), dtype=numpy.float32)
    if n <= 3:
        return anc, geom, anc_b, geom_b
    ics = _icoor_map(rt)
    conn = {c: i for i, c in enumerate(rt.connection_names)}

    def uaid(name):
        if name in rt.atom_index:
            return (rt.atom_index[name], -1, -1)
        return (-1, conn.get(name, -1), 0)

    def is_atom_h(name):
        return name in rt.atom_index and bool(is_h[rt.atom_index[name]])

    def is_atom_leaf(name):
        return name in rt.atom_index and bool(is_leaf[rt.atom_index[name]])

    def inter_res(name):
        return name not in rt.atom_index

    for j, name in enumerate(rt.atom_names):
        ic = ics[name]
        p_u, gp_u = uaid(ic["parent"]), uaid(ic["grand_parent"])
        phi = math.radians(ic["phi"])
        theta = math.pi - math.radians(ic["theta"])
        d = ic["d"]
        if is_atom_h(ic["great_grand_parent"]):
            seen = {name}
            while is_atom_leaf(ic["great_grand_parent"]):
                g = ic["great_grand_parent"]
                if g in seen:
                    break
                seen.add(g)
                ic = ics[g]
                phi += math.radians(ic["phi"])
        ggp_u = uaid(ic["great_grand_parent"])
        backup = False
        phi_b = phi
        if inter_res(ic["great_grand_parent"]):
            while inter_res(ic["great_grand_parent"]):
                ic = ics[ic["great_grand_parent"]]
                backup = True
                phi_b += math.radians(ic["phi"])
        elif not is_atom_h(ic["great_grand_parent"]):
            seen_b = set()
            while is_atom_leaf(ic["great_grand_parent"]):
                g = ic["great_grand_parent"]
                backup = True
                if g in seen_b:
                    break
                seen_b.add(g)
                ic = ics[g]
                phi_b += math.radians(ic["phi"])
        anc[j] = (p_u, gp_u, ggp_u)
        geom[j] = (phi, theta, d)
        if backup:
            anc_b[j] = (p_u, gp_u, uaid(ic["great_grand_parent"]))
            geom_b[j] = (phi_b, theta, d)
    return anc, geom, anc_b, geom_b


MAINCHAIN = ("N", "CA", "C")


def downstream_of_conn(rt) -> numpy.ndarray:
    """[n_conn, n_atoms]: the atom j bonds into the residue from each
    connection (restypes.py:475-533): along the mainchain from "down",
    back along it from "up" (then its last atom repeated), and up the
    icoor parents from any other connection."""
    n = rt.n_atoms
    out = numpy.full((len(rt.connections), n), -1, dtype=numpy.int32)
    ics = _icoor_map(rt)
    mc = [a for a in MAINCHAIN if a in rt.atom_index]
    for c, (cname, catom) in enumerate((cn[0], cn[1]) for cn in rt.connections):
        if cname == "down" and mc:
            walk = mc
        elif cname == "up" and mc:
            walk = mc[::-1]
        else:
            walk, a = [], catom
            for _ in range(n):
                walk.append(a)
                a = ics[a]["parent"]
        for j in range(n):
            out[c, j] = rt.atom_index[walk[min(j, len(walk) - 1)]]
    return out


class LeafTables:
    """The leaf-atom op's block-type tables for a PackedBlockTypes:
    n_atoms [T], downstream [T, C, Amax] (the connection atom at 0),
    anc / anc_backup [T, Amax, 3, 3], geom / geom_backup [T, Amax, 3],
    is_leaf [T, Amax]."""

    def __init__(self, pbt, atom_is_h):
        T, Amax = pbt.n_types, pbt.max_n_atoms
        C = int(pbt.conn_atom.shape[1]) if pbt.conn_atom.dim() == 2 else 0
        self.is_leaf = numpy.zeros((T, Amax), dtype=bool)
        self.anc = numpy.full((T, Amax, 3, 3), -1, dtype=numpy.int32)
        self.anc_backup = numpy.full((T, Amax, 3, 3), -1, dtype=numpy.int32)
        self.geom = numpy.zeros((T, Amax, 3), dtype=numpy.float32)
        self.geom_backup = numpy.zeros((T, Amax, 3), dtype=numpy.float32)
        self.downstream = numpy.full((T, max(C, 1), Amax), -1, dtype=numpy.int32)
        for i, rt in enumerate(pbt.active_block_types):
            h = atom_is_h(rt)
            lf = leaf_atoms_of(rt, h)
            a, g, ab, gb = leaf_icoor_tables(rt, h, lf)
            n = rt.n_atoms
            self.is_leaf[i, :n] = lf
            self.anc[i, :n], self.geom[i, :n] = a, g
            self.anc_backup[i, :n], self.geom_backup[i, :n] = ab, gb
            self.downstream[i, : len(rt.connections), :n] = downstream_of_conn(rt)
        self.n_atoms = numpy.asarray(pbt.n_atoms.cpu() if torch.is_tensor(pbt.n_atoms)
                                     else pbt.n_atoms, dtype=numpy.int32)


# ---------------------------------------------------------------- device op


class CLeafArgs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "n_poses", "max_n_pose_atoms", "max_n_blocks", "max_n_block_atoms", "max_n_conn")] + [
        (n, ctypes.c_void_p) for n in (
            "block_atom_missing", "pose_atom_missing", "block_coord_offset", "block_type",
            "inter_block_connections", "bt_n_atoms", "bt_atom_downstream_of_conn",
            "bt_atom_ancestors", "bt_atom_icoors", "bt_atom_ancestors_backup",
            "bt_atom_icoors_backup")]


_bound = False
_lock = threading.Lock()


def _leaf_lib():
    global _bound
    L = ops.lib()
    with _lock:
        if not _bound:
            A = ctypes.POINTER(CLeafArgs)
            V = ctypes.c_void_p
            L.tmol_leaf_atoms_build.argtypes = [A, V, ctypes.c_int32, V, V]
            L.tmol_leaf_atoms_backward.argtypes = [A, V, ctypes.c_int32, V, V, V]
            _bound = True
    return L


def _failed(checks):
    """The messages of the checks whose condition holds anywhere; the
    conditions are reduced on their device and read back together (one
    synchronization)."""
    if not checks:
        return []
    flags = torch.stack([c.any() for _, c in checks]).tolist()
    return [m for (m, _), f in zip(checks, flags) if f]


def _validate_leaf_tables(k, P, NA, B, Amax, C):
    """Every index the kernels follow stays inside its table: block types,
    block offsets plus sizes inside the pose, connection partners, ancestor
    ids and the downstream atoms.  Two read-backs: the range checks, then
    the checks that index with the values the first ones cleared."""
    T = int(k["bt_n"].shape[0])
    bt = k["btype"]
    if tuple(k["down"].shape) != (T, C, Amax) or tuple(k["anc"].shape) != (T, Amax, 3, 3) \
            or tuple(k["anc_b"].shape) != (T, Amax, 3, 3) \
            or tuple(k["geom"].shape) != (T, Amax, 3) or tuple(k["geom_b"].shape) != (T, Amax, 3):
        raise ops.LJLKLibraryError("leaf-atom op: block-type table shapes")
    n_of_type = k["bt_n"].view(T, 1, 1)
    checks = [("block_type outside [-1, n_block_types)", (bt < -1) | (bt >= T)),
              ("bt_n_atoms outside [0, max_n_block_atoms]", (k["bt_n"] < 0) | (k["bt_n"] > Amax)),
              ("atom_downstream_of_conn outside [-1, n_atoms of its block type)",
               (k["down"] < -1) | (k["down"] >= n_of_type))]
    for key in ("anc", "anc_b"):
        u = k[key]
        checks.append(("%s outside the block-type tables" % key,
                       (u[..., 0] < -1) | (u[..., 0] >= n_of_type) | (u[..., 1] < -1)
                       | (u[..., 1] >= C) | ((u[..., 1] >= 0) & ((u[..., 2] < 0) | (u[..., 2] >= Amax)))))
    if C > 0:
        ob, oc = k["irc"][..., 0], k["irc"][..., 1]
        checks.append(("inter_block_connections outside [0, max_n_blocks) x [0, max_n_conn)",
                       (ob < -1) | (ob >= B) | ((ob >= 0) & ((oc < 0) | (oc >= C)))))
    bad = _failed(checks)
    if not bad:
        real = bt >= 0
        n = torch.where(real, k["bt_n"][bt.clamp(min=0).long()], torch.zeros_like(bt))
        checks = [("block_coord_offset + n_atoms outside the pose",
                   real & ((k["boff"] < 0) | (k["boff"] + n > NA)))]
        if C > 0:
            ob = k["irc"][..., 0]
            partner = bt.gather(1, ob.clamp(min=0).view(P, -1).long()).view(ob.shape)
            checks.append(("inter_block_connections naming a padding block",
                           (ob >= 0) & (partner < 0)))
        bad = _failed(checks)
    if bad:
        raise ops.LJLKLibraryError("leaf-atom op: " + "; ".join(bad))


class _LeafCall:
    def __init__(self, args):
        (orig, bmiss, pmiss, boff, btype, irc, bt_n, down, anc, geom, anc_b, geom_b) = args
        ops._require_device(orig)
        dev = orig.device
        if orig.dim() != 3 or orig.shape[2] != 3:
            raise ops.LJLKLibraryError("orig_coords must be [n_poses, n_atoms, 3]")
        P, NA = int(orig.shape[0]), int(orig.shape[1])
        P2, B, Amax = (int(x) for x in bmiss.shape)
        if P2 != P or tuple(pmiss.shape) != (P, NA) or tuple(btype.shape) != (P, B):
            raise ops.LJLKLibraryError("leaf-atom op: inconsistent pose shapes")
        u8 = lambda t: t.to(device=dev, dtype=torch.uint8).contiguous()  # noqa: E731
        i32 = lambda t: ops._i32(torch.as_tensor(t), dev)  # noqa: E731
        f = lambda t: torch.as_tensor(t).to(device=dev, dtype=orig.dtype).contiguous()  # noqa
        self.keep = dict(bmiss=u8(bmiss), pmiss=u8(pmiss), boff=i32(boff), btype=i32(btype),
                         irc=i32(irc), bt_n=i32(bt_n), down=i32(down), anc=i32(anc),
                         geom=f(geom), anc_b=i32(anc_b), geom_b=f(geom_b))
        k = self.keep
        C = int(irc.shape[2]) if irc.dim() == 4 else 0
        _validate_leaf_tables(k, P, NA, B, Amax, C)
        a = CLeafArgs()
        a.n_poses, a.max_n_pose_atoms, a.max_n_blocks, a.max_n_block_atoms = P, NA, B, Amax
        a.max_n_conn = int(irc.shape[2]) if irc.dim() == 4 else 0
        for name, key in (("block_atom_missing", "bmiss"), ("pose_atom_missing", "pmiss"),
                          ("block_coord_offset", "boff"), ("block_type", "btype"),
                          ("inter_block_connections", "irc"), ("bt_n_atoms", "bt_n"),
                          ("bt_atom_downstream_of_conn", "down"), ("bt_atom_ancestors", "anc"),
                          ("bt_atom_icoors", "geom"), ("bt_atom_ancestors_backup", "anc_b"),
                          ("bt_atom_icoors_backup", "geom_b")):
            setattr(a, name, k[key].data_ptr())
        self.cargs = a
        self.orig = orig.contiguous()
        self.dcode = ops._dtype_code(orig.dtype)
        self.device = dev

    def build(self):
        out = torch.empty_like(self.orig)
        ops._check(_leaf_lib().tmol_leaf_atoms_build(
            ctypes.byref(self.cargs), ops._p(self.orig), self.dcode, ops._p(out),
            ops._stream(self.device)), "leaf atoms")
        return out

    def backward(self, dnew):
        g = torch.empty_like(self.orig)
        ops._check(_leaf_lib().tmol_leaf_atoms_backward(
            ctypes.byref(self.cargs), ops._p(self.orig), self.dcode,
            ops._p(dnew.to(self.orig.dtype).contiguous()), ops._p(g),
            ops._stream(self.device)), "leaf atoms backward")
        return g


class PoseLeafAtomFunction(torch.autograd.Function):
    """gen_pose_leaf_atoms forward / backward (PoseLeafAtomGen,
    tmol/io/details/compiled/compiled.ops.cpp:19-160)."""

    @staticmethod
    def forward(ctx, *args):
        call = _LeafCall(args)
        ctx.call = call
        ctx.n_args = len(args)
        return call.build()

    @staticmethod
    def backward(ctx, dnew):
        if dnew is None:
            return (None,) * ctx.n_args
        return (ctx.call.backward(dnew),) + (None,) * (ctx.n_args - 1)


def gen_pose_leaf_atoms(*args):
    """Drop-in for ``torch.ops.tmol_io.gen_pose_leaf_atoms`` (12 arguments)."""
    if len(args) != 12:
        raise TypeError("gen_pose_leaf_atoms takes 12 arguments (%d given)" % len(args))
    return PoseLeafAtomFunction.apply(*args)


# ---------------------------------------------------------------- canonical ordering


class CanonicalOrdering:
    """Residue types by name3 (CYD under CYS) and, per type, the canonical
    atom list: the ordered union of its variants' atoms (base, N- and
    C-terminus, both, and the disulfide form), with PDB aliases."""

    def __init__(self, db=None):
        self.db = db or default_chemical_database()
        names = [n for n in self.db.canonical_names if n != "CYD"]
        self.restypes = names
        self.restype_index = {n: i for i, n in enumerate(names)}
        self.atoms: List[List[str]] = []
        self.atom_index: List[dict] = []
        for n in names:
            order, index = [], {}
            variants = [(n, a, b) for a in (False, True) for b in (False, True)]
            if n == "CYS":
                variants += [("CYD", a, b) for a in (False, True) for b in (False, True)]
            for nm, nt, ct in variants:
                rt = self.db.residue_type(nm, nterm=nt, cterm=ct)
                for an in rt.atom_names:
                    if an not in index:
                        index[an] = len(order)
                        order.append(an)
                for alt, canon in rt.aliases.items():
                    if canon in index and alt not in index:
                        index[alt] = index[canon]
            self.atoms.append(order)
            self.atom_index.append(index)
        self.max_n_canonical_atoms = max(len(a) for a in self.atoms)


def canonical_form_from_pdb(pdb_text: str, co: Optional[CanonicalOrdering] = None,
                            residue_start: Optional[int] = None,
                            residue_end: Optional[int] = None):
    """(chain_id [1, L] int32, res_types [1, L] int32, coords [1, L, A, 3]
    float32 with NaN for absent atoms) from the ATOM records of a PDB
    (canonical_ordering.py:464-614); unknown residues and atoms raise.
    ``residue_start`` / ``residue_end`` keep residues start .. end-1 of the
    file, counted in file order (canonical_ordering.py:490-517)."""
    co = co or CanonicalOrdering()
    recs = parse_pdb_residues(pdb_text)
    recs = recs[(residue_start or 0): (len(recs) if residue_end is None else residue_end)]
    if not recs:
        raise ValueError("no residues in the requested range")
    L, A = len(recs), co.max_n_canonical_atoms
    chain_id = numpy.zeros((1, L), dtype=numpy.int32)
    res_types = numpy.full((1, L), -1, dtype=numpy.int32)
    coords = numpy.full((1, L, A, 3), numpy.nan, dtype=numpy.float32)
    chains = {}
    for i, (chain, resseq, _icode, res3, atoms) in enumerate(recs):
        chain_id[0, i] = chains.setdefault(chain, len(chains))
        res3 = "CYS" if res3 == "CYD" else res3
        if res3 not in co.restype_index:
            raise ValueError("unknown residue %s %s %d" % (res3, chain, resseq))
        t = co.restype_index[res3]
        res_types[0, i] = t
        for nm, x in atoms.items():
            if nm not in co.atom_index[t]:
                raise ValueError("atom %s not in %s (%s %d)" % (nm, res3, chain, resseq))
            coords[0, i, co.atom_index[t][nm]] = x
    return (torch.tensor(chain_id), torch.tensor(res_types), torch.tensor(coords))


def pose_stack_from_pdb(pdb_lines_or_fname, device="cuda", *, residue_start=None,
                        residue_end=None, res_not_connected=None, dtype=torch.float32,
                        co: Optional[CanonicalOrdering] = None) -> PoseStack:
    """A PoseStack from the text of a PDB or its file name, missing leaf
    atoms built on ``device`` (tmol/io/__init__.py:24-60): the canonical
    form of residues start .. end-1, then pose_stack_from_canonical_form."""
    import os

    text = pdb_lines_or_fname
    if isinstance(text, (list, tuple)):
        text = "\n".join(text)
    elif "\n" not in text and os.path.exists(text):
        with open(text) as f:
            text = f.read()
    co = co or CanonicalOrdering()
    ch, rt, xyz = canonical_form_from_pdb(text, co, residue_start, residue_end)
    return pose_stack_from_canonical_form(co, ch, rt, xyz, res_not_connected, device=device,
                                          dtype=dtype)


def _disulfides(co, res_types, coords):
    """Pairs (i, j) of CYS residues whose SG atoms are each other's closest
    within 2.5 A (disulfide_search.py)."""
    cys = co.restype_index["CYS"]
    sg = co.atom_index[cys]["SG"]
    inds = [i for i in range(len(res_types)) if res_types[i] == cys and
            not numpy.isnan(coords[i, sg]).any()]
    pairs, used = [], set()
    cand = []
    for a in range(len(inds)):
        for b in range(a + 1, len(inds)):
            d = float(numpy.linalg.norm(coords[inds[a], sg] - coords[inds[b], sg]))
            if d < DISULFIDE_CUTOFF:
                cand.append((d, inds[a], inds[b]))
    for d, i, j in sorted(cand):
        if i not in used and j not in used:
            used.update((i, j))
            pairs.append((i, j))
    return pairs


def block_layout_from_canonical_form(co, chain_id, res_types, coords, res_not_connected=None):
    """Per pose: the residue types (termini, disulfides), the block-ordered
    coordinates (NaN where missing), missing flags, the connectivity and the
    disulfide bonds (pose_stack_construction.py:172-290).  Missing non-leaf
    atoms raise ValueError."""
    db = co.db
    chain_id = numpy.asarray(chain_id)
    res_types = numpy.asarray(res_types)
    coords = numpy.asarray(coords, dtype=numpy.float64)
    rnc = None if res_not_connected is None else numpy.asarray(res_not_connected, dtype=bool)
    out = []
    for p in range(res_types.shape[0]):
        keep = [i for i in range(res_types.shape[1]) if res_types[p, i] >= 0]
        rt_idx = res_types[p, keep]
        ch = chain_id[p, keep]
        xyz = coords[p, keep]
        ss = _disulfides(co, rt_idx, xyz)
        ss_res = {i for pair in ss for i in pair}
        rts, blk_xyz, missing, connected = [], [], [], []
        for k in range(len(keep)):
            name3 = co.restypes[rt_idx[k]]
            first = k == 0 or ch[k - 1] != ch[k]
            last = k == len(keep) - 1 or ch[k + 1] != ch[k]
            nc_lo = rnc is not None and bool(rnc[p, keep[k], 0])
            nc_hi = rnc is not None and bool(rnc[p, keep[k], 1])
            nterm = first and not nc_lo
            cterm = last and not nc_hi
            if name3 == "CYS" and k in ss_res:
                name3 = "CYD"
            if name3 == "HIS":
                t = co.restype_index["HIS"]
                hd1 = not numpy.isnan(xyz[k, co.atom_index[t]["HD1"]]).any() \
                    if "HD1" in co.atom_index[t] else False
                he2 = not numpy.isnan(xyz[k, co.atom_index[t]["HE2"]]).any()
                if hd1 and not he2:
                    raise NotImplementedError(
                        "HIS %d of pose %d is the HD1 tautomer, not in this database" % (k, p))
            rt = db.residue_type(name3, nterm=nterm, cterm=cterm)
            t = co.restype_index["CYS" if name3 == "CYD" else name3]
            x = numpy.full((rt.n_atoms, 3), numpy.nan)
            for a, an in enumerate(rt.atom_names):
                x[a] = xyz[k, co.atom_index[t][an]]
            rts.append(rt)
            blk_xyz.append(x)
            missing.append(numpy.isnan(x).any(axis=1))
            connected.append(k > 0 and not first and not nc_lo and not (
                rnc is not None and bool(rnc[p, keep[k - 1], 1])))
        out.append(dict(rts=rts, coords=blk_xyz, missing=missing, connected=connected,
                        extra_bonds=[(i, rts[i].connection_index("dslf"), j,
                                      rts[j].connection_index("dslf")) for i, j in ss]))
    return out


def pose_stack_from_canonical_form(co, chain_id, res_types, coords, res_not_connected=None,
                                   device="cuda", dtype=torch.float32):
    """A PoseStack from a canonical form, the missing leaf atoms built on
    ``device`` (the reference's pose_stack_from_canonical_form, for the
    canonical amino acids), returned on ``device`` as the reference's is.
    Missing non-leaf atoms raise ValueError; an atom whose ancestors are all
    missing stays NaN."""
    layouts = block_layout_from_canonical_form(co, chain_id, res_types, coords,
                                               res_not_connected)
    ps = pose_stack_from_layouts(layouts, dtype)
    args, tables = leaf_op_args(ps, co.db, layouts)
    for k, L in enumerate(layouts):
        leaf = [tables.is_leaf[ps.packed_block_types.index[rt.name]][: rt.n_atoms]
                for rt in L["rts"]]
        bad = [(b, L["rts"][b].atom_names[a]) for b in range(len(L["rts"]))
               for a in numpy.nonzero(L["missing"][b] & ~leaf[b])[0]]
        if bad:
            raise ValueError("pose %d: missing non-leaf atoms %s" % (k, bad[:8]))
    dev = torch.device(device)
    built = gen_pose_leaf_atoms(*[a.to(dev) if torch.is_tensor(a) else a for a in args])
    ps.coords = built.detach().cpu().to(dtype)
    for k, pose in enumerate(ps.poses):
        off = 0
        for b, rt in enumerate(pose.residue_types):
            pose.coords[b][:] = ps.coords[k, off: off + rt.n_atoms].double().numpy()
            off += rt.n_atoms
    if dev.type == "cpu":
        return ps
    return PoseStack(ps.poses, device=dev, dtype=dtype)


def pose_stack_from_layouts(layouts, dtype=torch.float32) -> PoseStack:
    """The padded PoseStack of per-pose block layouts, missing atoms at 0
    until built."""
    poses = [Pose(L["rts"], [numpy.nan_to_num(x) for x in L["coords"]], L["connected"],
                  extra_bonds=L["extra_bonds"]) for L in layouts]
    return PoseStack(poses, dtype=dtype)


class PaddedLeafLayout:
    """The padded tensors the leaf-atom op reads, for a list of poses,
    without the bond-separation tables of a full PoseStack: coords
    [P, NA, 3], block_coord_offset [P, B], block_type_ind [P, B] (-1 =
    padding) and inter_residue_connections [P, B, C, 2]."""

    def __init__(self, poses, packed_block_types=None, dtype=torch.float32):
        if packed_block_types is None:
            seen = {}
            for p in poses:
                for rt in p.residue_types:
                    seen.setdefault(rt.name, rt)
            from .pose import PackedBlockTypes

            packed_block_types = PackedBlockTypes(list(seen.values()))
        pbt = packed_block_types
        self.packed_block_types = pbt
        P, B = len(poses), max(p.n_blocks for p in poses)
        NA, C = max(p.n_atoms for p in poses), pbt.max_n_conn
        self.n_poses, self.max_n_blocks, self.max_n_atoms = P, B, NA
        coords = numpy.zeros((P, NA, 3), dtype=numpy.float64)
        boff = numpy.zeros((P, B), dtype=numpy.int32)
        btype = numpy.full((P, B), -1, dtype=numpy.int32)
        irc = numpy.full((P, B, C, 2), -1, dtype=numpy.int32)
        for k, pose in enumerate(poses):
            sizes = numpy.array([rt.n_atoms for rt in pose.residue_types])
            offs = numpy.concatenate(([0], numpy.cumsum(sizes)[:-1]))
            boff[k, : len(sizes)] = offs
            btype[k, : len(sizes)] = [pbt.index[rt.name] for rt in pose.residue_types]
            coords[k, : int(sizes.sum())] = numpy.concatenate(pose.coords)
            for (b1, c1), (b2, c2) in pose.bonds():
                irc[k, b1, c1] = (b2, c2)
                irc[k, b2, c2] = (b1, c1)
        self.coords = torch.tensor(coords, dtype=dtype)
        self.block_coord_offset = torch.tensor(boff)
        self.block_type_ind = torch.tensor(btype)
        self.inter_residue_connections = torch.tensor(irc)


def leaf_op_args(ps, db, layouts, tables=None):
    """The 12 arguments of ``tmol_io::gen_pose_leaf_atoms`` for a padded
    PoseStack or PaddedLeafLayout whose blocks miss the atoms flagged in
    ``layouts[p]["missing"]`` (build_missing_leaf_atoms.py:175-212), and the
    block-type tables."""
    pbt = ps.packed_block_types
    if tables is None:
        ats = db.atom_type_set
        tables = LeafTables(pbt, lambda rt: ats.is_hydrogen[rt.atom_types].astype(bool))
    P, B, Amax, NA = ps.n_poses, ps.max_n_blocks, pbt.max_n_atoms, ps.max_n_atoms
    bmiss = numpy.zeros((P, B, Amax), dtype=bool)
    pmiss = numpy.zeros((P, NA), dtype=bool)
    boff = ps.block_coord_offset.numpy()
    for p, L in enumerate(layouts):
        for b, m in enumerate(L["missing"]):
            bmiss[p, b, : len(m)] = m
            pmiss[p, boff[p, b]: boff[p, b] + len(m)] = m
    orig = ps.coords.clone()
    orig[torch.from_numpy(pmiss)] = float("nan")
    i32 = lambda x: torch.as_tensor(numpy.ascontiguousarray(x), dtype=torch.int32)  # noqa
    args = [orig, torch.from_numpy(bmiss), torch.from_numpy(pmiss), i32(boff),
            i32(ps.block_type_ind), i32(ps.inter_residue_connections), i32(tables.n_atoms),
            i32(tables.downstream), i32(tables.anc), torch.tensor(tables.geom),
            i32(tables.anc_backup), torch.tensor(tables.geom_backup)]
    return args, tables
