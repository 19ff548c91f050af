"""hbond energy term: the hbond potential's parameter tables, the block-type
tables of the reference's hbond ops, and the per-block-type donor /
acceptor lists of the HIP path.

Mirrors, for the one hbond score type:

* tmol/score/hbond/params.py:73-286 -- the donor x acceptor pair table
  ``[acceptor_hybridization, acceptor_weight, donor_weight]`` (float32),
  the polynomial table (float64; per pair the AHdist, cosBAH and cosAHD
  polynomials, each as coefficients c10..c0, a zero pad, [xmin, xmax],
  [min_val, max_val]) and the global row ``[hb_sp2_range_span,
  hb_sp2_BAH180_rise, hb_sp2_outer_width, hb_sp3_softmax_fade,
  threshold_distance, max_ha_dis]`` (float32; max_ha_dis = the largest
  xmax of the A-H distance polynomials);
* tmol/score/hbond/hbond_dependent_term.py:56-358 -- donor and acceptor
  types by atom type, the polar hydrogens of each donor in bond order, per
  tile of 32 atoms;
* tmol/score/hbond/hbond_energy_term.py:68-132 -- the argument lists of
  gen_hbond_bases and hbond_pose_scores.

The HIP path reads, per block type, the donor hydrogens (with their heavy
atom and donor type) and the acceptors (with bases B, B0 and acceptor
type).  The bases follow the reference's walk over the all-bonds table
(tmol/score/hbond/identification.hh:26-198); block types whose bases would
lie across an inter-block bond are refused (no canonical amino acid has
one).
"""

import numpy
import torch

from ..chemical import default_chemical_database
from .lk_ball_energy_term import HYBRIDIZATION, TILE, LKBallEnergyTerm

POLY_WIDTH = 16  # c10..c0, pad, xmin, xmax, min_val, max_val
DON_WIDTH = 4    # H, D, donor type, 0
ACC_WIDTH = 8    # A, B, B0, acceptor type, hybridization, 0, 0, 0
NEIGHBOR_REACH = 5.5  # hbond_pose_score.impl.hh:600 (detect_block_neighbors)


def _poly_row(p):
    """One bounded polynomial as the reference packs it (params.py:124-131,
    243-271): coefficients from c10 down to c0, a zero pad, range, bound."""
    return list(reversed(p["c"])) + [0.0, p["xmin"], p["xmax"], p["min_val"], p["max_val"]]


def acceptor_bases(A, hyb, all_bonds, ranges, is_h, name="block type"):
    """(B, B0) of an acceptor by the reference's walk
    (identification.hh:26-167): sp2 -- B the first heavy neighbor, B0 the
    first other neighbor of B; sp3 -- B0 a bonded hydrogen, else another
    neighbor, else B; ring -- B0 another heavy neighbor.  (-1, -1) when an
    sp2 or ring acceptor lacks either base, or for no hybridization.  A walk
    that would cross an inter-block bond raises NotImplementedError."""
    def bound(a):
        return [int(all_bonds[k, 1]) for k in range(ranges[a, 0], ranges[a, 1])]

    def cross():
        raise NotImplementedError(
            "hbond: acceptor %d of %s has a base across an inter-block bond" % (A, name))

    if hyb not in (1, 2, 3):
        return -1, -1
    nb = bound(A)
    B = -1
    for x in nb:
        if x < 0:
            cross()
        if not is_h[x]:
            B = x
            break
    if hyb == 1:
        if B < 0:
            return -1, -1
        for x in bound(B):
            if x != A:
                if x < 0:
                    cross()
                return B, x
        return -1, -1
    if hyb == 2:
        for x in nb:
            if x >= 0 and x != B and is_h[x]:
                return B, x
        for x in nb:
            if x != B:
                if x < 0:
                    cross()
                return B, x
        return B, B
    if B < 0:
        return -1, -1
    for x in nb:
        if x != B and (x < 0 or not is_h[x]):
            if x < 0:
                cross()
            return B, x
    return -1, -1


class HBondBlockType:
    """Per-block-type hbond tables of one residue type (numpy): the
    reference's tiled donor / acceptor types, and the HIP path's lists."""

    def __init__(self, rt, term):
        lkb = term.lkb.block_type(rt)
        ats = term.db.atom_type_set
        names = [ats.names[t] if t < len(ats.names) else None for t in rt.atom_types]
        self.lkb = lkb
        self.n_atoms, self.n_tiles = lkb.n_atoms, lkb.n_tiles
        nt = self.n_tiles
        don_t = numpy.array([term.don_type_index.get(term.atom_don_type.get(nm), -1)
                             for nm in names], dtype=numpy.int32)
        acc_t = numpy.array([term.acc_type_index.get(term.atom_acc_type.get(nm), -1)
                             for nm in names], dtype=numpy.int32)
        self.tile_donor_type = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_acceptor_type = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        for t in range(nt):
            for k in range(lkb.tile_n_donH[t]):
                h = t * TILE + lkb.tile_donH_inds[t, k]
                self.tile_donor_type[t, k] = don_t[lkb.D_for_H[lkb.donH.index(h)]]
            for k in range(lkb.tile_n_acc[t]):
                a = t * TILE + lkb.tile_acc_inds[t, k]
                self.tile_acceptor_type[t, k] = acc_t[a]
        # HIP lists, in tile order (one tile: ascending atom order)
        self.don = numpy.array(
            [[h, d, don_t[d], 0] for h, d in zip(lkb.donH, lkb.D_for_H)],
            dtype=numpy.int32).reshape(-1, DON_WIDTH)
        rows = []
        for a in lkb.acc:
            hyb = int(lkb.acc_hyb[a])
            B, B0 = acceptor_bases(a, hyb, lkb.all_bonds, lkb.atom_all_bond_ranges,
                                   lkb.is_hydrogen, rt.name)
            rows.append([a, B, B0, acc_t[a], hyb, 0, 0, 0])
        self.acc = numpy.array(rows, dtype=numpy.int32).reshape(-1, ACC_WIDTH)


class HBondEnergyTerm:
    tile_size = TILE

    def __init__(self, device="cpu", chemical_db=None):
        self.device = torch.device(device)
        self.db = chemical_db or default_chemical_database()
        self.lkb = LKBallEnergyTerm(chemical_db=self.db)
        raw = self.db.raw
        hb = raw["hbond"]
        self.don_type_names = [n for n, _ in hb["donor_type_params"]]
        self.acc_type_names = [n for n, _ in hb["acceptor_type_params"]]
        self.don_type_index = {n: i for i, n in enumerate(self.don_type_names)}
        self.acc_type_index = {n: i for i, n in enumerate(self.acc_type_names)}
        self.atom_don_type = {d: t for d, t in hb["donor_atom_types"]}
        self.atom_acc_type = {a: t for a, t in hb["acceptor_atom_types"]}
        atom_hyb = {a["name"]: HYBRIDIZATION[a["acceptor_hybridization"]]
                    for a in raw["atom_types"]}
        nD, nA = len(self.don_type_names), len(self.acc_type_names)
        # pair table (params.py:73-146, 216-229): NaN weights where unset,
        # hybridization 0 before the acceptor columns are filled
        pp = numpy.full((nD, nA, 3), numpy.nan, dtype=numpy.float32)
        pp[:, :, 0] = 0
        for name, w in hb["donor_type_params"]:
            pp[self.don_type_index[name], :, 2] = w
        acc_hyb = {t: atom_hyb[a] for a, t in hb["acceptor_atom_types"]}
        for name, w in hb["acceptor_type_params"]:
            j = self.acc_type_index[name]
            pp[:, j, 1] = w
            pp[:, j, 0] = acc_hyb[name]
        polys = {p["name"]: p for p in hb["polynomial_parameters"]}
        pt = numpy.full((nD, nA, 3 * POLY_WIDTH), numpy.nan, dtype=numpy.float64)
        for k in range(3):
            pt[:, :, k * POLY_WIDTH + 11] = 0.0  # the pad between coefficients and range
        for don, acc, ahd, bah, ahd_ang in hb["pair_parameters"]:
            i, j = self.don_type_index[don], self.acc_type_index[acc]
            for k, nm in enumerate((ahd, bah, ahd_ang)):
                pt[i, j, k * POLY_WIDTH:(k + 1) * POLY_WIDTH] = _poly_row(polys[nm])
        self._pair_params = pp
        self._pair_polys = pt
        g = hb["global_parameters"]
        max_ahdis = max(p["xmax"] for p in hb["polynomial_parameters"]
                        if p["dimension"] == "hbgd_AHdist")
        self._global = numpy.array(
            [[g["hb_sp2_range_span"], g["hb_sp2_BAH180_rise"], g["hb_sp2_outer_width"],
              g["hb_sp3_softmax_fade"], g["threshold_distance"], max_ahdis]], dtype=numpy.float32)
        self._bt = {}

    @classmethod
    def class_name(cls):
        return "HBond"

    @classmethod
    def score_types(cls):
        from .score_types import ScoreType
        return [ScoreType.hbond]

    def n_bodies(self):
        return 2

    @property
    def pair_param_table(self):
        return torch.tensor(self._pair_params, device=self.device)

    @property
    def pair_poly_table(self):
        return torch.tensor(self._pair_polys, device=self.device)

    @property
    def global_param_table(self):
        return torch.tensor(self._global, device=self.device)

    @property
    def max_ha_dis(self) -> float:
        return float(self._global[0, 5])

    def block_type(self, rt) -> HBondBlockType:
        if rt.name not in self._bt:
            self._bt[rt.name] = HBondBlockType(rt, self)
        return self._bt[rt.name]

    def packed_tables(self, pbt):
        """The reference ops' per-block-type tensors for a PackedBlockTypes
        (hbond_dependent_term.py:266-358 plus the all-bonds table and the
        path distances), cached on the pbt."""
        cache = getattr(pbt, "_hbond_tables", None)
        if cache is not None:
            return cache
        base = self.lkb.packed_tables(pbt)
        bts = [self.block_type(rt) for rt in pbt.active_block_types]
        T = pbt.n_types
        nt = base["tile_n_donH"].shape[1]
        dt = numpy.full((T, nt, TILE), -1, dtype=numpy.int32)
        at = numpy.full((T, nt, TILE), -1, dtype=numpy.int32)
        for i, b in enumerate(bts):
            dt[i, :b.n_tiles] = b.tile_donor_type
            at[i, :b.n_tiles] = b.tile_acceptor_type
        out = {k: base[k] for k in (
            "n_all_bonds", "all_bonds", "atom_all_bond_ranges", "tile_n_donH", "tile_n_acc",
            "tile_donH_inds", "tile_acc_inds", "tile_acceptor_hybridization", "is_hydrogen",
            "path_distance")}
        out["tile_donorH_type"] = torch.tensor(dt, device=pbt.device)
        out["tile_acceptor_type"] = torch.tensor(at, device=pbt.device)
        out["bts"] = bts
        pbt._hbond_tables = out
        return out


def bases_op_args(pose_stack, term: HBondEnergyTerm, coords=None):
    """The 20 arguments of ``tmol_hbond::gen_hbond_bases`` for a padded
    stack (hbond_energy_term.py:76-97)."""
    ps = pose_stack
    if coords is None:
        coords = ps.coords
    pbt = ps.packed_block_types
    t = term.packed_tables(pbt)
    return [
        coords.flatten(0, -2), ps.rot_coord_offset, ps.first_rot_for_block, ps.block_type_ind,
        ps.block_ind_for_rot, ps.pose_ind_for_rot, ps.block_type_ind_for_rot,
        ps.inter_residue_connections, pbt.n_atoms, pbt.n_conn, pbt.conn_atom,
        t["n_all_bonds"], t["all_bonds"], t["atom_all_bond_ranges"], t["tile_n_donH"],
        t["tile_n_acc"], t["tile_donH_inds"], t["tile_acc_inds"],
        t["tile_acceptor_hybridization"], t["is_hydrogen"],
    ]


def score_op_args(pose_stack, term: HBondEnergyTerm, derived_coords, derived_atom_inds,
                  coords=None, block_pair=False):
    """The 37 arguments of ``tmol_hbond::hbond_pose_scores`` for a padded
    stack (hbond_energy_term.py:98-132)."""
    ps = pose_stack
    if coords is None:
        coords = ps.coords
    pbt = ps.packed_block_types
    t = term.packed_tables(pbt)
    dev = pbt.device
    return [
        coords.flatten(0, -2), ps.rot_coord_offset, ps.pose_ind_for_atom,
        ps.first_rot_for_block, ps.block_type_ind, ps.block_ind_for_rot,
        ps.pose_ind_for_rot, ps.block_type_ind_for_rot, ps.n_rots_for_pose,
        ps.rot_offset_for_pose, ps.n_rots_for_block, ps.rot_offset_for_block,
        ps.max_n_rots_per_pose,
        ps.inter_residue_connections, ps.min_block_bondsep, ps.inter_block_bondsep,
        pbt.n_atoms, pbt.n_conn, pbt.conn_atom, t["n_all_bonds"], t["all_bonds"],
        t["atom_all_bond_ranges"], t["path_distance"],
        t["tile_n_donH"], t["tile_n_acc"], t["tile_donH_inds"], t["tile_acc_inds"],
        t["tile_donorH_type"], t["tile_acceptor_type"], t["tile_acceptor_hybridization"],
        t["is_hydrogen"],
        # in the coordinates' precision, the polynomials in float64
        # (hbond_energy_term.py:28-44)
        term.pair_param_table.to(dev, coords.dtype), term.pair_poly_table.to(dev),
        term.global_param_table.to(dev, coords.dtype),
        derived_coords, derived_atom_inds, bool(block_pair),
    ]
