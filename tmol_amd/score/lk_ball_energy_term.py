"""lk_ball energy term: hbond atom typing, the block-type tables of the
reference's lk_ball ops, and the water-building recipes of the HIP path.

Mirrors, for the four lk_ball score types (lk_ball_iso, lk_ball, lk_bridge,
lk_bridge_uncpl):

* tmol/score/lk_ball/lk_ball_energy_term.py:22-58 -- global parameter rows
  ``[lj_hbond_dis, lj_hbond_OH_donor_dis, lj_hbond_hdis, lkb_water_dist,
  max_dis]`` and ``[lkb_water_dist, lkb_water_angle_sp2, _sp3, _ring]`` as
  float32 (angles in radians), the torsion lists ``lkb_water_tors_*``;
* lk_ball_energy_term.py:69-198 -- per tile of 32 atoms, the "polar" atoms
  (donor heavy atoms and acceptors, which carry waters) listed before the
  other heavy atoms ("occluders"), with their 9 parameters
  ``lj_radius, lk_dgfree, lk_lambda, lk_volume, is_donor, is_hydroxyl,
  is_polarh, is_acceptor, is_carbon_lk``;
* tmol/score/hbond/hbond_dependent_term.py:23-53, 117-264 -- donors and
  acceptors by atom type (hbond.yaml's donor/acceptor atom types), the polar
  hydrogens of each donor in bond order, acceptor hybridization and attached
  hydrogen count, tiled;
* tmol/chemical/all_bonds.py:8-79 and bond_dependent_term.py:30-88 -- the
  all-bonds table: each atom's intra-block bonds in sorted order, then its
  inter-block connections.

The HIP path builds waters from per-block-type recipes: one row per water
slot, naming the atoms its frame is built from.  The acceptor bases follow
the reference's walk over the all-bonds table
(tmol/score/hbond/identification.hh:26-167); block types whose acceptor
bases would lie across an inter-block bond are refused (no canonical amino
acid has one).
"""

import math

import numpy
import torch

from ..chemical import default_chemical_database
from .ljlk_energy_term import LJLKEnergyTerm
from .score_types import ScoreType

TILE = 32
MAX_N_WATER = 4
HYBRIDIZATION = {None: 0, "sp2": 1, "sp3": 2, "ring": 3}
# water recipe kinds
W_DONOR, W_SP2, W_SP3, W_RING = 0, 1, 2, 3
RECIPE_WIDTH = 8  # atom, slot, kind, a1, a2, a3, torsion index, 0


def all_bonds_of(rt):
    """(all_bonds [n, 3], atom_all_bond_ranges [n_atoms, 2]) of a residue
    type: per atom, its intra-block bonds sorted by partner, then one row per
    inter-block connection on it (all_bonds.py:8-79)."""
    pairs = set()
    for a, b in rt.bonds:
        pairs.add((int(a), int(b)))
        pairs.add((int(b), int(a)))
    intra = sorted(pairs)
    n = rt.n_atoms
    rows = [[] for _ in range(n)]
    for a, b in intra:
        rows[a].append((a, b, -1))
    for c, ca in enumerate(rt.conn_atom):
        rows[int(ca)].append((int(ca), -1, c))
    bonds = numpy.zeros((len(intra) + len(rt.conn_atom), 3), dtype=numpy.int32)
    ranges = numpy.zeros((n, 2), dtype=numpy.int32)
    k = 0
    for a in range(n):
        ranges[a, 0] = k
        for r in rows[a]:
            bonds[k] = r
            k += 1
        ranges[a, 1] = k
    return bonds, ranges


def _tiles(inds, n_atoms):
    """Atom indices grouped per tile of 32 (left-justified, original
    indices), over the tiles of n_atoms atoms."""
    n_tiles = (n_atoms - 1) // TILE + 1
    out = [[] for _ in range(n_tiles)]
    for i in sorted(int(x) for x in inds):
        out[i // TILE].append(i)
    return out


class LKBallBlockType:
    """Per-block-type lk_ball / hbond tables of one residue type (numpy)."""

    def __init__(self, rt, term):
        n = rt.n_atoms
        ats = term.db.atom_type_set
        names = [ats.names[t] if t < len(ats.names) else None for t in rt.atom_types]
        self.n_atoms = n
        self.n_tiles = (n - 1) // TILE + 1
        self.is_hydrogen = ats.is_hydrogen[rt.atom_types].astype(numpy.int32)
        is_don = numpy.array([nm in term.donor_types for nm in names], dtype=bool)
        is_acc = numpy.array([nm in term.acceptor_types for nm in names], dtype=bool)
        hyb = numpy.array([term.hybridization.get(nm, 0) for nm in names], dtype=numpy.int32)
        self.all_bonds, self.atom_all_bond_ranges = all_bonds_of(rt)

        # polar hydrogens of each donor, in ascending atom order
        # (hbond_dependent_term.py:23-53 over the sorted intra-block bonds)
        nbrs = [[] for _ in range(n)]
        for a, b, c in self.all_bonds:
            if b >= 0:
                nbrs[a].append(int(b))
        D_idx = numpy.nonzero(is_don)[0]
        A_idx = numpy.nonzero(is_acc)[0]
        donH, D_for_H, which = [], [], []
        n_H = numpy.zeros(n, dtype=numpy.int32)
        for d in D_idx:
            k = 0
            for b in nbrs[d]:
                if self.is_hydrogen[b]:
                    donH.append(b)
                    D_for_H.append(int(d))
                    which.append(k)
                    k += 1
            n_H[d] = k
        order = numpy.argsort(numpy.array(donH, dtype=numpy.int64), kind="stable")
        donH = [donH[i] for i in order]
        D_for_H = [D_for_H[i] for i in order]
        which = [which[i] for i in order]
        self.donH, self.D_for_H, self.which_H = donH, D_for_H, which
        self.don_hvy, self.acc = [int(x) for x in D_idx], [int(x) for x in A_idx]
        self.acc_hyb = {int(a): int(hyb[a]) for a in A_idx}
        self.n_H = n_H

        nt = self.n_tiles
        self.tile_n_donH = numpy.zeros(nt, dtype=numpy.int32)
        self.tile_n_acc = numpy.zeros(nt, dtype=numpy.int32)
        self.tile_donH_inds = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_donH_hvy_inds = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_which_donH = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_acc_inds = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_acc_hyb = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_acc_nH = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        pos = {h: i for i, h in enumerate(donH)}
        for t, tl in enumerate(_tiles(donH, n)):
            self.tile_n_donH[t] = len(tl)
            for k, h in enumerate(tl):
                self.tile_donH_inds[t, k] = h % TILE
                self.tile_donH_hvy_inds[t, k] = D_for_H[pos[h]]
                self.tile_which_donH[t, k] = which[pos[h]]
        for t, tl in enumerate(_tiles(A_idx, n)):
            self.tile_n_acc[t] = len(tl)
            for k, a in enumerate(tl):
                self.tile_acc_inds[t, k] = a % TILE
                self.tile_acc_hyb[t, k] = hyb[a]
                self.tile_acc_nH[t, k] = n_H[a]

        # lk_ball: polar (donor heavy or acceptor) then heavy apolar, per tile
        # (lk_ball_energy_term.py:69-198)
        is_polar = is_don | is_acc
        heavy = self.is_hydrogen == 0
        self.is_polar = is_polar
        self.is_occluder = heavy | is_polar
        pol_t = _tiles(numpy.nonzero(is_polar)[0], n)
        apo_t = _tiles(numpy.nonzero(heavy & ~is_polar)[0], n)
        self.tile_n_polar = numpy.zeros(nt, dtype=numpy.int32)
        self.tile_n_occ = numpy.zeros(nt, dtype=numpy.int32)
        self.tile_pol_occ_inds = numpy.full((nt, TILE), -1, dtype=numpy.int32)
        self.tile_params = numpy.zeros((nt, TILE, 9), dtype=numpy.float32)
        tp = term.lkb_type_params.numpy()
        for t in range(nt):
            lst = pol_t[t] + apo_t[t]
            self.tile_n_polar[t] = len(pol_t[t])
            self.tile_n_occ[t] = len(lst)
            for k, a in enumerate(lst):
                self.tile_pol_occ_inds[t, k] = a % TILE
                self.tile_params[t, k] = tp[rt.atom_types[a]]
        # compact occluder list of the HIP path: polars first, then the other
        # heavy atoms, each group in atom order
        pol = [int(a) for a in numpy.nonzero(is_polar)[0]]
        apo = [int(a) for a in numpy.nonzero(heavy & ~is_polar)[0]]
        self.occ_list = pol + apo
        self.n_polar = len(pol)
        self.recipes = self._recipes(rt, term)

    def _recipes(self, rt, term):
        return water_recipes(
            self.donH, self.D_for_H, self.which_H, self.acc, self.acc_hyb, self.n_H,
            self.all_bonds, self.atom_all_bond_ranges, self.is_hydrogen,
            (len(term.tors_sp2), len(term.tors_sp3), len(term.tors_ring)), rt.name)


def water_recipes(donH, D_for_H, which_H, acc, acc_hyb, n_H, all_bonds, ranges, is_h, n_tors,
                  name="block type"):
    """Rows [atom, slot, kind, a1, a2, a3, torsion index, 0], one per water
    of a block type: donor waters (a1 = D, a2 = H) in slot which_H; acceptor
    waters (a1 = A, a2 = B, a3 = B0) in slots n_H[A] + torsion index, the
    bases by the reference's walk over the all-bonds table
    (tmol/score/hbond/identification.hh:26-167).  Acceptors with no water
    frame (no heavy neighbor; sp3 without a second distinct base) get no
    rows.  A walk that would read an inter-block partner raises
    NotImplementedError."""
    rows = []
    for h, d, w in zip(donH, D_for_H, which_H):
        rows.append([int(d), int(w), W_DONOR, int(d), int(h), -1, 0, 0])

    def bound(a):
        return [int(all_bonds[k, 1]) for k in range(ranges[a, 0], ranges[a, 1])]

    def fail(a):
        raise NotImplementedError(
            "lk_ball: acceptor %d of %s has its base across an inter-block bond" % (a, name))

    for a in acc:
        hyb = int(acc_hyb[a])
        if hyb not in (1, 2, 3):
            continue
        nb = bound(a)
        B = B0 = -1
        # B: first heavy neighbor (identification.hh:40-45, 85-90, 125-130)
        for x in nb:
            if x < 0:
                fail(a)
            if not is_h[x]:
                B = x
                break
        if B < 0:
            continue
        if hyb == 1:  # sp2: first neighbor of B other than A
            for x in bound(B):
                if x != a:
                    if x < 0:
                        fail(a)
                    B0 = x
                    break
            if B0 < 0:
                continue
        elif hyb == 2:  # sp3: a bonded hydrogen, else another neighbor, else B
            for x in nb:
                if x >= 0 and x != B and is_h[x]:
                    B0 = x
                    break
            if B0 < 0:
                for x in nb:
                    if x != B:
                        if x < 0:
                            fail(a)
                        B0 = x
                        break
            if B0 < 0 or B0 == B:
                continue  # no distinct second base: no waters
        else:  # ring: another heavy neighbor
            for x in nb:
                if x != B and (x < 0 or not is_h[x]):
                    if x < 0:
                        fail(a)
                    B0 = x
                    break
            if B0 < 0:
                continue
        kind = {1: W_SP2, 2: W_SP3, 3: W_RING}[hyb]
        for w in range(n_tors[hyb - 1]):
            slot = w + int(n_H[a])
            if slot >= MAX_N_WATER:
                raise ValueError("lk_ball: more than 4 waters on atom %d of %s" % (a, name))
            rows.append([int(a), slot, kind, int(a), B, B0, w, 0])
    return numpy.array(rows, dtype=numpy.int32).reshape(-1, RECIPE_WIDTH)


class LKBallEnergyTerm:
    tile_size = TILE

    def __init__(self, device="cpu", chemical_db=None):
        self.device = torch.device(device)
        self.db = chemical_db or default_chemical_database()
        raw = self.db.raw
        self.donor_types = {d for d, _ in raw["hbond"]["donor_atom_types"]}
        self.acceptor_types = {a for a, _ in raw["hbond"]["acceptor_atom_types"]}
        self.hybridization = {
            a["name"]: HYBRIDIZATION[a["acceptor_hybridization"]] for a in raw["atom_types"]
        }
        ljlk = LJLKEnergyTerm(chemical_db=self.db)
        tp = ljlk._type_params  # lj_radius, lj_wdepth, lk_dgfree, ... is_carbon_lk
        self.lkb_type_params = tp[:, [0, 2, 3, 4, 5, 6, 7, 8, 9]].contiguous()
        g = self.db.ljlk["global"]
        lw = self.db.ljlk["lkb_water"]
        self.global_values = [g["lj_hbond_dis"], g["lj_hbond_OH_donor_dis"], g["lj_hbond_hdis"],
                              g["lkb_water_dist"], g["max_dis"]]
        rad = math.radians
        self.watergen_values = [g["lkb_water_dist"], rad(lw["lkb_water_angle_sp2"]),
                                rad(lw["lkb_water_angle_sp3"]), rad(lw["lkb_water_angle_ring"])]
        self.tors_sp2 = [rad(x) for x in lw["lkb_water_tors_sp2"]]
        self.tors_sp3 = [rad(x) for x in lw["lkb_water_tors_sp3"]]
        self.tors_ring = [rad(x) for x in lw["lkb_water_tors_ring"]]
        self._bt = {}

    @classmethod
    def class_name(cls):
        return "LKBall"

    @classmethod
    def score_types(cls):
        return [ScoreType.lk_ball_iso, ScoreType.lk_ball, ScoreType.lk_bridge,
                ScoreType.lk_bridge_uncpl]

    def n_bodies(self):
        return 2

    # float32 values, as the reference stacks them (lk_ball_energy_term.py:32-58)
    @property
    def global_params(self):
        return torch.tensor([self.global_values], dtype=torch.float32, device=self.device)

    @property
    def watergen_params(self):
        return torch.tensor([self.watergen_values], dtype=torch.float32, device=self.device)

    def water_tors(self):
        f = lambda v: torch.tensor(v, dtype=torch.float32, device=self.device)  # noqa: E731
        return f(self.tors_sp2), f(self.tors_sp3), f(self.tors_ring)

    @property
    def max_dis(self) -> float:
        return float(numpy.float32(self.global_values[4]))

    def block_type(self, rt) -> LKBallBlockType:
        if rt.name not in self._bt:
            self._bt[rt.name] = LKBallBlockType(rt, self)
        return self._bt[rt.name]

    def packed_tables(self, pbt):
        """The reference op's per-block-type tensors for a PackedBlockTypes
        (hbond_dependent_term.py:266-340, lk_ball_energy_term.py:199-233,
        bond_dependent_term.py:30-88), cached on the pbt."""
        cache = getattr(pbt, "_lk_ball_tables", None)
        if cache is not None:
            return cache
        bts = [self.block_type(rt) for rt in pbt.active_block_types]
        T, Amax = pbt.n_types, pbt.max_n_atoms
        nt = (Amax - 1) // TILE + 1
        nb = max(b.all_bonds.shape[0] for b in bts)
        i32 = numpy.int32
        out = {
            "n_all_bonds": numpy.array([b.all_bonds.shape[0] for b in bts], dtype=i32),
            "all_bonds": numpy.full((T, nb, 3), -1, dtype=i32),
            "atom_all_bond_ranges": numpy.full((T, Amax, 2), -1, dtype=i32),
            "tile_n_donH": numpy.zeros((T, nt), dtype=i32),
            "tile_n_acc": numpy.zeros((T, nt), dtype=i32),
            "tile_donH_inds": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_donH_hvy_inds": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_which_donH_of_donH_hvy": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_acc_inds": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_acceptor_hybridization": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_acceptor_n_attached_H": numpy.full((T, nt, TILE), -1, dtype=i32),
            "is_hydrogen": numpy.zeros((T, Amax), dtype=i32),
            "tile_n_polar_atoms": numpy.zeros((T, nt), dtype=i32),
            "tile_n_occluder_atoms": numpy.zeros((T, nt), dtype=i32),
            "tile_pol_occ_inds": numpy.full((T, nt, TILE), -1, dtype=i32),
            "tile_lk_ball_params": numpy.zeros((T, nt, TILE, 9), dtype=numpy.float32),
            "path_distance": numpy.full((T, Amax, Amax), 6, dtype=i32),
        }
        for i, (b, rt) in enumerate(zip(bts, pbt.active_block_types)):
            k, n, t = b.all_bonds.shape[0], b.n_atoms, b.n_tiles
            out["all_bonds"][i, :k] = b.all_bonds
            out["atom_all_bond_ranges"][i, :n] = b.atom_all_bond_ranges
            out["tile_n_donH"][i, :t] = b.tile_n_donH
            out["tile_n_acc"][i, :t] = b.tile_n_acc
            out["tile_donH_inds"][i, :t] = b.tile_donH_inds
            out["tile_donH_hvy_inds"][i, :t] = b.tile_donH_hvy_inds
            out["tile_which_donH_of_donH_hvy"][i, :t] = b.tile_which_donH
            out["tile_acc_inds"][i, :t] = b.tile_acc_inds
            out["tile_acceptor_hybridization"][i, :t] = b.tile_acc_hyb
            out["tile_acceptor_n_attached_H"][i, :t] = b.tile_acc_nH
            out["is_hydrogen"][i, :n] = b.is_hydrogen
            out["tile_n_polar_atoms"][i, :t] = b.tile_n_polar
            out["tile_n_occluder_atoms"][i, :t] = b.tile_n_occ
            out["tile_pol_occ_inds"][i, :t] = b.tile_pol_occ_inds
            out["tile_lk_ball_params"][i, :t] = b.tile_params
            # unmodified path distances (pbt.bond_separation of the reference)
            out["path_distance"][i, :n, :n] = rt.path_distance
        tens = {k: torch.tensor(v, device=pbt.device) for k, v in out.items()}
        tens["bts"] = bts
        pbt._lk_ball_tables = tens
        return tens


def watergen_op_args(pose_stack, term: LKBallEnergyTerm, coords=None):
    """The 33 arguments of ``tmol_lk_ball::gen_pose_waters`` for a padded
    stack (lk_ball_energy_term.py:244-268)."""
    ps = pose_stack
    if coords is None:
        coords = ps.coords
    pbt = ps.packed_block_types
    t = term.packed_tables(pbt)
    tors = term.water_tors()
    return [
        coords.flatten(0, -2), ps.rot_coord_offset, ps.pose_ind_for_atom,
        ps.first_rot_for_block, ps.block_type_ind, ps.block_ind_for_rot,
        ps.pose_ind_for_rot, ps.block_type_ind_for_rot, ps.n_rots_for_pose,
        ps.rot_offset_for_pose, ps.n_rots_for_block, ps.rot_offset_for_block,
        ps.max_n_rots_per_pose,
        ps.inter_residue_connections, pbt.n_atoms, pbt.n_conn, pbt.conn_atom,
        t["n_all_bonds"], t["all_bonds"], t["atom_all_bond_ranges"],
        t["tile_n_donH"], t["tile_n_acc"], t["tile_donH_inds"], t["tile_donH_hvy_inds"],
        t["tile_which_donH_of_donH_hvy"], t["tile_acc_inds"],
        t["tile_acceptor_hybridization"], t["tile_acceptor_n_attached_H"], t["is_hydrogen"],
        term.watergen_params.to(pbt.device), *[x.to(pbt.device) for x in tors],
    ]


def score_op_args(pose_stack, term: LKBallEnergyTerm, waters, coords=None, block_pair=False):
    """The 28 arguments of ``tmol_lk_ball::lk_ball_pose_score`` for a padded
    stack (lk_ball_energy_term.py:272-296)."""
    ps = pose_stack
    if coords is None:
        coords = ps.coords
    pbt = ps.packed_block_types
    t = term.packed_tables(pbt)
    return [
        coords.flatten(0, -2), ps.rot_coord_offset, ps.pose_ind_for_atom,
        ps.first_rot_for_block, ps.block_type_ind, ps.block_ind_for_rot,
        ps.pose_ind_for_rot, ps.block_type_ind_for_rot, ps.n_rots_for_pose,
        ps.rot_offset_for_pose, ps.n_rots_for_block, ps.rot_offset_for_block,
        ps.max_n_rots_per_pose,
        ps.inter_residue_connections, ps.min_block_bondsep, ps.inter_block_bondsep,
        pbt.n_atoms, pbt.n_conn, pbt.conn_atom,
        t["tile_n_polar_atoms"], t["tile_n_occluder_atoms"], t["tile_pol_occ_inds"],
        t["tile_lk_ball_params"], t["path_distance"],
        term.global_params.to(pbt.device), term.max_dis, waters, bool(block_pair),
    ]
