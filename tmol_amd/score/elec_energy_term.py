"""Elec energy term: parameter resolution and the op's term attributes.

Mirrors tmol/score/elec/elec_energy_term.py:13-186 and
tmol/score/elec/params.py:28-175:

* partial charges per block-type atom, looked up by residue name and patch
  variant ("ALA:nterm" tries the "nterm" entry, then the unpatched one;
  params.py:44-69); residues absent from the table get zero charges;
* count-pair representatives (params.py:71-108): each "outer" atom listed for
  an "inner" atom makes the outer atom the inner atom's representative (the
  last listed entry wins);
* ``elec_inter_repr_path_distance[t, i, j] = path_distance[i, rep(j)]`` and
  ``elec_intra_repr_path_distance[t, i, j] = path_distance[rep(i), rep(j)]``,
  with 3 and 4 counted as 5 for non-polymer types (elec_energy_term.py:62-91);
* ``global_params [1, 5] = D, D0, S, min_dis, max_dis``, and ``max_dis`` as a
  host scalar for the neighbor pass (elec_energy_term.py:158-186).
"""

import numpy
import torch

from ..chemical import default_chemical_database
from .score_types import ScoreType


class ElecParamResolver:
    """Charges and count-pair representatives keyed [res][atom][variant]
    (params.py:122-175)."""

    def __init__(self, elec_db):
        g = elec_db["global"]
        self.global_params = {k: float(v) for k, v in g.items()}
        self.partial_charges = {}
        for res, var, atom, q in elec_db["atom_charge_parameters"]:
            self.partial_charges.setdefault(res, {}).setdefault(atom, {})[var] = float(q)
        self.cp_reps = {}
        for res, var, inner, outer in elec_db["atom_cp_reps_parameters"]:
            self.cp_reps.setdefault(res, {}).setdefault(outer, {})[var] = inner

    @staticmethod
    def _res_vars(block_type):
        res, *variants = block_type.name.split(":")
        return res, variants + [""]

    def get_partial_charges_for_block(self, block_type) -> numpy.ndarray:
        res, variants = self._res_vars(block_type)
        out = numpy.zeros(block_type.n_atoms, dtype=numpy.float32)
        if res not in self.partial_charges:
            return out
        table = self.partial_charges[res]
        for i, name in enumerate(block_type.atom_names):
            if name not in table:
                raise KeyError("Elec charge for atom %s,%s not found" % (block_type.name, name))
            for v in variants:
                if v in table[name]:
                    out[i] = table[name][v]
                    break
        return out

    def get_bonded_path_length_mapping_for_block(self, block_type) -> numpy.ndarray:
        rep = numpy.arange(block_type.n_atoms, dtype=numpy.int32)
        res, variants = self._res_vars(block_type)
        if res not in self.cp_reps:
            return rep
        table = self.cp_reps[res]
        for outer in block_type.atom_index:
            if outer not in table:
                continue
            inner = next((table[outer][v] for v in variants if v in table[outer]), None)
            if inner is None:
                continue
            if inner not in block_type.atom_index:
                raise KeyError("Invalid elec cp mapping: %s %s->%s" % (res, outer, inner))
            rep[block_type.atom_index[inner]] = block_type.atom_index[outer]
        return rep


class ElecEnergyTerm:
    def __init__(self, device="cpu", chemical_db=None):
        self.device = torch.device(device)
        self.db = chemical_db or default_chemical_database()
        self.param_resolver = ElecParamResolver(self.db.raw["elec"])
        self._tables = {}

    @classmethod
    def class_name(cls):
        return "Elec"

    @classmethod
    def score_types(cls):
        return [ScoreType.fa_elec]

    def n_bodies(self):
        return 2

    def set_options(self, options: dict):
        pass

    def setup_block_type(self, block_type):
        """(partial charges, inter repr path distance, intra repr path
        distance) of one block type (elec_energy_term.py:36-91)."""
        key = block_type.name
        if key not in self._tables:
            q = self.param_resolver.get_partial_charges_for_block(block_type)
            rep = self.param_resolver.get_bonded_path_length_mapping_for_block(block_type)
            inter = block_type.path_distance[:, rep]
            intra = inter[rep, :].copy()
            if not block_type.is_polymer:
                intra[(intra == 3) | (intra == 4)] = 5
            self._tables[key] = (q, inter, intra)
        return self._tables[key]

    def packed_tables(self, pbt):
        """[T, Amax] charges and [T, Amax, Amax] path-distance tables for a
        PackedBlockTypes (elec_energy_term.py:93-148); padding is zero."""
        cache = getattr(pbt, "_elec_tables", None)
        if cache is not None:
            return cache
        T, A = pbt.n_types, pbt.max_n_atoms
        q = numpy.zeros((T, A), dtype=numpy.float32)
        inter = numpy.zeros((T, A, A), dtype=numpy.int32)
        intra = numpy.zeros((T, A, A), dtype=numpy.int32)
        for i, bt in enumerate(pbt.active_block_types):
            qi, ei, ai = self.setup_block_type(bt)
            n = bt.n_atoms
            q[i, :n] = qi
            inter[i, :n, :n] = ei
            intra[i, :n, :n] = ai
        lig = numpy.array([int(bt.is_ligand_fragment) for bt in pbt.active_block_types],
                          dtype=numpy.int32)
        cache = {
            "partial_charge": torch.tensor(q, device=pbt.device),
            "inter_repr_path_distance": torch.tensor(inter, device=pbt.device),
            "intra_repr_path_distance": torch.tensor(intra, device=pbt.device),
            "is_ligand_fragment": torch.tensor(lig, device=pbt.device),
        }
        pbt._elec_tables = cache
        return cache

    @property
    def global_params(self) -> torch.Tensor:
        g = self.param_resolver.global_params
        vals = [
            g["elec_sigmoidal_die_D"],
            g["elec_sigmoidal_die_D0"],
            g["elec_sigmoidal_die_S"],
            g["elec_min_dis"],
            g["elec_max_dis"],
        ]
        return torch.tensor([vals], dtype=torch.float32, device=self.device)

    @property
    def max_dis(self) -> float:
        # float32 value of elec_max_dis, handed to the op as a host scalar
        return float(numpy.float32(self.param_resolver.global_params["elec_max_dis"]))

    def get_pose_score_term_function(self):
        from ..ops import elec_pose_scores

        return elec_pose_scores

    def get_score_term_attributes(self, pose_stack):
        """The 11 term-specific op arguments, in the op's order
        (elec_energy_term.py:158-186)."""
        pbt = pose_stack.packed_block_types
        t = self.packed_tables(pbt)
        return [
            pose_stack.min_block_bondsep,
            pose_stack.inter_block_bondsep,
            pbt.n_atoms,
            t["partial_charge"],
            pbt.n_conn,
            pbt.conn_atom,
            t["inter_repr_path_distance"],
            t["intra_repr_path_distance"],
            t["is_ligand_fragment"],
            self.global_params.to(pose_stack.device),
            self.max_dis,
        ]

    def render_whole_pose_scoring_module(self, pose_stack):
        from .scoring_module import TermWholePoseScoringModule

        return TermWholePoseScoringModule(
            self.class_name(),
            pose_stack,
            self.get_score_term_attributes(pose_stack),
            self.get_pose_score_term_function(),
        )

    def render_block_pair_scoring_module(self, pose_stack):
        from .scoring_module import TermBlockPairScoringModule

        return TermBlockPairScoringModule(
            self.class_name(),
            pose_stack,
            self.get_score_term_attributes(pose_stack),
            self.get_pose_score_term_function(),
        )

    def get_rotamer_score_term_function(self):
        from ..ops import elec_rotamer_scores

        return elec_rotamer_scores

    def render_rotamer_scoring_module(self, pose_stack, rotamer_set):
        """Rotamer-pair scoring for packing (the term's
        render_rotamer_scoring_module): the pose stack supplies the
        bond-separation tables, the rotamer set the coordinates' layout."""
        from .scoring_module import TermRotamerScoringModule

        return TermRotamerScoringModule(
            self.class_name(),
            rotamer_set,
            self.get_score_term_attributes(pose_stack),
            self.get_rotamer_score_term_function(),
        )


def elec_op_args(pose_stack, term: ElecEnergyTerm, coords=None, block_pair=False):
    """Full 25-argument list of ``elec_pose_scores`` for a padded stack
    (tmol/score/common/scoring_module.py:56-72 + the term attributes)."""
    ps = pose_stack
    if coords is None:
        coords = ps.coords
    return [
        coords.flatten(0, -2),
        ps.rot_coord_offset,
        ps.pose_ind_for_atom,
        ps.first_rot_for_block,
        ps.block_type_ind,
        ps.block_ind_for_rot,
        ps.pose_ind_for_rot,
        ps.block_type_ind_for_rot,
        ps.n_rots_for_pose,
        ps.rot_offset_for_pose,
        ps.n_rots_for_block,
        ps.rot_offset_for_block,
        ps.max_n_rots_per_pose,
    ] + term.get_score_term_attributes(ps) + [bool(block_pair)]
