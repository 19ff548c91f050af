"""LJ+LK energy term: parameter resolution and the op's term attributes.

Mirrors tmol/score/ljlk/ljlk_energy_term.py:16-178 and
tmol/score/ljlk/params.py:63-163:

* ``type_params [n_types, 10]`` columns ``lj_radius, lj_wdepth, lk_dgfree,
  lk_lambda, lk_volume, is_donor, is_hydroxyl, is_polarh, is_acceptor,
  is_carbon_lk`` (ljlk_energy_term.py:130-146); the trailing "None" atom type
  row is NaN for the float columns and 0 for the flags (params.py:125-131);
* ``global_params [1, 5]`` = ``max_dis, lj_dlin_sigma_factor (0.91 with the
  soft_rep option), lj_hbond_dis, lj_hbond_OH_donor_dis, lj_hbond_hdis``
  (ljlk_energy_term.py:147-162);
* ``is_carbon_lk`` = atom types CH1, CH2, CH3, Caro (params.py:131-136).
"""

import numpy
import torch

from ..chemical import default_chemical_database
from .score_types import ScoreType


class LJLKEnergyTerm:
    tile_size = 32

    def __init__(self, device="cpu", chemical_db=None):
        self.device = torch.device(device)
        self.db = chemical_db or default_chemical_database()
        self.soft_repulsive = False
        ats = self.db.atom_type_set
        lj = {p["name"]: p for p in self.db.ljlk["atom_type_parameters"]}
        rows = []
        for i, name in enumerate(ats.names + [None]):
            if name is None or name not in lj:
                rows.append([numpy.nan] * 5 + [0.0] * 5)
                continue
            p = lj[name]
            rows.append(
                [
                    p["lj_radius"],
                    p["lj_wdepth"],
                    p["lk_dgfree"],
                    p["lk_lambda"],
                    p["lk_volume"],
                    float(ats.is_donor[i]),
                    float(ats.is_hydroxyl[i]),
                    float(ats.is_polarh[i]),
                    float(ats.is_acceptor[i]),
                    float(name in ("CH1", "CH2", "CH3", "Caro")),
                ]
            )
        # float32 values of the YAML numbers, as torch.tensor(...) makes them
        self._type_params = torch.tensor(
            numpy.array(rows, dtype=numpy.float64), dtype=torch.float32
        )
        self.global_param_values = self.db.ljlk["global"]

    @classmethod
    def class_name(cls):
        return "LJLK"

    @classmethod
    def score_types(cls):
        return [ScoreType.fa_ljatr, ScoreType.fa_ljrep, ScoreType.fa_lk]

    def n_bodies(self):
        return 2

    def set_options(self, options: dict):
        if "soft_rep" in options:
            self.soft_repulsive = bool(options["soft_rep"])

    @property
    def type_params(self) -> torch.Tensor:
        return self._type_params.to(self.device)

    @property
    def global_params(self) -> torch.Tensor:
        g = self.global_param_values
        vals = [
            g["max_dis"],
            g["lj_dlin_sigma_factor_soft"] if self.soft_repulsive else g["lj_dlin_sigma_factor"],
            g["lj_hbond_dis"],
            g["lj_hbond_OH_donor_dis"],
            g["lj_hbond_hdis"],
        ]
        return torch.tensor([vals], dtype=torch.float32, device=self.device)

    @property
    def max_dis(self) -> float:
        # float32 value of max_dis, handed to the op as a host scalar
        return float(numpy.float32(self.global_param_values["max_dis"]))

    def get_pose_score_term_function(self):
        from ..ops import ljlk_pose_scores

        return ljlk_pose_scores

    def get_score_term_attributes(self, pose_stack):
        """The 13 term-specific op arguments, in the op's order
        (ljlk_energy_term.py:163-178)."""
        pbt = pose_stack.packed_block_types
        return [
            pose_stack.min_block_bondsep,
            pose_stack.inter_block_bondsep,
            pbt.n_atoms,
            pbt.ljlk_n_heavy_atoms_in_tile,
            pbt.ljlk_heavy_atoms_in_tile,
            pbt.atom_types,
            pbt.n_conn,
            pbt.conn_atom,
            pbt.ljlk_bond_separation,
            pbt.ljlk_is_ligand_fragment,
            self.type_params.to(pose_stack.device),
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
        from ..ops import ljlk_rotamer_scores

        return ljlk_rotamer_scores

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


def pose_op_args(pose_stack, term: LJLKEnergyTerm, coords=None, block_pair=False):
    """Full 27-argument list of ``ljlk_pose_scores`` for a padded stack
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
