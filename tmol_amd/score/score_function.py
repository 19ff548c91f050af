"""ScoreFunction / WholePoseScoringModule slice for the two-body terms built
here, LJ+LK and elec (tmol/score/score_function.py:26-165, 321-355).  A term
is created the first time one of its score types gets a weight; terms keep
their creation order, and each contributes its score types in its own order.
"""

import torch

from .elec_energy_term import ElecEnergyTerm
from .ljlk_energy_term import LJLKEnergyTerm
from .score_types import ScoreType

_TERM_FOR_SCORE_TYPE = {
    ScoreType.fa_ljatr: LJLKEnergyTerm,
    ScoreType.fa_ljrep: LJLKEnergyTerm,
    ScoreType.fa_lk: LJLKEnergyTerm,
    ScoreType.fa_elec: ElecEnergyTerm,
}


class ScoreFunction:
    def __init__(self, param_db=None, device="cpu"):
        self._device = torch.device(device)
        self._weights = torch.zeros(len(ScoreType), device=self._device)
        self._terms = {}
        self._param_db = param_db
        self.term_options = {}

    def set_weight(self, st: ScoreType, weight: float):
        cls = _TERM_FOR_SCORE_TYPE[ScoreType(st)]
        if cls not in self._terms:
            self._terms[cls] = cls(device=self._device, chemical_db=self._param_db)
        self._weights[int(st)] = weight

    def get_weight(self, st: ScoreType):
        return self._weights[int(st)]

    def all_terms(self):
        return list(self._terms.values())

    def set_option(self, key, value):
        self.term_options[key] = value

    def set_options(self, options):
        self.term_options = dict(options)

    def weights_tensor(self):
        return torch.tensor(
            [float(self._weights[int(st)]) for t in self.all_terms() for st in t.score_types()],
            dtype=torch.float32,
            device=self._device,
        )

    def pre_work_initialization(self, pose_stack):
        for t in self.all_terms():
            t.set_options(self.term_options)

    def render_whole_pose_scoring_module(self, pose_stack):
        self.pre_work_initialization(pose_stack)
        mods = [t.render_whole_pose_scoring_module(pose_stack) for t in self.all_terms()]
        return WholePoseScoringModule(self.weights_tensor(), mods)

    def render_block_pair_scoring_module(self, pose_stack):
        self.pre_work_initialization(pose_stack)
        mods = [t.render_block_pair_scoring_module(pose_stack) for t in self.all_terms()]
        return BlockPairScoringModule(self.weights_tensor(), mods)


class WholePoseScoringModule:
    def __init__(self, weights, term_modules):
        self.weights = torch.nn.Parameter(weights.unsqueeze(1), requires_grad=False)
        self.term_modules = term_modules

    def __call__(self, coords, sum_terms=True, apply_weights=True):
        unweighted = self.unweighted_scores(coords)
        weighted = self.weights * unweighted if apply_weights else unweighted
        return torch.sum(weighted, dim=0) if sum_terms else weighted

    def unweighted_scores(self, coords):
        return torch.cat([term(coords) for term in self.term_modules], dim=0)


class BlockPairScoringModule:
    def __init__(self, weights, term_modules):
        self.weights = torch.nn.Parameter(
            weights.view(-1, 1, 1, 1), requires_grad=False
        )
        self.term_modules = term_modules

    def __call__(self, coords, sum_terms=True, apply_weights=True):
        unweighted = self.unweighted_scores(coords)
        weighted = self.weights * unweighted if apply_weights else unweighted
        return torch.sum(weighted, dim=0) if sum_terms else weighted

    def unweighted_scores(self, coords):
        return torch.cat([term(coords) for term in self.term_modules], dim=0)
