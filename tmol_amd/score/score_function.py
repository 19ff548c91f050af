"""ScoreFunction / WholePoseScoringModule slice for the LJ+LK term
(tmol/score/score_function.py:26-165, 321-355)."""

import torch

from .ljlk_energy_term import LJLKEnergyTerm
from .score_types import ScoreType


class ScoreFunction:
    def __init__(self, param_db=None, device="cpu"):
        self._device = torch.device(device)
        self._weights = torch.zeros(len(ScoreType), device=self._device)
        self._term = None
        self._param_db = param_db
        self.term_options = {}

    def set_weight(self, st: ScoreType, weight: float):
        if self._term is None:
            self._term = LJLKEnergyTerm(device=self._device, chemical_db=self._param_db)
        self._weights[int(st)] = weight

    def get_weight(self, st: ScoreType):
        return self._weights[int(st)]

    def all_terms(self):
        return [] if self._term is None else [self._term]

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
