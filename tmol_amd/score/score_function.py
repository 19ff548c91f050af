"""ScoreFunction / WholePoseScoringModule slice for the two-body terms built
here, LJ+LK and elec (tmol/score/score_function.py:26-199, 321-415).  A term
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

    def render_rotamer_scoring_module(self, pose_stack, rotamer_set):
        """Weighted rotamer-pair energies of a rotamer set, summed over the
        terms, as a sparse (n_poses, n_rots, n_rots) tensor
        (score_function.py:182-198)."""
        self.pre_work_initialization(pose_stack)
        mods = [t.render_rotamer_scoring_module(pose_stack, rotamer_set)
                for t in self.all_terms()]
        return RotamerScoringModule(self.weights_tensor(), mods)


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


class RotamerScoringModule:
    """Sum of the terms' weighted rotamer-pair energies (score_function.py:
    362-415): each term's [n_subterms, nnz] scores are weighted and summed
    at the nnz level, the terms' entries concatenated, and one sparse COO
    tensor (n_poses, n_rots, n_rots) built; entries of a rotamer pair listed
    by several terms add up when coalesced."""

    def __init__(self, weights, term_modules):
        self.weights = torch.nn.Parameter(weights.view(-1, 1, 1, 1), requires_grad=False)
        self.term_modules = term_modules

    def __call__(self, coords):
        values, indices = [], []
        n_poses = n_rots = None
        off = 0
        for term in self.term_modules:
            scores, idx = term.forward(coords)
            n_sub = scores.shape[0]
            w = self.weights[off: off + n_sub, 0, 0, 0].to(scores.dtype)
            values.append((w[:, None] * scores).sum(dim=0))
            indices.append(idx)
            off += n_sub
            if n_poses is None:
                n_poses, n_rots = term.n_poses, term.n_rots
        if n_poses is None:
            return torch.sparse_coo_tensor(
                torch.zeros((3, 0), dtype=torch.int32, device=coords.device),
                torch.zeros(0, dtype=torch.float32, device=coords.device),
                size=(0, 0, 0),
            )
        return torch.sparse_coo_tensor(torch.cat(indices, dim=1), torch.cat(values),
                                       size=(n_poses, n_rots, n_rots))
