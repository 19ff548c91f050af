"""Scoring API slice for the LJ+LK and elec terms: ScoreFunction, score types,
terms."""

from .score_types import ScoreType  # noqa: F401
from .ljlk_energy_term import LJLKEnergyTerm  # noqa: F401
from .elec_energy_term import ElecEnergyTerm  # noqa: F401
from .score_function import ScoreFunction, WholePoseScoringModule  # noqa: F401


def beta2016_ljlk_score_function(device="cpu"):
    """ScoreFunction with the LJ+LK weights of beta2016
    (tmol/database/default/score_functions/beta2016.sfxn:6-8)."""
    sfxn = ScoreFunction(device=device)
    sfxn.set_weight(ScoreType.fa_ljatr, 1.0)
    sfxn.set_weight(ScoreType.fa_ljrep, 0.55)
    sfxn.set_weight(ScoreType.fa_lk, 1.0)
    return sfxn
