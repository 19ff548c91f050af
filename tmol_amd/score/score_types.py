"""Score types produced by the LJ+LK term, in output order
(tmol/score/score_types.py; tmol/score/terms/ljlk_creator.py:7-19)."""

import enum


class ScoreType(enum.IntEnum):
    fa_ljatr = 0
    fa_ljrep = 1
    fa_lk = 2
