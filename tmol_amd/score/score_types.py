"""Score types of the two-body terms built here, in output order within each
term (tmol/score/score_types.py; tmol/score/terms/ljlk_creator.py:7-19,
elec_creator.py, lk_ball_creator.py, hbond_creator.py)."""

import enum


class ScoreType(enum.IntEnum):
    fa_ljatr = 0
    fa_ljrep = 1
    fa_lk = 2
    fa_elec = 3
    lk_ball_iso = 4
    lk_ball = 5
    lk_bridge = 6
    lk_bridge_uncpl = 7
    hbond = 8
