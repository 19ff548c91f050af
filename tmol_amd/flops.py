"""Algorithmic work model of LJ+LK energy + gradient.

The work unit is one atom-pair evaluation inside a neighbor block pair (LJ:
all atom pairs, LK: heavy-atom pairs; n1*n2 between blocks, n(n-1)/2 within a
block).  FLOPs are counted per potential region with +, -, *, /, sqrt, exp
and floor as one each; compares and selects are free; quantities that depend
only on the atom-type pair (sigma, epsilon, d_lin, spline knots, spline
end-point values and Hermite coefficients, f_desolv at the knots) are
treated as precomputed, so the table is the minimum work of the function, not
the instruction count of any implementation.

Per pair, energy + gradient:
  distance           dx, dy, dz, d^2                         8
  sqrt                                                       1   (d < max_dis)
  gradient scatter   1/d, g/d, 3 mul, 3 + 3 accumulate      11   (dV != 0)
  energy accumulate  LJ: atr, rep                            2;  LK: 1
LJ regions (lj.hh:115-177)
  linear             V_lin + V'_lin (d - d_lin), split, w    6
  12-6               sigma/d, powers, V, dV, split, w       14
  Hermite to zero    t, Horner V and dV, 1/w, w             15
LK regions, both directions (lk_isotropic.hh:167-235, 328-360)
  constant           sum, w                                  2   (no gradient)
  close Hermite      2 x 13 + sum + w                       30
  f_desolv           2 x 12 (exp) + sum + w                 28
  far Hermite        2 x 13 + sum + w                       30
Pairs beyond max_dis cost the distance only (8); pairs with count-pair
weight 0 cost nothing.
"""

CENSUS_FIELDS = [
    "lj_pairs", "lj_cp0", "lj_far", "lj_linear", "lj_126", "lj_spline",
    "lk_pairs", "lk_cp0", "lk_far", "lk_const", "lk_close", "lk_fdesolv", "lk_spline",
    "items", "intra_items",
]

DIST, SQRT, GRAD = 8, 1, 11
FLOPS_ENERGY_GRAD = {
    "lj_far": DIST,
    "lj_linear": DIST + SQRT + 6 + GRAD + 2,
    "lj_126": DIST + SQRT + 14 + GRAD + 2,
    "lj_spline": DIST + SQRT + 15 + GRAD + 2,
    "lk_far": DIST,
    "lk_const": DIST + SQRT + 2 + 1,
    "lk_close": DIST + SQRT + 30 + GRAD + 1,
    "lk_fdesolv": DIST + SQRT + 28 + GRAD + 1,
    "lk_spline": DIST + SQRT + 30 + GRAD + 1,
}


# Energy only (the rotamer-pair forward, lj_atom_energy / lk_atom_energy):
# no gradient scatter; LJ linear: V_lin + V'_lin (d - d_lin), split, w (6);
# 12-6: sigma/d, powers, V, split, w (11); Hermite: t, Horner V, w (10);
# LK per direction 8 (close/far Hermite: t, Horner; f_desolv: 1/d^2,
# (d - r)^2 / lambda^2, exp, scale), both + sum + w = 18.
FLOPS_ENERGY = {
    "lj_far": DIST,
    "lj_linear": DIST + SQRT + 6 + 2,
    "lj_126": DIST + SQRT + 11 + 2,
    "lj_spline": DIST + SQRT + 10 + 2,
    "lk_far": DIST,
    "lk_const": DIST + SQRT + 2 + 1,
    "lk_close": DIST + SQRT + 18 + 1,
    "lk_fdesolv": DIST + SQRT + 18 + 1,
    "lk_spline": DIST + SQRT + 18 + 1,
}


# Elec (potentials.hh:55-201), per enumerated atom pair, in the same
# convention (energy + gradient):
#   pairs with a zero charge or count-pair weight 0       0
#   beyond max_dis                                          8 (distance)
#   flat (d < min_dis - 0.25)   d, q1 q2 scale, w, accumulate          DIST+SQRT+3
#   low / high Hermite fade     t, Horner V and dV, q1 q2 (2), w (2),
#                               gradient, accumulate                   DIST+SQRT+13+4+GRAD+1
#   Coulomb                     eps: exp, x S, polynomial, product,
#                               D - (8); V: C1 / (d eps) - C2 (3);
#                               dV: eps' (4), eps + d eps' (2),
#                               -C1 (..) / (d^2 eps^2) (5); q1 q2 (2),
#                               w (2), gradient, accumulate             DIST+SQRT+26+GRAD+1
ELEC_CENSUS_FIELDS = [
    "items", "intra_items", "pairs", "charged", "cp0",
    "flat", "low", "coulomb", "high", "beyond",
]

ELEC_FLOPS_ENERGY_GRAD = {
    "far": DIST,
    "flat": DIST + SQRT + 3,
    "low": DIST + SQRT + 13 + 4 + GRAD + 1,
    "coulomb": DIST + SQRT + 26 + GRAD + 1,
    "high": DIST + SQRT + 13 + 4 + GRAD + 1,
}


def elec_far_pairs(census: dict) -> int:
    """Charged pairs of non-zero weight beyond max_dis (culled before the
    evaluation, or evaluated to zero)."""
    inside = census["flat"] + census["low"] + census["coulomb"] + census["high"]
    return census["charged"] - census["cp0"] - inside


def elec_algorithmic_flops(census: dict) -> int:
    n = dict(census)
    n["far"] = elec_far_pairs(census)
    return sum(n[k] * f for k, f in ELEC_FLOPS_ENERGY_GRAD.items())


def census_dict(values):
    return {k: int(v) for k, v in zip(CENSUS_FIELDS, values)}


def algorithmic_flops(census: dict, energy_only: bool = False) -> int:
    table = FLOPS_ENERGY if energy_only else FLOPS_ENERGY_GRAD
    return sum(census[k] * f for k, f in table.items())


def pair_evaluations(census: dict) -> int:
    """Atom-pair evaluations (LJ all-atom + LK heavy-atom pairs)."""
    return census["lj_pairs"] + census["lk_pairs"]


def algorithmic_bytes(n_atoms: int, n_blocks: int, n_poses: int, n_items: int,
                      real_bytes: int = 4) -> dict:
    """Compulsory HBM bytes of one energy+gradient step, per kernel:
    neighbors: coords in, block metadata in, spheres out/in, work list out;
    score: coords in, work list in, block metadata in, gradient out,
    per-pose energies out."""
    coords = 3 * real_bytes * n_atoms
    meta = 8 * n_blocks
    spheres = 4 * real_bytes * n_blocks
    items = 8 * n_items
    return {
        "neighbors": coords + meta + 2 * spheres + items,
        "score": coords + items + 16 * n_items + 3 * real_bytes * n_atoms + 3 * 8 * n_poses,
    }


# lk_ball (tmol/score/lk_ball/potentials/lk_ball.hh:42-494), per enumerated
# (polar atom, occluder) pair of a neighbor block pair, same convention,
# energy + weighted gradient of the four score types:
#   separation <= 3                                                   0
#   beyond max_dis, or a polar atom without waters                    8 (distance)
#   evaluated   distance + sqrt 9; isotropic LK one direction with
#               derivative 14; lk_fraction: 4 waters x (3 sub, 3 mul,
#               2 add, - d2_low, exp, sum = 11) = 44, -log, ramp
#               polynomial 7; its derivative 4 x 10 + 6 = 46; the four
#               energies 4; weighted gradient combination 12; I, J
#               gradient 6                                           142
#   + bridge    (occluder a donor or acceptor with waters) 16 water
#               pairs x 11 = 176, -log, overlap polynomial 7, base
#               length fraction 8; derivative 16 x 11 + 10          +377
# Energy only: evaluated 9 + 13 + 51 + 4 = 77, bridge +191.
LKB_CENSUS_FIELDS = ["candidates", "cp0", "evaluated", "bridge"]
LKB_FLOPS_ENERGY_GRAD = {"far": DIST, "evaluated": 142, "bridge": 377}
LKB_FLOPS_ENERGY = {"far": DIST, "evaluated": 77, "bridge": 191}


def lkb_algorithmic_flops(census: dict, gradient: bool = True) -> int:
    t = LKB_FLOPS_ENERGY_GRAD if gradient else LKB_FLOPS_ENERGY
    far = census["candidates"] - census["cp0"] - census["evaluated"]
    return int(far * t["far"] + census["evaluated"] * t["evaluated"]
               + census["bridge"] * t["bridge"])


# hbond (tmol/score/hbond/potentials/potentials.hh:36-508), per donor H x
# acceptor candidate of a neighbor block pair, same convention (one FLOP per
# add, multiply, divide, square root or transcendental; the double-precision
# polynomial steps count as FLOPs too), energy + weighted gradient:
#   separation < 5                                                    0
#   H-A distance test                                                 9
#   evaluated   A-H distance with its polynomial (10 Horner steps of
#               value and slope, 40) and gradient 56; A-H-D angle
#               (two differences, cross product, normalisation, atan2,
#               gradients of both arms) with polynomial 150; B-A-H
#               polynomial of the cosine with gradient 95 (sp3: twice
#               plus the soft-max, 230); sp2 chi: angle 100, dihedral
#               with Blondel-Karplus gradient 110, the chi energy's
#               trigonometry 120; weight, fade and the 15 scaled
#               gradient components 35.  Backbone O (sp2) dominates the
#               acceptors, so the sp2 total is used for every pair      670
# Energy only: 9 + 45 + 90 + 60 + 200 + 10 = 414.
HB_CENSUS_FIELDS = ["candidates", "cp_excluded", "evaluated"]
HB_FLOPS_ENERGY_GRAD = {"test": DIST + 1, "evaluated": 670}
HB_FLOPS_ENERGY = {"test": DIST + 1, "evaluated": 414}


def hb_algorithmic_flops(census: dict, gradient: bool = True) -> int:
    t = HB_FLOPS_ENERGY_GRAD if gradient else HB_FLOPS_ENERGY
    tested = census["candidates"] - census["cp_excluded"]
    return int(tested * t["test"] + census["evaluated"] * t["evaluated"])
