"""Pins for the lk_ball oracle (CPU): tmol's lk_ball term baselines for 1ubq,
the potential-level golden values of tmol/tests/score/lk_ball/potentials/
test_compiled_lk_ball.py, central finite differences of the water builders,
the pair potential and the whole chain, the host-side block-type tables,
and the reference's own CPU lk_ball ops where oracle/_ref has been built
(tmol/tests/score/lk_ball/test_lk_ball_energy_term.py)."""

import math
import os

import numpy
import pytest
import torch

from conftest import load_term_baseline
from oracle import oracle as orc
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack
from tmol_amd.score.lk_ball_energy_term import (
    LKBallEnergyTerm, score_op_args, watergen_op_args)

R = math.radians


@pytest.fixture(scope="module")
def lkb_term():
    return LKBallEnergyTerm()


def _args(ps, term, block_pair=False, waters=None):
    """Both ops' arguments, float parameters in the coordinates' dtype (as
    the reference's convert_float64 does, lk_ball_energy_term.py:270, 298)."""
    wa = watergen_op_args(ps, term)
    if ps.coords.dtype == torch.float64:
        for i in (29, 30, 31, 32):
            wa[i] = wa[i].double()
    if waters is None:
        waters = torch.tensor(orc.oracle_lkb_waters(*wa))
    sa = score_op_args(ps, term, waters, block_pair=block_pair)
    if ps.coords.dtype == torch.float64:
        sa[22], sa[24] = sa[22].double(), sa[24].double()
    return wa, sa


def _scores(ps, term, block_pair=False):
    wa, sa = _args(ps, term, block_pair)
    s, _, _, _ = orc.oracle_lkb_score(wa, sa)
    return s


# ---------------------------------------------------------------- goldens


def test_whole_pose_scoring_10(ubq_pdb, lkb_term):
    # test_energy_term.py whole-pose baseline of LKBallEnergyTerm
    ps = PoseStack([pose_from_pdb(ubq_pdb)] * 10)
    s = _scores(ps, lkb_term)
    numpy.testing.assert_allclose(s, load_term_baseline("test_whole_pose_scoring_10", "lk_ball"),
                                  atol=1e-5, rtol=1e-3)
    # test_lk_ball_energy_term.py:49-71, the scoring-module smoke values
    numpy.testing.assert_allclose(s[:, :1], [[422.0388], [172.1965], [1.581745], [11.031567]],
                                  atol=1e-3, rtol=1e-3)


def test_whole_pose_scoring_jagged(ubq_pdb, lkb_term):
    ps = PoseStack([pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (0, 50)),
                    pose_from_pdb(ubq_pdb, (20, 50))])
    numpy.testing.assert_allclose(
        _scores(ps, lkb_term), load_term_baseline("test_whole_pose_scoring_jagged", "lk_ball"),
        atol=1e-5, rtol=1e-3)


def test_block_scoring(ubq_pdb, lkb_term):
    ps = PoseStack([pose_from_pdb(ubq_pdb, (0, 4))])
    s = _scores(ps, lkb_term, block_pair=True)
    numpy.testing.assert_allclose(s, load_term_baseline("test_block_scoring", "lk_ball"),
                                  atol=1e-5, rtol=1e-3)
    w = _scores(ps, lkb_term)
    numpy.testing.assert_allclose(w[:, 0], s[:, 0].sum((1, 2)), atol=1e-5, rtol=1e-5)


# ---------------------------------------------------------------- potential


def test_water_builders_golden_and_vjp():
    # test_compiled_lk_ball.py:26-75
    A, B, B0 = (-6.071, -0.619, -3.193), (-5.250, -1.595, -2.543), (-5.489, 0.060, -3.542)
    refs = [[-7.42086525, -1.79165583, -5.14882262], [-7.75428876, 0.40906314, -1.4232189]]
    for t, ref in zip((120.0, 240.0), refs):
        W = orc.oracle_lkb_build_water(2, A, B, B0, 2.65, R(109.0), R(t))
        numpy.testing.assert_allclose(W, ref, atol=1e-7)
    D, H = (-6.007, 4.706, -0.074), (-6.747, 4.361, 0.549)
    numpy.testing.assert_allclose(orc.oracle_lkb_build_water(0, D, H, None, 2.65),
                                  [-7.91642236, 3.81579633, 1.5335272], atol=1e-7)
    # closed forms: the acceptor water lies dist from A at angle (A, W, B)
    # = the build angle; the donor water lies on the D -> H ray
    W = orc.oracle_lkb_build_water(2, A, B, B0, 2.828, R(109.3), R(120.0))
    a, b, w = map(numpy.asarray, (A, B, W))
    assert abs(numpy.linalg.norm(w - a) - 2.828) < 1e-12
    cosang = numpy.dot(w - a, b - a) / (numpy.linalg.norm(w - a) * numpy.linalg.norm(b - a))
    assert abs(math.degrees(math.acos(cosang)) - 109.3) < 1e-9
    # vector-Jacobian products against central differences, every kind
    rng = numpy.random.default_rng(5)
    h = 1e-6
    for kind, ang in ((0, 0.0), (1, R(134.5)), (2, R(109.3)), (3, R(180.0))):
        pts = [numpy.asarray(x, dtype=numpy.float64) for x in (A, B, B0)]
        g = rng.normal(size=3)
        W, grads = orc.oracle_lkb_build_water(kind, *pts, 2.828, ang, R(240.0), g=g)
        for which in range(3 if kind else 2):
            for k in range(3):
                pp = [p.copy() for p in pts]
                pm = [p.copy() for p in pts]
                pp[which][k] += h
                pm[which][k] -= h
                fd = (numpy.dot(g, orc.oracle_lkb_build_water(kind, *pp, 2.828, ang, R(240.0)))
                      - numpy.dot(g, orc.oracle_lkb_build_water(kind, *pm, 2.828, ang, R(240.0))))
                assert abs(fd / (2 * h) - grads[which][k]) < 1e-7, (kind, which, k)


def _acc_waters(A, B, B0, tors, dist=2.65, ang=R(109.0)):
    return numpy.stack([orc.oracle_lkb_build_water(2, A, B, B0, dist, ang, R(t)) for t in tors])


def test_lk_fraction_and_bridge_golden():
    # test_compiled_lk_ball.py:78-204
    WI = _acc_waters((0, 0, 0), (0, 0, -1), (1, 0, 0), (120.0, 240.0))
    assert abs(orc.oracle_lkb_fraction(WI, (-2.5, 0.1, 2.5), 1.8) - 0.65) < 0.01
    WI = _acc_waters((0, 3, 3), (0, 3, 4), (1, 0, 0), (60.0, 300.0))
    WJ = _acc_waters((0, 0, 0), (0, 0, -1), (1, 0, 0), (120.0, 240.0))
    assert abs(orc.oracle_lkb_bridge((0, 3, 3), (0, 0, 0), WI, WJ, 2.65) - 0.025) < 0.001
    # strongly overlapping waters clamp the overlap fraction to 1
    f = orc.oracle_lkb_bridge((0, 0, 0), (4.32664, 0, 0), [[0, 0, 0], [0.1, 0, 0]],
                              [[0.05, 0.05, 0], [0, 0.1, 0.05]], 2.65)
    assert abs(f - 1.0) < 1e-3
    # no waters at all: fraction 0 (the -log of an empty sum is +inf)
    assert orc.oracle_lkb_fraction(numpy.full((4, 3), numpy.nan), (1, 2, 3), 1.8) == 0.0


def _tp(term, name):
    ats = term.db.atom_type_set
    return term.lkb_type_params[ats.type_index(name)].double().numpy()


def _gp(term):
    return numpy.float32(term.global_values).astype(numpy.float64)


def _pair(term, I, J, WI, WJ, ti, tj, sep=5):
    return orc.oracle_lkb_pair(I, J, orc._waters4(WI, numpy.float64),
                               orc._waters4(WJ, numpy.float64), sep, _tp(term, ti),
                               _tp(term, tj), _gp(term))


def test_pair_spotchecks_golden(lkb_term):
    # test_compiled_lk_ball.py:267-517 at bonded path length 5
    T = lkb_term
    gp = _gp(T)
    wd = gp[3]
    c = numpy.array([[-6.007, 4.706, -0.074], [-6.747, 4.361, 0.549], [-5.791, 5.657, 0.240],
                     [-6.305, 4.706, -1.040], [-10.018, 6.062, -2.221], [-9.160, 5.711, -2.665],
                     [-9.745, 6.899, -1.697], [-10.429, 5.372, -1.610]])
    wi = numpy.stack([orc.oracle_lkb_build_water(0, c[0], c[k], None, wd) for k in (1, 2, 3)])
    wj = numpy.stack([orc.oracle_lkb_build_water(0, c[4], c[k], None, wd) for k in (5, 6, 7)])
    s = _pair(T, c[0], c[4], wi, wj, "Nlys", "Nlys") + _pair(T, c[4], c[0], wj, wi, "Nlys", "Nlys")
    numpy.testing.assert_allclose(s, [0.3355, 0.0, 0.2649, 0.7896], atol=1e-4, rtol=1e-4)

    c = numpy.array([[-5.282, -0.190, -0.858], [-6.520, 0.686, -0.931], [-6.652, 1.379, -1.961],
                     [-6.932, 4.109, -5.683]])
    ws = numpy.stack([orc.oracle_lkb_build_water(2, c[2], c[1], c[0], wd, R(134.5), t)
                      for t in T.tors_sp2])
    s = _pair(T, c[2], c[3], ws, numpy.full((2, 3), numpy.nan), "OOC", "CH3")
    numpy.testing.assert_allclose(s, [0.14107985, 0.04765878, 0.0, 0.0], atol=1e-5, rtol=1e-5)

    c = numpy.array([[-5.250, -1.595, -2.543], [-6.071, -0.619, -3.193], [-5.489, 0.060, -3.542],
                     [-10.628, 2.294, -1.933], [-9.991, 1.160, -1.435], [-10.715, 0.960, -0.319]])
    sp3 = numpy.stack([orc.oracle_lkb_build_water(2, c[1], c[0], c[2], wd, R(109.3), t)
                       for t in T.tors_sp3] + [orc.oracle_lkb_build_water(0, c[1], c[2], None, wd)])
    ring = numpy.stack([orc.oracle_lkb_build_water(3, c[4], 0.5 * (c[3] + c[5]), c[5], wd,
                                                   R(180.0), t) for t in T.tors_ring])
    s = _pair(T, c[1], c[4], sp3, ring, "OH", "NhisDDepro") + _pair(
        T, c[4], c[1], ring, sp3, "NhisDDepro", "OH")
    numpy.testing.assert_allclose(s, [0.09018922, 0.0901892, 0.0441963, 0.4900393], atol=1e-4)
    nan = numpy.full((3, 3), numpy.nan)
    numpy.testing.assert_allclose(_pair(T, c[1], c[3], sp3, nan, "OH", "CH0"),
                                  [0.00385956, 0.0001626, 0.0, 0.0], atol=1e-4, rtol=1e-4)
    numpy.testing.assert_allclose(_pair(T, c[1], c[5], sp3, nan, "OH", "Caro"),
                                  [0.00369549, 0.0028072, 0.0, 0.0], atol=1e-4, rtol=1e-4)
    numpy.testing.assert_allclose(_pair(T, c[4], c[0], ring, nan, "NhisDDepro", "CH2"),
                                  [0.01360676, 0.0135272, 0.0, 0.0], atol=1e-4, rtol=1e-4)


def test_pair_rules(lkb_term):
    T = lkb_term
    # two acceptors 4.6 A apart (near the bridge's base-length peak, sqrt(8/3)
    # water distances) sharing a water between them
    I, J = numpy.array([0.0, 0, 0]), numpy.array([4.6, 0.0, 0.1])
    WI = numpy.array([[2.25, 0.4, 0.0], [-1.5, 2.0, 0.5]])
    WJ = numpy.array([[2.35, 0.2, 0.1], [6.5, -1.8, 0.0]])
    e5 = _pair(T, I, J, WI, WJ, "OH", "OCbb", sep=5)
    e4 = _pair(T, I, J, WI, WJ, "OH", "OCbb", sep=4)
    # separation 4: the isotropic part weighs 0.2, the uncoupled bridge does not
    numpy.testing.assert_allclose(e4[:3], 0.2 * e5[:3], rtol=1e-12)
    assert e4[3] == e5[3] and e5[3] > 0
    # an occluder that is neither donor nor acceptor has no bridge terms
    assert _pair(T, I, J, WI, WJ, "OH", "CH2")[2:].tolist() == [0.0, 0.0]
    # a polar atom without waters scores nothing against anyone
    nan = numpy.full((4, 3), numpy.nan)
    assert _pair(T, I, J, nan, WJ, "OH", "OCbb").tolist() == [0.0] * 4


def test_pair_derivatives_central_differences(lkb_term):
    T = lkb_term
    rng = numpy.random.default_rng(11)
    gp = _gp(T)
    for trial in range(6):
        I = rng.normal(size=3)
        J = I + rng.normal(size=3) * 1.5 + 2.0
        WI = I + rng.normal(size=(3, 3)) * 1.6
        WJ = J + rng.normal(size=(2, 3)) * 1.6
        pi, pj = _tp(T, "OH"), _tp(T, "ONH2" if trial % 2 else "Nlys")
        WI4, WJ4 = orc._waters4(WI, numpy.float64), orc._waters4(WJ, numpy.float64)
        E, (dI, dJ, dWI, dWJ) = orc.oracle_lkb_pair(I, J, WI4, WJ4, 5, pi, pj, gp, deriv=True)
        h = 1e-6
        for name, arr, grad in (("I", I, dI), ("J", J, dJ), ("WI", WI4, dWI), ("WJ", WJ4, dWJ)):
            flat = arr.reshape(-1)
            for k in range(flat.size):
                if not numpy.isfinite(flat[k]):
                    continue
                old = flat[k]
                flat[k] = old + h
                ep = orc.oracle_lkb_pair(I, J, WI4, WJ4, 5, pi, pj, gp)
                flat[k] = old - h
                em = orc.oracle_lkb_pair(I, J, WI4, WJ4, 5, pi, pj, gp)
                flat[k] = old
                fd = (ep - em) / (2 * h)
                an = grad.reshape(4, -1)[:, k]
                assert numpy.abs(fd - an).max() < 2e-6 * max(1.0, numpy.abs(an).max()), (
                    trial, name, k, fd, an)


def test_pose_gradient_central_differences(ubq_pdb, lkb_term):
    """The whole chain (waters, pair derivatives, water backward) against
    central differences of the weighted energy, residues 0..4 of 1ubq
    (test_lk_ball_energy_term.py gradcheck on the same residues)."""
    ps = PoseStack([pose_from_pdb(ubq_pdb, (0, 4))], dtype=torch.float64)
    wa, sa = _args(ps, lkb_term)
    w = numpy.array([[1.0], [0.7], [1.3], [0.4]])
    _, _, dX, dW = orc.oracle_lkb_score(wa, sa, dTdV=w)
    g = dX + orc.oracle_lkb_waters_backward(wa, dW)
    X0 = ps.coords.flatten(0, -2).clone()

    def energy(X):
        wa2 = list(wa)
        wa2[0] = X
        W = torch.tensor(orc.oracle_lkb_waters(*wa2))
        sa2 = list(sa)
        sa2[0], sa2[26] = X, W
        return float((orc.oracle_lkb_score(wa2, sa2)[0] * w).sum())

    h = 1e-6
    rng = numpy.random.default_rng(2)
    atoms = rng.choice(X0.shape[0], 24, replace=False)
    for a in atoms:
        for k in range(3):
            Xp, Xm = X0.clone(), X0.clone()
            Xp[a, k] += h
            Xm[a, k] -= h
            fd = (energy(Xp) - energy(Xm)) / (2 * h)
            assert abs(fd - g[a, k]) < 1e-6 * max(1.0, abs(g[a, k])), (a, k, fd, g[a, k])


# ---------------------------------------------------------------- host tables


def test_block_type_tables(lkb_term):
    db = lkb_term.db
    ser = lkb_term.block_type(db.residue_type("SER"))
    rt = db.residue_type("SER")
    nm = rt.atom_names
    # donors: N (1 H) and OG (HG); acceptors: O (sp2) and OG (sp3)
    assert sorted(nm[d] for d in ser.don_hvy) == ["N", "OG"]
    assert sorted(nm[a] for a in ser.acc) == ["O", "OG"]
    # OG's waters: HG's donor water in slot 0, two sp3 waters after it
    rec = {(nm[r[0]], int(r[1])): r for r in ser.recipes}
    assert set(rec) == {("N", 0), ("OG", 0), ("OG", 1), ("OG", 2), ("O", 0), ("O", 1)}
    assert nm[rec[("OG", 1)][4]] == "CB" and nm[rec[("OG", 1)][5]] == "HG"
    assert nm[rec[("O", 0)][4]] == "C" and nm[rec[("O", 0)][5]] == "CA"
    # HIS: ND1 a ring acceptor built from CG and CE1; NE2 a donor
    his = db.residue_type("HIS")
    hb = lkb_term.block_type(his)
    nm = his.atom_names
    ring = [r for r in hb.recipes if r[2] == 3]
    assert [(nm[r[0]], nm[r[4]], nm[r[5]]) for r in ring] == [("ND1", "CG", "CE1")]
    # polar atoms listed before the other heavy atoms, every heavy atom once
    for name in db.canonical_names:
        rt = db.residue_type(name)
        b = lkb_term.block_type(rt)
        heavy = set(numpy.nonzero(b.is_hydrogen == 0)[0].tolist())
        assert set(b.occ_list) == heavy and len(b.occ_list) == len(heavy)
        assert all(b.is_polar[a] for a in b.occ_list[: b.n_polar])
        assert not any(b.is_polar[a] for a in b.occ_list[b.n_polar:])
        # at most four waters per atom, slots distinct
        slots = {}
        for r in b.recipes:
            slots.setdefault(int(r[0]), set()).add(int(r[1]))
        assert sum(len(v) for v in slots.values()) == len(b.recipes)
        assert all(max(v) < 4 for v in slots.values())


def test_waters_match_recipes(ubq_pdb, lkb_term):
    """The recipes the device path uses place every water the oracle's walk
    over the reference tables places, in the same slot."""
    ps = PoseStack([pose_from_pdb(ubq_pdb)], dtype=torch.float64)
    wa, _ = _args(ps, lkb_term)
    W = orc.oracle_lkb_waters(*wa)
    X = ps.coords.flatten(0, -2).numpy()
    built = numpy.full_like(W, numpy.nan)
    pbt = ps.packed_block_types
    off = ps.rot_coord_offset.numpy()
    bts = ps.block_type_ind_for_rot.numpy()
    wd, angs = lkb_term.watergen_values[0], lkb_term.watergen_values
    tors = {1: lkb_term.tors_sp2, 2: lkb_term.tors_sp3, 3: lkb_term.tors_ring}
    f32 = lambda v: float(numpy.float32(v))  # noqa: E731
    for g, bt in enumerate(bts):
        if bt < 0:
            continue
        rt = pbt.active_block_types[bt]
        for a, slot, kind, a1, a2, a3, ti, _ in lkb_term.block_type(rt).recipes:
            o = off[g]
            if kind == 0:
                built[o + a, slot] = orc.oracle_lkb_build_water(0, X[o + a1], X[o + a2], None,
                                                                f32(wd))
            else:
                Bx = X[o + a2] if kind != 3 else 0.5 * (X[o + a2] + X[o + a3])
                built[o + a, slot] = orc.oracle_lkb_build_water(
                    kind, X[o + a1], Bx, X[o + a3], f32(wd), f32(angs[kind]), f32(tors[kind][ti]))
    assert numpy.array_equal(numpy.isnan(built), numpy.isnan(W))
    numpy.testing.assert_allclose(built[~numpy.isnan(W)], W[~numpy.isnan(W)], rtol=0, atol=1e-12)


# ---------------------------------------------------------------- reference op

needs_ref = pytest.mark.skipif(
    not os.path.exists(os.path.join(os.path.dirname(orc.__file__), "_ref",
                                    "tmol_lk_ball_ref.so")),
    reason="oracle/_ref not built (needs the reference tree)",
)


@needs_ref
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_oracle_matches_reference_lk_ball_ops(ubq_pdb, lkb_term, dt):
    """Waters, energies (whole pose and block pair), neighbor flags and the
    full gradient (both ops' backward) against the reference's CPU ops."""
    from tmol_amd.generator import generate_poses

    ops = orc.load_reference_lk_ball_ops()
    poses = [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        3, (20, 60), seed=9)
    ps = PoseStack(poses, dtype=dt)
    tol = 1e-12 if dt == torch.float64 else 2e-5
    for bp in (False, True):
        wa, _ = _args(ps, lkb_term)
        X = ps.coords.flatten(0, -2).clone().requires_grad_(True)
        wa2 = list(wa)
        wa2[0] = X
        Wr = ops.gen_pose_waters(*wa2)
        Wo = orc.oracle_lkb_waters(*wa)
        assert numpy.array_equal(numpy.isnan(Wr.detach().numpy()), numpy.isnan(Wo))
        m = ~numpy.isnan(Wo)
        assert numpy.abs(Wr.detach().numpy()[m] - Wo[m]).max() <= tol * 10
        _, sa = _args(ps, lkb_term, bp, waters=Wr)
        sa2 = list(sa)
        sa2[0] = X
        Er, nbr = ops.lk_ball_pose_score(*sa2)
        sa_o = list(sa)
        sa_o[26] = Wr.detach()
        Eo, nbo, _, _ = orc.oracle_lkb_score(wa, sa_o)
        assert numpy.array_equal(nbr.numpy(), nbo)
        scale = max(1.0, numpy.abs(Eo).max())
        assert numpy.abs(Er.detach().numpy() - Eo).max() <= tol * scale
        dTdV = torch.rand(Er.shape, dtype=dt, generator=torch.Generator().manual_seed(4))
        Er.backward(dTdV)
        _, _, dX, dW = orc.oracle_lkb_score(wa, sa_o, dTdV=dTdV.numpy())
        g = dX + orc.oracle_lkb_waters_backward(wa, dW)
        gr = X.grad.numpy()
        assert numpy.abs(g - gr).max() <= tol * 10 * max(1.0, numpy.abs(gr).max())
