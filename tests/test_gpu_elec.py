"""GPU parity of the elec term (``tmol_amd.ops.elec_pose_scores``, the
drop-in for tmol_elec::elec_pose_scores) through the C ABI:

* tmol's elec baselines for 1ubq through the term's scoring modules;
* neighbor flags bit-identical to the oracle; energies and dV/dx within
  1e-10 (f64) and 1e-4 of max(1, |E|) (f32) of the f64 oracle, which is
  itself bit-identical to the reference CPU op in f64;
* block-pair energies against the oracle, and their weighted backward by
  its defining properties (all-ones weights give the whole-pose gradient,
  linearity in the weights, central differences of the f64 kernels);
* the weighted-gradient mode against dV/dx scaled per pose;
* the 64-pose padded drop-in workload: f32 against f64 per pose, sampled
  poses against the oracle;
* a ScoreFunction with LJ+LK and elec weights.
"""

import sys

import numpy
import pytest
import torch

from conftest import load_term_baseline
from oracle import oracle as orc
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack
from tmol_amd.score.elec_energy_term import ElecEnergyTerm, elec_op_args

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
TERM = ElecEnergyTerm()


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _args(ps, dt, device=None, block_pair=False):
    a = elec_op_args(ps, TERM, block_pair=block_pair)
    if dt == torch.float64:
        a[16], a[22] = a[16].double(), a[22].double()
    if device is not None:
        a = [x.to(device) if isinstance(x, torch.Tensor) else x for x in a]
    return a


def _poses(ubq_pdb):
    from tmol_amd.generator import generate_poses

    return [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        4, (20, 120), seed=11)


def test_ubq_term_baselines(ubq_pdb):
    from tmol_amd.ops import elec_pose_scores  # noqa: F401

    _say("elec ubq baselines")
    ps = PoseStack([pose_from_pdb(ubq_pdb)] * 10, device=DEV)
    m = TERM.render_whole_pose_scoring_module(ps)
    s = m(ps.coords).detach().cpu().numpy()
    numpy.testing.assert_allclose(s, load_term_baseline("test_whole_pose_scoring_10", "elec"),
                                  atol=1e-5, rtol=1e-3)
    ps = PoseStack([pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (0, 50)),
                    pose_from_pdb(ubq_pdb, (20, 50))], device=DEV)
    s = TERM.render_whole_pose_scoring_module(ps)(ps.coords).detach().cpu().numpy()
    numpy.testing.assert_allclose(s, load_term_baseline("test_whole_pose_scoring_jagged", "elec"),
                                  atol=1e-5, rtol=1e-3)
    ps = PoseStack([pose_from_pdb(ubq_pdb, (0, 4))], device=DEV)
    s = TERM.render_block_pair_scoring_module(ps)(ps.coords).detach().cpu().numpy()
    numpy.testing.assert_allclose(s, load_term_baseline("test_block_scoring", "elec"),
                                  atol=1e-5, rtol=1e-3)
    ps = PoseStack([pose_from_pdb(ubq_pdb, (0, 3))], device=DEV)
    s = TERM.render_whole_pose_scoring_module(ps)(ps.coords).detach().cpu().numpy()
    numpy.testing.assert_allclose(s, [[-0.419618]], atol=1e-5, rtol=1e-5)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_energies_gradients_neighbors_match_oracle(ubq_pdb, dt):
    from tmol_amd.ops import elec_pose_scores

    _say("elec vs oracle %s" % dt)
    poses = _poses(ubq_pdb)
    ps = PoseStack(poses, dtype=dt)
    ps64 = PoseStack(poses, dtype=torch.float64)
    E64, _, dv64, _ = orc.oracle_elec_scores(*_args(ps64, torch.float64), require_grad=True)
    # neighbor flags come from spheres in the coordinate dtype: oracle in the same dtype
    _, nbr_o, _, _ = orc.oracle_elec_scores(*_args(ps, dt))
    a = _args(ps, dt, DEV)
    c = a[0].clone().requires_grad_(True)
    a[0] = c
    s, nbr = elec_pose_scores(*a)
    assert tuple(s.shape) == (1, len(poses))
    numpy.testing.assert_array_equal(nbr.cpu().numpy(), nbr_o)
    tol = 1e-10 if dt == torch.float64 else 1e-4
    E = s.detach().double().cpu().numpy()
    err = numpy.abs(E - E64) / numpy.maximum(1.0, numpy.abs(E64))
    _say("max rel energy err %.3g" % err.max())
    assert err.max() < tol
    s.sum().backward()
    g = c.grad.double().cpu().numpy()
    gerr = numpy.abs(g - dv64[0]).max() / max(1.0, numpy.abs(dv64).max())
    _say("grad max err / scale %.3g" % gerr)
    assert gerr < tol


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_block_pair_forward_and_backward(ubq_pdb, dt):
    from tmol_amd.ops import elec_pose_scores

    _say("elec block pair %s" % dt)
    poses = _poses(ubq_pdb)[:4]
    ps = PoseStack(poses, dtype=dt)
    ps64 = PoseStack(poses, dtype=torch.float64)
    bp64, _, _, _ = orc.oracle_elec_scores(*_args(ps64, torch.float64, block_pair=True))
    a = _args(ps, dt, DEV, block_pair=True)
    c = a[0].clone().requires_grad_(True)
    a[0] = c
    s, _ = elec_pose_scores(*a)
    tol = 1e-10 if dt == torch.float64 else 1e-4
    err = numpy.abs(s.detach().double().cpu().numpy() - bp64).max() / max(1.0, numpy.abs(bp64).max())
    assert err < tol
    # all-ones weights: the whole-pose gradient
    s.backward(torch.ones_like(s))
    g1 = c.grad.clone()
    aw = _args(ps, dt, DEV)
    cw = aw[0].clone().requires_grad_(True)
    aw[0] = cw
    sw, _ = elec_pose_scores(*aw)
    sw.sum().backward()
    scale = max(1.0, cw.grad.abs().max().item())
    assert (g1 - cw.grad).abs().max().item() < tol * scale
    # linearity in the weights
    gen = torch.Generator().manual_seed(5)
    w1 = torch.rand(s.shape, generator=gen, dtype=torch.float64).to(DEV, dt)
    w2 = torch.rand(s.shape, generator=gen, dtype=torch.float64).to(DEV, dt)
    grads = []
    for w in (w1, w2, w1 + w2):
        c.grad = None
        s, _ = elec_pose_scores(*a)
        s.backward(w)
        grads.append(c.grad.clone())
    assert (grads[0] + grads[1] - grads[2]).abs().max().item() < tol * scale
    if dt == torch.float64:
        # central differences of sum(w1 * E_bp) on a few atoms
        x0 = a[0].detach().clone()
        rng = numpy.random.default_rng(2)
        for atom in rng.choice(x0.shape[0], 6, replace=False):
            for k in range(3):
                h = 1e-5
                vals = []
                for sgn in (1, -1):
                    x = x0.clone()
                    x[atom, k] += sgn * h
                    ap = list(a)
                    ap[0] = x
                    vals.append((elec_pose_scores(*ap)[0] * w1).sum().item())
                fd = (vals[0] - vals[1]) / (2 * h)
                assert grads[0][atom, k].item() == pytest.approx(fd, rel=1e-5, abs=1e-7)


def test_weighted_gradient_mode(ubq_pdb):
    """WGRAD with per-pose weights equals dV/dx scaled by the pose weight."""
    from tmol_amd.ops import DVDX, WGRAD, ElecPaddedCall

    _say("elec wgrad")
    ps = PoseStack(_poses(ubq_pdb), dtype=torch.float64)
    a = _args(ps, torch.float64, DEV)
    call = ElecPaddedCall(a, torch.float64)
    call.neighbors(None)
    P = call.P
    e = torch.empty((1, P), dtype=torch.float64, device=DEV)
    dvdx = torch.empty((1, a[0].shape[0], 3), dtype=torch.float64, device=DEV)
    call.score(DVDX, energies=e, dvdx=dvdx)
    w = torch.linspace(0.5, 2.0, P, dtype=torch.float64, device=DEV)
    grad = torch.empty_like(call.coords)
    e2 = torch.empty((1, P), dtype=torch.float64, device=DEV)
    call.score(WGRAD, weights=w, energies=e2, grad=grad)
    want = dvdx[0] * w[a[2].long()].unsqueeze(-1)
    assert (grad - want).abs().max().item() < 1e-10 * max(1.0, want.abs().max().item())
    assert (e - e2).abs().max().item() < 1e-10 * max(1.0, e.abs().max().item())


def test_padded_64_pose_workload():
    """The 64-pose padded drop-in workload (150..600 residues): f32 against
    f64 per pose, and 3 sampled poses against the oracle."""
    from tmol_amd.generator import workload_poses
    from tmol_amd.ops import elec_pose_scores

    _say("elec padded_64x150_600")
    poses = workload_poses("padded_64x150_600", 64)
    out = {}
    for dt in (torch.float64, torch.float32):
        ps = PoseStack(poses, dtype=dt, device=DEV)
        a = _args(ps, dt)
        c = a[0].clone().requires_grad_(True)
        a[0] = c
        s, nbr = elec_pose_scores(*a)
        s.sum().backward()
        out[dt] = (s.detach().double().cpu().numpy(), c.grad.double().cpu().numpy(),
                   nbr.cpu().numpy())
    E64, G64, N64 = out[torch.float64]
    E32, G32, N32 = out[torch.float32]
    # Elec totals are sums of block-pair terms of both signs that cancel
    # (pose totals down to 0.25 against sum |E_bp| ~ 2000 here), so float32
    # rounding is bounded relative to S = sum |E_bp|: the reference's own
    # float32 arithmetic (the f32 oracle) reaches 2.1e-7 S on this workload.
    # Bound: 1e-4 max(1, |E|), or 1e-6 S where that is larger.
    ps = PoseStack(poses, dtype=torch.float64, device=DEV)
    bp, _ = elec_pose_scores(*_args(ps, torch.float64, block_pair=True))
    S = bp.abs().sum((2, 3)).cpu().numpy()
    bound = numpy.maximum(1e-4 * numpy.maximum(1.0, numpy.abs(E64)), 1e-6 * S)
    err = numpy.abs(E32 - E64)
    _say("f32 vs f64: max err / bound %.3g, max rel %.3g"
         % ((err / bound).max(), (err / numpy.maximum(1.0, numpy.abs(E64))).max()))
    assert (err <= bound).all()
    assert numpy.linalg.norm(G32 - G64) < 1e-4 * numpy.linalg.norm(G64)
    for k in (0, 17, 63):
        ps1 = PoseStack([poses[k]], dtype=torch.float64)
        Eo, _, _, _ = orc.oracle_elec_scores(*_args(ps1, torch.float64))
        assert abs(E64[0, k] - Eo[0, 0]) < 1e-10 * max(1.0, abs(Eo[0, 0]))


def test_score_function_with_ljlk_and_elec(ubq_pdb):
    from tmol_amd.score import ScoreFunction, ScoreType
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

    _say("elec score function")
    ps = PoseStack([pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (10, 40))], device=DEV)
    sfxn = ScoreFunction(device=DEV)
    for st, w in ((ScoreType.fa_ljatr, 1.0), (ScoreType.fa_ljrep, 0.55), (ScoreType.fa_lk, 1.0),
                  (ScoreType.fa_elec, 1.0)):
        sfxn.set_weight(st, w)
    m = sfxn.render_whole_pose_scoring_module(ps)
    coords = ps.coords.clone().requires_grad_(True)
    tot = m(coords)
    per = m(coords, sum_terms=False, apply_weights=False).detach().cpu().numpy()
    assert per.shape == (4, 2)
    # beta2016 values of 1ubq (test_score_function.py)
    numpy.testing.assert_allclose(per[:, 0], [-417.02362, 240.7147, 301.93347, -134.03497],
                                  rtol=1e-4, atol=1e-4)
    want = per[0] + 0.55 * per[1] + per[2] + per[3]
    numpy.testing.assert_allclose(tot.detach().cpu().numpy(), want, rtol=1e-5)
    tot.sum().backward()
    assert torch.isfinite(coords.grad).all()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_ragged_scorer_matches_oracle(ubq_pdb, dtype):
    """The ragged batch path (count pair from chain positions, one work
    list over all poses): per-pose energies and the weighted gradient against
    the oracle on the padded stack of the same poses."""
    from tmol_amd.generator import generate_poses
    from tmol_amd.pose import RaggedPoseStack
    from tmol_amd.scorer import ElecRaggedScorer

    _say("elec ragged %s" % dtype)
    poses = [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        6, (10, 150), seed=5, log_uniform=True)
    stack = RaggedPoseStack(poses, dtype=dtype)
    sc = ElecRaggedScorer(stack, TERM, device=DEV, dtype=dtype)
    x = stack.coords.to(DEV)
    w = torch.linspace(0.5, 1.5, len(poses), dtype=torch.float64)
    e, g = sc.energy_and_grad(x, w)
    e = e.double().cpu().numpy()
    g = g.double().cpu().numpy()
    ps = PoseStack(poses, dtype=torch.float64)
    E64, _, dv64, _ = orc.oracle_elec_scores(*_args(ps, torch.float64), require_grad=True)
    A = ps.max_n_atoms
    keep = numpy.concatenate([p * A + numpy.arange(pose.n_atoms) for p, pose in enumerate(poses)])
    pose_of_atom = numpy.concatenate([numpy.full(pose.n_atoms, p) for p, pose in enumerate(poses)])
    og = dv64[0, keep] * w.numpy()[pose_of_atom][:, None]
    tol = 1e-10 if dtype == torch.float64 else 1e-4
    assert (numpy.abs(e - E64[0]) / numpy.maximum(1.0, numpy.abs(E64[0]))).max() < tol
    assert numpy.abs(g - og).max() / max(1.0, numpy.abs(og).max()) < tol
    # energies-only pass agrees with the fused one
    e2 = sc.energies_only(x).double().cpu().numpy()
    assert numpy.abs(e2 - e).max() <= 1e-12 * max(1.0, numpy.abs(e).max()) or dtype == torch.float32


def test_elec_census_counts(ubq_pdb):
    """Census of the ragged work list: pair counts from the block sizes,
    charged-pair counts from the charge tables, and the evaluated pairs
    inside max_dis against a brute-force count."""
    from tmol_amd.flops import elec_far_pairs
    from tmol_amd.pose import RaggedPoseStack
    from tmol_amd.scorer import ElecRaggedScorer

    _say("elec census")
    poses = [pose_from_pdb(ubq_pdb)]
    stack = RaggedPoseStack(poses)
    sc = ElecRaggedScorer(stack, TERM, device=DEV)
    x = stack.coords.to(DEV)
    c = sc.census(x)
    pose = poses[0]
    nb = pose.n_blocks
    q = [TERM.setup_block_type(rt)[0] for rt in pose.residue_types]
    n = [len(qq) for qq in q]
    nq = [int((qq != 0).sum()) for qq in q]
    cnt, items = sc.ws.rows(sc.cstack, sc.tables.cbt)
    cnt = cnt.cpu().numpy()
    items = items.cpu().numpy()
    want_pairs = want_charged = n_inter = 0
    for g1 in range(nb):
        for k in range(min(cnt[g1], sc.capacity)):
            g2 = items[g1, k]
            if g2 == g1:
                want_pairs += n[g1] * (n[g1] - 1) // 2
                want_charged += nq[g1] * (nq[g1] - 1) // 2
            else:
                n_inter += 1
                want_pairs += n[g1] * n[g2]
                want_charged += nq[g1] * nq[g2]
    assert c["items"] == n_inter and c["intra_items"] == nb
    assert c["pairs"] == want_pairs and c["charged"] == want_charged
    # brute force: charged pairs of count-pair weight > 0 closer than max_dis,
    # separations from the representative path distances and the pose's
    # connection-atom separations (elec_pose_score.impl.hh:47-66)
    ps = PoseStack(poses, dtype=torch.float64)
    xyz = ps.coords[0].numpy()
    tabs = [TERM.setup_block_type(rt) for rt in pose.residue_types]
    conn = [rt.conn_atom for rt in pose.residue_types]
    ibs = ps.inter_block_bondsep[0].numpy()
    mins = ps.min_block_bondsep[0].numpy()
    off = numpy.concatenate([[0], numpy.cumsum(n)])
    weighted = 0
    for b1 in range(nb):
        for b2 in range(b1, nb):
            d = numpy.linalg.norm(xyz[off[b1]:off[b1 + 1], None] - xyz[None, off[b2]:off[b2 + 1]],
                                  axis=-1)
            m = (q[b1][:, None] != 0) & (q[b2][None, :] != 0) & (d < 5.5)
            if b1 == b2:
                sep = tabs[b1][2]
                m &= numpy.triu(numpy.ones_like(m, dtype=bool), 1)
            elif mins[b1, b2] > 4:
                sep = numpy.full(m.shape, mins[b1, b2])
            else:
                sep = numpy.full(m.shape, 5)
                for c1, a1 in enumerate(conn[b1]):
                    for c2, a2 in enumerate(conn[b2]):
                        v = (tabs[b1][1][a1][:, None] + ibs[b1, b2, c1, c2]
                             + tabs[b2][1][a2][None, :])
                        sep = numpy.minimum(sep, v)
            weighted += int((m & (sep >= 4)).sum())
    evaluated = c["flat"] + c["low"] + c["coulomb"] + c["high"]
    _say("census %s; brute force %d" % (c, weighted))
    # float32 distances may move a pair across max_dis: allow a handful
    assert abs(evaluated - weighted) <= 3
    assert elec_far_pairs(c) >= 0
