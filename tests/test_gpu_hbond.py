"""hbond on the GPU against the oracle: derived base coordinates, energies
(whole pose and block pair) and gradients through both drop-in ops, float64
and float32, tmol's 1ubq baselines, and the ragged scorer on poses of
uneven length."""

import numpy
import pytest
import torch

from conftest import load_term_baseline
from oracle import oracle as orc
from tmol_amd.generator import generate_poses
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack, RaggedPoseStack
from tmol_amd.score.hbond_energy_term import HBondEnergyTerm, bases_op_args, score_op_args

pytestmark = pytest.mark.gpu

# float64: kernels and oracle evaluate the same formulas in a different
# order; float32: hardware atan2 / acos / exp / log and float atomics in the
# gradient sums (the polynomials are evaluated in float64 on both sides)
TOL = {torch.float64: dict(e=1e-10, g=1e-9), torch.float32: dict(e=2e-5, g=2e-4)}


@pytest.fixture(scope="module")
def term():
    return HBondEnergyTerm()


def _poses(ubq_pdb, n_gen=4, seed=21):
    return [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        n_gen, (30, 120), seed=seed)


def _dev(args, dev):
    return [a.to(dev) if torch.is_tensor(a) else a for a in args]


def _host(args):
    return [a.cpu() if torch.is_tensor(a) else a for a in args]


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_bases_match_oracle(ubq_pdb, term, dt):
    from tmol_amd.hbond import gen_hbond_bases

    print("hbond bases", dt, flush=True)
    ps = PoseStack(_poses(ubq_pdb), dtype=dt)
    ba = bases_op_args(ps, term)
    D, I = gen_hbond_bases(*_dev(ba, "cuda"))
    Do, Io = orc.oracle_hb_bases(*ba)
    assert numpy.array_equal(I.cpu().numpy(), Io)
    Dn = D.cpu().numpy()
    assert numpy.array_equal(numpy.isnan(Dn), numpy.isnan(Do))
    m = ~numpy.isnan(Do)
    assert m.sum() > 1000
    assert numpy.array_equal(Dn[m], Do[m])  # copies of coordinates or zeros


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("block_pair", [False, True])
def test_pose_score_and_gradients_match_oracle(ubq_pdb, term, dt, block_pair):
    """Energies, neighbor flags and the weighted coordinate gradient of the
    drop-in ops against the oracle."""
    from tmol_amd.hbond import gen_hbond_bases, hbond_pose_scores

    print("hbond score", dt, "block_pair", block_pair, flush=True)
    ps = PoseStack(_poses(ubq_pdb), dtype=dt)
    ba = bases_op_args(ps, term)
    D, I = gen_hbond_bases(*_dev(ba, "cuda"))
    sa = _dev(score_op_args(ps, term, D, I, block_pair=block_pair), "cuda")
    X = sa[0].clone().requires_grad_(True)
    sa2 = list(sa)
    sa2[0] = X
    E, nbr = hbond_pose_scores(*sa2)
    sa_h = _host(sa)
    Eo, nbo, _ = orc.oracle_hb_score(sa_h, ba)
    assert numpy.array_equal(nbr.cpu().numpy(), nbo)
    assert (Eo != 0).sum() >= ps.n_poses
    scale = max(1.0, numpy.abs(Eo).max())
    assert numpy.abs(E.detach().cpu().numpy() - Eo).max() <= TOL[dt]["e"] * scale
    dTdV = torch.rand(E.shape, dtype=dt, generator=torch.Generator().manual_seed(6))
    E.backward(dTdV.cuda())
    _, _, dX = orc.oracle_hb_score(sa_h, ba, dTdV=dTdV.numpy())
    gr = X.grad.cpu().numpy()
    assert numpy.abs(dX - gr).max() <= TOL[dt]["g"] * max(1.0, numpy.abs(dX).max())


def test_goldens_on_gpu(ubq_pdb, term):
    """tmol's 1ubq hbond baselines through the drop-in ops in float32:
    whole pose, and block pairs of residues 6-7 and 10-11."""
    from tmol_amd.hbond import gen_hbond_bases, hbond_pose_scores

    ps = PoseStack([pose_from_pdb(ubq_pdb)] * 3)
    D, I = gen_hbond_bases(*_dev(bases_op_args(ps, term), "cuda"))
    E, _ = hbond_pose_scores(*_dev(score_op_args(ps, term, D, I), "cuda"))
    gold = load_term_baseline("test_whole_pose_scoring_10", "hbond")[:, :3]
    numpy.testing.assert_allclose(E.cpu().numpy(), gold, atol=1e-5, rtol=1e-5)
    ps = PoseStack([pose_from_pdb(ubq_pdb, [(6, 8), (10, 12)])])
    D, I = gen_hbond_bases(*_dev(bases_op_args(ps, term), "cuda"))
    E, _ = hbond_pose_scores(*_dev(score_op_args(ps, term, D, I, block_pair=True), "cuda"))
    numpy.testing.assert_allclose(E.cpu().numpy(), load_term_baseline("test_block_scoring",
                                                                      "hbond"),
                                  atol=1e-5, rtol=1e-5)


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_ragged_scorer_matches_oracle(term, dt):
    """The ragged scorer (own neighbor pass, bases, WGRAD score) against the
    oracle on the padded stack of the same poses."""
    from tmol_amd.hbond import HBondRaggedScorer

    print("hbond ragged", dt, flush=True)
    poses = generate_poses(6, (20, 260), seed=5, log_uniform=True)
    rs = RaggedPoseStack(poses, dtype=dt)
    sc = HBondRaggedScorer(rs, term, device="cuda", dtype=dt)
    X = rs.coords.cuda()
    w = torch.rand(len(poses), dtype=dt, generator=torch.Generator().manual_seed(8))
    E, G = sc.energy_and_grad(X, w.cuda())
    ps = PoseStack(poses, dtype=dt)
    ba = bases_op_args(ps, term)
    Do, Io = orc.oracle_hb_bases(*ba)
    sa = score_op_args(ps, term, torch.tensor(Do), torch.tensor(Io))
    Eo, _, dX = orc.oracle_hb_score(sa, ba, dTdV=w.numpy()[None, :])
    scale = max(1.0, numpy.abs(Eo).max())
    assert numpy.abs(E.cpu().numpy() - Eo[0]).max() <= TOL[dt]["e"] * scale
    idx = []
    for p, pose in enumerate(poses):
        base = p * ps.max_n_atoms
        idx.extend(range(base, base + pose.n_atoms))
    go = dX[numpy.array(idx)]
    assert numpy.abs(G.cpu().numpy() - go).max() <= TOL[dt]["g"] * max(1.0, numpy.abs(go).max())
    E2 = sc.energies_only(X).clone()
    assert torch.allclose(E2, E, rtol=1e-12 if dt == torch.float64 else 1e-5, atol=1e-6)
    c = sc.census(X)
    assert c["candidates"] > c["evaluated"] > 0 and c["cp_excluded"] > 0


def test_single_block_and_glycine_poses(term):
    """Edge cases: a one-residue pose (intra-block pairs only), a glycine
    chain (backbone donors and acceptors only) and a short mixed chain."""
    from tmol_amd.chemical import default_chemical_database
    from tmol_amd.generator import build_chain, sample_torsions
    from tmol_amd.hbond import gen_hbond_bases, hbond_pose_scores
    from tmol_amd.pose import Pose

    db = default_chemical_database()
    rng = numpy.random.default_rng(3)
    poses = []
    for names in (["TRP"], ["GLY"] * 12, ["SER", "HIS", "ASP", "ASN", "TYR"]):
        rts = [db.residue_type(n, nterm=i == 0, cterm=i == len(names) - 1)
               for i, n in enumerate(names)]
        phis, psis, omegas, chis = sample_torsions(rts, rng)
        poses.append(Pose(rts, build_chain(rts, phis, psis, omegas, chis)))
    ps = PoseStack(poses, dtype=torch.float64)
    ba = bases_op_args(ps, term)
    D, I = gen_hbond_bases(*_dev(ba, "cuda"))
    E, _ = hbond_pose_scores(*_dev(score_op_args(ps, term, D, I), "cuda"))
    Eo, _, _ = orc.oracle_hb_score(score_op_args(ps, term, D.cpu(), I.cpu()), ba)
    assert numpy.abs(E.cpu().numpy() - Eo).max() <= 1e-10 * max(1.0, numpy.abs(Eo).max())
