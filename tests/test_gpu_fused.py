"""GPU tests of the fused LJ+LK+elec pass (``LJLKElecRaggedScorer``,
``tmol_ljlk_elec_score``): one neighbor pass and one scoring launch give the
same per-pose energies and weighted gradient as the two terms scored
separately, and as the two oracles; hipGraph replays match eager calls."""

import sys

import numpy
import pytest
import torch

from oracle import oracle as orc
from tmol_amd.generator import generate_poses, workload_poses
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack, RaggedPoseStack
from tmol_amd.score.elec_energy_term import ElecEnergyTerm, elec_op_args
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
W4 = numpy.array([1.0, 0.55, 1.0, 0.8])


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _poses(ubq_pdb):
    return [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        6, (10, 150), seed=21, log_uniform=True)


def _scorers(poses, dtype):
    from tmol_amd.scorer import ElecRaggedScorer, LJLKElecRaggedScorer, LJLKRaggedScorer

    st = RaggedPoseStack(poses, dtype=dtype)
    lj = LJLKEnergyTerm(device=DEV)
    el = ElecEnergyTerm()
    return (st, LJLKElecRaggedScorer(st, lj, el, device=DEV, dtype=dtype),
            LJLKRaggedScorer(st, lj, device=DEV, dtype=dtype),
            ElecRaggedScorer(st, el, device=DEV, dtype=dtype))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_equals_separate_terms_and_oracles(ubq_pdb, dtype):
    poses = _poses(ubq_pdb)
    _say("fused vs separate %s" % dtype)
    st, fused, lj, el = _scorers(poses, dtype)
    x = st.coords.to(DEV)
    w = torch.tensor(W4)
    e4, g4 = [t.double().cpu().numpy() for t in fused.energy_and_grad(x, w)]
    e3, g3 = [t.double().cpu().numpy() for t in lj.energy_and_grad(x, w[:3])]
    e1, g1 = [t.double().cpu().numpy() for t in el.energy_and_grad(x, float(W4[3]))]
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    scale = numpy.maximum(1.0, numpy.abs(e4))
    assert (numpy.abs(e4[:3] - e3) / scale[:3]).max() < tol
    assert (numpy.abs(e4[3] - e1) / scale[3]).max() < tol
    g = g3 + g1
    assert numpy.abs(g4 - g).max() / max(1.0, numpy.abs(g).max()) < tol
    # the gradient-only pass gives the same gradient
    grad = torch.empty_like(fused.grad)
    fused.neighbors(x)
    fused.score(x, 4 | 16, weights=fused._weights(w), grad=grad)
    assert numpy.abs(grad.double().cpu().numpy() - g4).max() <= 1e-12 * max(1.0, numpy.abs(g4).max()) \
        or dtype == torch.float32
    if dtype == torch.float64:
        # both oracles on the padded stack of the same poses
        ps = PoseStack(poses, dtype=torch.float64)
        a = pose_op_args(ps, LJLKEnergyTerm())
        a[23], a[24] = a[23].double(), a[24].double()
        s3, _, dv3, _ = orc.oracle_pose_scores(*a, require_grad=True)
        b = elec_op_args(ps, ElecEnergyTerm())
        b[16], b[22] = b[16].double(), b[22].double()
        s1, _, dv1, _ = orc.oracle_elec_scores(*b, require_grad=True)
        A = ps.max_n_atoms
        keep = numpy.concatenate([p * A + numpy.arange(pose.n_atoms) for p, pose in enumerate(poses)])
        og = (dv3[:, keep] * W4[:3, None, None]).sum(0) + W4[3] * dv1[0, keep]
        oe = numpy.concatenate([s3, s1], axis=0)
        assert (numpy.abs(e4 - oe) / numpy.maximum(1.0, numpy.abs(oe))).max() < 1e-9
        assert numpy.abs(g4 - og).max() / max(1.0, numpy.abs(og).max()) < 1e-9


def test_fused_graph_replays_match_eager(ubq_pdb):
    poses = _poses(ubq_pdb)
    _say("fused graph replay")
    st, fused, _, _ = _scorers(poses, torch.float32)
    x = st.coords.to(DEV)
    e0, g0 = [t.clone() for t in fused.energy_and_grad(x)]
    fused.capture(x)
    for k in range(3):
        e, g = fused.replay()
        torch.cuda.synchronize()
        assert torch.equal(e, e0) or (e - e0).abs().max().item() < 1e-4
        assert (g - g0).abs().max().item() <= 1e-5 * max(1.0, g0.abs().max().item())
        _ = fused.energy_and_grad(x)  # eager work between replays
        torch.cuda.synchronize()


def test_fused_on_the_512_pose_batch():
    """The headline batch: fused energies and gradient against the two
    separate passes (float32)."""
    poses = workload_poses("ragged_512x50_1200")
    _say("fused 512 poses")
    st, fused, lj, el = _scorers(poses, torch.float32)
    x = st.coords.to(DEV)
    e4, g4 = [t.double() for t in fused.energy_and_grad(x)]
    e3, g3 = [t.double() for t in lj.energy_and_grad(x, torch.tensor([1.0, 0.55, 1.0]))]
    e1, g1 = [t.double() for t in el.energy_and_grad(x, 1.0)]
    assert (e4[:3] - e3).abs().max().item() < 1e-5 * max(1.0, e3.abs().max().item())
    assert (e4[3] - e1).abs().max().item() < 1e-4
    g = g3 + g1
    assert ((g4 - g).norm() / g.norm()).item() < 1e-6
