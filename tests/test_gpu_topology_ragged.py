"""GPU parity of the ragged scorer -- the headline batched path -- on poses
whose bond topology is not a linear chain (tests/topology_cases.py: a
head-to-tail cycle, a disulfide, an inter-chain disulfide cross-link, and a
linear control), against the oracle on the padded stack of the same poses:

* count-pair separations from the kernel's own code, for every block pair
  within 5 bonds and a sample of the others: bit-exact;
* per-pose energies and the weighted gradient: f64 within 1e-9 (energies)
  and 1e-8 (gradient), f32 within 1e-4 relative of the f64 oracle;
* the extra bonds matter: the same poses scored as open chains differ.
"""

import sys

import numpy
import pytest
import torch

from oracle import oracle as orc
from tmol_amd.pose import Pose, PoseStack, RaggedPoseStack
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args
from topology_cases import general_topology_poses

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
W = numpy.array([1.0, 0.55, 1.0])


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _oracle(poses):
    ps = PoseStack(poses, dtype=torch.float64)
    a = pose_op_args(ps, LJLKEnergyTerm())
    a[23], a[24] = a[23].double(), a[24].double()
    s, _, dv, _ = orc.oracle_pose_scores(*a, require_grad=True, dtype=numpy.float64)
    A = ps.max_n_atoms
    keep = numpy.concatenate([p * A + numpy.arange(pose.n_atoms) for p, pose in enumerate(poses)])
    return s, dv[:, keep], a


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_ragged_general_topology_energy_and_grad(dtype):
    from tmol_amd.scorer import LJLKRaggedScorer

    poses = general_topology_poses()
    _say("ragged general topology %s" % dtype)
    stack = RaggedPoseStack(poses, dtype=dtype)
    sc = LJLKRaggedScorer(stack, LJLKEnergyTerm(device=DEV), device=DEV, dtype=dtype)
    x = stack.coords.to(DEV)
    e, g = sc.energy_and_grad(x, torch.tensor(W))
    e = e.double().cpu().numpy()
    g = g.double().cpu().numpy()
    s64, dv64, _ = _oracle(poses)
    og = (dv64 * W[:, None, None]).sum(0)
    if dtype == torch.float64:
        numpy.testing.assert_allclose(e, s64, rtol=1e-9, atol=1e-9)
        numpy.testing.assert_allclose(g, og, rtol=1e-8, atol=1e-8)
    else:
        rel = numpy.abs(e - s64) / numpy.maximum(numpy.abs(s64), 1.0)
        assert rel.max() < 1e-4, rel.max()
        assert numpy.abs(g - og).max() / numpy.abs(og).max() < 1e-5
    # without their extra bonds the cycle, disulfide and cross-link poses
    # score differently (the count pair of the bonded atoms changes)
    open_poses = [Pose(p.residue_types, p.coords, connected_to_prev=p.connected_to_prev)
                  for p in poses]
    s_open, _, _ = _oracle(open_poses)
    for p in range(3):
        assert abs(s_open[:, p] - s64[:, p]).max() > 1e-3, p
    numpy.testing.assert_allclose(s_open[:, 3], s64[:, 3], rtol=0, atol=0)


def test_ragged_general_topology_separations_bit_exact():
    from tmol_amd.scorer import LJLKRaggedScorer

    poses = general_topology_poses()
    _say("ragged general topology separations")
    stack = RaggedPoseStack(poses, dtype=torch.float64)
    sc = LJLKRaggedScorer(stack, LJLKEnergyTerm(device=DEV), device=DEV, dtype=torch.float64)
    _, _, args = _oracle(poses)
    pbs = stack.pose_block_start.numpy()
    ps = PoseStack(poses)
    mins = ps.min_block_bondsep.numpy()
    rng = numpy.random.default_rng(3)
    trip = []
    for p, pose in enumerate(poses):
        nb = pose.n_blocks
        near = [(p, b1, b2) for b1 in range(nb) for b2 in range(b1, nb) if mins[p, b1, b2] <= 5]
        far = [(p, int(b1), int(b2)) for b1, b2 in rng.integers(0, nb, (20, 2)) if b1 <= b2]
        trip += near + far
    items = torch.tensor([[pbs[p] + b1, pbs[p] + b2] for p, b1, b2 in trip], dtype=torch.int32)
    seps = sc.debug_pair_separations(stack.coords.to(DEV), items).cpu().numpy()
    ref = orc.oracle_pair_separations(args, trip)
    n_cp = 0
    for k, m in enumerate(ref):
        n1, n2 = m.shape
        assert (seps[k, :n1, :n2] == m).all(), trip[k]
        n_cp += int((m <= 4).sum())
    assert n_cp > 1000
