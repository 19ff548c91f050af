"""GPU parity: the HIP op against the C oracle and tmol's goldens.

Bars: neighbor flags, work list and count-pair separations bit-exact against
the oracle; energies within the reference's golden tolerances and within
1e-4 relative of the FP64 run of the same kernels; gradients against the
FP64 run, the oracle and finite differences (tolerances stated per test).
"""

import sys

import numpy
import pytest
import torch

from conftest import load_term_baseline
from oracle import oracle as orc
from tmol_amd.generator import generate_poses
from tmol_amd.io_pdb import pose_from_pdb
from tmol_amd.pose import PoseStack
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _stacks(ubq_pdb):
    poses = [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (20, 50))] + generate_poses(
        4, (30, 120), seed=11
    )
    return poses


def _gpu_args(poses, dtype=torch.float32, block_pair=False, term=None):
    term = term or LJLKEnergyTerm(device=DEV)
    ps = PoseStack(poses, device=DEV, dtype=dtype)
    args = pose_op_args(ps, term, block_pair=block_pair)
    if dtype == torch.float64:
        args[23] = args[23].double()
        args[24] = args[24].double()
    return ps, args


def _cpu_args(poses, dtype=torch.float32, block_pair=False):
    term = LJLKEnergyTerm()
    ps = PoseStack(poses, dtype=dtype)
    args = pose_op_args(ps, term, block_pair=block_pair)
    if dtype == torch.float64:
        args[23] = args[23].double()
        args[24] = args[24].double()
    return ps, args


def test_goldens_on_gpu(ubq_pdb):
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    _say("goldens whole_pose_10")
    _, args = _gpu_args([pose_from_pdb(ubq_pdb)] * 10)
    s, _ = ljlk_pose_scores(*args)
    numpy.testing.assert_allclose(
        s.cpu().numpy(), load_term_baseline("test_whole_pose_scoring_10"), atol=1e-5, rtol=1e-3
    )
    _say("goldens jagged")
    _, args = _gpu_args(
        [pose_from_pdb(ubq_pdb), pose_from_pdb(ubq_pdb, (0, 50)), pose_from_pdb(ubq_pdb, (20, 50))]
    )
    s, _ = ljlk_pose_scores(*args)
    numpy.testing.assert_allclose(
        s.cpu().numpy(), load_term_baseline("test_whole_pose_scoring_jagged"), atol=1e-5, rtol=1e-3
    )
    _say("goldens block pair")
    _, args = _gpu_args([pose_from_pdb(ubq_pdb, (0, 4))], block_pair=True)
    s, _ = ljlk_pose_scores(*args)
    numpy.testing.assert_allclose(
        s.cpu().numpy(), load_term_baseline("test_block_scoring"), atol=1e-5, rtol=1e-3
    )


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_neighbors_bit_exact(ubq_pdb, dtype):
    from tmol_amd.ops import padded_work_list

    poses = _stacks(ubq_pdb)
    _say("neighbors %s" % dtype)
    ps, args = _gpu_args(poses, dtype)
    items, nbr, overflow = padded_work_list(args)
    _, cargs = _cpu_args(poses, dtype)
    _, ref_nbr, _, _ = orc.oracle_pose_scores(*cargs)
    assert overflow == 0
    assert (nbr.cpu().numpy() == ref_nbr).all()
    P, B = ps.n_poses, ps.max_n_blocks
    got = set(map(tuple, items.numpy().tolist()))
    p, b1, b2 = numpy.nonzero(ref_nbr)
    want = set(zip((p * B + b1).tolist(), (p * B + b2).tolist()))
    assert got == want
    assert len(got) == items.shape[0]  # no duplicates


def test_count_pair_bit_exact(ubq_pdb):
    from tmol_amd.ops import debug_pair_separations, padded_work_list

    poses = _stacks(ubq_pdb)[:3]
    _say("count pair")
    ps, args = _gpu_args(poses)
    items, _, _ = padded_work_list(args)
    seps = debug_pair_separations(args, items).cpu().numpy()
    _, cargs = _cpu_args(poses)
    B = ps.max_n_blocks
    trip = [(int(g1) // B, int(g1) % B, int(g2) % B) for g1, g2 in items.numpy()]
    ref = orc.oracle_pair_separations(cargs, trip)
    n_checked = 0
    for k, m in enumerate(ref):
        n1, n2 = m.shape
        assert (seps[k, :n1, :n2] == m).all(), trip[k]
        n_checked += (m <= 4).sum()
    assert n_checked > 0  # the sequence-adjacent (count-pair) path was exercised


def test_energies_vs_fp64_and_oracle(ubq_pdb):
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    poses = _stacks(ubq_pdb)
    _say("energies f32/f64")
    _, a32 = _gpu_args(poses, torch.float32)
    _, a64 = _gpu_args(poses, torch.float64)
    s32, _ = ljlk_pose_scores(*a32)
    s64, _ = ljlk_pose_scores(*a64)
    s32 = s32.cpu().numpy().astype(numpy.float64)
    s64 = s64.cpu().numpy()
    rel = numpy.abs(s32 - s64) / numpy.maximum(numpy.abs(s64), 1.0)
    print("max rel err f32 vs f64 per (term, pose):", rel.max())
    assert rel.max() < 1e-4
    _, c64 = _cpu_args(poses, torch.float64)
    o64, _, _, _ = orc.oracle_pose_scores(*c64)
    numpy.testing.assert_allclose(s64, o64, rtol=1e-9, atol=1e-9)
    _, c32 = _cpu_args(poses, torch.float32)
    o32, _, _, _ = orc.oracle_pose_scores(*c32)
    numpy.testing.assert_allclose(s32, o32, rtol=1e-4, atol=1e-4)


def test_whole_pose_gradients(ubq_pdb):
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    poses = _stacks(ubq_pdb)
    _say("whole-pose gradients")
    w = torch.tensor([[1.0], [0.55], [1.0]], dtype=torch.float64)
    grads = {}
    for dt in (torch.float32, torch.float64):
        _, a = _gpu_args(poses, dt)
        c = a[0].clone().requires_grad_(True)
        a[0] = c
        s, _ = ljlk_pose_scores(*a)
        (s * w.to(DEV, dt)).sum().backward()
        grads[dt] = c.grad.double().cpu().numpy()
    g32, g64 = grads[torch.float32], grads[torch.float64]
    # per-atom error relative to the largest gradient component of the stack
    err = numpy.abs(g32 - g64).max() / numpy.abs(g64).max()
    print("max |g32 - g64| / max|g64|:", err)
    assert err < 1e-5
    _, c64 = _cpu_args(poses, torch.float64)
    _, _, dv, _ = orc.oracle_pose_scores(*c64, require_grad=True)
    og = (dv * w.numpy()[:, :, None]).sum(0)
    numpy.testing.assert_allclose(g64, og, rtol=1e-8, atol=1e-8)


def test_block_pair_backward(ubq_pdb):
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    poses = [pose_from_pdb(ubq_pdb, (0, 12)), pose_from_pdb(ubq_pdb, (40, 48))]
    _say("block pair backward gradcheck")
    _, a = _gpu_args(poses, torch.float64, block_pair=True)
    x = a[0].clone().requires_grad_(True)
    scale = None

    def f(c):
        nonlocal scale
        aa = list(a)
        aa[0] = c
        s, _ = ljlk_pose_scores(*aa)
        if scale is None:
            scale = 0.01 * torch.arange(s.numel(), device=DEV, dtype=s.dtype).reshape(s.shape)
        return (scale * s).sum()

    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-5, rtol=1e-3, nondet_tol=1e-6)


def test_whole_pose_gradcheck(ubq_pdb):
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    _say("whole pose gradcheck")
    _, a = _gpu_args([pose_from_pdb(ubq_pdb, (0, 4))], torch.float64)
    wt = torch.rand(3, 1, dtype=torch.float64, device=DEV)
    x = a[0].clone().requires_grad_(True)

    def f(c):
        aa = list(a)
        aa[0] = c
        s, _ = ljlk_pose_scores(*aa)
        return (wt * s).sum()

    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-5, rtol=1e-3, nondet_tol=1e-6)


def test_cpu_tensor_raises(ubq_pdb):
    from tmol_amd.ops import LJLKLibraryError

    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores

    _, cargs = _cpu_args([pose_from_pdb(ubq_pdb, (0, 4))])
    with pytest.raises(LJLKLibraryError):
        ljlk_pose_scores(*cargs)
