"""GPU parity of rotamer-pair scoring (``tmol_amd.ops.ljlk_rotamer_scores``,
the drop-in for tmol_ljlk::ljlk_rotamer_scores) through the C ABI:

* rotamer spheres, block spheres and the dispatch list bit-identical to the
  oracle (which matches the reference op exactly);
* energies within 1e-10 (f64) and 1e-4 of max(1, |E|) (f32) of the f64
  oracle, and against the reference's golden vectors;
* the dTdV-weighted backward against the oracle;
* one rotamer per block reproduces the block-pair op;
* edge cases: blocks without rotamers, several poses of different lengths;
* the 16 x 300-residue repacking workload: bit-identical dispatch, sampled
  energies, f32 against f64.
"""

import json
import os
import sys

import numpy
import pytest
import torch

from oracle import oracle as orc
from tmol_amd.rotamers import (current_rotamers, dispatch_checksum, repack_workload,
                               rotamer_op_args, subset_rotamers)
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _args(ps, rs, dt, device=None):
    a = rotamer_op_args(ps, rs, LJLKEnergyTerm())
    if dt == torch.float64:
        a[23], a[24] = a[23].double(), a[24].double()
    if device is not None:
        a = [x.to(device) if isinstance(x, torch.Tensor) else x for x in a]
    return a


def _small_case(n_res=24, seed=3):
    from tmol_amd.chemical import default_chemical_database
    from tmol_amd.generator import build_chains_fast, workload_specs
    from tmol_amd.pose import Pose, PoseStack
    from tmol_amd.rotamers import build_repack_rotamers

    db = default_chemical_database()
    specs = workload_specs("padded_64x150_600", 2)
    sp = []
    for k, n in enumerate((n_res, n_res - 7)):
        rts, phis, psis, omegas, chis = specs[k]
        r = list(rts[:n])
        r[-1] = db.residue_type(r[-1].base_name, cterm=True)
        sp.append((r, phis[:n], psis[:n], omegas[:n], chis[:n]))
    coords = build_chains_fast(sp)
    poses = [Pose(s[0], c) for s, c in zip(sp, coords)]
    out = {}
    for dt in (torch.float32, torch.float64):
        ps = PoseStack(poses, dtype=dt)
        out[dt] = (ps, build_repack_rotamers(sp, ps, seed=seed))
    return out


@pytest.fixture(scope="module")
def small():
    return _small_case()


def _gpu_call(args):
    from tmol_amd.ops import RotamerCall

    call = RotamerCall(args, args[0].dtype)
    idx = call.dispatch()
    return call, idx


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_dispatch_and_spheres_bit_exact(small, dt):
    ps, rs = small[dt]
    _say("rotamer dispatch %s, %d rotamers" % (dt, rs.n_rotamers_total))
    args = _args(ps, rs, dt)
    o, idx, rsph, bsph, nbr = orc.oracle_rotamer_dispatch(*args)
    call, gidx = _gpu_call(_args(ps, rs, dt, DEV))
    torch.cuda.synchronize()
    g_rs, g_bs = call.spheres()
    numpy.testing.assert_array_equal(g_rs.cpu().numpy(), rsph)
    numpy.testing.assert_array_equal(g_bs.cpu().numpy(), bsph)
    numpy.testing.assert_array_equal(gidx.cpu().numpy(), idx)
    assert idx.shape[1] > 10000


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_energies_match_oracle(small, dt):
    ps, rs = small[dt]
    _say("rotamer energies %s" % dt)
    ps64, rs64 = small[torch.float64]
    E64, idx, *_ = orc.oracle_rotamer_scores(*_args(ps64, rs64, torch.float64))
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores

    s, gidx = ljlk_rotamer_scores(*_args(ps, rs, dt, DEV))
    numpy.testing.assert_array_equal(gidx.cpu().numpy(), idx)
    E = s.double().cpu().numpy()
    err = numpy.abs(E - E64) / numpy.maximum(1.0, numpy.abs(E64))
    _say("max rel err %.3g" % err.max())
    assert err.max() < (1e-10 if dt == torch.float64 else 1e-4)
    assert numpy.count_nonzero(E64) > 1000


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_backward_matches_oracle(small, dt):
    ps, rs = small[dt]
    _say("rotamer backward %s" % dt)
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores

    a = _args(ps, rs, dt, DEV)
    c = a[0].clone().requires_grad_(True)
    a[0] = c
    s, idx = ljlk_rotamer_scores(*a)
    w = torch.rand(s.shape, generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    (s * w.to(DEV, dt)).sum().backward()
    ps64, rs64 = small[torch.float64]
    o, oidx, *_ = orc.oracle_rotamer_dispatch(*_args(ps64, rs64, torch.float64))
    g = o.score(oidx, dTdV=w.numpy())
    gg = c.grad.double().cpu().numpy()
    err = numpy.abs(gg - g).max() / max(1.0, numpy.abs(g).max())
    _say("grad max err / scale %.3g" % err)
    assert err < (1e-10 if dt == torch.float64 else 1e-4)
    assert numpy.linalg.norm(gg - g) <= (1e-10 if dt == torch.float64 else 1e-4) * numpy.linalg.norm(g)


@pytest.mark.parametrize("name,dt", [("float32", torch.float32), ("float64", torch.float64)])
def test_reference_golden(name, dt):
    g = json.load(open(os.path.join(HERE, "golden", "rotamer_ref_single100.json")))[name]
    ps, rs = repack_workload("cpu_single_100", seed=1, dtype=dt)
    _say("rotamer golden %s" % name)
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores

    a = _args(ps, rs, dt, DEV)
    c = a[0].clone().requires_grad_(True)
    a[0] = c
    s, idx = ljlk_rotamer_scores(*a)
    idx_h = idx.cpu().numpy()
    assert idx_h.shape[1] == g["n_pairs"]
    assert dispatch_checksum(idx_h) == int(g["dispatch_checksum"])
    pick = numpy.array(g["sample_pairs"])
    E = s.detach().double().cpu().numpy()
    ref = numpy.array(g["sample_energies"])
    err = numpy.abs(E[:, pick] - ref) / numpy.maximum(1.0, numpy.abs(ref))
    assert err.max() < (1e-10 if dt == torch.float64 else 1e-4)
    numpy.testing.assert_allclose(E.sum(1), g["totals"], rtol=1e-10 if dt == torch.float64 else 1e-5)
    gen = torch.Generator().manual_seed(20261203)
    w = torch.rand(s.shape, generator=gen, dtype=torch.float64).to(dt)
    (s * w.to(DEV)).sum().backward()
    atoms = numpy.array(g["sample_atoms"])
    gr = numpy.array(g["sample_grad"])
    gg = c.grad.double().cpu().numpy()[atoms]
    assert numpy.abs(gg - gr).max() / max(1.0, numpy.abs(gr).max()) < (1e-10 if dt == torch.float64 else 1e-4)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_single_rotamer_set_equals_block_pair_op(dt):
    """One rotamer per block: the rotamer op reproduces the block-pair op."""
    from tmol_amd.generator import workload_poses
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores
    ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores
    from tmol_amd.pose import PoseStack

    _say("rotamer vs block pair %s" % dt)
    poses = workload_poses("cpu_single_100") + workload_poses("padded_64x150_600", 1)
    ps = PoseStack(poses, dtype=dt)
    rs = current_rotamers(ps)
    s, idx = ljlk_rotamer_scores(*_args(ps, rs, dt, DEV))
    pa = pose_op_args(ps, LJLKEnergyTerm(), block_pair=True)
    if dt == torch.float64:
        pa[23], pa[24] = pa[23].double(), pa[24].double()
    pa = [x.to(DEV) if isinstance(x, torch.Tensor) else x for x in pa]
    bp, nbr = ljlk_pose_scores(*pa)
    idx = idx.long()
    blk = rs.block_ind_for_rot.to(DEV).long()
    want = bp[:, idx[0], blk[idx[1]], blk[idx[2]]]
    assert idx.shape[1] == int(nbr.sum())
    err = ((s - want).abs() / want.abs().clamp(min=1.0)).max().item()
    assert err < (1e-12 if dt == torch.float64 else 1e-5)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_blocks_without_rotamers(small, dt):
    """Some blocks lose all their rotamers, others keep one: dispatch and
    energies still match the oracle."""
    ps, rs = small[dt]
    _say("rotamer subset %s" % dt)
    blk = rs.block_ind_for_rot.numpy()
    keep = numpy.ones(rs.n_rotamers_total, dtype=bool)
    keep[blk % 5 == 2] = False
    keep[(blk % 5 == 4) & (numpy.arange(keep.size) % 2 == 1)] = False
    sub = subset_rotamers(rs, keep, ps.packed_block_types.n_atoms.numpy())
    assert int((sub.n_rots_for_block == 0).sum()) > 2 * 3
    args = _args(ps, sub, dt)
    o, idx, *_ = orc.oracle_rotamer_dispatch(*args)
    E = o.score(idx)
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores

    s, gidx = ljlk_rotamer_scores(*_args(ps, sub, dt, DEV))
    numpy.testing.assert_array_equal(gidx.cpu().numpy(), idx)
    err = numpy.abs(s.double().cpu().numpy() - E) / numpy.maximum(1.0, numpy.abs(E))
    assert err.max() < (1e-10 if dt == torch.float64 else 1e-4)


def test_repack_workload_16x300():
    """The benchmark workload: dispatch bit-identical to the oracle, 4000
    sampled pairs against the oracle (f64 and f32), and the f32 run's
    per-term totals against the f64 run's."""
    _say("repack 16x300 dispatch")
    out = {}
    for dt in (torch.float64, torch.float32):
        ps, rs = repack_workload("repack_16x300", seed=1, dtype=dt)
        import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
        ljlk_rotamer_scores = torch.ops.tmol_ljlk.ljlk_rotamer_scores

        s, idx = ljlk_rotamer_scores(*_args(ps, rs, dt, DEV))
        out[dt] = (ps, rs, s.double().cpu().numpy(), idx.cpu().numpy())
    ps, rs, E64, idx64 = out[torch.float64]
    _, _, E32, idx32 = out[torch.float32]
    o, idx, *_ = orc.oracle_rotamer_dispatch(*_args(ps, rs, torch.float64))
    numpy.testing.assert_array_equal(idx64, idx)
    numpy.testing.assert_array_equal(idx32, idx)
    _say("repack 16x300: %d rotamers, %d pairs" % (rs.n_rotamers_total, idx.shape[1]))
    assert idx.shape[1] > 5_000_000
    pick = numpy.sort(numpy.random.default_rng(1).choice(idx.shape[1], 4000, replace=False))
    Eo = o.score(idx[:, pick])
    assert (numpy.abs(E64[:, pick] - Eo) / numpy.maximum(1.0, numpy.abs(Eo))).max() < 1e-10
    assert (numpy.abs(E32[:, pick] - Eo) / numpy.maximum(1.0, numpy.abs(Eo))).max() < 1e-4
    err = numpy.abs(E32 - E64) / numpy.maximum(1.0, numpy.abs(E64))
    assert err.max() < 1e-4
