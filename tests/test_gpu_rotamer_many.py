"""GPU parity of the rotamer-pair kernels on blocks with 65 to 200 rotamers
(``rotamer_cases.dense_case``): the LJ+LK (``ljlk_rotamer_scores``) and elec
(``elec_rotamer_scores``) ops, forward and dTdV-weighted backward, float32
and float64, against the f64 oracles (themselves equal to the reference's
CPU ops on this set, tests/test_rotamer_many_oracle.py).  Blocks this large
run the kernels' later windows of 64 b2 rotamers, the slot carried across
windows, and output / dTdV indices past the first window."""

import sys

import numpy
import pytest
import torch

from oracle import oracle as orc
from rotamer_cases import dense_case
from tmol_amd.rotamers import rotamer_op_args
from tmol_amd.score.elec_energy_term import ElecEnergyTerm
from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
TOL = {torch.float64: 1e-10, torch.float32: 1e-4}


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dense():
    return dense_case()


def _args(term, ps, rs, dt, device=None):
    a = rotamer_op_args(ps, rs, term)
    if isinstance(term, ElecEnergyTerm):
        a[16], a[22] = a[16].to(dt), a[22].to(dt)
    else:
        a[23], a[24] = a[23].to(dt), a[24].to(dt)
    if device is not None:
        a = [x.to(device) if isinstance(x, torch.Tensor) else x for x in a]
    return a


def _ops(name):
    import tmol_amd.ops as ops  # registers torch.ops.tmol_ljlk

    if name == "ljlk":
        return LJLKEnergyTerm(), torch.ops.tmol_ljlk.ljlk_rotamer_scores, orc.oracle_rotamer_scores
    return ElecEnergyTerm(), ops.elec_rotamer_scores, orc.oracle_elec_rotamer_scores


@pytest.mark.parametrize("name", ["ljlk", "elec"])
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_many_rotamers_forward(dense, name, dt):
    term, op, oracle = _ops(name)
    ps, rs = dense[dt]
    _say("%s rotamer forward, %d rotamers (max %d per block), %s"
         % (name, rs.n_rotamers_total, int(rs.n_rots_for_block.max()), dt))
    ps64, rs64 = dense[torch.float64]
    out = oracle(*_args(term, ps64, rs64, torch.float64))
    E64, idx = out[0], out[1]
    s, gidx = op(*_args(term, ps, rs, dt, DEV))
    numpy.testing.assert_array_equal(gidx.cpu().numpy(), idx)
    E = s.double().cpu().numpy()
    err = numpy.abs(E - E64) / numpy.maximum(1.0, numpy.abs(E64))
    _say("pairs %d, max rel err %.3g" % (idx.shape[1], err.max()))
    assert err.max() < TOL[dt]
    assert numpy.count_nonzero(E64) > 1000
    # pairs whose second rotamer sits past the first 64 of its block
    rob = rs64.rot_offset_for_block.numpy()
    blk = rs64.block_ind_for_rot.numpy()
    pose = rs64.pose_for_rot.numpy()
    late = (idx[2] - rob[pose[idx[2]], blk[idx[2]]]) >= 64
    assert late.sum() > 1000 and numpy.count_nonzero(E64[..., late]) > 100


@pytest.mark.parametrize("name", ["ljlk", "elec"])
@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_many_rotamers_backward(dense, name, dt):
    term, op, oracle = _ops(name)
    ps, rs = dense[dt]
    _say("%s rotamer backward, %s" % (name, dt))
    a = _args(term, ps, rs, dt, DEV)
    c = a[0].clone().requires_grad_(True)
    a[0] = c
    s, idx = op(*a)
    w = torch.rand(s.shape, generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    (s * w.to(DEV, dt)).sum().backward()
    ps64, rs64 = dense[torch.float64]
    g = oracle(*_args(term, ps64, rs64, torch.float64), dTdV=w.numpy())[-1]
    gg = c.grad.double().cpu().numpy()
    err = numpy.abs(gg - g).max() / max(1.0, numpy.abs(g).max())
    _say("grad max err / scale %.3g" % err)
    assert err < TOL[dt]
    assert numpy.linalg.norm(gg - g) <= TOL[dt] * numpy.linalg.norm(g)
