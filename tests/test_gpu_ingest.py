"""Missing-leaf-atom building on the GPU (``gen_pose_leaf_atoms`` drop-in,
tmol_leaf_atoms_build / _backward) against the oracle, float32 and float64;
the canonical-form ingest end to end (heavy atoms of 1ubq and 6DMZ in,
hydrogens built, LJ+LK scored); and size-independent properties at a large
batch (idempotence, given atoms passing through)."""

import os

import numpy
import pytest
import torch

from oracle import leaf_oracle as lo
from tmol_amd import ingest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
# float64: the same formulas in another order; float32: hardware sin / cos
# and sqrt, float atomics in the backward sums
TOL = {torch.float64: 1e-11, torch.float32: 2e-5}


@pytest.fixture(scope="module")
def co():
    return ingest.CanonicalOrdering()


def _pdb(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def _random_missing(co, dtype, seed=11, frac=0.6, names=("1ubq.pdb", "6DMZ_A.pdb")):
    lays = []
    for name in names:
        ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb(name), co)
        lays += ingest.block_layout_from_canonical_form(co, ch, rt, xyz)
    ps = ingest.pose_stack_from_layouts(lays, dtype)
    _, tab = ingest.leaf_op_args(ps, co.db, lays)
    rng = numpy.random.default_rng(seed)
    for L in lays:
        for b, r in enumerate(L["rts"]):
            lf = tab.is_leaf[ps.packed_block_types.index[r.name]][: r.n_atoms]
            L["missing"][b][:] = lf & (rng.random(r.n_atoms) < frac)
    args, _ = ingest.leaf_op_args(ps, co.db, lays)
    return ps, args


def _dev(args):
    return [a.cuda() if torch.is_tensor(a) else a for a in args]


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("frac", [0.6, 1.0])
def test_leaf_atoms_match_oracle(co, dt, frac):
    """Forward (NaN placement exact) and backward of the HIP op against the
    oracle's autograd; frac=1 drops every leaf, O and OXT together (both
    unbuildable: NaN), every terminus on its backup ancestors."""
    print("leaf atoms", dt, frac, flush=True)
    ps, args = _random_missing(co, dt, frac=frac)
    args = [a.to(dt) if a.is_floating_point() else a for a in args]
    X = args[0].cuda().requires_grad_(True)
    da = _dev(args)
    da[0] = X
    out = ingest.gen_pose_leaf_atoms(*da)
    Xo = args[0].clone().requires_grad_(True)
    ao = list(args)
    ao[0] = Xo
    ref = lo.oracle_leaf_atoms(*ao)
    o = out.detach().cpu()
    assert torch.equal(torch.isnan(o), torch.isnan(ref))
    assert int(args[1].sum()) > 500
    scale = float(ref.detach().abs().nan_to_num(0).max())
    assert float((o - ref.detach()).abs().nan_to_num(0).max()) <= TOL[dt] * scale
    w = torch.randn(ref.shape, dtype=dt, generator=torch.Generator().manual_seed(4))
    (out.nan_to_num(0) * w.cuda()).sum().backward()
    (ref.nan_to_num(0) * w).sum().backward()
    g = X.grad.cpu()
    gs = max(1.0, float(Xo.grad.abs().max()))
    assert float((g - Xo.grad).abs().max()) <= 10 * TOL[dt] * gs


@pytest.mark.parametrize("pdb", ["1ubq.pdb", "6DMZ_A.pdb"])
def test_ingest_heavy_atoms_then_score(co, pdb):
    """Hydrogens stripped from the canonical form, rebuilt on the GPU by
    pose_stack_from_canonical_form: every rebuilt atom within 0.1 A of the
    deposited one (the reference test's bound) and the LJ+LK energies of
    the rebuilt pose close to those of the deposited one."""
    import tmol_amd.ops  # noqa: F401  (registers torch.ops.tmol_ljlk)
    from tmol_amd.score.ljlk_energy_term import LJLKEnergyTerm, pose_op_args

    ljlk_pose_scores = torch.ops.tmol_ljlk.ljlk_pose_scores
    print("ingest", pdb, flush=True)
    ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb(pdb), co)
    full = ingest.pose_stack_from_layouts(
        ingest.block_layout_from_canonical_form(co, ch, rt, xyz))
    heavy = xyz.clone()
    for r in range(rt.shape[1]):
        t = int(rt[0, r])
        for name, k in co.atom_index[t].items():
            if name.startswith("H") or (len(name) > 1 and name[0].isdigit() and name[1] == "H"):
                heavy[0, r, k] = float("nan")
    ps = ingest.pose_stack_from_canonical_form(co, ch, rt, heavy, device="cuda")
    assert ps.coords.is_cuda and ps.block_type_ind.is_cuda
    assert ps.coords.shape == full.coords.shape
    dev = (ps.coords[0].cpu() - full.coords[0]).norm(dim=-1)
    assert not torch.isnan(ps.coords).any()
    assert float(dev.max()) < 1e-1 and float(dev.mean()) < 1e-3
    from tmol_amd.pose import PoseStack

    term = LJLKEnergyTerm(device="cuda")
    e_full, _ = ljlk_pose_scores(*pose_op_args(PoseStack(full.poses, device="cuda"), term))
    e_built, _ = ljlk_pose_scores(*pose_op_args(ps, term))
    assert torch.allclose(e_built.cpu(), e_full.cpu(), rtol=2e-3, atol=2e-2)


def test_large_batch_properties(co):
    """512 poses (256 copies each of 1ubq and 6DMZ with different random
    masks): building twice is idempotent, given atoms pass through
    bit-exactly, and every pose's output matches that of the same pose
    built alone."""
    print("leaf atoms batch 512", flush=True)
    lays = []
    for name in ("1ubq.pdb", "6DMZ_A.pdb"):
        ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb(name), co)
        lays += ingest.block_layout_from_canonical_form(co, ch, rt, xyz)
    base = ingest.pose_stack_from_layouts(lays)
    _, tab = ingest.leaf_op_args(base, co.db, lays)
    rng = numpy.random.default_rng(5)
    big = []
    for k in range(512):
        L = lays[k % 2]
        m = []
        for b, r in enumerate(L["rts"]):
            lf = tab.is_leaf[base.packed_block_types.index[r.name]][: r.n_atoms]
            m.append(lf & (rng.random(r.n_atoms) < 0.8))
        big.append(dict(L, missing=m))
    ps = ingest.pose_stack_from_layouts(big)
    args, _ = ingest.leaf_op_args(ps, co.db, big)
    da = _dev(args)
    out = ingest.gen_pose_leaf_atoms(*da)
    given = ~args[2].cuda()
    real = torch.zeros_like(given)
    for p, L in enumerate(big):
        real[p, : sum(r.n_atoms for r in L["rts"])] = True
    keep = given & real
    assert torch.equal(out[keep], da[0][keep])
    da2 = list(da)
    da2[0] = out.clone()
    out2 = ingest.gen_pose_leaf_atoms(*da2)
    assert torch.equal(torch.isnan(out2), torch.isnan(out))
    assert torch.equal(out2.nan_to_num(0), out.nan_to_num(0))
    for p in (0, 1, 257, 510):
        one = [big[p]]
        ps1 = ingest.pose_stack_from_layouts(one)
        a1, _ = ingest.leaf_op_args(ps1, co.db, one)
        o1 = ingest.gen_pose_leaf_atoms(*_dev(a1))
        n = o1.shape[1]
        assert torch.equal(o1[0].nan_to_num(0), out[p, :n].nan_to_num(0))


@pytest.mark.parametrize("amax", [64, 100])
def test_wide_block_type_tables(co, amax):
    """The same poses with the block-type tables padded to ``amax`` atom
    slots (one block per wave pass, lanes striding over the slots when
    amax > 64): forward and backward identical to the unpadded call."""
    print("leaf atoms amax", amax, flush=True)
    ps, args = _random_missing(co, torch.float32, seed=7)
    A0 = args[1].shape[2]

    def pad(t, dim, value):
        shape = list(t.shape)
        shape[dim] = amax - A0
        return torch.cat([t, torch.full(shape, value, dtype=t.dtype)], dim=dim)

    wide = list(args)
    wide[1] = pad(args[1], 2, False)
    wide[7] = pad(args[7], 2, -1)
    wide[8] = pad(args[8], 1, -1)
    wide[9] = pad(args[9], 1, 0.0)
    wide[10] = pad(args[10], 1, -1)
    wide[11] = pad(args[11], 1, 0.0)
    outs, grads = [], []
    for a in (args, wide):
        X = a[0].cuda().requires_grad_(True)
        da = _dev(a)
        da[0] = X
        out = ingest.gen_pose_leaf_atoms(*da)
        w = torch.randn(out.shape, generator=torch.Generator().manual_seed(2)).cuda()
        (out.nan_to_num(0) * w).sum().backward()
        outs.append(out.detach().cpu())
        grads.append(X.grad.cpu())
    assert torch.equal(outs[0].nan_to_num(0), outs[1].nan_to_num(0))
    assert torch.equal(torch.isnan(outs[0]), torch.isnan(outs[1]))
    assert torch.allclose(grads[0], grads[1], rtol=1e-6, atol=1e-6)


def test_pose_stack_from_pdb(co):
    """The top-level entry from a file name with a residue range: the
    interior residues' atoms (hydrogens included, as given) equal the
    deposited ones and the new termini are built and finite."""
    print("pose_stack_from_pdb 1ubq 10..40", flush=True)
    ps = ingest.pose_stack_from_pdb(os.path.join(GOLDEN, "1ubq.pdb"), "cuda",
                                    residue_start=10, residue_end=40, co=co)
    assert ps.n_poses == 1 and ps.max_n_blocks == 30
    assert not torch.isnan(ps.coords).any()
    rts = ps.poses[0].residue_types
    assert rts[0].is_nterm and rts[-1].is_cterm
    full = ingest.pose_stack_from_layouts(ingest.block_layout_from_canonical_form(
        co, *ingest.canonical_form_from_pdb(_pdb("1ubq.pdb"), co)))
    for b in range(1, 29):
        assert numpy.allclose(ps.poses[0].coords[b], full.poses[0].coords[b + 10], atol=1e-4)
