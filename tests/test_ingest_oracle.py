"""Canonical-form ingest on the CPU: the leaf-atom tables and the oracle of
``gen_pose_leaf_atoms`` against the reference's own test (every leaf atom
of 1ubq rebuilt within 0.1 A of the deposited one,
tmol/tests/io/details/test_build_missing_leaf_atoms.py:52-173), the
reference C++ op built into oracle/_ref, gradcheck, disulfide detection on
6DMZ (tmol/tests/data/pdb/6DMZ_A.pdb, four disulfides), and the
error paths of the host side."""

import os

import numpy
import pytest
import torch

from oracle import leaf_oracle as lo
from tmol_amd import ingest
from tmol_amd.ops import LJLKLibraryError

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# tmol/tests/io/details/test_build_missing_leaf_atoms.py:29-49: leaves the
# reference test leaves in place (O / OXT are rebuilt from each other only)
NOT_REBUILT = {"O", "OXT"}
NOT_REBUILT_BY_RES = {"GLU": "OE2", "ASP": "OD2", "TRP": "CD2"}


@pytest.fixture(scope="module")
def co():
    return ingest.CanonicalOrdering()


def _pdb(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def _rebuild_all_leaves(co, pdb_text, dtype=torch.float32):
    """(ps, args, layout): every leaf atom but the reference test's
    exceptions flagged missing."""
    ch, rt, xyz = ingest.canonical_form_from_pdb(pdb_text, co)
    lay = ingest.block_layout_from_canonical_form(co, ch, rt, xyz)
    ps = ingest.pose_stack_from_layouts(lay, dtype)
    _, tab = ingest.leaf_op_args(ps, co.db, lay)
    L = lay[0]
    for b, r in enumerate(L["rts"]):
        lf = tab.is_leaf[ps.packed_block_types.index[r.name]][: r.n_atoms]
        for a, an in enumerate(r.atom_names):
            if lf[a] and an not in NOT_REBUILT and NOT_REBUILT_BY_RES.get(r.base_name) != an:
                L["missing"][b][a] = True
    args, _ = ingest.leaf_op_args(ps, co.db, lay)
    return ps, args, L


@pytest.mark.parametrize("pdb", ["1ubq.pdb", "6DMZ_A.pdb"])
def test_rebuilt_leaves_match_deposited_atoms(co, pdb):
    ps, args, L = _rebuild_all_leaves(co, _pdb(pdb))
    m = args[2][0]
    assert int(m.sum()) > 300  # hydrogens of every residue, OXT from O
    out = lo.oracle_leaf_atoms(*args)
    dev = (out[0] - ps.coords[0]).norm(dim=-1)
    assert not torch.isnan(out).any()
    assert float(dev[m].max()) < 1e-1  # the reference test's tolerance
    assert float(dev[m].mean()) < 2e-3  # the deposited hydrogens sit at ideal geometry
    assert float(dev[~m].max()) == 0.0  # given atoms pass through


def test_leaf_annotation(co):
    """Leaves are the atoms nothing is built from: hydrogens, O / OXT, the
    last side-chain heavy atoms that end no named torsion."""
    db = co.db
    ats = db.atom_type_set
    ser = db.residue_type("SER", cterm=True)
    lf = ingest.leaf_atoms_of(ser, ats.is_hydrogen[ser.atom_types].astype(bool))
    leaves = {n for n, x in zip(ser.atom_names, lf) if x}
    assert {"HA", "HB2", "HB3", "HG", "O", "OXT"} <= leaves
    assert not leaves & {"N", "CA", "C", "CB", "OG"}  # OG ends chi1: not a leaf
    glu = db.residue_type("GLU")
    lf = ingest.leaf_atoms_of(glu, ats.is_hydrogen[glu.atom_types].astype(bool))
    leaves = {n for n, x in zip(glu.atom_names, lf) if x}
    assert "OE2" in leaves and "OE1" not in leaves  # OE1 ends chi3


def test_downstream_of_connections(co):
    cyd = co.db.residue_type("CYD")
    d = ingest.downstream_of_conn(cyd)
    names = [[cyd.atom_names[i] for i in row[:4]] for row in d]
    assert names[0] == ["N", "CA", "C", "C"]  # down
    assert names[1] == ["C", "CA", "N", "N"]  # up
    assert names[2] == ["SG", "CB", "CA", "N"]  # dslf: up the icoor parents


def test_oracle_matches_reference_op_on_random_masks(co):
    """Two poses (1ubq, 6DMZ) padded together, random leaves missing
    (termini, OXT with and without O): forward and backward of the oracle
    against the reference op."""
    try:
        ref = lo.load_reference_io_ops()
    except FileNotFoundError:
        pytest.skip("reference ingest op not built")
    lays = []
    for name in ("1ubq.pdb", "6DMZ_A.pdb"):
        ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb(name), co)
        lays += ingest.block_layout_from_canonical_form(co, ch, rt, xyz)
    ps = ingest.pose_stack_from_layouts(lays)
    _, tab = ingest.leaf_op_args(ps, co.db, lays)
    rng = numpy.random.default_rng(11)
    for L in lays:
        for b, r in enumerate(L["rts"]):
            lf = tab.is_leaf[ps.packed_block_types.index[r.name]][: r.n_atoms]
            L["missing"][b][:] = lf & (rng.random(r.n_atoms) < 0.6)
    args, _ = ingest.leaf_op_args(ps, co.db, lays)
    assert int(args[1].sum()) > 500
    out = lo.oracle_leaf_atoms(*args)
    r = ref.gen_pose_leaf_atoms(*args)
    assert torch.equal(torch.isnan(out), torch.isnan(r))
    assert float((out - r).abs().nan_to_num(0).max()) < 2e-5
    X = args[0].clone().requires_grad_(True)
    a2 = list(args)
    a2[0] = X
    w = torch.randn(out.shape, generator=torch.Generator().manual_seed(3))
    (ref.gen_pose_leaf_atoms(*a2).nan_to_num(0) * w).sum().backward()
    g_ref = X.grad.clone()
    X.grad = None
    (lo.oracle_leaf_atoms(*a2).nan_to_num(0) * w).sum().backward()
    assert float((X.grad - g_ref).abs().max()) < 1e-4 * max(1.0, float(g_ref.abs().max()))


def test_oracle_gradcheck(co):
    ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb("6DMZ_A.pdb"), co)
    keep = 6
    lay = ingest.block_layout_from_canonical_form(co, ch[:, :keep], rt[:, :keep],
                                                  xyz[:, :keep])
    ps = ingest.pose_stack_from_layouts(lay, torch.float64)
    _, tab = ingest.leaf_op_args(ps, co.db, lay)
    for b, r in enumerate(lay[0]["rts"]):
        lay[0]["missing"][b][:] = tab.is_leaf[ps.packed_block_types.index[r.name]][: r.n_atoms]
    args, _ = ingest.leaf_op_args(ps, co.db, lay)
    args = [a.double() if a.is_floating_point() else a for a in args]
    args[0] = args[0].nan_to_num(0)  # built atoms read only present ancestors
    plan = lo.leaf_plan(args)
    assert len(plan[0]) > 40 and len(plan[4]) == 0

    def f(x):
        a = list(args)
        a[0] = x
        return lo.oracle_leaf_atoms(*a, plan=plan)

    X = args[0].clone().requires_grad_(True)
    assert torch.autograd.gradcheck(f, (X,), eps=1e-6, atol=1e-6)


def test_canonical_form_and_disulfides(co):
    ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb("6DMZ_A.pdb"), co)
    assert ch.shape == (1, 47) and xyz.shape == (1, 47, co.max_n_canonical_atoms, 3)
    cys = co.restype_index["CYS"]
    assert int((rt == cys).sum()) == 8
    lay = ingest.block_layout_from_canonical_form(co, ch, rt, xyz)[0]
    assert sorted((a, c) for a, _, c, _ in lay["extra_bonds"]) == [
        (2, 46), (13, 33), (19, 40), (23, 42)]
    assert sum(r.base_name == "CYD" for r in lay["rts"]) == 8
    assert lay["rts"][0].is_nterm and lay["rts"][-1].is_cterm
    ps = ingest.pose_stack_from_layouts([lay])
    irc = ps.inter_residue_connections[0]
    dslf = lay["rts"][2].connection_index("dslf")
    assert irc[2, dslf].tolist() == [46, lay["rts"][46].connection_index("dslf")]


def test_chain_breaks_and_termini(co):
    """Two chains: each gets its own termini; res_not_connected splits a
    chain without termini."""
    ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb("1ubq.pdb"), co)
    ch2 = ch.clone()
    ch2[0, 40:] = 1
    lay = ingest.block_layout_from_canonical_form(co, ch2, rt, xyz)[0]
    assert lay["rts"][39].is_cterm and lay["rts"][40].is_nterm
    assert not lay["connected"][40] and lay["connected"][41]
    rnc = torch.zeros((1, 76, 2), dtype=torch.bool)
    rnc[0, 20, 1] = True  # residue 20 not connected to 21
    lay = ingest.block_layout_from_canonical_form(co, ch, rt, xyz, rnc)[0]
    assert not lay["connected"][21] and not lay["rts"][20].is_cterm


def test_ingest_errors(co):
    pdb = _pdb("1ubq.pdb")
    bad = pdb.replace(" CB  MET A   1", " XX  MET A   1")
    with pytest.raises(ValueError, match="atom XX"):
        ingest.canonical_form_from_pdb(bad, co)
    ch, rt, xyz = ingest.canonical_form_from_pdb(pdb, co)
    xyz = xyz.clone()
    lys = co.restype_index[co.restypes[int(rt[0, 5])]]
    xyz[0, 5, co.atom_index[lys]["CA"]] = float("nan")  # a non-leaf atom
    with pytest.raises(ValueError, match="missing non-leaf"):
        ingest.pose_stack_from_canonical_form(co, ch, rt, xyz, device="cpu")


def test_table_validation_rejects_bad_indices(co):
    ch, rt, xyz = ingest.canonical_form_from_pdb(_pdb("6DMZ_A.pdb"), co)
    lay = ingest.block_layout_from_canonical_form(co, ch, rt, xyz)
    ps = ingest.pose_stack_from_layouts(lay)
    args, _ = ingest.leaf_op_args(ps, co.db, lay)
    keys = ("orig", "bmiss", "pmiss", "boff", "btype", "irc", "bt_n", "down", "anc",
            "geom", "anc_b", "geom_b")
    k = dict(zip(keys, args))
    P, NA = args[0].shape[:2]
    B, Amax = args[1].shape[1:]
    C = args[5].shape[2]
    ingest._validate_leaf_tables(k, P, NA, B, Amax, C)  # the real tables pass
    for key, edit in (("btype", lambda t: t.fill_(99)), ("boff", lambda t: t.add_(NA)),
                      ("irc", lambda t: t.fill_(B)), ("down", lambda t: t.fill_(Amax)),
                      ("anc", lambda t: t[..., 1].fill_(C))):
        k2 = dict(k)
        k2[key] = edit(k[key].clone())
        with pytest.raises(LJLKLibraryError):
            ingest._validate_leaf_tables(k2, P, NA, B, Amax, C)


def test_canonical_form_residue_range(co):
    """residue_start / residue_end keep residues start .. end-1 in file
    order; the range's first and last residues become the termini."""
    full = ingest.canonical_form_from_pdb(_pdb("1ubq.pdb"), co)
    sub = ingest.canonical_form_from_pdb(_pdb("1ubq.pdb"), co, residue_start=10, residue_end=40)
    assert sub[1].shape == (1, 30)
    assert torch.equal(sub[1][0], full[1][0, 10:40])
    assert torch.equal(sub[2][0].nan_to_num(0), full[2][0, 10:40].nan_to_num(0))
    lay = ingest.block_layout_from_canonical_form(co, *sub)[0]
    assert lay["rts"][0].is_nterm and lay["rts"][-1].is_cterm
    with pytest.raises(ValueError, match="no residues"):
        ingest.canonical_form_from_pdb(_pdb("1ubq.pdb"), co, residue_start=80)
