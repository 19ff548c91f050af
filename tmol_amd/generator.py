"""Seeded synthetic poses with ideal geometry.

Each residue is built from its internal coordinates (the ``icoors`` of the
residue-type database: bond length ``d``, ``theta`` = 180 deg - bond angle,
torsion ``phi`` relative to parent / grand parent / great grand parent),
using the natural-extension reference frame placement (NeRF).  Backbone
torsions are imposed through the connection pseudo-atoms (``down`` sets phi,
``up`` sets psi), side-chain chis through the icoor torsion of each chi's
fourth atom.  The N-CA-C backbone of the chain is placed the same way and each
residue's ideal template is superimposed on it.

Sampling per residue (the workload definition this package benchmarks):
70% helix phi,psi ~ N(-63, 10), N(-41, 10); 30% strand N(-120, 15),
N(130, 15); omega = 180; each chi ~ N(c, 10) with c uniform in
{60, 180, -60}; proline ring chis keep their ideal values.  Sequences are
i.i.d. uniform over the 20 canonical amino acids (HIS = HE2 tautomer).

Same role as tmol's extended_pose_stack_from_sequences + set_named_torsions
(tmol/pose/util.py:38-140); written independently and checked by geometry
invariants (tests/test_generator.py).
"""

from typing import List, Optional, Sequence

import numpy

from .chemical import ResidueType, default_chemical_database
from .pose import Pose

DEG = numpy.pi / 180.0
CANONICAL_ORDER = [
    "ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
    "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
]
MASTER_SEED = 20261202


def nerf(a, b, c, bond, angle, torsion):
    """Place d with |cd| = bond, angle(b, c, d) = angle,
    dihedral(a, b, c, d) = torsion (radians)."""
    bc = c - b
    bc = bc / numpy.linalg.norm(bc)
    n = numpy.cross(b - a, bc)
    n = n / numpy.linalg.norm(n)
    m = numpy.cross(n, bc)
    return c + bond * (
        -numpy.cos(angle) * bc
        + numpy.sin(angle) * numpy.cos(torsion) * m
        + numpy.sin(angle) * numpy.sin(torsion) * n
    )


def dihedral(a, b, c, d):
    """IUPAC dihedral angle (radians) of four points."""
    b1, b2, b3 = b - a, c - b, d - c
    n1 = numpy.cross(b1, b2)
    n2 = numpy.cross(b2, b3)
    return numpy.arctan2(numpy.linalg.norm(b2) * numpy.dot(b1, n2), numpy.dot(n1, n2))


def _angle(a, b, c):
    u, v = a - b, c - b
    return numpy.arccos(numpy.clip(numpy.dot(u, v) / numpy.linalg.norm(u) / numpy.linalg.norm(v), -1, 1))


def chi_definitions(rt: ResidueType):
    out = []
    for t in rt.torsions:
        if t["name"].startswith("chi"):
            out.append((t["name"], tuple(t[k]["atom"] for k in "abcd")))
    return out


def build_residue_template(rt: ResidueType, phi=None, psi=None, chis=None):
    """Coordinates of every atom (and the up/down pseudo-atoms) in a local
    frame; torsions in degrees (None keeps the ideal icoor value)."""
    overrides = {}
    if phi is not None:
        overrides["down"] = phi
    if psi is not None:
        overrides["up"] = psi
    if chis:
        defs = dict(chi_definitions(rt))
        for name, val in chis.items():
            a, b, c, d = defs[name]
            overrides[d] = val
    ics = rt.icoors
    pos = {}
    # root frame: first atom at the origin, second on +x, third in the xy plane
    n0, n1, n2 = ics[0]["name"], ics[1]["name"], ics[2]["name"]
    pos[n0] = numpy.zeros(3)
    pos[n1] = numpy.array([ics[1]["d"], 0.0, 0.0])
    A = numpy.pi - ics[2]["theta"] * DEG  # bond angle at the parent
    if ics[2]["parent"] == n1:
        # angle(n0, n1, n2) = A, with n1 -> n0 along -x
        pos[n2] = pos[n1] + ics[2]["d"] * numpy.array([-numpy.cos(A), numpy.sin(A), 0.0])
    else:
        # angle(n1, n0, n2) = A, with n0 -> n1 along +x
        pos[n2] = pos[n0] + ics[2]["d"] * numpy.array([numpy.cos(A), numpy.sin(A), 0.0])
    for ic in ics[3:]:
        name = ic["name"]
        tors = overrides.get(name, ic["phi"])
        pos[name] = nerf(
            pos[ic["great_grand_parent"]],
            pos[ic["grand_parent"]],
            pos[ic["parent"]],
            ic["d"],
            numpy.pi - ic["theta"] * DEG,
            tors * DEG,
        )
    return pos


def _frame(n, ca, c):
    x = ca - n
    x = x / numpy.linalg.norm(x)
    y = (c - ca) - numpy.dot(c - ca, x) * x
    y = y / numpy.linalg.norm(y)
    z = numpy.cross(x, y)
    return numpy.stack([x, y, z], axis=1), ca


def build_chain(rts: Sequence[ResidueType], phis, psis, omegas, chis_list) -> List[numpy.ndarray]:
    """Global coordinates for a single chain; torsions in degrees."""
    n = len(rts)
    temps = [
        build_residue_template(
            rts[i],
            phi=phis[i] if rts[i].down_connection_ind >= 0 else None,
            psi=psis[i] if rts[i].up_connection_ind >= 0 else None,
            chis=chis_list[i],
        )
        for i in range(n)
    ]
    N = numpy.zeros((n, 3))
    CA = numpy.zeros((n, 3))
    C = numpy.zeros((n, 3))
    t0 = temps[0]
    d_nca = numpy.linalg.norm(t0["CA"] - t0["N"])
    d_cac = numpy.linalg.norm(t0["C"] - t0["CA"])
    a_ncac = _angle(t0["N"], t0["CA"], t0["C"])
    N[0] = 0.0
    CA[0] = [d_nca, 0.0, 0.0]
    C[0] = CA[0] + d_cac * numpy.array([-numpy.cos(a_ncac), numpy.sin(a_ncac), 0.0])
    for i in range(1, n):
        tp, tc = temps[i - 1], temps[i]
        N[i] = nerf(N[i - 1], CA[i - 1], C[i - 1],
                    numpy.linalg.norm(tp["up"] - tp["C"]), _angle(tp["CA"], tp["C"], tp["up"]),
                    psis[i - 1] * DEG)
        CA[i] = nerf(CA[i - 1], C[i - 1], N[i],
                     numpy.linalg.norm(tc["CA"] - tc["N"]), _angle(tc["down"], tc["N"], tc["CA"]),
                     omegas[i - 1] * DEG)
        C[i] = nerf(C[i - 1], N[i], CA[i],
                    numpy.linalg.norm(tc["C"] - tc["CA"]), _angle(tc["N"], tc["CA"], tc["C"]),
                    phis[i] * DEG)
    out = []
    for i, rt in enumerate(rts):
        t = temps[i]
        Rt, ot = _frame(t["N"], t["CA"], t["C"])
        Rg, og = _frame(N[i], CA[i], C[i])
        R = Rg @ Rt.T
        xyz = numpy.stack([R @ (t[a] - ot) + og for a in rt.atom_names])
        out.append(xyz)
    return out


def sample_torsions(rts, rng):
    n = len(rts)
    helix = rng.random(n) < 0.7
    phis = numpy.where(helix, rng.normal(-63, 10, n), rng.normal(-120, 15, n))
    psis = numpy.where(helix, rng.normal(-41, 10, n), rng.normal(130, 15, n))
    omegas = numpy.full(n, 180.0)
    chis = []
    for rt in rts:
        d = {}
        if rt.base_name != "PRO":
            for name, _ in chi_definitions(rt):
                d[name] = float(rng.choice([60.0, 180.0, -60.0]) + rng.normal(0, 10))
        chis.append(d)
    return phis, psis, omegas, chis


def _nerf_v(a, b, c, bond, angle, torsion):
    """Vectorised nerf over leading axes (angles in radians)."""
    bc = c - b
    bc = bc / numpy.linalg.norm(bc, axis=-1, keepdims=True)
    n = numpy.cross(b - a, bc)
    n = n / numpy.linalg.norm(n, axis=-1, keepdims=True)
    m = numpy.cross(n, bc)
    bond = numpy.asarray(bond)[..., None]
    angle = numpy.asarray(angle)[..., None]
    torsion = numpy.asarray(torsion)[..., None]
    return c + bond * (
        -numpy.cos(angle) * bc
        + numpy.sin(angle) * numpy.cos(torsion) * m
        + numpy.sin(angle) * numpy.sin(torsion) * n
    )


def _templates_v(rt: ResidueType, phi, psi, chis):
    """Vectorised build_residue_template for m residues of one type.
    phi, psi: [m] degrees (ignored where the connection is absent);
    chis: dict name -> [m] degrees.  Returns name -> [m, 3]."""
    m = len(phi)
    overrides = {}
    if rt.down_connection_ind >= 0:
        overrides["down"] = numpy.asarray(phi)
    if rt.up_connection_ind >= 0:
        overrides["up"] = numpy.asarray(psi)
    defs = dict(chi_definitions(rt))
    for name, val in chis.items():
        overrides[defs[name][3]] = numpy.asarray(val)
    ics = rt.icoors
    pos = {}
    n0, n1, n2 = ics[0]["name"], ics[1]["name"], ics[2]["name"]
    pos[n0] = numpy.zeros((m, 3))
    pos[n1] = numpy.tile([ics[1]["d"], 0.0, 0.0], (m, 1))
    A = numpy.pi - ics[2]["theta"] * DEG
    if ics[2]["parent"] == n1:
        pos[n2] = pos[n1] + ics[2]["d"] * numpy.array([-numpy.cos(A), numpy.sin(A), 0.0])
    else:
        pos[n2] = pos[n0] + ics[2]["d"] * numpy.array([numpy.cos(A), numpy.sin(A), 0.0])
    for ic in ics[3:]:
        name = ic["name"]
        tors = overrides.get(name, numpy.full(m, ic["phi"]))
        pos[name] = _nerf_v(
            pos[ic["great_grand_parent"]], pos[ic["grand_parent"]], pos[ic["parent"]],
            numpy.full(m, ic["d"]), numpy.full(m, numpy.pi - ic["theta"] * DEG), tors * DEG,
        )
    return pos


def _frames_v(n, ca, c):
    x = ca - n
    x = x / numpy.linalg.norm(x, axis=-1, keepdims=True)
    y = (c - ca) - numpy.sum((c - ca) * x, axis=-1, keepdims=True) * x
    y = y / numpy.linalg.norm(y, axis=-1, keepdims=True)
    z = numpy.cross(x, y)
    return numpy.stack([x, y, z], axis=-1), ca  # R[..., :, k] = axis k


def _angle_v(a, b, c):
    u, v = a - b, c - b
    cosang = numpy.sum(u * v, -1) / numpy.linalg.norm(u, axis=-1) / numpy.linalg.norm(v, axis=-1)
    return numpy.arccos(numpy.clip(cosang, -1, 1))


def build_chains_fast(chains):
    """Vectorised build_chain over many chains.

    chains: list of (rts, phis, psis, omegas, chis_list).  Residue templates
    are built per residue type in one vectorised pass; each residue's
    backbone frame relative to its predecessor comes from a vectorised
    nerf; global frames are the running product of those relative
    transforms (one sequential 4x4 product per residue, batched over
    chains).  Same geometry as build_chain."""
    flat = []  # (chain, index)
    for ci, (rts, phis, psis, omegas, chis) in enumerate(chains):
        for i in range(len(rts)):
            flat.append((ci, i))
    nflat = len(flat)
    # per-residue local atom coordinates in its own N/CA/C frame
    local = [None] * nflat
    geo = numpy.zeros((nflat, 6))  # d_nca, d_cac, a_ncac, a_down_n_ca, d_c_up, a_ca_c_up
    upl = numpy.zeros((nflat, 3))
    cl = numpy.zeros((nflat, 3))
    by_type = {}
    for k, (ci, i) in enumerate(flat):
        by_type.setdefault(chains[ci][0][i].name, []).append(k)
    for name, ks in by_type.items():
        ks = numpy.array(ks)
        rt = chains[flat[ks[0]][0]][0][flat[ks[0]][1]]
        phi = numpy.array([chains[flat[k][0]][1][flat[k][1]] for k in ks], dtype=float)
        psi = numpy.array([chains[flat[k][0]][2][flat[k][1]] for k in ks], dtype=float)
        chi_names = [c for c, _ in chi_definitions(rt)]
        chis = {}
        for cn in chi_names:
            vals = [chains[flat[k][0]][4][flat[k][1]].get(cn) for k in ks]
            if all(v is not None for v in vals):
                chis[cn] = numpy.array(vals, dtype=float)
        t = _templates_v(rt, phi, psi, chis)
        R, o = _frames_v(t["N"], t["CA"], t["C"])
        Rt = numpy.transpose(R, (0, 2, 1))

        def to_local(x):
            return numpy.einsum("mij,mj->mi", Rt, x - o)

        atoms = numpy.stack([to_local(t[a]) for a in rt.atom_names], axis=1)
        geo[ks, 0] = numpy.linalg.norm(t["CA"] - t["N"], axis=-1)
        geo[ks, 1] = numpy.linalg.norm(t["C"] - t["CA"], axis=-1)
        geo[ks, 2] = _angle_v(t["N"], t["CA"], t["C"])
        if "down" in t:
            geo[ks, 3] = _angle_v(t["down"], t["N"], t["CA"])
        if "up" in t:
            geo[ks, 4] = numpy.linalg.norm(t["up"] - t["C"], axis=-1)
            geo[ks, 5] = _angle_v(t["CA"], t["C"], t["up"])
            upl[ks] = to_local(t["up"])
        cl[ks] = to_local(t["C"])
        for j, k in enumerate(ks):
            local[k] = atoms[j]
    out = []
    k0 = 0
    for ci, (rts, phis, psis, omegas, chis) in enumerate(chains):
        n = len(rts)
        ks = numpy.arange(k0, k0 + n)
        # relative transform residue i -> i+1 in frame i coordinates
        if n > 1:
            a = ks[:-1]
            b = ks[1:]
            Ni = numpy.zeros((n - 1, 3)) + upl[a]
            CAi = numpy.zeros((n - 1, 3))
            Ci = cl[a]
            Nn = Ni
            CAn = _nerf_v(CAi, Ci, Nn, geo[b, 0], geo[b, 3], numpy.asarray(omegas[:-1]) * DEG)
            Cn = _nerf_v(Ci, Nn, CAn, geo[b, 1], geo[b, 2], numpy.asarray(phis[1:]) * DEG)
            Rr, orr = _frames_v(Nn, CAn, Cn)
        Rg = numpy.zeros((n, 3, 3))
        og = numpy.zeros((n, 3))
        Rg[0] = numpy.eye(3)
        og[0] = [geo[ks[0], 0], 0.0, 0.0]  # N at the origin, CA on +x (as build_chain)
        for i in range(1, n):
            Rg[i] = Rg[i - 1] @ Rr[i - 1]
            og[i] = Rg[i - 1] @ orr[i - 1] + og[i - 1]
        out.append([local[k] @ Rg[i].T + og[i] for i, k in enumerate(ks)])
        k0 += n
    return out


def random_pose_spec(n_res: int, rng, db=None):
    db = db or default_chemical_database()
    seq = rng.integers(0, 20, n_res)
    names = [CANONICAL_ORDER[k] for k in seq]
    rts = [
        db.residue_type(nm, nterm=(i == 0), cterm=(i == n_res - 1))
        for i, nm in enumerate(names)
    ]
    phis, psis, omegas, chis = sample_torsions(rts, rng)
    return rts, phis, psis, omegas, chis


def random_pose(n_res: int, rng, db=None) -> Pose:
    rts, phis, psis, omegas, chis = random_pose_spec(n_res, rng, db)
    return Pose(rts, build_chain(rts, phis, psis, omegas, chis))


def generate_pose_specs(n_poses: int, res_range, seed: int = MASTER_SEED,
                        log_uniform: bool = False, db=None):
    """Residue types and torsions (rts, phis, psis, omegas, chis) of
    ``n_poses`` poses; residue counts uniform (or log-uniform) over the
    inclusive ``res_range``; pose k uses its own sub-seed."""
    lo, hi = res_range
    rng = numpy.random.default_rng(seed)
    if log_uniform:
        sizes = numpy.round(numpy.exp(rng.uniform(numpy.log(lo), numpy.log(hi), n_poses))).astype(int)
    else:
        sizes = rng.integers(lo, hi + 1, n_poses)
    return [random_pose_spec(int(s), numpy.random.default_rng([seed, k]), db)
            for k, s in enumerate(sizes)]


def generate_poses(n_poses: int, res_range, seed: int = MASTER_SEED,
                   log_uniform: bool = False, db=None) -> List[Pose]:
    """``n_poses`` poses built from generate_pose_specs."""
    specs = generate_pose_specs(n_poses, res_range, seed, log_uniform, db)
    coords = build_chains_fast(specs)
    return [Pose(sp[0], xyz) for sp, xyz in zip(specs, coords)]


# workload definitions (residue counts and pose counts) used by bench.py and
# the tests; sub-seeds per workload keep them independent
WORKLOADS = {
    "cpu_single_100": dict(n_poses=1, res_range=(100, 100), log_uniform=False, seed=MASTER_SEED + 0),
    "padded_64x150_600": dict(n_poses=64, res_range=(150, 600), log_uniform=False, seed=MASTER_SEED + 1),
    "ragged_512x50_1200": dict(n_poses=512, res_range=(50, 1200), log_uniform=True, seed=MASTER_SEED + 2),
    "diff_32x400": dict(n_poses=32, res_range=(400, 400), log_uniform=False, seed=MASTER_SEED + 3),
    "pruner_1024x300": dict(n_poses=1024, res_range=(300, 300), log_uniform=False, seed=MASTER_SEED + 4),
    # pruner training set: the 512-pose batch's distribution (log-uniform
    # 50-1200 residues); calibration set: poses like the held-out batch (300
    # residues); each with its own seed
    "pruner_train": dict(n_poses=192, res_range=(50, 1200), log_uniform=True, seed=MASTER_SEED + 5),
    "pruner_calib": dict(n_poses=128, res_range=(300, 300), log_uniform=False, seed=MASTER_SEED + 6),
    # rotamer-pair scoring (repacking): poses whose every residue gets
    # chi-grid rotamers (tmol_amd.rotamers)
    "repack_16x300": dict(n_poses=16, res_range=(300, 300), log_uniform=False, seed=MASTER_SEED + 7),
}


def workload_poses(name: str, n_poses: Optional[int] = None) -> List[Pose]:
    w = dict(WORKLOADS[name])
    if n_poses is not None:
        w["n_poses"] = n_poses
    return generate_poses(**w)


def workload_specs(name: str, n_poses: Optional[int] = None):
    """The torsion specs workload_poses builds its poses from."""
    w = dict(WORKLOADS[name])
    if n_poses is not None:
        w["n_poses"] = n_poses
    return generate_pose_specs(**w)
