"""CPU tests of the synthetic pose generator (tmol_amd/generator.py).

The vectorised builder that makes every workload (``build_chains_fast``) is
pinned three ways:

* against the per-residue builder ``build_chain`` on chains that hold every
  canonical residue type, both termini variants included, to 1e-9 A;
* against the geometry the residue-type database fixes: every bond length
  and bond angle of the icoor tree (``d`` and 180 deg - ``theta``), within
  and across residues (the peptide bond through the up/down connections);
* against the torsions it was asked for: phi, psi, omega and chi1..chi4
  measured on the built coordinates equal the sampled values.

The sampler's distributions (70 / 30 helix / strand, N(mean, sd) per state,
chi centres 60 / 180 / -60) are checked on a large seeded sample.
"""

import numpy

from tmol_amd.chemical import default_chemical_database
from tmol_amd.generator import (CANONICAL_ORDER, DEG, build_chain, build_chains_fast,
                                chi_definitions, dihedral, generate_pose_specs, sample_torsions)


def _angle(a, b, c):
    u, v = a - b, c - b
    return numpy.arccos(numpy.clip(numpy.dot(u, v) / numpy.linalg.norm(u) / numpy.linalg.norm(v), -1, 1))


def _angdiff(a, b):
    return abs((a - b + 180.0) % 360.0 - 180.0)


def _every_type_chains(seed=11):
    """Three chains that together hold every canonical residue type, each as
    N-terminus, interior residue and C-terminus, with sampled torsions."""
    db = default_chemical_database()
    rng = numpy.random.default_rng(seed)
    chains = []
    names = list(CANONICAL_ORDER)
    for shift in range(3):
        order = names[shift:] + names[:shift]
        seq = order + order[:7]
        rts = [db.residue_type(nm, nterm=(i == 0), cterm=(i == len(seq) - 1))
               for i, nm in enumerate(seq)]
        chains.append((rts,) + tuple(sample_torsions(rts, rng)))
    # every residue type also as an N-terminus and as a C-terminus
    for nm in names:
        for rts in ([db.residue_type(nm, nterm=True), db.residue_type("GLY", cterm=True)],
                    [db.residue_type("ALA", nterm=True), db.residue_type(nm, cterm=True)]):
            chains.append((rts,) + tuple(sample_torsions(rts, rng)))
    return chains


def test_fast_builder_equals_reference_builder():
    chains = _every_type_chains()
    fast = build_chains_fast(chains)
    seen = set()
    for (rts, phis, psis, omegas, chis), xyz_fast in zip(chains, fast):
        xyz = build_chain(rts, phis, psis, omegas, chis)
        assert len(xyz) == len(xyz_fast) == len(rts)
        for rt, a, b in zip(rts, xyz, xyz_fast):
            assert a.shape == b.shape == (rt.n_atoms, 3)
            assert numpy.abs(a - b).max() < 1e-9, rt.name
            seen.add(rt.name)
    # 20 interior types plus N- and C-terminal variants of each
    assert len(seen) >= 60


def test_ideal_bond_lengths_and_angles():
    """Every icoor edge (atom, parent) has the database's bond length and
    every (atom, parent, grand parent) its bond angle, inside a residue and
    across the peptide bond."""
    chains = _every_type_chains(seed=12)
    fast = build_chains_fast(chains)
    n_checked = 0
    for (rts, *_), xyz in zip(chains, fast):
        for i, (rt, X) in enumerate(zip(rts, xyz)):
            pos = {nm: X[k] for k, nm in enumerate(rt.atom_names)}
            if i > 0:
                # "down" is the previous residue's C; "up" the next one's N
                pos["down"] = xyz[i - 1][rts[i - 1].atom_index["C"]]
            if i + 1 < len(rts):
                pos["up"] = xyz[i + 1][rts[i + 1].atom_index["N"]]
            for ic in rt.icoors:
                a, p, gp = ic["name"], ic["parent"], ic["grand_parent"]
                if a == p or a not in pos or p not in pos:
                    continue
                d = numpy.linalg.norm(pos[a] - pos[p])
                # the peptide bond is built from the previous residue's "up"
                # icoor; residue types' up and down icoors agree to 1e-6 A
                tol = 5e-6 if "up" in (a, p) or "down" in (a, p) else 1e-9
                assert abs(d - ic["d"]) < tol, (rt.name, a, p, d, ic["d"])
                n_checked += 1
                if gp in pos and gp != a and ic["theta"] != 180.0:
                    ang = _angle(pos[gp], pos[p], pos[a]) / DEG
                    tol = 1e-4 if "up" in (a, p, gp) or "down" in (a, p, gp) else 1e-7
                    assert abs(ang - (180.0 - ic["theta"])) < tol, (rt.name, a, p, gp, ang)
                    n_checked += 1
    assert n_checked > 2000


def test_built_torsions_equal_sampled_torsions():
    chains = _every_type_chains(seed=13)
    fast = build_chains_fast(chains)
    n_checked = 0
    for (rts, phis, psis, omegas, chis), xyz in zip(chains, fast):
        def at(i, name):
            return xyz[i][rts[i].atom_index[name]]

        for i, rt in enumerate(rts):
            if i > 0:
                phi = dihedral(at(i - 1, "C"), at(i, "N"), at(i, "CA"), at(i, "C")) / DEG
                assert _angdiff(phi, phis[i]) < 1e-6
                n_checked += 1
            if i + 1 < len(rts):
                psi = dihedral(at(i, "N"), at(i, "CA"), at(i, "C"), at(i + 1, "N")) / DEG
                omega = dihedral(at(i, "CA"), at(i, "C"), at(i + 1, "N"), at(i + 1, "CA")) / DEG
                assert _angdiff(psi, psis[i]) < 1e-6
                assert _angdiff(omega, omegas[i]) < 1e-6
                n_checked += 2
            for name, quad in chi_definitions(rt):
                if name not in chis[i]:
                    continue
                got = dihedral(*[at(i, q) for q in quad]) / DEG
                assert _angdiff(got, chis[i][name]) < 1e-6, (rt.name, name)
                n_checked += 1
    assert n_checked > 300


def test_sampler_distributions():
    """Helix fraction 0.7; helix and strand phi/psi means and spreads; chi
    values centred on 60 / 180 / -60 with sd 10; proline ring chis kept."""
    specs = generate_pose_specs(40, (150, 150), seed=5)
    phis = numpy.concatenate([s[1] for s in specs])
    psis = numpy.concatenate([s[2] for s in specs])
    omegas = numpy.concatenate([s[3] for s in specs])
    helix = psis < 45.0  # helix psi ~ -41 +- 10, strand psi ~ 130 +- 15
    n = len(phis)
    assert n == 6000
    frac = helix.mean()
    assert abs(frac - 0.7) < 4 * numpy.sqrt(0.21 / n)
    for sel, mp, sp, ms, ss in ((helix, -63, 10, -41, 10), (~helix, -120, 15, 130, 15)):
        m = sel.sum()
        assert abs(phis[sel].mean() - mp) < 5 * sp / numpy.sqrt(m)
        assert abs(psis[sel].mean() - ms) < 5 * ss / numpy.sqrt(m)
        assert abs(phis[sel].std() / sp - 1) < 0.1
        assert abs(psis[sel].std() / ss - 1) < 0.1
    assert (omegas == 180.0).all()
    dev, centres = [], {60.0: 0, 180.0: 0, -60.0: 0}
    for s in specs:
        for rt, chis in zip(s[0], s[4]):
            if rt.base_name == "PRO":
                assert chis == {}
            for v in chis.values():
                c = min(centres, key=lambda c: _angdiff(v, c))
                centres[c] += 1
                dev.append((v - c + 180.0) % 360.0 - 180.0)
    dev = numpy.array(dev)
    assert abs(dev.std() / 10.0 - 1) < 0.1 and abs(dev.mean()) < 1.0
    tot = sum(centres.values())
    for c in centres.values():
        assert abs(c / tot - 1 / 3) < 0.05
