"""Extract the parameter tables the LJ+LK path needs from tmol's YAML database.

Reads (only when run by hand, in a checkout that has the reference tree):
  tmol/database/default/chemical/chemical.yaml  (atom types, canonical residues,
                                                 the alpha-amino-acid termini variants)
  tmol/database/default/scoring/ljlk.yaml       (LJ/LK global + per-atom-type params)
  tmol/database/default/scoring/elec.yaml       (elec global params, partial charges
                                                 and count-pair representatives)
  tmol/database/default/scoring/hbond.yaml      (donor / acceptor atom types, which
                                                 place lk_ball's waters; the hbond
                                                 potential's weights, polynomials and
                                                 donor x acceptor pairs)
and writes ``default_db.json`` next to this file.  The JSON is data only: the
same numbers, names and connectivity as the YAML, restricted to what the
20 canonical amino acids with termini need.  Nothing at run time reads the
reference tree.

Usage:  python -m tmol_amd.database.extract_reference_db [/path/to/reference]
"""

import json
import os
import sys

import yaml

CANONICAL = [
    "ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
    "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
]
# residue types beyond the 20 sampled ones: the disulfide-bonded cysteine
# (a third connection, "dslf" on SG), for poses with side-chain cross-links
EXTRA = ["CYD"]
VARIANTS = ["CarboxyTerminus", "AminoTerminus", "AminoTerminusPro"]


def _angle(v):
    # "63.8 deg" -> 63.8 (degrees); plain numbers are already degrees
    if isinstance(v, str):
        v = v.strip()
        assert v.endswith("deg"), v
        return float(v[:-3])
    return float(v)


def _icoor(ic):
    out = {
        "name": ic["name"],
        "phi": _angle(ic["phi"]),
        "theta": _angle(ic["theta"]),
        "d": float(ic["d"]),
        "parent": ic["parent"],
        "grand_parent": ic["grand_parent"],
        "great_grand_parent": ic["great_grand_parent"],
    }
    if "source" in ic:
        out["source"] = ic["source"]
    return out


def extract(ref_root):
    chem = yaml.safe_load(
        open(os.path.join(ref_root, "tmol/database/default/chemical/chemical.yaml"))
    )
    ljlk = yaml.safe_load(
        open(os.path.join(ref_root, "tmol/database/default/scoring/ljlk.yaml"))
    )

    atom_types = []
    for at in chem["atom_types"]:
        atom_types.append(
            {
                "name": at["name"],
                "element": at["element"],
                "is_acceptor": bool(at.get("is_acceptor", False)),
                "is_donor": bool(at.get("is_donor", False)),
                "is_hydroxyl": bool(at.get("is_hydroxyl", False)),
                "is_polarh": bool(at.get("is_polarh", False)),
                "acceptor_hybridization": at.get("acceptor_hybridization", None),
            }
        )

    residues = []
    by_name = {r["name"]: r for r in chem["residues"]}
    for name in CANONICAL + EXTRA:
        r = by_name[name]
        residues.append(
            {
                "name": r["name"],
                "name3": r["name3"],
                "one_letter_code": r.get("one_letter_code"),
                "atoms": [[a["name"], a["atom_type"]] for a in r["atoms"]],
                "atom_aliases": [[a["name"], a["alt_name"]] for a in r["atom_aliases"]],
                "bonds": [[b[0], b[1]] for b in r["bonds"]],
                "connections": [[c["name"], c["atom"]] for c in r["connections"]],
                "torsions": r["torsions"],
                "icoors": [_icoor(ic) for ic in r["icoors"]],
                "mainchain_atoms": r["properties"]["polymer"]["mainchain_atoms"],
            }
        )

    variants = []
    vby = {v["name"]: v for v in chem["variants"]}
    for name in VARIANTS:
        v = vby[name]
        variants.append(
            {
                "name": v["name"],
                "display_name": v["display_name"],
                "pattern": v["pattern"],
                "remove_atoms": v["remove_atoms"],
                "add_atoms": [[a["name"], a["atom_type"]] for a in v["add_atoms"]],
                "add_atom_aliases": [
                    [a["name"], a["alt_name"]] for a in v["add_atom_aliases"]
                ],
                "modify_atoms": [[a["name"], a["atom_type"]] for a in v["modify_atoms"]],
                "add_bonds": [[b[0], b[1]] for b in v["add_bonds"]],
                "icoors": [_icoor(ic) for ic in v["icoors"]],
            }
        )

    g = ljlk["global_parameters"]
    ljlk_out = {
        "global": {
            k: float(g[k])
            for k in (
                "max_dis",
                "lj_dlin_sigma_factor",
                "lj_dlin_sigma_factor_soft",
                "lj_hbond_OH_donor_dis",
                "lj_hbond_hdis",
                "lj_hbond_dis",
                "lk_min_dis2sigma",
                "lkb_water_dist",
            )
        },
        # lk_ball water geometry, degrees as written (ljlk.yaml: "134.5 deg")
        "lkb_water": {
            k: (
                [_angle(x) for x in g[k]] if isinstance(g[k], list) else _angle(g[k])
            )
            for k in (
                "lkb_water_angle_sp2",
                "lkb_water_angle_sp3",
                "lkb_water_angle_ring",
                "lkb_water_tors_sp2",
                "lkb_water_tors_sp3",
                "lkb_water_tors_ring",
            )
        },
        "atom_type_parameters": [
            {
                "name": p["name"],
                "lj_radius": float(p["lj_radius"]),
                "lj_wdepth": float(p["lj_wdepth"]),
                "lk_dgfree": float(p["lk_dgfree"]),
                "lk_lambda": float(p["lk_lambda"]),
                "lk_volume": float(p["lk_volume"]),
            }
            for p in ljlk["atom_type_parameters"]
        ],
    }
    elec = yaml.safe_load(
        open(os.path.join(ref_root, "tmol/database/default/scoring/elec.yaml"))
    )
    keep = set(CANONICAL + EXTRA)

    def _res(tag):
        # "ALA" / "ALA:cterm" -> ("ALA", "" / "cterm")
        parts = tag.split(":")
        return parts[0], parts[1] if len(parts) > 1 else ""

    elec_out = {
        "global": {k: float(v) for k, v in elec["global_parameters"].items()},
        "atom_charge_parameters": [
            [*_res(x["res"]), x["atom"], float(x["charge"])]
            for x in elec["atom_charge_parameters"]
            if _res(x["res"])[0] in keep
        ],
        "atom_cp_reps_parameters": [
            [*_res(x["res"]), x["atm_inner"], x["atm_outer"]]
            for x in elec["atom_cp_reps_parameters"]
            if _res(x["res"])[0] in keep
        ],
    }
    hb = yaml.safe_load(
        open(os.path.join(ref_root, "tmol/database/default/scoring/hbond.yaml"))
    )
    poly_keys = ["c_%d" % n for n in range(11)]
    hbond_out = {
        "donor_atom_types": [[x["d"], x["donor_type"]] for x in hb["donor_atom_types"]],
        "acceptor_atom_types": [
            [x["a"], x["acceptor_type"]] for x in hb["acceptor_atom_types"]
        ],
        # the hbond potential (hbond.yaml; tmol/database/scoring/hbond.py):
        # type weights in file order (the type indices), the polynomials
        # (coefficients c_0..c_10, absent ones 0), the donor x acceptor pairs
        "global_parameters": dict(hb["global_parameters"]),
        "donor_type_params": [[x["name"], x["weight"]] for x in hb["donor_type_params"]],
        "acceptor_type_params": [[x["name"], x["weight"]] for x in hb["acceptor_type_params"]],
        "polynomial_parameters": [
            {
                "name": x["name"], "dimension": x["dimension"], "xmin": x["xmin"],
                "xmax": x["xmax"], "min_val": x["min_val"], "max_val": x["max_val"],
                "degree": x["degree"], "c": [float(x.get(k, 0.0)) for k in poly_keys],
            }
            for x in hb["polynomial_parameters"]
        ],
        "pair_parameters": [
            [x["donor_type"], x["acceptor_type"], x["AHdist"], x["cosBAH"], x["cosAHD"]]
            for x in hb["pair_parameters"]
        ],
    }
    return {
        "source": "tmol/database/default (chemical.yaml, scoring/ljlk.yaml, scoring/elec.yaml,"
        " scoring/hbond.yaml)",
        "atom_types": atom_types,
        "residues": residues,
        "variants": variants,
        "ljlk": ljlk_out,
        "elec": elec_out,
        "hbond": hbond_out,
    }


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "default_db.json")
    with open(out, "w") as f:
        json.dump(extract(ref), f, indent=1)
    print("wrote", out)
