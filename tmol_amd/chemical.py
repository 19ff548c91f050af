"""Residue-type chemistry for the LJ+LK path: atom types, canonical residues,
termini variants, intra-residue bond separation.

The parameter tables are the reference's own (``tmol/database/default``),
extracted into ``database/default_db.json``.  The algorithms here are written
for this package:

* atom-type index = database order + one trailing "None" row
  (tmol/score/chemical_database.py:84-87);
* path distance = unweighted shortest path over intra-residue bonds, capped at
  MAX_SIG_BOND_SEPARATION = 6 (tmol/chemical/restypes.py:403-418,
  tmol/chemical/constants.py);
* heavy atoms = atoms whose type element is not H
  (tmol/score/atom_type_dependent_term.py:60-83);
* termini variants for alpha amino acids (chemical.yaml ``variants``:
  CarboxyTerminus, AminoTerminus, AminoTerminusPro) resolved for the
  N-CA-C(=O) backbone.
"""

import json
import os
from collections import deque
from typing import Dict, List, Optional, Tuple

import numpy

MAX_SIG_BOND_SEPARATION = 6
_DB_PATH = os.path.join(os.path.dirname(__file__), "database", "default_db.json")


class AtomTypeSet:
    """Atom types in database order plus a trailing ``None`` entry."""

    def __init__(self, atom_types: List[dict]):
        self.names = [a["name"] for a in atom_types]
        self.index = {n: i for i, n in enumerate(self.names)}
        self.n_types = len(self.names) + 1  # trailing None / NaN row
        self.element = [a["element"] for a in atom_types] + [None]

        def flag(key):
            return numpy.array(
                [bool(a[key]) for a in atom_types] + [False], dtype=bool
            )

        self.is_acceptor = flag("is_acceptor")
        self.is_donor = flag("is_donor")
        self.is_hydroxyl = flag("is_hydroxyl")
        self.is_polarh = flag("is_polarh")
        self.is_hydrogen = numpy.array(
            [e == "H" for e in self.element[:-1]] + [False], dtype=bool
        )

    def type_index(self, name: str) -> int:
        return self.index[name]


class ResidueType:
    """One block type: atoms, bonds, connections, internal coordinates."""

    def __init__(
        self,
        name: str,
        name3: str,
        atoms: List[Tuple[str, str]],
        bonds: List[Tuple[str, str]],
        connections: List[Tuple[str, str]],
        icoors: List[dict],
        torsions: List[dict],
        aliases: Dict[str, str],
        atom_type_set: AtomTypeSet,
        base_name: Optional[str] = None,
        is_nterm: bool = False,
        is_cterm: bool = False,
    ):
        self.name = name
        self.name3 = name3
        self.base_name = base_name or name
        self.is_nterm = is_nterm
        self.is_cterm = is_cterm
        self.atom_names = [a[0] for a in atoms]
        self.atom_type_names = [a[1] for a in atoms]
        self.n_atoms = len(atoms)
        self.atom_index = {n: i for i, n in enumerate(self.atom_names)}
        self.bonds = numpy.array(
            [[self.atom_index[a], self.atom_index[b]] for a, b in bonds],
            dtype=numpy.int32,
        ).reshape(-1, 2)
        self.connections = list(connections)
        self.connection_names = [c[0] for c in connections]
        self.conn_atom = numpy.array(
            [self.atom_index[c[1]] for c in connections], dtype=numpy.int32
        )
        self.icoors = icoors
        self.torsions = torsions
        # alias (as written in PDB files) -> canonical atom name
        self.aliases = dict(aliases)
        self.is_ligand_fragment = False
        self.is_polymer = True

        ats = atom_type_set
        self.atom_types = numpy.array(
            [ats.type_index(t) for t in self.atom_type_names], dtype=numpy.int32
        )
        self.heavy_atom_inds = numpy.nonzero(
            numpy.logical_not(ats.is_hydrogen[self.atom_types])
        )[0].astype(numpy.int32)
        self.path_distance = self._path_distance()

    def _path_distance(self) -> numpy.ndarray:
        n = self.n_atoms
        adj = [[] for _ in range(n)]
        for a, b in self.bonds:
            adj[a].append(b)
            adj[b].append(a)
        pd = numpy.full((n, n), MAX_SIG_BOND_SEPARATION, dtype=numpy.int32)
        for s in range(n):
            pd[s, s] = 0
            q = deque([s])
            while q:
                u = q.popleft()
                du = pd[s, u]
                if du >= MAX_SIG_BOND_SEPARATION:
                    continue
                for v in adj[u]:
                    if v != s and pd[s, v] == MAX_SIG_BOND_SEPARATION and du + 1 < pd[s, v]:
                        pd[s, v] = du + 1
                        q.append(v)
        return pd

    def connection_index(self, name: str) -> int:
        try:
            return self.connection_names.index(name)
        except ValueError:
            return -1

    @property
    def down_connection_ind(self) -> int:
        return self.connection_index("down")

    @property
    def up_connection_ind(self) -> int:
        return self.connection_index("up")

    def __repr__(self):
        return "ResidueType(%s, n_atoms=%d)" % (self.name, self.n_atoms)


def _resolve(token: str, mapping: Dict[str, str]) -> str:
    if token.startswith("<") and token.endswith(">"):
        return mapping[token]
    return token


class ChemicalDatabase:
    """Canonical residue types with termini variants, built on demand."""

    def __init__(self, path: str = _DB_PATH):
        with open(path) as f:
            db = json.load(f)
        self.raw = db
        self.atom_type_set = AtomTypeSet(db["atom_types"])
        self.ljlk = db["ljlk"]
        self._base = {r["name"]: r for r in db["residues"]}
        self._variants = {v["name"]: v for v in db["variants"]}
        self.one_letter = {r["one_letter_code"]: r["name"] for r in db["residues"]
                           if r.get("one_letter_code")}
        self._cache: Dict[Tuple[str, bool, bool], ResidueType] = {}

    @property
    def canonical_names(self) -> List[str]:
        return list(self._base.keys())

    def residue_type(self, name3: str, nterm: bool = False, cterm: bool = False):
        key = (name3, bool(nterm), bool(cterm))
        if key not in self._cache:
            self._cache[key] = self._build(name3, nterm, cterm)
        return self._cache[key]

    def _build(self, name3: str, nterm: bool, cterm: bool) -> ResidueType:
        r = self._base[name3]
        atoms = [tuple(a) for a in r["atoms"]]
        bonds = [tuple(b) for b in r["bonds"]]
        conns = [tuple(c) for c in r["connections"]]
        icoors = [dict(ic) for ic in r["icoors"]]
        aliases = {alt: nm for nm, alt in r["atom_aliases"]}
        suffix = []

        def remove_atom(nm):
            nonlocal atoms, bonds, icoors
            atoms = [a for a in atoms if a[0] != nm]
            bonds = [b for b in bonds if nm not in b]
            icoors = [ic for ic in icoors if ic["name"] != nm]

        def remove_conn(cname):
            nonlocal conns, icoors
            conns = [c for c in conns if c[0] != cname]
            icoors = [ic for ic in icoors if ic["name"] != cname]

        def apply(variant, mapping):
            nonlocal atoms, bonds, icoors, aliases
            for tok in variant["remove_atoms"]:
                t = _resolve(tok, mapping)
                if t in ("up", "down"):
                    remove_conn(t)
                else:
                    remove_atom(t)
            for nm, at in variant["modify_atoms"]:
                nm = _resolve(nm, mapping)
                atoms = [(a[0], at if a[0] == nm else a[1]) for a in atoms]
            for nm, at in variant["add_atoms"]:
                atoms.append((nm, at))
            for nm, alt in variant["add_atom_aliases"]:
                aliases[alt] = nm
            for a, b in variant["add_bonds"]:
                bonds.append((_resolve(a, mapping), _resolve(b, mapping)))
            for ic in variant["icoors"]:
                ic = dict(ic)
                for k in ("parent", "grand_parent", "great_grand_parent"):
                    ic[k] = _resolve(ic[k], mapping)
                ic.pop("source", None)
                icoors.append(ic)

        if nterm:
            if name3 == "PRO":
                # 'CN(C)[{down}]': N bonded to CA (C1) and CD (C2)
                apply(
                    self._variants["AminoTerminusPro"],
                    {"<{down}>": "down", "<N1>": "N", "<C1>": "CA", "<C2>": "CD"},
                )
            else:
                # '[*][*][NH][{down}]': *2 = CA, *1 = C, H1 = backbone H
                apply(
                    self._variants["AminoTerminus"],
                    {"<{down}>": "down", "<H1>": "H", "<N1>": "N", "<*2>": "CA", "<*1>": "C"},
                )
            suffix.append("nterm")
        if cterm:
            # '[*][*]C(=O)[{up}]': *1 = N, *2 = CA, C1 = C, O1 = O
            apply(
                self._variants["CarboxyTerminus"],
                {"<{up}>": "up", "<O1>": "O", "<C1>": "C", "<*2>": "CA", "<*1>": "N"},
            )
            suffix.append("cterm")

        name = name3 if not suffix else name3 + ":" + ":".join(suffix)
        return ResidueType(
            name=name,
            name3=r["name3"],
            atoms=atoms,
            bonds=bonds,
            connections=conns,
            icoors=icoors,
            torsions=r["torsions"],
            aliases=aliases,
            atom_type_set=self.atom_type_set,
            base_name=name3,
            is_nterm=nterm,
            is_cterm=cterm,
        )


_default_db: Optional[ChemicalDatabase] = None


def default_chemical_database() -> ChemicalDatabase:
    global _default_db
    if _default_db is None:
        _default_db = ChemicalDatabase()
    return _default_db
