"""Minimal PDB reader: canonical amino-acid chains -> ``Pose``.

Covers what the reference's PDB path does for fully-protonated canonical
proteins such as 1ubq (tmol/io/pdb_parsing.py, canonical_ordering.py:
437-614, pose_stack_construction.py:15-387): atom names resolved through the
residue types' ``atom_aliases``; the first residue of a chain gets the
N-terminus variant and the last the C-terminus variant; HIS is the HE2
tautomer.  Residue subsets follow ``res_not_connected``
(tmol/tests/score/common/test_energy_term.py:104-137): a subset start that is
not the chain start keeps its non-terminal type and is left unbonded below,
a subset end that is not the chain end keeps its non-terminal type and is left
unbonded above.  Missing atoms are not built: every atom of the chosen type
must be present.
"""

from typing import List, Sequence, Tuple

import numpy

from .chemical import default_chemical_database
from .pose import Pose


def parse_pdb_residues(pdb_text: str):
    """Returns a list of residues: (chain, resseq, icode, name3, {atom: xyz})."""
    residues = []
    index = {}
    for line in pdb_text.splitlines():
        if not (line.startswith("ATOM") or line.startswith("HETATM")):
            continue
        name = line[12:16].strip()
        alt = line[16]
        if alt not in (" ", "A"):
            continue
        res3 = line[17:20].strip()
        chain = line[21]
        resseq = int(line[22:26])
        icode = line[26]
        xyz = numpy.array(
            [float(line[30:38]), float(line[38:46]), float(line[46:54])],
            dtype=numpy.float64,
        )
        key = (chain, resseq, icode)
        if key not in index:
            index[key] = len(residues)
            residues.append([chain, resseq, icode, res3, {}])
        residues[index[key]][4][name] = xyz
    return residues


def pose_from_pdb(
    pdb_text: str,
    res_range=None,
    db=None,
) -> Pose:
    """Build one pose from a PDB; ``res_range=(i, j)`` keeps residues i..j-1
    (0-based over the whole file), a list of such ranges keeps each segment
    (tmol/tests/score/common/test_energy_term.py:104-137), with the
    reference's subset termini rule; residues of different segments are not
    bonded."""
    db = db or default_chemical_database()
    recs = parse_pdb_residues(pdb_text)
    n_all = len(recs)
    if res_range is None:
        ranges = [(0, n_all)]
    elif isinstance(res_range[0], (tuple, list)):
        ranges = [tuple(r) for r in res_range]
    else:
        ranges = [tuple(res_range)]
    sel = [(gi, k == 0) for i0, i1 in ranges for k, gi in enumerate(range(i0, i1))]
    rts = []
    coords = []
    connected = []
    for k, (gi, seg_start) in enumerate(sel):
        chain, resseq, icode, res3, atoms = recs[gi]
        chain_start = gi == 0 or recs[gi - 1][0] != chain
        chain_end = gi == n_all - 1 or recs[gi + 1][0] != chain
        nterm = chain_start
        cterm = chain_end
        rt = db.residue_type(res3, nterm=nterm, cterm=cterm)
        xyz = numpy.zeros((rt.n_atoms, 3), dtype=numpy.float64)
        found = numpy.zeros(rt.n_atoms, dtype=bool)
        for nm, x in atoms.items():
            canon = rt.aliases.get(nm, nm)
            if canon not in rt.atom_index:
                raise ValueError("atom %s not in %s (%s %d)" % (nm, rt.name, chain, resseq))
            xyz[rt.atom_index[canon]] = x
            found[rt.atom_index[canon]] = True
        if not found.all():
            missing = [rt.atom_names[a] for a in numpy.nonzero(~found)[0]]
            raise ValueError("missing atoms %s in %s %d" % (missing, rt.name, resseq))
        rts.append(rt)
        coords.append(xyz)
        connected.append(k > 0 and not chain_start and not seg_start)
    return Pose(rts, coords, connected)
