"""CPU oracle for the LJ+LK path -- test infrastructure only (see oracle.py)."""
