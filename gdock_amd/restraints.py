"""restraints.rs: residue-pair restraints between receptor and ligand C-alpha atoms."""
from __future__ import annotations

from dataclasses import dataclass

from .structure import Molecule


@dataclass(frozen=True)
class Restraint:
    """restraints::Restraint, kept as the residue numbers the reference matches on."""
    rec_resseq: int
    lig_resseq: int

    @property
    def pair(self) -> tuple[int, int]:
        return (self.rec_resseq, self.lig_resseq)


def _first_ca(m: Molecule, resseq: int):
    for i, a in enumerate(m.atoms):
        if a["name"].decode().strip() == "CA" and int(a["resseq"]) == resseq:
            return i
    return None


def create_restraints_from_pairs(mol1: Molecule, mol2: Molecule, pairs) -> list[Restraint]:
    """restraints.rs:44-77: keep pairs whose C-alpha exists on both sides."""
    out = []
    for r1, r2 in pairs:
        if _first_ca(mol1, r1) is not None and _first_ca(mol2, r2) is not None:
            out.append(Restraint(int(r1), int(r2)))
    return out
