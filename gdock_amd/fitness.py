"""fitness.rs energy functions on the device (one pose scored as given).

Each call uploads the pair as a complex and scores the ligand coordinates
unmoved, which is how commands/score.rs:35-46 evaluates a model.  Population
scoring goes through Complex.score_poses / Population.eval_fitness instead.
"""
from __future__ import annotations

from .device import Complex
from .restraints import Restraint
from .structure import Molecule


def _score(receptor: Molecule, ligand: Molecule, pairs=()):
    cx = Complex(receptor, ligand, pairs)
    try:
        out, ns = cx.score_coords(ligand.xyz[None])
        return out[0], int(ns[0]), cx.n_restraints
    finally:
        cx.close()


def vdw_energy(receptor: Molecule, ligand: Molecule) -> float:
    """fitness.rs:86-105."""
    return float(_score(receptor, ligand)[0][0])


def elec_energy(receptor: Molecule, ligand: Molecule) -> float:
    """fitness.rs:284-302."""
    return float(_score(receptor, ligand)[0][1])


def desolv_energy(receptor: Molecule, ligand: Molecule) -> float:
    """fitness.rs:122-184."""
    return float(_score(receptor, ligand)[0][2])


def air_energy(restraints: list[Restraint], receptor: Molecule, ligand: Molecule) -> float:
    """fitness.rs:205-249 (restraints located by residue number, as the reference does)."""
    return float(_score(receptor, ligand, [r.pair for r in restraints])[0][3])


def satisfaction_ratio(restraints: list[Restraint], receptor: Molecule, ligand: Molecule) -> float:
    """fitness.rs:251-267."""
    out, ns, n = _score(receptor, ligand, [r.pair for r in restraints])
    return ns / n if n else 0.0
