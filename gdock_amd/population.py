"""population.rs / chromosome.rs: chromosomes and device population scoring."""
from __future__ import annotations

from dataclasses import dataclass, field

from . import constants
from .device import Complex
from .restraints import Restraint
from .structure import Molecule


@dataclass
class Chromosome:
    """chromosome::Chromosome (chromosome.rs:12-24)."""
    genes: list[float]
    fitness: float = 0.0
    vdw: float = 0.0
    elec: float = 0.0
    desolv: float = 0.0
    air: float = 0.0
    restraint_penalty: float = 0.0
    rmsd: float = 0.0
    fnat: float = 0.0


@dataclass
class Population:
    """population::Population; eval_fitness is the device drop-in (population.rs:48-61)."""
    chromosomes: list[Chromosome]
    receptor: Molecule
    ligand: Molecule
    restraints: list[Restraint] = field(default_factory=list)
    weights: constants.EnergyWeights = field(default_factory=constants.EnergyWeights)
    generation: int = 0
    _complex: Complex | None = None

    def _cx(self) -> Complex:
        if self._complex is None:
            self._complex = Complex(self.receptor, self.ligand, [r.pair for r in self.restraints])
        return self._complex

    def eval_fitness(self) -> None:
        self._cx().eval_fitness(self.chromosomes, self.weights)

    def get_min_fittest(self) -> Chromosome:
        return min(self.chromosomes, key=lambda c: c.fitness)

    def size(self) -> int:
        return len(self.chromosomes)
