"""clustering.rs: FCC contacts and clustering on the device."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .device import Complex


@dataclass(frozen=True)
class ClusteringConfig:
    """clustering.rs:15-35."""
    contact_distance: float = 5.0
    fcc_cutoff: float = 0.60
    strictness: float = 0.75
    min_cluster_size: int = 4


@dataclass
class ClusterResult:
    """clustering.rs:39-46."""
    center_idx: int
    members: list[int] = field(default_factory=list)
    size: int = 0


def calculate_contacts(cx: Complex, genes, contact_distance: float = 5.0) -> list[set[str]]:
    """Contact strings of receptor + (ligand moved by each gene row), clustering.rs:62-115."""
    bits, _ = cx.contacts(np.asarray(genes).reshape(-1, 6), contact_distance)
    return [cx.contact_set(b, contact_distance) for b in bits]


def cluster_structures(cx: Complex, genes, config: ClusteringConfig | None = None) -> list[ClusterResult]:
    """clustering.rs:265-293 over the complexes receptor + (ligand moved by genes i)."""
    return cx.cluster(np.asarray(genes).reshape(-1, 6), config or ClusteringConfig())
