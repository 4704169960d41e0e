"""Seeded synthetic receptor/ligand complexes and pose sets.

A chain is a random-walk C-alpha trace (3.8 A steps) confined to a globule;
every residue gets the heavy atoms of its OPLS template (residue names drawn
uniformly from the 20 standard amino acids, HIS using the HSD atom names).
Backbone N/C/O and side-chain atoms are placed at bond-length offsets along
random directions biased away from the globule centre.  The geometry is
project-defined (the templates carry names and types, not coordinates);
every quantity is a function of the seed (numpy PCG64).

Complexes are emitted as PDB text (receptor chain A, ligand chain B, ligand
serials continuing after the receptor's) so that both the product and the
oracle parse and parameterize them independently.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Heavy atoms of the toppar residue templates (toppar.rs:152-581), backbone first.
HEAVY_TEMPLATES = {
    "ALA": ["N", "CA", "C", "O", "CB"],
    "ARG": ["N", "CA", "C", "O", "CB", "CG", "CD", "NE", "CZ", "NH1", "NH2"],
    "ASN": ["N", "CA", "C", "O", "CB", "CG", "OD1", "ND2"],
    "ASP": ["N", "CA", "C", "O", "CB", "CG", "OD1", "OD2"],
    "CYS": ["N", "CA", "C", "O", "CB", "SG"],
    "GLN": ["N", "CA", "C", "O", "CB", "CG", "CD", "OE1", "NE2"],
    "GLU": ["N", "CA", "C", "O", "CB", "CG", "CD", "OE1", "OE2"],
    "GLY": ["N", "CA", "C", "O"],
    "HIS": ["N", "CA", "C", "O", "CB", "ND1", "CG", "CE1", "NE2", "CD2"],
    "ILE": ["N", "CA", "C", "O", "CB", "CG2", "CG1", "CD"],
    "LEU": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CD2"],
    "LYS": ["N", "CA", "C", "O", "CB", "CG", "CD", "CE", "NZ"],
    "MET": ["N", "CA", "C", "O", "CB", "CG", "SD", "CE"],
    "PHE": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CE1", "CZ", "CD2", "CE2"],
    "PRO": ["N", "CA", "C", "O", "CD", "CB", "CG"],
    "SER": ["N", "CA", "C", "O", "CB", "OG"],
    "THR": ["N", "CA", "C", "O", "CB", "OG1", "CG2"],
    "TRP": ["N", "CA", "C", "O", "CB", "CG", "CD1", "NE1", "CE2", "CD2", "CE3", "CZ3", "CZ2", "CH2"],
    "TYR": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CE1", "CZ", "OH", "CD2", "CE2"],
    "VAL": ["N", "CA", "C", "O", "CB", "CG1", "CG2"],
}
RESIDUES = sorted(HEAVY_TEMPLATES)
MEAN_HEAVY = sum(len(v) for v in HEAVY_TEMPLATES.values()) / len(HEAVY_TEMPLATES)  # 8.35


def globule_radius(n_res: int) -> float:
    """14 A for the 60-76 residue chains of the GA configuration, scaled with volume."""
    return max(10.0, 14.0 * (n_res / 68.0) ** (1.0 / 3.0))


def _unit(rng: np.random.Generator, n: int | None = None) -> np.ndarray:
    v = rng.normal(size=(3,) if n is None else (n, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@dataclass
class Chain:
    chain: str
    resnames: list[str]
    names: list[str] = field(default_factory=list)
    resseq: list[int] = field(default_factory=list)
    xyz: np.ndarray = None  # (n, 3)
    ca: np.ndarray = None   # (n_res, 3)


def make_chain(rng: np.random.Generator, chain: str, n_res: int | None = None, n_atoms: int | None = None,
               radius: float | None = None) -> Chain:
    if n_res is None:  # draw residues until the templates hold n_atoms heavy atoms
        resnames, total = [], 0
        while total < n_atoms:
            resnames.append(RESIDUES[int(rng.integers(0, len(RESIDUES)))])
            total += len(HEAVY_TEMPLATES[resnames[-1]])
        n_res = len(resnames)
    else:
        resnames = [RESIDUES[i] for i in rng.integers(0, len(RESIDUES), size=n_res)]
    R = globule_radius(n_res) if radius is None else radius
    ca = np.zeros((n_res, 3))
    ca[0] = _unit(rng) * R * 0.5 * rng.random()
    for i in range(1, n_res):
        for _ in range(64):
            p = ca[i - 1] + 3.8 * _unit(rng)
            if np.linalg.norm(p) <= R:
                break
        else:
            p = ca[i - 1] - 3.8 * ca[i - 1] / max(np.linalg.norm(ca[i - 1]), 1e-9)
        ca[i] = p
    names, resseq, xyz = [], [], []
    for i, rn in enumerate(resnames):
        out = ca[i] / max(np.linalg.norm(ca[i]), 1e-9)
        prev = ca[i]
        pos = {"CA": ca[i]}
        for nm in HEAVY_TEMPLATES[rn]:
            if nm == "CA":
                p = ca[i]
            elif nm == "N":
                p = ca[i] + 1.46 * _unit(rng)
            elif nm == "C":
                p = ca[i] + 1.52 * _unit(rng)
            elif nm == "O":
                p = pos["C"] + 1.23 * _unit(rng)
            else:  # side chain grows outward from CA
                d = _unit(rng) + 0.8 * out
                p = prev + 1.52 * d / np.linalg.norm(d)
                prev = p
            pos[nm] = p
            names.append(nm)
            resseq.append(i + 1)
            xyz.append(p)
    c = Chain(chain, resnames, names, resseq, np.array(xyz), ca)
    if n_atoms is not None:
        assert len(c.names) >= n_atoms, (len(c.names), n_atoms)
        c.names, c.resseq, c.xyz = c.names[:n_atoms], c.resseq[:n_atoms], c.xyz[:n_atoms]
    return c


def chain_pdb(c: Chain, first_serial: int = 1, shift=(0.0, 0.0, 0.0)) -> str:
    lines = []
    for k, (nm, rs, p) in enumerate(zip(c.names, c.resseq, c.xyz)):
        res = c.resnames[rs - 1]
        nf = f" {nm:<3s}" if len(nm) < 4 else nm
        x, y, z = (p[0] + shift[0], p[1] + shift[1], p[2] + shift[2])
        lines.append(f"ATOM  {first_serial + k:5d} {nf:4s} {res:3s} {c.chain:1s}{rs:4d}    "
                     f"{x:8.3f}{y:8.3f}{z:8.3f}{1.0:6.2f}{0.0:6.2f}          {nm[0]:>2s}")
    return "\n".join(lines) + "\nEND\n"


def surface_residues(c: Chain, frac: float = 0.3) -> list[int]:
    """Residue numbers whose C-alpha lies in the outer `frac` by distance from the centroid."""
    present = sorted({rs for rs, nm in zip(c.resseq, c.names) if nm == "CA"})
    ca = np.array([c.ca[r - 1] for r in present])
    d = np.linalg.norm(ca - ca.mean(axis=0), axis=1)
    k = max(1, int(round(frac * len(present))))
    order = np.argsort(-d, kind="stable")[:k]
    return sorted(present[i] for i in order)


@dataclass
class Complex:
    seed: int
    receptor_pdb: str
    ligand_pdb: str
    restraints: list[tuple[int, int]]
    n_rec: int
    n_lig: int


def make_complex(seed: int, rec_res: int | None = None, lig_res: int | None = None, rec_atoms: int | None = None,
                 lig_atoms: int | None = None, n_restraints: int = 6) -> Complex:
    rng = np.random.Generator(np.random.PCG64(seed))
    rec = make_chain(rng, "A", n_res=rec_res, n_atoms=rec_atoms)
    lig = make_chain(rng, "B", n_res=lig_res, n_atoms=lig_atoms)
    rs, ls = surface_residues(rec), surface_residues(lig)
    pairs = [(r, l) for r in rs for l in ls]
    pick = rng.choice(len(pairs), size=min(n_restraints, len(pairs)), replace=False)
    restraints = [pairs[i] for i in sorted(pick)]
    off = float(globule_radius(len(rec.resnames)) + globule_radius(len(lig.resnames)) + 10.0)
    return Complex(seed, chain_pdb(rec, 1), chain_pdb(lig, len(rec.names) + 1, shift=(off, 0.0, 0.0)),
                   restraints, len(rec.names), len(lig.names))


def random_genes(seed: int, n: int, radius: float = 20.0) -> np.ndarray:
    """n x 6 genes: uniform rotations (Shoemake quaternions) expressed as the
    reference's Euler angles (R = Rz(a) Ry(b) Rx(g), chromosome.rs/structure.rs),
    and translations uniform in a ball of `radius` (the ligand COM starts on the
    receptor COM after position_ligand)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R00 = 1 - 2 * (y * y + z * z)
    R10 = 2 * (x * y + z * w)
    R20 = 2 * (x * z - y * w)
    R21 = 2 * (y * z + x * w)
    R22 = 1 - 2 * (x * x + y * y)
    beta = -np.arcsin(np.clip(R20, -1.0, 1.0))
    alpha = np.arctan2(R10, R00)
    gamma = np.arctan2(R21, R22)
    ang = np.stack([alpha, beta, gamma], axis=1) % (2 * np.pi)
    d = _unit(rng, n) * (radius * rng.random(n) ** (1.0 / 3.0))[:, None]
    return np.ascontiguousarray(np.concatenate([ang, d], axis=1))


def ga_genes(seed: int, n: int) -> np.ndarray:
    """Genes spread as the GA's initial population (angles U[0, 2pi], shifts U[-50, 50])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = np.empty((n, 6))
    g[:, :3] = rng.random((n, 3)) * 2 * np.pi
    g[:, 3:] = rng.random((n, 3)) * 100.0 - 50.0
    return g


# The benchmark workloads, restated as inputs (sizes, seeds).
CONFIGS = {
    "ga_dock": dict(seed=0, rec_res=60, lig_res=76, n_restraints=6),
    "batched_scoring": dict(seed=1, rec_atoms=2000, lig_atoms=1500, n_restraints=6, n_poses=4096),
    "fcc_10k": dict(seed=2, rec_res=300, lig_res=200, n_restraints=6, n_poses=10000),
}


def uneven_batch(n_complexes: int = 64, seed0: int = 100):
    """Complexes of the uneven persistent-launch workload: receptor atoms
    log-uniform in 500-8000, ligand atoms log-uniform in 300-3000, 4-12
    restraints, per-complex seeds seed0 + i."""
    out = []
    for i in range(n_complexes):
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        nr = int(round(np.exp(rng.uniform(np.log(500), np.log(8000)))))
        nl = int(round(np.exp(rng.uniform(np.log(300), np.log(3000)))))
        nrest = int(rng.integers(4, 13))
        out.append(make_complex(seed0 + i, rec_atoms=nr, lig_atoms=nl, n_restraints=nrest))
    return out
