"""Shared fixtures: reference data files, synthetic complexes, tolerances."""
from __future__ import annotations

import gzip
from functools import lru_cache
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
DATA = GOLDEN / "data"

# parity bar for every energy component: 1e-5 relative, 1e-3 kcal/mol absolute floor
REL_TOL = 1e-5
ABS_TOL = 1e-3
COMPONENTS = ["vdw", "elec", "desolv", "air", "restraint_penalty", "fitness"]

# README.md:43 restraints for 2oob
RESTRAINTS_2OOB = [(933, 6), (936, 8), (940, 42), (941, 44), (946, 45), (950, 46)]


def read_text(name: str) -> str:
    p = DATA / name
    if p.suffix == ".gz":
        with gzip.open(p, "rt") as f:
            return f.read()
    return p.read_text()


def within_tol(got: np.ndarray, ref: np.ndarray, rel=REL_TOL, abs_=ABS_TOL) -> np.ndarray:
    return np.abs(got - ref) <= np.maximum(rel * np.abs(ref), abs_)


def assert_energies_close(got: np.ndarray, ref: np.ndarray, what: str = ""):
    got, ref = np.asarray(got), np.asarray(ref)
    ok = within_tol(got, ref)
    if not ok.all():
        bad = np.argwhere(~ok)[:10]
        lines = [f"pose {i} comp {COMPONENTS[j] if got.ndim == 2 else j}: got {got[tuple([i, j])] if got.ndim == 2 else got[i]!r} "
                 f"ref {ref[tuple([i, j])] if ref.ndim == 2 else ref[i]!r}" for i, j in
                 (b if len(b) == 2 else (b[0], 0) for b in bad)]
        raise AssertionError(f"{what}: {int((~ok).sum())} values out of tolerance\n" + "\n".join(lines))


@lru_cache(maxsize=None)
def synth_complex(seed: int, **kw):
    from gdock_amd import synth
    return synth.make_complex(seed, **kw)


def oracle_atoms(text: str, model: int = 0):
    import oracle.oracle as orc
    return orc.read_pdb_text(text)[model]


def test_atom(x, y, z, element="C", charge=0.0, serial=1, name="CA", resseq=1, chainid="A", resname="ALA",
              eps=-0.1, rmin2=2.0, radius=1.7):
    """The reference tests' create_test_atom (fitness.rs:308-330) as a record."""
    from gdock_amd._lib import ATOM_DTYPE
    a = np.zeros(1, ATOM_DTYPE)
    a["x"], a["y"], a["z"] = x, y, z
    a["element"], a["charge"], a["serial"], a["name"] = element.encode(), charge, serial, name.encode()
    a["resseq"], a["chainid"], a["resname"] = resseq, chainid.encode(), resname.encode()
    a["epsilon"], a["rmin2"], a["vdw_radius"] = eps, rmin2, radius
    a["altloc"], a["icode"] = b" ", b" "
    return a


def mol(*atoms):
    from gdock_amd._lib import ATOM_DTYPE
    return np.concatenate(atoms) if atoms else np.zeros(0, ATOM_DTYPE)


def relabel(atoms: np.ndarray, rule) -> np.ndarray:
    """Copy of atom records with (chainid, resseq) replaced by
    rule(chain, resseq) -> (chain, resseq)."""
    out = atoms.copy()
    for i in range(len(out)):
        c, r = rule(out["chainid"][i].decode(), int(out["resseq"][i]))
        out["chainid"][i] = c.encode()
        out["resseq"][i] = r
    return out


def _same(c, r):
    return c, r


# Chain layouts of 2oob (receptor chain A, residues 929-972; ligand chain B,
# residues 1-72) as (receptor rule, ligand rule).  clustering.rs:62-115 keys
# residues by (chain, resseq) over the whole complex, so each layout brings
# in other candidate pairs: receptor-receptor, ligand-ligand, none at all
# (one chain ID for both), or residues holding receptor and ligand atoms.
TOPOLOGIES = {
    "receptor_two_chains": (lambda c, r: ("C", r) if r >= 951 else (c, r), _same),
    "shared_chain_id": (_same, lambda c, r: ("A", r)),
    "ligand_two_chains": (_same, lambda c, r: ("D", r) if r >= 37 else (c, r)),
    "both_two_chains": (lambda c, r: ("C", r) if r >= 951 else (c, r), lambda c, r: ("D", r) if r >= 37 else (c, r)),
    # ligand residues 1-36 -> (A, 929-964), 37-72 -> (C, 951-986): keys of
    # both chains hold receptor and ligand atoms
    "merged_keys": (lambda c, r: ("C", r) if r >= 951 else (c, r),
                    lambda c, r: ("A", r + 928) if r <= 36 else ("C", r + 914)),
}


def contacts_brute_force(atoms: np.ndarray, cutoff: float) -> set[str]:
    """clustering.rs:62-115 over every heavy atom pair: residue keys (chain,
    resseq), pairs of different chains, distance <= cutoff, smaller chain
    first."""
    heavy = atoms[atoms["element"] != b"H"]
    xyz = np.stack([heavy["x"], heavy["y"], heavy["z"]], 1)
    d = np.sqrt(((xyz[:, None] - xyz[None]) ** 2).sum(-1))
    ch = heavy["chainid"]
    ii, jj = np.nonzero((d <= cutoff) & (ch[:, None] != ch[None]))
    out = set()
    for i, j in zip(ii.tolist(), jj.tolist()):
        a, b = (i, j) if ch[i] < ch[j] else (j, i)
        out.add(f"{ch[a].decode()} {heavy['resseq'][a]} {ch[b].decode()} {heavy['resseq'][b]}")
    return out
