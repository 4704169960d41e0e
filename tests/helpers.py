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
