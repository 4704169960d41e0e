"""Atoms and molecules (structure.rs:9-327) as numpy records of gdock_atom.

Parsing and parameterization run in the C++ library (gdock_read_pdb), which
follows process_atom_line / read_pdb and the toppar lookups.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import ATOM_DTYPE, I32, P, check, lib


class Molecule:
    """structure::Molecule: an ordered list of parameterized atoms."""

    __slots__ = ("atoms",)

    def __init__(self, atoms: np.ndarray | None = None):
        self.atoms = np.zeros(0, ATOM_DTYPE) if atoms is None else np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)

    def __len__(self) -> int:
        return len(self.atoms)

    def copy(self) -> "Molecule":
        return Molecule(self.atoms.copy())

    @property
    def xyz(self) -> np.ndarray:
        return np.stack([self.atoms["x"], self.atoms["y"], self.atoms["z"]], axis=1)

    def set_xyz(self, xyz: np.ndarray) -> None:
        self.atoms["x"], self.atoms["y"], self.atoms["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]

    def center_of_mass(self) -> tuple[float, float, float]:
        """structure.rs:63-69 (sequential sums)."""
        n = float(len(self.atoms))
        return tuple(float(sum(self.atoms[c].tolist())) / n for c in ("x", "y", "z"))


def make_atom(x: float, y: float, z: float, *, serial: int = 1, name: str = "CA", resname: str = "ALA",
              chainid: str = "A", resseq: int = 1, element: str = "C", charge: float = 0.0,
              vdw_radius: float = 1.7, epsilon: float = -0.1, rmin2: float = 2.0) -> np.ndarray:
    """One gdock_atom record; the defaults are the reference tests' create_test_atom."""
    a = np.zeros(1, ATOM_DTYPE)
    a["x"], a["y"], a["z"] = x, y, z
    a["serial"], a["name"], a["resname"], a["chainid"] = serial, name.encode(), resname.encode(), chainid.encode()
    a["resseq"], a["element"], a["charge"] = resseq, element.encode(), charge
    a["vdw_radius"], a["epsilon"], a["rmin2"] = vdw_radius, epsilon, rmin2
    a["altloc"], a["icode"] = b" ", b" "
    return a


def molecule(*atoms: np.ndarray) -> Molecule:
    return Molecule(np.concatenate(atoms) if atoms else None)


def read_pdb_text(text: str) -> list[Molecule]:
    """structure::read_pdb on text; returns the Model as a list of molecules."""
    raw = text.encode()
    atoms = P()
    offs = I32()
    n_models = C.c_int32()
    L = lib()
    check(L.gdock_read_pdb(raw, len(raw), C.byref(atoms), C.byref(offs), C.byref(n_models)), what="read_pdb")
    try:
        off = np.ctypeslib.as_array(offs, shape=(n_models.value + 1,)).copy()
        total = int(off[-1])
        buf = (C.c_char * (max(total, 1) * ATOM_DTYPE.itemsize)).from_address(atoms.value)
        arr = np.frombuffer(buf, dtype=ATOM_DTYPE, count=total).copy()
    finally:
        L.gdock_free(atoms)
        L.gdock_free(C.cast(offs, P))
    return [Molecule(arr[off[m]:off[m + 1]]) for m in range(n_models.value)]


def read_pdb(path: str | Path) -> list[Molecule]:
    return read_pdb_text(Path(path).read_text())


def to_pdb_string(mol: Molecule) -> str:
    """Atom::to_pdb_string for every atom, concatenated (structure.rs:159-176)."""
    n = C.c_size_t()
    L = lib()
    check(L.gdock_format_pdb(mol.atoms.ctypes.data, len(mol), None, 0, C.byref(n)), what="format_pdb")
    buf = C.create_string_buffer(n.value + 1)
    check(L.gdock_format_pdb(mol.atoms.ctypes.data, len(mol), buf, n.value + 1, C.byref(n)), what="format_pdb")
    return buf.raw[:n.value].decode()


def write_pdb(mol: Molecule, path: str | Path) -> None:
    """structure.rs:303-310."""
    Path(path).write_text(to_pdb_string(mol))


def apply_genes(ligand: Molecule, genes) -> Molecule:
    """Chromosome::apply_genes (chromosome.rs:79-83): rotate about the COM, then displace."""
    g = np.ascontiguousarray(np.asarray(genes, dtype=np.float64).reshape(6))
    out = Molecule(np.zeros(len(ligand), ATOM_DTYPE))
    check(lib().gdock_apply_genes(ligand.atoms.ctypes.data, len(ligand), g.ctypes.data_as(C.POINTER(C.c_double)),
                                  out.atoms.ctypes.data), what="apply_genes")
    return out


def combine_molecules(receptor: Molecule, ligand: Molecule) -> Molecule:
    """run.rs:35-49."""
    return Molecule(np.concatenate([receptor.atoms, ligand.atoms]))


def position_ligand(receptor: Molecule, ligand: Molecule) -> Molecule:
    """utils.rs:77-95: move the ligand's COM onto the receptor's (returns a copy)."""
    out = ligand.copy()
    check(lib().gdock_position_ligand(receptor.atoms.ctypes.data, len(receptor), out.atoms.ctypes.data, len(out)),
          what="position_ligand")
    return out
