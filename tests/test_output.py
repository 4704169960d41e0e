"""Output stage on the host: PDB records (structure.rs:159-176,303-310) and
Chromosome::apply_genes (chromosome.rs:79-83).

Pins:
- the reference's own data files: every ATOM record of 2oob_A/B, 2oob and
  1h590A/B, parsed and written back, equals the file's record up to column 78
  except where to_pdb_string writes differently from the file's writer: a
  three-letter atom name is centred the Rust way ("CD1 ", the odd pad on the
  right) and a blank element column gets the element inferred from the name;
- an independent restatement of the same format string in Python's format
  mini-language (correctly rounded, ties to even, as Rust's exact-precision
  float formatting) on random atoms, including negative zero, long fields and
  values on rounding ties;
- apply_genes bit for bit against the FP64 oracle.
"""
from __future__ import annotations

import gzip
import math
from pathlib import Path

import numpy as np
import pytest

import oracle.oracle as orc
from gdock_amd import structure, synth
from gdock_amd._lib import ATOM_DTYPE

DATA = Path(__file__).resolve().parent / "golden" / "data"


def _text(name: str) -> str:
    p = DATA / name
    return gzip.decompress(p.read_bytes()).decode() if name.endswith(".gz") else p.read_text()


def _expected_record(line: str) -> str:
    name = line[12:16].strip()
    centred = f"{name:^4}"  # Rust's {:^4}: the odd pad on the right
    element = line[76:78] if line[76:78].strip() else f"{name[:1]:>2}"
    return line[:12] + centred + line[16:76] + element + "\n"


@pytest.mark.parametrize("name", ["2oob_A.pdb", "2oob_B.pdb", "2oob.pdb", "1h590A_0A.pdb", "1h590B_0A_poses.pdb.gz"])
def test_written_records_reproduce_reference_data(name):
    text = _text(name)
    models = structure.read_pdb_text(text)
    lines = [ln for ln in text.splitlines() if ln.startswith("ATOM")]
    got = "".join(structure.to_pdb_string(m) for m in models).splitlines(keepends=True)
    assert len(got) == len(lines)
    three = 0
    for g, ln in zip(got, lines):
        assert g == _expected_record(ln), (g, ln)
        three += len(ln[12:16].strip()) == 3
    assert three > 0 or name.startswith("1h590B")


def _python_record(a) -> str:
    s = lambda f: a[f].decode() if isinstance(a[f], bytes) else a[f].tobytes().split(b"\0")[0].decode()  # noqa: E731
    ch = lambda f: (a[f].decode() or " ")  # noqa: E731
    return (f"ATOM  {int(a['serial']):>5} {s('name'):^4}{ch('altloc'):>1}{s('resname'):3} {ch('chainid'):>1}"
            f"{int(a['resseq']):>4}{ch('icode'):>1}   {float(a['x']):>8.3f}{float(a['y']):>8.3f}{float(a['z']):>8.3f}"
            f"{float(a['occupancy']):>6.2f}{float(a['tempfactor']):>6.2f}          {s('element'):>2}\n")


def test_records_match_the_format_string_on_random_atoms():
    rng = np.random.default_rng(3)
    n = 2000
    a = np.zeros(n, ATOM_DTYPE)
    a["serial"] = rng.integers(-9999, 999999, n)
    a["resseq"] = rng.integers(-999, 32767, n)
    scale = 10.0 ** rng.integers(-4, 5, (n, 3))
    xyz = rng.uniform(-1, 1, (n, 3)) * scale
    xyz[::7] = np.round(xyz[::7] * 16.0) / 16.0       # binary ties at 3 decimals (x.0625, x.1875, ...)
    xyz[::11, 0] = -0.0
    xyz[::13, 1] = -0.0004                             # rounds to "-0.000"
    a["x"], a["y"], a["z"] = xyz.T
    a["occupancy"] = rng.choice([0.0, 0.5, 1.0, 0.125, 0.375], n).astype(np.float32)  # ties at 2 decimals
    a["tempfactor"] = rng.uniform(-99, 999, n).astype(np.float32)
    names = [b"N", b"CA", b"CD1", b"HD21", b"O"]
    a["name"] = [names[i] for i in rng.integers(0, 5, n)]
    a["resname"] = [[b"ALA", b"GL", b"HSD", b"X"][i] for i in rng.integers(0, 4, n)]
    a["element"] = [[b"C", b"N", b"FE", b""][i] for i in rng.integers(0, 4, n)]
    a["chainid"], a["altloc"], a["icode"] = b"B", b" ", b"A"
    got = structure.to_pdb_string(structure.Molecule(a)).splitlines(keepends=True)
    assert got == [_python_record(r) for r in a]
    assert any(" -0.000" in g for g in got) and any(" 0.125" in g or "  0.12" in g for g in got)


def test_non_finite_coordinates_spelled_as_rust():
    a = structure.make_atom(float("nan"), float("inf"), -float("inf"))
    line = structure.to_pdb_string(structure.Molecule(a))
    assert line[30:54] == "     NaN     inf    -inf"


def test_apply_genes_bit_exact_against_oracle():
    lig = structure.read_pdb_text(_text("2oob_B.pdb"))[0]
    for g in np.concatenate([synth.random_genes(4, 20, 15.0), synth.ga_genes(5, 20), np.zeros((1, 6))]):
        mine = structure.apply_genes(lig, g)
        ref = orc.apply_genes(lig.atoms, g)
        for f in ("x", "y", "z"):
            assert np.array_equal(mine.atoms[f], ref[f])
        for f in ("serial", "name", "resname", "element", "occupancy", "tempfactor", "charge"):
            assert np.array_equal(mine.atoms[f], lig.atoms[f])


def test_apply_genes_identity_and_translation():
    """chromosome.rs:256-279: zero rotation moves every atom by the displacement only."""
    lig = structure.read_pdb_text(_text("1h590A_0A.pdb"))[0]
    out = structure.apply_genes(lig, [0, 0, 0, 1.5, -2.0, 0.25])
    assert np.allclose(out.xyz - lig.xyz, [1.5, -2.0, 0.25], atol=1e-12, rtol=0)
    rot = structure.apply_genes(lig, [math.pi / 3, 0.4, -1.1, 0, 0, 0])
    assert np.allclose(rot.center_of_mass(), lig.center_of_mass(), atol=1e-9, rtol=0)
