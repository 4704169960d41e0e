"""Autotuner of the ligand atoms per lane (K) of the scoring kernels.

Run on a GPU as ``python -m gdock_amd.tune [--out PATH]``.  For every
receptor size class (half octaves of receptor atoms) and every 128-atom
ligand interval with more than one candidate K, it builds a seeded synthetic
complex of that shape (receptor atoms 2^(class/2), ligand atoms the
interval's midpoint), times each candidate (gdock_time_lanes: 512 probe
poses near the receptor, one warm-up launch and 5 timed ones) and writes the
chunk count of the fewest-chunk candidate, or of the fastest one when it is
more than 3% faster (near-ties would flip from run to run), to
gdock_amd/csrc/lanes_table.inc.  The
library reads that table when it is built; nothing is measured at run time,
so a complex's K, and with it the summation order of its scores, depends on
its shape alone.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
TABLE = PKG / "csrc" / "lanes_table.inc"
THREADS = 256
INTERVAL = 128


def candidates(n_lig: int) -> list[int]:
    """The K of 1, 2, 3, ... chunks (at most 8 atoms per lane), at most three."""
    ks = []
    for chunks in range(1, 17):
        k = -(-n_lig // (THREADS * chunks))
        if k < 1:
            break
        if k > 8:
            continue
        if k not in ks:
            ks.append(k)
        if len(ks) >= 3:
            break
    return ks or [1]


def chunks_of(n_lig: int, k: int) -> int:
    return -(-n_lig // (THREADS * k))


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--classes", default="17-26", help="receptor size classes, e.g. 17-26")
    ap.add_argument("--max-lig", type=int, default=3072)
    ap.add_argument("--poses", type=int, default=512)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--margin", type=float, default=0.03,
                    help="a candidate with more chunks must be this much faster to be chosen")
    args = ap.parse_args(argv)
    from . import _lib, structure, synth
    from .device import Complex, default_context

    lo, hi = (int(x) for x in args.classes.split("-"))
    ctx = default_context(0)
    L = _lib.lib()
    rows = []
    t0 = time.time()
    for cls in range(lo, hi + 1):
        n_rec = int(round(2.0 ** (cls / 2.0)))
        for i in range(args.max_lig // INTERVAL):
            n_lig = INTERVAL * i + INTERVAL // 2
            ks = candidates(n_lig)
            if len(ks) < 2:
                continue
            c = synth.make_complex(1000 * cls + i, rec_atoms=n_rec, lig_atoms=n_lig, n_restraints=4)
            rec = structure.read_pdb_text(c.receptor_pdb)[0]
            lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
            cx = Complex(rec, lig, c.restraints, ctx)
            times = {}
            for k in ks:
                ms = C.c_float()
                _lib.check(L.gdock_time_lanes(ctx.handle, cx.handle, k, args.poses, args.reps, C.byref(ms)), ctx.handle,
                           "gdock_time_lanes")
                times[k] = ms.value
            cx.close()
            # the fewest chunks (largest K) unless another candidate is faster
            # by more than the margin: near-ties would flip from run to run
            fastest = min(ks, key=lambda k: (times[k], k))
            best = fastest if times[fastest] < (1.0 - args.margin) * times[ks[0]] else ks[0]
            rows.append((n_lig, cls, chunks_of(n_lig, best)))
            print(f"class {cls} (receptor {n_rec}) ligand {n_lig}: " +
                  ", ".join(f"K={k} {times[k]:.3f} ms" for k in ks) + f" -> K={best} ({time.time() - t0:.0f} s)",
                  flush=True)
    lines = [
        "// Chunk counts of the scoring kernels by complex shape, written by the",
        "// autotuner (python -m gdock_amd.tune).  Rows: {ligand atoms (midpoint of a",
        "// 128-atom interval), receptor size class (round(2 log2(receptor atoms))),",
        "// chunks}.  The first row never matches.",
        "constexpr LanesRow kLanesTable[] = {",
        "    {0, -1, 1},",
    ]
    lines += [f"    {{{n}, {c}, {ch}}}," for n, c, ch in rows]
    lines.append("};")
    Path(args.out).write_text("\n".join(lines) + "\n")
    print(f"wrote {len(rows)} rows to {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
