"""ctypes access to the FP64 oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  Records use the same 80-byte layout as
gdock_atom so that the same numpy arrays feed both sides.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libgdock_oracle.so"

ATOM_DTYPE = np.dtype(
    [("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("charge", "<f8"), ("vdw_radius", "<f8"), ("epsilon", "<f8"),
     ("rmin2", "<f8"), ("occupancy", "<f4"), ("tempfactor", "<f4"), ("serial", "<i4"), ("resseq", "<i2"), ("chainid", "S1"), ("altloc", "S1"), ("icode", "S1"),
     ("name", "S5"), ("resname", "S5"), ("element", "S5")], align=True)

_lib = None
P, D, I32, I64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)


def build() -> Path:
    subprocess.run(["make", "-s", "-j8", "-C", str(HERE)], check=True)
    return LIB


MUTANTS = HERE / "mutants"


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = load(LIB)
    return _lib


_loaded: dict = {}


def load(path) -> C.CDLL:
    """The oracle library at `path` (the restatement or one of its mutant
    builds, oracle/mutants/libgdock_oracle_m<k>.so) with its signatures set."""
    path = str(path)
    if path in _loaded:
        return _loaded[path]
    L = C.CDLL(path)
    sig = {
        "or_pdb_count_models": (C.c_int, [C.c_char_p]),
        "or_pdb_model": (C.c_int, [C.c_char_p, C.c_int, P, C.c_int]),
        "or_position_ligand": (None, [P, C.c_int, P, C.c_int, P]),
        "or_apply_genes": (None, [P, C.c_int, D, P]),
        "or_center_of_mass": (None, [P, C.c_int, D]),
        "or_vdw_energy": (C.c_double, [P, C.c_int, P, C.c_int]),
        "or_elec_energy": (C.c_double, [P, C.c_int, P, C.c_int]),
        "or_desolv_energy": (C.c_double, [P, C.c_int, P, C.c_int]),
        "or_air_energy": (C.c_double, [P, C.c_int, P, C.c_int, I32, C.c_int]),
        "or_satisfaction_ratio": (C.c_double, [P, C.c_int, P, C.c_int, I32, C.c_int]),
        "or_restraints_resolved": (C.c_int, [P, C.c_int, P, C.c_int, I32, C.c_int]),
        "or_score_poses": (None, [P, C.c_int, P, C.c_int, I32, C.c_int, D, C.c_int, D, C.c_int, C.c_int, D, I32]),
        "or_burial_counts": (None, [P, C.c_int, P, C.c_int, D, I32, I32]),
        "or_contacts": (C.c_int, [P, C.c_int, C.c_double, I32, C.c_int]),
        "or_calculate_fcc": (None, [I32, C.c_int, I32, C.c_int, D]),
        "or_cluster_elements": (C.c_int, [C.POINTER(C.c_uint8), C.c_int, C.c_int, I64, I64, I64]),
        "or_cluster_poses": (C.c_int, [P, C.c_int, P, C.c_int, D, C.c_int, D, C.c_int, I64, I64, I64, I32]),
        "or_ga_run": (P, [P, C.c_int, P, C.c_int, I32, C.c_int, D, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int]),
        "or_ga_free": (None, [P]),
        "or_ga_run_dockq": (P, [P, C.c_int, P, C.c_int, P, C.c_int, C.c_uint64, C.c_int, C.c_int, C.c_int]),
        "or_ga_generations": (C.c_int, [P]),
        "or_ga_converged_at": (C.c_int, [P]),
        "or_ga_trace": (None, [P, D, D, I32]),
        "or_ga_hof": (C.c_int, [P, D, D, I32]),
        "or_ga_clusters": (C.c_int, [P, I64, I64]),
        "or_ga_selected": (C.c_int, [P, I64, I64]),
        "or_ga_final_population": (C.c_int, [P, D]),
        "or_hof_new": (P, [C.c_int]),
        "or_hof_try_add": (C.c_int, [P, D, C.c_double, C.c_int]),
        "or_hof_entries": (C.c_int, [P, D]),
        "or_hof_angular_difference": (C.c_double, [C.c_double, C.c_double]),
        "or_hof_free": (None, [P]),
        "or_rng_stream": (None, [C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]),
        "or_rng_normals": (None, [C.c_uint64, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)]),
        "or_stream_math": (None, [C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "or_evolve": (None, [D, D, C.c_int, C.c_uint64, D, D]),
        "or_evaluator_new": (P, [P, C.c_int, P, C.c_int]),
        "or_evaluator_free": (None, [P]),
        "or_evaluator_info": (None, [P, I32, I32, I32, C.POINTER(C.c_int16), C.POINTER(C.c_int16)]),
        "or_calc_metrics": (None, [P, P, C.c_int, C.c_int, D]),
        "or_eval_contacts": (C.c_int, [P, C.c_int, P, C.c_int]),
        "or_evolve_scripted": (C.c_int, [D, D, C.c_int, D, C.c_int, D, D, I32, D, C.c_int]),
        "or_ga_fitness_sequence": (C.c_int, [C.c_int, C.c_int, D, C.c_int, C.c_uint64, I32]),
        "or_select_models": (C.c_int, [D, C.c_int, I64, I64, C.c_int, I64, I64]),
        "or_hof_add_population": (None, [P, D, D, C.c_int, C.c_int]),
        "or_hof_genes": (C.c_int, [P, D]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype, f.argtypes = r, a
    _loaded[path] = L
    return L


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _v(a):
    return C.c_void_p(a.ctypes.data)


def _pairs(pairs):
    return np.ascontiguousarray(np.array(list(pairs), np.int32).reshape(-1, 2))


def read_pdb_text(text: str) -> list[np.ndarray]:
    raw = text.encode()
    L = lib()
    out = []
    for m in range(L.or_pdb_count_models(raw)):
        n = L.or_pdb_model(raw, m, None, 0)
        a = np.zeros(n, ATOM_DTYPE)
        L.or_pdb_model(raw, m, _v(a), n)
        out.append(a)
    return out


def position_ligand(rec: np.ndarray, lig: np.ndarray) -> np.ndarray:
    out = np.zeros_like(lig)
    lib().or_position_ligand(_v(rec), len(rec), _v(lig), len(lig), _v(out))
    return out


def apply_genes(lig: np.ndarray, genes) -> np.ndarray:
    g = np.ascontiguousarray(genes, np.float64)
    out = np.zeros_like(lig)
    lib().or_apply_genes(_v(lig), len(lig), _p(g), _v(out))
    return out


def center_of_mass(m: np.ndarray) -> tuple:
    o = np.zeros(3)
    lib().or_center_of_mass(_v(m), len(m), _p(o))
    return tuple(o)


def vdw_energy(r, l):
    return lib().or_vdw_energy(_v(r), len(r), _v(l), len(l))


def elec_energy(r, l):
    return lib().or_elec_energy(_v(r), len(r), _v(l), len(l))


def desolv_energy(r, l):
    return lib().or_desolv_energy(_v(r), len(r), _v(l), len(l))


def air_energy(pairs, r, l):
    p = _pairs(pairs)
    return lib().or_air_energy(_v(r), len(r), _v(l), len(l), _p(p, C.c_int32), len(p))


def satisfaction_ratio(pairs, r, l):
    p = _pairs(pairs)
    return lib().or_satisfaction_ratio(_v(r), len(r), _v(l), len(l), _p(p, C.c_int32), len(p))


def restraints_resolved(pairs, r, l):
    p = _pairs(pairs)
    return lib().or_restraints_resolved(_v(r), len(r), _v(l), len(l), _p(p, C.c_int32), len(p))


def score_poses(rec, lig, pairs, genes, weights=(0.4, 0.05, 3.4, 1.0), transform=True, nthreads=8):
    """Population::eval_fitness in FP64: (out (n, 6), n_satisfied (n,))."""
    p = _pairs(pairs)
    g = np.ascontiguousarray(genes, np.float64).reshape(-1, 6)
    w = np.ascontiguousarray(weights, np.float64)
    out = np.zeros((len(g), 6))
    ns = np.zeros(len(g), np.int32)
    lib().or_score_poses(_v(rec), len(rec), _v(lig), len(lig), _p(p, C.c_int32), len(p), _p(g), len(g), _p(w),
                         1 if transform else 0, nthreads, _p(out), _p(ns, C.c_int32))
    return out, ns


def burial_counts(rec, lig, genes=None):
    rc = np.zeros(len(rec), np.int32)
    lc = np.zeros(len(lig), np.int32)
    g = None if genes is None else np.ascontiguousarray(genes, np.float64)
    lib().or_burial_counts(_v(rec), len(rec), _v(lig), len(lig), None if g is None else _p(g), _p(rc, C.c_int32),
                           _p(lc, C.c_int32))
    return rc, lc


def contacts(mol: np.ndarray, cutoff=5.0) -> set[str]:
    n = lib().or_contacts(_v(mol), len(mol), cutoff, None, 0)
    buf = np.zeros((max(n, 1), 4), np.int32)
    lib().or_contacts(_v(mol), len(mol), cutoff, _p(buf, C.c_int32), n)
    return {f"{chr(a)} {b} {chr(c)} {d}" for a, b, c, d in buf[:n].tolist()}


def calculate_fcc(x: set[str], y: set[str]):
    def enc(s):
        rows = []
        for t in s:
            ca, ra, cb, rb = t.split()
            rows.append((ord(ca), int(ra), ord(cb), int(rb)))
        return np.ascontiguousarray(np.array(rows, np.int32).reshape(-1, 4))
    a, b = enc(x), enc(y)
    o = np.zeros(2)
    lib().or_calculate_fcc(_p(a, C.c_int32), len(a), _p(b, C.c_int32), len(b), _p(o))
    return o[0], o[1]


def cluster_elements(nb: np.ndarray, min_size: int):
    n = len(nb)
    m = np.ascontiguousarray(nb, np.uint8)
    cen = np.zeros(max(n, 1), np.int64)
    off = np.zeros(n + 1, np.int64)
    mem = np.zeros(max(n, 1), np.int64)
    k = lib().or_cluster_elements(_p(m, C.c_uint8), n, min_size, _p(cen, C.c_int64), _p(off, C.c_int64),
                                  _p(mem, C.c_int64))
    return [(int(cen[i]), mem[off[i]:off[i + 1]].tolist()) for i in range(k)]


def cluster_poses(rec, lig, genes, config=(5.0, 0.60, 0.75, 4), nthreads=8):
    """cluster_structures over receptor + (lig moved by genes i): ([(center, members)], set sizes)."""
    g = np.ascontiguousarray(genes, np.float64).reshape(-1, 6)
    n = len(g)
    cfg = np.ascontiguousarray(config, np.float64)
    cen = np.zeros(max(n, 1), np.int64)
    off = np.zeros(n + 1, np.int64)
    mem = np.zeros(max(n, 1), np.int64)
    sizes = np.zeros(max(n, 1), np.int32)
    k = lib().or_cluster_poses(_v(rec), len(rec), _v(lig), len(lig), _p(g), n, _p(cfg), nthreads,
                               _p(cen, C.c_int64), _p(off, C.c_int64), _p(mem, C.c_int64), _p(sizes, C.c_int32))
    return [(int(cen[i]), mem[off[i]:off[i + 1]].tolist()) for i in range(k)], sizes[:n]


class GaRun:
    """Full GA run (run.rs:361-733 restated) with its per-generation trace.
    With `reference`, the debug mode: fitness = -DockQ (chromosome.rs:97-108)."""

    def __init__(self, rec, lig_positioned, pairs, weights=(0.4, 0.05, 3.4, 1.0), seed=42, pop_size=150,
                 max_gen=250, cluster=True, nthreads=8, reference=None):
        p = _pairs(pairs)
        w = np.ascontiguousarray(weights, np.float64)
        L = lib()
        if reference is None:
            self.h = L.or_ga_run(_v(rec), len(rec), _v(lig_positioned), len(lig_positioned), _p(p, C.c_int32),
                                 len(p), _p(w), seed, pop_size, max_gen, 1 if cluster else 0, nthreads)
        else:
            self.h = L.or_ga_run_dockq(_v(rec), len(rec), _v(lig_positioned), len(lig_positioned), _v(reference),
                                       len(reference), seed, pop_size, max_gen, 1 if cluster else 0)
        self.generations = L.or_ga_generations(self.h)
        self.converged_at = L.or_ga_converged_at(self.h)
        g = self.generations
        self.genes = np.zeros((g, pop_size, 6))
        self.fitness = np.zeros((g, pop_size))
        self.best_idx = np.zeros(g, np.int32)
        L.or_ga_trace(self.h, _p(self.genes), _p(self.fitness), _p(self.best_idx, C.c_int32))
        nh = L.or_ga_hof(self.h, None, None, None)
        self.hof_genes = np.zeros((nh, 6))
        self.hof_fitness = np.zeros(nh)
        self.hof_generation = np.zeros(nh, np.int32)
        L.or_ga_hof(self.h, _p(self.hof_genes), _p(self.hof_fitness), _p(self.hof_generation, C.c_int32))
        nc = L.or_ga_clusters(self.h, None, None)
        self.cluster_center = np.zeros(nc, np.int64)
        self.cluster_size = np.zeros(nc, np.int64)
        L.or_ga_clusters(self.h, _p(self.cluster_center, C.c_int64), _p(self.cluster_size, C.c_int64))
        si = np.zeros(5, np.int64)
        ss = np.zeros(5, np.int64)
        ns = L.or_ga_selected(self.h, _p(si, C.c_int64), _p(ss, C.c_int64))
        self.selected = list(zip(si[:ns].tolist(), ss[:ns].tolist()))
        self.final_genes = np.zeros((pop_size, 6))
        L.or_ga_final_population(self.h, _p(self.final_genes))
        L.or_ga_free(self.h)
        self.h = None


def rng_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().or_rng_stream(seed, n, _p(out, C.c_uint64))
    return out


def rng_normals(seed: int, n: int, mean=0.0, sd=1.0) -> np.ndarray:
    out = np.zeros(n)
    lib().or_rng_normals(seed, n, mean, sd, _p(out))
    return out


def stream_math(x, which: int) -> np.ndarray:
    """The project stream's ln (which 0) or cos(2 pi u) (which 1)."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().or_stream_math(_p(x), len(x), which, _p(out))
    return out


def evolve(genes, fitness, seed):
    g = np.ascontiguousarray(genes, np.float64).reshape(-1, 6)
    f = np.ascontiguousarray(fitness, np.float64)
    og = np.zeros_like(g)
    of = np.zeros_like(f)
    lib().or_evolve(_p(g), _p(f), len(f), seed, _p(og), _p(of))
    return og, of


class Hof:
    def __init__(self, max_size=500):
        self.h = lib().or_hof_new(max_size)

    def try_add(self, genes, fitness, gen):
        g = np.ascontiguousarray(genes, np.float64)
        return bool(lib().or_hof_try_add(self.h, _p(g), fitness, gen))

    def fitness(self):
        n = lib().or_hof_entries(self.h, None)
        o = np.zeros(n)
        lib().or_hof_entries(self.h, _p(o))
        return o

    def genes(self):
        n = lib().or_hof_entries(self.h, None)
        o = np.zeros((max(n, 1), 6))
        lib().or_hof_genes(self.h, _p(o))
        return o[:n]

    def add_population(self, genes, fitness, gen):
        """HallOfFame::add_from_population (run.rs:194-216)."""
        g = np.ascontiguousarray(genes, np.float64).reshape(-1, 6)
        f = np.ascontiguousarray(fitness, np.float64)
        lib().or_hof_add_population(self.h, _p(g), _p(f), len(f), gen)

    def __del__(self):
        try:
            lib().or_hof_free(self.h)
        except Exception:
            pass


def angular_difference(a, b):
    return lib().or_hof_angular_difference(a, b)


class Evaluator:
    """evaluator.rs Evaluator restated: DockQ of models against a reference ligand."""

    def __init__(self, receptor, reference):
        self.h = lib().or_evaluator_new(_v(receptor), len(receptor), _v(reference), len(reference))
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        lib().or_evaluator_info(self.h, C.byref(a), C.byref(b), C.byref(c), None, None)
        ir, il = np.zeros(max(a.value, 1), np.int16), np.zeros(max(b.value, 1), np.int16)
        lib().or_evaluator_info(self.h, C.byref(a), C.byref(b), C.byref(c), _p(ir, C.c_int16), _p(il, C.c_int16))
        self.receptor_interface = set(ir[:a.value].tolist())
        self.reference_interface = set(il[:b.value].tolist())
        self.n_native = c.value

    def calc_metrics(self, models: list) -> np.ndarray:
        """models: ligand atom arrays of one topology -> (n, 4) rmsd, irmsd, fnat, dockq."""
        if not models:
            return np.zeros((0, 4))
        m = np.ascontiguousarray(np.concatenate(models))
        out = np.zeros((len(models), 4))
        lib().or_calc_metrics(self.h, _v(m), len(models[0]), len(models), _p(out))
        return out

    def __del__(self):
        try:
            lib().or_evaluator_free(self.h)
        except Exception:
            pass


def eval_contacts(a, b) -> int:
    """evaluator.rs calculate_contacts: number of atom pairs within 5 A."""
    return lib().or_eval_contacts(_v(a), len(a), _v(b), len(b))


# ---- GA decisions on scripted streams (the oracle's Rng test hook)
DRAW_KINDS = {0: "unit", 1: "unit_incl", 2: "below", 3: "normal"}


def evolve_scripted(genes, fitness, script, L=None):
    """One Population::evolve step (population.rs:74-121) whose random draws
    replay `script` in order (unit / unit_incl: the value; below(n): the
    index; normal: the standard deviate).  Returns (genes, fitness, draws,
    ok), draws being the (kind, n) of every draw taken."""
    L = L or lib()
    genes = np.ascontiguousarray(genes, np.float64).reshape(-1, 6)
    fitness = np.ascontiguousarray(fitness, np.float64)
    script = np.ascontiguousarray(script, np.float64)
    n = len(fitness)
    og, of = np.zeros((n, 6)), np.zeros(n)
    cap = 4096
    kinds, args = np.zeros(cap, np.int32), np.zeros(cap)
    m = L.or_evolve_scripted(_p(genes), _p(fitness), n, _p(script), len(script), _p(og), _p(of), _p(kinds, C.c_int32),
                             _p(args), cap)
    taken = m if m >= 0 else cap
    draws = [(DRAW_KINDS[int(kinds[i])], int(args[i])) for i in range(min(taken, cap)) if kinds[i] in DRAW_KINDS]
    return og, of, draws, m >= 0 and m == len(script)


def ga_fitness_sequence(fseq, pop_size=6, max_gen=40, seed=1, L=None):
    """The generation loop with every individual of generation g scored
    fseq[min(g, len - 1)]: (generations evaluated, early-stop generation or -1)."""
    L = L or lib()
    f = np.ascontiguousarray(fseq, np.float64)
    conv = C.c_int32()
    g = L.or_ga_fitness_sequence(pop_size, max_gen, _p(f), len(f), seed, C.byref(conv))
    return g, conv.value


def select_models(hof_fitness, clusters, L=None):
    """Model selection (run.rs:694-733) from Hall of Fame fitness values and
    clusters [(centre, size), ...] in cluster_structures' order."""
    L = L or lib()
    hf = np.ascontiguousarray(hof_fitness, np.float64)
    cen = np.ascontiguousarray([c for c, _ in clusters] or [0], np.int64)
    siz = np.ascontiguousarray([s for _, s in clusters] or [0], np.int64)
    oi, os_ = np.zeros(16, np.int64), np.zeros(16, np.int64)
    k = L.or_select_models(_p(hf), len(hf), _p(cen, C.c_int64), _p(siz, C.c_int64), len(clusters),
                           _p(oi, C.c_int64), _p(os_, C.c_int64))
    return [(int(oi[i]), int(os_[i])) for i in range(k)]
