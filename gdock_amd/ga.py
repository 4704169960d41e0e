"""GA docking job (commands/run.rs): host GA driver in the library, every
generation scored on the device (gdock_ga_run)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import constants
from ._lib import GaParams, check, lib, ptr
from .device import Complex, weights_c


def default_params(**overrides) -> GaParams:
    p = GaParams()
    check(lib().gdock_ga_default_params(C.byref(p)), what="gdock_ga_default_params")
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


@dataclass
class GaResult:
    generations: int
    converged_at: int
    genes: np.ndarray = field(repr=False)       # (generations, pop, 6) when traced
    fitness: np.ndarray = field(repr=False)     # (generations, pop)
    best_idx: np.ndarray = field(repr=False)
    hof_genes: np.ndarray = field(repr=False)
    hof_fitness: np.ndarray = field(repr=False)
    hof_generation: np.ndarray = field(repr=False)
    cluster_center: np.ndarray = field(repr=False)
    cluster_size: np.ndarray = field(repr=False)
    selected: list
    final_genes: np.ndarray = field(repr=False)  # the population the generation loop leaves behind (pop, 6)


def run(cx: Complex, seed: int = 42, weights: constants.EnergyWeights | None = None, **params) -> GaResult:
    """One job through the plain path: one scoring launch per generation."""
    p = default_params(**params)
    w = weights_c(weights)
    h = C.c_void_p()
    check(lib().gdock_ga_run(cx.ctx.handle, cx.handle, C.byref(w), seed, C.byref(p), C.byref(h)), cx.ctx.handle,
          "gdock_ga_run")
    return take_result(h, p)


def run_device(cx: Complex, seed: int = 42, weights: constants.EnergyWeights | None = None, **params) -> GaResult:
    """The same job with the GA generation on the device (gdock_ga_run_device)."""
    p = default_params(**params)
    w = weights_c(weights)
    h = C.c_void_p()
    check(lib().gdock_ga_run_device(cx.ctx.handle, cx.handle, C.byref(w), seed, C.byref(p), C.byref(h)),
          cx.ctx.handle, "gdock_ga_run_device")
    return take_result(h, p)


def run_dockq(cx: Complex, evaluator, seed: int = 42, **params) -> GaResult:
    """Debug mode (chromosome.rs:97-108): fitness = -DockQ against the evaluator's reference."""
    p = default_params(**params)
    h = C.c_void_p()
    check(lib().gdock_ga_run_dockq(cx.ctx.handle, cx.handle, evaluator.handle, seed, C.byref(p), C.byref(h)),
          cx.ctx.handle, "gdock_ga_run_dockq")
    return take_result(h, p)


def take_result(h, p: GaParams) -> GaResult:
    """Reads a gdock_ga_result and frees it."""
    L = lib()
    try:
        g, conv, nh, nc, ns = (C.c_int32() for _ in range(5))
        check(L.gdock_ga_info(h, C.byref(g), C.byref(conv), C.byref(nh), C.byref(nc), C.byref(ns)), what="ga_info")
        pop, G = p.population_size, g.value
        traced = bool(p.keep_trace)
        genes = np.zeros((G if traced else 0, pop, 6))
        fit = np.zeros((G if traced else 0, pop))
        best = np.zeros(G if traced else 0, np.int32)
        if traced:
            check(L.gdock_ga_trace(h, ptr(genes), ptr(fit), ptr(best, C.c_int32)), what="ga_trace")
        hg, hf, hgen = np.zeros((nh.value, 6)), np.zeros(nh.value), np.zeros(nh.value, np.int32)
        L.gdock_ga_hof(h, ptr(hg), ptr(hf), ptr(hgen, C.c_int32))
        cc, cs = np.zeros(nc.value, np.int64), np.zeros(nc.value, np.int64)
        L.gdock_ga_clusters(h, ptr(cc, C.c_int64), ptr(cs, C.c_int64))
        si, ss = np.zeros(max(ns.value, 1), np.int64), np.zeros(max(ns.value, 1), np.int64)
        L.gdock_ga_selected(h, ptr(si, C.c_int64), ptr(ss, C.c_int64))
        fin = np.zeros((pop, 6))
        k = L.gdock_ga_final_population(h, ptr(fin))
        fin = fin[:max(k, 0)]  # an odd population breeds one child fewer
        return GaResult(G, conv.value, genes, fit, best, hg, hf, hgen, cc, cs,
                        list(zip(si[:ns.value].tolist(), ss[:ns.value].tolist())), fin)
    finally:
        L.gdock_ga_free(h)


def write_outputs(cx: Complex, result: GaResult, out_dir, evaluator=None) -> None:
    """commands::run's output files after clustering (run.rs:738-880):
    model_<i>.pdb, ranked_<i>.pdb and metrics.tsv in out_dir."""
    hg = np.ascontiguousarray(result.hof_genes, dtype=np.float64).reshape(-1, 6)
    hf = np.ascontiguousarray(result.hof_fitness, dtype=np.float64)
    si = np.ascontiguousarray([s[0] for s in result.selected] or [0], dtype=np.int64)
    ss = np.ascontiguousarray([s[1] for s in result.selected] or [0], dtype=np.int64)
    check(lib().gdock_write_run_outputs(cx.ctx.handle, cx.handle, evaluator.handle if evaluator else None, ptr(hg),
                                        ptr(hf), len(hf), ptr(si, C.c_int64), ptr(ss, C.c_int64),
                                        len(result.selected), str(out_dir).encode()),
          cx.ctx.handle, "gdock_write_run_outputs")


def write_best_outputs(cx: Complex, result: GaResult, out_dir, evaluator=None,
                       weights: constants.EnergyWeights | None = None, debug: bool = False) -> None:
    """commands::run's `--no-clust` output (run.rs:549-637): best_by_score.pdb,
    and with an evaluator best_by_dockq.pdb and metrics.tsv, from the final
    population.  debug: the reference's debug mode, where a pose's fitness is
    -DockQ (chromosome.rs:97-108); it needs the evaluator."""
    g = np.ascontiguousarray(result.final_genes, dtype=np.float64).reshape(-1, 6)
    if debug:
        if evaluator is None:
            raise ValueError("debug outputs rank poses by DockQ: an evaluator is required")
        check(lib().gdock_write_best_outputs_dockq(cx.ctx.handle, cx.handle, evaluator.handle, ptr(g), len(g),
                                                   str(out_dir).encode()),
              cx.ctx.handle, "gdock_write_best_outputs_dockq")
        return
    w = weights_c(weights)
    check(lib().gdock_write_best_outputs(cx.ctx.handle, cx.handle, evaluator.handle if evaluator else None,
                                         C.byref(w), ptr(g), len(g), str(out_dir).encode()),
          cx.ctx.handle, "gdock_write_best_outputs")


def rng_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().gdock_rng_stream(seed, n, ptr(out, C.c_uint64))
    return out
