"""ctypes binding of libgdock_hip.so (the C ABI in include/gdock_hip.h).

The library is built in-tree by ``python -m gdock_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgdock_hip.so"

# gdock_atom (structure.rs:28-49 fields), 88 bytes, 8-byte aligned
ATOM_DTYPE = np.dtype(
    [
        ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("charge", "<f8"), ("vdw_radius", "<f8"),
        ("epsilon", "<f8"), ("rmin2", "<f8"), ("occupancy", "<f4"), ("tempfactor", "<f4"),
        ("serial", "<i4"), ("resseq", "<i2"),
        ("chainid", "S1"), ("altloc", "S1"), ("icode", "S1"), ("name", "S5"), ("resname", "S5"),
        ("element", "S5"),
    ],
    align=True,
)
assert ATOM_DTYPE.itemsize == 88, ATOM_DTYPE.itemsize

OK, INVALID_ARG, OOM, DEVICE, BUSY, NOT_FOUND = range(6)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "DEVICE", 4: "BUSY", 5: "NOT_FOUND"}


class Weights(C.Structure):
    _fields_ = [("vdw", C.c_double), ("elec", C.c_double), ("desolv", C.c_double), ("air", C.c_double)]


class ChromosomeC(C.Structure):
    _fields_ = [("genes", C.c_double * 6), ("fitness", C.c_double), ("vdw", C.c_double), ("elec", C.c_double),
                ("desolv", C.c_double), ("air", C.c_double), ("restraint_penalty", C.c_double),
                ("rmsd", C.c_double), ("fnat", C.c_double)]


class ClusteringConfigC(C.Structure):
    _fields_ = [("contact_distance", C.c_double), ("fcc_cutoff", C.c_double), ("strictness", C.c_double),
                ("min_cluster_size", C.c_int64)]


class ClusterResultC(C.Structure):
    _fields_ = [("center_idx", C.c_int64), ("size", C.c_int64), ("members_offset", C.c_int64)]


class GaParams(C.Structure):
    _fields_ = [("population_size", C.c_int32), ("max_generations", C.c_int32), ("elitism", C.c_int32),
                ("tournament_size", C.c_int32), ("convergence_window", C.c_int32), ("early_stopping", C.c_int32),
                ("cluster", C.c_int32), ("keep_trace", C.c_int32), ("mutation_rate", C.c_double),
                ("crossover_rate", C.c_double), ("convergence_threshold", C.c_double),
                ("max_displacement", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("max_jobs", C.c_int32), ("max_population", C.c_int32), ("max_receptor_atoms", C.c_int32),
                ("ga_on_device", C.c_int32), ("memory_budget", C.c_int64), ("weights", Weights)]


class JobStats(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("generations", C.c_int32), ("slot", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("poses_scored", C.c_int64), ("jobs_finished", C.c_int64),
                ("jobs_queued", C.c_int64), ("peak_active", C.c_int32), ("reserved", C.c_int32),
                ("busy_s", C.c_double), ("bytes_in_use", C.c_int64), ("memory_budget", C.c_int64),
                ("kernel_s", C.c_double), ("admit_s", C.c_double), ("ga_s", C.c_double),
                ("finish_s", C.c_double)]


class GdockError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None

P = C.c_void_p
D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gdock_amd.build` "
                          "(the scoring path has no CPU fallback)")
    # One HIP runtime per process.  PyTorch ships its own libamdhip64 with the
    # same soname as /opt/rocm's: once torch is loaded, this library binds to
    # that copy, so torch streams and tensors can be handed to it.  Loaded the
    # other way round, the process would hold two runtimes and torch's would
    # find no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "gdock_version": (C.c_char_p, []),
        "gdock_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "gdock_ctx_destroy": (None, [P]),
        "gdock_last_error": (C.c_char_p, [P]),
        "gdock_read_pdb": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(P), C.POINTER(I32), I32]),
        "gdock_free": (None, [P]),
        "gdock_position_ligand": (C.c_int, [P, C.c_int32, P, C.c_int32]),
        "gdock_complex_create": (C.c_int, [P, P, C.c_int32, P, C.c_int32, I32, C.c_int32, C.POINTER(P)]),
        "gdock_complex_destroy": (None, [P, P]),
        "gdock_complex_n_restraints": (C.c_int32, [P]),
        "gdock_eval_fitness": (C.c_int, [P, P, C.POINTER(ChromosomeC), C.c_int32, C.POINTER(Weights)]),
        "gdock_score_poses": (C.c_int, [P, P, D, C.c_int32, C.POINTER(Weights), D, I32]),
        "gdock_score_coords": (C.c_int, [P, P, D, C.c_int32, C.POINTER(Weights), D, I32]),
        "gdock_prepare_poses": (C.c_int, [D, C.c_int32, D]),
        "gdock_score_poses_device": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Weights), P, P, P]),
        "gdock_score_poses_device_noalloc": (C.c_int, [P, P, P, C.c_int32, C.POINTER(Weights), P, P, P]),
        "gdock_count_pairs": (C.c_int, [P, P, D, C.c_int32, C.POINTER(C.c_int64)]),
        "gdock_contact_words": (C.c_int32, [P, C.c_double]),
        "gdock_contact_candidates": (C.c_int32, [P, C.c_double]),
        "gdock_contact_pair": (C.c_int, [P, C.c_double, C.c_int32, C.c_char_p, C.POINTER(C.c_int16), C.c_char_p,
                                         C.POINTER(C.c_int16)]),
        "gdock_contacts": (C.c_int, [P, P, D, C.c_int32, C.c_double, C.POINTER(C.c_uint64), I32]),
        "gdock_cluster_structures": (C.c_int, [P, P, D, C.c_int32, C.POINTER(ClusteringConfigC),
                                               C.POINTER(ClusterResultC), C.POINTER(C.c_int64), I32]),
        "gdock_batch_create": (C.c_int, [P, C.POINTER(P), C.c_int32, I32, C.POINTER(P)]),
        "gdock_batch_destroy": (None, [P, P]),
        "gdock_batch_n_poses": (C.c_int32, [P]),
        "gdock_batch_plan": (C.c_int, [P, I32, I32, I32, I32]),
        "gdock_batch_last_items": (C.c_int, [P, I32]),
        "gdock_batch_score": (C.c_int, [P, P, D, C.POINTER(Weights), D, I32]),
        "gdock_batch_device_buffers": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "gdock_batch_launch": (C.c_int, [P, P, C.POINTER(Weights), P]),
        "gdock_choose_lanes": (C.c_int32, [P]),
        "gdock_time_lanes": (C.c_int, [P, P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
        "gdock_ga_default_params": (C.c_int, [C.POINTER(GaParams)]),
        "gdock_ga_run": (C.c_int, [P, P, C.POINTER(Weights), C.c_uint64, C.POINTER(GaParams), C.POINTER(P)]),
"gdock_ga_run_device": (C.c_int, [P, P, C.POINTER(Weights), C.c_uint64, C.POINTER(GaParams), C.POINTER(P)]),
        "gdock_ga_free": (None, [P]),
        "gdock_ga_run_dockq": (C.c_int, [P, P, P, C.c_uint64, C.POINTER(GaParams), C.POINTER(P)]),
        "gdock_ga_info": (C.c_int, [P, I32, I32, I32, I32, I32]),
        "gdock_ga_trace": (C.c_int, [P, D, D, I32]),
        "gdock_ga_hof": (C.c_int32, [P, D, D, I32]),
        "gdock_ga_clusters": (C.c_int32, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "gdock_ga_selected": (C.c_int32, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "gdock_rng_stream": (None, [C.c_uint64, C.c_int32, C.POINTER(C.c_uint64)]),
        "gdock_rng_normals": (None, [C.c_uint64, C.c_int32, C.c_double, C.c_double, D]),
        "gdock_stream_math": (None, [D, C.c_int32, C.c_int32, D]),
        "gdock_evaluator_create": (C.c_int, [P, P, P, C.c_int32, C.POINTER(P)]),
        "gdock_evaluator_destroy": (None, [P, P]),
        "gdock_evaluator_info": (C.c_int, [P, I32, I32, I32]),
        "gdock_calc_metrics": (C.c_int, [P, P, D, C.c_int32, D]),
        "gdock_format_pdb": (C.c_int, [P, C.c_int32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "gdock_apply_genes": (C.c_int, [P, C.c_int32, D, P]),
        "gdock_write_best_outputs": (C.c_int, [P, P, P, C.POINTER(Weights), D, C.c_int32, C.c_char_p]),
        "gdock_write_best_outputs_dockq": (C.c_int, [P, P, P, D, C.c_int32, C.c_char_p]),
        "gdock_ga_final_population": (C.c_int32, [P, D]),
        "gdock_write_run_outputs": (C.c_int, [P, P, P, D, D, C.c_int32, C.POINTER(C.c_int64),
                                              C.POINTER(C.c_int64), C.c_int32, C.c_char_p]),
        "gdock_server_default_config": (C.c_int, [C.POINTER(ServerConfig)]),
        "gdock_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfig), C.POINTER(P)]),
        "gdock_server_destroy": (None, [P]),
        "gdock_server_submit": (C.c_int, [P, P, C.c_int32, P, C.c_int32, I32, C.c_int32, C.c_uint64,
                                          C.POINTER(GaParams), C.POINTER(C.c_int64)]),
        "gdock_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(P), C.POINTER(JobStats)]),
        "gdock_server_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def vptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def check(status: int, ctx=None, what: str = "") -> None:
    if status != OK:
        msg = lib().gdock_last_error(ctx).decode() if ctx else what
        raise GdockError(status, f"{what}: {msg}" if what else msg)
