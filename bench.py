"""Benchmark: population scoring of the 2000 x 1500-atom complex, 4096 poses per launch.

One step = one launch of the gfx950 scoring kernel over 4096 poses whose
records already sit in HBM (gdock_score_poses_device).  Prints one JSON line:
pose-evals/s for the whole job, the roofline entry of the scoring kernel (26
algorithmic FP32 FLOP per receptor-ligand pair over the kernel's event-timed
average duration, against the MI355X FP32 peak), the pairs the kernel steps
through after culling beside the brute-force pairs, and the CPU baseline (the
FP64 oracle restatement of the reference, timed on this host at its usable
thread count and at two fewer, on a bounded sample of the same poses).  The
same line carries the other two parts of the metric, measured after the
timed region: `uneven_batch` (64 uneven complexes x 150 poses: batched
persistent launch against the plain path, one complex after another) and
`served` (GA docking jobs arriving as a Poisson process at 20 jobs/s:
jobs/min and latency percentiles, served and plain).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pose-evals/s and served docking jobs/min with p50/p99 latency vs plain GPU"
FLOP_PER_PAIR = 26          # FP32: 3 sub, 5 for d^2, 15 soft-core LJ, 3 Coulomb (division = 1)
PEAK_FP32 = 157.3e12        # MI355X published FP32 peak
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched_scoring", choices=["batched_scoring", "uneven_batch", "serve", "fcc", "ga"])
    ap.add_argument("--jobs", type=int, default=160, help="serve: number of docking jobs")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrival rate, jobs/s (0: all at once)")
    ap.add_argument("--ga-on-device", action="store_true", help="serve: advance the GAs with the device GA")
    ap.add_argument("--poses", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=5.0,
                    help="target wall time of each CPU baseline sample (two thread counts)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="batched_scoring: skip the uneven-batch and served-stream measurements")
    ap.add_argument("--extra-jobs", type=int, default=32, help="served-stream jobs measured with the default line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference kernel for PMC byte counters)")
    return ap.parse_args()


TRAFFIC_FILE = ROOT / "profiles" / "score_poses_kernel_2000x1500x4096_traffic.txt"
VALU_FILE = ROOT / "profiles" / "score_poses_kernel_2000x1500x4096_valu.txt"


def measured_traffic():
    """HBM bytes per launch of the scoring kernel, from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known
    size); None when no such summary is present."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def measured_valu_issue():
    """Fraction of the chip's VALU issue slots the scoring kernel fills, from
    the committed rocprofv3 counter summary (SQ_INSTS_VALU x 4 cycles over
    1024 SIMDs x GRBM_GUI_ACTIVE cycles); None when no summary is present."""
    try:
        for line in VALU_FILE.read_text().splitlines():
            if line.startswith("VALU issue ="):
                return float(line.rsplit("=", 1)[1].strip().rstrip("%")) / 100.0
    except (OSError, ValueError):
        pass
    return None


def copy_calibration():
    import torch
    n = 1 << 30
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n}), flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (jobs would shard "
              "across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven_batch":
        bench_uneven(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "fcc":
        bench_fcc(args)
        return
    if args.workload == "ga":
        bench_ga(args)
        return
    import torch
    from gdock_amd import _lib, structure, synth
    from gdock_amd.device import Complex, default_context

    cfg = synth.CONFIGS["batched_scoring"]
    c = synth.make_complex(cfg["seed"], rec_atoms=cfg["rec_atoms"], lig_atoms=cfg["lig_atoms"],
                           n_restraints=cfg["n_restraints"])
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    n = args.poses
    genes = synth.random_genes(cfg["seed"], n)
    poses = np.zeros((n, 16))
    _lib.check(_lib.lib().gdock_prepare_poses(_lib.ptr(genes), n, _lib.ptr(poses)), what="prepare_poses")

    ctx = default_context(0)
    cx = Complex(rec, lig, c.restraints, ctx)
    dev = torch.device("cuda", 0)
    d_poses = torch.from_numpy(poses).to(dev)
    d_out = torch.zeros((n, 6), dtype=torch.float64, device=dev)
    d_ns = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    w = _lib.Weights(0.4, 0.05, 3.4, 1.0)
    L = _lib.lib()

    def launch():
        st = L.gdock_score_poses_device(ctx.handle, cx.handle, C.c_void_p(d_poses.data_ptr()), n, C.byref(w),
                                        C.c_void_p(d_out.data_ptr()), C.c_void_p(d_ns.data_ptr()),
                                        C.c_void_p(stream.cuda_stream))
        if st != 0:
            raise RuntimeError(L.gdock_last_error(ctx.handle).decode())

    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            launch()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ev0.record(stream)
        for _ in range(args.steps):
            launch()
        ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    out = d_out.cpu().numpy()
    assert np.isfinite(out).all()

    n_rec, n_lig = len(rec), len(lig)
    pairs = float(n_rec) * n_lig * n
    ms_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    flop = FLOP_PER_PAIR * pairs
    achieved = flop / (kern_ms * 1e-3)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pose-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "precision": "fp32 pair math; fp64 cutoff re-decisions, reductions, desolvation and AIR terms",
        "data": "synthetic (seeded complex and poses)",
        "config": {
            "workload": f"batched_scoring: receptor {n_rec} x ligand {n_lig} heavy atoms, {n} poses per launch "
                        f"(seed {cfg['seed']}), inputs resident in HBM",
            "poses_per_step": n, "receptor_atoms": n_rec, "ligand_atoms": n_lig,
            "parallelism": "single GPU",
        },
        "pair_interactions_per_s": pairs * args.steps / wall,
        "kernel_ms": kern_ms,
        "roofline": {
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": measured_traffic(),
            "hbm_gb_per_s": (measured_traffic() or 0.0) / (kern_ms * 1e-3) / 1e9 if measured_traffic() else None,
            "hbm_frac": (measured_traffic() or 0.0) / (kern_ms * 1e-3) / PEAK_HBM if measured_traffic() else None,
            "valu_issue_frac": measured_valu_issue(),
            "kernel": "score_poses_kernel",
            "flop_per_launch": flop,
        },
    }
    # culling: lane-pairs the kernel steps through (counting instantiation, untimed)
    stepped, within = cx.count_pairs(genes)
    result["pairs_brute_force_per_launch"] = pairs
    result["pairs_evaluated_per_launch"] = stepped
    result["pairs_within_15A_per_launch"] = within
    result["pairs_evaluated_frac"] = stepped / pairs
    result["pairs_evaluated_per_s"] = stepped / (kern_ms * 1e-3)
    cx.close()
    if not args.no_extras:
        u = measure_uneven(max(5, args.steps // 2), 2)
        result["uneven_batch"] = {k: u[k] for k in ("workload", "pose_evals_per_s", "plain_pose_evals_per_s",
                                                   "speedup_vs_plain", "kernel_ms", "kernel_pose_evals_per_s",
                                                   "ms_per_step", "plain_ms_per_step")}
        v = measure_serve(args.extra_jobs, 20.0)
        result["served"] = {k: v[k] for k in ("workload", "jobs_per_min", "latency_p50_s", "latency_p99_s",
                                              "generations", "peak_active_jobs", "span_s", "mean_step_ms",
                                              "plain", "speedup_jobs_per_min")}
        d = measure_serve(args.extra_jobs, 20.0, ga_on_device=True, plain=False)
        result["served"]["ga_on_device"] = {k: d[k] for k in ("jobs_per_min", "latency_p50_s", "latency_p99_s",
                                                               "mean_step_ms", "mean_step_kernel_ms",
                                                               "mean_step_ga_ms")}
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(rec, lig, c.restraints, genes, args.cpu_seconds)
        result["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


def bench_uneven(args):
    u = measure_uneven(args.steps, args.warmup)
    result = {
        "metric": METRIC,
        "value": u["pose_evals_per_s"],
        "unit": "pose-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": u["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded complexes and poses)",
        "config": {"workload": u["workload"], "parallelism": "single GPU"},
    }
    result.update({k: v for k, v in u.items() if k not in ("workload", "pose_evals_per_s", "ms_per_step")})
    print(json.dumps(result), flush=True)


def measure_uneven(steps, warmup):
    """64 complexes (receptor 500-8000, ligand 300-3000 atoms, seeds 100-163),
    150 poses each (the GA's initial population: angles U[0, 2 pi], shifts
    U[-50, 50] A).  Batched: one replay of the captured generation graph
    (persistent kernel) per step.  Plain: the drop-in path one job after
    another (gdock_score_poses per complex).  Both take genes from the host and
    return results to the host; kernel-only times come from HIP events."""
    import torch
    from gdock_amd import _lib, structure, synth
    from gdock_amd.device import Batch, Complex, default_context

    per = 150
    steps = max(1, steps)
    cs = synth.uneven_batch(64, 100)
    ctx = default_context(0)
    cxs, genes, pairs = [], [], 0.0
    for i, c in enumerate(cs):
        rec = structure.read_pdb_text(c.receptor_pdb)[0]
        lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
        cxs.append(Complex(rec, lig, c.restraints, ctx))
        genes.append(synth.ga_genes(100 + i, per))
        pairs += float(len(rec)) * len(lig) * per
    t0 = time.perf_counter()
    b = Batch(cxs, [per] * len(cxs), ctx)
    create_s = time.perf_counter() - t0
    g = np.concatenate(genes)
    for _ in range(warmup):
        b.score(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out, _ = b.score(g)
    wall_b = (time.perf_counter() - t0) / steps
    assert np.isfinite(out).all()
    # kernel-only: graph replays on one stream, inputs already in HBM
    stream = torch.cuda.Stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        b.launch(None, stream.cuda_stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    kern_b = ev0.elapsed_time(ev1) / steps * 1e-3
    # plain drop-in, one job after another
    for cx, gi in zip(cxs, genes):
        cx.score_poses(gi)
    t0 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        for cx, gi in zip(cxs, genes):
            cx.score_poses(gi)
    wall_p = (time.perf_counter() - t0) / max(1, steps // 2)
    n = per * len(cxs)
    flop = FLOP_PER_PAIR * pairs
    plan = b.plan()
    b.close()
    for cx in cxs:
        cx.close()
    return {
        "workload": f"uneven_batch: 64 complexes (seeds 100-163, receptor 500-8000 and ligand "
                    f"300-3000 atoms), {per} GA-initial poses each, host genes in / results out",
        "pose_evals_per_s": n / wall_b,
        "ms_per_step": wall_b * 1e3,
        "plain_ms_per_step": wall_p * 1e3,
        "plain_pose_evals_per_s": n / wall_p,
        "speedup_vs_plain": wall_p / wall_b,
        "kernel_ms": kern_b * 1e3,
        "kernel_pose_evals_per_s": n / kern_b,
        "pair_interactions_per_s": pairs / kern_b,
        "batch_create_s": create_s,
        "plan": {"grid": plan["grid"], "lds_bytes": plan["lds_bytes"],
                 "lanes_k_histogram": {str(k): plan["lanes_k"].count(k) for k in sorted(set(plan["lanes_k"]))}},
        "roofline": {"bound": "valu", "achieved": flop / kern_b / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": flop / kern_b / PEAK_FP32, "traffic": None, "kernel": "score_batch_kernel",
                     "flop_per_launch": flop,
                     "basis": "brute-force pairs; the culling skips most pairs of GA-spread poses, so this "
                              "algorithmic rate can exceed what the ALUs execute"},
    }


def bench_serve(args):
    v = measure_serve(args.jobs, args.rate, ga_on_device=args.ga_on_device)
    result = {
        "metric": METRIC,
        "value": v["jobs_per_min"],
        "unit": "jobs/min",
        "n_gpus": 1, "steps": args.jobs, "warmup": 1,
        "ms_per_step": v["span_s"] / args.jobs * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded complexes, Poisson arrivals)",
        "config": {"workload": v["workload"], "parallelism": "single GPU"},
    }
    result.update({k: x for k, x in v.items() if k not in ("workload", "jobs_per_min")})
    print(json.dumps(result), flush=True)


def measure_serve(n_jobs, rate, ga_on_device=False, plain=True):
    """Served docking stream: jobs (complexes drawn as in the uneven batch,
    per-job seeds 1000+i, 4-12 restraints) arrive as a Poisson process and run
    the full GA (population 150, early stop after 10 flat generations) and the
    FCC clustering of their Hall of Fame.  Server: continuous batching with
    memory admission, the GAs advanced on host threads or (ga_on_device) by
    the device GA.  Plain: the same jobs one at a time through gdock_ga_run
    (one scoring launch per generation); its latencies apply the same arrival
    times to the measured service times (first come, first served)."""
    import torch  # noqa: F401  (one HIP runtime, see gdock_amd._lib)
    from gdock_amd import ga, structure, synth
    from gdock_amd.device import Complex, default_context
    from gdock_amd.serve import Server

    cs = synth.uneven_batch(n_jobs, 1000)
    jobs = []
    for c in cs:
        rec = structure.read_pdb_text(c.receptor_pdb)[0]
        lig = structure.read_pdb_text(c.ligand_pdb)[0]
        jobs.append((rec, lig, c.restraints))
    rng = np.random.default_rng(7)
    if rate > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(jobs)))
        arrivals -= arrivals[0]
    else:  # burst: every job queued at once (the server's capacity)
        arrivals = np.zeros(len(jobs))
    ctx = default_context(0)

    srv = Server(max_jobs=16, max_receptor_atoms=8192, ga_on_device=ga_on_device)
    warm = srv.submit(*jobs[0], seed=1, max_generations=2, cluster=0)
    srv.wait(warm, 300.0)
    t0 = time.perf_counter()
    ids = []
    for (rec, lig, pairs), a in zip(jobs, arrivals):
        d = a - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        ids.append(srv.submit(rec, lig, pairs, seed=42))
    res = [srv.wait(j, 600.0) for j in ids]
    st = srv.stats()
    srv.close()
    sub = np.array([t.submit_s for _, t in res])
    fin = np.array([t.finish_s for _, t in res])
    lat = fin - sub
    span = fin.max() - sub.min()
    gens = sum(r.generations for r, _ in res)

    service, pgens = [], 0
    for rec, lig, pairs in (jobs if plain else []):
        t = time.perf_counter()
        cx = Complex(rec, structure.position_ligand(rec, lig), pairs, ctx)
        r = ga.run(cx, seed=42)
        cx.close()
        service.append(time.perf_counter() - t)
        pgens += r.generations
    free_at, plat = 0.0, []
    for a, sv in zip(arrivals, service):
        start = max(a, free_at)
        free_at = start + sv
        plat.append(free_at - a)
    plat = np.array(plat) if plain else np.zeros(1)
    pspan = free_at - arrivals[0] if plain else float("nan")
    return {
        "workload": f"serve: {len(jobs)} GA docking jobs (receptor 500-8000, ligand 300-3000 atoms, "
                    f"population 150, early stop 10, FCC of the Hall of Fame), "
                    + (f"Poisson arrivals {rate:g} jobs/s" if rate > 0 else "all submitted at once"),
        "jobs_per_min": len(jobs) / span * 60.0,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "pose_evals_per_s": st["poses_scored"] / span,
        "generations": gens, "peak_active_jobs": st["peak_active"], "steps_run": st["steps"],
        "server_busy_s": st["busy_s"], "mean_step_ms": st["busy_s"] / max(st["steps"], 1) * 1e3,
        "mean_step_kernel_ms": st["kernel_s"] / max(st["steps"], 1) * 1e3,
        "mean_step_ga_ms": st["ga_s"] / max(st["steps"], 1) * 1e3,
        "ga_on_device": bool(ga_on_device),
        "server_admit_s": st["admit_s"], "server_ga_s": st["ga_s"], "server_finish_s": st["finish_s"],
        "span_s": span,
        "plain": ({"jobs_per_min": len(jobs) / pspan * 60.0, "latency_p50_s": float(np.percentile(plat, 50)),
                   "latency_p99_s": float(np.percentile(plat, 99)), "mean_service_s": float(np.mean(service)),
                   "generations": pgens, "pose_evals_per_s": pgens * 150 / float(np.sum(service))}
                  if plain else None),
        "speedup_jobs_per_min": pspan / span,
    }


def bench_fcc(args):
    """FCC clustering of 10,000 poses of one complex (receptor 300, ligand 200
    residues, seed 2; rotations uniform, translations in a 20 A ball):
    contact bitsets at 5 A, the pairwise FCC overlap, the neighbour graph and
    the greedy clustering (gdock_cluster_structures), genes from the host."""
    import torch  # noqa: F401
    from gdock_amd import structure, synth
    from gdock_amd.device import Complex, default_context

    cfg = synth.CONFIGS["fcc_10k"]
    c = synth.make_complex(cfg["seed"], rec_res=cfg["rec_res"], lig_res=cfg["lig_res"],
                           n_restraints=cfg["n_restraints"])
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    n = cfg["n_poses"]
    genes = synth.random_genes(cfg["seed"], n)
    cx = Complex(rec, lig, c.restraints, default_context(0))
    for _ in range(args.warmup):
        clusters = cx.cluster(genes)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        clusters = cx.cluster(genes)
    wall = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    bits, sizes = cx.contacts(genes)
    t_contacts = time.perf_counter() - t0
    result = {
        "metric": METRIC,
        "value": n / wall, "unit": "poses clustered/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic (seeded complex and poses)",
        "config": {"workload": f"fcc_10k: {n} poses, receptor {len(rec)} / ligand {len(lig)} atoms "
                               f"({cfg['rec_res']} x {cfg['lig_res']} residues), 5 A contacts, FCC 0.60 / 0.75, "
                               f"min cluster 4", "parallelism": "single GPU"},
        "fcc_pairs_per_s": n * (n - 1) / wall,
        "contacts_only_s": t_contacts,
        "mean_contacts_per_pose": float(sizes.mean()),
        "bitset_words64": int(bits.shape[1]),
        "n_clusters": len(clusters),
        "largest_clusters": [cl.size for cl in clusters[:5]],
    }
    print(json.dumps(result), flush=True)


def bench_ga(args):
    """Plain GA dock of the 60 x 76-residue complex (seed 0, 6 restraints): the
    full run (population 150, early stop after 10 flat generations, seed 42,
    FCC of the Hall of Fame and model selection) through gdock_ga_run, one
    scoring launch per generation.  The CPU leg runs the oracle's GA (same
    random stream, so the same decisions) on the host's cores."""
    import torch  # noqa: F401
    from gdock_amd import ga, structure, synth
    from gdock_amd.device import Complex, default_context

    cfg = synth.CONFIGS["ga_dock"]
    c = synth.make_complex(cfg["seed"], rec_res=cfg["rec_res"], lig_res=cfg["lig_res"],
                           n_restraints=cfg["n_restraints"])
    rec = structure.read_pdb_text(c.receptor_pdb)[0]
    lig = structure.position_ligand(rec, structure.read_pdb_text(c.ligand_pdb)[0])
    cx = Complex(rec, lig, c.restraints, default_context(0))
    for _ in range(args.warmup):
        ga.run(cx, seed=42)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = ga.run(cx, seed=42)
    wall = (time.perf_counter() - t0) / args.steps
    evals = res.generations * 150
    result = {
        "metric": METRIC,
        "value": evals / wall, "unit": "pose-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded complex)",
        "config": {"workload": f"ga_dock: receptor {len(rec)} / ligand {len(lig)} atoms (60 x 76 residues, seed 0), "
                               f"population 150, early stop 10, seed 42, FCC of the Hall of Fame",
                   "parallelism": "single GPU"},
        "job_wall_s": wall, "generations": res.generations, "hof": len(res.hof_fitness),
        "clusters": len(res.cluster_center), "selected": res.selected,
    }
    # the same job with the GA generation on the device (gdock_ga_run_device)
    for _ in range(args.warmup):
        ga.run_device(cx, seed=42)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dres = ga.run_device(cx, seed=42)
    dwall = (time.perf_counter() - t0) / args.steps
    result["device_ga"] = {"job_wall_s": dwall, "pose_evals_per_s": dres.generations * 150 / dwall,
                           "generations": dres.generations, "selected": dres.selected,
                           "same_run_as_host_driver": bool(dres.generations == res.generations and
                                                           dres.selected == res.selected)}
    if not args.no_cpu:
        import oracle.oracle as orc
        cpu = host_cpus()
        threads = max(1, min(cpu["nproc"], cpu["share"] or cpu["nproc"]))
        t0 = time.perf_counter()
        ora = orc.GaRun(rec.atoms, lig.atoms, c.restraints, seed=42, nthreads=threads)
        cpu = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": ora.generations * 150 / cpu, "unit": "pose-evals/s", "cores": threads,
                                  "kind": "port", "sample": f"the whole GA run ({ora.generations} generations), "
                                  f"{cpu:.1f} s wall, FP64 oracle"}
        result["cpu_job_wall_s"] = cpu
        result["same_run_as_oracle"] = bool(ora.generations == res.generations and ora.selected == res.selected)
    print(json.dumps(result), flush=True)


def host_cpus():
    """This host's CPU model, nproc (the CPUs this process may run on), the
    logical CPU count, and the CPU share granted to it (the cgroup quota, or
    OMP_NUM_THREADS where the host sets that instead)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    share = None
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            share = int(quota) // int(period)
    except (OSError, ValueError):
        pass
    if share is None and os.environ.get("OMP_NUM_THREADS", "").isdigit():
        share = int(os.environ["OMP_NUM_THREADS"])
    return {"model": model, "nproc": nproc, "logical_cpus": os.cpu_count(), "share": share}


def cpu_baseline(rec, lig, pairs, genes, target_s):
    """FP64 oracle (reference restatement, one pose per task on a thread pool
    as rayon's par_iter_mut) on a sample of the same poses, at the host's
    usable thread count and at two fewer (the reference's default thread
    count is the CPU count minus 2, main.rs:195-198).  The usable count is
    nproc, capped at the CPU share this process is granted."""
    import oracle.oracle as orc
    cpu = host_cpus()
    full = max(1, min(cpu["nproc"], cpu["share"] or cpu["nproc"]))
    runs = []
    for threads in dict.fromkeys([full, max(1, full - 2)]):
        t0 = time.perf_counter()
        orc.score_poses(rec.atoms, lig.atoms, pairs, genes[:threads], nthreads=threads)
        dt = time.perf_counter() - t0
        m = int(min(len(genes), max(threads, threads * round(target_s / max(dt, 1e-3)))))
        t0 = time.perf_counter()
        orc.score_poses(rec.atoms, lig.atoms, pairs, genes[:m], nthreads=threads)
        dt = time.perf_counter() - t0
        runs.append({"threads": threads, "poses": m, "wall_s": dt, "value": m / dt})
    return {"value": runs[0]["value"], "unit": "pose-evals/s", "cores": full, "kind": "port",
            "sample": f"first {runs[0]['poses']} of the same poses on {full} threads, {runs[0]['wall_s']:.1f} s wall "
                      f"(FP64 oracle, -O3 generic x86-64)",
            "runs": runs, "cpu_model": cpu["model"], "nproc": cpu["nproc"], "logical_cpus": cpu["logical_cpus"],
            "cpu_share": cpu["share"]}


if __name__ == "__main__":
    main()
