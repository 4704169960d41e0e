"""Benchmark: population scoring of the 2000 x 1500-atom complex, 4096 poses per launch.

One step = one launch of the gfx950 scoring kernel over 4096 poses whose
records already sit in HBM (gdock_score_poses_device).  Prints one JSON line:
pose-evals/s for the whole job, the roofline entry of the scoring kernel (26
algorithmic FP32 FLOP per receptor-ligand pair over the kernel's event-timed
average duration, against the MI355X FP32 peak) and the CPU baseline (the FP64
oracle restatement of the reference, timed on this host's cores on a bounded
sample of the same poses).
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
    ap.add_argument("--poses", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target wall time of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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
            "valu_issue_frac": measured_valu_issue(),
            "kernel": "score_poses_kernel",
            "flop_per_launch": flop,
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(rec, lig, c.restraints, genes, args.cpu_seconds)
    print(json.dumps(result), flush=True)


def bench_uneven(args):
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
    for _ in range(args.warmup):
        b.score(g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, _ = b.score(g)
    wall_b = (time.perf_counter() - t0) / args.steps
    assert np.isfinite(out).all()
    # kernel-only: graph replays on one stream, inputs already in HBM
    stream = torch.cuda.Stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        b.launch(None, stream.cuda_stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    kern_b = ev0.elapsed_time(ev1) / args.steps * 1e-3
    # plain drop-in, one job after another
    for cx, gi in zip(cxs, genes):
        cx.score_poses(gi)
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        for cx, gi in zip(cxs, genes):
            cx.score_poses(gi)
    wall_p = (time.perf_counter() - t0) / max(1, args.steps // 2)
    n = per * len(cxs)
    flop = FLOP_PER_PAIR * pairs
    plan = b.plan()
    result = {
        "metric": METRIC,
        "value": n / wall_b,
        "unit": "pose-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall_b * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded complexes and poses)",
        "config": {"workload": f"uneven_batch: 64 complexes (seeds 100-163, receptor 500-8000 and ligand "
                               f"300-3000 atoms), {per} GA-initial poses each, host genes in / results out",
                   "parallelism": "single GPU"},
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
    print(json.dumps(result), flush=True)


def bench_serve(args):
    """Served docking stream: jobs (complexes drawn as in the uneven batch,
    per-job seeds 1000+i, 4-12 restraints) arrive as a Poisson process and run
    the full GA (population 150, early stop after 10 flat generations) and the
    FCC clustering of their Hall of Fame.  Server: continuous batching with
    memory admission.  Plain: the same jobs one at a time through gdock_ga_run
    (one scoring launch per generation); its latencies apply the same arrival
    times to the measured service times (first come, first served)."""
    import torch  # noqa: F401  (one HIP runtime, see gdock_amd._lib)
    from gdock_amd import ga, structure, synth
    from gdock_amd.device import Batch, Complex, default_context
    from gdock_amd.serve import Server

    cs = synth.uneven_batch(args.jobs, 1000)
    jobs = []
    for c in cs:
        rec = structure.read_pdb_text(c.receptor_pdb)[0]
        lig = structure.read_pdb_text(c.ligand_pdb)[0]
        jobs.append((rec, lig, c.restraints))
    rng = np.random.default_rng(7)
    if args.rate > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(jobs)))
        arrivals -= arrivals[0]
    else:  # burst: every job queued at once (the server's capacity)
        arrivals = np.zeros(len(jobs))
    ctx = default_context(0)
    # lane choices are measured once per complex shape and cached for the
    # process, as a deployed autotuning table would be
    Batch([Complex(r, structure.position_ligand(r, l), p, ctx) for r, l, p in jobs], [1] * len(jobs), ctx).close()

    srv = Server(max_jobs=64, max_receptor_atoms=8192)
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
    for rec, lig, pairs in jobs:
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
    plat = np.array(plat)
    pspan = free_at - arrivals[0]
    result = {
        "metric": METRIC,
        "value": len(jobs) / span * 60.0,
        "unit": "jobs/min",
        "n_gpus": 1, "steps": len(jobs), "warmup": 1,
        "ms_per_step": span / len(jobs) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded complexes, Poisson arrivals)",
        "config": {"workload": f"serve: {len(jobs)} GA docking jobs (receptor 500-8000, ligand 300-3000 atoms, "
                               f"population 150, early stop 10, FCC of the Hall of Fame), "
                               + (f"Poisson arrivals {args.rate:g} jobs/s" if args.rate > 0 else "all submitted at once"),
                   "parallelism": "single GPU"},
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "pose_evals_per_s": st["poses_scored"] / span,
        "generations": gens, "peak_active_jobs": st["peak_active"], "steps_run": st["steps"],
        "server_busy_s": st["busy_s"], "mean_step_ms": st["busy_s"] / max(st["steps"], 1) * 1e3,
        "mean_step_kernel_ms": st["kernel_s"] / max(st["steps"], 1) * 1e3,
        "server_admit_s": st["admit_s"], "server_ga_s": st["ga_s"], "server_finish_s": st["finish_s"],
        "span_s": span,
        "plain": {"jobs_per_min": len(jobs) / pspan * 60.0, "latency_p50_s": float(np.percentile(plat, 50)),
                  "latency_p99_s": float(np.percentile(plat, 99)), "mean_service_s": float(np.mean(service)),
                  "generations": pgens, "pose_evals_per_s": pgens * 150 / float(np.sum(service))},
        "speedup_jobs_per_min": pspan / span,
    }
    print(json.dumps(result), flush=True)


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
    if not args.no_cpu:
        import oracle.oracle as orc
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        ora = orc.GaRun(rec.atoms, lig.atoms, c.restraints, seed=42, nthreads=threads)
        cpu = time.perf_counter() - t0
        result["cpu_baseline"] = {"value": ora.generations * 150 / cpu, "unit": "pose-evals/s", "cores": threads,
                                  "kind": "port", "sample": f"the whole GA run ({ora.generations} generations), "
                                  f"{cpu:.1f} s wall, FP64 oracle"}
        result["cpu_job_wall_s"] = cpu
        result["same_run_as_oracle"] = bool(ora.generations == res.generations and ora.selected == res.selected)
    print(json.dumps(result), flush=True)


def cpu_baseline(rec, lig, pairs, genes, target_s):
    """FP64 oracle (reference restatement, one pose per task on a thread pool as
    rayon's par_iter_mut) on a sample of the same poses."""
    import oracle.oracle as orc
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    k = threads
    orc.score_poses(rec.atoms, lig.atoms, pairs, genes[:k], nthreads=threads)
    dt = time.perf_counter() - t0
    m = int(min(len(genes), max(k, k * round(target_s / max(dt, 1e-3)))))
    t0 = time.perf_counter()
    orc.score_poses(rec.atoms, lig.atoms, pairs, genes[:m], nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": m / dt, "unit": "pose-evals/s", "cores": threads, "kind": "port",
            "sample": f"first {m} of the same poses, {dt:.1f} s wall on {threads} threads (FP64 oracle, -O3 generic x86-64)"}


if __name__ == "__main__":
    main()
