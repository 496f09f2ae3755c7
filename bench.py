"""MDDF benchmark on one MI355X: frames/s on the 5k-atom solute / 40k-water system.

Workload (one step = one full mddf job over the trajectory, inputs resident in
HBM): cubic 110 A box, 5,000-atom random-walk globule (radius ~25 A), 40,000
rigid 3-site waters outside it, 200 jittered frames, FP32 coordinates,
Options(): cutoff = dbulk = 10 A, binstep 0.02 A, n_random_samples 10, all
5,000 per-atom solute groups.

Prints one JSON line.  roofline: FP32 VALU work of the minimum-distance
kernels (12 FP32 ops per useful pair, W pairs per frame) over their measured
time (HIP events on the job's stream).  cpu_baseline: the CPU FP64
restatement of the reference (oracle/, cell lists, the reference's chunked
parallel setup) timed on a bounded sample of the same frames.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as G  # noqa: E402

METRIC = "frames/s (5k-atom solute, 40k waters) GPU vs CPU; served jobs/s + p99 latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_k_eval_tiles_config2.md")


def measured_traffic_per_frame():
    """HBM bytes per frame of the evaluation kernels (real + random passes),
    from the FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/
    (each scaled by the same counter's reading for a copy kernel of known
    size); None when that summary is absent."""
    try:
        txt = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    import re
    m = re.search(r"traffic_bytes_per_frame:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def work_per_frame(n_u, n_v_atoms, volume, n_rand, cutoff=10.0):
    """Useful (within-cutoff at uniform density) pair evaluations per frame:
    W = N_u rho_v V_c + n_rand n_u rho_v V_c (single-molecule solute)."""
    rho = n_v_atoms / volume
    vc = 4.0 / 3.0 * math.pi * cutoff ** 3
    return n_u * rho * vc * (1 + n_rand)


def bytes_per_frame(n_u, n_v_atoms, m_v, n_rand, coord_bytes=4):
    """Algorithmic HBM bytes per frame: coordinates in, placed solvent written
    and read per random pass, one 16-byte record per solvent molecule per pass."""
    return 3 * coord_bytes * (n_u + n_v_atoms) + n_rand * 24 * n_v_atoms + 16 * m_v * (1 + n_rand)


def usable_cores():
    """CPU cores this process may run on: the affinity mask, capped by the
    cgroup CPU quota and by the host's declared CPU share (OMP_NUM_THREADS)
    when they are set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    try:
        n = min(n, max(1, int(os.environ.get("OMP_NUM_THREADS", n))))
    except ValueError:
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pkg, O, s, opts, tm, nframes_sample, threads):
    """The oracle (FP64, cell lists, the reference's chunked threads) over the
    first frames; returns (frames/s, seconds, raw accumulators)."""
    from complexmixtures_amd.trajectory import convert_unitcell
    job = O.Job(1, s["solute"].natomspermol, s["solvent"].nmols, s["solvent"].natomspermol,
                irefatom=tm.irefatom - 1, cutoff=opts.cutoff, dbulk=opts.dbulk,
                n_random_samples=opts.n_random_samples, seed=opts.seed, nthreads=threads)
    f = nframes_sample
    xu = s["x_solute"][:f].astype(np.float64)
    xv = s["x_solvent"][:f].astype(np.float64)
    cells = np.stack([convert_unitcell(c, True) for c in s["cells"][:f]])
    t = time.perf_counter()
    raw = O.run(job, list(range(1, f + 1)), xu, xv, cells, np.ones(f))
    dt = time.perf_counter() - t
    return f / dt, dt, raw


SQ_FILE = os.path.join(ROOT, "profiles", "sq_counters_config2_200frames.md")


def sq_valu_summary():
    """VALU issue fraction and SQ_INSTS_VALU per frame of the random-pass
    evaluation kernel from the SQ counter summary in profiles/ (None when
    absent)."""
    try:
        txt = open(SQ_FILE).read()
    except OSError:
        return None
    import re
    m = re.search(r"valu_issue_fraction:\s*([0-9.eE+]+)", txt)
    n = re.search(r"valu_wave_instructions_per_frame:\s*([0-9.eE+]+)", txt)
    if not (m and n):
        return None
    return float(m.group(1)), float(n.group(1))


def config3_rates(args):
    """Frames/s of the triclinic mixed case (a = b = c = 90 A, 80 degree angles,
    FP64 input): 3,000-atom polymer vs 12,000 urea and vs 15,000 waters, and
    urea vs urea (self-MDDF), n_random_samples 10."""
    import complexmixtures_amd as cm
    import complexmixtures_amd.synthetic as S
    from complexmixtures_amd import _lib
    from complexmixtures_amd.mddf import TrajectoryMetaData, _job_for
    from complexmixtures_amd.results import make_result
    from complexmixtures_amd.trajectory import convert_unitcell
    out = {}
    nf = args.config3_frames
    for name in ("urea", "water", "self"):
        s = S.config3(nframes=nf, solvent=name)
        opts = cm.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261205, silent=True)
        traj = cm.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])
        tm = TrajectoryMetaData(traj, opts)
        R = make_result(traj, opts, tm)
        du = _lib.DeviceArray(s["x_solute"])
        dv = du if name == "self" else _lib.DeviceArray(s["x_solvent"])
        cells = np.stack([convert_unitcell(c, tm.orthorhombic) for c in s["cells"]])
        rate = None
        for rep in range(2):
            job = _job_for(R, opts, tm, False)
            job.set_profiling(True)
            t = time.perf_counter()
            job.submit_device(np.arange(1, nf + 1), du, dv, 0, cells, np.ones(nf))
            job.finish()
            el = time.perf_counter() - t
            st = job.stats()
            job.close()
            rate = {"frames_per_s": round(nf / el, 2), "eval_ms_per_frame": round(st["eval_ms"] / nf, 4),
                    "fp64_exhaustive_resolves": st["fp64_rechecks"]}
        out["polymer_vs_" + name if name != "self" else "urea_self"] = rate
        du.free()
        if dv is not du:
            dv.free()
    return {"frames": nf, "triclinic": True, "input": "f64", **out}


def self_water_rate(args):
    """Self-MDDF of water, coordination numbers: 100 A cubic box, 33,400 3-site
    waters (the largest self requests of the served stream), 40 frames, FP32."""
    from complexmixtures_amd import _lib
    import complexmixtures_amd.serve as SV
    from complexmixtures_amd.synthetic import served_system
    nf = 40
    s = served_system(7, 2000, 100.0, nf)
    sysm = SV.System(n_solute=2000, n_water=s["n_water"], box=100.0, nframes=nf,
                     d_solute=_lib.DeviceArray(s["x_solute"]), d_water=_lib.DeviceArray(s["x_solvent"]),
                     cells=s["cells"], irefatom=SV.default_irefatom(s["x_solvent"][0]))
    kw, du, dv, _ = SV.job_kwargs(sysm, SV.Request(0.0, 0, nf, False, True, 99))
    for rep in range(3):
        job = _lib.Job(**kw)
        job.set_profiling(True)
        t = time.perf_counter()
        job.submit_device(np.arange(1, nf + 1), du, dv, 0, sysm.cells[:nf], np.ones(nf))
        raw = job.finish()
        el = time.perf_counter() - t
        st = job.stats()
        job.close()
    sysm.d_solute.free()
    sysm.d_water.free()
    # each unordered pair within the cutoff is two records (both orientations)
    return {"workload": f"self-MDDF coordination numbers, 100 A cubic box, {s['n_water']} waters, {nf} frames, FP32",
            "frames_per_s": round(nf / el, 2), "eval_ms_per_frame": round(st["eval_ms"] / nf, 4),
            "pairs_within_cutoff_per_frame": round(float(raw["md_count"].sum()) / 2 / nf),
            "fp64_exhaustive_resolves": st["fp64_rechecks"]}


def serve_capacity_once(args):
    """One capacity run of one serving path (every request of the served
    stream queued at once), after the same pool, autotuning and first use of
    the kernels as served_stream: the workload of a kernel trace of that path."""
    import complexmixtures_amd.serve as SV
    pool = SV.make_pool(args.serve_systems, 20261206, args.serve_frames)
    for s in pool:
        SV.autotune(s, frames=4)
    rng = np.random.default_rng(20261207)
    reqs = SV.request_stream(rng, args.serve_requests, len(pool), args.serve_frames, rate=20.0)
    sat = [SV.Request(0.0, r.system, r.nframes, r.mddf, r.self_, r.seed) for r in reqs]
    SV.run_plain(pool, [r for r in sat if r.self_][:1] + sat[:1])
    _lib_sync()
    path = args.serve_capacity_path
    r = SV.run_served(pool, sat, slots=args.slots) if path == "served" else SV.run_plain(pool, sat)
    return {"path": path, "requests": len(sat), "jobs_per_s": round(r["jobs_per_s"], 3),
            "makespan_s": round(r["makespan_s"], 4), "p99_s": round(r["p99_s"], 4)}


def batched_systems(args):
    """Many uneven systems in one launch per stage: independent cubic systems
    (solute 200-20,000 atoms log-uniform, water at 0.0334 molecules/A^3 in
    boxes of 30-150 A), 50 jittered frames each, FP32, cutoff 10 A, binstep
    0.02 A, n_random_samples 10.  Plain: every system as its own job, one after
    the other (create, submit, finish).  Batched: all jobs in one
    cm_jobs_submit_device (rounds of one batch of every job, each stage one
    launch over all of them, replayed from a HIP graph), then every job
    finished.  Both paths' counts must be identical."""
    import complexmixtures_amd.serve as SV
    from complexmixtures_amd import _lib
    t0 = time.perf_counter()
    pool = SV.make_pool(args.batch_systems, 20261208, args.batch_frames)
    gen_s = time.perf_counter() - t0
    nf = args.batch_frames
    reqs = [SV.Request(0.0, k, nf, True, False, 40000 + k) for k in range(len(pool))]
    specs = [SV.job_kwargs(pool[r.system], r) for r in reqs]

    def plain():
        out = []
        for (kw, du, dv, _), r in zip(specs, reqs):
            job = _lib.Job(**kw)
            job.submit_device(np.arange(1, nf + 1), du, dv, 0, pool[r.system].cells[:nf], np.ones(nf))
            out.append(job.finish())
            job.close()
        return out

    def batched():
        jobs = [_lib.Job(**kw) for kw, _, _, _ in specs]
        _lib.submit_many([(j, np.arange(1, nf + 1), du, dv, 0, pool[r.system].cells[:nf], np.ones(nf))
                          for j, (kw, du, dv, _), r in zip(jobs, specs, reqs)])
        out = [j.finish() for j in jobs]
        for j in jobs:
            j.close()
        return out

    plain()
    batched()
    _lib_sync()
    tp, tb, same = [], [], True
    for _ in range(3):
        t = time.perf_counter()
        a = plain()
        tp.append(time.perf_counter() - t)
        s0 = _lib.multi_stats()
        t = time.perf_counter()
        b = batched()
        tb.append(time.perf_counter() - t)
        s1 = _lib.multi_stats()
        same = same and all(SV.digest(x) == SV.digest(y) for x, y in zip(a, b))
    mp, mb = sorted(tp)[1], sorted(tb)[1]
    frames = nf * len(pool)
    for s_ in pool:
        s_.d_solute.free()
        s_.d_water.free()
    return {"workload": f"{len(pool)} cubic systems (solute 200-20,000 atoms log-uniform, water boxes 30-150 A), "
                        f"{nf} frames each, FP32, n_random_samples 10",
            "solute_atoms": [p.n_solute for p in pool], "waters": [p.n_water for p in pool],
            "plain": {"jobs_per_s": round(len(pool) / mp, 3), "frames_per_s": round(frames / mp, 1),
                      "seconds": round(mp, 4)},
            "batched": {"jobs_per_s": round(len(pool) / mb, 3), "frames_per_s": round(frames / mb, 1),
                        "seconds": round(mb, 4), "rounds": s1["rounds"] - s0["rounds"],
                        "stage_launches": s1["launches"] - s0["launches"]},
            "ratio": round(mp / mb, 3), "counts_identical": bool(same), "runs": "median of 3 per path",
            "gen_s": round(gen_s, 2)}


def _lib_sync():
    from complexmixtures_amd import _lib
    _lib.check(_lib.lib().cm_device_synchronize())


def served_stream(args):
    """Served mixed-size stream (Poisson arrivals at 20 requests/s) through the
    request server vs the plain one-job-at-a-time path, plus both paths'
    capacity with every request queued at once; each request's counts must be
    the same on both paths."""
    import complexmixtures_amd.serve as SV
    t0 = time.perf_counter()
    pool = SV.make_pool(args.serve_systems, 20261206, args.serve_frames)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tunings = [SV.autotune(s, frames=4)[0] for s in pool]
    tune_s = time.perf_counter() - t0
    rng = np.random.default_rng(20261207)
    reqs = SV.request_stream(rng, args.serve_requests, len(pool), args.serve_frames, rate=20.0)
    sat = [SV.Request(0.0, r.system, r.nframes, r.mddf, r.self_, r.seed) for r in reqs]
    SV.run_plain(pool, [r for r in sat if r.self_][:1] + sat[:1])          # first use of every kernel
    plain = SV.run_plain(pool, reqs)
    served = SV.run_served(pool, reqs, slots=args.slots)
    # capacity: every request queued at once, three runs of each path
    # alternated; the run with the median rate is reported
    caps_p, caps_s = [], []
    for _ in range(3):
        caps_p.append(SV.run_plain(pool, sat))
        caps_s.append(SV.run_served(pool, sat, slots=args.slots))
    med = lambda runs: sorted(runs, key=lambda d: d["jobs_per_s"])[len(runs) // 2]
    plain_cap, served_cap = med(caps_p), med(caps_s)
    same = all(plain["digests"] == d["digests"] for d in [served] + caps_p + caps_s)
    keep = ("jobs_per_s", "p50_s", "p99_s", "mean_s", "makespan_s")
    r3 = lambda d: {k: round(d[k], 4) for k in keep}
    frames = sum(r.nframes for r in reqs)
    out = {"workload": f"{len(reqs)} requests over {len(pool)} systems (solute 200-20,000 atoms log-uniform, cubic "
                       "water boxes 30-150 A), 10-200 frames each, 30% with 10 ideal-gas passes, 20% self-MDDF, "
                       "Poisson arrivals 20/s",
           "frames_total": frames, "slots": args.slots, "counts_identical": bool(same),
           "served": r3(served), "plain": r3(plain),
           "capacity": {"served_jobs_per_s": round(served_cap["jobs_per_s"], 3),
                        "plain_jobs_per_s": round(plain_cap["jobs_per_s"], 3),
                        "ratio": round(served_cap["jobs_per_s"] / plain_cap["jobs_per_s"], 3),
                        "served_p99_s": round(served_cap["p99_s"], 4), "plain_p99_s": round(plain_cap["p99_s"], 4),
                        "runs": "median of 3 per path",
                        "served_runs_jobs_per_s": [round(d["jobs_per_s"], 2) for d in caps_s],
                        "plain_runs_jobs_per_s": [round(d["jobs_per_s"], 2) for d in caps_p]},
           "plain_service_by_class": SV.service_breakdown(sat, plain_cap["service_s"]),
           "tunings": tunings, "pool_gen_s": round(gen_s, 2), "autotune_s": round(tune_s, 2)}
    for s in pool:
        s.d_solute.free()
        s.d_water.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--cpu-frames", type=int, default=0, help="CPU baseline sample (0: one frame per thread)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true", help="only time the evaluation kernels per tuning, then exit")
    ap.add_argument("--cell-fraction", type=float, default=0.0, help="tuning: fine cell side / cutoff (0: default)")
    ap.add_argument("--columns", type=int, default=0, help="tuning: columns per solvent atom row (0: default)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--serve-only", action="store_true", help="only the served-stream measurement")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-systems", type=int, default=12)
    ap.add_argument("--serve-frames", type=int, default=200)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--serve-capacity-path", choices=("served", "plain"), default=None,
                    help="only one capacity run of that serving path (for kernel traces)")
    ap.add_argument("--no-config3", action="store_true", help="skip the triclinic mixed-case rates")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-systems measurement")
    ap.add_argument("--batched-only", action="store_true", help="only the many-systems measurement")
    ap.add_argument("--batch-systems", type=int, default=64)
    ap.add_argument("--batch-frames", type=int, default=50)
    ap.add_argument("--self-water-only", action="store_true", help="only the self-MDDF water rate")
    ap.add_argument("--config3-frames", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain read-N/write-N elementwise kernel over MIB MiB three times (the "
                         "reference for scaling FETCH_SIZE/WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; the path shards by frames across GPUs "
                                   "but multi-GPU runs are not part of it"}))
        sys.exit(2)
    if args.copy_probe:
        import torch
        n = args.copy_probe * 2 ** 20 // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 3}))
        return
    pkg = G.load_package()
    if args.serve_capacity_path:
        print(json.dumps({"serve_capacity": serve_capacity_once(args)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"served_stream": served_stream(args)}), flush=True)
        return
    if args.self_water_only:
        print(json.dumps({"self_water": self_water_rate(args)}), flush=True)
        return
    if args.batched_only:
        print(json.dumps({"batched_systems": batched_systems(args)}), flush=True)
        return
    import complexmixtures_amd.synthetic as S
    from complexmixtures_amd import _lib
    from complexmixtures_amd.mddf import TrajectoryMetaData, _job_for
    from complexmixtures_amd.results import make_result
    from complexmixtures_amd.trajectory import convert_unitcell

    t0 = time.perf_counter()
    s = S.config2(nframes=args.frames)
    gen_s = time.perf_counter() - t0
    opts = pkg.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261204, silent=True)
    traj = pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])
    tm = TrajectoryMetaData(traj, opts)
    R = make_result(traj, opts, tm)
    nf = args.frames
    du = _lib.DeviceArray(s["x_solute"])
    dv = _lib.DeviceArray(s["x_solvent"])
    cells = np.stack([convert_unitcell(c, tm.orthorhombic) for c in s["cells"]])
    frames = np.arange(1, nf + 1)
    weights = np.ones(nf)

    if args.cell_fraction or args.columns:
        tuning = dict(cell_fraction=args.cell_fraction or 0.5, columns=args.columns)
    else:
        # the autotuner's choice on the first 40 frames (one batch; untimed)
        import complexmixtures_amd.serve as SV
        from complexmixtures_amd.mddf import job_kwargs
        tuning, _ = SV.autotune_job(job_kwargs(R, opts, tm, False), du, dv, cells, frames=40)

    def one_job(profile=False, nframes=nf):
        job = _job_for(R, opts, tm, False, tuning=tuning)
        if profile:
            job.set_profiling(True)
        job.submit_device(frames[:nframes], du, dv, 0, cells[:nframes], weights[:nframes])
        raw = job.finish()
        st = None
        if profile:
            st = job.stats()
            st["launches"] = job.launch_times()
        job.close()
        return raw, st

    if args.sweep:
        # eval-kernel time per tuning over the same frames (results are the
        # same for every tuning; only the work decomposition changes)
        base = None
        for cf in (0.25, 0.35, 0.5, 0.75, 1.0):
            for c in (1, 2, 4, 8):
                job = _job_for(R, opts, tm, False, tuning=dict(cell_fraction=cf, columns=c))
                job.set_profiling(True)
                t1 = time.perf_counter()
                job.submit_device(frames, du, dv, 0, cells, weights)
                raw = job.finish()
                dt = time.perf_counter() - t1
                st = job.stats()
                job.close()
                same = base is None or all(np.array_equal(raw[k], base[k]) for k in raw if k != "volume_total")
                base = base or raw
                print(json.dumps({"cell_fraction": cf, "columns": c, "eval_ms": round(st["eval_ms"], 3),
                                  "total_ms": round(st["total_ms"], 3), "wall_ms": round(1e3 * dt, 3),
                                  "fp64_rechecks": st["fp64_rechecks"], "same_counts": same}), flush=True)
        return
    for _ in range(args.warmup):
        one_job(profile=True)
    _lib.check(_lib.lib().cm_device_synchronize())
    # every timed job records HIP events around its evaluation launches (no
    # host wait; read after the job), so the roofline is the timed region's own
    sts = []
    t = time.perf_counter()
    for _ in range(args.steps):
        raw, st = one_job(profile=True)
        sts.append(st)
    _lib.check(_lib.lib().cm_device_synchronize())
    el = time.perf_counter() - t
    fps = nf * args.steps / el
    st = {"eval_ms": sum(x["eval_ms"] for x in sts) / len(sts),
          "eval_launches": sts[0]["eval_launches"], "fp64_rechecks": sts[-1]["fp64_rechecks"]}
    real_ms = np.concatenate([x["launches"][0] for x in sts])
    rand_ms = np.concatenate([x["launches"][1] for x in sts])
    # the same job from host buffers (cm_job_submit: coordinates cross PCIe per batch)
    job = _job_for(R, opts, tm, False, tuning=tuning)
    t1 = time.perf_counter()
    job.submit(frames, s["x_solute"], s["x_solvent"], cells, weights)
    job.finish()
    host_fps = nf / (time.perf_counter() - t1)
    job.close()

    n_u = s["solute"].natomspermol
    n_v = s["solvent"].nmols * s["solvent"].natomspermol
    vol = float(np.linalg.det(cells[0]))
    W = work_per_frame(n_u, n_v, vol, opts.n_random_samples, opts.dbulk)
    ops = 12.0 * W * nf
    eval_s = st["eval_ms"] / 1e3
    achieved = ops / eval_s / 1e12
    bpf = bytes_per_frame(n_u, n_v, s["solvent"].nmols, opts.n_random_samples)
    tpf = measured_traffic_per_frame()
    # frames per launch pair (real + random evaluation) of a job
    launch_pairs = max(1, len(real_ms) // args.steps)
    frames_per_launch = -(-nf // launch_pairs)
    # the same from the median launch pair instead of the mean
    med_pair_s = (float(np.median(real_ms)) + float(np.median(rand_ms))) / 1e3
    achieved_med = 12.0 * W * frames_per_launch / med_pair_s / 1e12
    sq = sq_valu_summary()
    out = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * el / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "mddf: 110 A cubic box, 5,000-atom globule solute, 40,000 3-site waters, "
                               f"{nf} frames, FP32 coordinates, cutoff 10 A, binstep 0.02 A, n_random_samples 10",
                   "frames_per_step": nf, "solute_atoms": n_u, "solvent_molecules": s["solvent"].nmols},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                     "traffic": round(tpf * frames_per_launch) if tpf else None,
                     "traffic_note": f"HBM bytes per {frames_per_launch}-frame launch pair (real + random "
                                     "evaluation), from the counter summary in profiles/",
                     "algorithmic_bytes_per_launch": round(bpf * frames_per_launch),
                     "kernel": "k_eval_tiles (real + random passes)", "pairs_per_frame": round(W),
                     "eval_ms_per_frame": round(st["eval_ms"] / nf, 5),
                     "eval_ms_job": round(st["eval_ms"], 3), "eval_launches_job": st["eval_launches"],
                     "timing": f"HIP events around every evaluation launch of the {args.steps} timed jobs",
                     "frac_from_median_launch": round(achieved_med / PEAK_FP32_TFLOPS, 5),
                     "launch_us_real": {"median": round(1e3 * float(np.median(real_ms)), 1),
                                        "mean": round(1e3 * float(np.mean(real_ms)), 1),
                                        "max": round(1e3 * float(np.max(real_ms)), 1)},
                     "launch_us_random": {"median": round(1e3 * float(np.median(rand_ms)), 1),
                                          "mean": round(1e3 * float(np.mean(rand_ms)), 1),
                                          "max": round(1e3 * float(np.max(rand_ms)), 1)},
                     "valu_issue_fraction_random_pass": sq[0] if sq else None,
                     "valu_lane_instructions_per_W_pair": round(64 * sq[1] / W, 1) if sq else None,
                     "hbm_GBps_algorithmic": round(bpf * nf / eval_s / 1e9, 2),
                     "fp64_exhaustive_resolves": st["fp64_rechecks"]},
        "tuning": tuning,
        "host_input_frames_per_s": round(host_fps, 2),
        "gen_s": round(gen_s, 2),
    }
    if not args.no_cpu:
        O = G.load_oracle()
        threads = args.cpu_threads or usable_cores()
        f = args.cpu_frames or min(nf, 10 * threads)
        v, dt, oraw = cpu_baseline(pkg, O, s, opts, tm, f, threads)
        # the GPU's counts over the same frames, bit for bit
        graw, _ = one_job(nframes=f)
        keys = {"md_count": "md_count", "md_count_random": "md_count_random", "rdf_count": "rdf_count",
                "rdf_count_random": "rdf_count_random", "solute_group_count": "gu", "solvent_group_count": "gv",
                "solute_group_count_random": "gu_random", "solvent_group_count_random": "gv_random"}
        exact = all(np.array_equal(graw[k], oraw[o]) for k, o in keys.items())
        out["cpu_baseline"] = {"value": round(v, 4), "unit": "frames/s", "cores": threads, "kind": "port",
                               "cpu": cpu_model(), "cores_usable": usable_cores(), "cpu_count": os.cpu_count(),
                               "sample": f"first {f} frames of the same trajectory, {dt:.1f} s",
                               "gpu_over_cpu": round(fps / v, 1)}
        out["bit_exact"] = {"value": bool(exact),
                            "what": f"every raw accumulator of the GPU job over the first {f} frames equals the "
                                    "FP64 oracle's (the CPU baseline's own run)"}
    du.free()
    dv.free()
    if not args.no_batched:
        out["batched_systems"] = batched_systems(args)
    if not args.no_serve:
        out["served_stream"] = served_stream(args)
    if not args.no_config3:
        out["triclinic_mixed"] = config3_rates(args)
        out["self_water"] = self_water_rate(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
