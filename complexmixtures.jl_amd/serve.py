"""Serving MDDF requests on one GPU.

Server: continuous batching through the request server of libcmgpu
(cm_server_*): requests are admitted in arrival order as slots and memory free
up, each slot drives its own stream, so the work of admitted requests
overlaps on the device.  Plain path: one job at a time (cm_job_create,
cm_job_submit_device, cm_job_finish, cm_job_destroy), the way a single mddf
call runs.  Both read the same device-resident systems; each request's counts
are the same either way.

The request stream is the uneven mixed-size workload: Poisson arrivals,
systems whose solute size is log-uniform over 200-20,000 atoms in cubic water
boxes of 30-150 A, 10-200 frames per request, 30% with the ideal-gas reference
passes (n_random_samples 10; the others are coordination-number requests) and
20% self-MDDF (water vs water).  The autotuner picks the cell size and the
tile (columns per row) of the single-solute kernels per system by timing
them on a few frames.
"""
import hashlib
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .results import setbin
from .synthetic import served_system

CUTOFF = 10.0
BINSTEP = 0.02
N_RANDOM = 10


@dataclass
class System:
    n_solute: int
    n_water: int
    box: float
    nframes: int
    d_solute: object
    d_water: object
    cells: np.ndarray
    irefatom: int
    tuning: dict = field(default_factory=dict)
    host: dict = None


def default_irefatom(x_water_frame0):
    """The reference's default irefatom (src/Trajectory.jl:206-211): the atom of
    the first solvent molecule closest to that molecule's centre, frame 1."""
    m = np.asarray(x_water_frame0[:3], dtype=np.float64)
    return int(np.argmin(np.linalg.norm(m - m.mean(axis=0), axis=1))) + 1


def make_pool(n_systems, seed, nframes, n_solute_range=(200, 20000), box_range=(30.0, 150.0), dtype=np.float32,
              keep_host=False):
    """Systems of the uneven distribution, uploaded once to device memory
    (keep_host: also keep the host coordinates)."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_systems):
        lo, hi = np.log(n_solute_range[0]), np.log(n_solute_range[1])
        nu = int(round(np.exp(rng.uniform(lo, hi))))
        box = float(rng.uniform(*box_range))
        s = served_system(seed * 1000 + k, nu, box, nframes, dtype=dtype)
        pool.append(System(n_solute=nu, n_water=s["n_water"], box=box, nframes=nframes,
                           d_solute=_lib.DeviceArray(s["x_solute"]), d_water=_lib.DeviceArray(s["x_solvent"]),
                           cells=s["cells"], irefatom=default_irefatom(s["x_solvent"][0]),
                           host=s if keep_host else None))
    return pool


@dataclass
class Request:
    arrival: float
    system: int
    nframes: int
    mddf: bool
    self_: bool
    seed: int


def request_stream(rng, n, n_systems, max_frames, rate=20.0, frames=(10, 200), p_mddf=0.3, p_self=0.2,
                   saturate=False):
    """n requests; arrivals are a Poisson process of the given rate (all at 0
    when saturate)."""
    t = 0.0
    out = []
    for i in range(n):
        if not saturate:
            t += rng.exponential(1.0 / rate)
        out.append(Request(arrival=t, system=int(rng.integers(n_systems)),
                           nframes=int(min(max_frames, rng.integers(frames[0], frames[1] + 1))),
                           mddf=bool(rng.random() < p_mddf), self_=bool(rng.random() < p_self),
                           seed=int(rng.integers(1, 2 ** 62))))
    return out


def job_kwargs(system, req):
    """Job arguments and device inputs of one request (default Options with
    cutoff = dbulk = 10 A, binstep 0.02 A; default groups)."""
    common = dict(nbins=setbin(CUTOFF, BINSTEP), binstep=BINSTEP, cutoff=CUTOFF, dbulk=CUTOFF, usecutoff=False,
                  n_random_samples=N_RANDOM, coordination_number_only=not req.mddf, triclinic=False, seed=req.seed,
                  irefatom=system.irefatom)
    if req.self_:
        kw = dict(nmols_u=system.n_water, natoms_per_mol_u=3, nmols_v=system.n_water, natoms_per_mol_v=3,
                  autocorrelation=True, n_groups_u=3, n_groups_v=3, **common)
        return kw, system.d_water, system.d_water, {}
    kw = dict(nmols_u=1, natoms_per_mol_u=system.n_solute, nmols_v=system.n_water, natoms_per_mol_v=3,
              autocorrelation=False, n_groups_u=system.n_solute, n_groups_v=3, **common)
    return kw, system.d_solute, system.d_water, system.tuning


def digest(raw):
    """SHA-1 over every raw accumulator (bit patterns) and the volume."""
    h = hashlib.sha1()
    for k in sorted(raw):
        if k == "volume_total":
            h.update(np.float64(raw[k]).tobytes())
        else:
            h.update(np.ascontiguousarray(raw[k]).tobytes())
    return h.hexdigest()


def _summary(lat, t_first, t_last, n):
    lat = np.sort(np.asarray(lat))
    span = max(t_last - t_first, 1e-9)
    return {"jobs_per_s": n / span, "makespan_s": span, "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean())}


class _Digests:
    """SHA-1 digests computed on a worker thread, off both paths' clocks (the
    hash releases the GIL, as the ctypes calls do)."""

    def __init__(self, n, keep_raw):
        self.out = [None] * n
        self.raw = [None] * n if keep_raw else None
        self.q = queue.Queue()
        self.th = threading.Thread(target=self._run)
        self.th.start()

    def _run(self):
        while True:
            item = self.q.get()
            if item is None:
                return
            i, raw = item
            self.out[i] = digest(raw)
            if self.raw is not None:
                self.raw[i] = raw

    def put(self, i, raw):
        self.q.put((i, raw))

    def close(self):
        self.q.put(None)
        self.th.join()


def run_served(pool, reqs, slots=4, memory_budget=0, keep_raw=False):
    """Requests enqueued at their arrival times on a cm_server.  A request is
    done when its counts are in the caller's arrays (cm_server_wait returned);
    latency is that time minus its arrival, on the same clock as run_plain."""
    srv = _lib.Server(slots, memory_budget)
    handoff = queue.Queue()
    done = [None] * len(reqs)
    dg = _Digests(len(reqs), keep_raw)
    t0 = time.perf_counter()

    def collect():
        while True:
            item = handoff.get()
            if item is None:
                return
            i, tk = item
            raw, _ = srv.wait(tk)
            done[i] = time.perf_counter() - t0
            dg.put(i, raw)

    th = threading.Thread(target=collect)
    th.start()
    try:
        for i, r in enumerate(reqs):
            dt = t0 + r.arrival - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            s = pool[r.system]
            kw, du, dv, tuning = job_kwargs(s, r)
            tk = srv.enqueue(np.arange(1, r.nframes + 1), du, dv, 0, s.cells[:r.nframes], np.ones(r.nframes),
                             tuning=tuning, **kw)
            handoff.put((i, tk))
    finally:
        handoff.put(None)
        th.join()
        srv.close()
        dg.close()
    lat = [done[i] - r.arrival for i, r in enumerate(reqs)]
    out = _summary(lat, min(r.arrival for r in reqs), max(done), len(reqs))
    out["digests"] = dg.out
    if keep_raw:
        out["raw"] = dg.raw
    return out


def run_plain(pool, reqs, keep_raw=False):
    """The same requests one at a time (each starts when it has arrived and the
    previous one is done; done = counts in the caller's arrays)."""
    dg = _Digests(len(reqs), keep_raw)
    t0 = time.perf_counter()
    lat, fin, service = [], [], []
    try:
        for i, r in enumerate(reqs):
            dt = t0 + r.arrival - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t_start = time.perf_counter()
            s = pool[r.system]
            kw, du, dv, tuning = job_kwargs(s, r)
            job = _lib.Job(tuning=tuning or None, **kw)
            job.submit_device(np.arange(1, r.nframes + 1), du, dv, 0, s.cells[:r.nframes], np.ones(r.nframes))
            raw = job.finish()
            job.close()
            t = time.perf_counter() - t0
            service.append(time.perf_counter() - t_start)
            fin.append(t)
            lat.append(t - r.arrival)
            dg.put(i, raw)
    finally:
        dg.close()
    out = _summary(lat, min(r.arrival for r in reqs), max(fin), len(reqs))
    out["digests"] = dg.out
    out["service_s"] = service
    if keep_raw:
        out["raw"] = dg.raw
    return out


def service_breakdown(reqs, service):
    """Plain-path service time per request class: self-MDDF, cross with
    ideal-gas passes, cross coordination numbers."""
    out = {}
    for name, sel in (("self", lambda r: r.self_), ("cross_mddf", lambda r: not r.self_ and r.mddf),
                      ("cross_cn", lambda r: not r.self_ and not r.mddf)):
        idx = [i for i, r in enumerate(reqs) if sel(r)]
        if idx:
            tot = sum(service[i] for i in idx)
            out[name] = {"requests": len(idx), "seconds": round(tot, 3),
                         "ms_per_frame": round(1e3 * tot / sum(reqs[i].nframes for i in idx), 3)}
    return out


def tuning_candidates(natoms_per_mol_v):
    """Cell fractions x columns per row: a row is a whole molecule for 3-atom
    solvent molecules (2-8 columns: 32-8 molecules per chunk), one solvent
    atom otherwise (1-2 columns)."""
    cols = (2, 4, 8) if natoms_per_mol_v == 3 else (1, 2)
    return [dict(cell_fraction=cf, columns=c) for cf in (0.35, 0.5, 0.75, 1.0) for c in cols]


TUNING_CANDIDATES = tuning_candidates(3)


def autotune_job(job_kw, d_solute, d_solvent, cells, frames=4, candidates=None, repeats=2):
    """Times the single-solute kernels of one job description on its first
    frames for each candidate tuning (kernel time from HIP events, the fastest
    of `repeats` runs, after one untimed run that takes the first-use costs of
    the process) and returns (best tuning, [(ms, tuning)]); every candidate
    must give the same counts."""
    if candidates is None:
        candidates = tuning_candidates(job_kw["natoms_per_mol_v"])
    nf = min(frames, d_solute.shape[0])

    def run(cand):
        job = _lib.Job(tuning=cand, **job_kw)
        job.set_profiling(True)
        job.submit_device(np.arange(1, nf + 1), d_solute, d_solvent, 0, cells[:nf], np.ones(nf))
        raw = job.finish()
        ms = job.stats()["total_ms"]
        job.close()
        return ms, digest(raw)

    run(candidates[0])
    best, ref, timing = None, None, []
    for cand in candidates:
        ms = None
        for _ in range(max(1, repeats)):
            t, d = run(cand)
            if ref is None:
                ref = d
            elif d != ref:
                raise RuntimeError(f"tuning {cand} changed the counts")
            ms = t if ms is None else min(ms, t)
        timing.append((ms, cand))
        if best is None or ms < best[0]:
            best = (ms, cand)
    return dict(best[1]), timing


def autotune(system, frames=4, candidates=None):
    """Autotunes the cross (single-solute) requests of a system with the
    ideal-gas passes on and keeps the result in system.tuning."""
    req = Request(arrival=0.0, system=0, nframes=min(frames, system.nframes), mddf=True, self_=False, seed=12345)
    kw, du, dv, _ = job_kwargs(system, req)
    system.tuning, timing = autotune_job(kw, du, dv, system.cells, frames, candidates)
    return system.tuning, timing
