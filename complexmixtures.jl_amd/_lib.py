"""ctypes binding of libcmgpu.so (include/cmgpu.h).

The device library is the only compute path: if it is not built, or no GPU
is visible, calls raise -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from .options import ArgumentError

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libcmgpu.so")

CM_OK, CM_EINVAL, CM_EHIP, CM_ENOMEM, CM_EUNSUPPORTED, CM_ESTATE = 0, -1, -2, -3, -4, -5
CM_F64, CM_F32 = 0, 1

_P = ctypes.POINTER
_i32, _i64, _u64, _d = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double


class JobDesc(ctypes.Structure):
    _fields_ = [("nmols_u", _i32), ("natoms_per_mol_u", _i32), ("nmols_v", _i32), ("natoms_per_mol_v", _i32),
                ("autocorrelation", _i32), ("irefatom", _i32), ("nbins", _i32), ("usecutoff", _i32),
                ("binstep", _d), ("cutoff", _d), ("dbulk", _d),
                ("n_random_samples", _i32), ("coordination_number_only", _i32), ("triclinic", _i32), ("lcell", _i32),
                ("seed", _u64), ("n_groups_u", _i32), ("n_groups_v", _i32),
                ("groups_u_ptr", _P(_i32)), ("groups_u_idx", _P(_i32)),
                ("groups_v_ptr", _P(_i32)), ("groups_v_idx", _P(_i32)),
                ("max_batch_frames", _i64), ("memory_budget", _i64)]


class RawOut(ctypes.Structure):
    _fields_ = [("md_count", _P(_d)), ("md_count_random", _P(_d)), ("rdf_count", _P(_d)),
                ("rdf_count_random", _P(_d)), ("solute_group_count", _P(_d)), ("solvent_group_count", _P(_d)),
                ("solute_group_count_random", _P(_d)), ("solvent_group_count_random", _P(_d)),
                ("volume_total", _d),
                ("hits_md", _P(_u64)), ("hits_md_random", _P(_u64)), ("hits_rdf", _P(_u64)),
                ("hits_rdf_random", _P(_u64))]


class KernelStats(ctypes.Structure):
    _fields_ = [("eval_ms", _d), ("eval_launches", _i64), ("total_ms", _d), ("total_launches", _i64),
                ("frames", _i64), ("candidate_pairs", _i64), ("fp64_rechecks", _i64)]


class MDRecord(ctypes.Structure):
    """cm_md_record: MinimumDistance, src/minimum_distances.jl:12-19 (0-based i, j)."""
    _fields_ = [("within_cutoff", _i32), ("i", _i32), ("j", _i64), ("d", _d),
                ("ref_atom_within_cutoff", _i32), ("pad", _i32), ("d_ref_atom", _d)]


class Frames(ctypes.Structure):
    """cm_frames: one job's frames for cm_jobs_submit_device."""
    _fields_ = [("nframes", _i32), ("dtype", _i32), ("frame_index", _P(_i64)), ("d_x_solute", ctypes.c_void_p),
                ("d_x_solvent", ctypes.c_void_p), ("cells", _P(_d)), ("weights", _P(_d))]


class ServerDesc(ctypes.Structure):
    _fields_ = [("slots", _i32), ("pad", _i32), ("memory_budget", _i64)]


class Request(ctypes.Structure):
    _fields_ = [("job", _P(JobDesc)), ("nframes", _i32), ("dtype", _i32), ("frame_index", _P(_i64)),
                ("d_x_solute", ctypes.c_void_p), ("d_x_solvent", ctypes.c_void_p), ("cells", _P(_d)),
                ("weights", _P(_d)), ("cell_fraction", _d), ("columns", _i32), ("pad", _i32)]


class RequestTimes(ctypes.Structure):
    _fields_ = [("enqueued", _d), ("started", _d), ("finished", _d)]


EXPORTS = ["cm_job_create", "cm_job_submit", "cm_job_submit_device", "cm_job_finish", "cm_job_set_profiling",
           "cm_job_set_tuning", "cm_job_kernel_stats", "cm_job_launch_times", "cm_job_set_verify", "cm_job_records",
           "cm_job_placement", "cm_job_reference_draws", "cm_job_destroy", "cm_last_error", "cm_device_alloc", "cm_device_upload",
           "cm_device_free", "cm_device_synchronize", "cm_job_footprint", "cm_server_create", "cm_server_enqueue",
           "cm_server_wait", "cm_server_destroy", "cm_release_cached_memory", "cm_device_download",
           "cm_mddf_noalloc", "cm_jobs_submit_device", "cm_multi_set_graphs", "cm_multi_stats",
           "cm_set_readback_pack_limit"]

_lib = None


def release_cached_memory():
    """Returns the device and page-locked host blocks that finished jobs left
    cached for later jobs (cm_release_cached_memory)."""
    check(lib().cm_release_cached_memory())


def set_readback_pack_limit(limit=0xffffffff):
    """Largest hit count cm_job_finish reads back as a 32-bit word; a lower
    limit sends jobs down the FP64 read-back (cm_set_readback_pack_limit)."""
    check(lib().cm_set_readback_pack_limit(int(limit)))


class GPUError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GPUError(f"{LIB_PATH} is not built: run __graft_entry__.build() (the HIP library is the only "
                           "compute path)")
        L = ctypes.CDLL(LIB_PATH)
        L.cm_job_create.argtypes = [_P(JobDesc), _P(ctypes.c_void_p)]
        L.cm_job_submit.argtypes = [ctypes.c_void_p, _i32, _P(_i64), ctypes.c_void_p, ctypes.c_void_p, _i32, _P(_d),
                                    _P(_d)]
        L.cm_job_submit_device.argtypes = L.cm_job_submit.argtypes
        L.cm_job_finish.argtypes = [ctypes.c_void_p, _P(RawOut)]
        L.cm_job_set_profiling.argtypes = [ctypes.c_void_p, _i32]
        L.cm_job_set_tuning.argtypes = [ctypes.c_void_p, ctypes.c_double, _i32, _i32]
        L.cm_job_kernel_stats.argtypes = [ctypes.c_void_p, _P(KernelStats)]
        L.cm_job_launch_times.argtypes = [ctypes.c_void_p, _i32, _P(_d), _P(_d), _P(_i32)]
        L.cm_job_set_verify.argtypes = [ctypes.c_void_p, _i32]
        L.cm_job_records.argtypes = [ctypes.c_void_p, _i64, _i32, _P(MDRecord)]
        L.cm_job_placement.argtypes = [ctypes.c_void_p, _i64, _i32, _P(_d), _P(ctypes.c_uint8)]
        L.cm_job_reference_draws.argtypes = [ctypes.c_void_p, _i64, _P(_i32)]
        L.cm_job_destroy.argtypes = [ctypes.c_void_p]
        L.cm_job_destroy.restype = None
        L.cm_last_error.restype = ctypes.c_char_p
        L.cm_device_alloc.argtypes = [ctypes.c_size_t, _P(ctypes.c_void_p)]
        L.cm_device_upload.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.cm_device_free.argtypes = [ctypes.c_void_p]
        L.cm_job_footprint.argtypes = [_P(JobDesc), _i32, _i32, _P(_d), _P(_i64)]
        L.cm_server_create.argtypes = [_P(ServerDesc), _P(ctypes.c_void_p)]
        L.cm_server_enqueue.argtypes = [ctypes.c_void_p, _P(Request), _P(_u64)]
        L.cm_server_wait.argtypes = [ctypes.c_void_p, _u64, _P(RawOut), _P(RequestTimes)]
        L.cm_server_destroy.argtypes = [ctypes.c_void_p]
        L.cm_server_destroy.restype = None
        L.cm_release_cached_memory.argtypes = []
        L.cm_set_readback_pack_limit.argtypes = [_u64]
        L.cm_device_download.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.cm_jobs_submit_device.argtypes = [_i32, _P(ctypes.c_void_p), _P(Frames)]
        L.cm_multi_set_graphs.argtypes = [_i32]
        L.cm_multi_stats.argtypes = [_P(_i64), _P(_i64)]
        L.cm_mddf_noalloc.argtypes = [_P(JobDesc), _i32, _P(_i64), ctypes.c_void_p, ctypes.c_void_p, _i32, _P(_d),
                                      _P(_d), ctypes.c_void_p, _P(_d), ctypes.c_void_p]
        for f in ("cm_job_create", "cm_job_submit", "cm_job_submit_device", "cm_job_finish", "cm_job_set_profiling",
                  "cm_job_set_tuning", "cm_job_kernel_stats", "cm_job_launch_times", "cm_job_set_verify",
                  "cm_job_records", "cm_job_placement", "cm_job_reference_draws", "cm_device_alloc", "cm_device_upload", "cm_device_free",
                  "cm_device_synchronize", "cm_job_footprint", "cm_server_create", "cm_server_enqueue",
                  "cm_server_wait", "cm_device_download", "cm_mddf_noalloc", "cm_jobs_submit_device",
                  "cm_multi_set_graphs", "cm_multi_stats", "cm_set_readback_pack_limit"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def check(code):
    if code == CM_OK:
        return
    msg = lib().cm_last_error().decode()
    if code == CM_EINVAL:
        raise ArgumentError(msg)
    raise GPUError(f"libcmgpu error {code}: {msg}")


def _dp(a):
    return a.ctypes.data_as(_P(_d))


class DeviceArray:
    """A host array copied once into device memory (HBM-resident input)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype = host.dtype
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        check(lib().cm_device_alloc(self.nbytes, ctypes.byref(self.ptr)))
        check(lib().cm_device_upload(self.ptr, host.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        return cls(np.zeros(shape, dtype=dtype))

    def offset(self, nbytes):
        return ctypes.c_void_p(self.ptr.value + nbytes)

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().cm_device_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            check(lib().cm_device_free(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def job_desc(*, nmols_u, natoms_per_mol_u, nmols_v, natoms_per_mol_v, autocorrelation, irefatom, nbins, binstep,
             cutoff, dbulk, usecutoff, n_random_samples, coordination_number_only, triclinic, seed, n_groups_u,
             n_groups_v, groups_u=(None, None), groups_v=(None, None), lcell=1, max_batch_frames=0, memory_budget=0):
    """JobDesc for the C ABI; returns (desc, arrays the desc points into)."""
    keep = [g for g in (*groups_u, *groups_v) if g is not None]
    nul = _P(_i32)()

    def ip(a):
        return a.ctypes.data_as(_P(_i32)) if a is not None else nul

    d = JobDesc(int(nmols_u), int(natoms_per_mol_u), int(nmols_v), int(natoms_per_mol_v), int(bool(autocorrelation)),
                int(irefatom), int(nbins), int(bool(usecutoff)), float(binstep), float(cutoff), float(dbulk),
                int(n_random_samples), int(bool(coordination_number_only)), int(bool(triclinic)), int(lcell),
                int(seed) & 0xFFFFFFFFFFFFFFFF, int(n_groups_u), int(n_groups_v),
                ip(groups_u[0]), ip(groups_u[1]), ip(groups_v[0]), ip(groups_v[1]),
                int(max_batch_frames), int(memory_budget))
    return d, keep


def _cells9(cells):
    return np.ascontiguousarray(np.transpose(np.asarray(cells, dtype=np.float64), (0, 2, 1)).reshape(-1, 9))


def _raw_buffers(nbins, ngu, ngv):
    out = {k: np.zeros(nbins) for k in ("md_count", "md_count_random", "rdf_count", "rdf_count_random")}
    for k, g in (("solute_group_count", ngu), ("solvent_group_count", ngv),
                 ("solute_group_count_random", ngu), ("solvent_group_count_random", ngv)):
        out[k] = np.zeros((g, nbins))
    hits = {k: np.zeros(nbins, dtype=np.uint64) for k in ("hits_md", "hits_md_random", "hits_rdf", "hits_rdf_random")}
    r = RawOut(*[_dp(out[k]) for k in ("md_count", "md_count_random", "rdf_count", "rdf_count_random",
                                       "solute_group_count", "solvent_group_count", "solute_group_count_random",
                                       "solvent_group_count_random")], 0.0,
               *[hits[k].ctypes.data_as(_P(_u64)) for k in ("hits_md", "hits_md_random", "hits_rdf", "hits_rdf_random")])
    return out, hits, r


def footprint(frame_cells, dtype=np.float32, **job_kw):
    """Device bytes a job holds while it runs these frames (cm_job_footprint)."""
    d, _keep = job_desc(**job_kw)
    c9 = _cells9(frame_cells)
    b = _i64()
    check(lib().cm_job_footprint(ctypes.byref(d), len(c9), CM_F32 if np.dtype(dtype) == np.float32 else CM_F64,
                                 _dp(c9), ctypes.byref(b)))
    return b.value


class Job:
    """One accumulation job on the device (cm_job_*)."""

    def __init__(self, *, tuning=None, **job_kw):
        d, self._keep = job_desc(**job_kw)
        self.nbins, self.ngu, self.ngv = int(job_kw["nbins"]), int(job_kw["n_groups_u"]), int(job_kw["n_groups_v"])
        self.h = ctypes.c_void_p()
        check(lib().cm_job_create(ctypes.byref(d), ctypes.byref(self.h)))
        if tuning:
            self.set_tuning(**tuning)

    def submit(self, frame_index, x_solute, x_solvent, cells, weights):
        """Host arrays: frame_index [F] (1-based), x_* [F][N][3] float64/float32,
        cells [F][3][3] (lattice vectors as columns), weights [F]."""
        xu = np.ascontiguousarray(x_solute)
        xv = np.ascontiguousarray(x_solvent)
        if xu.dtype != xv.dtype or xu.dtype not in (np.float64, np.float32):
            raise ArgumentError("coordinates must be both float64 or both float32")
        dt = CM_F32 if xu.dtype == np.float32 else CM_F64
        fi = np.ascontiguousarray(frame_index, dtype=np.int64)
        c9 = _cells9(cells)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        check(lib().cm_job_submit(self.h, len(fi), fi.ctypes.data_as(_P(_i64)), xu.ctypes.data_as(ctypes.c_void_p),
                                  xv.ctypes.data_as(ctypes.c_void_p), dt, _dp(c9), _dp(w)))

    def submit_device(self, frame_index, d_solute, d_solvent, frame0, cells, weights):
        """Device-resident coordinates (DeviceArray of [F][N][3]); frames
        frame0 .. frame0+len(frame_index)-1 of those arrays.  The job keeps
        references to both arrays until it is closed: its kernels read them
        after this call returns."""
        self._inputs = getattr(self, "_inputs", []) + [d_solute, d_solvent]
        dt = CM_F32 if d_solute.dtype == np.float32 else CM_F64
        fi = np.ascontiguousarray(frame_index, dtype=np.int64)
        c9 = _cells9(cells)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        fu = d_solute.nbytes // d_solute.shape[0]
        fv = d_solvent.nbytes // d_solvent.shape[0]
        check(lib().cm_job_submit_device(self.h, len(fi), fi.ctypes.data_as(_P(_i64)), d_solute.offset(frame0 * fu),
                                         d_solvent.offset(frame0 * fv), dt, _dp(c9), _dp(w)))

    def finish(self):
        out, hits, r = _raw_buffers(self.nbins, self.ngu, self.ngv)
        check(lib().cm_job_finish(self.h, ctypes.byref(r)))
        out["volume_total"] = r.volume_total
        out.update(hits)
        return out

    def set_profiling(self, on=True):
        check(lib().cm_job_set_profiling(self.h, int(on)))

    def set_tuning(self, cell_fraction=0.5, columns=0, eval_workgroups=0):
        check(lib().cm_job_set_tuning(self.h, float(cell_fraction), int(columns), int(eval_workgroups)))

    def stats(self):
        s = KernelStats()
        check(lib().cm_job_kernel_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in KernelStats._fields_}

    def set_verify(self, on=True):
        """Keep records, placements and reference draws of every frame (cm_job_set_verify)."""
        check(lib().cm_job_set_verify(self.h, int(on)))
        self.nmols_v = None

    def records(self, frame, pass_, nmols_v):
        """MinimumDistance records of one pass (0 real, 1..K random) of a frame:
        list of (within, i, j, d, ref, d_ref)."""
        out = (MDRecord * nmols_v)()
        check(lib().cm_job_records(self.h, int(frame), int(pass_), out))
        return [(m.within_cutoff, m.i, m.j, m.d, m.ref_atom_within_cutoff, m.d_ref_atom) for m in out]

    def placement(self, frame, pass_, nmols_v, natoms_per_mol_v):
        """(coordinates [nmols_v * natoms][3], placed flags [nmols_v]) of random pass pass_ of a frame."""
        out = np.zeros((nmols_v * natoms_per_mol_v, 3))
        fl = np.zeros(nmols_v, dtype=np.uint8)
        check(lib().cm_job_placement(self.h, int(frame), int(pass_), _dp(out), fl.ctypes.data_as(_P(ctypes.c_uint8))))
        return out, fl.astype(bool)

    def reference_draws(self, frame, n_random_samples):
        out = np.zeros(max(n_random_samples, 1), dtype=np.int32)
        check(lib().cm_job_reference_draws(self.h, int(frame), out.ctypes.data_as(_P(_i32))))
        return out[:n_random_samples]

    def launch_times(self):
        """Per profiled batch: (real-pass, random-pass) evaluation launch
        durations in ms (-1 without random passes)."""
        n = _i32()
        check(lib().cm_job_launch_times(self.h, 0, None, None, ctypes.byref(n)))
        real = np.zeros(max(n.value, 1))
        rnd = np.zeros(max(n.value, 1))
        check(lib().cm_job_launch_times(self.h, n.value, _dp(real), _dp(rnd), ctypes.byref(n)))
        return real[:n.value], rnd[:n.value]

    def close(self):
        if self.h:
            lib().cm_job_destroy(self.h)
            self.h = ctypes.c_void_p()
        self._inputs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server (cm_server_*): continuous batching of jobs over
    device-resident frames.  enqueue() returns a ticket; wait(ticket) blocks
    until that request completes and returns (raw, times)."""

    def __init__(self, slots=4, memory_budget=0):
        self.h = ctypes.c_void_p()
        check(lib().cm_server_create(ctypes.byref(ServerDesc(int(slots), 0, int(memory_budget))), ctypes.byref(self.h)))
        self._meta = {}

    def enqueue(self, frame_index, d_solute, d_solvent, frame0, cells, weights, tuning=None, **job_kw):
        d, keep = job_desc(**job_kw)
        dt = CM_F32 if d_solute.dtype == np.float32 else CM_F64
        fi = np.ascontiguousarray(frame_index, dtype=np.int64)
        c9 = _cells9(cells)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        fu = d_solute.nbytes // d_solute.shape[0]
        fv = d_solvent.nbytes // d_solvent.shape[0]
        t = tuning or {}
        req = Request(ctypes.pointer(d), len(fi), dt, fi.ctypes.data_as(_P(_i64)), d_solute.offset(frame0 * fu),
                      d_solvent.offset(frame0 * fv), _dp(c9), _dp(w), float(t.get("cell_fraction", 0.0)),
                      int(t.get("columns", 0)), 0)
        ticket = _u64()
        check(lib().cm_server_enqueue(self.h, ctypes.byref(req), ctypes.byref(ticket)))
        # the request's kernels read both device arrays whenever a slot takes
        # it: they stay referenced until wait() has returned its counts (or
        # the server is closed)
        self._meta[ticket.value] = (int(job_kw["nbins"]), int(job_kw["n_groups_u"]), int(job_kw["n_groups_v"]),
                                    d_solute, d_solvent)
        return ticket.value

    def wait(self, ticket):
        nb, ngu, ngv, _du, _dv = self._meta[ticket]
        out, hits, r = _raw_buffers(nb, ngu, ngv)
        tm = RequestTimes()
        try:
            check(lib().cm_server_wait(self.h, _u64(ticket), ctypes.byref(r), ctypes.byref(tm)))
        finally:
            # the request is complete (or refused) once cm_server_wait returns
            self._meta.pop(ticket, None)
        out["volume_total"] = r.volume_total
        out.update(hits)
        return out, {"enqueued": tm.enqueued, "started": tm.started, "finished": tm.finished}

    def close(self):
        if self.h:
            lib().cm_server_destroy(self.h)
            self.h = ctypes.c_void_p()
        self._meta = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def acc_len(nbins, n_groups_u, n_groups_v):
    """Doubles of the FP64 accumulator block of cm_mddf_noalloc."""
    return 4 * nbins + 2 * (n_groups_u + n_groups_v) * nbins


def mddf_noalloc(frame_index, d_solute, d_solvent, frame0, cells, weights, d_acc, stream=None, **job_kw):
    """cm_mddf_noalloc: one accumulation over device-resident frames whose
    device buffers all come from the library's fixed scratch; d_acc is a
    DeviceArray of acc_len(...) doubles (the output).  Returns (raw dict read
    back from d_acc, volume_total)."""
    d, _keep = job_desc(**job_kw)
    nb, ngu, ngv = int(job_kw["nbins"]), int(job_kw["n_groups_u"]), int(job_kw["n_groups_v"])
    if d_acc.dtype != np.float64 or d_acc.nbytes < 8 * acc_len(nb, ngu, ngv):
        raise ArgumentError("d_acc must hold acc_len(nbins, n_groups_u, n_groups_v) float64 values")
    dt = CM_F32 if d_solute.dtype == np.float32 else CM_F64
    fi = np.ascontiguousarray(frame_index, dtype=np.int64)
    c9 = _cells9(cells)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    fu = d_solute.nbytes // d_solute.shape[0]
    fv = d_solvent.nbytes // d_solvent.shape[0]
    vol = _d()
    check(lib().cm_mddf_noalloc(ctypes.byref(d), len(fi), fi.ctypes.data_as(_P(_i64)), d_solute.offset(frame0 * fu),
                                d_solvent.offset(frame0 * fv), dt, _dp(c9), _dp(w), d_acc.ptr, ctypes.byref(vol),
                                stream))
    a = d_acc.download().reshape(-1)[:acc_len(nb, ngu, ngv)]
    out, o = {}, 0
    for k, n in (("md_count", nb), ("md_count_random", nb), ("rdf_count", nb), ("rdf_count_random", nb),
                 ("solute_group_count", ngu * nb), ("solvent_group_count", ngv * nb),
                 ("solute_group_count_random", ngu * nb), ("solvent_group_count_random", ngv * nb)):
        out[k] = a[o:o + n].reshape(-1, nb) if "group" in k else a[o:o + n].copy()
        o += n
    out["volume_total"] = vol.value
    return out


def submit_many(items):
    """cm_jobs_submit_device: items are (job, frame_index, d_solute, d_solvent,
    frame0, cells, weights) as for Job.submit_device; the frames of all jobs
    run together (one launch per stage per round of batches).  Each job keeps
    references to its device arrays until it is closed."""
    n = len(items)
    handles = (ctypes.c_void_p * max(n, 1))()
    fr = (Frames * max(n, 1))()
    keep = []
    for k, (job, frame_index, d_solute, d_solvent, frame0, cells, weights) in enumerate(items):
        job._inputs = getattr(job, "_inputs", []) + [d_solute, d_solvent]
        fi = np.ascontiguousarray(frame_index, dtype=np.int64)
        c9 = _cells9(cells)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        keep += [fi, c9, w]
        fu = d_solute.nbytes // d_solute.shape[0]
        fv = d_solvent.nbytes // d_solvent.shape[0]
        handles[k] = job.h
        fr[k] = Frames(len(fi), CM_F32 if d_solute.dtype == np.float32 else CM_F64, fi.ctypes.data_as(_P(_i64)),
                       d_solute.offset(frame0 * fu), d_solvent.offset(frame0 * fv), _dp(c9), _dp(w))
    check(lib().cm_jobs_submit_device(n, handles, fr))


def set_multi_graphs(on):
    """Rounds of submit_many replay a captured HIP graph (on) or launch their stages directly."""
    check(lib().cm_multi_set_graphs(int(bool(on))))


def multi_stats():
    r, l = _i64(), _i64()
    check(lib().cm_multi_stats(ctypes.byref(r), ctypes.byref(l)))
    return {"rounds": r.value, "launches": l.value}
