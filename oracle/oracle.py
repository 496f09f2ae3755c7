"""ctypes wrapper of the CPU FP64 oracle (cm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product never imports
this module.

The oracle restates the reference's per-frame algorithm (see the header of
cm_oracle.c for the file:line list).  This wrapper only marshals arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.POINTER
_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


class MD(ctypes.Structure):
    """MinimumDistance record, src/minimum_distances.jl:12-19 (0-based i, j)."""
    _fields_ = [("within_cutoff", _i32), ("i", _i32), ("j", _i64), ("d", _d),
                ("ref_atom_within_cutoff", _i32), ("pad", _i32), ("d_ref_atom", _d)]


class Desc(ctypes.Structure):
    _fields_ = [("nmols_u", _i32), ("nu", _i32), ("nmols_v", _i32), ("nv", _i32),
                ("autocorrelation", _i32), ("irefatom", _i32), ("nbins", _i32), ("usecutoff", _i32),
                ("binstep", _d), ("cutoff", _d), ("dbulk", _d),
                ("n_random_samples", _i32), ("coordination_only", _i32), ("triclinic", _i32),
                ("use_cells", _i32), ("seed", _u64), ("ngroups_u", _i32), ("ngroups_v", _i32),
                ("gu_ptr", _P(_i32)), ("gu_idx", _P(_i32)), ("gv_ptr", _P(_i32)), ("gv_idx", _P(_i32)),
                ("nthreads", _i32), ("pad", _i32)]


class Raw(ctypes.Structure):
    _fields_ = [("md_count", _P(_d)), ("md_count_random", _P(_d)), ("rdf_count", _P(_d)),
                ("rdf_count_random", _P(_d)), ("gu", _P(_d)), ("gv", _P(_d)), ("gu_random", _P(_d)),
                ("gv_random", _P(_d)), ("volume_total", _d)]


class Final(ctypes.Structure):
    _fields_ = [(n, _P(_d)) for n in (
        "d", "md_count", "md_count_random", "coordination_number", "coordination_number_random",
        "mddf", "kb", "gu", "gv", "gu_random", "gv_random", "rdf_count", "rdf_count_random",
        "sum_rdf_count", "sum_rdf_count_random", "rdf", "kb_rdf")] + [
        ("density_solute", _d), ("density_solvent", _d), ("density_solvent_bulk", _d),
        ("volume_total", _d), ("volume_bulk", _d), ("volume_domain", _d), ("volume_shell", _P(_d))]


def _declare(L):
    L.orc_philox4x32_10.argtypes = [_P(_u32), _P(_u32), _P(_u32)]
    L.orc_sincos2pi.argtypes = [_d, _P(_d), _P(_d)]
    L.orc_eulermat.argtypes = [_d, _d, _d, _P(_d)]
    L.orc_eulermat_turns.argtypes = [_d, _d, _d, _P(_d)]
    L.orc_move.argtypes = [_P(_d), ctypes.c_int, _P(_d), _d, _d, _d]
    L.orc_random_move.argtypes = [_P(_d), ctypes.c_int, ctypes.c_int, _P(_d), ctypes.c_int, _P(_d), _P(_d)]
    L.orc_canon_d2.argtypes = [_P(_d), ctypes.c_int, _P(_d), _P(_d)]
    L.orc_canon_d2.restype = _d
    L.orc_setbin.argtypes = [_d, _d]
    L.orc_setbin.restype = _i64
    L.orc_sphericalshellvolume.argtypes = [_i64, _d]
    L.orc_sphericalshellvolume.restype = _d
    L.orc_shellradius.argtypes = [_i64, _d]
    L.orc_shellradius.restype = _d
    L.orc_update_md.argtypes = [MD, MD]
    L.orc_update_md.restype = MD
    L.orc_run.argtypes = [_P(Desc), ctypes.c_int, _P(_i64), _P(_d), _P(_d), _P(_d), _P(_d), _P(Raw)]
    L.orc_run.restype = ctypes.c_int
    L.orc_records.argtypes = [_P(Desc), _P(_d), _P(_d), _P(_d), ctypes.c_int, _P(MD)]
    L.orc_records.restype = ctypes.c_int
    L.orc_placement.argtypes = [_P(Desc), _P(_d), _P(_d), _i64, _P(_i64), _u32, _u32, _P(_d)]
    L.orc_ref_solute.argtypes = [_u64, _u32, _u32, _i64]
    L.orc_ref_solute.restype = _i64
    L.orc_finalize.argtypes = [_P(Desc), _P(Raw), _d, _P(Final)]


def _ptr(a, ct=_d):
    return a.ctypes.data_as(_P(ct))


# ---------------------------------------------------------------------------
# small functions
# ---------------------------------------------------------------------------
def philox(ctr, key):
    c = (_u32 * 4)(*ctr)
    k = (_u32 * 2)(*key)
    o = (_u32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def sincos2pi(u):
    s, c = _d(), _d()
    lib().orc_sincos2pi(u, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def eulermat(beta, gamma, theta):
    A = np.zeros(9)
    lib().orc_eulermat(beta, gamma, theta, _ptr(A))
    return A.reshape(3, 3)


def eulermat_turns(ub, ug, ut):
    A = np.zeros(9)
    lib().orc_eulermat_turns(ub, ug, ut, _ptr(A))
    return A.reshape(3, 3)


def move(x, newcm, beta, gamma, theta):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    cm = np.ascontiguousarray(newcm, dtype=np.float64)
    lib().orc_move(_ptr(x), x.shape[0], _ptr(cm), beta, gamma, theta)
    return x


def random_move(x, iref, cell, triclinic, uc, ua):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    cell9 = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).T.reshape(9))
    uc = np.ascontiguousarray(uc, dtype=np.float64)
    ua = np.ascontiguousarray(ua, dtype=np.float64)
    lib().orc_random_move(_ptr(x), x.shape[0], iref, _ptr(cell9), int(triclinic), _ptr(uc), _ptr(ua))
    return x


def canon_d2(cell, triclinic, x, y):
    cell9 = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).T.reshape(9))
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_canon_d2(_ptr(cell9), int(triclinic), _ptr(x), _ptr(y))


def setbin(d, step):
    return lib().orc_setbin(d, step)


def shellradius(i, step):
    return lib().orc_shellradius(i, step)


def sphericalshellvolume(i, step):
    return lib().orc_sphericalshellvolume(i, step)


def update_md(m1, m2):
    return lib().orc_update_md(MD(*m1), MD(*m2))


def ref_solute(seed, frame, sample, nmols_u):
    return lib().orc_ref_solute(seed, frame, sample, nmols_u)


# ---------------------------------------------------------------------------
# job-level entry points
# ---------------------------------------------------------------------------
class Job:
    """Oracle job description.  Field meanings follow cm_job_desc (include/cmgpu.h).

    groups_u / groups_v: None (atom-type groups) or a list of lists of group
    ids: groups_u[i] for the molecule-local solute atom i, groups_v[j] for the
    global solvent atom j.  In autocorrelation mode groups_v maps into the
    solute groups.
    """

    def __init__(self, nmols_u, nu, nmols_v, nv, *, autocorrelation=False, irefatom=0, binstep=0.02,
                 cutoff=10.0, dbulk=10.0, usecutoff=False, n_random_samples=10, coordination_only=False,
                 triclinic=False, seed=321, ngroups_u=None, ngroups_v=None, groups_u=None, groups_v=None,
                 use_cells=True, nthreads=1):
        self.nmols_u, self.nu, self.nmols_v, self.nv = int(nmols_u), int(nu), int(nmols_v), int(nv)
        self.autocorrelation = bool(autocorrelation)
        self.irefatom = int(irefatom)
        self.binstep, self.cutoff, self.dbulk = float(binstep), float(cutoff), float(dbulk)
        self.usecutoff = bool(usecutoff)
        self.nbins = int(setbin(self.cutoff, self.binstep))
        self.n_random_samples = int(n_random_samples)
        self.coordination_only = bool(coordination_only)
        self.triclinic = bool(triclinic)
        self.seed = int(seed)
        self.ngroups_u = int(ngroups_u if ngroups_u is not None else nu)
        self.ngroups_v = int(ngroups_v if ngroups_v is not None else (nu if autocorrelation else nv))
        self.groups_u, self.groups_v = groups_u, groups_v
        self.use_cells = bool(use_cells)
        self.nthreads = int(nthreads)

    @staticmethod
    def _csr(groups):
        if groups is None:
            return None, None
        ptr = np.zeros(len(groups) + 1, dtype=np.int32)
        for k, g in enumerate(groups):
            ptr[k + 1] = ptr[k] + len(g)
        idx = np.array([x for g in groups for x in g] or [0], dtype=np.int32)
        return ptr, idx

    def cdesc(self):
        gu_ptr, gu_idx = self._csr(self.groups_u)
        gv_ptr, gv_idx = self._csr(self.groups_v)
        keep = [gu_ptr, gu_idx, gv_ptr, gv_idx]
        nul = _P(_i32)()
        d = Desc(self.nmols_u, self.nu, self.nmols_v, self.nv, int(self.autocorrelation), self.irefatom,
                 self.nbins, int(self.usecutoff), self.binstep, self.cutoff, self.dbulk,
                 self.n_random_samples, int(self.coordination_only), int(self.triclinic), int(self.use_cells),
                 self.seed & 0xFFFFFFFFFFFFFFFF, self.ngroups_u, self.ngroups_v,
                 _ptr(gu_ptr, _i32) if gu_ptr is not None else nul,
                 _ptr(gu_idx, _i32) if gu_idx is not None else nul,
                 _ptr(gv_ptr, _i32) if gv_ptr is not None else nul,
                 _ptr(gv_idx, _i32) if gv_idx is not None else nul,
                 self.nthreads, 0)
        return d, keep


def _cells9(cells):
    cells = np.asarray(cells, dtype=np.float64)
    # [f][3][3] with lattice vectors as columns -> column-major 9 values per frame
    return np.ascontiguousarray(np.transpose(cells, (0, 2, 1)).reshape(cells.shape[0], 9))


def run(job, frame_index, xu, xv, cells, weights):
    """Raw accumulators of the frames given (FP64 sequential sums of weights)."""
    xu = np.ascontiguousarray(xu, dtype=np.float64)
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    nf = xu.shape[0]
    fi = np.ascontiguousarray(frame_index, dtype=np.int64)
    c9 = _cells9(cells)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    nb = job.nbins
    out = {k: np.zeros(nb) for k in ("md_count", "md_count_random", "rdf_count", "rdf_count_random")}
    out["gu"] = np.zeros((job.ngroups_u, nb))
    out["gv"] = np.zeros((job.ngroups_v, nb))
    out["gu_random"] = np.zeros((job.ngroups_u, nb))
    out["gv_random"] = np.zeros((job.ngroups_v, nb))
    raw = Raw(_ptr(out["md_count"]), _ptr(out["md_count_random"]), _ptr(out["rdf_count"]),
              _ptr(out["rdf_count_random"]), _ptr(out["gu"]), _ptr(out["gv"]), _ptr(out["gu_random"]),
              _ptr(out["gv_random"]), 0.0)
    d, keep = job.cdesc()
    st = lib().orc_run(ctypes.byref(d), nf, _ptr(fi, _i64), _ptr(xu), _ptr(xv), _ptr(c9), _ptr(w),
                       ctypes.byref(raw))
    del keep
    if st != 0:
        raise RuntimeError("oracle run failed")
    out["volume_total"] = raw.volume_total
    return out


def records(job, xu, xv, cell, isolute):
    """MinimumDistance list of one real pass (one frame, one solute molecule)."""
    xu = np.ascontiguousarray(xu, dtype=np.float64)
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    c9 = _cells9(np.asarray(cell)[None])[0].copy()
    lst = (MD * max(job.nmols_v, 1))()
    d, keep = job.cdesc()
    st = lib().orc_records(ctypes.byref(d), _ptr(xu), _ptr(xv), _ptr(c9), int(isolute), lst)
    del keep
    if st != 0:
        raise RuntimeError("oracle records failed")
    return [(m.within_cutoff, m.i, m.j, m.d, m.ref_atom_within_cutoff, m.d_ref_atom) for m in lst[:job.nmols_v]]


def placement(job, xv, cell, bulk, frame, sample):
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    c9 = _cells9(np.asarray(cell)[None])[0].copy()
    b = np.ascontiguousarray(bulk, dtype=np.int64) if len(bulk) else np.zeros(1, dtype=np.int64)
    out = np.zeros((job.nmols_v * job.nv, 3))
    d, keep = job.cdesc()
    lib().orc_placement(ctypes.byref(d), _ptr(xv), _ptr(c9), len(bulk), _ptr(b, _i64), frame, sample, _ptr(out))
    del keep
    return out


def finalize(job, raw, Q):
    """finalresults! over raw accumulators; returns the Result fields."""
    nb = job.nbins
    r = Raw(_ptr(raw["md_count"]), _ptr(raw["md_count_random"]), _ptr(raw["rdf_count"]),
            _ptr(raw["rdf_count_random"]), _ptr(np.ascontiguousarray(raw["gu"])),
            _ptr(np.ascontiguousarray(raw["gv"])), _ptr(np.ascontiguousarray(raw["gu_random"])),
            _ptr(np.ascontiguousarray(raw["gv_random"])), raw["volume_total"])
    names = ["d", "md_count", "md_count_random", "coordination_number", "coordination_number_random", "mddf",
             "kb", "gu", "gv", "gu_random", "gv_random", "rdf_count", "rdf_count_random", "sum_rdf_count",
             "sum_rdf_count_random", "rdf", "kb_rdf"]
    arrs = {}
    for n in names:
        if n in ("gu", "gu_random"):
            arrs[n] = np.zeros((job.ngroups_u, nb))
        elif n in ("gv", "gv_random"):
            arrs[n] = np.zeros((job.ngroups_v, nb))
        else:
            arrs[n] = np.zeros(nb)
    arrs["volume_shell"] = np.zeros(nb)
    fin = Final(*[_ptr(arrs[n]) for n in names], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, _ptr(arrs["volume_shell"]))
    d, keep = job.cdesc()
    lib().orc_finalize(ctypes.byref(d), ctypes.byref(r), float(Q), ctypes.byref(fin))
    del keep
    for n in ("density_solute", "density_solvent", "density_solvent_bulk", "volume_total", "volume_bulk",
              "volume_domain"):
        arrs[n] = getattr(fin, n)
    return arrs
