"""The entry point that allocates nothing (cm_mddf_noalloc) against the job
API (cm_job_create + cm_job_submit_device + cm_job_finish) on the same
inputs, through the C ABI: two calls in a row, the second with a smaller
input, then a self-MDDF call; and a job too large for the fixed scratch is
refused before any device work."""
import numpy as np
import pytest

from conftest import RAW_KEYS, assert_raw_equal

pytestmark = pytest.mark.gpu


def _kw(s, pkg, opts, auto=False):
    from complexmixtures_amd.mddf import TrajectoryMetaData, job_kwargs
    from complexmixtures_amd.results import make_result
    traj = pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])
    tm = TrajectoryMetaData(traj, opts)
    R = make_result(traj, opts, tm)
    return job_kwargs(R, opts, tm, False)


def _both(pkg, s, opts, dtype):
    from complexmixtures_amd import _lib
    kw = _kw(s, pkg, opts)
    xu = s["x_solute"].astype(dtype)
    xv = s["x_solvent"].astype(dtype)
    nf = xu.shape[0]
    du = _lib.DeviceArray(xu)
    dv = du if s["x_solvent"] is s["x_solute"] else _lib.DeviceArray(xv)
    fi = np.arange(1, nf + 1)
    w = np.ones(nf)
    job = _lib.Job(**kw)
    job.submit_device(fi, du, dv, 0, s["cells"], w)
    ref = job.finish()
    job.close()
    n = _lib.acc_len(kw["nbins"], kw["n_groups_u"], kw["n_groups_v"])
    d_acc = _lib.DeviceArray(np.full(n, np.nan))     # the call must overwrite every value
    got = _lib.mddf_noalloc(fi, du, dv, 0, s["cells"], w, d_acc, **kw)
    return ref, got


def test_no_alloc_equals_job_api(pkg):
    import complexmixtures_amd.synthetic as S
    # 1: 40 A box, 1,000-atom chain, 2,000 waters, 6 frames, FP64, 10 random passes
    s1 = S.config1(nframes=6)
    o1 = pkg.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261203, silent=True)
    ref, got = _both(pkg, s1, o1, np.float64)
    assert ref["md_count_random"].sum() > 0
    assert_raw_equal(ref, got, keys=RAW_KEYS)
    # 2: smaller, FP32
    s2 = S.config1(nframes=2, n_solute=300, n_water=900, box=30.0, seed=5)
    o2 = pkg.Options(dbulk=8.0, n_random_samples=4, seed=77, silent=True)
    ref, got = _both(pkg, s2, o2, np.float32)
    assert_raw_equal(ref, got, keys=RAW_KEYS)
    # 3: self-MDDF of waters (molecule-pair kernels)
    rng = np.random.default_rng(44)
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, 30.0, (800, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    sel = pkg.AtomSelection(list(range(1, 2401)), natomspermol=3)
    s3 = {"x_solute": xv, "x_solvent": xv, "cells": np.repeat(np.diag([30.0] * 3)[None], 2, axis=0),
          "solute": sel, "solvent": sel}
    ref, got = _both(pkg, s3, pkg.Options(dbulk=9.0, n_random_samples=3, seed=52, silent=True), np.float64)
    assert_raw_equal(ref, got, keys=RAW_KEYS)


def test_no_alloc_refuses_jobs_beyond_the_scratch(pkg):
    """6,000,000 solvent atoms with 10 random passes need about 2.4 GB of batch
    buffers per frame: refused on the host, before any device work (the
    one-atom coordinate arrays passed are never read)."""
    from complexmixtures_amd import _lib
    from complexmixtures_amd.results import setbin
    du = _lib.DeviceArray(np.zeros((1, 1, 3), dtype=np.float32))
    kw = dict(nmols_u=1, natoms_per_mol_u=1, nmols_v=2_000_000, natoms_per_mol_v=3, autocorrelation=False,
              irefatom=1, nbins=setbin(10.0, 0.02), binstep=0.02, cutoff=10.0, dbulk=10.0, usecutoff=False,
              n_random_samples=10, coordination_number_only=False, triclinic=False, seed=3, n_groups_u=1,
              n_groups_v=3)
    d_acc = _lib.DeviceArray(np.zeros(_lib.acc_len(kw["nbins"], 1, 3)))
    with pytest.raises(_lib.GPUError, match="too large for the no-allocation entry point"):
        _lib.mddf_noalloc([1], du, du, 0, np.diag([400.0] * 3)[None], [1.0], d_acc, **kw)
