"""Many jobs in one launch per stage (cm_jobs_submit_device): every job's raw
accumulators equal those of the same job submitted alone, bit for bit, with
the rounds replayed from a captured HIP graph and launched directly, over
jobs of different sizes, frame counts, weights, tunings and with and without
ideal-gas passes (a self-MDDF job in the same call goes its own way)."""
import numpy as np
import pytest

from conftest import RAW_KEYS, assert_raw_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool(pkg):
    import complexmixtures_amd.serve as SV
    return SV.make_pool(5, seed=31, nframes=14, n_solute_range=(200, 3000), box_range=(30.0, 60.0))


def _jobs(pool):
    import complexmixtures_amd.serve as SV
    out = []
    specs = [(0, 14, True, None, None), (1, 9, False, dict(cell_fraction=0.75, columns=4), None),
             (2, 14, True, dict(cell_fraction=0.35, columns=4), 3), (3, 5, True, None, 2),
             (4, 12, False, None, None), (1, 6, True, None, 4)]
    for k, (si, nf, mddf, tuning, bmax) in enumerate(specs):
        s = pool[si]
        r = SV.Request(0.0, si, nf, mddf, False, 1000 + k)
        kw, du, dv, _ = SV.job_kwargs(s, r)
        if bmax:
            kw["max_batch_frames"] = bmax
        w = np.ones(nf)
        if k == 2:
            w[5:9] = 2.0          # weight changes inside the job: flushes between rounds
        if k == 3:
            w[:] = 0.5
        out.append((kw, tuning, du, dv, s.cells[:nf], w, nf))
    # a self-MDDF job (many solute molecules): submitted on its own path
    s = pool[0]
    r = SV.Request(0.0, 0, 4, True, True, 77)
    kw, du, dv, _ = SV.job_kwargs(s, r)
    out.append((kw, None, du, dv, s.cells[:4], np.ones(4), 4))
    return out


def _alone(jobs):
    from complexmixtures_amd import _lib
    res = []
    for kw, tuning, du, dv, cells, w, nf in jobs:
        job = _lib.Job(tuning=tuning, **kw)
        job.submit_device(np.arange(1, nf + 1), du, dv, 0, cells, w)
        res.append(job.finish())
        job.close()
    return res


def _together(jobs):
    from complexmixtures_amd import _lib
    handles = [_lib.Job(tuning=tuning, **kw) for kw, tuning, *_ in jobs]
    _lib.submit_many([(h, np.arange(1, nf + 1), du, dv, 0, cells, w)
                      for h, (kw, tuning, du, dv, cells, w, nf) in zip(handles, jobs)])
    res = [h.finish() for h in handles]
    for h in handles:
        h.close()
    return res


@pytest.mark.parametrize("graphs", [True, False])
def test_many_jobs_equal_jobs_alone(pkg, pool, graphs):
    from complexmixtures_amd import _lib
    jobs = _jobs(pool)
    ref = _alone(jobs)
    _lib.set_multi_graphs(graphs)
    try:
        before = _lib.multi_stats()
        # several calls in a row (graph replays), with eager work between them
        for rep in range(3):
            got = _together(jobs)
            for k, (a, b) in enumerate(zip(ref, got)):
                assert_raw_equal(a, b, keys=RAW_KEYS)
                assert a["volume_total"] == b["volume_total"], k
            _alone(jobs[:1])
        after = _lib.multi_stats()
        # rounds: the largest number of batches of a job (14 frames in 3-frame batches: 5)
        assert after["rounds"] - before["rounds"] >= 3 * 5
    finally:
        _lib.set_multi_graphs(True)


def test_many_jobs_uniform_batches(pkg, pool):
    """Twelve jobs, one per (system, kind), all in one round per batch."""
    import complexmixtures_amd.serve as SV
    from complexmixtures_amd import _lib
    jobs = []
    for k in range(12):
        s = pool[k % 5]
        kw, du, dv, _ = SV.job_kwargs(s, SV.Request(0.0, k % 5, 14, k % 3 != 0, False, 500 + k))
        jobs.append((kw, None, du, dv, s.cells[:14], np.ones(14), 14))
    ref = _alone(jobs)
    got = _together(jobs)
    for a, b in zip(ref, got):
        assert_raw_equal(a, b, keys=RAW_KEYS)
