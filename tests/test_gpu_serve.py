"""Request server on the GPU: every request's counts are the same alone (the
plain one-job path) and in the mix (continuous batching), and match the
oracle."""
import numpy as np
import pytest

from conftest import ORACLE_KEY, RAW_KEYS, assert_raw_equal

pytestmark = pytest.mark.gpu


def _serve():
    import complexmixtures_amd.serve as S
    return S


@pytest.fixture(scope="module")
def pool(pkg):
    S = _serve()
    return S.make_pool(3, seed=7, nframes=12, n_solute_range=(200, 1500), box_range=(30.0, 42.0), keep_host=True)


def _requests(S, n, pool, seed, saturate=True):
    rng = np.random.default_rng(seed)
    return S.request_stream(rng, n, len(pool), pool[0].nframes, frames=(2, 12), p_mddf=0.5, p_self=0.3,
                            saturate=saturate)


def test_served_equals_plain(pkg, pool):
    S = _serve()
    reqs = _requests(S, 14, pool, 3)
    assert any(r.self_ for r in reqs) and any(r.mddf for r in reqs) and any(not r.self_ for r in reqs)
    served = S.run_served(pool, reqs, slots=3, keep_raw=True)
    plain = S.run_plain(pool, reqs, keep_raw=True)
    for a, b in zip(plain["raw"], served["raw"]):
        assert_raw_equal(a, b)
    assert served["digests"] == plain["digests"]


def test_served_fp64_readback_equals_packed(pkg, pool):
    """A request whose frames share one weight comes back as 32-bit hits;
    with a pack limit of 0 every request falls back to the FP64 accumulators.
    The counts are the same."""
    from complexmixtures_amd import _lib
    S = _serve()
    reqs = _requests(S, 6, pool, 5)
    packed = S.run_served(pool, reqs, slots=2, keep_raw=True)
    _lib.set_readback_pack_limit(0)
    try:
        wide = S.run_served(pool, reqs, slots=2, keep_raw=True)
    finally:
        _lib.set_readback_pack_limit()
    for a, b in zip(packed["raw"], wide["raw"]):
        assert_raw_equal(a, b)
    assert packed["digests"] == wide["digests"]


def test_served_matches_oracle(pkg, oracle, pool):
    """One cross and one self request of the mix against the FP64 oracle."""
    S = _serve()
    reqs = [S.Request(0.0, 0, 5, True, False, 11), S.Request(0.0, 1, 4, True, True, 12),
            S.Request(0.0, 2, 6, False, False, 13)]
    served = S.run_served(pool, reqs, slots=2, keep_raw=True)
    for r, raw in zip(reqs, served["raw"]):
        s = pool[r.system]
        kw, _, _, _ = S.job_kwargs(s, r)
        h = s.host
        xv = h["x_solvent"][:r.nframes].astype(np.float64)
        xu = xv if r.self_ else h["x_solute"][:r.nframes].astype(np.float64)
        job = oracle.Job(kw["nmols_u"], kw["natoms_per_mol_u"], kw["nmols_v"], kw["natoms_per_mol_v"],
                         autocorrelation=kw["autocorrelation"], irefatom=kw["irefatom"] - 1, cutoff=kw["cutoff"],
                         dbulk=kw["dbulk"], n_random_samples=kw["n_random_samples"],
                         coordination_only=kw["coordination_number_only"], seed=kw["seed"])
        ref = oracle.run(job, list(range(1, r.nframes + 1)), xu, xv, s.cells[:r.nframes], np.ones(r.nframes))
        ref = {k: ref[ORACLE_KEY.get(k, k)] for k in RAW_KEYS} | {"volume_total": ref["volume_total"]}
        assert_raw_equal(ref, raw)


def test_memory_budget_admits_one_at_a_time(pkg, pool):
    """A budget below every footprint: requests run alone, in order, with the
    same counts."""
    S = _serve()
    reqs = _requests(S, 6, pool, 5)
    served = S.run_served(pool, reqs, slots=4, memory_budget=1, keep_raw=False)
    plain = S.run_plain(pool, reqs)
    assert served["digests"] == plain["digests"]


def test_server_rejects_and_reports(pkg, pool):
    from complexmixtures_amd import _lib
    S = _serve()
    srv = _lib.Server(2)
    try:
        r = S.Request(0.0, 0, 3, True, False, 1)
        kw, du, dv, _ = S.job_kwargs(pool[0], r)
        bad = dict(kw, irefatom=9)
        with pytest.raises(pkg.ArgumentError):
            srv.enqueue(np.arange(1, 4), du, dv, 0, pool[0].cells[:3], np.ones(3), **bad)
        with pytest.raises(pkg.ArgumentError):
            srv.enqueue(np.arange(1, 4), du, dv, 0, pool[0].cells[:3], -np.ones(3), **kw)
        tk = srv.enqueue(np.arange(1, 4), du, dv, 0, pool[0].cells[:3], np.ones(3), **kw)
        raw, tm = srv.wait(tk)
        assert tm["enqueued"] <= tm["started"] <= tm["finished"]
        assert raw["md_count"].sum() > 0
    finally:
        srv.close()


def test_autotuner_keeps_counts(pkg, pool):
    S = _serve()
    tuning, timing = S.autotune(pool[0], frames=2)
    assert tuning in S.TUNING_CANDIDATES and len(timing) == len(S.TUNING_CANDIDATES)
