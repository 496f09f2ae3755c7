import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as G  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return G.load_package()


@pytest.fixture(scope="session")
def oracle():
    return G.load_oracle()


def oracle_job(pkg, O, traj, options, coordination_number_only=False, use_cells=True, nthreads=1):
    """Oracle job with the parameters the product derives for the same inputs."""
    from complexmixtures_amd.mddf import TrajectoryMetaData
    from complexmixtures_amd.results import isautocorrelation
    from complexmixtures_amd.selection import group_csr
    tm = TrajectoryMetaData(traj, options)
    auto = isautocorrelation(traj)
    su, sv = traj.solute, traj.solvent

    def lists(ptr_idx):
        ptr, idx = ptr_idx
        if ptr is None:
            return None
        return [list(idx[ptr[k]:ptr[k + 1]]) for k in range(len(ptr) - 1)]

    gu = lists(group_csr(su, list(range(su.natomspermol)), su))
    gv = lists(group_csr(sv, list(range(sv.nmols * sv.natomspermol)), su if auto else sv))
    return O.Job(su.nmols, su.natomspermol, sv.nmols, sv.natomspermol, autocorrelation=auto,
                 irefatom=tm.irefatom - 1, binstep=options.binstep, cutoff=options.cutoff, dbulk=options.dbulk,
                 usecutoff=options.usecutoff, n_random_samples=options.n_random_samples,
                 coordination_only=coordination_number_only, triclinic=not tm.orthorhombic, seed=options.seed,
                 ngroups_u=tm.n_groups_solute, ngroups_v=tm.n_groups_solute if auto else tm.n_groups_solvent,
                 groups_u=gu, groups_v=gv, use_cells=use_cells, nthreads=nthreads), tm


def selected_frames(options, tm, weights):
    fr = [i for i in range(options.firstframe, tm.lastframe_read + 1, options.stride) if weights[i - 1] != 0.0]
    return fr


RAW_KEYS = ("md_count", "md_count_random", "rdf_count", "rdf_count_random", "solute_group_count",
            "solvent_group_count", "solute_group_count_random", "solvent_group_count_random")
ORACLE_KEY = {"solute_group_count": "gu", "solvent_group_count": "gv", "solute_group_count_random": "gu_random",
              "solvent_group_count_random": "gv_random"}


def oracle_raw(pkg, O, traj, options, frame_weights=None, coordination_number_only=False, **kw):
    job, tm = oracle_job(pkg, O, traj, options, coordination_number_only, **kw)
    w = np.ones(tm.lastframe_read) if frame_weights is None else np.asarray(frame_weights, dtype=np.float64)
    fr = selected_frames(options, tm, w)
    f0 = np.array(fr) - 1
    if hasattr(traj, "xu"):
        xu, xv, cells = traj.xu[f0], traj.xv[f0], traj.cells[f0]
    else:
        traj.opentraj()
        traj.firstframe()
        allf = traj.read_frames(tm.lastframe_read)
        traj.closetraj()
        xu, xv, cells = allf[0][f0], allf[1][f0], allf[2][f0]
    from complexmixtures_amd.trajectory import convert_unitcell
    cells = np.stack([convert_unitcell(c, tm.orthorhombic) for c in cells])
    raw = O.run(job, fr, np.asarray(xu, np.float64), np.asarray(xv, np.float64), cells, w[f0])
    return {k: raw[ORACLE_KEY.get(k, k)] for k in RAW_KEYS} | {"volume_total": raw["volume_total"]}, job


def gpu_raw(pkg, traj, options, frame_weights=None, coordination_number_only=False, **kw):
    from complexmixtures_amd.mddf import _accumulate
    R, tm, raw = _accumulate(traj, options, frame_weights, coordination_number_only, **kw)
    return raw


def assert_raw_equal(a, b, keys=RAW_KEYS, rtol=0.0):
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (k, x.shape, y.shape)
        if rtol == 0.0:
            if not np.array_equal(x, y):
                bad = np.argwhere(x != y)
                raise AssertionError(f"{k}: {len(bad)} entries differ, first at {bad[:5].tolist()}: "
                                     f"{x[tuple(bad[0])]} vs {y[tuple(bad[0])]}")
        else:
            np.testing.assert_allclose(x, y, rtol=rtol, atol=0, err_msg=k)
    assert a["volume_total"] == b["volume_total"] or rtol and np.isclose(a["volume_total"], b["volume_total"],
                                                                         rtol=rtol)


def assert_group_split(raw, autocorrelation):
    """Default groups: every record adds its weight once to md_count and, over
    the groups, once to the solute side (self: w/2 for atom i and w/2 for atom
    j, src/update_counters.jl:48-53) and, for cross jobs, once to the solvent
    side; the same for the ideal-gas arrays."""
    for md, gu, gv in (("md_count", "solute_group_count", "solvent_group_count"),
                       ("md_count_random", "solute_group_count_random", "solvent_group_count_random")):
        assert np.array_equal(np.asarray(raw[gu]).sum(axis=0), raw[md]), gu
        if autocorrelation:
            assert not np.asarray(raw[gv]).any(), gv
        else:
            assert np.array_equal(np.asarray(raw[gv]).sum(axis=0), raw[md]), gv
