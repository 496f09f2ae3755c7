"""Geometry the molecule-pair kernels do not handle, detected on the device:
rigid molecules so long that the cutoff plus two molecule radii exceed a plane
spacing of the cell (while the cutoff itself is below half of it).  A 50 A box
holds the same 22.5 A rods (whole under the minimum image, 10 + 2 x 11.25 A
below the spacing) and is accepted.  The
kernels still run within their buffers (the molecule grid degenerates to one
cell per dimension); the job reports CM_EUNSUPPORTED when it is collected,
from cm_job_finish (mddf) and from cm_server_wait (request server)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MSG = "molecules too large for the unit cell"


def _rods(nmol=20, natoms=16, box=30.0, nframes=2, seed=70):
    """Straight rods of natoms atoms 1.5 A apart (22.5 A long for 16 atoms)."""
    rng = np.random.default_rng(seed)
    rod = np.zeros((natoms, 3))
    rod[:, 0] = 1.5 * np.arange(natoms)
    rod -= rod.mean(axis=0)
    import complexmixtures_amd.synthetic as S
    x = S.place_molecules(rng, rod, rng.uniform(0, box, (nmol, 3)))
    xs = np.stack([x + rng.normal(scale=0.05, size=x.shape) for _ in range(nframes)])
    return xs, np.repeat(np.diag([box] * 3)[None], nframes, axis=0)


def test_long_molecules_rejected_by_finish(pkg):
    from complexmixtures_amd import _lib
    xs, cells = _rods()
    sel = pkg.AtomSelection(list(range(1, xs.shape[1] + 1)), natomspermol=16)
    traj = pkg.MemoryTrajectory(xs, xs, cells, sel, sel)
    with pytest.raises(_lib.GPUError, match=MSG):
        pkg.mddf(traj, pkg.Options(dbulk=10.0, n_random_samples=2, seed=3, silent=True))
    # a cell with room for the molecules is accepted
    xs2, cells2 = _rods(box=50.0)
    traj2 = pkg.MemoryTrajectory(xs2, xs2, cells2, sel, sel)
    R = pkg.mddf(traj2, pkg.Options(dbulk=10.0, n_random_samples=2, seed=3, silent=True))
    assert R.md_count.sum() > 0


def test_long_molecules_rejected_by_server(pkg):
    from complexmixtures_amd import _lib
    from complexmixtures_amd.results import setbin
    xs, cells = _rods()
    d = _lib.DeviceArray(xs.astype(np.float32))
    kw = dict(nmols_u=20, natoms_per_mol_u=16, nmols_v=20, natoms_per_mol_v=16, autocorrelation=True, irefatom=1,
              nbins=setbin(10.0, 0.02), binstep=0.02, cutoff=10.0, dbulk=10.0, usecutoff=False, n_random_samples=2,
              coordination_number_only=False, triclinic=False, seed=5, n_groups_u=16, n_groups_v=16)
    srv = _lib.Server(2)
    try:
        tk = srv.enqueue(np.arange(1, 3), d, d, 0, cells, np.ones(2), **kw)
        with pytest.raises(_lib.GPUError, match=MSG):
            srv.wait(tk)
        # the slot serves the next request normally
        xs2, cells2 = _rods(box=50.0)
        d2 = _lib.DeviceArray(xs2.astype(np.float32))
        tk = srv.enqueue(np.arange(1, 3), d2, d2, 0, cells2, np.ones(2), **kw)
        raw, _ = srv.wait(tk)
        assert raw["md_count"].sum() > 0
    finally:
        srv.close()
