"""GPU parity at the benchmark sizes: the 110 A / 5,000-atom globule / 40,000
water system (FP32 coordinates, every solute atom its own group) and the
triclinic 90 A / 80 degree systems (3,000-atom polymer vs 12,000 urea and vs
15,000 waters, urea vs urea), a few frames each, bit for bit against the FP64
oracle through the C ABI."""
import numpy as np
import pytest

from conftest import assert_group_split, assert_raw_equal, gpu_raw, oracle_raw

pytestmark = pytest.mark.gpu


def _traj(pkg, s):
    return pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])


@pytest.fixture(scope="module")
def globule_water():
    import complexmixtures_amd.synthetic as S
    return S.config2(nframes=4)


@pytest.mark.parametrize("tuning", [None, dict(cell_fraction=0.35, columns=8), dict(cell_fraction=0.75, columns=2)])
def test_globule_water_full_size(pkg, oracle, globule_water, tuning):
    """110 A cubic box, 5,000-atom globule, 40,000 waters, FP32 input,
    n_random_samples 10, 5,000 solute groups: 4 frames, default and two other
    tunings of the tile kernels."""
    s = globule_water
    assert s["x_solute"].dtype == np.float32 and s["x_solute"].shape == (4, 5000, 3)
    assert s["x_solvent"].shape == (4, 120000, 3)
    traj = _traj(pkg, s)
    opts = pkg.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261204, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts, nthreads=4)
    b = gpu_raw(pkg, traj, opts, job_kw={"tuning": tuning} if tuning else None)
    assert a["md_count"].sum() > 1000 and a["md_count_random"].sum() > 1e5
    assert_raw_equal(a, b)
    assert_group_split(b, False)


@pytest.mark.parametrize("solvent,nframes", [("urea", 2), ("water", 1), ("self", 2)])
def test_triclinic_full_size(pkg, oracle, solvent, nframes):
    """a = b = c = 90 A, 80 degree angles, FP64 input, n_random_samples 10:
    3,000-atom polymer vs 12,000 8-site urea and vs 15,000 waters, and the
    12,000 urea as both solute and solvent."""
    import complexmixtures_amd.synthetic as S
    s = S.config3(nframes=nframes, solvent=solvent)
    traj = _traj(pkg, s)
    opts = pkg.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261205, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts, nthreads=4)
    b = gpu_raw(pkg, traj, opts)
    assert a["md_count"].sum() > 1000
    assert_raw_equal(a, b)
    assert_group_split(b, solvent == "self")
