"""Host side of the serving path: the request stream, the reference-atom
default and the footprint estimate (no GPU work)."""
import numpy as np
import pytest


def test_request_stream_mix(pkg):
    import complexmixtures_amd.serve as S
    rng = np.random.default_rng(1)
    reqs = S.request_stream(rng, 4000, 7, 200, rate=20.0)
    arr = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert np.all(gaps >= 0) and abs(gaps.mean() - 1 / 20.0) < 0.004          # Poisson, rate 20/s
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.08                         # exponential gaps
    fr = np.array([r.nframes for r in reqs])
    assert fr.min() >= 10 and fr.max() <= 200 and abs(fr.mean() - 105) < 4
    assert abs(np.mean([r.mddf for r in reqs]) - 0.3) < 0.03
    assert abs(np.mean([r.self_ for r in reqs]) - 0.2) < 0.03
    assert {r.system for r in reqs} == set(range(7))
    sat = S.request_stream(rng, 10, 3, 50, saturate=True)
    assert all(r.arrival == 0.0 for r in sat) and max(r.nframes for r in sat) <= 50


def test_default_irefatom(pkg):
    """Water: the oxygen (atom 1 of the template) is closest to the centre."""
    import complexmixtures_amd.serve as S
    import complexmixtures_amd.synthetic as syn
    rng = np.random.default_rng(2)
    w = syn.place_molecules(rng, syn.water_template(), rng.uniform(0, 30, (4, 3)))
    assert S.default_irefatom(w) == 1
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.2, 0, 0]])
    assert S.default_irefatom(x) == 2


def test_served_system_shapes(pkg):
    import complexmixtures_amd.synthetic as syn
    s = syn.served_system(3, 250, 32.0, 10, noise_frames=4)
    assert s["x_solute"].shape == (10, 250, 3) and s["x_solvent"].shape == (10, 3 * s["n_water"], 3)
    assert s["n_water"] == int(0.0334 * 32.0 ** 3)
    d = s["x_solvent"][1:] - s["x_solvent"][:1]
    assert 0.3 < np.std(d) < 0.6                                                # two N(0, 0.3) draws apart
    assert not np.array_equal(s["x_solute"][0], s["x_solute"][4])               # rolled noise differs


def test_footprint_grows_with_frames(pkg):
    from complexmixtures_amd import _lib
    kw = dict(nmols_u=1, natoms_per_mol_u=500, nmols_v=1000, natoms_per_mol_v=3, autocorrelation=False, irefatom=1,
              nbins=500, binstep=0.02, cutoff=10.0, dbulk=10.0, usecutoff=False, n_random_samples=10,
              coordination_number_only=False, triclinic=False, seed=1, n_groups_u=500, n_groups_v=3)
    cells = np.repeat(np.diag([40.0] * 3)[None], 64, axis=0)
    a = _lib.footprint(cells[:1], **kw)
    b = _lib.footprint(cells[:8], **kw)
    c = _lib.footprint(cells, **kw)
    assert 0 < a < b < c
    assert c > 2 * 500 * 500 * 8                                                # both solute group arrays
    with pytest.raises(pkg.ArgumentError):
        _lib.footprint(cells[:1], **dict(kw, irefatom=4))
