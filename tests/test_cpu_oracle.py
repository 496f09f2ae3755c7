"""CPU tests pinning the FP64 oracle.

Each function of the oracle is checked against the reference's own tests and
known answers (cited), against mathematics (closed forms, invariants, brute
force on tiny inputs), or against published test vectors.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

# ---------------------------------------------------------------------------
# random numbers and trigonometry
# ---------------------------------------------------------------------------
# Known-answer vectors of Philox4x32-10 published with Random123 (kat_vectors).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(oracle, ctr, key, expect):
    assert oracle.philox(ctr, key) == expect


def test_sincos2pi_against_libm(oracle):
    rng = np.random.default_rng(0)
    us = np.concatenate([rng.random(20000), np.arange(0, 1, 1 / 64), [0.0, 0.125, 0.375, 1 - 2 ** -53]])
    for u in us:
        u = float(u)
        s, c = oracle.sincos2pi(u)
        assert abs(s - math.sin(2 * math.pi * u)) <= 1e-15
        assert abs(c - math.cos(2 * math.pi * u)) <= 1e-15
        # libm on the exactly reduced argument (|2 pi r| <= pi/4, rounding of 2 pi r below 6e-17)
        q = round(4 * u)
        r = (4 * u - q) / 4
        sr, cr = math.sin(2 * math.pi * r), math.cos(2 * math.pi * r)
        ss, cc = [(sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)][q % 4]
        assert abs(s - ss) <= 2.5e-16 and abs(c - cc) <= 2.5e-16
    assert oracle.sincos2pi(0.0) == (0.0, 1.0)
    assert oracle.sincos2pi(0.25) == (1.0, -0.0)
    assert oracle.sincos2pi(0.5) == (-0.0, -1.0)
    assert oracle.sincos2pi(0.75) == (-1.0, 0.0)


# ---------------------------------------------------------------------------
# rigid-body moves, src/rigid_body.jl
# ---------------------------------------------------------------------------
def test_eulermat_closed_forms(oracle):
    """src/rigid_body.jl:59-65"""
    np.testing.assert_allclose(oracle.eulermat(0.0, 0.0, 0.0), np.eye(3), atol=1e-15)
    np.testing.assert_allclose(oracle.eulermat(math.pi, 0.0, 0.0), [[1, 0, 0], [0, -1, 0], [0, 0, -1]], atol=1e-15)
    np.testing.assert_allclose(oracle.eulermat(0.0, math.pi, 0.0), [[-1, 0, 0], [0, 1, 0], [0, 0, -1]], atol=1e-15)
    np.testing.assert_allclose(oracle.eulermat(0.0, 0.0, math.pi), [[-1, 0, 0], [0, -1, 0], [0, 0, 1]], atol=1e-15)
    # angles in turns give the same matrices, and they are rotations
    np.testing.assert_array_equal(oracle.eulermat_turns(0.5, 0.0, 0.0), [[1, 0, 0], [0, -1, 0], [0, 0, -1]])
    rng = np.random.default_rng(1)
    for _ in range(200):
        u = rng.random(3)
        A = oracle.eulermat_turns(*u)
        np.testing.assert_allclose(A @ A.T, np.eye(3), atol=2e-15)
        assert abs(np.linalg.det(A) - 1) < 2e-15
        np.testing.assert_allclose(A, oracle.eulermat(*(2 * math.pi * u)), atol=2e-15)


def test_move_closed_forms(oracle):
    """src/rigid_body.jl:82-96"""
    x = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    np.testing.assert_allclose(oracle.move(x, [0, 0, 0], 0, 0, 0), [[0.5, 0, 0], [-0.5, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(oracle.move(x, [1, 1, 1], 0, 0, 0), [[1.5, 1, 1], [0.5, 1, 1]], atol=1e-15)
    np.testing.assert_allclose(oracle.move(x, [0, 0, 0], math.pi, 0, 0), [[0.5, 0, 0], [-0.5, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(oracle.move(x, [0, 0, 0], 0, math.pi, 0), [[-0.5, 0, 0], [0.5, 0, 0]], atol=1e-15)
    np.testing.assert_allclose(oracle.move(x, [0, 0, 0], 0, 0, math.pi), [[-0.5, 0, 0], [0.5, 0, 0]], atol=1e-15)


def _internal(x):
    return np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)


@pytest.mark.parametrize("cell", [np.diag([10.0, 10.0, 10.0]),
                                  np.array([[10.0, 5.0, 0.0], [0.0, 10.0, 0.0], [0.0, 0.0, 10.0]])])
def test_random_move_rigid(oracle, cell):
    """src/rigid_body.jl:139-189: internal distances are preserved, also for a
    molecule split across the cell boundary (made whole around its reference atom)."""
    rng = np.random.default_rng(2)
    tric = not np.allclose(cell, np.diag(np.diag(cell)))
    for shift in (-1.0, -9.0, 4.0):
        x = shift + 2 * rng.random((5, 3))
        y = oracle.random_move(x, 0, cell, tric, rng.random(3), rng.random(3))
        np.testing.assert_allclose(_internal(x), _internal(y), atol=1e-12)
    # a molecule split by the boundary: atom 1 one lattice vector away
    x = np.array([[0.5, 0.5, 0.5], [1.2, 0.5, 0.5], [0.5, 1.3, 0.5]])
    x[1] += cell[:, 0]
    y = oracle.random_move(x, 0, cell, tric, [0.3, 0.4, 0.5], [0.1, 0.2, 0.3])
    whole = x.copy()
    whole[1] -= cell[:, 0]
    np.testing.assert_allclose(_internal(whole), _internal(y), atol=1e-12)
    # the centre lands at H * u
    np.testing.assert_allclose(y.mean(axis=0), cell @ np.array([0.3, 0.4, 0.5]), atol=1e-12)


# ---------------------------------------------------------------------------
# binning and shells, src/results.jl
# ---------------------------------------------------------------------------
def test_bins_and_shells(oracle, pkg):
    """src/results.jl:259-281 and setbin (src/results.jl:28)"""
    for f in (oracle.sphericalshellvolume, pkg.sphericalshellvolume):
        assert math.isclose(f(1, 1.0), 4 * math.pi / 3)
        assert math.isclose(f(2, 1.0), 4 * math.pi / 3 * (8 - 1))
        assert math.isclose(f(3, 1.0), 4 * math.pi / 3 * (27 - 8))
    for f in (oracle.shellradius, pkg.shellradius):
        assert math.isclose(f(1, 0.1), 0.07937005259840998, rel_tol=1e-12)
        assert math.isclose(f(5, 0.3), 1.3664650373440481, rel_tol=1e-12)
    for d, b in ((0.0, 1), (0.005, 1), (0.02, 1), (0.0200001, 2), (9.0, 450), (10.0, 500)):
        assert oracle.setbin(d, 0.02) == b == pkg.setbin(d, 0.02)
    for i in range(1, 501):
        assert oracle.shellradius(i, 0.02) == pkg.shellradius(i, 0.02)


# ---------------------------------------------------------------------------
# minimum distance records, src/minimum_distances.jl
# ---------------------------------------------------------------------------
def test_update_md(oracle):
    """src/minimum_distances.jl:41-46 plus the tie rule (smallest (i, j))."""
    r = oracle.update_md((1, 1, 2, 1.0, 1, 0, 1.0), (1, 1, 2, 0.5, 1, 0, 0.5))
    assert (r.within_cutoff, r.i, r.j, r.d, r.ref_atom_within_cutoff, r.d_ref_atom) == (1, 1, 2, 0.5, 1, 0.5)
    zero = (0, -1, -1, math.inf, 0, 0, math.inf)
    r = oracle.update_md(zero, (1, 3, 7, 2.0, 0, 0, math.inf))
    assert (r.i, r.j, r.d, r.ref_atom_within_cutoff, r.d_ref_atom) == (3, 7, 2.0, 0, math.inf)
    # ref flag and d_ref are merged independently of the winner
    r = oracle.update_md((1, 3, 7, 2.0, 0, 0, math.inf), (1, 4, 9, 3.0, 1, 0, 3.0))
    assert (r.i, r.j, r.d, r.ref_atom_within_cutoff, r.d_ref_atom) == (3, 7, 2.0, 1, 3.0)
    # ties on d: smallest (i, j) whatever the order
    a, b = (1, 4, 9, 2.0, 0, 0, math.inf), (1, 3, 12, 2.0, 0, 0, math.inf)
    assert oracle.update_md(a, b).i == 3 and oracle.update_md(b, a).i == 3


def _brute_min_image_d2(cell, x, y, reach=5):
    n = np.stack(np.meshgrid(*[np.arange(-reach, reach + 1)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return float(np.min(np.sum((y - x - n @ cell.T) ** 2, axis=1)))


@pytest.mark.parametrize("tric", [False, True])
def test_canonical_distance_is_minimum_image(oracle, pkg, tric):
    """Within half the smallest plane spacing the recorded distance is the
    minimum over all lattice images (checked by enumeration)."""
    cell = pkg.unitcell_matrix(30.0, 32.0, 34.0, 80.0, 75.0, 85.0) if tric else np.diag([30.0, 32.0, 34.0])
    rng = np.random.default_rng(3)
    n_in = 0
    for _ in range(400):
        x = rng.uniform(-60, 60, 3)
        y = x + rng.normal(size=3) * 6 + cell @ rng.integers(-3, 4, 3)
        d2 = oracle.canon_d2(cell, tric, x, y)
        b = _brute_min_image_d2(cell, x, y)
        if b < 14.0 ** 2:
            n_in += 1
            assert abs(d2 - b) <= 1e-9 * max(b, 1.0)
        else:
            assert d2 >= b - 1e-9 * b
    assert n_in > 100


# ---------------------------------------------------------------------------
# whole-frame algorithm
# ---------------------------------------------------------------------------
def _toy_cross():
    xv = np.array([[9, 0, 0], [9, 2, 0], [9, 0, 2], [11, 0, 0], [11, 2, 0], [11, 0, 2], [12, 0, 0], [12, 2, 0],
                   [12, 0, 2]], float)
    return np.zeros((1, 3)), xv, np.diag([30.0] * 3)


@pytest.mark.parametrize("use_cells", [False, True])
def test_toy_cross_oracle(oracle, use_cells):
    """src/mddf.jl:587-624 (known answers of the toy cross system)"""
    xu, xv, cell = _toy_cross()
    for nf in (1, 2):
        job = oracle.Job(1, 1, 3, 3, irefatom=0, n_random_samples=10 ** 4, seed=321, use_cells=use_cells)
        raw = oracle.run(job, list(range(1, nf + 1)), np.stack([xu] * nf), np.stack([xv] * nf),
                         np.stack([cell] * nf), [1.0] * nf)
        fin = oracle.finalize(job, raw, float(nf))
        assert fin["volume_total"] == 27000.0
        assert math.isclose(fin["md_count"].sum(), 1.0)
        assert math.isclose(fin["coordination_number"].sum(), 51.0)
        assert math.isclose(fin["density_solvent"], 3 / 27000)
        assert math.isclose(fin["density_solvent_bulk"], 2 / fin["volume_bulk"])
        assert math.isclose(fin["volume_domain"], fin["volume_total"] - fin["volume_bulk"])
        assert math.isclose(fin["volume_domain"], 4 * math.pi / 3 * 1000, rel_tol=0.03)
    rec = oracle.records(job, xu, xv, cell, 0)
    assert rec[0][:5] == (1, 0, 0, 9.0, 1) and rec[1][0] == 0 and rec[2][0] == 0


def test_toy_self_oracle(oracle):
    """src/mddf.jl:629-654: two atoms 5 A apart, self correlation"""
    xv = np.array([[[0.0, 0, 0], [5.0, 0, 0]]])
    cell = np.diag([30.0] * 3)[None]
    job = oracle.Job(2, 1, 2, 1, autocorrelation=True, irefatom=0, n_random_samples=10 ** 4, seed=321)
    raw = oracle.run(job, [1], xv, xv, cell, [1.0])
    fin = oracle.finalize(job, raw, 1.0)
    assert math.isclose(fin["md_count"].sum(), 1.0)
    assert math.isclose(fin["density_solute"], 2 / 27000)
    assert math.isclose(fin["volume_domain"], 4 * math.pi / 3 * 1000, rel_tol=0.1)


def test_dupdcd_known_answers_oracle(oracle, pkg):
    """src/tools/coordination_number.jl:108-135: 7 and 14 TMAO molecules within 3
    and 5 A of the protein in frame 1; the O1 contribution sums to 1171."""
    z = np.load(os.path.join(GOLDEN, "dupdcd_atoms_1_4012.npz"))
    xyz = z["xyz"].astype(np.float64)
    A, B, C = z["dcd_cell"][0][[0, 2, 5]]
    cell = pkg.unitcell_matrix(A, B, C)
    xu, xv = xyz[:1, 0:1463], xyz[:1, 1478:4012]
    job = oracle.Job(1, 1463, 181, 14, irefatom=0, n_random_samples=200, seed=321)
    raw = oracle.run(job, [1], xu, xv, cell[None], [1.0])
    fin = oracle.finalize(job, raw, 1.0)
    d, cn = fin["d"], fin["coordination_number"]
    assert cn[np.argmax(d > 3)] == 7.0
    assert cn[np.argmax(d > 5)] == 14.0
    assert np.cumsum(fin["gv"][4]).sum() == 1171.0
    assert fin["gu"].sum() == fin["gv"].sum()
    brute = oracle.Job(1, 1463, 181, 14, irefatom=0, n_random_samples=200, seed=321, use_cells=False)
    raw_b = oracle.run(brute, [1], xu, xv, cell[None], [1.0])
    for k in ("md_count", "rdf_count", "gu", "gv", "md_count_random", "gv_random"):
        assert np.array_equal(raw[k], raw_b[k])


def _small_systems(pkg):
    import complexmixtures_amd.synthetic as S
    out = []
    s = S.config1(nframes=2, n_solute=150, n_water=400, box=24.0, seed=31)
    out.append(("cross", s, dict(nmols_u=1, nu=150, nmols_v=400, nv=3), False))
    s = S.config3(nframes=2, n_solute=120, n_urea=250, n_water=5, seed=32, a=30.0)
    out.append(("triclinic", s, dict(nmols_u=1, nu=120, nmols_v=250, nv=8), True))
    s = S.config3(nframes=1, n_solute=5, n_urea=200, n_water=5, solvent="self", seed=33, a=28.0)
    out.append(("self", s, dict(nmols_u=200, nu=8, nmols_v=200, nv=8, autocorrelation=True), True))
    return out


@pytest.mark.parametrize("case", [0, 1, 2])
def test_cells_equal_brute_force(oracle, pkg, case):
    """The cell-list enumeration finds exactly the pairs of the all-pairs loop."""
    name, s, kw, tric = _small_systems(pkg)[case]
    nf = s["x_solute"].shape[0]
    for usecutoff in (False, True):
        res = []
        for use_cells in (False, True):
            job = oracle.Job(**kw, irefatom=1, cutoff=8.0, dbulk=6.0 if usecutoff else 8.0, usecutoff=usecutoff,
                             n_random_samples=3, triclinic=tric, seed=5, use_cells=use_cells)
            res.append(oracle.run(job, list(range(1, nf + 1)), s["x_solute"], s["x_solvent"], s["cells"],
                                  [1.0] * nf))
        for k in res[0]:
            assert np.array_equal(res[0][k], res[1][k]), (name, k)


def test_parallel_chunks_equal_serial(oracle, pkg):
    """The reference's chunked frame parallelism (src/mddf.jl:285-338) with sum!
    gives the serial result for integer weights."""
    import complexmixtures_amd.synthetic as S
    s = S.config1(nframes=5, n_solute=120, n_water=300, box=24.0, seed=34)
    r = []
    for nth in (1, 3):
        job = oracle.Job(1, 120, 300, 3, irefatom=0, cutoff=8.0, dbulk=8.0, n_random_samples=3, seed=9, nthreads=nth)
        r.append(oracle.run(job, [1, 2, 3, 4, 5], s["x_solute"], s["x_solvent"], s["cells"], [1.0] * 5))
    for k in r[0]:
        assert np.array_equal(r[0][k], r[1][k])


def test_placement_statistics(oracle, pkg):
    """Placed molecules are rigid copies of bulk molecules with centres uniform
    in the cell (randomize_solvent!, src/mddf.jl:65-88)."""
    import complexmixtures_amd.synthetic as S
    s = S.config1(nframes=1, n_solute=10, n_water=2000, box=30.0, seed=35)
    job = oracle.Job(1, 10, 2000, 3, irefatom=0, cutoff=8.0, dbulk=8.0, n_random_samples=1, seed=4)
    bulk = np.arange(0, 2000, 2)
    placed = oracle.placement(job, s["x_solvent"][0], s["cells"][0], bulk, 1, 0).reshape(2000, 3, 3)
    src = s["x_solvent"][0].reshape(2000, 3, 3)
    keys = {tuple(np.round(_internal(src[b])[np.triu_indices(3, 1)], 8)) for b in bulk}
    for m in range(2000):
        assert tuple(np.round(_internal(placed[m])[np.triu_indices(3, 1)], 8)) in keys
    cen = placed.mean(axis=1) / 30.0
    assert np.all((cen > -0.01) & (cen < 1.01))
    assert abs(cen.mean() - 0.5) < 0.02 and abs(cen.std() - math.sqrt(1 / 12)) < 0.02


# ---------------------------------------------------------------------------
# finalisation against the reference's golden Result files
# ---------------------------------------------------------------------------
def _goldens():
    meta = json.load(open(os.path.join(GOLDEN, "result_goldens.json")))
    arr = np.load(os.path.join(GOLDEN, "result_goldens.npz"))
    return meta, arr


@pytest.mark.parametrize("key", ["NAMD__tmao_tmao", "NAMD__water_tmao", "NAMD__water_water", "PDB__tmao_tmao",
                                 "Gromacs__EMI_DCA", "Gromacs__EMI_EMI"])
def test_finalresults_reproduces_goldens(oracle, pkg, key):
    """Raw counts recovered from a golden Result, finalised again by the oracle
    and by the host, reproduce every derived field of the golden
    (src/results.jl:311-428) to 1e-10 relative."""
    meta, arr = _goldens()
    m = meta[key]
    g = {k.split(".", 1)[1]: arr[k] for k in arr.files if k.startswith(key + ".")}
    o = m["options"]
    Q = float(sum(m["frame_weights"][i - 1] for i in range(o["firstframe"], m["lastframe_read"] + 1, o["stride"])))
    nu = m["solute"]["nmols"]
    nr = o["n_random_samples"]
    fix = m["density"]["solvent_bulk"] / m["density"]["solvent"]
    raw = {
        "md_count": g["md_count"] * (nu * Q),
        "md_count_random": g["md_count_random"] / fix * (nr * Q),
        "rdf_count": g["rdf_count"] * (nu * Q),
        "rdf_count_random": g["rdf_count_random"] / fix * (nr * Q),
        "gu": g["solute_group_count"] * (nu * Q),
        "gv": g["solvent_group_count"] * (nu * Q),
        "gu_random": g["solute_group_count_random"] / fix * (nr * Q),
        "gv_random": g["solvent_group_count_random"] / fix * (nr * Q),
        "volume_total": m["volume"]["total"] * Q,
    }
    # raw counts of a trajectory are sums of frame weights: integers here
    np.testing.assert_allclose(raw["md_count"], np.round(raw["md_count"]), atol=1e-6)
    job = oracle.Job(nu, m["solute"]["natomspermol"], m["solvent"]["nmols"], m["solvent"]["natomspermol"],
                     autocorrelation=m["autocorrelation"], binstep=o["binstep"], cutoff=o["cutoff"], dbulk=o["dbulk"],
                     usecutoff=o["usecutoff"], n_random_samples=nr,
                     ngroups_u=len(g["solute_group_count"]), ngroups_v=len(g["solvent_group_count"]))
    fin = oracle.finalize(job, raw, Q)
    for k in ("d", "md_count", "coordination_number", "coordination_number_random", "mddf", "kb", "rdf",
              "kb_rdf", "sum_rdf_count", "sum_rdf_count_random"):
        np.testing.assert_allclose(fin[k], g[k], rtol=1e-10, atol=1e-10 * np.max(np.abs(g[k])), err_msg=k)
    np.testing.assert_allclose(fin["volume_shell"], g["volume_shell"], rtol=1e-10, atol=1e-9)
    for k, gk in (("volume_domain", m["volume"]["domain"]), ("volume_bulk", m["volume"]["bulk"]),
                  ("density_solvent_bulk", m["density"]["solvent_bulk"]),
                  ("density_solvent", m["density"]["solvent"]), ("density_solute", m["density"]["solute"])):
        assert math.isclose(fin[k], gk, rel_tol=1e-10), (k, fin[k], gk)
    # the host's finalresults gives the oracle's bits
    from complexmixtures_amd.options import Options, TrajectoryFileOptions
    from complexmixtures_amd.results import Result, finalresults
    opts = Options.from_dict(o)
    sel = lambda s: pkg.AtomSelection(list(range(1, s["nmols"] * s["natomspermol"] + 1)), nmols=s["nmols"])
    R = Result(nbins=m["nbins"], dbulk=m["dbulk"], cutoff=m["cutoff"], autocorrelation=m["autocorrelation"],
               solute=sel(m["solute"]), solvent=sel(m["solvent"]), n_groups_solute=len(g["solute_group_count"]),
               n_groups_solvent=len(g["solvent_group_count"]),
               files=[TrajectoryFileOptions("golden", opts, m["irefatom"], m["lastframe_read"], m["nframes_read"],
                                            m["frame_weights"])])
    R.md_count[:], R.md_count_random[:] = raw["md_count"], raw["md_count_random"]
    R.rdf_count[:], R.rdf_count_random[:] = raw["rdf_count"], raw["rdf_count_random"]
    R.solute_group_count[:], R.solute_group_count_random[:] = raw["gu"], raw["gu_random"]
    R.solvent_group_count[:], R.solvent_group_count_random[:] = raw["gv"], raw["gv_random"]
    R.volume.total = raw["volume_total"]
    finalresults(R, opts, False)
    for k in ("md_count", "md_count_random", "coordination_number", "mddf", "kb", "rdf", "kb_rdf"):
        assert np.array_equal(getattr(R, k), fin[k]), k
    assert R.volume.domain == fin["volume_domain"] and R.density.solvent_bulk == fin["density_solvent_bulk"]


# ---------------------------------------------------------------------------
# group-count split (update_counters!, src/update_counters.jl:43-88)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("key", ["NAMD__tmao_tmao", "NAMD__water_tmao", "NAMD__water_water", "PDB__tmao_tmao",
                                 "Gromacs__EMI_DCA", "Gromacs__EMI_EMI"])
def test_group_split_against_goldens(key):
    """Every record within the cutoff adds its weight w to md_count once and,
    summed over the groups, w to the solute groups: cross MDDFs add w for
    atom i to the solute groups and w for atom j to the solvent groups; self
    MDDFs add w/2 for atom i and w/2 for atom j, both to the solute groups
    (src/update_counters.jl:48-53, 72-77; finalresults! copies them to the
    solvent side, src/results.jl:335-337).  The reference's own Result files
    fix this: sum_g solute_group_count[g] == md_count bin by bin, and the same
    for the ideal-gas arrays.  A rule with w instead of w/2 in self mode gives
    twice md_count; one that sends atom j to the solvent groups gives half."""
    meta, arr = _goldens()
    m = meta[key]
    assert not m["solute"]["custom_groups"] and not m["solvent"]["custom_groups"]
    g = {k.split(".", 1)[1]: arr[k] for k in arr.files if k.startswith(key + ".")}
    for md, gu, gv in (("md_count", "solute_group_count", "solvent_group_count"),
                       ("md_count_random", "solute_group_count_random", "solvent_group_count_random")):
        scale = np.max(g[md])
        assert scale > 0
        np.testing.assert_allclose(g[gu].sum(axis=0), g[md], rtol=0, atol=1e-12 * scale, err_msg=f"{key} {gu}")
        np.testing.assert_allclose(g[gv].sum(axis=0), g[md], rtol=0, atol=1e-12 * scale, err_msg=f"{key} {gv}")
        if m["autocorrelation"]:
            assert np.array_equal(g[gu], g[gv])


@pytest.mark.parametrize("auto", [False, True])
def test_group_split_oracle_raw(oracle, pkg, auto):
    """The oracle's raw accumulators satisfy the same split: with unit weights
    sum_g gu == md_count exactly (self: half units for atom i and atom j),
    sum_g gv == md_count for cross jobs and gv untouched for self jobs (the
    reference copies the solute groups at finalisation)."""
    import complexmixtures_amd.synthetic as S
    rng = np.random.default_rng(61)
    box = 24.0
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (300, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    cells = np.repeat(np.diag([box] * 3)[None], 2, axis=0)
    if auto:
        job = oracle.Job(300, 3, 300, 3, autocorrelation=True, cutoff=8.0, dbulk=8.0, n_random_samples=3, seed=9)
        raw = oracle.run(job, [1, 2], xv, xv, cells, np.ones(2))
    else:
        xu = xv[:, :30]
        job = oracle.Job(10, 3, 290, 3, cutoff=8.0, dbulk=8.0, n_random_samples=3, seed=9)
        raw = oracle.run(job, [1, 2], xu, xv[:, 30:], cells, np.ones(2))
    for md, gu, gv in (("md_count", "gu", "gv"), ("md_count_random", "gu_random", "gv_random")):
        assert raw[md].sum() > 0
        assert np.array_equal(raw[gu].sum(axis=0), raw[md])
        if auto:
            assert not raw[gv].any()
            assert np.all(raw[gu] * 2 == np.round(raw[gu] * 2))
        else:
            assert np.array_equal(raw[gv].sum(axis=0), raw[md])
