"""GPU parity: libcmgpu.so against the FP64 oracle and the reference's known answers.

Every test runs the HIP library through the C ABI (via the package) and
compares raw accumulators with the oracle bit for bit (integer hit counts
times dyadic weights are exact), or checks the closed-form answers of the
reference's own tests.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_group_split, assert_raw_equal, gpu_raw, oracle_raw

pytestmark = pytest.mark.gpu

TOY = os.path.join(GOLDEN, "toy")


def mem(pkg, s):
    return pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])


def syn():
    import complexmixtures_amd.synthetic as S
    return S


# ---------------------------------------------------------------------------
# known answers of the reference's tests
# ---------------------------------------------------------------------------
def test_toy_cross_known_answers(pkg, oracle):
    """src/mddf.jl:587-624"""
    protein = pkg.AtomSelection([10], nmols=1)
    water = pkg.AtomSelection(list(range(1, 10)), natomspermol=3)
    f = os.path.join(TOY, "cross.pdb")
    for lastframe in (1, 2):
        opts = pkg.Options(seed=321, StableRNG=True, nthreads=1, silent=True, n_random_samples=10 ** 5,
                           lastframe=lastframe)
        R = pkg.mddf(f, protein, water, opts, trajectory_format="PDBTraj")
        assert R.volume.total == 27000.0
        assert math.isclose(R.volume.domain, R.volume.total - R.volume.bulk)
        assert math.isclose(R.volume.domain, (4 * math.pi / 3) * R.dbulk ** 3, rel_tol=0.01)
        assert math.isclose(R.density.solute, 1 / R.volume.total)
        assert math.isclose(R.density.solvent, 3 / R.volume.total)
        assert math.isclose(R.density.solvent_bulk, 2 / R.volume.bulk)
        assert math.isclose(R.md_count.sum(), 1)
        assert math.isclose(R.coordination_number.sum(), 51)
        C = pkg.coordination_number(f, protein, water, opts, trajectory_format="PDBTraj")
        assert C.volume.total == R.volume.total
        assert C.volume.domain == 0.0
        assert math.isclose(C.density.solute, 1 / C.volume.total)
        assert math.isclose(C.density.solvent, 3 / C.volume.total)
        assert C.density.solvent_bulk == 0.0
        assert np.array_equal(C.md_count, R.md_count)
        assert np.array_equal(pkg.coordination_number(C), pkg.coordination_number(R))
        traj = pkg.Trajectory(f, protein, water, format="PDBTraj")
        a, _ = oracle_raw(pkg, oracle, traj, opts)
        b = gpu_raw(pkg, pkg.Trajectory(f, protein, water, format="PDBTraj"), opts)
        assert_raw_equal(a, b)


def test_toy_self_known_answers(pkg, oracle):
    """src/mddf.jl:629-654 (self correlation, atoms 5 A apart)"""
    atom = pkg.AtomSelection([1, 2], natomspermol=1)
    f = os.path.join(TOY, "self_monoatomic.pdb")
    opts = pkg.Options(seed=321, StableRNG=True, nthreads=1, silent=True, n_random_samples=10 ** 5, lastframe=1)
    R = pkg.mddf(f, atom, opts, trajectory_format="PDBTraj")
    assert R.volume.total == 27000.0
    assert math.isclose(R.volume.domain, R.volume.total - R.volume.bulk)
    assert math.isclose(R.volume.domain, (4 * math.pi / 3) * R.dbulk ** 3, rel_tol=0.1)
    assert math.isclose(R.density.solute, 2 / R.volume.total)
    assert math.isclose(R.density.solvent, 2 / R.volume.total)
    assert math.isclose(R.md_count.sum(), 1.0)
    C = pkg.coordination_number(f, atom, opts, trajectory_format="PDBTraj")
    assert math.isclose(C.md_count.sum(), 1.0)
    a, _ = oracle_raw(pkg, oracle, pkg.Trajectory(f, atom, format="PDBTraj"), opts)
    b = gpu_raw(pkg, pkg.Trajectory(f, atom, format="PDBTraj"), opts)
    assert_raw_equal(a, b)


def test_toy_frame_weights(pkg):
    """src/mddf.jl:666-757"""
    atom = pkg.AtomSelection([1, 2], natomspermol=1)
    f = os.path.join(TOY, "self_monoatomic.pdb")
    fd = os.path.join(TOY, "self_monoatomic_duplicated_first_frame.pdb")
    kw = dict(trajectory_format="PDBTraj")
    R1 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True, firstframe=2), **kw)
    R2 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True), frame_weights=[0.0, 1.0], **kw)
    assert np.array_equal(R1.md_count, R2.md_count) and np.array_equal(R1.rdf_count, R2.rdf_count)
    R2 = pkg.mddf(f, atom, pkg.Options(silent=True), frame_weights=np.array([[0, 1]]), **kw)
    assert np.array_equal(R1.md_count, R2.md_count)
    with pytest.raises(pkg.ArgumentError):
        pkg.mddf(f, atom, pkg.Options(silent=True), frame_weights=np.array([[0, 1], [0, 1]]), **kw)
    R1 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True, lastframe=1), **kw)
    R2 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True, lastframe=1), frame_weights=[1.0, 0.0], **kw)
    assert np.array_equal(R1.md_count, R2.md_count) and np.array_equal(R1.rdf_count, R2.rdf_count)
    R1 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True), **kw)
    R2 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True), frame_weights=[0.3, 0.3], **kw)
    assert np.array_equal(R1.md_count, R2.md_count) and np.array_equal(R1.rdf_count, R2.rdf_count)
    R1 = pkg.mddf(fd, atom, pkg.Options(seed=321, silent=True, firstframe=1, lastframe=3), **kw)
    R2 = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True), frame_weights=[2.0, 1.0], **kw)
    assert np.array_equal(R1.md_count, R2.md_count) and np.array_equal(R1.rdf_count, R2.rdf_count)
    R2 = pkg.mddf(f, atom, pkg.Options(silent=True), frame_weights=[2.0, 1.0], **kw)
    assert math.isclose(R2.md_count.sum(), 2 / 3)
    R2 = pkg.mddf(f, atom, pkg.Options(silent=True), frame_weights=[1.0, 2.0], **kw)
    assert math.isclose(R2.md_count.sum(), 1 / 3)
    # only with atoms in the bulk, then both frames
    for firstframe, dens in ((2, 1.0), (1, 0.5)):
        R = pkg.mddf(f, atom, pkg.Options(seed=321, silent=True, n_random_samples=10 ** 5,
                                                           firstframe=firstframe), **kw)
        assert R.volume.total == 27000.0
        assert math.isclose(R.volume.domain, R.volume.total - R.volume.bulk)
        assert math.isclose(R.volume.domain, (4 * math.pi / 3) * R.dbulk ** 3, rel_tol=0.1)
        assert math.isclose(R.density.solvent_bulk, dens / R.volume.bulk)


def test_stop_file_interrupt(pkg, tmp_path, monkeypatch):
    """src/mddf.jl:656-664"""
    atom = pkg.AtomSelection([1, 2], natomspermol=1)
    f = os.path.join(TOY, "self_monoatomic.pdb")
    monkeypatch.chdir(tmp_path)
    (tmp_path / "stop_complexmixtures").write_text("")
    R = pkg.mddf(f, atom, pkg.Options(silent=True), trajectory_format="PDBTraj")
    assert R.md_count.sum() == 0.0


def _dupdcd(pkg):
    z = np.load(os.path.join(GOLDEN, "dupdcd_atoms_1_4012.npz"))
    xyz = z["xyz"]
    cells = np.stack([pkg.unitcell_matrix(*c[[0, 2, 5]]) for c in z["dcd_cell"]])
    protein = pkg.AtomSelection(list(range(1, 1464)), nmols=1)
    tmao = pkg.AtomSelection(list(range(1479, 4013)), natomspermol=14,
                             group_names=["N", "C1", "C2", "C3", "O1", "H11", "H12", "H13", "H21", "H22", "H23",
                                          "H31", "H32", "H33"])
    return xyz, cells, protein, tmao


def test_dupdcd_known_answers(pkg, oracle):
    """src/tools/coordination_number.jl:108-135 on frame 1 of the NAMD test trajectory."""
    xyz, cells, protein, tmao = _dupdcd(pkg)
    xu = xyz[:, 0:1463].astype(np.float64)
    xv = xyz[:, 1478:4012].astype(np.float64)
    traj = pkg.MemoryTrajectory(xu, xv, cells, protein, tmao)
    opts = pkg.Options(lastframe=1, seed=321, StableRNG=True, nthreads=1, silent=True, n_random_samples=200)
    R = pkg.mddf(traj, opts)
    assert R.files[0].irefatom == 1
    assert math.isclose(R.solute_group_count.sum(), R.solvent_group_count.sum())
    assert R.coordination_number[np.argmax(R.d > 3)] == 7.0
    assert R.coordination_number[np.argmax(R.d > 5)] == 14.0
    assert np.cumsum(R.solvent_group_count[tmao.group_names.index("O1")]).sum() == 1171.0
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_dupdcd_self_tmao_and_weights(pkg, oracle):
    """Self MDDF of TMAO (mols family) over the 3 frames; weights [2, 1] on frames
    (1, 3) equal the duplicated first frame (src/mddf.jl:860-879)."""
    xyz, cells, protein, tmao = _dupdcd(pkg)
    xv = xyz[:, 1478:4012].astype(np.float64)
    traj = pkg.MemoryTrajectory(xv, xv, cells, tmao, tmao)
    opts = pkg.Options(seed=321, silent=True, n_random_samples=4)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_group_split(b, True)
    R1 = pkg.mddf(traj, pkg.Options(silent=True, n_random_samples=4))
    R2 = pkg.mddf(traj, pkg.Options(silent=True, n_random_samples=4), frame_weights=[2.0, 0.0, 1.0])
    assert np.array_equal(R1.md_count, R2.md_count)
    assert np.array_equal(R1.solute_group_count, R2.solute_group_count)
    assert np.array_equal(R1.solvent_group_count, R2.solvent_group_count)
    assert R1.volume.total == R2.volume.total


# ---------------------------------------------------------------------------
# synthetic systems against the oracle
# ---------------------------------------------------------------------------
def test_parity_config1_reduced(pkg, oracle):
    s = syn().config1(nframes=4)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=10.0, binstep=0.02, n_random_samples=10, seed=20261203, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_group_split(b, False)
    assert b["hits_md"].sum() == round(a["md_count"].sum())


def test_parity_float32_input(pkg, oracle):
    s = syn().config1(nframes=3, n_solute=300, n_water=900, box=30.0, seed=5)
    s["x_solute"] = s["x_solute"].astype(np.float32)
    s["x_solvent"] = s["x_solvent"].astype(np.float32)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=8.0, n_random_samples=6, seed=77, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_parity_triclinic_urea(pkg, oracle):
    s = syn().config3(nframes=2, n_solute=400, n_urea=800, n_water=10, seed=3, a=40.0)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=10.0, n_random_samples=5, seed=9, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_parity_self_urea_triclinic(pkg, oracle):
    s = syn().config3(nframes=2, n_solute=10, n_urea=600, n_water=10, solvent="self", seed=4, a=36.0)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=9.0, n_random_samples=3, seed=10, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_group_split(b, True)


def test_parity_cross_many_solute_molecules(pkg, oracle):
    S = syn()
    rng = np.random.default_rng(12)
    box = 28.0
    wt = S.water_template()
    xu = S.place_molecules(rng, wt, rng.uniform(0, box, (40, 3)))
    xv = S.place_molecules(rng, wt, rng.uniform(0, box, (700, 3)))
    xu = np.stack([xu + rng.normal(scale=0.3, size=xu.shape) for _ in range(3)])
    xv = np.stack([xv + rng.normal(scale=0.3, size=xv.shape) for _ in range(3)])
    su = pkg.AtomSelection(list(range(1, 121)), natomspermol=3)
    sv = pkg.AtomSelection(list(range(121, 121 + 2100)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([box] * 3)[None], 3, axis=0), su, sv)
    for opts in (pkg.Options(dbulk=9.0, n_random_samples=7, seed=3, silent=True),
                 pkg.Options(bulk_range=(6.0, 9.0), n_random_samples=7, seed=3, silent=True)):
        a, _ = oracle_raw(pkg, oracle, traj, opts)
        b = gpu_raw(pkg, traj, opts)
        assert_raw_equal(a, b)
        assert_group_split(b, False)


def test_parity_custom_groups_and_usecutoff(pkg, oracle):
    s = syn().config1(nframes=2, n_solute=150, n_water=500, box=26.0, seed=21)
    n_at = 150
    g1 = list(range(1, 60))
    g2 = list(range(40, 151))          # overlaps g1: atoms count in both groups
    g3 = [3, 77, 149]
    su = pkg.AtomSelection(list(range(1, n_at + 1)), nmols=1, group_atom_indices=[g1, g2, g3],
                           group_names=["a", "b", "c"])
    ov = list(range(n_at + 1, n_at + 1501, 3))
    hv = [i for i in range(n_at + 1, n_at + 1501) if i not in set(ov)]
    sv = pkg.AtomSelection(list(range(n_at + 1, n_at + 1501)), natomspermol=3,
                           group_atom_indices=[ov, hv[:500]], group_names=["O", "Hsome"])
    traj = pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], su, sv)
    for opts in (pkg.Options(bulk_range=(6.0, 9.0), n_random_samples=5, seed=4, silent=True),
                 pkg.Options(dbulk=9.0, n_random_samples=5, seed=4, silent=True, irefatom=2)):
        a, _ = oracle_raw(pkg, oracle, traj, opts)
        b = gpu_raw(pkg, traj, opts)
        assert_raw_equal(a, b)


def test_parity_weights_stride_coordination(pkg, oracle):
    s = syn().config1(nframes=6, n_solute=200, n_water=600, box=26.0, seed=8)
    traj = mem(pkg, s)
    w = [2.0, 0.5, 0.0, 1.0, 0.25, 4.0]
    opts = pkg.Options(dbulk=8.0, n_random_samples=4, seed=15, silent=True, firstframe=1, stride=1)
    a, _ = oracle_raw(pkg, oracle, traj, opts, frame_weights=w)
    b = gpu_raw(pkg, traj, opts, frame_weights=w)
    assert_raw_equal(a, b)                     # dyadic weights: exact
    w2 = [0.3, 0.7, 0.1, 0.9, 0.3, 0.3]
    a, _ = oracle_raw(pkg, oracle, traj, opts, frame_weights=w2)
    b = gpu_raw(pkg, traj, opts, frame_weights=w2)
    assert_raw_equal(a, b, rtol=1e-13)         # integer hits times the weight vs sequential sums
    opts = pkg.Options(dbulk=8.0, n_random_samples=4, seed=15, silent=True, firstframe=2, lastframe=6, stride=2)
    for co in (False, True):
        a, _ = oracle_raw(pkg, oracle, traj, opts, coordination_number_only=co)
        b = gpu_raw(pkg, traj, opts, coordination_number_only=co)
        assert_raw_equal(a, b)


def test_batching_does_not_change_results(pkg):
    """Results depend on the frame numbers only, not on how frames are batched."""
    s = syn().config1(nframes=5, n_solute=200, n_water=600, box=26.0, seed=8)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=8.0, n_random_samples=4, seed=15, silent=True)
    a = gpu_raw(pkg, traj, opts)
    b = gpu_raw(pkg, traj, opts, batch_frames=2)
    c = gpu_raw(pkg, traj, opts, job_kw={"max_batch_frames": 1})
    assert_raw_equal(a, b)
    assert_raw_equal(a, c)


def test_exact_ties(pkg, oracle):
    """Equidistant atoms (exact FP ties) resolve like the oracle: smallest (i, j)."""
    # solute: two atoms symmetric about the x axis; solvent: waters on the axis
    xu = np.array([[[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]]])
    wt = np.array([[4.0, 0.0, 0.0], [4.5, 0.5, 0.5], [4.5, -0.5, 0.5],
                   [0.0, 0.0, 5.0], [0.5, 0.5, 5.5], [-0.5, 0.5, 5.5]])
    xv = wt[None]
    su = pkg.AtomSelection([1, 2], nmols=1)
    sv = pkg.AtomSelection([3, 4, 5, 6, 7, 8], natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.diag([20.0] * 3)[None], su, sv)
    opts = pkg.Options(dbulk=8.0, n_random_samples=2, seed=1, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert b["solute_group_count"][0].sum() == 2 and b["solute_group_count"][1].sum() == 0


@pytest.mark.parametrize("tuning", [dict(cell_fraction=0.5, columns=1), dict(cell_fraction=0.3, columns=2),
                                    dict(cell_fraction=1.0, columns=4), dict(cell_fraction=0.25, columns=8),
                                    dict(cell_fraction=0.2, columns=1, eval_workgroups=7),
                                    dict(cell_fraction=0.5, columns=4, eval_workgroups=1),
                                    dict(cell_fraction=0.35, columns=2, eval_workgroups=13)])
def test_tuning_does_not_change_results(pkg, oracle, tuning):
    """Cell size, columns and grid of the single-solute kernels change the work
    decomposition, never the records.  Small grids give each workgroup many
    chunks, so the chunks handed out on request at the end of a launch (several
    per request) are most of the work: every chunk must still be evaluated
    exactly once."""
    s = syn().config1(nframes=2, n_solute=300, n_water=1200, box=30.0, seed=41)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=9.0, n_random_samples=4, seed=6, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts, job_kw={"tuning": tuning})
    assert_raw_equal(a, b)


def test_dense_solute_staging_rounds(pkg, oracle):
    """A solute dense enough that a chunk's candidates exceed one LDS staging
    round, and a cell holding more atoms than the staging buffer."""
    S = syn()
    rng = np.random.default_rng(5)
    box = 30.0
    clump = 15.0 + rng.uniform(-1.5, 1.5, size=(1500, 3))       # > 1024 atoms in one 10 A cell
    shell = 15.0 + rng.normal(scale=4.0, size=(600, 3))
    xu = np.concatenate([clump, shell])[None]
    xv = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (900, 3)))[None]
    su = pkg.AtomSelection(list(range(1, 2101)), nmols=1)
    sv = pkg.AtomSelection(list(range(2101, 2101 + 2700)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.diag([box] * 3)[None], su, sv)
    opts = pkg.Options(dbulk=10.0, n_random_samples=3, seed=8, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    for tuning in (dict(cell_fraction=1.0, columns=1), dict(cell_fraction=0.5, columns=4),
                   dict(cell_fraction=0.5, columns=8)):
        b = gpu_raw(pkg, traj, opts, job_kw={"tuning": tuning})
        assert_raw_equal(a, b)


def _edge_distances():
    # bin edges (multiples of binstep), dbulk, the cutoff, and values one ulp-ish around them
    base = [2.0, 2.02, 3.0, 4.46, 6.0, 8.0, 9.98, 10.0]
    out = []
    for d in base:
        out += [d, np.nextafter(d, 0.0), np.nextafter(d, 20.0), d + 1e-9, d - 1e-9]
    return out


def test_bin_edges_cutoff_and_dbulk(pkg, oracle):
    """Minimum distances exactly on and next to bin edges, dbulk and the
    cutoff (single-atom solute, waters along the axes, FP64 input): the bins,
    cutoff and bulk decisions taken from FP32 bounds or FP64 recomputation
    match the oracle."""
    c = np.array([15.0, 15.0, 15.0])
    axes = [np.array(v, float) for v in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1])]
    ds = _edge_distances()
    nframes = (len(ds) + 5) // 6
    xv = np.zeros((nframes, 6 * 3, 3))
    for f in range(nframes):
        for k, e in enumerate(axes):
            d = ds[(f * 6 + k) % len(ds)]
            p = np.roll(e, 1)
            o = c + d * e
            xv[f, 3 * k] = o
            xv[f, 3 * k + 1] = c + (d + 0.8) * e + 0.6 * p
            xv[f, 3 * k + 2] = c + (d + 0.8) * e - 0.6 * p
    xu = np.repeat(c[None, None], nframes, axis=0)
    su = pkg.AtomSelection([1], nmols=1)
    sv = pkg.AtomSelection(list(range(2, 20)), natomspermol=3)
    cells = np.repeat(np.diag([40.0] * 3)[None], nframes, axis=0)
    traj = pkg.MemoryTrajectory(xu, xv, cells, su, sv)
    for opts in (pkg.Options(dbulk=8.0, cutoff=10.0, usecutoff=True, n_random_samples=3, seed=3, silent=True),
                 pkg.Options(dbulk=10.0, n_random_samples=2, seed=4, silent=True)):
        a, _ = oracle_raw(pkg, oracle, traj, opts)
        b = gpu_raw(pkg, traj, opts)
        assert_raw_equal(a, b)


def test_bin_edges_self(pkg, oracle):
    """The same edge distances between single-atom molecules of a self-MDDF
    (many-solute-molecule kernels): per frame a centre atom and six atoms at
    the listed distances along the axes."""
    ds = _edge_distances()
    c = np.array([20.0, 20.0, 20.0])
    axes = [np.array(v, float) for v in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1])]
    nframes = (len(ds) + 5) // 6
    xv = np.zeros((nframes, 7, 3))
    for f in range(nframes):
        xv[f, 0] = c
        for k, e in enumerate(axes):
            xv[f, 1 + k] = c + ds[(f * 6 + k) % len(ds)] * e
    sel = pkg.AtomSelection(list(range(1, 8)), natomspermol=1)
    cells = np.repeat(np.diag([40.0] * 3)[None], nframes, axis=0)
    traj = pkg.MemoryTrajectory(xv, xv, cells, sel, sel)
    opts = pkg.Options(dbulk=8.0, cutoff=10.0, usecutoff=True, n_random_samples=2, seed=5, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_molecule_pair_queue_cross(pkg, oracle):
    """Many solute molecules in a box with three molecule cells per side:
    the queued pair evaluation (cross, with ideal-gas passes)."""
    S = syn()
    rng = np.random.default_rng(31)
    box = 40.0
    wt = S.water_template()
    xu = S.place_molecules(rng, wt, rng.uniform(0, box, (60, 3)))
    xv = S.place_molecules(rng, wt, rng.uniform(0, box, (1500, 3)))
    xu = np.stack([xu + rng.normal(scale=0.3, size=xu.shape) for _ in range(2)])
    xv = np.stack([xv + rng.normal(scale=0.3, size=xv.shape) for _ in range(2)])
    su = pkg.AtomSelection(list(range(1, 181)), natomspermol=3)
    sv = pkg.AtomSelection(list(range(181, 181 + 4500)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([box] * 3)[None], 2, axis=0), su, sv)
    opts = pkg.Options(dbulk=10.0, n_random_samples=5, seed=13, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_molecule_pair_queue_self_many_groups(pkg, oracle):
    """Self-MDDF of 24-atom rigid molecules (24 groups: counts too large for
    the LDS copy, so they go straight to global memory) in a box with three
    molecule cells per side, and the same molecules as water-sized groups."""
    S = syn()
    rng = np.random.default_rng(32)
    mol = S.random_walk(rng, 24, bond=1.5, rmin=3.0, start=np.zeros(3))
    box = 66.0
    x = S.place_molecules(rng, mol, rng.uniform(0, box, (260, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    sel = pkg.AtomSelection(list(range(1, 24 * 260 + 1)), natomspermol=24)
    traj = pkg.MemoryTrajectory(xv, xv, np.repeat(np.diag([box] * 3)[None], 2, axis=0), sel, sel)
    opts = pkg.Options(dbulk=10.0, n_random_samples=3, seed=14, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_group_split(b, True)


@pytest.mark.parametrize("nv", [64, 65])
def test_large_solvent_molecules(pkg, oracle, nv):
    """Solvent molecules of 64 atoms (the largest the single-solute tile
    kernels take) and of 65 atoms (molecule-pair kernels with one solute
    molecule)."""
    S = syn()
    rng = np.random.default_rng(40 + nv)
    mol = S.random_walk(rng, nv, bond=1.2, rmin=1.5, start=np.zeros(3))
    box = 60.0
    xv = S.place_molecules(rng, mol, rng.uniform(0, box, (40, 3)))
    xu = S.random_walk(rng, 120, bond=1.5, rmin=3.0, start=np.full(3, box / 2))
    xu = np.stack([xu + rng.normal(scale=0.2, size=xu.shape) for _ in range(2)])
    xv = np.stack([xv + rng.normal(scale=0.2, size=xv.shape) for _ in range(2)])
    su = pkg.AtomSelection(list(range(1, 121)), nmols=1)
    sv = pkg.AtomSelection(list(range(121, 121 + 40 * nv)), natomspermol=nv)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([box] * 3)[None], 2, axis=0), su, sv)
    opts = pkg.Options(dbulk=10.0, n_random_samples=4, seed=41, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_ions_and_empty_frames(pkg, oracle):
    """Single-atom solute and solvent, with a frame where no solvent atom is
    near the solute and one where the only solvent atom sits on the solute."""
    xu = np.array([[[10.0, 10.0, 10.0]], [[10.0, 10.0, 10.0]], [[10.0, 10.0, 10.0]]])
    xv = np.array([[[35.0, 35.0, 35.0], [36.0, 35.0, 35.0]],
                   [[10.0, 10.0, 10.0], [35.0, 35.0, 35.0]],
                   [[13.0, 10.0, 10.0], [10.0, 18.5, 10.0]]])
    su = pkg.AtomSelection([1], nmols=1)
    sv = pkg.AtomSelection([2, 3], natomspermol=1)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([50.0] * 3)[None], 3, axis=0), su, sv)
    for opts in (pkg.Options(dbulk=10.0, n_random_samples=3, seed=42, silent=True),
                 pkg.Options(dbulk=6.0, cutoff=10.0, usecutoff=True, n_random_samples=2, seed=43, silent=True)):
        a, _ = oracle_raw(pkg, oracle, traj, opts)
        b = gpu_raw(pkg, traj, opts)
        assert_raw_equal(a, b)


def test_self_coordination_numbers(pkg, oracle):
    """Self coordination numbers (no ideal-gas passes) of waters."""
    S = syn()
    rng = np.random.default_rng(44)
    box = 34.0
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (1200, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    sel = pkg.AtomSelection(list(range(1, 3601)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xv, xv, np.repeat(np.diag([box] * 3)[None], 2, axis=0), sel, sel)
    opts = pkg.Options(dbulk=10.0, seed=45, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts, coordination_number_only=True)
    b = gpu_raw(pkg, traj, opts, coordination_number_only=True)
    assert_raw_equal(a, b)
    assert_group_split(b, True)


@pytest.mark.parametrize("irefatom", [2, 3])
def test_self_water_reference_atom(pkg, oracle, irefatom):
    """Self-MDDF of waters with a hydrogen as the reference atom: the rdf
    counts and the bulk flags of the ideal-gas passes come from that atom in
    both orientations of every pair, against the oracle."""
    S = syn()
    rng = np.random.default_rng(50 + irefatom)
    box = 34.0
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (1200, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    sel = pkg.AtomSelection(list(range(1, 3601)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xv, xv, np.repeat(np.diag([box] * 3)[None], 2, axis=0), sel, sel)
    opts = pkg.Options(dbulk=9.0, n_random_samples=3, seed=52, silent=True, irefatom=irefatom)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_parity_self_urea_reference_atom(pkg, oracle):
    """Self-MDDF of 8-atom molecules with the sixth atom as the reference
    atom, against the oracle."""
    s = syn().config3(nframes=2, n_solute=10, n_urea=600, n_water=10, solvent="self", seed=5, a=36.0)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=9.0, n_random_samples=3, seed=11, silent=True, irefatom=6)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)


def test_self_water_half_shell_with_random_passes(pkg, oracle):
    """Self-MDDF of waters with three or more cells per dimension (the
    half-shell walk of the molecule-pair kernel, both orientations from one
    evaluation, FP64 parts resolved from the deferred queue) and ideal-gas
    passes (the dbulk flags written for both orientations), against the
    oracle."""
    S = syn()
    rng = np.random.default_rng(46)
    box = 40.0
    n = 2100
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (n, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(3)])
    sel = pkg.AtomSelection(list(range(1, 3 * n + 1)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xv, xv, np.repeat(np.diag([box] * 3)[None], 3, axis=0), sel, sel)
    opts = pkg.Options(dbulk=10.0, n_random_samples=4, seed=47, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_group_split(b, True)


def test_self_water_walk_split_does_not_change_results():
    """Full-size self-MDDF water (100 A box, 33,400 molecules): one job over
    8 frames walks one lane per molecule; 8 one-frame jobs split each
    molecule's neighbour walk over several lanes.  The summed counts must be
    identical (integer hits, unit weights)."""
    from complexmixtures_amd import _lib
    import complexmixtures_amd.serve as SV
    from complexmixtures_amd.synthetic import served_system
    nf = 8
    s = served_system(7, 2000, 100.0, nf)
    du = _lib.DeviceArray(s["x_solute"])
    dw = _lib.DeviceArray(s["x_solvent"])
    sysm = SV.System(n_solute=2000, n_water=s["n_water"], box=100.0, nframes=nf, d_solute=du, d_water=dw,
                     cells=s["cells"], irefatom=SV.default_irefatom(s["x_solvent"][0]))
    kw, _, _, _ = SV.job_kwargs(sysm, SV.Request(0.0, 0, nf, False, True, 99))
    job = _lib.Job(**kw)
    job.submit_device(np.arange(1, nf + 1), dw, dw, 0, s["cells"], np.ones(nf))
    together = job.finish()
    job.close()
    parts = None
    for f in range(nf):
        job = _lib.Job(**kw)
        job.submit_device(np.array([f + 1]), dw, dw, f, s["cells"][f:f + 1], np.ones(1))
        raw = job.finish()
        job.close()
        parts = raw if parts is None else {k: parts[k] + raw[k] for k in raw}
    du.free()
    dw.free()
    assert float(together["md_count"].sum()) > 1e6
    for k in together:
        if k == "volume_total":
            assert np.isclose(together[k], parts[k], rtol=1e-14)
        else:
            assert np.array_equal(together[k], parts[k]), k


def test_cached_memory_reuse_and_release(pkg):
    """Jobs take the device and page-locked blocks that earlier jobs left
    cached; the counts do not depend on whether a block was reused or fresh
    (released in between)."""
    from complexmixtures_amd import _lib
    s = syn().config1(nframes=3)
    traj = mem(pkg, s)
    opts = pkg.Options(dbulk=10.0, n_random_samples=3, seed=48, silent=True)
    a = gpu_raw(pkg, traj, opts)
    b = gpu_raw(pkg, traj, opts)
    _lib.release_cached_memory()
    c = gpu_raw(pkg, traj, opts)
    assert_raw_equal(a, b)
    assert_raw_equal(a, c)


@pytest.mark.parametrize("self_mddf", [False, True])
def test_packed_and_fp64_readback_agree(pkg, oracle, self_mddf):
    """cm_job_finish sends the hits of a job whose frames share one weight as
    32-bit words and forms (double)hits * w on the host (w/2 in the self-MDDF
    group range); a pack limit of 0 sends the same job down the FP64 path
    after the packing (the md/rdf hits added once).  Both read-backs agree bit
    for bit, and with the oracle (exactly for a dyadic weight, to 1e-13 for
    0.3, whose per-record sums round differently from the product)."""
    from complexmixtures_amd import _lib
    S = syn()
    if self_mddf:
        s = S.config3(nframes=2, n_solute=10, n_urea=600, n_water=10, solvent="self", seed=4, a=36.0)
        opts = pkg.Options(dbulk=9.0, n_random_samples=3, seed=10, silent=True)
    else:
        s = S.config1(nframes=3, n_solute=300, n_water=900, box=30.0, seed=5)
        opts = pkg.Options(dbulk=8.0, n_random_samples=6, seed=77, silent=True)
    traj = mem(pkg, s)
    nf = s["x_solute"].shape[0]
    for w in (0.75, 0.3):
        fw = np.full(nf, w)
        packed = gpu_raw(pkg, traj, opts, frame_weights=fw)
        _lib.set_readback_pack_limit(0)
        try:
            wide = gpu_raw(pkg, traj, opts, frame_weights=fw)
        finally:
            _lib.set_readback_pack_limit()
        assert_raw_equal(packed, wide)
        for k in ("hits_md", "hits_md_random", "hits_rdf", "hits_rdf_random"):
            assert np.array_equal(packed[k], wide[k]), k
        a, _ = oracle_raw(pkg, oracle, traj, opts, frame_weights=fw)
        assert_raw_equal(a, packed, rtol=0.0 if w == 0.75 else 1e-13)
        if w == 0.75:
            assert_group_split(packed, self_mddf)
        assert packed["md_count"].sum() > 0


def test_packed_readback_overflow_list(pkg, oracle):
    """Group counts of 65,535 and more go out beside the 16-bit words of the
    packed read-back: a point solute with 5 A bins and 600 ideal-gas passes
    puts about 70,000 hits in one solute-group bin.  The packed read-back
    equals the FP64 one and the oracle."""
    from complexmixtures_amd import _lib
    S = syn()
    rng = np.random.default_rng(61)
    box, n, nf = 30.0, 900, 60
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (n, 3)))
    xv = np.stack([x + rng.normal(scale=0.3, size=x.shape) for _ in range(nf)])
    xu = np.full((nf, 1, 3), box / 2)
    su = pkg.AtomSelection([1], nmols=1)
    sv = pkg.AtomSelection(list(range(2, 3 * n + 2)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([box] * 3)[None], nf, axis=0), su, sv)
    opts = pkg.Options(binstep=5.0, n_random_samples=10, seed=62, silent=True)
    packed = gpu_raw(pkg, traj, opts)
    assert packed["solute_group_count_random"].max() >= 65535
    _lib.set_readback_pack_limit(0)
    try:
        wide = gpu_raw(pkg, traj, opts)
    finally:
        _lib.set_readback_pack_limit()
    assert_raw_equal(packed, wide)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    assert_raw_equal(a, packed)


def test_large_cell_grid_direct_near_marks(pkg, oracle):
    """A 330 A box with 5 A cells has 287,496 cells, more than the LDS bitmap
    of the near-cell marks holds: the marks go straight to memory.  Counts
    against the oracle, ideal-gas passes included."""
    S = syn()
    rng = np.random.default_rng(63)
    box, n, nf = 330.0, 3000, 2
    xu0 = box / 2 + rng.normal(scale=4.0, size=(60, 3))
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (n, 3)))
    xu = np.stack([xu0 + rng.normal(scale=0.2, size=xu0.shape) for _ in range(nf)])
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(nf)])
    su = pkg.AtomSelection(list(range(1, 61)), nmols=1)
    sv = pkg.AtomSelection(list(range(61, 61 + 3 * n)), natomspermol=3)
    traj = pkg.MemoryTrajectory(xu, xv, np.repeat(np.diag([box] * 3)[None], nf, axis=0), su, sv)
    opts = pkg.Options(n_random_samples=40, seed=64, silent=True)
    a, _ = oracle_raw(pkg, oracle, traj, opts)
    b = gpu_raw(pkg, traj, opts, job_kw=dict(tuning=dict(cell_fraction=0.5, columns=4)))
    assert_raw_equal(a, b)
    assert b["md_count_random"].sum() > 0
