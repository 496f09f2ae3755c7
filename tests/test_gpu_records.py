"""Record-level parity through the verification getters of the C ABI
(cm_job_set_verify, cm_job_records, cm_job_placement,
cm_job_reference_draws): every MinimumDistance record (within, i, j, the FP64
bits of d, ref flag, d_ref) of the real pass and of random passes, the placed
solvent coordinates of the random passes, and the drawn reference solute
molecules, against the FP64 oracle (orc_records, orc_placement,
orc_ref_solute) on the reference's toy system, the NAMD test trajectory and the
40 A chain / water system."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle_job

pytestmark = pytest.mark.gpu


def _job(pkg, traj, opts, coordination_number_only=False):
    from complexmixtures_amd import _lib
    from complexmixtures_amd.mddf import TrajectoryMetaData, job_kwargs
    from complexmixtures_amd.results import make_result
    tm = TrajectoryMetaData(traj, opts)
    R = make_result(traj, opts, tm)
    kw = job_kwargs(R, opts, tm, coordination_number_only)
    job = _lib.Job(**kw)
    job.set_verify(True)
    return job, kw, tm


def _frames(traj, tm):
    from complexmixtures_amd.trajectory import convert_unitcell
    if hasattr(traj, "xu"):
        xu, xv, cells = traj.xu, traj.xv, traj.cells
    else:
        traj.opentraj()
        traj.firstframe()
        xu, xv, cells = traj.read_frames(tm.lastframe_read)
        traj.closetraj()
    cells = np.stack([convert_unitcell(c, tm.orthorhombic) for c in cells])
    return np.asarray(xu, np.float64), np.asarray(xv, np.float64), cells


def _inbulk(rec, usecutoff, dbulk):
    within, i, j, d, ref, dref = rec
    return (within and d > dbulk) if usecutoff else not within


def _check_job(pkg, oracle, traj, opts, frames, passes):
    job, kw, tm = _job(pkg, traj, opts)
    xu, xv, cells = _frames(traj, tm)
    nf = len(xu)
    job.submit(np.arange(1, nf + 1), xu, xv, cells, np.ones(nf))
    job.finish()
    O, _ = oracle_job(pkg, oracle, traj, opts)
    Mv, nv = kw["nmols_v"], kw["natoms_per_mol_v"]
    nchecked = nplaced = 0
    for fr in frames:
        f0 = fr - 1
        ref_real = oracle.records(O, xu[f0], xv[f0], cells[f0], 0)
        got_real = job.records(fr, 0, Mv)
        assert got_real == ref_real, f"frame {fr} real pass"
        nchecked += sum(r[0] for r in ref_real)
        bulk = [v for v, r in enumerate(ref_real) if _inbulk(r, opts.usecutoff, opts.dbulk)]
        for p in passes:
            placed, flag = job.placement(fr, p, Mv, nv)
            ref_pl = oracle.placement(O, xv[f0], cells[f0], bulk, fr, p - 1)
            sel = np.repeat(flag, nv)
            nplaced += int(flag.sum())
            assert np.array_equal(placed[sel], ref_pl[sel]), f"frame {fr} pass {p} placement"
            assert np.all(np.isnan(placed[~sel]))
            ref_rand = oracle.records(O, xu[f0], ref_pl, cells[f0], 0)
            got_rand = job.records(fr, p, Mv)
            # slots the device did not place cannot reach the cutoff
            assert all(not ref_rand[v][0] for v in range(Mv) if not flag[v])
            assert got_rand == ref_rand, f"frame {fr} pass {p} records"
            nchecked += sum(r[0] for r in ref_rand)
        assert np.array_equal(job.reference_draws(fr, kw["n_random_samples"]), np.zeros(kw["n_random_samples"]))
    job.close()
    assert nplaced > 0
    return nchecked


def test_records_toy_cross(pkg, oracle):
    """The reference's toy system (src/mddf.jl:587-624): one atom vs three waters."""
    protein = pkg.AtomSelection([10], nmols=1)
    water = pkg.AtomSelection(list(range(1, 10)), natomspermol=3)
    traj = pkg.Trajectory(os.path.join(GOLDEN, "toy", "cross.pdb"), protein, water, format="PDBTraj")
    opts = pkg.Options(seed=321, silent=True, n_random_samples=50)
    assert _check_job(pkg, oracle, traj, opts, [1, 2], [1, 2, 50]) > 0


def test_records_dupdcd(pkg, oracle):
    """NAMD test trajectory, protein vs TMAO (14-atom molecules), frames 1-3."""
    z = np.load(os.path.join(GOLDEN, "dupdcd_atoms_1_4012.npz"))
    cells = np.stack([pkg.unitcell_matrix(*c[[0, 2, 5]]) for c in z["dcd_cell"]])
    protein = pkg.AtomSelection(list(range(1, 1464)), nmols=1)
    tmao = pkg.AtomSelection(list(range(1479, 4013)), natomspermol=14)
    traj = pkg.MemoryTrajectory(z["xyz"][:, 0:1463].astype(np.float64), z["xyz"][:, 1478:4012].astype(np.float64),
                                cells, protein, tmao)
    opts = pkg.Options(seed=321, silent=True, n_random_samples=3)
    assert _check_job(pkg, oracle, traj, opts, [1, 3], [1, 3]) > 100


@pytest.mark.parametrize("usecutoff", [False, True])
def test_records_chain_water(pkg, oracle, usecutoff):
    """40 A box, 1,000-atom chain, 2,000 waters (3-atom molecules: one molecule per lane)."""
    import complexmixtures_amd.synthetic as S
    s = S.config1(nframes=2)
    traj = pkg.MemoryTrajectory(s["x_solute"], s["x_solvent"], s["cells"], s["solute"], s["solvent"])
    opts = (pkg.Options(bulk_range=(7.0, 10.0), n_random_samples=4, seed=9, silent=True, irefatom=2) if usecutoff
            else pkg.Options(dbulk=10.0, n_random_samples=4, seed=9, silent=True))
    assert _check_job(pkg, oracle, traj, opts, [1, 2], [1, 4]) > 1000


def test_reference_draws_many_solute_molecules(pkg, oracle):
    """Self-MDDF (molecule-pair kernels): the drawn reference molecule of
    every random pass equals the oracle's draw, and the placed solvent of a
    pass equals orc_placement for that frame's bulk list."""
    import complexmixtures_amd.synthetic as S
    rng = np.random.default_rng(62)
    box = 24.0
    x = S.place_molecules(rng, S.water_template(), rng.uniform(0, box, (200, 3)))
    xv = np.stack([x + rng.normal(scale=0.2, size=x.shape) for _ in range(2)])
    sel = pkg.AtomSelection(list(range(1, 601)), natomspermol=3)
    cells = np.repeat(np.diag([box] * 3)[None], 2, axis=0)
    traj = pkg.MemoryTrajectory(xv, xv, cells, sel, sel)
    opts = pkg.Options(dbulk=8.0, n_random_samples=5, seed=63, silent=True)
    job, kw, tm = _job(pkg, traj, opts)
    job.submit(np.arange(1, 3), xv, xv, cells, np.ones(2))
    job.finish()
    O, _ = oracle_job(pkg, oracle, traj, opts)
    for fr in (1, 2):
        d = job.reference_draws(fr, 5)
        assert [int(x) for x in d] == [oracle.ref_solute(kw["seed"], fr, k, 200) for k in range(5)]
        for k in (0, 4):
            u = int(d[k])
            recs = oracle.records(O, xv[fr - 1], xv[fr - 1], cells[fr - 1], u)
            # the bulk pool of the drawn molecule excludes the molecule itself (src/mddf.jl:406-415)
            bulk = [v for v, r in enumerate(recs) if not r[0] and v != u]
            placed, flag = job.placement(fr, k + 1, 200, 3)
            assert flag.all()
            assert np.array_equal(placed, oracle.placement(O, xv[fr - 1], cells[fr - 1], bulk, fr, k))
    with pytest.raises(Exception, match="one solute molecule"):
        job.records(1, 0, 200)
    job.close()
