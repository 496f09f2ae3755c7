"""CPU tests of the host layer: options, selections, readers, metadata, Result
layout, the C ABI exports of libcmgpu.so (loaded, not called: no GPU here)."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REFERENCE, ROOT

TOY = os.path.join(GOLDEN, "toy")


def test_options_defaults_and_errors(pkg):
    """src/Options.jl:470-513"""
    O, E = pkg.Options, pkg.ArgumentError
    o = O(silent=True)
    assert (o.dbulk, o.cutoff, o.usecutoff) == (10.0, 10.0, False)
    o = O(bulk_range=(10.0, 14.0))
    assert (o.dbulk, o.cutoff, o.usecutoff) == (10.0, 14.0, True)
    o = O(dbulk=10.0, silent=True)
    assert (o.dbulk, o.cutoff, o.usecutoff) == (10.0, 10.0, False)
    o = O(dbulk=10.0, usecutoff=False)
    assert (o.dbulk, o.cutoff, o.usecutoff) == (10.0, 10.0, False)
    assert O(cutoff=12.0, usecutoff=True, silent=True).cutoff == 12.0
    assert O(dbulk=10.0, usecutoff=True, silent=True).cutoff == 14.0
    bad = [dict(dbulk=10.0, binstep=0.3), dict(dbulk=10.0, cutoff=10.0, usecutoff=True),
           dict(dbulk=6.0, cutoff=10.0, binstep=0.3, usecutoff=True),
           dict(dbulk=8.0, cutoff=10.0, binstep=0.3, usecutoff=True), dict(bulk_range=(10.0, 12.0), binstep=0.3),
           dict(stride=0), dict(lastframe=1, firstframe=2), dict(bulk_range=(12.0, 10.0)),
           dict(bulk_range=(12.0, 10.0), dbulk=8.0), dict(bulk_range=(12.0, 10.0), cutoff=15.0),
           dict(bulk_range=(12.0, 10.0), usecutoff=False), dict(dbulk=10.0, cutoff=15.0, usecutoff=False),
           dict(n_random_samples=0), dict(bulk_range=[1.0]), dict(cutoff=12.0)]
    for kw in bad:
        with pytest.raises(E):
            O(silent=True, **kw)
    assert O(silent=True) == O(silent=True)
    assert O.from_dict(o.to_dict()) == o


def test_atom_selection(pkg):
    """src/AtomSelection.jl:5-37, 290-372"""
    A, E = pkg.AtomSelection, pkg.ArgumentError
    s = A(list(range(1, 13)), natomspermol=3)
    assert (s.nmols, s.natomspermol, s.custom_groups, s.ngroups) == (4, 3, False, 3)
    s = A(list(range(1, 13)), nmols=2)
    assert (s.nmols, s.natomspermol) == (2, 6)
    for kw in (dict(), dict(nmols=5), dict(natomspermol=5)):
        with pytest.raises(E):
            A(list(range(1, 13)), **kw)
    with pytest.raises(E):
        A([], nmols=1)
    with pytest.raises(E):
        A([1, 2, 3], nmols=1, group_names=["a"])
    g = A([1, 2, 3], natomspermol=1, group_atom_indices=[[1, 2], [3]], group_names=["G1", "G2"])
    assert g.custom_groups and g.ngroups == 2
    with pytest.raises(E):
        A([1, 2, 3], natomspermol=1, group_atom_indices=[[1, 1]], group_names=["a"])
    with pytest.raises(E):
        A([1, 2, 3], natomspermol=1, group_atom_indices=[[4]], group_names=["a"])
    with pytest.raises(E):
        A([1, 2, 3], natomspermol=1, group_atom_indices=[[1], [2]], group_names=["a"])
    with pytest.warns(UserWarning):
        w = A([1, 2, 3], natomspermol=1, group_atom_indices=[[2, 1]], group_names=["a"])
    assert w.group_atom_indices == [[1, 2]]


def test_group_map_quirks(pkg):
    """update_group_count!, src/update_counters.jl:22-36: the solute map uses
    indices[i] with a molecule-local i (molecule 1's atoms); an atom may be in
    several groups."""
    from complexmixtures_amd.selection import group_csr
    s = pkg.AtomSelection([5, 6, 7, 8, 9, 10], nmols=2, group_atom_indices=[[5, 9], [6, 9], [10]],
                          group_names=["a", "b", "c"])
    ptr, idx = group_csr(s, [0, 1, 2], s)
    assert [list(idx[ptr[k]:ptr[k + 1]]) for k in range(3)] == [[0], [1], []]
    ptr, idx = group_csr(s, list(range(6)), s)
    assert [list(idx[ptr[k]:ptr[k + 1]]) for k in range(6)] == [[0], [1], [], [], [0, 1], [2]]
    assert group_csr(pkg.AtomSelection([1, 2], nmols=1), [0, 1], pkg.AtomSelection([1, 2], nmols=1)) == (None, None)


def test_unitcell_conversion(pkg):
    """src/Trajectory.jl:83-102 and the Chemfiles cell convention"""
    from complexmixtures_amd.trajectory import convert_unitcell, is_orthorhombic
    assert is_orthorhombic(np.eye(3))
    assert not is_orthorhombic(np.array([[1.0, 1.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 2.0]]))
    np.testing.assert_array_equal(convert_unitcell([1.0, 2.0, 3.0]), np.diag([1.0, 2.0, 3.0]))
    m = pkg.unitcell_matrix(30.0, 30.0, 30.0)
    np.testing.assert_array_equal(m, np.diag([30.0] * 3))
    m = pkg.unitcell_matrix(90.0, 90.0, 90.0, 80.0, 80.0, 80.0)
    for k in range(3):
        assert math.isclose(np.linalg.norm(m[:, k]), 90.0)
    cos = lambda u, v: u @ v / np.linalg.norm(u) / np.linalg.norm(v)
    assert math.isclose(cos(m[:, 1], m[:, 2]), math.cos(math.radians(80.0)))
    assert m[1, 0] == m[2, 0] == m[2, 1] == 0.0


def test_pdb_reader(pkg):
    """src/trajectory_formats/PDBTraj.jl on the toy files"""
    t = pkg.Trajectory(os.path.join(TOY, "cross.pdb"), pkg.AtomSelection([10], nmols=1),
                       pkg.AtomSelection(list(range(1, 10)), natomspermol=3), format="PDBTraj")
    assert t.nframes == 2
    t.opentraj()
    t.firstframe()
    t.nextframe()
    np.testing.assert_array_equal(t.x_solute, [[0.0, 0.0, 0.0]])
    np.testing.assert_array_equal(t.x_solvent[3], [11.0, 0.0, 0.0])
    np.testing.assert_array_equal(t.getunitcell(), np.diag([30.0] * 3))
    t.closetraj()
    with pytest.raises(pkg.ArgumentError):
        pkg.Trajectory("x.pdb", pkg.AtomSelection([1], nmols=1), format="XXX")


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference checkout not present")
def test_dcd_reader_matches_fixture(pkg):
    """NAMD DCD reader (src/trajectory_formats/NamdDCD.jl) on the reference's
    duplicated-first-frame trajectory: 62,026 atoms, 3 frames, 84.42188 A cube."""
    path = os.path.join(REFERENCE, "test", "data", "NAMD", "traj_duplicated_first_frame.dcd")
    z = np.load(os.path.join(GOLDEN, "dupdcd_atoms_1_4012.npz"))
    prot = pkg.AtomSelection(list(range(1, 1464)), nmols=1)
    tmao = pkg.AtomSelection(list(range(1479, 4013)), natomspermol=14)
    t = pkg.Trajectory(path, prot, tmao)
    assert t.nframes == 3 and t.natoms_file == 62026
    t.opentraj()
    t.firstframe()
    for f in range(3):
        t.nextframe()
        np.testing.assert_array_equal(t.x_solute, z["xyz"][f, :1463].astype(np.float64))
        np.testing.assert_array_equal(t.x_solvent, z["xyz"][f, 1478:4012].astype(np.float64))
        np.testing.assert_array_equal(np.diag(t.getunitcell()), z["dcd_cell"][f][[0, 2, 5]])
    assert z["dcd_cell"][0][0] == 84.42188262939453
    t.closetraj()
    assert pkg.Trajectory(path, prot, tmao, lastframe=2).nframes == 2


def _dup_traj(pkg, nframes=3):
    z = np.load(os.path.join(GOLDEN, "dupdcd_atoms_1_4012.npz"))
    xyz = z["xyz"].astype(np.float64)[:nframes]
    cells = np.stack([pkg.unitcell_matrix(*c[[0, 2, 5]]) for c in z["dcd_cell"][:nframes]])
    prot = pkg.AtomSelection(list(range(1, 1464)), nmols=1)
    tmao = pkg.AtomSelection(list(range(1479, 4013)), natomspermol=14)
    return pkg.MemoryTrajectory(xyz[:, :1463], xyz[:, 1478:4012], cells, prot, tmao)


def test_metadata_and_result_layout(pkg):
    """src/Trajectory.jl:251-277 and src/results.jl:471-513"""
    from complexmixtures_amd.results import make_result
    t = _dup_traj(pkg)
    tm = pkg.TrajectoryMetaData(t, pkg.Options(silent=True))
    assert (tm.irefatom, tm.lastframe_read, tm.nframes_read, tm.n_groups_solute, tm.n_groups_solvent) == \
        (1, 3, 3, 1463, 14)
    assert tm.orthorhombic
    tm2 = pkg.TrajectoryMetaData(t, pkg.Options(irefatom=2, lastframe=3, stride=2, silent=True))
    assert (tm2.irefatom, tm2.lastframe_read, tm2.nframes_read) == (2, 3, 2)
    with pytest.raises(pkg.ArgumentError):
        pkg.TrajectoryMetaData(t, pkg.Options(irefatom=1000, silent=True))
    with pytest.raises(pkg.ArgumentError):
        pkg.TrajectoryMetaData(t, pkg.Options(lastframe=100, silent=True))
    R = make_result(t, pkg.Options(silent=True), tm)
    assert R.autocorrelation is False and R.cutoff == 10.0 and R.dbulk == 10.0 and R.nbins == 500
    for k in ("d", "kb", "kb_rdf", "md_count", "md_count_random", "mddf", "rdf_count", "rdf_count_random",
              "coordination_number", "coordination_number_random", "sum_rdf_count_random"):
        assert len(getattr(R, k)) == 500
    assert R.solute_group_count.shape == (1463, 500) and R.solvent_group_count.shape == (14, 500)
    assert (R.volume.total, R.volume.bulk, R.volume.domain, len(R.volume.shell)) == (0.0, 0.0, 0.0, 500)
    assert (R.density.solute, R.density.solvent, R.density.solvent_bulk) == (0.0, 0.0, 0.0)
    assert R.files[0].irefatom == 1 and R.files[0].nframes_read == 3 and R.weights == [1.0]
    # frame_weights validation, src/results.jl:195-208
    opts = pkg.Options(silent=True)
    with pytest.raises(pkg.ArgumentError):
        make_result(t, opts, tm, frame_weights=[1.0, 1.0])
    with pytest.raises(pkg.ArgumentError):
        make_result(t, pkg.Options(firstframe=2, lastframe=3, stride=2, silent=True),
                    pkg.TrajectoryMetaData(t, pkg.Options(firstframe=2, lastframe=3, stride=2, silent=True)),
                    frame_weights=[1.0, 0.0, 1.0])
    with pytest.raises(pkg.ArgumentError):
        pkg.coordination_number(t, opts, coordination_number_only=True)


def _header_functions():
    src = open(os.path.join(ROOT, "include", "cmgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(pkg):
    """libcmgpu.so loads without a GPU and exports every function of include/cmgpu.h."""
    from complexmixtures_amd import _lib
    L = _lib.lib()
    names = _header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n
    assert isinstance(L.cm_last_error(), bytes)


def test_oracle_is_not_linked_by_the_product():
    """The product never loads the oracle: no reference to it in the package."""
    pkg_dir = os.path.join(ROOT, "complexmixtures.jl_amd")
    for dirpath, _, files in os.walk(pkg_dir):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "liborc" not in txt and "import oracle" not in txt and "cm_oracle" not in txt, f


def test_synthetic_geometry(pkg):
    import complexmixtures_amd.synthetic as S
    w = S.water_template()
    assert math.isclose(np.linalg.norm(w[1] - w[0]), 0.96) and math.isclose(np.linalg.norm(w[2] - w[0]), 0.96)
    ang = math.degrees(math.acos((w[1] - w[0]) @ (w[2] - w[0]) / 0.96 ** 2))
    assert math.isclose(ang, 104.5)
    s = S.config1(nframes=2, n_solute=100, n_water=50, box=20.0, seed=1)
    assert s["x_solute"].shape == (2, 100, 3) and s["x_solvent"].shape == (2, 150, 3)
    d = np.linalg.norm(np.diff(S.random_walk(np.random.default_rng(0), 50), axis=0), axis=1)
    np.testing.assert_allclose(d, 1.5)
    u = S.urea_template()
    assert u.shape == (8, 3) and np.allclose(u[:, 2], 0.0)


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference checkout not present")
def test_json_load_save_reference_files(pkg, tmp_path):
    """src/results.jl:590-621: the reference's Result files load, round-trip
    through save/load unchanged, and the version gate rejects too new / too old
    files and broken content."""
    data = os.path.join(REFERENCE, "test", "data")
    for rel in ("NAMD/tmao_tmao.json", "NAMD/water_tmao.json", "NAMD/water_water.json", "PDB/tmao_tmao.json",
                "Gromacs/EMI_DCA.json", "Gromacs/EMI_EMI.json"):
        R = pkg.load(os.path.join(data, rel))
        f = tmp_path / "r.json"
        pkg.save(str(f), R)
        assert pkg.results_equal(R, pkg.load(str(f)))
        pkg.save(R, str(f))
        assert pkg.results_equal(R, pkg.load(str(f)))
        assert open(f).read().startswith('{"Version":"2.18.0-DEV","nbins":500,')
    with pytest.warns(UserWarning):
        L = pkg.load(os.path.join(data, "legacy", "tmao_tmao.json"))
    assert L.md_count.shape == (500,)
    wv = os.path.join(data, "legacy", "wrong_version_jsons")
    with pytest.raises(pkg.ArgumentError, match="newer version"):
        pkg.load(os.path.join(wv, "too_new.json"))
    with pytest.raises(pkg.ArgumentError, match="older"):
        pkg.load(os.path.join(wv, "too_old.json"))
    with pytest.raises(pkg.ArgumentError, match="Result object"):
        pkg.load(os.path.join(wv, "broken_json_2.18.0.json"))
    f = tmp_path / "v.json"
    f.write_text('{"Version":"2.18.3-DEV"}')
    with pytest.raises(pkg.ArgumentError):
        pkg.load(str(f))


def test_json_round_trip_of_a_fresh_result(pkg, tmp_path):
    from complexmixtures_amd.results import make_result
    t = _dup_traj(pkg)
    R = make_result(t, pkg.Options(silent=True), pkg.TrajectoryMetaData(t, pkg.Options(silent=True)))
    R.md_count[:] = np.linspace(0, 1, 500)
    R.kb[3] = np.inf
    R.mddf[4] = np.nan
    f = tmp_path / "x.json"
    pkg.save(str(f), R)
    R2 = pkg.load(str(f))
    assert pkg.results_equal(R, R2)
    assert np.isinf(R2.kb[3]) and np.isnan(R2.mddf[4])


def test_dcd_writer_round_trip(pkg, tmp_path):
    """Synthetic trajectories written as NAMD DCD read back exactly (FP32)."""
    import complexmixtures_amd.synthetic as S
    s = S.config3(nframes=3, n_solute=50, n_urea=40, n_water=5, seed=2, a=40.0)
    xyz = np.concatenate([s["x_solute"], s["x_solvent"]], axis=1)
    f = tmp_path / "t.dcd"
    pkg.write_dcd(str(f), xyz, s["cells"])
    t = pkg.Trajectory(str(f), s["solute"], s["solvent"])
    assert t.nframes == 3
    t.opentraj()
    t.firstframe()
    for k in range(3):
        t.nextframe()
        np.testing.assert_array_equal(t.x_solute, s["x_solute"][k].astype(np.float32).astype(np.float64))
        np.testing.assert_array_equal(t.x_solvent, s["x_solvent"][k].astype(np.float32).astype(np.float64))
        np.testing.assert_allclose(t.getunitcell(), s["cells"][k], atol=1e-10)
    t.closetraj()
    box = S.config1(nframes=2, n_solute=20, n_water=10, box=25.0, seed=3)
    pkg.write_dcd(str(f), np.concatenate([box["x_solute"], box["x_solvent"]], axis=1), box["cells"])
    t = pkg.Trajectory(str(f), box["solute"], box["solvent"])
    t.opentraj(); t.firstframe(); t.nextframe()
    np.testing.assert_array_equal(t.getunitcell(), np.diag([25.0] * 3))
    t.closetraj()
