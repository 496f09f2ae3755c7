"""mddf / coordination_number, mirroring src/mddf.jl:186-347 and :539-585.

The entry points, keyword handling, metadata, frame selection, stop-file
interrupt and finalisation follow the reference.  The chunked frame loop
(src/mddf.jl:285-338) is replaced by batches of frames submitted to the
device library.
"""
import os
import secrets

import numpy as np

from . import _lib
from .options import ArgumentError, Options
from .results import finalresults, make_result
from .selection import AtomSelection, group_csr
from .trajectory import MemoryTrajectory, Trajectory, TrajectoryBase, convert_unitcell, is_orthorhombic


class TrajectoryMetaData:
    """src/Trajectory.jl:180-249"""

    def __init__(self, trajectory, options):
        if options.irefatom > trajectory.solvent.natomspermol:
            raise ArgumentError(f"in MDDF options: Reference atom index {options.irefatom} is greater than number of "
                                "atoms of the solvent molecule. ")
        if options.lastframe > trajectory.nframes:
            raise ArgumentError("in MDDF options: lastframe is greater than trajectory.nframes. ")
        trajectory.opentraj()
        trajectory.firstframe()
        if options.irefatom == -1:
            trajectory.nextframe()
            self.unitcell = np.asarray(trajectory.getunitcell(), dtype=np.float64)
            first_mol = np.asarray(trajectory.x_solvent[:trajectory.solvent.natomspermol], dtype=np.float64)
            cm = first_mol.mean(axis=0)
            dist = np.sqrt(((first_mol - cm) ** 2).sum(axis=1))
            self.irefatom = int(np.argmin(dist)) + 1
        else:
            trajectory.nextframe()
            self.unitcell = np.asarray(trajectory.getunitcell(), dtype=np.float64)
            self.irefatom = options.irefatom
        self.orthorhombic = is_orthorhombic(self.unitcell)
        self.lastframe_read = trajectory.nframes if options.lastframe == -1 else options.lastframe
        self.nframes_read = len(range(options.firstframe, self.lastframe_read + 1, options.stride))
        trajectory.closetraj()
        self.n_groups_solute = trajectory.solute.ngroups
        self.n_groups_solvent = trajectory.solvent.ngroups


def job_kwargs(R, options, tmeta, coordination_number_only):
    """Arguments of a device job (cm_job_desc) for Result R of this trajectory."""
    solute, solvent = R.solute, R.solvent
    nu = solute.natomspermol
    Nv = solvent.nmols * solvent.natomspermol
    groups_u = group_csr(solute, range(nu), solute)
    # self: the solvent-side atom j counts into the solute groups (update_counters.jl:48-53)
    groups_v = group_csr(solvent, range(Nv), solute if R.autocorrelation else solvent)
    seed = options.seed if options.seed > 0 else secrets.randbits(63) + 1
    return dict(nmols_u=solute.nmols, natoms_per_mol_u=nu, nmols_v=solvent.nmols,
                natoms_per_mol_v=solvent.natomspermol, autocorrelation=R.autocorrelation,
                irefatom=tmeta.irefatom, nbins=R.nbins, binstep=options.binstep, cutoff=options.cutoff,
                dbulk=options.dbulk, usecutoff=options.usecutoff, n_random_samples=options.n_random_samples,
                coordination_number_only=coordination_number_only, triclinic=not tmeta.orthorhombic, seed=seed,
                n_groups_u=tmeta.n_groups_solute,
                n_groups_v=tmeta.n_groups_solute if R.autocorrelation else tmeta.n_groups_solvent,
                groups_u=groups_u, groups_v=groups_v, lcell=options.lcell)


def _job_for(R, options, tmeta, coordination_number_only, **kw):
    return _lib.Job(**job_kwargs(R, options, tmeta, coordination_number_only), **kw)


def _store_raw(R, raw):
    R.md_count[:] = raw["md_count"]
    R.md_count_random[:] = raw["md_count_random"]
    R.rdf_count[:] = raw["rdf_count"]
    R.rdf_count_random[:] = raw["rdf_count_random"]
    R.solute_group_count[:] = raw["solute_group_count"]
    R.solute_group_count_random[:] = raw["solute_group_count_random"]
    if not R.autocorrelation:
        R.solvent_group_count[:] = raw["solvent_group_count"]
        R.solvent_group_count_random[:] = raw["solvent_group_count_random"]
    R.volume.total = raw["volume_total"]


def _accumulate(trajectory, options, frame_weights=None, coordination_number_only=False, batch_frames=None,
                job_kw=None):
    """Metadata, Result allocation and the frame loop; returns (R, tmeta, raw)."""
    tmeta = TrajectoryMetaData(trajectory, options)
    R = make_result(trajectory, options, tmeta, frame_weights)
    fw = R.files[0].frame_weights
    job = _job_for(R, options, tmeta, coordination_number_only, **(job_kw or {}))
    try:
        ortho = tmeta.orthorhombic
        to_compute = set(range(options.firstframe, tmeta.lastframe_read + 1, options.stride))
        nu_at = len(trajectory.solute.indices)
        nv_at = len(trajectory.solvent.indices)
        if batch_frames is None:
            per = 8 * 3 * (nu_at + nv_at)
            batch_frames = int(max(1, min(256, (256 << 20) // max(per, 1))))
        if isinstance(trajectory, MemoryTrajectory):
            # frames are already in memory: submit contiguous runs without copying
            sel = [i for i in range(options.firstframe, tmeta.lastframe_read + 1)
                   if i in to_compute and fw[i - 1] != 0.0]
            k = 0
            while k < len(sel):
                if os.path.isfile("stop_complexmixtures"):
                    break
                run = [sel[k]]
                while (k + len(run) < len(sel) and len(run) < batch_frames
                       and sel[k + len(run)] == run[-1] + 1):
                    run.append(sel[k + len(run)])
                xu, xv, cells = trajectory.frame_block(run[0] - 1, len(run))
                cells = np.stack([convert_unitcell(c, ortho) for c in cells])
                job.submit(run, xu, xv, cells, [fw[i - 1] for i in run])
                k += len(run)
        else:
            trajectory.opentraj()
            trajectory.firstframe()
            for _ in range(options.firstframe - 1):
                trajectory.nextframe()
            iframe = options.firstframe - 1
            bu = np.empty((batch_frames, nu_at, 3))
            bv = np.empty((batch_frames, nv_at, 3))
            bc = np.empty((batch_frames, 3, 3))
            idx, ws = [], []

            def flush():
                if idx:
                    n = len(idx)
                    job.submit(idx, bu[:n], bv[:n], bc[:n], ws)
                    idx.clear()
                    ws.clear()

            while iframe < tmeta.lastframe_read:
                if os.path.isfile("stop_complexmixtures"):
                    break
                trajectory.nextframe()
                iframe += 1
                w = fw[iframe - 1]
                if iframe in to_compute and w != 0.0:
                    n = len(idx)
                    bu[n] = trajectory.x_solute
                    bv[n] = trajectory.x_solvent
                    bc[n] = convert_unitcell(trajectory.getunitcell(), ortho)
                    idx.append(iframe)
                    ws.append(w)
                    if len(idx) == batch_frames:
                        flush()
            flush()
            trajectory.closetraj()
        raw = job.finish()
    finally:
        job.close()
    return R, tmeta, raw


def _resolve_args(args, kw):
    """Dispatch of mddf(file, solute, solvent[, options]), mddf(file, sel[, options]), mddf(traj[, options])."""
    trajectory_format = kw.pop("trajectory_format", "")
    chemfiles = kw.pop("chemfiles", False)
    if not args:
        raise ArgumentError("missing trajectory")
    first = args[0]
    if isinstance(first, TrajectoryBase):
        options = args[1] if len(args) > 1 else Options()
        return first, options
    if isinstance(first, str):
        sels = [a for a in args[1:] if isinstance(a, AtomSelection)]
        opts = [a for a in args[1:] if isinstance(a, Options)]
        if len(sels) not in (1, 2) or len(opts) > 1:
            raise ArgumentError("No method matching these arguments")
        options = opts[0] if opts else Options()
        solute = sels[0]
        solvent = sels[1] if len(sels) == 2 else sels[0]
        traj = Trajectory(first, solute, solvent, format=trajectory_format, chemfiles=chemfiles,
                          lastframe=options.lastframe)
        return traj, options
    raise ArgumentError("No method matching these arguments")


def mddf(*args, frame_weights=None, coordination_number_only=False, low_memory=False, **kw):
    """mddf(trajectory_file, solute, solvent, options; ...) / mddf(trajectory_file, solute_and_solvent, options)
    / mddf(trajectory, options; frame_weights, coordination_number_only, low_memory)

    src/mddf.jl:186-347.  `low_memory` is accepted and has no effect (the
    device path keeps one set of accumulators per job)."""
    job_kw = kw.pop("job_kw", None)
    batch_frames = kw.pop("batch_frames", None)
    traj, options = _resolve_args(args, kw)
    if kw:
        raise ArgumentError(f"unexpected keywords: {sorted(kw)}")
    R, tmeta, raw = _accumulate(traj, options, frame_weights, coordination_number_only, batch_frames, job_kw)
    _store_raw(R, raw)
    finalresults(R, options, coordination_number_only)
    return R


def coordination_number(*args, **kw):
    """coordination_number(trajectory_file, solute, solvent, options; ...), src/mddf.jl:539-574"""
    if "coordination_number_only" in kw:
        raise ArgumentError("\n\nThe keyword argument `coordination_number_only` is not allowed in the "
                            "`coordination_number` function.\nIt is, by definition, set to `true` in this function. \n")
    if len(args) == 1 and not isinstance(args[0], (str, TrajectoryBase)):
        from .results import coordination_number as cn
        return cn(args[0])
    return mddf(*args, coordination_number_only=True, **kw)
