"""Result layout and FP64 finalisation, mirroring src/results.jl.

Result fields in the reference's order (src/results.jl:61-118).  Group counts
are [group][bin] arrays (the reference's Vector{Vector{Float64}}).
finalresults follows src/results.jl:311-469 operation by operation; sums over
bins are sequential.
"""
import math
from dataclasses import dataclass

import numpy as np

from .options import ArgumentError, TrajectoryFileOptions

VERSION = "2.18.3-DEV"   # layout version of the reference's Result (Project.toml:4)
Angs3tocm3permol = 6.022140857e23 / 1e24     # units, src/io.jl:6-14
SitesperAngs3tomolperL = 1e27 / 6.022140857e23


def setbin(d, step):
    """src/results.jl:28"""
    return max(1, int(math.ceil(d / step)))


def _cube(x):
    return x * x * x   # Julia's x^3 with a literal exponent


def sphericalshellvolume(i, step):
    """src/results.jl:254-257"""
    rmin = (i - 1) * step
    return (4 * math.pi / 3) * (_cube(rmin + step) - _cube(rmin))


def shellradius(i, step):
    """src/results.jl:272-275"""
    rmin = (i - 1) * step
    return (0.5 * (_cube(rmin + step) + _cube(rmin))) ** (1 / 3)


@dataclass
class Density:
    solute: float = 0.0
    solvent: float = 0.0
    solvent_bulk: float = 0.0


@dataclass
class Volume:
    total: float
    bulk: float
    domain: float
    shell: np.ndarray


def _seqsum(a):
    a = np.asarray(a, dtype=np.float64)
    return np.float64(np.cumsum(a)[-1]) if a.size else np.float64(0.0)


class Result:
    """src/results.jl:61-118"""

    _vec_fields = ("d", "md_count", "md_count_random", "coordination_number", "coordination_number_random", "mddf",
                   "kb", "rdf_count", "rdf_count_random", "sum_rdf_count", "sum_rdf_count_random", "rdf", "kb_rdf")
    _group_fields = ("solute_group_count", "solvent_group_count", "solute_group_count_random",
                     "solvent_group_count_random")

    def __init__(self, *, nbins, dbulk, cutoff, autocorrelation, solute, solvent, n_groups_solute, n_groups_solvent,
                 files, weights=None):
        self.Version = VERSION
        self.nbins = int(nbins)
        self.dbulk = float(dbulk)
        self.cutoff = float(cutoff)
        for n in self._vec_fields:
            setattr(self, n, np.zeros(self.nbins))
        self.autocorrelation = bool(autocorrelation)
        self.solute = solute
        self.solvent = solvent
        self.solute_group_count = np.zeros((n_groups_solute, self.nbins))
        self.solvent_group_count = np.zeros((n_groups_solvent, self.nbins))
        self.solute_group_count_random = np.zeros((n_groups_solute, self.nbins))
        self.solvent_group_count_random = np.zeros((n_groups_solvent, self.nbins))
        self.density = Density()
        self.volume = Volume(0.0, 0.0, 0.0, np.zeros(self.nbins))
        self.files = files
        self.weights = [1.0] if weights is None else list(weights)

    def __repr__(self):
        return (f"Result(nbins={self.nbins}, dbulk={self.dbulk}, cutoff={self.cutoff}, "
                f"autocorrelation={self.autocorrelation}, frames={self.files[0].nframes_read if self.files else 0})")


def make_result(trajectory, options, trajectory_data, frame_weights=None):
    """Result(trajectory, options; trajectory_data, frame_weights), src/results.jl:124-193"""
    nbins = setbin(options.cutoff, options.binstep)
    if frame_weights is not None and np.size(frame_weights) > 0:
        fw = np.asarray(frame_weights)
        if sum(1 for s in fw.shape if s > 1) > 1:
            raise ArgumentError(f"\nThe frame_weights provided must be one-dimensional, but a {fw.shape}-dimensional "
                                "array was given.\n")
        fw = fw.reshape(-1).astype(np.float64)
        if len(fw) < trajectory_data.lastframe_read:
            raise ArgumentError("\nThe length of the frame_weights vector provided must at least the number of frames "
                                f"to be read.\n\n    Input given: length(frame_weights) = {len(fw)}\n"
                                f"                 last frame to be read: {trajectory_data.lastframe_read}\n")
        rng = range(options.firstframe, trajectory_data.lastframe_read + 1, options.stride)
        if sum(fw[i - 1] for i in rng) <= 0.0:
            raise ArgumentError("\nThe sum of the frame weights must be greater than zero for the frames that will be "
                                f"considered: {options.firstframe}:{options.stride}:{trajectory_data.lastframe_read}\n")
    else:
        fw = np.ones(trajectory_data.lastframe_read)
    files = [TrajectoryFileOptions(filename=trajectory.filename, options=options, irefatom=trajectory_data.irefatom,
                                   lastframe_read=trajectory_data.lastframe_read,
                                   nframes_read=trajectory_data.nframes_read, frame_weights=[float(x) for x in fw])]
    return Result(nbins=nbins, dbulk=options.dbulk, cutoff=options.cutoff,
                  autocorrelation=isautocorrelation(trajectory), solute=trajectory.solute, solvent=trajectory.solvent,
                  n_groups_solute=trajectory_data.n_groups_solute, n_groups_solvent=trajectory_data.n_groups_solvent,
                  files=files, weights=[1.0])


def isautocorrelation(trajectory):
    """src/results.jl:240-243"""
    return list(trajectory.solute.indices) == list(trajectory.solvent.indices)


def set_samples(R):
    """src/results.jl:230-237"""
    n = R.files[0].options.n_random_samples
    return (R.solvent.nmols - 1 if R.autocorrelation else R.solvent.nmols), n


def sum_frame_weights(R):
    """src/results.jl:288-297"""
    Q = np.float64(0.0)
    for file in R.files:
        o = file.options
        for i in range(o.firstframe, file.lastframe_read + 1, o.stride):
            Q += file.frame_weights[i - 1]
    return Q


def finalresults(R, options, coordination_number_only):
    """finalresults!, src/results.jl:311-318.  Scalars are numpy float64, so that
    divisions by zero give Inf/NaN as Julia's Float64 arithmetic does."""
    R.volume.total = np.float64(R.volume.total)
    with np.errstate(divide="ignore", invalid="ignore"):
        if not coordination_number_only:
            return _mddf_final_results(R, options)
        return _coordination_number_final_results(R, options)


def _mddf_final_results(R, options):
    """src/results.jl:320-376"""
    R.d[:] = [shellradius(i, options.binstep) for i in range(1, R.nbins + 1)]
    solvent_nmols, nrandom = set_samples(R)
    Q = sum_frame_weights(R)
    nu = R.solute.nmols
    R.md_count[:] = R.md_count / (nu * Q)
    R.solute_group_count[:] = R.solute_group_count / (nu * Q)
    R.solute_group_count_random[:] = R.solute_group_count_random / (nrandom * Q)
    if R.autocorrelation:
        R.solvent_group_count[:] = R.solute_group_count
        R.solvent_group_count_random[:] = R.solute_group_count_random
    else:
        R.solvent_group_count[:] = R.solvent_group_count / (nu * Q)
        R.solvent_group_count_random[:] = R.solvent_group_count_random / (nrandom * Q)
    R.md_count_random[:] = R.md_count_random / (nrandom * Q)
    R.rdf_count[:] = R.rdf_count / (nu * Q)
    R.rdf_count_random[:] = R.rdf_count_random / (nrandom * Q)
    R.volume.total = R.volume.total / Q
    R.volume.shell[:] = R.volume.total * (R.rdf_count_random / solvent_nmols)
    step = R.files[0].options.binstep
    ibulk = setbin(R.dbulk + 0.5 * step, step)
    R.volume.domain = _seqsum(R.volume.shell[:ibulk - 1])
    if not R.files[0].options.usecutoff:
        R.volume.bulk = R.volume.total - R.volume.domain
        n_solvent_in_bulk = solvent_nmols - _seqsum(R.rdf_count)
    else:
        n_solvent_in_bulk = _seqsum(R.rdf_count[ibulk - 1:R.nbins])
        R.volume.bulk = _seqsum(R.volume.shell[ibulk - 1:R.nbins])
    R.density.solvent = R.solvent.nmols / R.volume.total
    R.density.solute = R.solute.nmols / R.volume.total
    R.density.solvent_bulk = n_solvent_in_bulk / R.volume.bulk
    density_fix = R.density.solvent_bulk / R.density.solvent
    return renormalize(R, density_fix, silent=options.silent)


def renormalize(R, density_fix, silent=True):
    """renormalize!, src/results.jl:378-428"""
    R.md_count_random *= density_fix
    R.rdf_count_random *= density_fix
    R.solute_group_count_random *= density_fix
    R.solvent_group_count_random *= density_fix
    R.coordination_number[:] = np.cumsum(R.md_count)
    R.coordination_number_random[:] = np.cumsum(R.md_count_random)
    pos = R.md_count_random > 0.0
    if not silent and not np.all(pos):
        import warnings
        warnings.warn("\nIdeal-gas histogram bins with zero samples. \n"
                      "Increase n_random_samples, number of trajectory frames, and/or bin size.\n")
    R.mddf[pos] = R.md_count[pos] / R.md_count_random[pos]
    inv = np.float64(1) / R.density.solvent_bulk
    R.kb[:] = Angs3tocm3permol * inv * (R.coordination_number - R.coordination_number_random)
    posr = R.rdf_count_random > 0.0
    R.rdf[posr] = R.rdf_count[posr] / R.rdf_count_random[posr]
    R.sum_rdf_count[:] = np.cumsum(R.rdf_count)
    R.sum_rdf_count_random[:] = np.cumsum(R.rdf_count_random)
    R.kb_rdf[:] = Angs3tocm3permol * inv * (R.sum_rdf_count - R.sum_rdf_count_random)
    return R


def _coordination_number_final_results(R, options):
    """src/results.jl:430-469"""
    if not options.silent:
        import warnings
        warnings.warn("\ncoordination_number_only was set to true, so the MDDF and KB integrals were not computed. \n"
                      "(to remove this warning use `Options(silent=true)`)\n")
    R.d[:] = [shellradius(i, options.binstep) for i in range(1, R.nbins + 1)]
    Q = sum_frame_weights(R)
    nu = R.solute.nmols
    R.md_count[:] = R.md_count / (nu * Q)
    R.solute_group_count[:] = R.solute_group_count / (nu * Q)
    if R.autocorrelation:
        R.solvent_group_count[:] = R.solute_group_count
    else:
        R.solvent_group_count[:] = R.solvent_group_count / (nu * Q)
    R.rdf_count[:] = R.rdf_count / (nu * Q)
    R.volume.total = R.volume.total / Q
    R.density.solvent = R.solvent.nmols / R.volume.total
    R.density.solute = R.solute.nmols / R.volume.total
    R.coordination_number[:] = np.cumsum(R.md_count)
    R.sum_rdf_count[:] = np.cumsum(R.rdf_count)
    return R


def coordination_number(R, group=None):
    """coordination_number(R), src/tools/coordination_number.jl:105"""
    if group is None:
        return R.coordination_number
    raise NotImplementedError("group contributions are outside the accelerated path")
