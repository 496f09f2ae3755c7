"""MI355X MDDF path of ComplexMixtures.jl.

Host-side mirror of the reference interface for the minimum-distance
distribution function:

    R = mddf(trajectory_file, solute, solvent, Options(...))
    R = mddf(MemoryTrajectory(x_solute, x_solvent, unitcells, solute, solvent), options)
    C = coordination_number(trajectory_file, solute, solvent, options)

with the accumulation running in libcmgpu.so (HIP, gfx950) behind the C ABI
of include/cmgpu.h.
"""
from .options import ArgumentError, Options, TrajectoryFileOptions
from .selection import AtomSelection
from .trajectory import (MemoryTrajectory, NamdDCD, PDBTraj, Trajectory, convert_unitcell, unitcell_matrix,
                         write_dcd)
from .results import (Density, Result, Volume, coordination_number as _cn_of_result, finalresults, setbin,
                      shellradius, sphericalshellvolume, sum_frame_weights)
from .mddf import TrajectoryMetaData, coordination_number, mddf
from .persistence import load, results_equal, save
from . import _lib

__all__ = ["ArgumentError", "Options", "TrajectoryFileOptions", "AtomSelection", "MemoryTrajectory", "NamdDCD",
           "PDBTraj", "Trajectory", "convert_unitcell", "unitcell_matrix", "Density", "Result", "Volume",
           "finalresults", "setbin", "shellradius", "sphericalshellvolume", "sum_frame_weights",
           "TrajectoryMetaData", "coordination_number", "mddf", "load", "save", "results_equal", "write_dcd"]
