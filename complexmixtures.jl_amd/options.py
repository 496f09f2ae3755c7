"""Options and TrajectoryFileOptions, mirroring src/Options.jl.

Same keywords, defaults, validation and messages as the reference
(src/Options.jl:68-192, TrajectoryFileOptions src/Options.jl:274-283).
"""
import warnings
from dataclasses import dataclass, field
from typing import List


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


@dataclass(frozen=True)
class _OptionsData:
    firstframe: int
    lastframe: int
    stride: int
    irefatom: int
    n_random_samples: int
    binstep: float
    dbulk: float
    cutoff: float
    usecutoff: bool
    lcell: int
    GC: bool
    GC_threshold: float
    seed: int
    StableRNG: bool
    nthreads: int
    silent: bool


class Options(_OptionsData):
    """Options(; firstframe=1, lastframe=-1, stride=1, irefatom=-1,
    n_random_samples=10, binstep=0.02, dbulk=None, cutoff=None,
    usecutoff=None, bulk_range=None, lcell=1, GC=True, GC_threshold=0.3,
    seed=321, StableRNG=False, nthreads=0, silent=False)

    src/Options.jl:68-192.  `StableRNG`, `GC`, `GC_threshold` and `nthreads`
    are kept for the record; the device path draws from Philox4x32-10 keyed
    by `seed` and needs no host threads.
    """

    def __init__(self, *, firstframe=1, lastframe=-1, stride=1, irefatom=-1, n_random_samples=10, binstep=0.02,
                 dbulk=None, cutoff=None, usecutoff=None, bulk_range=None, lcell=1, GC=True, GC_threshold=0.3,
                 seed=321, StableRNG=False, nthreads=0, silent=False):
        warn = False
        if stride < 1:
            raise ArgumentError("in MDDF options: stride cannot be less than 1. ")
        if lastframe > 0 and lastframe < firstframe:
            raise ArgumentError("in MDDF options: lastframe must be greater or equal to firstframe. ")
        if n_random_samples < 1:
            raise ArgumentError(
                "\n in MDDF options: n_random_samples must be greater than 0. \n"
                "To skip the normalization of the distribution, use the coordination_number(...) "
                "function instead of mddf(...).\n")
        if bulk_range is not None and any(x is not None for x in (dbulk, cutoff, usecutoff)):
            raise ArgumentError("\nThe bulk_range argument implies that dbulk, cutoff, and usecutoff are not needed. \n")
        if all(x is None for x in (bulk_range, dbulk, cutoff, usecutoff)):
            dbulk, cutoff, usecutoff = 10.0, 10.0, False
            warn = True
        elif bulk_range is not None:
            if len(bulk_range) != 2:
                raise ArgumentError(
                    "\nbulk_range must be a tuple or vector with two elements, corresponding to dbulk and cutoff.\n"
                    "Example: Options(;bulk_range = (8.0, 12.0))\n")
            dbulk, cutoff = bulk_range
            usecutoff = True
        else:
            if dbulk is None:
                dbulk = 10.0
                warn = True
            if usecutoff is None:
                usecutoff = False
                warn = True
            if cutoff is None:
                warn = True
                cutoff = dbulk + 4.0 if usecutoff else dbulk
            elif not usecutoff:
                raise ArgumentError("in MDDF options: cutoff was defined with usecutoff set to false")
        if warn and not silent:
            warnings.warn(f"\nUsing default values for dbulk, cutoff and/or usecutoff: \n\n"
                          f"    dbulk = {dbulk}\n    cutoff = {cutoff}\n    usecutoff = {usecutoff}\n\n"
                          "It is recommended to set bulk_range manually, according to the \n"
                          "system size and correlations of the distribution function with, \n"
                          "for example: Options(bulk_range = (8.0, 12.0))\n", stacklevel=2)
        dbulk, cutoff = float(dbulk), float(cutoff)
        if usecutoff and dbulk >= cutoff:
            raise ArgumentError(" in MDDF options: The bulk volume is zero (dbulk must be smaller than cutoff). ")
        if (cutoff / binstep) % 1 > 1.e-5:
            raise ArgumentError("in MDDF options: cutoff must be a multiple of binstep.")
        if (dbulk / binstep) % 1 > 1.e-5:
            raise ArgumentError("in MDDF options: dbulk must be a multiple of binstep.")
        super().__init__(int(firstframe), int(lastframe), int(stride), int(irefatom), int(n_random_samples),
                         float(binstep), dbulk, cutoff, bool(usecutoff), int(lcell), bool(GC), float(GC_threshold),
                         int(seed), bool(StableRNG), int(nthreads), bool(silent))

    def to_dict(self):
        return {k: getattr(self, k) for k in _OptionsData.__dataclass_fields__}

    @classmethod
    def from_dict(cls, d):
        o = object.__new__(cls)
        _OptionsData.__init__(o, *[d[k] for k in _OptionsData.__dataclass_fields__])
        return o


@dataclass
class TrajectoryFileOptions:
    """src/Options.jl:274-283"""
    filename: str
    options: Options
    irefatom: int
    lastframe_read: int
    nframes_read: int
    frame_weights: List[float] = field(default_factory=list)
