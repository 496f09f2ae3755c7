"""Trajectories: the interface of src/Trajectory.jl:11-63 and two readers.

Interface (src/Trajectory.jl, src/trajectory_formats/PDBTraj.jl:120-174):
opentraj(), closetraj(), firstframe(), nextframe() (fills x_solute and
x_solvent, [natoms][3]), getunitcell() (3x3, lattice vectors as columns),
nframes, solute, solvent, filename.

Readers: NAMD DCD (src/trajectory_formats/NamdDCD.jl:53-229) and multi-model
PDB (src/trajectory_formats/PDBTraj.jl:16-198).  MemoryTrajectory holds all
frames in arrays (for synthetic inputs); the reference has only file-backed
trajectories.
"""
import math
import struct

import numpy as np

from .options import ArgumentError

trajectory_formats = ["dcd", "PDBTraj", "memory"]


def _cosd(angle):
    # exact zero for right angles, as the cell matrix of Chemfiles.UnitCell
    return 0.0 if angle == 90.0 else math.cos(math.radians(angle))


def _sind(angle):
    return 1.0 if angle == 90.0 else math.sin(math.radians(angle))


def unitcell_matrix(a, b, c, alpha=90.0, beta=90.0, gamma=90.0):
    """Cell matrix of lengths and angles (degrees): lattice vectors as columns,
    a along x, b in the xy plane (the convention of Chemfiles.UnitCell used by
    src/trajectory_formats/NamdDCD.jl:187 and PDBTraj.jl:197)."""
    ca, cb, cg, sg = _cosd(alpha), _cosd(beta), _cosd(gamma), _sind(gamma)
    m = np.zeros((3, 3))
    m[:, 0] = (a, 0.0, 0.0)
    m[:, 1] = (b * cg, b * sg, 0.0)
    cy = c * (ca - cb * cg) / sg
    cz2 = c * c - (c * cb) ** 2 - cy * cy
    m[:, 2] = (c * cb, cy, math.sqrt(max(cz2, 0.0)))
    return m


def is_orthorhombic(unitcell, tol=1e-10):
    """convert_unitcell, src/Trajectory.jl:72-77: diagonal up to tol * min|diag|"""
    m = np.asarray(unitcell, dtype=np.float64)
    if m.shape == (3,):
        return True
    if m.shape != (3, 3):
        raise ValueError("Unit cell must be a 3x3 matrix.")
    s = abs(np.min(np.diag(m)))
    return all(abs(m[i, j]) < tol * s for i in range(3) for j in range(3) if i != j)


def convert_unitcell(unitcell, orthorhombic=None):
    """Returns the 3x3 cell; with orthorhombic=True only the diagonal is kept
    (convert_unitcell(::SVector, m), src/Trajectory.jl:80-81)."""
    m = np.asarray(unitcell, dtype=np.float64)
    if m.shape == (3,):
        return np.diag(m)
    if orthorhombic is None:
        orthorhombic = is_orthorhombic(m)
    return np.diag(np.diag(m)) if orthorhombic else m.copy()


class TrajectoryBase:
    filename = ""

    def __init__(self, solute, solvent):
        self.solute = solute
        self.solvent = solvent
        self._ix_u = np.asarray(solute.indices, dtype=np.int64) - 1
        self._ix_v = np.asarray(solvent.indices, dtype=np.int64) - 1

    def read_frames(self, nframes, out_u=None, out_v=None, out_cells=None):
        """Reads the next nframes frames into arrays (default float64)."""
        nu, nv = len(self.solute.indices), len(self.solvent.indices)
        out_u = np.empty((nframes, nu, 3)) if out_u is None else out_u
        out_v = np.empty((nframes, nv, 3)) if out_v is None else out_v
        out_cells = np.empty((nframes, 3, 3)) if out_cells is None else out_cells
        for f in range(nframes):
            self.nextframe()
            out_u[f] = self.x_solute
            out_v[f] = self.x_solvent
            out_cells[f] = self.getunitcell()
        return out_u, out_v, out_cells


class NamdDCD(TrajectoryBase):
    """NAMD DCD trajectory: Fortran unformatted records, FP32 coordinates,
    unit cell record [A, gamma, B, beta, alpha, C] (angles 0 -> 90 degrees).
    src/trajectory_formats/NamdDCD.jl:53-229."""

    def __init__(self, filename, solute, solvent, lastframe=-1):
        super().__init__(solute, solvent)
        self.filename = filename
        self._fh = None
        with open(filename, "rb") as fh:
            self._parse_header(fh)
            nframes = 0
            while lastframe < 0 or nframes < lastframe:
                if not self._skip_frame(fh):
                    break
                nframes += 1
        self.nframes = nframes
        self.lastatom = max(max(solute.indices), max(solvent.indices))
        if self.lastatom > self.natoms_file:
            raise ArgumentError("atom indices beyond the number of atoms of the DCD file")
        self.x_solute = np.zeros((len(solute.indices), 3))
        self.x_solvent = np.zeros((len(solvent.indices), 3))
        self.unitcell_read = np.zeros(6)

    @staticmethod
    def _record(fh):
        h = fh.read(4)
        if len(h) < 4:
            return None
        (n,) = struct.unpack("<i", h)
        data = fh.read(n)
        t = fh.read(4)
        if len(data) < n or len(t) < 4 or struct.unpack("<i", t)[0] != n:
            return None
        return data

    def _parse_header(self, fh):
        h = self._record(fh)
        if h is None or h[:4] != b"CORD":
            raise ArgumentError(f"{self.filename} is not a DCD file")
        self._record(fh)                     # title
        n = self._record(fh)
        self.natoms_file = struct.unpack("<i", n)[0]
        self._data_start = fh.tell()
        c = self._record(fh)
        if c is None or len(c) != 48:
            raise ArgumentError("DCD file does not contain unit cell information.")
        fh.seek(self._data_start)

    def _skip_frame(self, fh):
        for nbytes in (48, 4 * self.natoms_file, 4 * self.natoms_file, 4 * self.natoms_file):
            h = fh.read(4)
            if len(h) < 4 or struct.unpack("<i", h)[0] != nbytes:
                return False
            fh.seek(nbytes, 1)
            t = fh.read(4)
            if len(t) < 4:
                return False
        return True

    def opentraj(self):
        self._fh = open(self.filename, "rb")

    def closetraj(self):
        if self._fh:
            self._fh.close()
        self._fh = None

    def firstframe(self):
        self._fh.seek(self._data_start)

    def nextframe(self):
        fh = self._fh
        self.unitcell_read[:] = np.frombuffer(self._record(fh), "<f8")
        x = np.frombuffer(self._record(fh), "<f4")
        y = np.frombuffer(self._record(fh), "<f4")
        z = np.frombuffer(self._record(fh), "<f4")
        xyz = np.stack([x, y, z], axis=1).astype(np.float64)
        self.x_solute[:] = xyz[self._ix_u]
        self.x_solvent[:] = xyz[self._ix_v]
        return self

    def getunitcell(self):
        A, g, B, b, a, C = self.unitcell_read
        if a == 0.0 or b == 0.0 or g == 0.0:
            a = b = g = 90.0
        return unitcell_matrix(A, B, C, a, b, g)


class PDBTraj(TrajectoryBase):
    """Multi-model PDB trajectory: frames end with END, CRYST1 gives the cell.
    src/trajectory_formats/PDBTraj.jl:16-198."""

    def __init__(self, filename, solute, solvent):
        super().__init__(solute, solvent)
        self.filename = filename
        nframes = natoms_ = 0
        cell = None
        with open(filename) as fh:
            for line in fh:
                fields = line.split()
                if not fields:
                    continue
                if fields[0] == "CRYST1":
                    cell = self._readcell(fields)
                if fields[0] == "END":
                    nframes += 1
                if nframes == 0 and fields[0] in ("ATOM", "HETATM"):
                    natoms_ += 1
        if cell is None or not np.any(cell):
            raise ValueError("Could not read unit cell from PDB file. Each frame must contain a CRYST1 field.")
        if nframes == 0:
            raise ValueError("Could not read any frame from PDB file. Each frame must end with the END specifier")
        if natoms_ == 0:
            raise ValueError("Could not read any ATOM from the trajectory file.")
        self.nframes = nframes
        self.unitcell = cell
        self.x_solute = np.zeros((len(solute.indices), 3))
        self.x_solvent = np.zeros((len(solvent.indices), 3))
        self._fh = None

    @staticmethod
    def _readcell(fields):
        a, b, c, al, be, ga = (float(x) for x in fields[1:7])
        return unitcell_matrix(a, b, c, al, be, ga)

    def opentraj(self):
        self._fh = open(self.filename)

    def closetraj(self):
        if self._fh:
            self._fh.close()
        self._fh = None

    def firstframe(self):
        self._fh.seek(0)

    def nextframe(self):
        iatom = i_u = i_v = 0
        nu, nv = len(self.solute.indices), len(self.solvent.indices)
        while True:
            rec = self._fh.readline()
            if rec == "":
                break
            rec = rec.rstrip("\n")
            if (0 < len(rec) < 3) or rec[:3] == "END":
                if rec[:3] == "END":
                    break
                continue
            if len(rec) >= 6:
                if rec[:6] == "CRYST1":
                    self.unitcell = self._readcell(rec.split())
                if rec[:4] == "ATOM" or rec[:6] == "HETATM":
                    iatom += 1
                    xyz = (float(rec[30:38]), float(rec[38:46]), float(rec[46:54]))
                    if i_u < nu and iatom == self.solute.indices[i_u]:
                        self.x_solute[i_u] = xyz
                        i_u += 1
                    if i_v < nv and iatom == self.solvent.indices[i_v]:
                        self.x_solvent[i_v] = xyz
                        i_v += 1
        return self

    def getunitcell(self):
        return self.unitcell.copy()


class MemoryTrajectory(TrajectoryBase):
    """All frames held in arrays: x_solute [nframes][natoms_u][3],
    x_solvent [nframes][natoms_v][3] (float64 or float32) and unitcells
    [nframes][3][3] (lattice vectors as columns) or [nframes][3] (box sides)."""

    def __init__(self, x_solute, x_solvent, unitcells, solute, solvent=None, filename="memory"):
        solvent = solute if solvent is None else solvent
        super().__init__(solute, solvent)
        self.filename = filename
        self.xu = np.asarray(x_solute)
        self.xv = self.xu if x_solvent is None else np.asarray(x_solvent)
        uc = np.asarray(unitcells, dtype=np.float64)
        nf = self.xu.shape[0]
        if uc.shape == (nf, 3):
            uc = np.stack([np.diag(r) for r in uc])
        if uc.shape != (nf, 3, 3):
            raise ArgumentError("unitcells must have shape (nframes, 3, 3) or (nframes, 3)")
        self.cells = uc
        self.nframes = nf
        if self.xu.shape[1] != len(solute.indices) or self.xv.shape[1] != len(solvent.indices):
            raise ArgumentError("coordinate arrays do not match the selections")
        self._i = 0
        self.x_solute = None
        self.x_solvent = None

    def opentraj(self):
        self._i = 0

    def closetraj(self):
        pass

    def firstframe(self):
        self._i = 0

    def nextframe(self):
        self.x_solute = self.xu[self._i]
        self.x_solvent = self.xv[self._i]
        self._cur = self._i
        self._i += 1
        return self

    def getunitcell(self):
        return self.cells[self._cur].copy()

    def frame_block(self, start, count):
        """Zero-copy views of `count` consecutive frames from 0-based `start`."""
        return self.xu[start:start + count], self.xv[start:start + count], self.cells[start:start + count]


def _cell_params(m):
    """Lengths and angles (degrees) of a cell matrix with lattice vectors as columns."""
    m = np.asarray(m, dtype=np.float64)
    a, b, c = (np.linalg.norm(m[:, k]) for k in range(3))
    ang = lambda u, v: math.degrees(math.acos(max(-1.0, min(1.0, float(u @ v) / (np.linalg.norm(u) * np.linalg.norm(v))))))
    return a, b, c, ang(m[:, 1], m[:, 2]), ang(m[:, 0], m[:, 2]), ang(m[:, 0], m[:, 1])


def write_dcd(filename, xyz, unitcells):
    """Writes a NAMD-style DCD (FP32 coordinates, unit-cell record
    [A, gamma, B, beta, alpha, C] per frame) readable by NamdDCD and by the
    reference's reader (src/trajectory_formats/NamdDCD.jl:53-229).
    xyz: [nframes][natoms][3]; unitcells: [nframes][3][3] or [nframes][3]."""
    xyz = np.asarray(xyz)
    nf, na = xyz.shape[0], xyz.shape[1]
    uc = np.asarray(unitcells, dtype=np.float64)
    if uc.shape == (nf, 3):
        uc = np.stack([np.diag(r) for r in uc])

    def rec(fh, payload):
        fh.write(struct.pack("<i", len(payload)))
        fh.write(payload)
        fh.write(struct.pack("<i", len(payload)))

    with open(filename, "wb") as fh:
        icntrl = [0] * 20
        icntrl[0] = nf          # number of frames
        icntrl[1] = 1           # first step
        icntrl[2] = 1           # step interval
        icntrl[3] = nf
        icntrl[10] = 1          # unit cell present
        icntrl[19] = 24         # CHARMM version (NAMD writes 24)
        hdr = b"CORD" + struct.pack("<9i", *icntrl[:9]) + struct.pack("<f", 1.0) + struct.pack("<10i", *icntrl[10:])
        rec(fh, hdr)
        title = b"REMARKS written by complexmixtures_amd".ljust(80)
        rec(fh, struct.pack("<i", 1) + title)
        rec(fh, struct.pack("<i", na))
        for f in range(nf):
            a, b, c, al, be, ga = _cell_params(uc[f])
            if abs(al - 90.0) < 1e-9 and abs(be - 90.0) < 1e-9 and abs(ga - 90.0) < 1e-9:
                al = be = ga = 90.0
            rec(fh, struct.pack("<6d", a, ga, b, be, al, c))
            for k in range(3):
                rec(fh, np.ascontiguousarray(xyz[f, :, k], dtype="<f4").tobytes())


def Trajectory(filename, solute, solvent=None, *, format="", chemfiles=False, show_progress=False, lastframe=-1):
    """src/Trajectory.jl:31-63"""
    solvent = solute if solvent is None else solvent
    filename = str(filename)
    format = str(format)
    if format and format not in trajectory_formats:
        raise ArgumentError(f"\nTrajectory format not properly set: {format} \n"
                            f"Available trajectory formats: {', '.join(trajectory_formats)}\n")
    ext = filename[filename.rfind(".") + 1:] if "." in filename else ""
    if not chemfiles and (format == "dcd" or ext == "dcd"):
        return NamdDCD(filename, solute, solvent, lastframe=lastframe)
    if not chemfiles and (format == "PDBTraj" or ext == "pdb"):
        return PDBTraj(filename, solute, solvent)
    raise ArgumentError("Only NAMD DCD and PDB trajectories are readable here (Chemfiles is not available).")
