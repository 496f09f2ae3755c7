"""Writes dupdcd_atoms_1_4012.npz from the reference's test trajectory.

Source: test/data/NAMD/traj_duplicated_first_frame.dcd of ComplexMixtures.jl
(62,026 atoms, 3 frames; frames 1 and 2 are frame 1 of the NAMD protein/TMAO
trajectory, frame 3 is its frame 2 -- comment at src/mddf.jl:860-862).
Kept: atoms 1..4012 (protein 1:1463, TMAO 1479:4012), FP32 as stored in the
file, and the unit-cell records.  Run from the repository root:
    python tests/golden/make_dupdcd_fixture.py /root/reference
"""
import os
import struct
import sys

import numpy as np


def records(buf):
    off = 0
    while off < len(buf):
        (n,) = struct.unpack_from("<i", buf, off)
        yield buf[off + 4: off + 4 + n]
        off += 8 + n


def main(ref_root):
    path = os.path.join(ref_root, "test", "data", "NAMD", "traj_duplicated_first_frame.dcd")
    recs = list(records(open(path, "rb").read()))
    natoms = struct.unpack("<i", recs[2])[0]
    frames, cells = [], []
    for f in range((len(recs) - 3) // 4):
        r = recs[3 + 4 * f: 7 + 4 * f]
        cells.append(np.frombuffer(r[0], "<f8"))
        xyz = np.stack([np.frombuffer(r[k], "<f4") for k in (1, 2, 3)], axis=1)
        frames.append(xyz[:4012])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dupdcd_atoms_1_4012.npz")
    np.savez_compressed(out, xyz=np.stack(frames).astype(np.float32), dcd_cell=np.stack(cells), natoms_file=natoms)
    print(out, np.stack(frames).shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
