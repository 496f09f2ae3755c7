"""Seeded synthetic systems of the benchmark configurations.

config1: cubic 40 A box, 1,000-atom self-avoiding chain (1.5 A bonds, 3.0 A
         nonbonded spacing), 2,000 rigid 3-site waters, 20 jittered frames, FP64.
config2: cubic 110 A box, 5,000-atom compact random-walk globule (radius
         ~25 A), 40,000 waters outside the globule, 200 jittered frames, FP32.
config3: triclinic a=b=c=90 A, alpha=beta=gamma=80 deg, 3,000-atom polymer,
         12,000 planar 8-site urea-like molecules and 15,000 waters.
Frames: base coordinates plus N(0, 0.3 A) jitter per atom per frame.
"""
import math

import numpy as np

from .selection import AtomSelection
from .trajectory import unitcell_matrix


def water_template():
    """O, H, H: O-H 0.96 A, H-O-H 104.5 deg, O at the origin."""
    a = math.radians(104.5) / 2
    return np.array([[0.0, 0.0, 0.0], [0.96 * math.sin(a), 0.96 * math.cos(a), 0.0],
                     [-0.96 * math.sin(a), 0.96 * math.cos(a), 0.0]])


def urea_template():
    """Planar 8-site urea-like molecule: C, O, N1, N2, H11, H12, H21, H22
    (C=O 1.23 A, C-N 1.35 A, N-H 1.0 A, 120 deg angles)."""
    c30, s30 = math.cos(math.radians(30)), math.sin(math.radians(30))
    n1 = np.array([1.35 * c30, -1.35 * s30, 0.0])
    n2 = np.array([-1.35 * c30, -1.35 * s30, 0.0])
    return np.array([[0.0, 0.0, 0.0], [0.0, 1.23, 0.0], n1, n2,
                     n1 + [c30, s30, 0.0], n1 + [0.0, -1.0, 0.0],
                     n2 + [-c30, s30, 0.0], n2 + [0.0, -1.0, 0.0]])


def random_rotations(rng, n):
    """Uniform random rotation matrices (unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def place_molecules(rng, template, centres):
    rot = random_rotations(rng, len(centres))
    t = template - template.mean(axis=0)
    return (np.einsum("mij,aj->mai", rot, t) + centres[:, None, :]).reshape(-1, 3)


class _Grid:
    """Spatial hash for rejection sampling of chains."""

    def __init__(self, h):
        self.h = h
        self.cells = {}

    def key(self, p):
        return tuple(np.floor(p / self.h).astype(int))

    def add(self, p, i):
        self.cells.setdefault(self.key(p), []).append(i)

    def near(self, p, pts, rmin, skip):
        k = self.key(p)
        r2 = rmin * rmin
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for i in self.cells.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if i < skip and ((pts[i] - p) ** 2).sum() < r2:
                            return True
        return False


def random_walk(rng, n, bond=1.5, rmin=3.0, centre=None, radius=None, start=None, max_tries=200):
    """Self-avoiding random walk (atoms three or more bonds apart at least rmin
    apart), optionally confined to a sphere; backtracks when trapped."""
    pts = np.zeros((n, 3))
    pts[0] = start if start is not None else (centre if centre is not None else np.zeros(3))
    g = _Grid(max(rmin, 1e-3))
    g.add(pts[0], 0)
    i = 1
    fails = 0
    while i < n:
        for _ in range(max_tries):
            d = rng.normal(size=3)
            p = pts[i - 1] + bond * d / np.linalg.norm(d)
            if radius is not None and np.linalg.norm(p - centre) > radius:
                continue
            if rmin > 0 and g.near(p, pts, rmin, i - 2):
                continue
            break
        else:
            fails += 1
            if fails > 100 * n:
                raise RuntimeError("random walk stuck; relax rmin or radius")
            back = min(i - 1, int(rng.integers(1, 10)))
            for k in range(i - back, i):
                g.cells[g.key(pts[k])].remove(k)
            i -= back
            continue
        pts[i] = p
        g.add(p, i)
        i += 1
    return pts


def jitter_frames(rng, base, nframes, sigma=0.3, dtype=np.float64):
    out = np.empty((nframes,) + base.shape, dtype=dtype)
    for f in range(nframes):
        out[f] = base + rng.normal(scale=sigma, size=base.shape)
    return out


def _pack(xu_base, xv_base, cell, nframes, nmols_u, nmols_v, rng, dtype):
    nu_at, nv_at = len(xu_base), len(xv_base)
    xu = jitter_frames(rng, xu_base, nframes, dtype=dtype)
    xv = jitter_frames(rng, xv_base, nframes, dtype=dtype)
    solute = AtomSelection(list(range(1, nu_at + 1)), nmols=nmols_u)
    solvent = AtomSelection(list(range(nu_at + 1, nu_at + nv_at + 1)), nmols=nmols_v)
    return {"x_solute": xu, "x_solvent": xv, "cells": np.repeat(cell[None], nframes, axis=0),
            "solute": solute, "solvent": solvent}


def config1(nframes=20, n_solute=1000, n_water=2000, box=40.0, seed=20261203):
    rng = np.random.default_rng(seed)
    chain = random_walk(rng, n_solute, bond=1.5, rmin=3.0, start=rng.uniform(0, box, 3))
    cent = rng.uniform(0, box, size=(n_water, 3))
    water = place_molecules(rng, water_template(), cent)
    return _pack(chain, water, np.diag([box] * 3), nframes, 1, n_water, rng, np.float64)


def config2(nframes=200, n_solute=5000, n_water=40000, box=110.0, radius=25.0, seed=20261204, dtype=np.float32):
    rng = np.random.default_rng(seed)
    centre = np.full(3, box / 2)
    glob = random_walk(rng, n_solute, bond=1.5, rmin=1.2, centre=centre, radius=radius, start=centre)
    cent = []
    need = n_water
    while need > 0:
        c = rng.uniform(0, box, size=(2 * need, 3))
        c = c[np.linalg.norm(c - centre, axis=1) > radius + 1.5][:need]
        cent.append(c)
        need -= len(c)
    water = place_molecules(rng, water_template(), np.concatenate(cent))
    return _pack(glob, water, np.diag([box] * 3), nframes, 1, n_water, rng, dtype)


def triclinic_cell(a=90.0, angle=80.0):
    return unitcell_matrix(a, a, a, angle, angle, angle)


def config3(nframes=100, n_solute=3000, n_urea=12000, n_water=15000, seed=20261205, solvent="urea", dtype=np.float64,
            a=90.0):
    """solvent: 'urea' (polymer vs urea), 'water' (polymer vs water) or 'self'
    (urea vs urea); the other molecules are spectators and are not returned."""
    rng = np.random.default_rng(seed)
    H = triclinic_cell(a)
    poly = random_walk(rng, n_solute, bond=1.5, rmin=3.0, start=H @ rng.uniform(0, 1, 3))
    urea = place_molecules(rng, urea_template(), (H @ rng.uniform(0, 1, size=(n_urea, 3)).T).T)
    water = place_molecules(rng, water_template(), (H @ rng.uniform(0, 1, size=(n_water, 3)).T).T)
    if solvent == "urea":
        return _pack(poly, urea, H, nframes, 1, n_urea, rng, dtype)
    if solvent == "water":
        return _pack(poly, water, H, nframes, 1, n_water, rng, dtype)
    xv = jitter_frames(rng, urea, nframes, dtype=dtype)
    sel = AtomSelection(list(range(1, len(urea) + 1)), nmols=n_urea)
    return {"x_solute": xv, "x_solvent": xv, "cells": np.repeat(H[None], nframes, axis=0), "solute": sel,
            "solvent": sel}


def water_box(nframes, n_solute_atoms, box, seed, density=0.0334, dtype=np.float64):
    """One system of the batched distribution: a random-walk solute and water
    at the given molecular density outside it."""
    rng = np.random.default_rng(seed)
    centre = np.full(3, box / 2)
    sol = random_walk(rng, n_solute_atoms, bond=1.5, rmin=2.0, start=centre)
    n_water = max(1, int(density * box ** 3))
    cent = rng.uniform(0, box, size=(n_water, 3))
    water = place_molecules(rng, water_template(), cent)
    return _pack(sol, water, np.diag([box] * 3), nframes, 1, n_water, rng, dtype)


def served_system(seed, n_solute, box, nframes, density=0.0334, sigma=0.3, dtype=np.float32, noise_frames=8):
    """A system of the uneven batched distribution: a random-walk solute of
    n_solute atoms started at the box centre and water at the given molecular
    density in a cubic box.  Frames are the base coordinates plus N(0, sigma)
    jitter; the jitter of frame f is one of noise_frames draws, rolled over the
    atoms by a frame-dependent offset (so frames differ without drawing
    nframes x natoms normals)."""
    rng = np.random.default_rng(seed)
    centre = np.full(3, box / 2)
    sol = random_walk(rng, n_solute, bond=1.5, rmin=2.0, start=centre)
    n_water = max(2, int(density * box ** 3))
    water = place_molecules(rng, water_template(), rng.uniform(0, box, size=(n_water, 3)))
    base = np.concatenate([sol, water]).astype(dtype)
    k = min(noise_frames, nframes)
    noise = rng.normal(scale=sigma, size=(k,) + base.shape).astype(dtype)
    x = np.empty((nframes,) + base.shape, dtype=dtype)
    for f in range(nframes):
        x[f] = base + np.roll(noise[f % k], (f // k) * 7919, axis=0)
    nu = len(sol)
    return {"x_solute": np.ascontiguousarray(x[:, :nu]), "x_solvent": np.ascontiguousarray(x[:, nu:]),
            "cells": np.repeat(np.diag([float(box)] * 3)[None], nframes, axis=0), "n_solute": nu, "n_water": n_water,
            "box": float(box)}
