"""Seeded synthetic systems for the se_e2_a path.

Water: molecules on an n x n x n grid (n = ceil(cbrt(N_mol)), the first N_mol
sites in lexicographic order) in a cubic box of side
L = (N_mol / 0.033428 A^-3)^(1/3) (1.0 g/cm^3), O at the site centres,
O-H 0.9572 A, H-O-H 104.52 deg, uniform-random orientation, then i.i.d.
N(0, sigma^2) jitter on every coordinate.  Atom order: all O (type 0), then
all H (type 1).

Copper: fcc, a = 3.615 A, n^3 conventional cells, N(0, 0.05^2) jitter,
single type 0.

All generators use numpy's PCG64 (``default_rng(seed)``).
"""
from __future__ import annotations

import math

import numpy as np

WATER_NUMBER_DENSITY = 0.033428  # molecules per A^3 at 1.0 g/cm^3
OH = 0.9572
HOH = math.radians(104.52)


def water_box(n_mol: int) -> float:
    """Cubic box side for n_mol molecules at 1.0 g/cm^3 (12.42 A for 64,
    49.68 A for 4096, rounded to 0.01 A)."""
    return round((n_mol / WATER_NUMBER_DENSITY) ** (1.0 / 3.0), 2)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    # uniform random unit quaternions (Shoemake)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack(
        [
            np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
            np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
            np.sqrt(u1) * np.sin(2 * np.pi * u3),
            np.sqrt(u1) * np.cos(2 * np.pi * u3),
        ],
        axis=1,
    )
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water(n_mol: int, seed: int = 0, sigma: float = 0.1, box: float | None = None):
    """Returns (coord [3*n_mol, 3] float64, atype [3*n_mol] int32, box [9])."""
    rng = np.random.default_rng(seed)
    L = water_box(n_mol) if box is None else float(box)
    n = int(math.ceil(round(n_mol ** (1.0 / 3.0), 9)))
    while n ** 3 < n_mol:
        n += 1
    g = np.arange(n)
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)[:n_mol]
    o = (sites + 0.5) * (L / n)
    half = HOH / 2
    local = np.array(
        [[OH * math.sin(half), OH * math.cos(half), 0.0], [-OH * math.sin(half), OH * math.cos(half), 0.0]]
    )
    R = _random_rotations(rng, n_mol)
    h = o[:, None, :] + np.einsum("nij,kj->nki", R, local)
    coord = np.concatenate([o, h.reshape(-1, 3)], axis=0)
    coord = coord + rng.normal(0.0, sigma, size=coord.shape)
    atype = np.concatenate([np.zeros(n_mol, np.int32), np.ones(2 * n_mol, np.int32)])
    boxm = np.diag([L, L, L]).reshape(9).astype(np.float64)
    return coord.astype(np.float64), atype, boxm


def fcc_copper(ncell: int, seed: int = 0, sigma: float = 0.05, a: float = 3.615):
    """Returns (coord [4*ncell^3, 3], atype, box [9])."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(ncell)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    coord = ((cells[:, None, :] + basis[None]) * a).reshape(-1, 3)
    coord = coord + rng.normal(0.0, sigma, size=coord.shape)
    atype = np.zeros(len(coord), np.int32)
    L = a * ncell
    boxm = np.diag([L, L, L]).reshape(9).astype(np.float64)
    return coord.astype(np.float64), atype, boxm


def uneven_water_frames(nframes: int, seed: int = 0, lo: int = 96, hi: int = 3000):
    """Batch of water frames with atom counts uniform over {lo, lo+3, ..., hi}
    (whole molecules), each in its own 1.0 g/cm^3 cubic box."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(lo // 3, hi // 3 + 1, size=nframes)
    return [water(int(c), seed=seed * 100003 + k + 1) for k, c in enumerate(counts)]
