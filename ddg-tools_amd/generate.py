"""Seeded synthetic meshes of the benchmark configurations.

``icosphere`` and ``torus`` follow ``crates/ddg-core/src/generators.rs`` (``icosphere`` :108, the
midpoint subdivision order :152-171; ``torus`` :194 with ``quad_grid_faces`` :362) and
``SplitMix64`` :60.  The perturbations are this package's (the reference has no Gaussian
generator): Box–Muller normal deviates from SplitMix64 uniforms and Gaussian-profile spherical
bumps.  Every function is deterministic in its seed.
"""
from __future__ import annotations

import math

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    """generators.rs:60 ``SplitMix64``."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def next_f64(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.next_f64()

    def normal(self) -> float:
        """Box–Muller: sqrt(−2 ln(1 − u1)) cos(2π u2)."""
        u1 = self.next_f64()
        u2 = self.next_f64()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def _normalize_rows(p: np.ndarray) -> np.ndarray:
    length = np.sqrt(p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1] + p[:, 2] * p[:, 2])
    return p / length[:, None]


def icosphere(level: int):
    """Unit icosphere: 10·4^level + 2 vertices, 20·4^level faces (generators.rs:108)."""
    if not 0 <= level <= 9:
        raise ValueError(f"icosphere level must be in 0..=9, got {level}")
    t = (1.0 + math.sqrt(5.0)) / 2.0
    pos = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
         [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    pos = _normalize_rows(pos)
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
             [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
             [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(level):
        mids: dict = {}
        new_pairs = []
        nv = pos.shape[0]

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            k = mids.get(key)
            if k is None:
                k = nv + len(new_pairs)
                mids[key] = k
                new_pairs.append((i, j))
            return k

        nxt = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nxt += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        ij = np.array(new_pairs, dtype=np.int64)
        m = pos[ij[:, 0]] + pos[ij[:, 1]]
        pos = np.vstack([pos, _normalize_rows(m)])
        faces = nxt
    return pos, np.array(faces, dtype=np.uint32)


def torus(n_major: int, n_minor: int, major_radius: float, minor_radius: float):
    """Torus around z: n_major·n_minor vertices, twice as many faces (generators.rs:194)."""
    if n_major < 3 or n_minor < 3 or not 0 < minor_radius < major_radius:
        raise ValueError("torus needs ≥ 3 segments each way and 0 < r < R")
    pos = np.empty((n_major * n_minor, 3))
    k = 0
    for i in range(n_major):
        theta = 2.0 * math.pi * i / n_major
        for j in range(n_minor):
            phi = 2.0 * math.pi * j / n_minor
            ring = major_radius + minor_radius * math.cos(phi)
            pos[k] = (ring * math.cos(theta), ring * math.sin(theta), minor_radius * math.sin(phi))
            k += 1
    faces = []
    for i in range(n_major):
        for j in range(n_minor):
            a = i * n_minor + j
            b = ((i + 1) % n_major) * n_minor + j
            c = ((i + 1) % n_major) * n_minor + (j + 1) % n_minor
            d = i * n_minor + (j + 1) % n_minor
            faces += [[a, b, c], [a, c, d]]
    return pos, np.array(faces, dtype=np.uint32)


def vertex_normals(pos: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (unit)."""
    p0, p1, p2 = pos[faces[:, 0]], pos[faces[:, 1]], pos[faces[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    vn = np.zeros_like(pos)
    for c in range(3):
        np.add.at(vn, faces[:, c], n)
    return _normalize_rows(vn)


def normal_jitter(pos: np.ndarray, faces: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    """Displaces every vertex along its normal by sigma·N(0,1)."""
    rng = SplitMix64(seed)
    z = np.array([rng.normal() for _ in range(pos.shape[0])])
    return pos + (sigma * z)[:, None] * vertex_normals(pos, faces)


def iso_jitter(pos: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    """Adds sigma·N(0,1) to every coordinate."""
    rng = SplitMix64(seed)
    z = np.array([[rng.normal() for _ in range(3)] for _ in range(pos.shape[0])])
    return pos + sigma * z


def bumps(pos: np.ndarray, count: int, amplitude: tuple, width: tuple, seed: int) -> np.ndarray:
    """Radial displacement of a unit-sphere mesh by `count` Gaussian spherical bumps:
    r = 1 + Σ_b A_b exp(−θ_b²/(2 w_b²)), θ_b the geodesic angle to the bump's centre (uniform on
    the sphere), A_b ~ U[amplitude], w_b ~ U[width] (radians)."""
    rng = SplitMix64(seed)
    unit = _normalize_rows(pos)
    r = np.ones(pos.shape[0])
    for _ in range(count):
        z = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        a = rng.uniform(*amplitude)
        w = rng.uniform(*width)
        s = math.sqrt(max(0.0, 1.0 - z * z))
        c = np.array([s * math.cos(phi), s * math.sin(phi), z])
        theta = np.arccos(np.clip(unit @ c, -1.0, 1.0))
        r = r + a * np.exp(-(theta * theta) / (2.0 * w * w))
    return unit * r[:, None]


# ---- the benchmark configurations --------------------------------------------------------------

def config_icosphere_l3(seed: int = 1):
    """Icosphere level 3 (642 vertices, 1,280 faces), normal jitter sigma 0.01."""
    p, f = icosphere(3)
    return normal_jitter(p, f, 0.01, seed), f


def config_torus_16k(seed: int = 2):
    """Torus R=1, r=0.4, 128 segments around the axis × 64 around the tube (8,192 vertices,
    16,384 faces), isotropic jitter sigma 0.005."""
    p, f = torus(128, 64, 1.0, 0.4)
    return iso_jitter(p, 0.005, seed), f


def config_bumpy_l6(seed: int = 3):
    """Icosphere level 6 (40,962 vertices, 81,920 faces) with 16 bumps, A ~ U[0, 0.1],
    w ~ U[0.1, 0.4] rad."""
    p, f = icosphere(6)
    return bumps(p, 16, (0.0, 0.1), (0.1, 0.4), seed), f


def config_bumpy_l7(seed: int = 5):
    """Icosphere level 7 (163,842 vertices, 327,680 faces) with 64 bumps, A ~ U[0, 0.05],
    w ~ U[0.05, 0.2] rad."""
    p, f = icosphere(7)
    return bumps(p, 64, (0.0, 0.05), (0.05, 0.2), seed), f
