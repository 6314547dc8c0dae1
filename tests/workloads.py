"""Seeded synthetic workloads shared by the tests and bench.py.

PRNG: SplitMix64 (Steele, Lea, Flood 2014), output k (k = 1, 2, ...) of a
stream with seed s is mix(s + k * 0x9E3779B97F4A7C15). Uniform
U[-1, 1) = 2 * (u >> 11) * 2^-53 - 1; normals by Box-Muller on the same
stream. Index order of a sphere: centred triplets, x ascending, then y, then
z, over [-R, R].
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int) -> np.ndarray:
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int) -> np.ndarray:
    """U[-1, 1) doubles."""
    u = splitmix64(seed, n)
    return 2.0 * ((u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0


def uniform01(seed: int, n: int) -> np.ndarray:
    u = splitmix64(seed, n)
    return (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed: int, n: int) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = np.maximum(u[0::2], 2.0 ** -60)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    return np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]


def sphere_triplets(radius: float, half: bool = False, kshift=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Centred triplets with |G + k|^2 <= r^2, x-major. half=True keeps the
    R2C half space x > 0, or x = 0 and y > 0, or x = y = 0 and z >= 0."""
    R = int(np.ceil(radius + 1))
    g = np.arange(-R, R + 1)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    kx, ky, kz = kshift
    mask = (x + kx) ** 2 + (y + ky) ** 2 + (z + kz) ** 2 <= radius * radius
    if half:
        mask &= (x > 0) | ((x == 0) & (y > 0)) | ((x == 0) & (y == 0) & (z >= 0))
    t = np.stack([x[mask], y[mask], z[mask]], axis=1).astype(np.int32)
    return np.ascontiguousarray(t)


def complex_values(seed: int, n: int) -> np.ndarray:
    u = uniform(seed, 2 * n)
    return u[0::2] + 1j * u[1::2]


def random_index_set(seed: int, dims, r2c: bool, stick_fraction=0.7, fill_fraction=0.7,
                     centered=False) -> np.ndarray:
    """Random sticks and z fill like the reference tests
    (tests/test_util/generate_indices.hpp:38-85): R2C uses the hermitian half
    (x <= nx/2, x = 0 only for y <= ny/2, (0,0) stick only for z <= nz/2).
    Order: x, then y, then z ascending (storage indices)."""
    nx, ny, nz = dims
    nxf = nx // 2 + 1 if r2c else nx
    nyf = ny // 2 + 1 if r2c else ny
    nzf = nz // 2 + 1 if r2c else nz
    u_stick = uniform01(seed, nxf * ny).reshape(nxf, ny)
    u_fill = uniform01(seed ^ 0x5A5A5A5A, nxf * ny * nz).reshape(nxf, ny, nz)
    x, y, z = np.meshgrid(np.arange(nxf), np.arange(ny), np.arange(nz), indexing="ij")
    keep = (u_stick[:, :, None] < stick_fraction) & (u_fill < fill_fraction)
    if r2c:
        keep &= ~((x == 0) & (y >= nyf))
        keep &= ~((x == 0) & (y == 0) & (z >= nzf))
    t = np.stack([x[keep], y[keep], z[keep]], axis=1).astype(np.int32)
    if centered:
        for d, n in enumerate(dims):
            t[:, d] = np.where(t[:, d] >= n // 2 + 1, t[:, d] - n, t[:, d])
    return np.ascontiguousarray(t)
