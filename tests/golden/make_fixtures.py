"""Writes the golden fixtures of the sparse 3D FFT (tests/golden/*.npz).

The expected outputs come from numpy's pocketfft applied to the definition of
the transform (reference docs/source/details.rst:4-44): backward = scatter the
values into the dense frequency grid (storage index = i mod N) and take the
unnormalised inverse 3D DFT; forward = full 3D DFT of the space field sampled
at the triplets, times 1/(Nx Ny Nz) with full scaling. R2C sets are half-space
sets without mirrored pairs, so the reference's symmetry fill equals the
hermitian completion f(-k) = conj f(k) and C2R keeps the real part.

The C oracle and the HIP library are both checked against these files; this
script depends on neither. Rerun it only to regenerate (deterministic):
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform01  # noqa: E402

SAMPLE = 4096  # grids above 16^3 keep a seeded sample of backward outputs


def scatter(dims, t, v, hermitian):
    nx, ny, nz = dims
    g = np.zeros((nz, ny, nx), np.complex128)
    x, y, z = t[:, 0] % nx, t[:, 1] % ny, t[:, 2] % nz
    g[z, y, x] = v
    if hermitian:
        g[(-z) % nz, (-y) % ny, (-x) % nx] = np.conj(v)
        g[z, y, x] = v  # self-mirrored entries keep their value
    return g


def case(name, dims, t, v, r2c):
    nx, ny, nz = dims
    ng = nx * ny * nz
    g = scatter(dims, t, v, r2c)
    space = np.fft.ifftn(g) * ng
    if r2c:
        space = space.real
    f = np.fft.fftn(space.astype(np.complex128))
    fwd = f[t[:, 2] % nz, t[:, 1] % ny, t[:, 0] % nx] * (1.0 / float(ng))
    out = dict(dims=np.array(dims, np.int32), r2c=np.array(int(r2c), np.int32),
               triplets=t.astype(np.int32), values=v.astype(np.complex128),
               forward_scaled=fwd.astype(np.complex128),
               backward_norm=np.array(np.linalg.norm(space.ravel())))
    flat = space.ravel()
    if ng > 16 ** 3:
        idx = np.sort((uniform01(0xF1C5 + ng, SAMPLE) * ng).astype(np.int64))
        out["backward_index"] = idx
        out["backward_sample"] = flat[idx]
    else:
        out["backward"] = flat
    np.savez(os.path.join(HERE, name + ".npz"), **out)
    print(name, dims, "r2c" if r2c else "c2c", len(t), "values")


def main():
    d = (11, 12, 13)
    for centered in (False, True):
        t = random_index_set(0x601D, d, False, centered=centered)
        case(f"c2c_11x12x13_{'centered' if centered else 'storage'}", d, t,
             complex_values(0x601E + centered, len(t)), False)
    t = sphere_triplets(6.0)
    case("c2c_32_sphere_r6", (32, 32, 32), t, complex_values(0x6020, len(t)), False)
    t = sphere_triplets(10.0)
    case("c2c_64_sphere_r10", (64, 64, 64), t, complex_values(0x6021, len(t)), False)
    t = sphere_triplets(5.0, half=True)
    case("r2c_12x11x13_halfsphere_r5", (12, 11, 13), t, complex_values(0x6022, len(t)), True)
    t = sphere_triplets(6.0, half=True)
    case("r2c_32_halfsphere_r6", (32, 32, 32), t, complex_values(0x6023, len(t)), True)
    t = sphere_triplets(10.0, half=True)
    case("r2c_64_halfsphere_r10", (64, 64, 64), t, complex_values(0x6024, len(t)), True)


if __name__ == "__main__":
    main()
