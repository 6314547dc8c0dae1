"""CPU oracle of the sparse 3D FFT path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / reported CPU baseline.
The product library never imports it.

Two independent layers:

* ``OraclePlan`` wraps ``liboracle.so`` (``spfft_oracle.c``), the C
  restatement of the reference's SPFFT_PU_HOST pipeline (see that file's
  header for the reference file:line of every stage).
* ``dense_*`` / ``brute_*`` helpers restate the *definition* of the
  transform (``docs/source/details.rst:4-44`` of the reference) with numpy's
  pocketfft or an O(N^2) sum, used to pin the C restatement.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

INVALID_PARAMETER = 5
INVALID_INDICES = 7


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (fast, a second or two)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_index_map.argtypes = [ctypes.c_int] * 5 + [ip, ip, ip, ip]
        L.oracle_index_map.restype = ctypes.c_int
        L.oracle_plan_create.argtypes = [ctypes.c_int] * 5 + [ip, ctypes.c_int, ip]
        L.oracle_plan_create.restype = ctypes.c_void_p
        L.oracle_plan_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_plan_num_sticks.argtypes = [ctypes.c_void_p]
        L.oracle_plan_num_sticks.restype = ctypes.c_int
        L.oracle_backward.argtypes = [ctypes.c_void_p, dp, dp]
        L.oracle_forward.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_int]
        L.oracle_fft1d.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.oracle_use_fft_library.argtypes = [ctypes.c_char_p]
        L.oracle_use_fft_library.restype = ctypes.c_int
        _lib = L
    return _lib


# Libraries exporting the FFTW3 C API the reference's host path calls
# (src/fft/fftw_interface.hpp:41-101): FFTW itself, or MKL's FFTW wrappers.
FFT_LIBRARIES = ["libfftw3.so.3", "/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"]
_provider = "built-in mixed-radix FFT (oracle)"


def use_fft_library(candidates=None):
    """Route the 1D lines of plans created from now on through the first
    loadable FFTW3-API library (for the CPU baseline timing). MKL runs with
    MKL_THREADING_LAYER=SEQUENTIAL, as the reference's CMake prefers. Returns
    the provider name, or None (built-in FFT kept)."""
    global _provider
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    for path in candidates or FFT_LIBRARIES:
        if lib().oracle_use_fft_library(path.encode()) == 0:
            _provider = f"FFTW3 C API of {path}"
            return _provider
    return None


def use_builtin_fft():
    global _provider
    lib().oracle_use_fft_library(None)
    _provider = "built-in mixed-radix FFT (oracle)"


def fft_provider() -> str:
    return _provider


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def index_map(r2c: bool, dims, triplets):
    """Returns (valueSlot, stickKey) int32 arrays, or an int error code."""
    nx, ny, nz = dims
    t = np.ascontiguousarray(np.asarray(triplets, dtype=np.int32).reshape(-1))
    n = t.size // 3
    slots = np.zeros(max(n, 1), np.int32)
    keys = np.zeros(max(n, 1), np.int32)
    ns = ctypes.c_int(0)
    err = lib().oracle_index_map(int(r2c), nx, ny, nz, n, _iptr(t), _iptr(slots), _iptr(keys),
                                 ctypes.byref(ns))
    if err:
        return err
    return slots[:n].copy(), keys[: ns.value].copy()


def fft1d(x, sign: int):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    out = np.zeros_like(x)
    lib().oracle_fft1d(x.size, _dptr(x.view(np.float64)), _dptr(out.view(np.float64)), sign)
    return out


class OraclePlan:
    """C restatement of the reference host transform for one index set."""

    def __init__(self, r2c: bool, dims, triplets, nthreads: int = 0):
        self.r2c = bool(r2c)
        self.dims = tuple(int(d) for d in dims)
        t = np.ascontiguousarray(np.asarray(triplets, dtype=np.int32).reshape(-1))
        self.num_values = t.size // 3
        err = ctypes.c_int(0)
        self._p = lib().oracle_plan_create(int(r2c), *self.dims, self.num_values, _iptr(t),
                                           nthreads, ctypes.byref(err))
        if not self._p:
            raise ValueError(f"oracle rejected the index set (SpfftError {err.value})")

    def __del__(self):
        if getattr(self, "_p", None):
            lib().oracle_plan_destroy(self._p)
            self._p = None

    @property
    def num_sticks(self) -> int:
        return lib().oracle_plan_num_sticks(self._p)

    def space_shape(self):
        nx, ny, nz = self.dims
        return (nz, ny, nx)

    def backward(self, values):
        """values: complex128[num_values] -> space domain (nz, ny, nx)."""
        v = np.ascontiguousarray(np.asarray(values, dtype=np.complex128))
        assert v.size == self.num_values
        dt = np.float64 if self.r2c else np.complex128
        out = np.zeros(self.space_shape(), dtype=dt)
        lib().oracle_backward(self._p, _dptr(v.view(np.float64)), _dptr(out.view(np.float64)))
        return out

    def forward(self, space, scaling: bool):
        dt = np.float64 if self.r2c else np.complex128
        s = np.ascontiguousarray(np.asarray(space, dtype=dt).reshape(self.space_shape()))
        out = np.zeros(self.num_values, dtype=np.complex128)
        lib().oracle_forward(self._p, _dptr(s.view(np.float64)), _dptr(out.view(np.float64)),
                             int(bool(scaling)))
        return out


# ---------------------------------------------------------------------------
# Definition-level references (numpy), independent of the C restatement.


def storage_index(dim: int, i):
    i = np.asarray(i)
    return np.where(i < 0, i + dim, i)


def dense_backward_c2c(dims, triplets, values):
    """Backward C2C by definition: scatter into the dense frequency grid and
    apply the unnormalised inverse 3D DFT (numpy ifftn * N)."""
    nx, ny, nz = dims
    t = np.asarray(triplets).reshape(-1, 3)
    g = np.zeros((nz, ny, nx), np.complex128)
    g[storage_index(nz, t[:, 2]), storage_index(ny, t[:, 1]), storage_index(nx, t[:, 0])] = values
    return np.fft.ifftn(g) * (nx * ny * nz)


def dense_forward(dims, triplets, space, scaling: bool, r2c: bool = False):
    """Forward by definition: full 3D DFT sampled at the triplets."""
    nx, ny, nz = dims
    t = np.asarray(triplets).reshape(-1, 3)
    f = np.fft.fftn(np.asarray(space).reshape(nz, ny, nx).astype(np.complex128))
    out = f[storage_index(nz, t[:, 2]), storage_index(ny, t[:, 1]), storage_index(nx, t[:, 0])]
    if scaling:
        out = out * (1.0 / float(nx * ny * nz))
    return out


def dense_backward_r2c_hermitian(dims, triplets, values):
    """Backward R2C for a value set that is already hermitian-consistent:
    complete the grid with f(-k) = conj f(k) and take the real inverse DFT."""
    nx, ny, nz = dims
    t = np.asarray(triplets).reshape(-1, 3)
    g = np.zeros((nz, ny, nx), np.complex128)
    x, y, z = (storage_index(nx, t[:, 0]), storage_index(ny, t[:, 1]), storage_index(nz, t[:, 2]))
    g[z, y, x] = values
    g[(-z) % nz, (-y) % ny, (-x) % nx] = np.conj(values)
    return (np.fft.ifftn(g) * (nx * ny * nz)).real


def brute_dft3(dims, triplets, values, sign=+1):
    """O(Ns * Ng) sum of plane waves, long double accumulation: tiny grids only."""
    nx, ny, nz = dims
    t = np.asarray(triplets).reshape(-1, 3)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    acc = np.zeros((nz, ny, nx), np.complex128)
    for (kx, ky, kz), v in zip(t, values):
        ph = (np.longdouble(kx) * xx / nx + np.longdouble(ky) * yy / ny + np.longdouble(kz) * zz / nz)
        ph = ph - np.floor(ph)
        ang = sign * 2 * np.pi * ph.astype(np.float64)
        acc += v * np.exp(1j * ang)
    return acc


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)
