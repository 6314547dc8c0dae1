"""Python mirror of the SpFFT C interface, bound with ctypes to the in-tree
``spfft_amd/lib/libspfft.so`` (HIP kernels for gfx950 + C++ host code).

The classes mirror the reference's C++ API (``include/spfft/grid.hpp``,
``include/spfft/transform.hpp``, ``include/spfft/multi_transform.hpp``):
same names, argument meaning and error codes. Data pointers are passed as
integers (``torch.Tensor.data_ptr()``) or numpy arrays (host memory). There
is no fallback: if the library is missing, importing this package fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspfft.so")

# ---- enums (include/spfft/types.h) ----
SPFFT_PU_HOST = 1
SPFFT_PU_GPU = 2
SPFFT_INDEX_TRIPLETS = 0
SPFFT_TRANS_C2C = 0
SPFFT_TRANS_R2C = 1
SPFFT_NO_SCALING = 0
SPFFT_FULL_SCALING = 1
SPFFT_EXEC_SYNCHRONOUS = 0
SPFFT_EXEC_ASYNCHRONOUS = 1

# ---- error codes (include/spfft/errors.h), in declaration order ----
ERROR_NAMES = [
    "SPFFT_SUCCESS", "SPFFT_UNKNOWN_ERROR", "SPFFT_INVALID_HANDLE_ERROR", "SPFFT_OVERFLOW_ERROR",
    "SPFFT_ALLOCATION_ERROR", "SPFFT_INVALID_PARAMETER_ERROR", "SPFFT_DUPLICATE_INDICES_ERROR",
    "SPFFT_INVALID_INDICES_ERROR", "SPFFT_MPI_SUPPORT_ERROR", "SPFFT_MPI_ERROR",
    "SPFFT_MPI_PARAMETER_MISMATCH_ERROR", "SPFFT_HOST_EXECUTION_ERROR", "SPFFT_FFTW_ERROR",
    "SPFFT_GPU_ERROR", "SPFFT_GPU_PRECEDING_ERROR", "SPFFT_GPU_SUPPORT_ERROR",
    "SPFFT_GPU_ALLOCATION_ERROR", "SPFFT_GPU_LAUNCH_ERROR", "SPFFT_GPU_NO_DEVICE_ERROR",
    "SPFFT_GPU_INVALID_VALUE_ERROR", "SPFFT_GPU_INVALID_DEVICE_PTR_ERROR", "SPFFT_GPU_COPY_ERROR",
    "SPFFT_GPU_FFT_ERROR",
]
for _code, _name in enumerate(ERROR_NAMES):
    globals()[_name] = _code


class SpfftError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        self.code = int(code)
        name = ERROR_NAMES[code] if 0 <= code < len(ERROR_NAMES) else f"code {code}"
        super().__init__(f"{where}: {name}" if where else name)


def _load():
    # PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 and
    # links them by a different file name. If libspfft.so were loaded first,
    # the system copies would be mapped and torch would then map a second HIP
    # runtime that finds no GPU. Importing torch first makes libspfft.so bind
    # to the runtime already in the process (same SONAME).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C spfft_amd` or "
            "`python -c 'import __graft_entry__; __graft_entry__.build()'`")
    return ctypes.CDLL(LIB_PATH)


lib = _load()

_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int


def _declare(prefix: str, real):
    rp = ctypes.POINTER(real)
    rpp = ctypes.POINTER(rp)
    sig = {
        "grid_create": [ctypes.POINTER(_vp), _i, _i, _i, _i, _i, _i],
        "grid_destroy": [_vp],
        "grid_max_dim_x": [_vp, _ip], "grid_max_dim_y": [_vp, _ip], "grid_max_dim_z": [_vp, _ip],
        "grid_max_num_local_z_columns": [_vp, _ip], "grid_max_local_z_length": [_vp, _ip],
        "grid_processing_unit": [_vp, _ip], "grid_device_id": [_vp, _ip],
        "grid_num_threads": [_vp, _ip],
        "transform_create": [ctypes.POINTER(_vp), _vp, _i, _i, _i, _i, _i, _i, _i, _i, _ip],
        "transform_create_independent": [ctypes.POINTER(_vp), _i, _i, _i, _i, _i, _i, _i, _i, _ip],
        "transform_destroy": [_vp],
        "transform_clone": [_vp, ctypes.POINTER(_vp)],
        "transform_forward": [_vp, _i, _vp, _i],
        "transform_forward_ptr": [_vp, _vp, _vp, _i],
        "transform_backward": [_vp, _vp, _i],
        "transform_backward_ptr": [_vp, _vp, _vp],
        "transform_get_space_domain": [_vp, _i, ctypes.POINTER(_vp)],
        "transform_dim_x": [_vp, _ip], "transform_dim_y": [_vp, _ip], "transform_dim_z": [_vp, _ip],
        "transform_local_z_length": [_vp, _ip], "transform_local_slice_size": [_vp, _ip],
        "transform_local_z_offset": [_vp, _ip],
        "transform_global_size": [_vp, ctypes.POINTER(ctypes.c_longlong)],
        "transform_num_local_elements": [_vp, _ip],
        "transform_num_global_elements": [_vp, ctypes.POINTER(ctypes.c_longlong)],
        "transform_device_id": [_vp, _ip], "transform_num_threads": [_vp, _ip],
        "transform_execution_mode": [_vp, _ip], "transform_set_execution_mode": [_vp, _i],
        "multi_transform_forward": [_i, ctypes.POINTER(_vp), _ip, ctypes.POINTER(_vp), _ip],
        "multi_transform_forward_ptr": [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp), _ip],
        "multi_transform_backward": [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip],
        "multi_transform_backward_ptr": [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(_vp)],
    }
    fns = {}
    for name, args in sig.items():
        f = getattr(lib, prefix + name)
        f.argtypes = args
        f.restype = ctypes.c_int
        fns[name] = f
    return fns


_API = {"double": _declare("spfft_", ctypes.c_double),
        "float": _declare("spfft_float_", ctypes.c_float)}


def _check(code: int, where: str):
    if code != SPFFT_SUCCESS:
        raise SpfftError(code, where)


def _ptr(x):
    """Device/host address of a torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"cannot take the address of {type(x)!r}")


def _indices(indices):
    arr = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1))
    return arr, arr.ctypes.data_as(_ip)


class _GridBase:
    _precision = "double"

    def __init__(self, max_dim_x, max_dim_y, max_dim_z, max_num_local_z_columns,
                 processing_unit=SPFFT_PU_GPU, max_num_threads=-1):
        self._api = _API[self._precision]
        h = _vp()
        _check(self._api["grid_create"](ctypes.byref(h), max_dim_x, max_dim_y, max_dim_z,
                                        max_num_local_z_columns, processing_unit, max_num_threads),
               "grid_create")
        self.handle = h

    def _get(self, name):
        v = _i()
        _check(self._api["grid_" + name](self.handle, ctypes.byref(v)), "grid_" + name)
        return v.value

    def max_dim_x(self): return self._get("max_dim_x")
    def max_dim_y(self): return self._get("max_dim_y")
    def max_dim_z(self): return self._get("max_dim_z")
    def max_num_local_z_columns(self): return self._get("max_num_local_z_columns")
    def max_local_z_length(self): return self._get("max_local_z_length")
    def processing_unit(self): return self._get("processing_unit")
    def device_id(self): return self._get("device_id")
    def num_threads(self): return self._get("num_threads")

    def create_transform(self, processing_unit, transform_type, dim_x, dim_y, dim_z,
                         local_z_length, num_local_elements, index_format, indices):
        arr, p = _indices(indices)
        h = _vp()
        _check(self._api["transform_create"](ctypes.byref(h), self.handle, processing_unit,
                                             transform_type, dim_x, dim_y, dim_z, local_z_length,
                                             num_local_elements, index_format, p),
               "transform_create")
        return self._transform_cls._from_handle(h, self._precision, transform_type,
                                                num_local_elements)

    def close(self):
        if getattr(self, "handle", None):
            self._api["grid_destroy"](self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class _TransformBase:
    _precision = "double"

    def __init__(self, max_num_threads, processing_unit, transform_type, dim_x, dim_y, dim_z,
                 num_local_elements, index_format, indices):
        self._api = _API[self._precision]
        arr, p = _indices(indices)
        h = _vp()
        _check(self._api["transform_create_independent"](
            ctypes.byref(h), max_num_threads, processing_unit, transform_type, dim_x, dim_y,
            dim_z, num_local_elements, index_format, p), "transform_create_independent")
        self.handle = h
        self.transform_type = transform_type
        self.num_values = num_local_elements

    @classmethod
    def _from_handle(cls, handle, precision, transform_type, num_values):
        obj = cls.__new__(cls)
        obj._api = _API[precision]
        obj.handle = handle
        # The C ABI has no accessor for the type, and its num_local_elements
        # returns the local z length (reference transform.cpp:433-445), so
        # both are kept from creation.
        obj.transform_type = transform_type
        obj.num_values = num_values
        return obj

    def _get(self, name, ctype=_i):
        v = ctype()
        _check(self._api["transform_" + name](self.handle, ctypes.byref(v)), "transform_" + name)
        return v.value

    def dim_x(self): return self._get("dim_x")
    def dim_y(self): return self._get("dim_y")
    def dim_z(self): return self._get("dim_z")
    def local_z_length(self): return self._get("local_z_length")
    def local_slice_size(self): return self._get("local_slice_size")
    def local_z_offset(self): return self._get("local_z_offset")
    def global_size(self): return self._get("global_size", ctypes.c_longlong)
    def num_local_elements(self): return self._get("num_local_elements")
    def num_global_elements(self): return self._get("num_global_elements", ctypes.c_longlong)
    def device_id(self): return self._get("device_id")
    def num_threads(self): return self._get("num_threads")
    def execution_mode(self): return self._get("execution_mode")

    def set_execution_mode(self, mode):
        _check(self._api["transform_set_execution_mode"](self.handle, mode), "set_execution_mode")

    def clone(self):
        h = _vp()
        _check(self._api["transform_clone"](self.handle, ctypes.byref(h)), "transform_clone")
        return type(self)._from_handle(h, self._precision, self.transform_type, self.num_values)

    def space_domain_data(self, location):
        p = _vp()
        _check(self._api["transform_get_space_domain"](self.handle, location, ctypes.byref(p)),
               "get_space_domain")
        return p.value

    def forward(self, input_or_location, output, scaling=SPFFT_NO_SCALING):
        """forward(location:int, out, scaling) or forward(input_ptr, out, scaling)."""
        if isinstance(input_or_location, int) and input_or_location in (SPFFT_PU_HOST, SPFFT_PU_GPU):
            _check(self._api["transform_forward"](self.handle, input_or_location, _ptr(output),
                                                  scaling), "transform_forward")
        else:
            _check(self._api["transform_forward_ptr"](self.handle, _ptr(input_or_location),
                                                      _ptr(output), scaling),
                   "transform_forward_ptr")

    def backward(self, input, output_or_location):
        """backward(input, location:int) or backward(input, out_ptr)."""
        if isinstance(output_or_location, int) and output_or_location in (SPFFT_PU_HOST,
                                                                          SPFFT_PU_GPU):
            _check(self._api["transform_backward"](self.handle, _ptr(input), output_or_location),
                   "transform_backward")
        else:
            _check(self._api["transform_backward_ptr"](self.handle, _ptr(input),
                                                       _ptr(output_or_location)),
                   "transform_backward_ptr")

    def close(self):
        if getattr(self, "handle", None):
            self._api["transform_destroy"](self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Transform(_TransformBase):
    _precision = "double"


class TransformFloat(_TransformBase):
    _precision = "float"


class Grid(_GridBase):
    _precision = "double"
    _transform_cls = Transform


class GridFloat(_GridBase):
    _precision = "float"
    _transform_cls = TransformFloat


def _multi(precision, name, transforms, *arrays):
    api = _API[precision]
    n = len(transforms)
    hs = (_vp * n)(*[t.handle for t in transforms])
    conv = []
    for kind, arr in arrays:
        if kind == "ptr":
            conv.append((_vp * n)(*[_ptr(a) for a in arr]))
        else:
            conv.append((_i * n)(*[int(a) for a in arr]))
    _check(api[name](n, hs, *conv), name)


def multi_transform_forward(transforms, inputs, outputs, scalings):
    """inputs: list of pointers, or list of processing-unit locations (ints)."""
    prec = transforms[0]._precision if transforms else "double"
    if inputs and all(isinstance(x, int) and x in (SPFFT_PU_HOST, SPFFT_PU_GPU) for x in inputs):
        _multi(prec, "multi_transform_forward", transforms, ("int", inputs), ("ptr", outputs),
               ("int", scalings))
    else:
        _multi(prec, "multi_transform_forward_ptr", transforms, ("ptr", inputs), ("ptr", outputs),
               ("int", scalings))


lib.spfft_amd_multi_transform_noalloc.argtypes = [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  ctypes.POINTER(_vp), _i, _ip]
lib.spfft_amd_multi_transform_noalloc.restype = ctypes.c_int


def multi_transform_noalloc(transforms, inputs, outputs, forward, scalings=None):
    """Allocation-free batched call (include/spfft_amd/no_alloc.h): double
    GPU transforms of one shape, device pointers only; `scalings` is read for
    forward calls."""
    n = len(transforms)
    hs = (_vp * max(n, 1))(*[t.handle for t in transforms])
    ins = (_vp * max(n, 1))(*[_ptr(a) for a in inputs])
    outs = (_vp * max(n, 1))(*[_ptr(a) for a in outputs])
    sc = (_i * max(n, 1))(*[int(a) for a in (scalings or [SPFFT_NO_SCALING] * n)])
    _check(lib.spfft_amd_multi_transform_noalloc(n, hs, ins, outs, 1 if forward else 0, sc),
           "spfft_amd_multi_transform_noalloc")


def multi_transform_backward(transforms, inputs, outputs):
    """outputs: list of pointers, or list of processing-unit locations (ints)."""
    prec = transforms[0]._precision if transforms else "double"
    if outputs and all(isinstance(x, int) and x in (SPFFT_PU_HOST, SPFFT_PU_GPU) for x in outputs):
        _multi(prec, "multi_transform_backward", transforms, ("ptr", inputs), ("int", outputs))
    else:
        _multi(prec, "multi_transform_backward_ptr", transforms, ("ptr", inputs),
               ("ptr", outputs))
