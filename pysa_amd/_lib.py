"""ctypes binding of libpysa_amd.so (the C ABI declared in include/pysa_amd.h).

The library is built in-tree (``make -C pysa_amd/csrc``) and loaded from the
package directory.  There is no CPU fallback: if the library or a GPU is
missing, every compute entry point raises ``RuntimeError``.
"""

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpysa_amd.so")

FORM_PLAIN, FORM_BLOCKED, FORM_AUTO, FORM_BLOCKED_EXACT, FORM_SPARSE = 0, 1, 2, 3, 4
ORDER = {"sequential": 0, "random": 1, "random_shared": 2, "block_random": 3}

_lib = None
_lock = threading.Lock()

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p
_c_dbl = ctypes.c_double

# name -> argtypes; every function returns int32 status unless noted.
_SIGNATURES = {
    "pysa_amd_problem_create": [_c_i32, _c_i32, _c_i32, _vp, _vp, ctypes.POINTER(_c_i64)],
    "pysa_amd_problem_destroy": [_c_i64],
    "pysa_amd_metropolis": [_c_i64, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32,
                            _vp, _vp, _c_u64, _c_u64, _c_i32, _c_i32, _vp, _vp, _vp,
                            _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_c_dbl)],
    "pysa_amd_metropolis_device": [_c_i64, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32,
                                   _c_i32, _vp, _vp, _c_u64, _c_u64, _c_i32, _c_i32, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   ctypes.POINTER(_c_dbl), _vp],
    "pysa_amd_metropolis_device_noalloc": [_c_i64, _c_i32, _c_i32, _c_i32, _c_i32, _c_i32,
                                           _c_i32, _c_i32, _vp, _vp, _c_u64, _c_u64, _c_i32,
                                           _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pysa_amd_ais": [_c_i64, _vp, _c_i32, _c_i32, _c_i32, _c_dbl, _c_i32, _c_i32, _c_u64,
                     _vp, _vp, _vp, _vp, _vp, _c_i32, ctypes.POINTER(_c_dbl)],
    "pysa_amd_energy": [_c_i64, _vp, _c_i64, _vp],
    "pysa_amd_plain_fields": [_c_i64, _c_i64, _vp],
    "pysa_amd_problem_set_shape": [_c_i64, _c_i32, _c_i32],
    "pysa_amd_problem_create_device": [_c_i32, _c_i32, _c_i32, _vp, _vp, ctypes.POINTER(_c_i64)],
    "pysa_amd_device_info": [ctypes.POINTER(_c_i32), ctypes.c_char_p, _c_i32],
    "pysa_amd_batched_device": [_c_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i32, _c_i32,
                                _c_i32, _c_i32, _c_i32, _c_i32, _c_i32, _vp, _c_u64, _c_u64,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_i32,
                                ctypes.POINTER(_c_dbl), _vp],
    "pysa_amd_malloc": [_c_u64, ctypes.POINTER(_vp)],
    "pysa_amd_free": [_vp],
    "pysa_amd_memcpy_h2d": [_vp, _vp, _c_u64],
    "pysa_amd_memcpy_d2h": [_vp, _vp, _c_u64],
    "pysa_amd_memcpy_d2d": [_vp, _vp, _c_u64],
}

EXPORTED = sorted(list(_SIGNATURES) + ["pysa_amd_last_error", "pysa_amd_version"])


def load(path=LIB_PATH):
    """Load the shared library (once) and declare its signatures."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"libpysa_amd.so not found at {path}; build it with `make -C pysa_amd/csrc` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = _c_i32
        lib.pysa_amd_last_error.restype = ctypes.c_char_p
        lib.pysa_amd_last_error.argtypes = []
        lib.pysa_amd_version.restype = ctypes.c_char_p
        lib.pysa_amd_version.argtypes = []
        _lib = lib
        return lib


def device_info():
    """(compute units, architecture name) of the current device."""
    cus = _c_i32(0)
    buf = ctypes.create_string_buffer(64)
    check(load().pysa_amd_device_info(ctypes.byref(cus), buf, 64))
    return int(cus.value), buf.value.decode()


def check(rc):
    """Raise RuntimeError carrying the library's message for a failed call."""
    if rc != 0:
        msg = _lib.pysa_amd_last_error().decode() if _lib is not None else ""
        raise RuntimeError(f"pysa_amd error {rc}: {msg}")


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the engine must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def dtype_code(float_type):
    dt = np.dtype(float_type)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise ValueError(f"float_type={float_type!r} not supported (float32 or float64).")


class Problem:
    """A dense problem resident on the current HIP device."""

    def __init__(self, couplings, local_fields, problem_type, float_type):
        lib = load()
        self.float_type = np.dtype(float_type)
        self.n = int(len(local_fields))
        self.ptype = 0 if problem_type == "ising" else 1
        self.J = np.ascontiguousarray(couplings, dtype=self.float_type)
        self.h = np.ascontiguousarray(local_fields, dtype=self.float_type)
        handle = _c_i64(0)
        check(lib.pysa_amd_problem_create(self.n, self.ptype, dtype_code(self.float_type),
                                          ptr(self.J), ptr(self.h), ctypes.byref(handle)))
        self.handle = handle.value

    def close(self):
        if getattr(self, "handle", 0):
            load().pysa_amd_problem_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def energy(self, states):
        """f64 energies of int8 states (count x n)."""
        states = np.ascontiguousarray(states, dtype=np.int8).reshape(-1, self.n)
        out = np.empty(len(states), dtype=np.float64)
        check(load().pysa_amd_energy(self.handle, ptr(states), len(states), ptr(out)))
        return out

    def set_shape(self, kernel=0, tiles=0):
        """Launch shape of the MFMA form on this problem: kernel 0 = by size,
        1 = pipelined, 2 = split-K; tiles of 32 replicas per workgroup 0 = by
        occupancy, 1 or 2.  Results do not depend on it."""
        check(load().pysa_amd_problem_set_shape(self.handle, int(kernel), int(tiles)))

    def plain_fields(self, count):
        """Local fields J s of the first `count` chains after the last
        plain-form run (count x n, problem dtype)."""
        out = np.empty((count, self.n), dtype=self.float_type)
        check(load().pysa_amd_plain_fields(self.handle, count, ptr(out)))
        return out


class DeviceArray:
    """A device buffer owned by the library's HIP runtime, shaped like a
    numpy array (for callers that keep inputs resident in HBM)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(load().pysa_amd_malloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(load().pysa_amd_memcpy_h2d(self.ptr, ptr(a), self.nbytes))

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        check(load().pysa_amd_memcpy_d2d(self.ptr, other.ptr, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(load().pysa_amd_memcpy_d2h(ptr(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            load().pysa_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def blocked_available():
    """Whether this build carries the MFMA-blocked form."""
    return BLOCKED_READY


# The blocked form is selected by form='auto' only once it is built.
BLOCKED_READY = True
