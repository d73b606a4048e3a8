"""Exhaustive check of the MFMA form's certified acceptance brackets.

The blocked kernels decide a proposal from a single-precision bracket
[lo, hi] around det_log(u) / beta (the hardware logarithm with a proven error
margin) and fall back to PySA's exact test log_r < beta * dn
(pysa/ising.py:52) only inside it.  Trajectories are exact iff the bracket
contains det_log(u) / beta for every 32-bit Philox word: this runs all 2^32
words on the GPU (tests/gpu_probe, a test-only library) for the betas of the
benchmark ladders and extreme ones, for both bracket forms
(accept_bounds: split-K kernel, accept_bounds_f: pipelined kernel).
"""

import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_probe", "_build",
                   "libbracket_check.so")


def _ladder(R, tmin, tmax):
    r = np.exp(1 / (R - 1) * np.log(tmax / tmin))
    return np.array([1 / (tmin * r**k) for k in range(R)], dtype=np.float32)


BETAS = sorted(set(
    [float(b) for b in _ladder(16, 1.0, 3.5)]        # C1
    + [float(b) for b in _ladder(32, 0.1, 5.0)[::4]]  # C2
    + [float(b) for b in _ladder(32, 0.3, 1.5)[::4]]  # C3
    + [1.0, 1e-30, 1e-12, 1e-6, 3e-3, 0.7071067, 1234.5, 1e6, 1e12, 1e30]))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (make -C tests/gpu_probe)")
    lib = ctypes.CDLL(LIB)
    lib.bracket_check.argtypes = [ctypes.c_float, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_ulonglong),
                                  ctypes.POINTER(ctypes.c_double)]
    lib.bracket_check.restype = ctypes.c_int
    return lib


@pytest.mark.parametrize("variant", [0, 1])
def test_brackets_contain_exact_threshold_for_every_word(lib, variant):
    widths = {}
    for beta in BETAS:
        bad = ctypes.c_ulonglong(0)
        wid = ctypes.c_double(0)
        rc = lib.bracket_check(ctypes.c_float(beta), variant, ctypes.byref(bad), ctypes.byref(wid))
        assert rc == 0, rc
        assert bad.value == 0, (beta, variant, bad.value)
        widths[beta] = wid.value
    # widest bracket relative to |det_log(u)| (for words near u = 1 the
    # absolute part of the margin dominates, so this is large by design)
    print({f"{b:.4g}": f"{w:.3g}" for b, w in widths.items()})
