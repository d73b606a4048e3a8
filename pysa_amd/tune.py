"""Autotuner for the MFMA form's launch shape.

What is tuned.  The block size B of the MFMA form is the MFMA tile, 32: the
proposals of a block run one after another per replica with B(B-1)/2
in-block field updates, so a larger block lengthens the sequential chain
(N(B-1)/2 updates per sweep) while the matrix work per sweep stays N^2 per
replica; B = 32 is therefore fixed and `block_size` only accepts it.  The
whole sweep loop (sweeps, exchanges, best tracking) runs in one persistent
launch, so there is no sweeps-per-graph count either.  What the tuner
chooses is how a batch is laid on the device:

  kernel  the pipelined kernel (proposal wave, three matrix waves, four
          helper waves) or the split-K kernel (sequential order only);
  tiles   replicas per workgroup of the pipelined kernel, one or two tiles
          of 32: two tiles share each coupling fragment between 64 replicas,
          one tile doubles the workgroups when a batch would leave compute
          units idle.

Results do not depend on the shape (the kernels make the same decisions on
the same Philox stream; tests check this bit for bit), only the speed.

How.  For a (n, float type, R, reads, order, device) key the tuner times
every valid shape on the caller's problem and batch size (random spins, a
few sweeps, device time of the engine call, best of two), keeps the fastest
and stores it in a JSON cache: $PYSA_AMD_TUNE_CACHE, else
~/.cache/pysa_amd/tune.json.  PYSA_AMD_TUNE=0 skips tuning and leaves the
library's occupancy rule (one tile when two-tile workgroups would leave
compute units idle).
"""

import json
import os
import threading

import numpy as np

from . import _lib

_lock = threading.Lock()
_mem = {}  # cache file -> its entries, as loaded or written by this process


def cache_path():
    """The JSON cache file."""
    p = os.environ.get("PYSA_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "pysa_amd", "tune.json")


def device_key():
    cus, arch = _lib.device_info()
    return f"{arch}:{cus}cu"


def shape_key(n, R, reads, order, float_type="float32"):
    return f"n={n}/{float_type}/R={R}/reads={reads}/order={order}/{device_key()}"


def candidates(n, R, order):
    """Valid (kernel, tiles) shapes: kernel 1 = pipelined, 2 = split-K."""
    nblk = (n + 31) // 32
    tiles = (1, 2) if R <= 32 else (2,)
    out = []
    if nblk >= 3 or order == _lib.ORDER["block_random"]:
        out += [(1, t) for t in tiles]
    if order == _lib.ORDER["sequential"]:
        out.append((2, 0))
    return out


def _load(path):
    if path in _mem:
        return _mem[path]
    try:
        with open(path) as f:
            d = json.load(f)
        d = d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        d = {}
    _mem[path] = d
    return d


def _store(path, d):
    _mem[path] = d
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = f"{path}.tmp{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune in the next process


def timing_sweeps(n, R, reads):
    """Sweeps per timed call: about 2e9 spin updates, between 2 and 50."""
    return int(max(2, min(50, 2e9 // max(1, n * n * R * reads))))


def time_shape(problem, R, reads, order, kernel, tiles, sweeps, reps=2, seed=0):
    """Device milliseconds of one MFMA-form engine call in this shape."""
    problem.set_shape(kernel, tiles)
    n = problem.n
    rng = np.random.default_rng(seed)
    ptype = problem.ptype
    b = rng.integers(0, 2, (reads, R, n))
    states0 = (b if ptype == 1 else 2 * b - 1).astype(np.int8)
    betas = (1.0 / np.geomspace(0.3, 1.5, R)).astype(np.float32)
    bsort = np.argsort(betas, kind="quicksort").astype(np.int32)
    best = float("inf")
    for k in range(reps):
        st = states0.copy()
        en = np.zeros((reads, R), np.float32)
        bi = np.empty((reads, R), np.int64)
        bs = np.empty((reads, n), np.int8)
        be = np.empty(reads, np.float32)
        bw = np.empty(reads, np.int32)
        kms = _lib._c_dbl(0.0)
        _lib.check(_lib.load().pysa_amd_metropolis(
            problem.handle, reads, R, sweeps, order, 1, 0, 0, _lib.ptr(betas), _lib.ptr(bsort),
            1234 + k, 0, _lib.FORM_BLOCKED, 32, None, None, None, _lib.ptr(st), _lib.ptr(en),
            _lib.ptr(bi), _lib.ptr(bs), _lib.ptr(be), _lib.ptr(bw), None, None,
            _lib.ctypes.byref(kms)))
        best = min(best, kms.value)
    return best


def tune(problem, R, reads, order, force=False, path=None):
    """The fastest shape for this problem and batch (from the cache unless
    force), applied to the problem; returns its cache entry."""
    path = path or cache_path()
    key = shape_key(problem.n, R, reads, order, np.dtype(problem.float_type).name)
    with _lock:
        cache = _load(path)
        entry = cache.get(key)
        if entry is None or force:
            sweeps = timing_sweeps(problem.n, R, reads)
            ms = {}
            for kernel, tiles in candidates(problem.n, R, order):
                ms[f"{kernel}/{tiles}"] = time_shape(problem, R, reads, order, kernel, tiles,
                                                     sweeps)
            best = min(ms, key=ms.get)
            kernel, tiles = (int(v) for v in best.split("/"))
            entry = {"kernel": kernel, "tiles": tiles, "ms": ms, "sweeps": sweeps}
            cache = dict(cache)
            cache[key] = entry
            _store(path, cache)
        problem.set_shape(entry["kernel"], entry["tiles"])
        return entry


def apply(problem, R, reads, order):
    """Set the problem's launch shape for a run: the tuned one, or the
    library's occupancy rule when PYSA_AMD_TUNE=0."""
    if os.environ.get("PYSA_AMD_TUNE", "1") == "0":
        problem.set_shape(0, 0)
        return None
    return tune(problem, R, reads, order)
