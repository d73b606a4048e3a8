"""The allocation-free entry point (pysa_amd_metropolis_device_noalloc)
against pysa_amd_metropolis_device on the same inputs: bit for bit, on two
calls in a row (the second with fewer reads), float couplings (two bf16
pieces with piece-sum in-block couplings) and integer couplings, both visit
orders of the MFMA form, both best semantics."""

import ctypes

import numpy as np
import pytest

from pysa_amd import _lib

pytestmark = pytest.mark.gpu

KEYS = ("states", "energies", "beta_idx", "best_state", "best_energy", "best_sweeps", "accepted")


def _problem(kind, n, rng):
    if kind == "float":
        W = rng.normal(size=(n, n // 2))
        J = (W @ W.T / n).astype(np.float32)
    else:
        A = np.triu(rng.integers(-9, 10, (n, n)), 1)
        J = (A + A.T).astype(np.float32)
    np.fill_diagonal(J, 0)
    h = (rng.normal(size=n) if kind == "float" else rng.integers(-2, 3, n)).astype(np.float32)
    return J, h


class Batch:
    """Device buffers of one call (allocated by the test, not the engine)."""

    def __init__(self, prob, reads, R, rng, ptype):
        D = _lib.DeviceArray
        n = prob.n
        b = rng.integers(0, 2, (reads, R, n))
        st = (b if ptype == 1 else 2 * b - 1).astype(np.int8)
        en = prob.energy(st.reshape(-1, n)).astype(np.float32).reshape(reads, R)
        self.st0, self.en0 = st, en
        self.betas = (1 / np.geomspace(0.3, 3.0, R)).astype(np.float32)
        self.d = {
            "betas": D.from_numpy(self.betas),
            "bsort": D.from_numpy(np.argsort(self.betas, kind="quicksort").astype(np.int32)),
            "states": D.from_numpy(st), "energies": D.from_numpy(en),
            "beta_idx": D((reads, R), np.int64), "best_state": D((reads, n), np.int8),
            "best_energy": D(reads, np.float32), "best_sweeps": D(reads, np.int32),
            "accepted": D(reads, np.uint64),
        }
        self.reads, self.R = reads, R

    def reset(self):
        self.d["states"].upload(self.st0)
        self.d["energies"].upload(self.en0)

    def args(self):
        d = self.d
        return (d["states"].ptr, d["energies"].ptr, d["beta_idx"].ptr, d["best_state"].ptr,
                d["best_energy"].ptr, d["best_sweeps"].ptr)

    def outputs(self):
        return {k: self.d[k].numpy() for k in KEYS}


def run_alloc(prob, bt, S, order, best_sem, seed, form):
    bt.reset()
    kms = _lib._c_dbl(0)
    d = bt.d
    _lib.check(_lib.load().pysa_amd_metropolis_device(
        prob.handle, bt.reads, bt.R, S, order, 1, 0, best_sem, d["betas"].ptr, d["bsort"].ptr,
        seed, 0, form, 0, None, None, None, *bt.args(), None, d["accepted"].ptr,
        ctypes.byref(kms), None))
    return bt.outputs()


def run_noalloc(prob, bt, S, order, best_sem, seed, form):
    bt.reset()
    d = bt.d
    _lib.check(_lib.load().pysa_amd_metropolis_device_noalloc(
        prob.handle, bt.reads, bt.R, S, order, 1, 0, best_sem, d["betas"].ptr, d["bsort"].ptr,
        seed, 0, form, *bt.args(), None, d["accepted"].ptr, None))
    return bt.outputs()  # device-to-host copies on the default stream wait for the run


@pytest.mark.parametrize("kind", ["float", "int"])
@pytest.mark.parametrize("order", [0, 3])
@pytest.mark.parametrize("best_sem", [0, 1])
def test_noalloc_matches_allocating_entry(kind, order, best_sem):
    rng = np.random.default_rng(71 + (kind == "int") + 2 * order + 7 * best_sem)
    n, R, S = 320, 32, 40
    J, h = _problem(kind, n, rng)
    prob = _lib.Problem(J, h, "ising", np.float32)
    form = _lib.FORM_BLOCKED
    # a fresh problem has no MFMA encoding yet: refused, nothing launched
    small = Batch(prob, 2, R, np.random.default_rng(5), 0)
    rc = _lib.load().pysa_amd_metropolis_device_noalloc(
        prob.handle, 2, R, S, order, 1, 0, best_sem, small.d["betas"].ptr, small.d["bsort"].ptr,
        9, 0, form, *small.args(), None, None, None)
    assert rc == -3, rc  # PYSA_AMD_ENOTSUP
    for reads in (12, 5):  # two calls in a row, the second smaller
        bt = Batch(prob, reads, R, np.random.default_rng(reads), 0)
        want = run_alloc(prob, bt, S, order, best_sem, 1234 + reads, form)
        got = run_noalloc(prob, bt, S, order, best_sem, 1234 + reads, form)
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (reads, k)
        assert want["accepted"].sum() > 0
    prob.close()


def test_noalloc_rejects_oversized_batch():
    rng = np.random.default_rng(3)
    n, R = 2048, 64
    J, h = _problem("int", n, rng)
    prob = _lib.Problem(J, h, "ising", np.float32)
    bt = Batch(prob, 2, R, rng, 0)
    run_alloc(prob, bt, 1, 0, 0, 1, _lib.FORM_BLOCKED)  # encodes the couplings
    reads = (32 << 20) // (R * (n + 8)) + 1  # one read past the 32 MiB scratch
    D = _lib.DeviceArray
    st = D((reads, R, n), np.int8)
    rc = _lib.load().pysa_amd_metropolis_device_noalloc(
        prob.handle, reads, R, 1, 0, 1, 0, 0, bt.d["betas"].ptr, bt.d["bsort"].ptr, 1, 0,
        _lib.FORM_BLOCKED, st.ptr, st.ptr, st.ptr, st.ptr, st.ptr, st.ptr, None, None, None)
    assert rc == -1, rc  # PYSA_AMD_EINVAL, nothing launched
    assert b"exceeds" in _lib.load().pysa_amd_last_error()
    st.free()
    prob.close()
