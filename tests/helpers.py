"""Shared test helpers: problem generators and engine calls through the C ABI."""

import numpy as np

from pysa_amd import _lib


def pm1_problem(n, rng, h_range=0):
    """Symmetric zero-diagonal +-1 couplings and integer fields."""
    A = np.triu(rng.integers(0, 2, (n, n)) * 2 - 1, 1)
    J = A + A.T
    h = rng.integers(-h_range, h_range + 1, n) if h_range else np.zeros(n, dtype=np.int64)
    return J, h


def int_problem(n, rng, lo=-3, hi=3, h_range=2):
    A = np.triu(rng.integers(lo, hi + 1, (n, n)), 1)
    J = A + A.T
    h = rng.integers(-h_range, h_range + 1, n)
    return J, h


def float_problem(n, rng):
    """The reference tests' generator (tests_ising_sequential.py:86-97)."""
    c = 2 * rng.random((n, n)) - 1
    c = (c + c.T) / 2
    h = np.copy(np.diagonal(c))
    np.fill_diagonal(c, 0)
    return c, h


def energies_of(J, h, states, ptype, dtype):
    """f64 energies rounded to dtype (states int8 (..., n))."""
    s = states.astype(np.float64)
    Jd = J.astype(np.float64)
    hd = h.astype(np.float64)
    e = 0.5 * np.einsum("...i,ij,...j->...", s, Jd, s) + s @ hd
    return e.astype(dtype)


def random_spins(shape, rng, ptype):
    b = rng.integers(0, 2, shape)
    return (b if ptype == 1 else 2 * b - 1).astype(np.int8)


def engine_metropolis(J, h, ptype, betas, states, energies, sweeps, *, order=0, use_pt=True,
                      log_omega=False, best_sem=0, seed=0, sweep_offset=0, replay=None,
                      form=_lib.FORM_PLAIN, block_b=0, problem=None):
    """Run the HIP engine through the C ABI; returns the same dict as oracle.metropolis."""
    ft = np.dtype(J.dtype)
    own = problem is None
    prob = problem or _lib.Problem(J, h, "ising" if ptype == 0 else "qubo", ft)
    st = np.ascontiguousarray(states, dtype=np.int8).copy()
    reads, R, n = st.shape
    en = np.ascontiguousarray(energies, dtype=ft).copy()
    betas = np.ascontiguousarray(betas, dtype=ft)
    bsort = np.argsort(betas, kind="quicksort").astype(np.int32)
    bi = np.empty((reads, R), np.int64)
    bs = np.empty((reads, n), np.int8)
    be = np.empty(reads, ft)
    bw = np.empty(reads, np.int32)
    lw = np.zeros((reads, sweeps), np.float64)
    acc = np.zeros(reads, np.uint64)
    lr = pm = pt = None
    if replay is not None:
        lr = np.ascontiguousarray(replay["log_r"], dtype=np.float64)
        if replay.get("perm") is not None:
            pm = np.ascontiguousarray(replay["perm"], dtype=np.int32)
        if replay.get("pt_logu") is not None:
            pt = np.ascontiguousarray(replay["pt_logu"], dtype=np.float64)
    kms = _lib._c_dbl(0)
    try:
        _lib.check(_lib.load().pysa_amd_metropolis(
            prob.handle, reads, R, sweeps, order, int(use_pt), int(log_omega), best_sem,
            _lib.ptr(betas), _lib.ptr(bsort), seed & (2**64 - 1), sweep_offset, form, block_b,
            _lib.ptr(lr), _lib.ptr(pm), _lib.ptr(pt), _lib.ptr(st), _lib.ptr(en), _lib.ptr(bi),
            _lib.ptr(bs), _lib.ptr(be), _lib.ptr(bw), _lib.ptr(lw) if log_omega else None,
            _lib.ptr(acc), _lib.ctypes.byref(kms)))
    finally:
        if own:
            prob.close()
    return dict(states=st, energies=en, beta_idx=bi, best_state=bs, best_energy=be,
                best_sweeps=bw, log_omegas=lw, accepted=acc, kernel_ms=kms.value)


def replay_arrays(rng, reads, S, R, n, random_order=False, pt=True):
    rep = dict(log_r=np.log(rng.random((reads, S, R, n))))
    if random_order:
        rep["perm"] = np.argsort(rng.random((reads, S, R, n)), axis=-1).astype(np.int32)
    if pt and R > 1:
        rep["pt_logu"] = np.log(rng.random((reads, S, R - 1)))
    return rep


def assert_same(a, b, keys=("states", "energies", "beta_idx", "best_state", "best_energy",
                            "best_sweeps", "accepted")):
    for k in keys:
        assert np.array_equal(a[k], b[k]), f"{k} differs"


def engine_ais(J, h, ptype, temps, states, *, ref=0, p=0.5, order=0, ordering=0, seed=0,
               replay=None, energies=False, form=_lib.FORM_PLAIN):
    """AIS through the C ABI; returns (states, log_w) like oracle.ais, and the
    tracked final energies as a third item when `energies`."""
    ft = np.dtype(J.dtype)
    prob = _lib.Problem(J, h, "ising" if ptype == 0 else "qubo", ft)
    st = np.ascontiguousarray(states, dtype=np.int8).copy()
    chains = st.shape[0]
    temps = np.ascontiguousarray(temps, dtype=ft)
    lw = np.empty(chains, np.float64)
    en = np.empty(chains, np.float64)
    lr = pm = None
    if replay is not None:
        lr = np.ascontiguousarray(replay["log_r"], dtype=np.float64)
        if replay.get("perm") is not None:
            pm = np.ascontiguousarray(replay["perm"], dtype=np.int32)
    kms = _lib._c_dbl(0)
    try:
        _lib.check(_lib.load().pysa_amd_ais(
            prob.handle, _lib.ptr(temps), len(temps), chains, ref, p, order, ordering,
            seed & (2**64 - 1), _lib.ptr(lr), _lib.ptr(pm), _lib.ptr(st), _lib.ptr(lw),
            _lib.ptr(en), form, _lib.ctypes.byref(kms)))
    finally:
        prob.close()
    return (st, lw, en) if energies else (st, lw)
