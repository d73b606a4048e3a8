"""The benchmark configurations as parity cases (sizes reduced only where the
oracle would take minutes).

  C1  dense +-1 Ising N=64, 32 reads x 16 replicas, T 1.0 -> 3.5, 1000
      sweeps, sequential: full size, plain and blocked forms bit for bit
      against the oracle.
  C2  QUBO N=512, Gaussian couplings, random order, T 0.1 -> 5.0: 4 reads x
      32 replicas x 60 sweeps; float couplings, so energies are checked
      against f64 recomputation (1e-5 relative), for both forms.
  C3  planted Wishart N=1024: full N and replica count, 8 reads x 1000
      sweeps; tracked energies within 1e-5 of f64 for both forms, best
      energies never below the planted ground state, the two forms' final
      cold-slot energies from the same distribution.
"""

import numpy as np
import pytest

import oracle
from pysa_amd import _lib, instances

from helpers import engine_metropolis

pytestmark = pytest.mark.gpu


def ladder(R, tmin, tmax):
    r = np.exp(1 / (R - 1) * np.log(tmax / tmin))
    return np.array([1 / (tmin * r**k) for k in range(R)], dtype=np.float32)


def split(P):
    J = P.astype(np.float32).copy()
    np.fill_diagonal(J, 0)
    return J, np.diag(P).astype(np.float32).copy()


def f64_energies(J, h, st):
    return oracle.energy_f64(J, h, st.reshape(-1, st.shape[-1])).reshape(st.shape[:-1])


def test_c1_full_size_bit_exact():
    J, h = split(instances.pm1_ising(64, 1001))
    reads, R, S = 32, 16, 1000
    rng = np.random.default_rng(1001)
    st = (2 * rng.integers(0, 2, (reads, R, 64)) - 1).astype(np.int8)
    en = f64_energies(J, h, st).astype(np.float32)
    betas = ladder(R, 1.0, 3.5)
    want = oracle.metropolis(J, h, 0, betas, st, en, S, order=0, seed=1001, nthreads=8)
    for form in (_lib.FORM_PLAIN, _lib.FORM_BLOCKED):
        got = engine_metropolis(J, h, 0, betas, st, en, S, order=0, seed=1001, form=form)
        for k in ("states", "energies", "beta_idx", "best_state", "best_energy", "best_sweeps",
                  "accepted"):
            assert np.array_equal(got[k], want[k]), (form, k)


def test_c2_qubo_random_order():
    P = instances.gaussian_qubo(512, 1002)
    J, h = split(P)
    reads, R, S = 4, 32, 60
    rng = np.random.default_rng(1002)
    st = rng.integers(0, 2, (reads, R, 512)).astype(np.int8)
    en = f64_energies(J, h, st).astype(np.float32)
    betas = ladder(R, 0.1, 5.0)
    for form, order in ((_lib.FORM_PLAIN, 1), (_lib.FORM_BLOCKED, 3)):
        got = engine_metropolis(J, h, 1, betas, st, en, S, order=order, seed=1002, form=form)
        e64 = f64_energies(J, h, got["states"])
        rel = np.abs(got["energies"] - e64) / np.maximum(1.0, np.abs(e64))
        assert rel.max() < 1e-5, (form, rel.max())
        be64 = f64_energies(J, h, got["best_state"][:, None, :])[:, 0]
        assert np.all(np.abs(got["best_energy"] - be64) <= 1e-5 * np.maximum(1.0, np.abs(be64)))
    # (the forms' energy distributions are compared in test_gpu_statistics)


@pytest.mark.parametrize("strategy", ["sequential", "random"])
def test_c3_wishart_full_n(strategy):
    P, egs = instances.planted_wishart(1024, 0.5, 1003)
    J, h = split(P)
    reads, R, S = 8, 32, 1000
    rng = np.random.default_rng(1003)
    st = (2 * rng.integers(0, 2, (reads, R, 1024)) - 1).astype(np.int8)
    en = f64_energies(J, h, st).astype(np.float32)
    betas = ladder(R, 0.3, 1.5)
    errs = {}
    # 'random': the plain form's per-replica permutation, the MFMA form's
    # block-random order (what Solver's update_strategy='random' runs)
    orders = {_lib.FORM_PLAIN: 0 if strategy == "sequential" else 1,
              _lib.FORM_BLOCKED: 0 if strategy == "sequential" else 3}
    for form in (_lib.FORM_PLAIN, _lib.FORM_BLOCKED):
        got = engine_metropolis(J, h, 0, betas, st, en, S, order=orders[form], seed=1003,
                                form=form)
        e64 = f64_energies(J, h, got["states"])
        errs[form] = np.abs(got["energies"] - e64) / np.maximum(1.0, np.abs(e64))
        be64 = f64_energies(J, h, got["best_state"][:, None, :])[:, 0]
        assert np.all(be64 >= egs - 1e-6 * abs(egs)), form
        errs[(form, "best")] = np.abs(got["best_energy"] - be64) / np.abs(be64)
    print({k: float(v.max()) for k, v in errs.items()})
    for form in (_lib.FORM_PLAIN, _lib.FORM_BLOCKED):
        rel = errs[form]
        # The tracked energy adds each sweep's accepted dE and rounds to f32
        # once per sweep (pysa/simulation.py:157-158).  The plain form resets
        # it to the f64 energy with every field refresh (128 sweeps), so the
        # drift of its f32 fields does not accumulate.  The blocked form's
        # products and in-block updates use the same hi+lo bf16 couplings,
        # so its energy changes telescope to the energy of those couplings,
        # within ~1e-7 of the f64 energy.
        assert rel.max() < 1e-5, (form, rel.max())
        assert errs[(form, "best")].max() < 1e-5, form


def test_c3_energy_error_full_length():
    """The benchmark's full length, 5000 sweeps of the planted Wishart
    instance with update_strategy='random': tracked and best energies
    within 1e-5 of f64 for both forms (the plain form resets its energy
    with each f64 field refresh, the MFMA form in a dry recomputation pass
    every 256 sweeps)."""
    P, egs = instances.planted_wishart(1024, 0.5, 1003)
    J, h = split(P)
    reads, R, S = 4, 32, 5000
    rng = np.random.default_rng(2003)
    st = (2 * rng.integers(0, 2, (reads, R, 1024)) - 1).astype(np.int8)
    en = f64_energies(J, h, st).astype(np.float32)
    betas = ladder(R, 0.3, 1.5)
    for form, order in ((_lib.FORM_PLAIN, 1), (_lib.FORM_BLOCKED, 3)):
        got = engine_metropolis(J, h, 0, betas, st, en, S, order=order, seed=2003, form=form)
        e64 = f64_energies(J, h, got["states"])
        rel = np.abs(got["energies"] - e64) / np.abs(e64)
        assert rel.max() < 1e-5, (form, rel.max())
        be64 = f64_energies(J, h, got["best_state"][:, None, :])[:, 0]
        assert np.all(np.abs(got["best_energy"] - be64) / np.abs(be64) < 1e-5), form
