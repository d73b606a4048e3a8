"""Plain HIP kernel vs the oracle (bit for bit on integer couplings).

Replay mode: host-supplied log-uniforms / permutations / PT uniforms drive both
sides.  Philox mode: both sides draw from the same counter-based stream.
"""

import numpy as np
import pytest

import oracle
from pysa_amd import _lib

from helpers import (assert_same, energies_of, engine_metropolis, float_problem, int_problem,
                     pm1_problem, random_spins, replay_arrays)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("best_sem", [0, 1])
@pytest.mark.parametrize("n", [5, 64, 130])
def test_replay_bit_exact(ptype, order, dtype, best_sem, n):
    rng = np.random.default_rng(1000 + 7 * n + 3 * ptype + order)
    J, h = int_problem(n, rng)
    J, h = J.astype(dtype), h.astype(dtype)
    reads, R, S = 3, 5, 12
    betas = np.geomspace(3.0, 0.1, R).astype(dtype)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, dtype)
    rep = replay_arrays(rng, reads, S, R, n, random_order=(order == 1))
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=order, best_sem=best_sem,
                            replay=rep)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=order, best_sem=best_sem,
                            replay=rep)
    assert_same(got, ref)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("ptype", [0, 1])
def test_philox_bit_exact(order, ptype):
    rng = np.random.default_rng(1001)
    n, reads, R, S = 64, 32, 16, 200
    J, h = pm1_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(1.0, 3.5, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=order, seed=1001, nthreads=8)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=order, seed=1001)
    assert_same(got, ref)


def test_log_omega_matches_oracle():
    rng = np.random.default_rng(5)
    n, reads, R, S = 20, 4, 6, 30
    J, h = int_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = np.append(np.geomspace(2.0, 0.2, R - 1), 0.0).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    ref = oracle.metropolis(J, h, 0, betas, st, en, S, log_omega=True, seed=9)
    got = engine_metropolis(J, h, 0, betas, st, en, S, log_omega=True, seed=9)
    assert_same(got, ref)
    assert np.array_equal(got["log_omegas"], ref["log_omegas"])


def test_float_couplings_close_to_f64():
    """Float couplings: incremental fields differ from the row dot in the last
    bits only; tracked energies stay within 1e-5 relative of f64 energies of
    the returned states."""
    rng = np.random.default_rng(11)
    n, reads, R, S = 200, 4, 8, 100
    J, h = float_problem(n, rng)
    J32, h32 = J.astype(np.float32), h.astype(np.float32)
    betas = np.geomspace(5.0, 0.5, R).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J32, h32, st, 0, np.float32)
    got = engine_metropolis(J32, h32, 0, betas, st, en, S, seed=3)
    e64 = oracle.energy_f64(J32, h32, got["states"].reshape(-1, n)).reshape(reads, R)
    rel = np.abs(got["energies"] - e64) / np.maximum(1.0, np.abs(e64))
    assert rel.max() < 1e-5, rel.max()


def test_energy_kernel_matches_oracle():
    rng = np.random.default_rng(2)
    n = 300
    J, h = float_problem(n, rng)
    st = random_spins((17, n), rng, 0)
    prob = _lib.Problem(J, h, "ising", np.float64)
    got = prob.energy(st)
    prob.close()
    ref = oracle.energy_f64(J, h, st)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)
