"""MFMA-blocked form vs the oracle and the plain form.

Integer couplings are encoded exactly (int8 or bf16 pieces), so with the same
Philox stream the blocked trajectories equal the oracle's and the plain
form's bit for bit.  Real couplings use split-bf16 pieces: checked against
f64 energies of the returned states.
"""

import numpy as np
import pytest

import oracle
from pysa_amd import _lib

from helpers import (assert_same, energies_of, engine_metropolis, float_problem, int_problem,
                     pm1_problem, random_spins)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("n,R,reads", [(32, 16, 4), (64, 16, 5), (100, 8, 9), (257, 32, 3),
                                       (1024, 32, 2), (2048, 24, 3)])
@pytest.mark.parametrize("kind", ["pm1", "int"])
def test_blocked_bit_exact_vs_oracle(ptype, n, R, reads, kind):
    rng = np.random.default_rng(2000 + n + 7 * R + ptype)
    J, h = pm1_problem(n, rng, h_range=1) if kind == "pm1" else int_problem(n, rng, -9, 9, 3)
    J, h = J.astype(np.float32), h.astype(np.float32)
    S = 30 if n < 1000 else (6 if n < 2000 else 3)
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=0, seed=77, nthreads=8)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=0, seed=77,
                            form=_lib.FORM_BLOCKED)
    assert_same(got, ref)


@pytest.mark.parametrize("best_sem", [0, 1])
def test_blocked_equals_plain(best_sem):
    rng = np.random.default_rng(5)
    n, R, reads, S = 96, 12, 11, 40
    J, h = int_problem(n, rng, -4, 4, 2)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = np.append(np.geomspace(3.0, 0.3, R - 1), 0.0).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    kw = dict(order=0, seed=123, best_sem=best_sem, log_omega=True)
    a = engine_metropolis(J, h, 0, betas, st, en, S, form=_lib.FORM_PLAIN, **kw)
    b = engine_metropolis(J, h, 0, betas, st, en, S, form=_lib.FORM_BLOCKED, **kw)
    assert_same(a, b)
    assert np.array_equal(a["log_omegas"], b["log_omegas"])


@pytest.mark.parametrize("form", [_lib.FORM_BLOCKED, _lib.FORM_BLOCKED_EXACT])
def test_blocked_float_couplings(form):
    rng = np.random.default_rng(9)
    n, R, reads, S = 512, 32, 4, 50
    J, h = float_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(0.1, 2.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    got = engine_metropolis(J, h, 0, betas, st, en, S, order=0, seed=4, form=form)
    e64 = oracle.energy_f64(J, h, got["states"].reshape(-1, n)).reshape(reads, R)
    rel = np.abs(got["energies"] - e64) / np.abs(e64)
    assert rel.max() < 1e-5, rel.max()
    b64 = oracle.energy_f64(J, h, got["best_state"])
    assert np.allclose(got["best_energy"], b64, rtol=1e-5)


@pytest.mark.parametrize("ptype", [0, 1])
def test_blocked_bf16_pieces_integer_bit_exact(ptype):
    """Integer couplings outside int8 take the bf16 path (hi + lo pieces, and
    for QUBO the +-1 operands with device row sums); their sums are exact, so
    the trajectories still equal the oracle's bit for bit."""
    rng = np.random.default_rng(31 + ptype)
    n, R, reads, S = 200, 16, 3, 20
    J, h = int_problem(n, rng, -300, 300, 40)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(20.0, 400.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=0, seed=91, nthreads=8)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=0, seed=91,
                            form=_lib.FORM_BLOCKED)
    assert ref["accepted"].sum() > 0
    assert_same(got, ref)
