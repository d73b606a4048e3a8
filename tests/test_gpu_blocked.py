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


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("n,R,reads", [(65, 16, 5), (100, 32, 3), (257, 8, 9), (1024, 32, 2)])
@pytest.mark.parametrize("best_sem", [0, 1])
def test_block_random_order_bit_exact(ptype, n, R, reads, best_sem):
    """The block-random order (update_strategy='random' of the MFMA form):
    with integer couplings the plain and the blocked kernels both equal the
    oracle's chain bit for bit."""
    rng = np.random.default_rng(3000 + n + R + ptype + 5 * best_sem)
    J, h = int_problem(n, rng, -9, 9, 3)
    J, h = J.astype(np.float32), h.astype(np.float32)
    S = 25 if n < 1000 else 6
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    kw = dict(order=3, seed=55 + n, best_sem=best_sem)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, nthreads=8, **kw)
    assert ref["accepted"].sum() > 0
    for form in (_lib.FORM_PLAIN, _lib.FORM_BLOCKED):
        got = engine_metropolis(J, h, ptype, betas, st, en, S, form=form, **kw)
        assert_same(got, ref)


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 3])
def test_headline_kernel_bit_exact(ptype, order):
    """The benchmarked instantiation (n = 1024: two bf16 pieces, eight spin
    words per matrix wave, three K ranges) on integer couplings in
    [-300, 300], which two bf16 pieces hold exactly: bit for bit against the
    oracle."""
    rng = np.random.default_rng(77 + ptype + 3 * order)
    n, R, reads, S = 1024, 32, 3, 8
    J, h = int_problem(n, rng, -300, 300, 60)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(40.0, 900.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    kw = dict(order=order, seed=9001 + order)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, nthreads=8, **kw)
    acc = ref["accepted"].sum() / (reads * R * n * S)
    assert 0.02 < acc < 0.98, acc
    got = engine_metropolis(J, h, ptype, betas, st, en, S, form=_lib.FORM_BLOCKED, **kw)
    assert_same(got, ref)


def _qubo_float(kind, n, rng):
    if kind == "gauss":
        Q = np.triu(rng.normal(size=(n, n)), 1)
        J = Q + Q.T
        h = rng.normal(size=n)
    else:  # nonnegative couplings, strongly negative fields: E far from 0
        Q = np.triu(rng.random((n, n)), 1)
        J = Q + Q.T
        h = -rng.uniform(0.2 * n, 0.3 * n, n)
    return J.astype(np.float32), h.astype(np.float32)


@pytest.mark.parametrize("kind", ["gauss", "nonneg"])
@pytest.mark.parametrize("order", [0, 3])
@pytest.mark.parametrize("form", [_lib.FORM_BLOCKED, _lib.FORM_BLOCKED_EXACT])
def test_blocked_float_qubo(kind, order, form):
    """QUBO with real couplings on the MFMA form (+-1 operands and row sums of
    the couplings the products use): tracked and best energies within 1e-5
    of the f64 energies of the returned states."""
    rng = np.random.default_rng(11 + order + (kind == "nonneg"))
    n, R, reads, S = 384, 32, 4, 400
    J, h = _qubo_float(kind, n, rng)
    betas = (1 / np.geomspace(0.05, 5.0, R)).astype(np.float32)
    if kind == "nonneg":
        betas = (1 / np.geomspace(0.5, 50.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, 1)
    en = energies_of(J, h, st, 1, np.float32)
    got = engine_metropolis(J, h, 1, betas, st, en, S, order=order, seed=4, form=form)
    e64 = oracle.energy_f64(J, h, got["states"].reshape(-1, n)).reshape(reads, R)
    rel = np.abs(got["energies"] - e64) / np.maximum(1.0, np.abs(e64))
    assert rel.max() < 1e-5, rel.max()
    b64 = oracle.energy_f64(J, h, got["best_state"])
    assert np.all(np.abs(got["best_energy"] - b64) <= 1e-5 * np.maximum(1.0, np.abs(b64)))


@pytest.mark.parametrize("order", [0, 3])
@pytest.mark.parametrize("ptype", [0, 1])
def test_blocked_energy_recompute_keeps_caller_offset(order, ptype):
    """The MFMA form recomputes the energy in dry passes (every 64 sweeps for
    float couplings; forced here every 30 sweeps with integer couplings,
    which otherwise skip them) and resets the tracked energy to it plus the
    offset the caller's initial energy had from it.  The run stays bit-exact
    against the oracle, whose tracked energy simply carries the offset
    (pysa/simulation.py:157-158), over several recomputations."""
    import os
    os.environ["PYSA_AMD_ENERGY_REFRESH"] = "30"
    try:
        _check_recompute_offset(order, ptype)
    finally:
        del os.environ["PYSA_AMD_ENERGY_REFRESH"]


def _check_recompute_offset(order, ptype):
    rng = np.random.default_rng(808 + order + 2 * ptype)
    n, R, reads, S = 130, 16, 4, 200
    J, h = int_problem(n, rng, -5, 5, 2)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32) + np.float32(7.0)  # caller's offset
    kw = dict(order=order, seed=31)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, nthreads=8, **kw)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, form=_lib.FORM_BLOCKED, **kw)
    assert_same(got, ref)
    e64 = oracle.energy_f64(J, h, got["states"].reshape(-1, n)).reshape(reads, R)
    assert np.array_equal(got["energies"].astype(np.float64), e64 + 7.0)


@pytest.mark.parametrize("form,order", [(_lib.FORM_PLAIN, 1), (_lib.FORM_BLOCKED, 3)])
def test_sweep_counter_past_32_bits(form, order):
    """sweep_offset + sweep past 2^32 keeps a distinct stream (the counter
    carries 48 sweep bits): equal to the oracle, and different from the run
    whose offset is 2^32 higher (a 32-bit counter would repeat it)."""
    rng = np.random.default_rng(61 + order)
    n, R, reads, S = 96, 8, 3, 4
    J, h = int_problem(n, rng, -3, 3, 1)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    off = (1 << 32) - 2
    ref = oracle.metropolis(J, h, 0, betas, st, en, S, order=order, seed=5, sweep_offset=off,
                            nthreads=8)
    got = engine_metropolis(J, h, 0, betas, st, en, S, order=order, seed=5, sweep_offset=off,
                            form=form)
    assert_same(got, ref)
    low = engine_metropolis(J, h, 0, betas, st, en, S, order=order, seed=5,
                            sweep_offset=off - (1 << 32) + (1 << 33), form=form)
    assert not np.array_equal(got["states"], low["states"])


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 3])
@pytest.mark.parametrize("kind", ["int", "bf16"])
def test_zero_fields_bit_exact(ptype, order, kind):
    """All local fields zero (the field sums then skip h): the MFMA form
    equals the oracle bit for bit, with int8 couplings and with couplings
    that take two bf16 pieces."""
    rng = np.random.default_rng(2400 + 4 * ptype + 2 * (order == 3) + (kind == "bf16"))
    n, R, reads, S = 300, 16, 6, 25
    # integers up to 3000 need two bf16 pieces; up to 9, int8
    big = kind == "bf16"
    J, _ = int_problem(n, rng, -3000, 3000, 0) if big else int_problem(n, rng, -9, 9, 0)
    J = J.astype(np.float32)
    h = np.zeros(n, np.float32)
    scale = 330.0 if big else 1.0
    betas = (1 / (scale * np.geomspace(8.0, 60.0, R))).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=order, seed=81, nthreads=8)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=order, seed=81,
                            form=_lib.FORM_BLOCKED)
    assert_same(got, ref)
