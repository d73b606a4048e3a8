"""Launch shapes of the MFMA form and the autotuner that picks them.

The shape (pipelined or split-K kernel, one or two tiles of 32 replicas per
workgroup) changes only how the work is laid on the device: with exact
couplings every shape gives the oracle's trajectories bit for bit, and with
float couplings the two tile counts give identical bits (each replica's
products and chain are the same instructions).
"""

import json

import numpy as np
import pytest

import oracle
from pysa_amd import _lib, tune
from pysa_amd.sa import Solver

from helpers import (assert_same, energies_of, engine_metropolis, float_problem, int_problem,
                     random_spins)

pytestmark = pytest.mark.gpu


def _run(prob, J, h, ptype, betas, st, en, S, order, seed, kernel, tiles):
    prob.set_shape(kernel, tiles)
    return engine_metropolis(J, h, ptype, betas, st, en, S, order=order, seed=seed,
                             form=_lib.FORM_BLOCKED, problem=prob)


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("n,R,reads", [(257, 32, 5), (1024, 16, 6), (300, 48, 3)])
@pytest.mark.parametrize("order", [0, 3])
def test_every_shape_equals_oracle(ptype, n, R, reads, order):
    rng = np.random.default_rng(41 + n + R + 3 * ptype + order)
    J, h = int_problem(n, rng, -9, 9, 3)
    J, h = J.astype(np.float32), h.astype(np.float32)
    S = 12 if n < 1000 else 5
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=order, seed=19, nthreads=8)
    prob = _lib.Problem(J, h, "ising" if ptype == 0 else "qubo", np.float32)
    try:
        shapes = tune.candidates(n, R, order) + [(0, 0)]
        for kernel, tiles in shapes:
            got = _run(prob, J, h, ptype, betas, st, en, S, order, 19, kernel, tiles)
            assert_same(got, ref)
    finally:
        prob.close()


def test_tiles_identical_bits_float_couplings():
    rng = np.random.default_rng(77)
    n, R, reads, S = 512, 16, 6, 40
    J, h = float_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(0.1, 2.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    prob = _lib.Problem(J, h, "ising", np.float32)
    try:
        a = _run(prob, J, h, 0, betas, st, en, S, 3, 5, 1, 1)
        b = _run(prob, J, h, 0, betas, st, en, S, 3, 5, 1, 2)
    finally:
        prob.close()
    assert_same(a, b)


def test_block_size_other_than_32_is_refused():
    rng = np.random.default_rng(3)
    n, R, reads = 96, 8, 2
    J, h = int_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    betas = (1 / np.geomspace(0.8, 4.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    with pytest.raises(RuntimeError, match="block size must be 32"):
        engine_metropolis(J, h, 0, betas, st, en, 4, form=_lib.FORM_BLOCKED, block_b=64)
    engine_metropolis(J, h, 0, betas, st, en, 4, form=_lib.FORM_BLOCKED, block_b=32)
    with pytest.raises(ValueError, match="block_size"):
        Solver(J + np.diag(h), "ising", "float32").metropolis_update(
            num_sweeps=2, num_reads=2, num_replicas=R, block_size=64)


def test_tuner_caches_a_valid_shape(tmp_path):
    path = str(tmp_path / "tune.json")
    rng = np.random.default_rng(8)
    n, R, reads = 256, 32, 8
    J, h = int_problem(n, rng)
    prob = _lib.Problem(J.astype(np.float32), h.astype(np.float32), "ising", np.float32)
    try:
        order = _lib.ORDER["block_random"]
        entry = tune.tune(prob, R, reads, order, path=path)
        assert (entry["kernel"], entry["tiles"]) in tune.candidates(n, R, order)
        assert set(entry["ms"]) == {f"{k}/{t}" for k, t in tune.candidates(n, R, order)}
        with open(path) as f:
            saved = json.load(f)
        key = tune.shape_key(n, R, reads, order)
        assert saved[key] == entry
        # a second call reads the cache (no new timings)
        assert tune.tune(prob, R, reads, order, path=path) is not None
    finally:
        prob.close()


def test_solver_tuned_and_untuned_agree():
    """The Solver's default (tuned shape) and block_size=32 (the library's
    shape) give the same samples for integer couplings."""
    rng = np.random.default_rng(12)
    n = 160
    J, h = int_problem(n, rng, -5, 5, 2)
    P = (J + np.diag(h)).astype(np.float32)
    out = []
    for bs in (0, 32):
        np.random.seed(5)
        df = Solver(P, "ising", "float32").metropolis_update(
            num_sweeps=20, num_reads=6, num_replicas=16, min_temp=0.5, max_temp=3.0,
            update_strategy="random", seed=99, block_size=bs, form="blocked")
        out.append(df)
    for col in ("energies", "best_energy"):
        assert np.array_equal(np.stack(out[0][col].to_numpy()), np.stack(out[1][col].to_numpy()))
