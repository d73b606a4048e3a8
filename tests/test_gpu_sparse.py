"""Sparse form: the plain kernel with an accepted flip updating only the
nonzeros of its coupling row (ELL copy of a symmetric J).  A zero coupling
adds exactly nothing to a field, and the periodic f64 refresh sums the
nonzeros in the dense order, so the sparse form's outputs equal the dense
plain form's bit for bit for any couplings (and the oracle's for integer
ones)."""

import numpy as np
import pytest

import oracle
from pysa_amd import _lib
from pysa_amd.sa import Solver

from helpers import assert_same, energies_of, engine_metropolis, random_spins

pytestmark = pytest.mark.gpu


def lattice(L, rng, kind):
    """Periodic L x L lattice, 4 neighbours: +-1 or Gaussian couplings, and
    integer or Gaussian fields."""
    n = L * L
    J = np.zeros((n, n))
    for i in range(L):
        for j in range(L):
            a = i * L + j
            for b in (((i + 1) % L) * L + j, i * L + (j + 1) % L):
                v = (2 * rng.integers(0, 2) - 1) if kind == "pm1" else rng.normal()
                J[a, b] = J[b, a] = v
    h = rng.integers(-1, 2, n).astype(float) if kind == "pm1" else 0.3 * rng.normal(size=n)
    return J, h


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("L,kind", [(8, "pm1"), (16, "gauss"), (64, "gauss")])
def test_sparse_equals_plain(dtype, ptype, order, L, kind):
    if dtype == np.float64 and L * L > 2048:
        pytest.skip("float64 above n = 2048 is refused by the plain kernel (a memory fault "
                    "of that instantiation, cause not found)")
    rng = np.random.default_rng(L + 3 * order + 7 * ptype + (dtype == np.float64))
    J, h = lattice(L, rng, kind)
    J, h = J.astype(dtype), h.astype(dtype)
    n = L * L
    R, reads = 8, 3
    S = 150 if L <= 16 else 8  # past the float32 refresh at sweep 128 on the small lattices
    betas = (1 / np.geomspace(0.5, 3.0, R)).astype(dtype)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, dtype)
    kw = dict(order=order, seed=21, log_omega=False, best_sem=order % 2)
    a = engine_metropolis(J, h, ptype, betas, st, en, S, form=_lib.FORM_PLAIN, **kw)
    b = engine_metropolis(J, h, ptype, betas, st, en, S, form=_lib.FORM_SPARSE, **kw)
    assert_same(a, b)


@pytest.mark.parametrize("ptype", [0, 1])
def test_sparse_equals_oracle_integer(ptype):
    rng = np.random.default_rng(5 + ptype)
    J, h = lattice(12, rng, "pm1")
    J, h = J.astype(np.float32), h.astype(np.float32)
    n, R, reads, S = 144, 8, 4, 40
    betas = (1 / np.geomspace(0.5, 3.0, R)).astype(np.float32)
    st = random_spins((reads, R, n), rng, ptype)
    en = energies_of(J, h, st, ptype, np.float32)
    ref = oracle.metropolis(J, h, ptype, betas, st, en, S, order=1, seed=8, nthreads=8)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=1, seed=8,
                            form=_lib.FORM_SPARSE)
    assert_same(got, ref)


def test_float64_above_2048_refused():
    rng = np.random.default_rng(2)
    J, h = lattice(48, rng, "gauss")  # n = 2304
    st = random_spins((1, 2, 48 * 48), rng, 0)
    en = energies_of(J, h, st, 0, np.float64)
    for form in (_lib.FORM_PLAIN, _lib.FORM_SPARSE):
        with pytest.raises(RuntimeError, match="float64 above n = 2048"):
            engine_metropolis(J, h, 0, np.ones(2), st, en, 2, form=form)


def test_sparse_refused_for_dense_problem():
    rng = np.random.default_rng(1)
    n = 64
    A = rng.normal(size=(n, n))
    J = ((A + A.T) / 2).astype(np.float32)
    np.fill_diagonal(J, 0)
    h = np.zeros(n, np.float32)
    st = random_spins((1, 4, n), rng, 0)
    en = energies_of(J, h, st, 0, np.float32)
    with pytest.raises(RuntimeError, match="sparse form needs"):
        engine_metropolis(J, h, 0, np.ones(4, np.float32), st, en, 2, form=_lib.FORM_SPARSE)


def test_solver_sparse_form_equals_plain():
    rng = np.random.default_rng(4)
    J, h = lattice(10, rng, "gauss")
    P = J + np.diag(h)
    out = {}
    for form in ("sparse", "plain"):
        stats = {}
        np.random.seed(3)
        df = Solver(P, "ising", "float64").metropolis_update(
            num_sweeps=30, num_reads=4, num_replicas=8, update_strategy="random", seed=5,
            form=form, engine_stats=stats)
        out[form] = (df, stats["form"])
    assert out["sparse"][1] == "sparse" and out["plain"][1] == "plain"
    for col in ("energies", "best_energy"):
        assert np.array_equal(np.stack(out["sparse"][0][col].to_numpy()),
                              np.stack(out["plain"][0][col].to_numpy()))
