"""Host-side checks that need no GPU: the C ABI exports, the drop-in Solver's
argument validation and temperature ladder, the partition-function helpers
and the synthetic instances."""

import math
import os
import re
import subprocess

import numpy as np
import pandas as pd
import pytest
from scipy.special import logsumexp

from pysa_amd import _lib, ais, instances
from pysa_amd.sa import Solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pysa_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pysa_amd_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pysa_amd", "csrc")], check=True)
    return _lib.load()


def test_header_matches_binding():
    assert header_functions() == _lib.EXPORTED


def test_library_exports_every_declared_symbol(lib):
    for name in header_functions():
        assert hasattr(lib, name), name
    assert lib.pysa_amd_version().decode().startswith("pysa_amd 0.")
    assert lib.pysa_amd_last_error() == b""


def test_library_has_no_cpu_compute_path(lib):
    """No symbol of the library exists to run the sweeps on the host."""
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert {n for n in names if n.startswith("pysa_amd_")} == set(header_functions())
    assert not any("oracle" in n for n in names)


def sym(n, rng):
    a = rng.normal(size=(n, n))
    return (a + a.T) / 2


def test_solver_constructor_validation():
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError, match="problem_type='spin' not supported."):
        Solver(sym(4, rng), "spin")
    with pytest.raises(ValueError, match="Problem must be a square matrix."):
        Solver(np.zeros((3, 4)), "ising")
    with pytest.raises(ValueError, match="Problem must be a symmetric matrix."):
        Solver(np.triu(np.ones((4, 4))), "ising")
    P = sym(5, rng)
    s = Solver(P, "qubo", "float64")
    assert np.array_equal(s.local_fields, np.diag(P))
    assert np.array_equal(np.diag(s.couplings), np.zeros(5))
    off = ~np.eye(5, dtype=bool)
    assert np.array_equal(s.couplings[off], P[off])


def test_metropolis_update_argument_errors():
    s = Solver(sym(6, np.random.default_rng(1)), "ising")
    with pytest.raises(ValueError, match=re.escape("len('temps') != 'num_replicas'.")):
        s.metropolis_update(10, num_replicas=3, temps=[1, 2])
    with pytest.raises(ValueError, match=re.escape("len('initialize_strategy') != 'num_replicas'.")):
        s.metropolis_update(10, num_replicas=3, initialize_strategy=np.ones((2, 6)))
    with pytest.raises(ValueError, match=re.escape("len('init_energies') != 'num_replicas'.")):
        s.metropolis_update(10, initialize_strategy=np.ones((2, 6)), init_energies=[1.0])
    with pytest.raises(ValueError, match="require the temps array to include np.inf"):
        s.metropolis_update(10, temps=[1.0, 2.0], get_part_fun=True)


def test_ladder_matches_reference_formula():
    """pysa/sa.py:167-201: geometric temperatures, slot 0 coldest; beta0 puts
    the ratio over R-2 and appends 0; R = 1 and R = 2 are special."""
    s = Solver(sym(4, np.random.default_rng(2)), "ising")
    b = s._ladder(None, None, 0.3, 1.5, False)
    assert b.dtype == np.float32 and len(b) == 4
    r = (1.5 / 0.3) ** (1 / 3)
    assert np.allclose(b, [1 / (0.3 * r**k) for k in range(4)], rtol=1e-6)
    b = s._ladder(5, None, 1.0, 3.5, True)
    r = 3.5 ** (1 / 3)
    assert np.allclose(b, [1, 1 / r, 1 / r**2, 1 / 3.5, 0], rtol=1e-6)
    assert s._ladder(1, None, 0.5, 2.0, False).tolist() == [2.0]
    assert np.allclose(s._ladder(2, None, 0.5, 2.0, False), [2.0, 0.5])
    b = s._ladder(None, [1.0, 2.0, np.inf], 0, 0, True)
    assert b.tolist() == [1.0, 0.5, 0.0]


def test_omegas_to_partition_closed_form():
    w = np.array([1.0, 2.0, 3.0])
    got = ais.omegas_to_partition(w, 5 * math.log(2))
    assert np.isclose(got, logsumexp(w) - math.log(3) + 5 * math.log(2), rtol=1e-15)
    assert np.isclose(ais.omegas_to_partition(np.full(7, 1e3), 0.0), 1e3, rtol=1e-15)


def test_combine_logzf():
    df = pd.DataFrame({"log_Zf": [10.0, 10.0, 10.0 + math.log(4)]})
    assert np.isclose(ais.combine_logZf(df), 10.0 + math.log(2), rtol=1e-15)


def test_get_log_omega_hand_ladder():
    """Ladder (1, 0.5, 0) held by replicas (2, 0, 1) with energies (-1, 3, -4):
    sorted by beta: 0 -> replica 1 (E 3), 0.5 -> replica 0 (E -1),
    1 -> replica 2 (E -4): log w = -(0.5 - 0)(-1) - (1 - 0.5)(-4) = 2.5."""
    assert ais.get_log_omega([1.0, 0.5, 0.0], [2, 0, 1], [-1.0, 3.0, -4.0]) == 2.5
    with pytest.raises(ValueError, match="requires a zero beta"):
        ais.get_log_omega([1.0, 0.5], [0, 1], [0.0, 0.0])


def test_reference_probabilities():
    assert ais.uniform_ref_prob(np.ones(5)) == 1
    assert ais.uniform_partition_fun(7) == 7 * np.log(2)
    s = np.array([1, -1, 1, 1])
    assert ais.bernoulli_ref_prob(s, 0.25) == pytest.approx(0.25 * 0.75**3)
    assert ais.bernoulli_ref_prob(s, 0.25, pos=0) == pytest.approx(0.25**2 * 0.75**2)
    assert ais.bernoulli_partition_fun(9) == 0


def test_instances():
    P = instances.pm1_ising(64, 1001)
    assert P.dtype == np.float32 and np.array_equal(P, P.T)
    assert set(np.unique(P[~np.eye(64, dtype=bool)])) == {-1.0, 1.0} and not P.diagonal().any()
    P = instances.gaussian_qubo(40, 1002)
    rng = np.random.default_rng(1002)
    Q = np.triu(rng.normal(size=(40, 40)), 1) + np.diag(rng.normal(size=40))
    x = np.random.default_rng(0).integers(0, 2, (20, 40)).astype(np.float64)
    J = P.astype(np.float64).copy()
    np.fill_diagonal(J, 0)
    e = 0.5 * np.einsum("ki,ij,kj->k", x, J, x) + x @ np.diag(P).astype(np.float64)
    assert np.allclose(e, np.einsum("ki,ij,kj->k", x, Q, x), rtol=1e-5, atol=1e-4)
    P, egs = instances.planted_wishart(64, 0.5, 1003)
    assert not P.diagonal().any() and np.array_equal(P, P.T)
    assert np.isclose(egs, 0.5 * P.astype(np.float64).sum(), rtol=1e-12)


def test_rbm_log_z_matches_enumeration():
    import itertools
    nv = 6
    P = instances.rbm_ising(nv, 4, wstd=0.5, bstd=0.3, seed=3)
    n = P.shape[0]
    J = P.astype(np.float64).copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float64)
    s = 2 * np.array(list(itertools.product([0, 1], repeat=n))) - 1
    E = 0.5 * np.einsum("ki,ij,kj->k", s, J, s) + s @ h
    for beta in (1.0, 0.37):
        assert np.isclose(instances.rbm_log_z(P, nv, beta), logsumexp(-beta * E), rtol=1e-12)


def test_batched_and_ais_form_argument_errors():
    """Argument checks that run before any device call."""
    from pysa_amd import batched
    rng = np.random.default_rng(3)
    A = np.triu(2 * rng.integers(0, 2, (40, 40)) - 1, 1)
    inst = [((A + A.T).astype(np.int8), np.float32(0.2))]
    with pytest.raises(ValueError, match="form='bogus' not recognized"):
        batched.solve_batch(inst, num_sweeps=2, form="bogus")
    with pytest.raises(ValueError, match="update_strategy='shuffled' not recognized"):
        batched.solve_batch(inst, num_sweeps=2, update_strategy="shuffled")
    s = Solver(sym(6, rng), "ising")
    with pytest.raises(ValueError, match="form='bogus' not recognized"):
        s.annealed_importance(5, 4, "uniform", form="bogus")
