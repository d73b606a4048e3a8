"""AIS on the GPU (pysa/simulation.py:266-317, pysa/sa.py:432-553): the
kernel against the oracle, and log Z against exact enumeration."""

import math

import numpy as np
import pytest
from scipy.special import logsumexp

import oracle
from pysa_amd import instances
from pysa_amd.sa import Solver

from helpers import engine_ais, int_problem, random_spins

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("ref", [0, 1])
@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("order", [0, 1])
def test_ais_replay_matches_oracle(ptype, ref, ordering, order):
    rng = np.random.default_rng(50 + 8 * ptype + 4 * ref + 2 * ordering + order)
    n, chains, nt = 37, 9, 25
    J, h = int_problem(n, rng)
    J, h = J.astype(np.float64), h.astype(np.float64)
    temps = np.array([(t + 1) / nt for t in range(nt)])
    st = random_spins((chains, n), rng, ptype)
    rep = dict(log_r=np.log(rng.random((chains, nt, n))))
    if order:
        rep["perm"] = np.argsort(rng.random((chains, nt, n)), axis=-1).astype(np.int32)
    want = oracle.ais(J, h, ptype, temps, st, ref=ref, p=0.3, order=order, ordering=ordering,
                      replay=rep)
    got = engine_ais(J, h, ptype, temps, st, ref=ref, p=0.3, order=order, ordering=ordering,
                     replay=rep)
    assert np.array_equal(got[0], want[0])
    assert np.allclose(got[1], want[1], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_ais_philox_matches_oracle(dtype, order):
    rng = np.random.default_rng(7)
    n, chains, nt = 64, 40, 30
    J, h = int_problem(n, rng)
    J, h = J.astype(dtype), h.astype(dtype)
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=dtype)
    st = random_spins((chains, n), rng, 0)
    want = oracle.ais(J, h, 0, temps, st, order=order, seed=123, nthreads=8)
    got = engine_ais(J, h, 0, temps, st, order=order, seed=123)
    assert np.array_equal(got[0], want[0])
    assert np.allclose(got[1], want[1], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("ordering", ["reference", "neal"])
def test_ais_log_z_exact_rbm_14_10(ordering):
    """The exact-enumeration case: 14 visible + 10 hidden spins with the
    benchmark's weight distributions; exact log Z sums the 2^14 visible states
    with the hidden units summed out in closed form (that formula is checked
    against full enumeration in test_host).  AIS with a uniform reference must
    agree within its bootstrap error."""
    nv, nh = 14, 10
    P = instances.rbm_ising(nv, nh, seed=1005)
    logz = instances.rbm_log_z(P, nv)
    np.random.seed(3)
    s = Solver(P.astype(np.float64), "ising", "float64")
    chains = 4096
    logzf, samples, omegas = s.annealed_importance(1000, chains, "uniform", "random", seed=77,
                                                   ordering=ordering)
    assert samples.shape == (chains, nv + nh)
    assert np.isclose(np.sum(omegas), 1.0)
    lw = np.log(np.asarray(omegas))
    rng = np.random.default_rng(0)
    boots = [logsumexp(lw[rng.integers(0, chains, chains)]) - math.log(chains)
             for _ in range(200)]
    se = float(np.std(boots))
    assert abs(logzf - logz) < 5 * se + 5e-3, (logzf, logz, se)


def test_ais_log_z_bernoulli_reference():
    """Bernoulli reference (pysa/ais.py:175-226, log Z0 = 0) on a small RBM at
    beta = 0.8 (couplings scaled as pysa/sa.py:511-512)."""
    nv, nh = 8, 6
    P = instances.rbm_ising(nv, nh, wstd=0.3, bstd=0.2, seed=9).astype(np.float64)
    logz = instances.rbm_log_z(P, nv, beta=0.8)
    np.random.seed(4)
    s = Solver(P, "ising", "float64")
    logzf, _, omegas = s.annealed_importance(400, 4096, "bernoulli", "sequential",
                                             reference_args=[0.4], beta=0.8, seed=5,
                                             ordering="neal")
    lw = np.log(np.asarray(omegas))
    rng = np.random.default_rng(1)
    se = float(np.std([logsumexp(lw[rng.integers(0, 4096, 4096)]) for _ in range(200)]))
    assert abs(logzf - logz) < 5 * se + 5e-3, (logzf, logz, se)


@pytest.mark.parametrize("ordering", ["reference", "neal"])
def test_ais_log_z_exact_rbm_14_10_float32(ordering):
    """The benchmark's AIS configuration at the exact-enumeration size:
    float32, 14 visible + 10 hidden spins, 4096 chains, 10,000 linear betas,
    uniform reference.  log Z within the bootstrap error of exact
    enumeration; the tracked final energies (recomputed in f64 every 64
    steps, updated from the f32 fields in between) within 1e-5 of the f64
    energies of the returned states."""
    nv, nh = 14, 10
    P = instances.rbm_ising(nv, nh, seed=1005)
    logz = instances.rbm_log_z(P.astype(np.float64), nv)
    np.random.seed(5)
    s = Solver(P.astype(np.float32), "ising", "float32")
    chains, nt = 4096, 10000
    logzf, samples, omegas = s.annealed_importance(nt, chains, "uniform", "random", seed=78,
                                                   ordering=ordering)
    lw = np.log(np.asarray(omegas))
    rng = np.random.default_rng(2)
    boots = [logsumexp(lw[rng.integers(0, chains, chains)]) - math.log(chains)
             for _ in range(200)]
    se = float(np.std(boots))
    assert abs(logzf - logz) < 5 * se + 5e-3, (logzf, logz, se)
    # tracked energies of the same problem through the C ABI
    J = P.astype(np.float32).copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float32).copy()
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    st0 = random_spins((chains, nv + nh), np.random.default_rng(9), 0)
    st, _, en = engine_ais(J, h, 0, temps, st0, order=1, seed=79,
                           ordering=0 if ordering == "reference" else 1, energies=True)
    e64 = oracle.energy_f64(J, h, st)
    assert np.max(np.abs(en - e64) / np.maximum(1.0, np.abs(e64))) < 1e-5


def test_ais_float32_matches_oracle_weights():
    """Float couplings: the kernel's weights (fields updated incrementally,
    f64 refresh every 256 steps) against the oracle's (energy recomputed every
    step, as pysa/simulation.py:309) on the same Philox stream."""
    rng = np.random.default_rng(12)
    n, chains, nt = 40, 64, 300
    P = instances.rbm_ising(24, 16, wstd=0.3, bstd=0.2, seed=12)
    J = P.astype(np.float32).copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float32).copy()
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    st = random_spins((chains, n), rng, 0)
    want = oracle.ais(J, h, 0, temps, st, order=1, seed=31, nthreads=8)
    got = engine_ais(J, h, 0, temps, st, order=1, seed=31)
    # decisions from f32 fields may differ in rare near-ties; most chains agree
    same = np.all(got[0] == want[0], axis=1)
    assert same.mean() > 0.9
    assert np.allclose(got[1][same], want[1][same], rtol=1e-6, atol=1e-4)


# ---- MFMA form (uniform reference, sequential or block-random order)

@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("ordering", [0, 1])
@pytest.mark.parametrize("order", [0, 3])
def test_ais_blocked_matches_oracle(ptype, ordering, order):
    """Integer couplings (exact in every encoding): the MFMA form's states,
    weights and tracked energies equal the oracle's, including chains that
    leave a tile partly empty and a size that is not a multiple of 32."""
    from pysa_amd import _lib
    rng = np.random.default_rng(900 + 4 * ptype + 2 * ordering + (order == 3))
    n, chains, nt = 150, 70, 20
    J, h = int_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    st = random_spins((chains, n), rng, ptype)
    want = oracle.ais(J, h, ptype, temps, st, order=order, ordering=ordering, seed=321, nthreads=8)
    got = engine_ais(J, h, ptype, temps, st, order=order, ordering=ordering, seed=321,
                     energies=True, form=_lib.FORM_BLOCKED)
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])
    Jd = J.astype(np.float64)
    s = got[0].astype(np.float64)
    e64 = 0.5 * np.sum((s @ Jd) * s, axis=1) + s @ h.astype(np.float64)
    assert np.array_equal(got[2], e64)


def test_ais_blocked_equals_plain_block_random():
    """Both forms in the block-random order give the same chains (integer
    couplings, int8 encoding, 2 x 32-spin blocks per tile of a large size)."""
    from pysa_amd import _lib
    rng = np.random.default_rng(910)
    n, chains, nt = 300, 128, 12
    J, h = int_problem(n, rng, -40, 40, 5)
    J, h = J.astype(np.float32), h.astype(np.float32)
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    st = random_spins((chains, n), rng, 0)
    a = engine_ais(J, h, 0, temps, st, order=3, seed=5, form=_lib.FORM_PLAIN)
    b = engine_ais(J, h, 0, temps, st, order=3, seed=5, form=_lib.FORM_BLOCKED)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])


def _ring(n, coupling):
    """Periodic Ising chain: J_{i,i+1} = coupling; log Z in closed form
    (transfer matrix: Z = (2 cosh J)^n + (2 sinh J)^n at beta = 1, with the
    reference's E = s.J.s/2 summing each bond once)."""
    J = np.zeros((n, n), np.float32)
    for i in range(n):
        J[i, (i + 1) % n] = J[(i + 1) % n, i] = coupling
    # E = sum_bonds J s_i s_j, Boltzmann weight exp(-E): Z = (2 cosh J)^n + (2 sinh(-J))^n
    c = float(np.float32(coupling))
    logz = n * math.log(2 * math.cosh(c)) + math.log1p((math.tanh(-c)) ** n)
    return J, np.zeros(n, np.float32), logz


@pytest.mark.parametrize("form", ["blocked", "blocked_exact"])
def test_ais_blocked_log_z_ring(form):
    """Float couplings (0.37 is not a sum of two bf16 values): log Z of a
    100-spin periodic chain against the closed form, through the Solver."""
    J, h, logz = _ring(100, 0.37)
    P = J + np.diag(h)
    s = Solver(P, "ising", "float32")
    stats = {}
    logzf, _, _ = s.annealed_importance(400, 4096, "uniform", "random", seed=91, form=form,
                                        engine_stats=stats)
    assert abs(logzf - logz) < 0.05, (logzf, logz)


def test_ais_blocked_log_z_ring_qubo():
    """The same ring as a QUBO (x = (s+1)/2: couplings 4J, fields -4J, energy
    shifted by nJ), which the MFMA form runs with +-1 operands and row sums."""
    n, c = 100, 0.37
    J, _, logz_ising = _ring(n, c)
    Jq = 4 * J
    hq = np.full(n, -4 * np.float32(c), np.float32)
    logz = logz_ising + n * float(np.float32(c))
    s = Solver(Jq + np.diag(hq), "qubo", "float32")
    logzf, _, _ = s.annealed_importance(400, 4096, "uniform", "random", seed=92, form="blocked")
    assert abs(logzf - logz) < 0.05, (logzf, logz)


@pytest.mark.parametrize("chains,temps", [(3, [0.0, 0.25, 1.0]), (65, [0.5]),
                                          (1, [0.0, 0.0, 0.7, 1.3])])
def test_ais_blocked_edge_cases(chains, temps):
    """A tile with few chains, a schedule starting at beta = 0 (every
    proposal accepted), a single step, repeated betas: the MFMA form equals
    the oracle bit for bit (integer couplings, block-random order)."""
    from pysa_amd import _lib
    rng = np.random.default_rng(920 + chains)
    n = 97
    J, h = int_problem(n, rng)
    J, h = J.astype(np.float32), h.astype(np.float32)
    temps = np.array(temps, dtype=np.float32)
    st = random_spins((chains, n), rng, 0)
    for ordering in (0, 1):
        want = oracle.ais(J, h, 0, temps, st, order=3, ordering=ordering, seed=17)
        got = engine_ais(J, h, 0, temps, st, order=3, ordering=ordering, seed=17,
                         form=_lib.FORM_BLOCKED)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
