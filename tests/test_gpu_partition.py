"""The DataFrame post-processors of pysa/ais.py on real engine output:
combine_logZf (pysa/ais.py:59-76) over a get_part_fun=True run, and
partition_function_post (pysa/ais.py:29-56) over a beta0=True run with
sort_output_temps=True (the call of examples/example_ais.py:116-126; with
unsorted output the temps column is the reference's per-replica quirk and
the pairing with energies does not hold).

PySA's PT estimate weighs each rung by the state sampled at the higher beta,
so its expectation is not log Z but the closed form
  n log 2 + sum_r [log Z(2 b_r - b_(r-1)) - log Z(b_r)]
(the stationary law of PT is the product of the slots' Boltzmann laws),
enumerated exactly on a 12-spin core.  The MFMA form needs n >= 32: the core
is embedded in n = 100 spins whose other spins are free (no couplings, no
field), which leaves every ratio of partition functions unchanged."""

import itertools

import numpy as np
import pytest
from scipy.special import logsumexp

from pysa_amd import ais
from pysa_amd.sa import Solver

pytestmark = pytest.mark.gpu

CORE = np.array([3, 17, 30, 33, 45, 58, 64, 70, 81, 90, 96, 99])


def _problem(n_total, seed=5):
    rng = np.random.default_rng(seed)
    k = len(CORE)
    c = np.triu(rng.normal(0, 0.5, (k, k)), 1)
    Jc = (c + c.T).astype(np.float32)
    hc = rng.normal(0, 0.3, k).astype(np.float32)
    if n_total == k:
        P = Jc.astype(np.float64) + np.diag(hc.astype(np.float64))
        return P, Jc, hc
    P = np.zeros((n_total, n_total))
    P[np.ix_(CORE, CORE)] = Jc
    P[CORE, CORE] = hc
    return P, Jc, hc


def _predicted(Jc, hc, betas, n_total):
    s = np.array(list(itertools.product([-1, 1], repeat=len(hc))), np.float64)
    E = 0.5 * np.einsum("ki,ij,kj->k", s, Jc.astype(np.float64), s) + s @ hc.astype(np.float64)
    b = np.sort(np.asarray(betas, np.float64))
    lz = lambda x: logsumexp(-x * E)  # noqa: E731
    return n_total * np.log(2) + sum(lz(2 * b[r] - b[r - 1]) - lz(b[r]) for r in range(1, len(b)))


@pytest.mark.parametrize("n_total,form,strategy", [(12, "plain", "sequential"),
                                                   (12, "plain", "random"),
                                                   (100, "blocked", "sequential"),
                                                   (100, "blocked", "random")])
def test_combine_logzf_on_engine_dataframe(n_total, form, strategy):
    P, Jc, hc = _problem(n_total)
    np.random.seed(11)
    solver = Solver(P.astype(np.float32), "ising", "float32")
    R, reads, S = 24, 512, 2000
    df = solver.metropolis_update(S, num_reads=reads, num_replicas=R, min_temp=0.5,
                                  max_temp=20.0, update_strategy=strategy,
                                  get_part_fun=True, sort_output_temps=True, form=form,
                                  seed=123)
    assert "log_Zf" in df.columns
    est = ais.combine_logZf(df)
    betas = 1 / np.asarray(df["temps"][0], np.float64)
    pred = _predicted(Jc, hc, betas, n_total)
    lz = df["log_Zf"].to_numpy(np.float64)
    rng = np.random.default_rng(0)
    boots = [logsumexp(lz[rng.integers(0, reads, reads)]) - np.log(reads) for _ in range(200)]
    se = float(np.std(boots))
    assert abs(est - pred) < 5 * se + 0.01, (est, pred, se)


@pytest.mark.parametrize("n_total,form", [(12, "plain"), (100, "blocked")])
def test_partition_function_post_on_engine_dataframe(n_total, form):
    P, Jc, hc = _problem(n_total)
    np.random.seed(12)
    solver = Solver(P.astype(np.float32), "ising", "float32")
    R, reads, S = 24, 4096, 200
    df = solver.metropolis_update(S, num_reads=reads, num_replicas=R, min_temp=0.5,
                                  max_temp=20.0, update_strategy="sequential", beta0=True,
                                  sort_output_temps=True, form=form, seed=321)
    est = ais.partition_function_post(df)
    temps = np.asarray(df["temps"][0], np.float64)
    with np.errstate(divide="ignore"):
        betas = 1 / temps
    assert betas.min() == 0.0
    pred = _predicted(Jc, hc, betas, n_total)
    # the same estimate from the columns directly (one log-omega per read)
    lw = np.array([ais.get_log_omega(1 / np.asarray(t, np.float64), np.arange(R), e)
                   for t, e in zip(df["temps"], df["energies"])])
    assert np.isclose(est, ais.omegas_to_partition(lw, n_total * np.log(2)))
    rng = np.random.default_rng(1)
    boots = [logsumexp(lw[rng.integers(0, reads, reads)]) - np.log(reads) + n_total * np.log(2)
             for _ in range(200)]
    se = float(np.std(boots))
    assert abs(est - pred) < 5 * se + 0.01, (est, pred, se)
