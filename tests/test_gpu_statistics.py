"""Distributional checks with real (float) couplings, where trajectories of
the two forms and the oracle part ways after the first near-tie decision.

* Per-slot Boltzmann statistics: a 12-spin core with Gaussian couplings and
  fields, spread over the four 32-spin blocks of an n = 100 problem whose
  other spins are free (no couplings, no field).  After PT sweeps the
  replica in slot k must sample exp(-beta_k E) of the core: chi-square of
  its energy over equiprobable bins of the exact distribution (2^12 states
  enumerated), for both forms and every visit order.
* Two-sample Kolmogorov-Smirnov between the plain and the MFMA form on the
  QUBO configuration (N = 512, Gaussian Q) with the same visit order: final
  energies of the coldest and a middle slot over 256 reads.
"""

import itertools

import numpy as np
import pytest
from scipy import stats

from pysa_amd import _lib, instances

from helpers import engine_metropolis, random_spins

pytestmark = pytest.mark.gpu

CORE = np.array([3, 17, 30, 33, 45, 58, 64, 70, 81, 90, 96, 99])


def _core_problem(ptype, seed):
    rng = np.random.default_rng(seed)
    n, k = 100, len(CORE)
    c = rng.normal(0, 1.0, (k, k))
    c = np.triu(c, 1)
    c = c + c.T
    hc = rng.normal(0, 0.5, k)
    J = np.zeros((n, n), np.float32)
    J[np.ix_(CORE, CORE)] = c
    h = np.zeros(n, np.float32)
    h[CORE] = hc
    return J, h, c.astype(np.float32).astype(np.float64), hc.astype(np.float32).astype(np.float64)


def _core_energies(c, hc, ptype):
    k = len(hc)
    b = np.array(list(itertools.product([0, 1], repeat=k)), np.float64)
    s = b if ptype == 1 else 2 * b - 1
    return 0.5 * np.einsum("ki,ij,kj->k", s, c, s) + s @ hc


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("form,order", [(_lib.FORM_PLAIN, 0), (_lib.FORM_PLAIN, 1),
                                        (_lib.FORM_PLAIN, 3), (_lib.FORM_BLOCKED, 0),
                                        (_lib.FORM_BLOCKED, 3)])
def test_float_couplings_boltzmann_per_slot(ptype, form, order):
    J, h, c, hc = _core_problem(ptype, 400 + ptype)
    n, R, reads, S = 100, 8, 1024, 150
    betas = np.array([1.6, 1.1, 0.8, 0.6, 0.45, 0.33, 0.22, 0.12], np.float32)
    rng = np.random.default_rng(401 + 7 * order + ptype)
    st = random_spins((reads, R, n), rng, ptype)
    s64 = st.astype(np.float64)
    en = (0.5 * np.einsum("...i,ij,...j->...", s64, J.astype(np.float64), s64)
          + s64 @ h.astype(np.float64)).astype(np.float32)
    got = engine_metropolis(J, h, ptype, betas, st, en, S, order=order, seed=402 + order,
                            form=form)
    allE = _core_energies(c, hc, ptype)
    weights_bits = 1 << np.arange(len(CORE))[::-1]
    for k in range(R):
        rep = got["beta_idx"][:, k]
        core = got["states"][np.arange(reads), rep][:, CORE]
        idx = ((core > 0).astype(np.int64) * weights_bits).sum(axis=1)
        samp = allE[idx]
        w = np.exp(-float(betas[k]) * (allE - allE.min()))
        p = w / w.sum()
        order_e = np.argsort(allE)
        cdf = np.cumsum(p[order_e])
        nb = 12
        edges_idx = np.searchsorted(cdf, np.arange(1, nb) / nb)
        edges = allE[order_e][np.minimum(edges_idx, len(allE) - 1)]
        bins = np.concatenate([[-np.inf], np.unique(edges), [np.inf]])
        exp_p = np.array([p[(allE > lo) & (allE <= hi)].sum() for lo, hi in zip(bins[:-1], bins[1:])])
        obs = np.array([((samp > lo) & (samp <= hi)).sum() for lo, hi in zip(bins[:-1], bins[1:])])
        keep = exp_p * reads >= 5
        e_ = exp_p[keep] * reads
        o = obs[keep]
        chi2 = ((o - e_ * o.sum() / e_.sum()) ** 2 / (e_ * o.sum() / e_.sum())).sum()
        pv = stats.chi2.sf(chi2, keep.sum() - 1)
        assert pv > 1e-4, (k, chi2, pv)


def test_qubo_forms_same_distribution_ks():
    P = instances.gaussian_qubo(512, 1002)
    J = P.astype(np.float32).copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float32).copy()
    reads, R, S = 256, 32, 300
    r = np.exp(1 / (R - 1) * np.log(5.0 / 0.1))
    betas = np.array([1 / (0.1 * r**k) for k in range(R)], dtype=np.float32)
    rng = np.random.default_rng(1002)
    st = rng.integers(0, 2, (reads, R, 512)).astype(np.int8)
    s64 = st.astype(np.float64)
    en = (0.5 * np.einsum("...i,ij,...j->...", s64, J.astype(np.float64), s64)
          + s64 @ h.astype(np.float64)).astype(np.float32)
    out = {}
    Jd, hd = J.astype(np.float64), h.astype(np.float64)
    for form, seed in ((_lib.FORM_PLAIN, 5), (_lib.FORM_BLOCKED, 6)):
        got = engine_metropolis(J, h, 1, betas, st, en, S, order=3, seed=seed, form=form)
        out[form] = {}
        for k in (0, 16):
            # f64 energies of the states (the cold slot's distribution is
            # nearly discrete; tracked values of one state differ between the
            # forms by their ~1e-6 rounding)
            x = got["states"][np.arange(reads), got["beta_idx"][:, k]].astype(np.float64)
            out[form][k] = 0.5 * np.einsum("ki,ij,kj->k", x, Jd, x) + x @ hd
    for k in (0, 16):
        a, b = out[_lib.FORM_PLAIN][k], out[_lib.FORM_BLOCKED][k]
        # asymptotic p-value (the exact one is not computable with the ties
        # of a nearly discrete distribution)
        pv = stats.ks_2samp(a, b, method="asymp").pvalue
        assert pv > 1e-3, (k, pv, a.mean(), b.mean())
