"""Synthetic problem instances of the benchmark configurations.

Each generator draws from ``np.random.default_rng(1000 + k)`` for config k and
returns the problem in PySA's convention: a symmetric matrix P whose diagonal
holds the local fields and whose off-diagonal holds the couplings, with
energy E(s) = s.(J s / 2 + h) (pysa/ising.py:37).
"""

import numpy as np


def pm1_ising(n=64, seed=1001):
    """C1: dense +-1 Ising, zero diagonal, h = 0."""
    rng = np.random.default_rng(seed)
    A = np.triu(2 * rng.integers(0, 2, (n, n)) - 1, 1)
    return (A + A.T).astype(np.float32)


def gaussian_qubo(n=512, seed=1002):
    """C2: QUBO with upper-triangular Q, off-diagonal and diagonal N(0,1).
    Returned as P = Q + Q^T - diag(Q) so that PySA's energy equals x^T Q x."""
    rng = np.random.default_rng(seed)
    Q = np.triu(rng.normal(size=(n, n)), 1) + np.diag(rng.normal(size=n))
    P = Q + Q.T - np.diag(np.diag(Q))
    return P.astype(np.float32)


def planted_wishart(n=1024, alpha=0.5, seed=1003):
    """C3: planted Wishart Ising, M = alpha*n Gaussian columns projected
    orthogonal to the all-ones vector; P = +W W^T / n with a zero diagonal.

    E(s) = (|W^T s|^2 - tr(W W^T)) / (2n) >= -|W|_F^2 / (2n), with equality
    at s = all-ones (W^T 1 = 0), so all-ones is a ground state.  (With the
    Hamze sign convention H = -1/2 s^T J s this is J = -W W^T / n.)
    Returns (P float32, E_gs) with E_gs the f64 energy of all-ones under the
    float32 couplings."""
    rng = np.random.default_rng(seed)
    m = int(round(alpha * n))
    W = rng.normal(size=(n, m))
    W -= W.mean(axis=0)
    P = (W @ W.T / n).astype(np.float32)
    np.fill_diagonal(P, 0)
    Pd = P.astype(np.float64)
    e_gs = 0.5 * Pd.sum()
    return P, e_gs


def sk_batch(count=2048, nmin=32, nmax=768, seed=1004):
    """C4: Sherrington-Kirkpatrick instances with N uniform in [nmin, nmax],
    J_ij = +-1/sqrt(N).  Returns a list of (signs int8, scale float32)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        A = np.triu(2 * rng.integers(0, 2, (n, n)) - 1, 1)
        out.append(((A + A.T).astype(np.int8), np.float32(1.0 / np.sqrt(n))))
    return out


def rbm_ising(nv=784, nh=500, wstd=0.05, bstd=0.1, seed=1005):
    """C5: bipartite (RBM-shaped) Ising: P = [[diag b, W], [W^T, diag c]]."""
    rng = np.random.default_rng(seed)
    W = rng.normal(0.0, wstd, size=(nv, nh))
    b = rng.normal(0.0, bstd, size=nv)
    c = rng.normal(0.0, bstd, size=nh)
    n = nv + nh
    P = np.zeros((n, n))
    P[:nv, nv:] = W
    P[nv:, :nv] = W.T
    P[np.arange(nv), np.arange(nv)] = b
    P[nv + np.arange(nh), nv + np.arange(nh)] = c
    return P.astype(np.float32)


def rbm_log_z(P, nv, beta=1.0):
    """Exact log Z of the bipartite Ising model P (spins +-1) at `beta` by
    summing out the hidden layer: E(v,h) = v.W h + b.v + c.h, so
    Z = sum_v exp(-beta b.v) prod_j 2 cosh(beta (W^T v + c)_j)."""
    P = np.asarray(P, dtype=np.float64)
    W = P[:nv, nv:]
    b = np.diag(P)[:nv]
    c = np.diag(P)[nv:]
    v = ((np.arange(2**nv)[:, None] >> np.arange(nv)) & 1) * 2.0 - 1.0
    x = beta * (v @ W + c)
    terms = -beta * (v @ b) + np.sum(np.logaddexp(x, -x), axis=1)
    mx = terms.max()
    return mx + np.log(np.sum(np.exp(terms - mx)))
