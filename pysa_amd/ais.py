"""Partition-function helpers around the engine's log-weights (host side, f64).

Mirrors the public helpers of pysa/ais.py:
  omegas_to_partition      pysa/ais.py:115-131
  combine_logZf            pysa/ais.py:59-76
  partition_function_post  pysa/ais.py:29-56 (the reference calls the 3-argument
                           get_log_omega with two arguments at :51; here the
                           call is completed, the rest is unchanged)
  get_log_omega            pysa/ais.py:79-112
  uniform_*/bernoulli_*    pysa/ais.py:137-226

The per-sweep log-omega of a metropolis_update run is computed on the GPU;
these functions only combine the results.
"""

import numpy as np

__all__ = [
    "omegas_to_partition", "combine_logZf", "partition_function_post", "get_log_omega",
    "uniform_prob_initialization", "uniform_ref_prob", "uniform_partition_fun",
    "bernoulli_prob_initialization", "bernoulli_ref_prob", "bernoulli_partition_fun",
]


def omegas_to_partition(log_omegas, logZ0):
    """log Z = max + log(sum(exp(w - max))) - log(len(w)) + log Z0, in f64
    (pysa/ais.py:115-131)."""
    w = np.asarray(log_omegas, dtype=np.float64)
    smax = w.max()
    expsum = np.sum(np.exp(w - smax))
    return smax - np.log(len(w)) + logZ0 + np.log(expsum)


def combine_logZf(solution):
    """Combine the 'log_Zf' column of a metropolis_update DataFrame over reads
    (pysa/ais.py:59-76)."""
    log_Zf = np.asarray(solution["log_Zf"].to_numpy(), dtype=np.float64)
    smax = log_Zf.max()
    return smax - np.log(len(log_Zf)) + np.log(np.sum(np.exp(log_Zf - smax)))


def get_log_omega(betas, beta_idx, energies):
    """log-omega of one PT ladder (pysa/ais.py:79-112): -sum_r (b_r - b_{r-1}) E_r
    over the ladder sorted by ascending beta, which must start at beta = 0."""
    betas = np.asarray(betas)
    energies = np.asarray(energies)
    order = np.argsort(betas, kind="quicksort")
    cur_b = betas[order]
    cur_e = energies[np.asarray(beta_idx)[order]]
    if cur_b[0] != 0:
        raise ValueError(
            '''Calculating the Partition Function requires a zero beta.  The most likely solution is that when you are calling metrpolis_update you need to include np.inf in your temperature list (if using user defined temperatures) or use the beta0 = True option (if using min_temp and max_temp).'''
        )
    lw, bold = 0.0, 0.0
    for r in range(1, len(cur_b)):
        lw += float(-cur_b[r] * cur_e[r]) - (-bold * float(cur_e[r]))
        bold = float(cur_b[r])
    return lw


def partition_function_post(solution):
    """log Z of the coldest ladder slot from a DataFrame whose temps include
    infinity (pysa/ais.py:29-56, with the get_log_omega call completed by
    passing the identity slot map, since the DataFrame's temps are already
    per state)."""
    n = len(solution["states"][0][0])
    log_omegas = []
    for temps, energies in zip(solution["temps"], solution["energies"]):
        with np.errstate(divide="ignore"):
            betas = 1 / np.asarray(temps)
        log_omegas.append(get_log_omega(betas, np.arange(len(betas)), energies))
    return omegas_to_partition(np.array(log_omegas), n * np.log(2))


def uniform_prob_initialization(n, problem_type, initial_args=None):
    """A uniformly random state from the global NumPy RNG (pysa/ais.py:137-149)."""
    if problem_type == "qubo":
        state = np.random.randint(2, size=n)
    elif problem_type == "ising":
        state = 2 * np.random.randint(2, size=n) - 1
    else:
        raise ValueError(f"problem_type=='{problem_type}' not supported.")
    return np.array(state)


def uniform_ref_prob(state, pos=None):
    """Unnormalised uniform probability (pysa/ais.py:152-160)."""
    return 1


def uniform_partition_fun(n):
    """log Z0 = n log 2 (pysa/ais.py:163-169)."""
    return n * np.log(2)


def bernoulli_prob_initialization(n, problem_type, initial_args=(0.5,)):
    """A state whose entries are -1/0 with probability p = initial_args[0]
    (pysa/ais.py:175-195), drawn element by element as the reference does."""
    p = initial_args[0]
    if problem_type == "qubo":
        state = [np.random.binomial(1, 1 - p) for _ in range(n)]
    elif problem_type == "ising":
        state = [2 * np.random.binomial(1, 1 - p) - 1 for _ in range(n)]
    else:
        raise ValueError(f"problem_type=='{problem_type}' not supported.")
    return np.array(state)


def bernoulli_ref_prob(state, p, pos=-1):
    """p^(n-c) (1-p)^c with c the number of entries equal to 1 after flipping
    `pos` (pysa/ais.py:198-217)."""
    state = np.asarray(state)
    bits = (state == 1).astype(np.int64)
    if 0 <= pos < len(state):
        bits[pos] = 1 - bits[pos]
    c = int(bits.sum())
    n = len(state)
    return (p**(n - c)) * ((1 - p)**c)


def bernoulli_partition_fun(n):
    """log Z0 = 0: the Bernoulli reference is normalised (pysa/ais.py:220-226)."""
    return 0
