"""Problem-instance input: PySA's interaction-list builder and the text
formats its tutorials read.

* ``get_problem_matrix`` -- pysa/utils.py:61-89: a list of (i, j, J)
  interactions to a dense symmetric matrix over the sorted variable labels
  (a (i, i, h) entry sets a local field).
* ``load_coupling_list`` -- the whitespace-separated ``i j J`` files of the
  tutorials (tutorials/hpo_demo.ipynb, cell 5, on
  tutorials/wishart_planting_N_50_alpha_0.50_inst_1.txt): entry (i, j) of an
  n x n matrix W, returned in the symmetric form (W + W^T) / 2 the tutorial
  passes to ``Solver``.
* ``load_gs_energy`` -- the ``name<TAB>energy`` lines of
  tutorials/gs_energies.txt.
* ``to_csr`` -- CSR arrays of a problem matrix (the engine's sparse form
  builds its own ELL copy from the dense matrix it is given).
"""

import warnings
from collections import Counter
from typing import Iterable, List, Tuple

import numpy as np


def get_problem_matrix(instance: Iterable[Tuple[int, int, float]]) -> np.ndarray:
    """Dense problem matrix from interactions (pysa/utils.py:61-89).

    Variables are the sorted set of labels that appear; entry (x, y) and
    (y, x) take J.  A pair listed more than once warns, and the last value
    listed wins (the reference's loop overwrites in order)."""
    instance = list(instance)
    if instance and set(Counter(tuple(sorted(x[:2])) for x in instance).values()) != {1}:
        warnings.warn("Duplicated couplings are ignored.")
    labels = sorted({v for x in instance for v in x[:2]})
    index = {v: i for i, v in enumerate(labels)}
    problem = np.zeros((len(labels), len(labels)))
    for x, y, J in instance:
        a, b = index[x], index[y]
        problem[a][b] = problem[b][a] = J
    return problem


def load_coupling_list(path, n=None, symmetrize=True) -> np.ndarray:
    """Dense matrix from an ``i j J`` text file (0-based indices).

    Repeated (i, j) entries add up, as a sparse COO-to-CSR conversion does
    (tutorials/hpo_demo.ipynb, cell 5); `n` defaults to the largest index
    + 1.  With `symmetrize` the result is (W + W^T) / 2."""
    data = np.loadtxt(path, ndmin=2)
    if data.size == 0:
        return np.zeros((n or 0, n or 0))
    rows = data[:, 0].astype(np.int64)
    cols = data[:, 1].astype(np.int64)
    vals = data[:, 2].astype(np.float64)
    if n is None:
        n = int(max(rows.max(), cols.max())) + 1
    if rows.min() < 0 or cols.min() < 0 or max(rows.max(), cols.max()) >= n:
        raise ValueError(f"indices out of range for n={n}")
    W = np.zeros((n, n))
    np.add.at(W, (rows, cols), vals)
    return (W + W.T) / 2.0 if symmetrize else W


def load_gs_energy(path, name=None) -> float:
    """Ground-state energy from ``name<TAB>energy`` lines: the first line, or
    the line whose first field is `name`."""
    with open(path) as f:
        for line in f:
            parts = line.strip().split()
            if len(parts) < 2:
                continue
            if name is None or parts[0] == name:
                return float(parts[1])
    raise KeyError(name)


def to_csr(problem: np.ndarray):
    """(indptr int32, indices int32, data, diag) of the off-diagonal part of
    a dense problem, in the problem's float type; rows list columns in
    increasing order."""
    P = np.asarray(problem)
    off = P.copy()
    np.fill_diagonal(off, 0)
    nz = off != 0
    counts = nz.sum(axis=1)
    indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    rows, cols = np.nonzero(nz)
    return indptr, cols.astype(np.int32), off[rows, cols], np.diag(P).copy()


def csr_from_lists(rows: List[int], cols: List[int], vals: List[float], n: int):
    """Symmetric CSR (both (i, j) and (j, i)) from an upper-triangle list."""
    W = np.zeros((n, n))
    for i, j, v in zip(rows, cols, vals):
        W[i, j] = W[j, i] = v
    return to_csr(W)
