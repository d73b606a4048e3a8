"""Instance I/O (pysa/utils.py:61-89, tutorials/hpo_demo.ipynb cell 5) and
the ladder heuristics / TTS objective / parameter search of the HPO tutorial
(cells 8-20): known answers worked by hand and the reference's own data
fixture."""

import math
import os
import warnings

import numpy as np
import pytest

from pysa_amd import hpo, io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_get_problem_matrix_known_answer():
    inst = [(10, 3, 1.5), (3, 7, -2.0), (7, 7, 0.25)]
    P = io.get_problem_matrix(inst)
    # labels sorted: 3 -> 0, 7 -> 1, 10 -> 2
    want = np.array([[0.0, -2.0, 1.5], [-2.0, 0.25, 0.0], [1.5, 0.0, 0.0]])
    assert np.array_equal(P, want)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        P2 = io.get_problem_matrix(inst + [(3, 10, 4.0)])
    assert any("Duplicated" in str(x.message) for x in w)
    assert P2[0, 2] == P2[2, 0] == 4.0  # the last listed value


def test_coupling_list_loader_on_tutorial_fixture():
    path = os.path.join(GOLDEN, "wishart_planting_N_50_alpha_0.50_inst_1.txt")
    P = io.load_coupling_list(path, n=50)
    raw = np.loadtxt(path)
    W = np.zeros((50, 50))
    for i, j, v in raw:
        W[int(i), int(j)] += v
    assert np.array_equal(P, (W + W.T) / 2)
    assert np.allclose(P, P.T)
    gs = io.load_gs_energy(os.path.join(GOLDEN, "wishart_N_50_gs_energies.txt"))
    assert gs == -1534000.0  # the fixture's listed value (its sum over i < j convention)


def test_to_csr_round_trip():
    rng = np.random.default_rng(3)
    A = np.triu(rng.integers(-2, 3, (9, 9)) * (rng.random((9, 9)) < 0.4), 1)
    P = (A + A.T).astype(np.float32)
    P[np.arange(9), np.arange(9)] = rng.normal(size=9)
    indptr, idx, data, diag = io.to_csr(P)
    D = np.zeros_like(P)
    for i in range(9):
        D[i, idx[indptr[i]:indptr[i + 1]]] = data[indptr[i]:indptr[i + 1]]
        assert np.all(np.diff(idx[indptr[i]:indptr[i + 1]]) > 0)
    np.fill_diagonal(D, diag)
    assert np.array_equal(D, P)


def test_delta_e_and_temperature_range_known_answer():
    J = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, -3.0], [0.0, -3.0, 0.0]])
    h = np.array([0.0, 0.5, 0.0])
    hot, cold, cnt = hpo.get_delta_e_ising(J, h)
    assert hot == 2 * (1 + 3 + 0.5)  # variable 1
    assert cold == 2 * 0.5           # |h_1|
    assert cnt == 1                  # only variable 1 has 0.5 as its smallest coefficient
    P = J + np.diag(h)
    tmin, tmax = hpo.temperature_range(P, "ising", phot=50, pcold=1)
    assert math.isclose(tmax, -9.0 / math.log(0.5))
    assert math.isclose(tmin, -1.0 / math.log(0.01))
    # QUBO bounds go through the Ising form J = (Q - diag Q) / 4, h = 1^T Q / 2
    Q = np.array([[2.0, -4.0], [-4.0, 6.0]])
    hot_q, cold_q, cnt_q = hpo.get_delta_e_qubo(Q)
    Jq = np.array([[0.0, -1.0], [-1.0, 0.0]])
    hq = np.array([-1.0, 1.0])
    assert (hot_q, cold_q, cnt_q) == hpo.get_delta_e_ising(Jq, hq)


def test_tts_objective_closed_form():
    e = np.array([-10.0, -9.0, -10.0, -5.0])
    # gap 0: success iff e <= -10 -> p = 0.5
    assert math.isclose(hpo.tts_objective(e, 2.0, -10.0, opt_gap=0.0),
                        2.0 * math.log(0.01) / math.log(0.5))
    assert hpo.tts_objective(e, 2.0, -20.0, opt_gap=0.0) == 1e10
    assert hpo.tts_objective(e, 2.0, -4.0, opt_gap=0.0) == 2.0


def test_random_search_finds_minimum():
    def obj(p):
        return (math.log(p["num_sweeps"]) - 2.0) ** 2 + (p["num_replicas"] - 9) ** 2

    best, loss, trials = hpo.search(obj, max_evals=200, seed=1, use_hyperopt=False)
    assert len(trials) == 200
    assert loss == min(t[1] for t in trials)
    assert best["num_replicas"] == 9 and abs(math.log(best["num_sweeps"]) - 2.0) < 0.5
