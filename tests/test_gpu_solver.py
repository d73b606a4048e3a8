"""The drop-in Solver end to end on the GPU: the reference's DataFrame
(pysa/sa.py:358-430), its options, and agreement with the oracle when the
randomness is replayed."""

import numpy as np
import pandas as pd
import pytest

import oracle
from pysa_amd import ais as pais
from pysa_amd.sa import Solver

from helpers import int_problem

pytestmark = pytest.mark.gpu

COLUMNS = ["states", "energies", "best_state", "best_energy", "temps", "num_sweeps",
           "min_sweeps", "init_time (us)", "runtime (us)", "problem_type", "float_type"]


def problem(n, seed, ptype="ising"):
    rng = np.random.default_rng(seed)
    J, h = int_problem(n, rng)
    P = J.astype(np.float64)
    np.fill_diagonal(P, h)
    return P


def test_dataframe_layout():
    P = problem(40, 1)
    s = Solver(P, "ising", "float32")
    df = s.metropolis_update(num_sweeps=50, num_reads=3, num_replicas=6, update_strategy="random")
    assert isinstance(df, pd.DataFrame)
    assert list(df.columns) == COLUMNS
    assert len(df) == 3
    for r in range(3):
        assert df["states"][r].shape == (6, 40) and df["states"][r].dtype == np.float32
        assert set(np.unique(df["states"][r])) <= {-1.0, 1.0}
        assert df["energies"][r].shape == (6,)
        assert df["best_state"][r].shape == (40,)
        assert df["num_sweeps"][r] == 50
        assert 0 <= df["min_sweeps"][r] < 50
        assert df["problem_type"][r] == "ising" and df["float_type"][r] == "float32"
        e = [s.get_energy(x) for x in df["states"][r]]
        assert np.array_equal(np.asarray(e, np.float32), df["energies"][r])
        assert s.get_energy(df["best_state"][r]) == df["best_energy"][r]
    lst = s.metropolis_update(num_sweeps=5, num_reads=2, num_replicas=3, return_dataframe=False)
    assert isinstance(lst, list) and list(lst[0].keys()) == COLUMNS


@pytest.mark.parametrize("ptype", ["ising", "qubo"])
@pytest.mark.parametrize("strategy", ["sequential", "random"])
@pytest.mark.parametrize("parallel", [True, False])
def test_replay_through_solver_matches_oracle(ptype, strategy, parallel):
    """Initial states come from the global NumPy RNG exactly as in PySA
    (one draw of shape (R, n) per read); with the sweeps' and exchanges'
    uniforms replayed, every output equals the oracle's."""
    n, reads, R, S = 33, 3, 5, 20
    P = problem(n, 2)
    s = Solver(P, ptype, "float32")
    rng = np.random.default_rng(3)
    rep = dict(log_r=np.log(rng.random((reads, S, R, n))),
               pt_logu=np.log(rng.random((reads, S, R - 1))))
    if strategy == "random":
        rep["perm"] = np.argsort(rng.random((reads, S, R, n)), axis=-1).astype(np.int32)
    np.random.seed(11)
    df = s.metropolis_update(num_sweeps=S, num_reads=reads, num_replicas=R,
                             update_strategy=strategy, parallel=parallel, replay=rep,
                             get_part_fun=True, min_temp=0.5, max_temp=4.0)
    np.random.seed(11)
    init = []
    for _ in range(reads):
        b = np.random.randint(2, size=(R, n))
        init.append(b if ptype == "qubo" else 2 * b - 1)
    st0 = np.array(init, dtype=np.int8)
    J = s.couplings.astype(np.float32)
    h = s.local_fields.astype(np.float32)
    en0 = oracle.energy_f64(J, h, st0.reshape(-1, n)).reshape(reads, R).astype(np.float32)
    betas = s._ladder(R, None, 0.5, 4.0, True)
    want = oracle.metropolis(J, h, 0 if ptype == "ising" else 1, betas, st0, en0, S,
                             order=1 if strategy == "random" else 0, log_omega=True,
                             best_sem=0 if parallel else 1, replay=rep)
    for r in range(reads):
        bi = want["beta_idx"][r]
        assert np.array_equal(df["states"][r], want["states"][r].astype(np.float32))
        assert np.array_equal(df["energies"][r], want["energies"][r])
        assert np.array_equal(df["best_state"][r], want["best_state"][r].astype(np.float32))
        assert df["best_energy"][r] == want["best_energy"][r]
        assert df["min_sweeps"][r] == want["best_sweeps"][r]
        with np.errstate(divide="ignore"):
            assert np.array_equal(df["temps"][r], 1 / betas[bi])
        assert df["log_Zf"][r] == pais.omegas_to_partition(want["log_omegas"][r], n * np.log(2))


def test_sort_output_temps_and_recompute_energy():
    n = 30
    P = problem(n, 4)
    s = Solver(P, "ising", "float64")
    df = s.metropolis_update(num_sweeps=20, num_reads=2, num_replicas=4, sort_output_temps=True,
                             recompute_energy=True, seed=5)
    betas = s._ladder(4, None, 0.3, 1.5, False)
    for r in range(2):
        assert np.array_equal(df["temps"][r], 1 / betas)
        e = np.array([s.get_energy(x) for x in df["states"][r]])
        assert np.array_equal(df["energies"][r], e)


def test_send_background_and_async():
    P = problem(20, 5)
    s = Solver(P, "qubo")
    fut = s.metropolis_update(num_sweeps=10, num_reads=2, num_replicas=3, send_background=True)
    df = fut.result(timeout=60)
    assert list(df.columns) == COLUMNS and len(df) == 2
    df2 = s.metropolis_update_async(num_sweeps=10, num_reads=1, num_replicas=3)
    assert len(df2) == 1


def test_initialize_strategies():
    n, R = 24, 4
    P = problem(n, 6)
    s = Solver(P, "qubo", "float32")
    df = s.metropolis_update(num_sweeps=1, num_replicas=R, initialize_strategy="zeros",
                             temps=[1e-9] * R, use_pt=False, seed=1,
                             update_strategy="sequential")
    assert df["best_energy"][0] <= 0
    st = np.random.default_rng(0).integers(0, 2, (R, n)).astype(np.float32)
    df = s.metropolis_update(num_sweeps=3, num_replicas=R, initialize_strategy=st, seed=2)
    assert df["states"][0].shape == (R, n)
    with pytest.raises(ValueError, match="not recognized"):
        s.metropolis_update(num_sweeps=3, num_replicas=R, initialize_strategy="bogus")


def test_forms_agree_on_integer_couplings():
    """Sequential order, Philox randomness: the blocked MFMA form and the
    plain form produce identical results for integer couplings."""
    n, reads, R, S = 96, 4, 16, 60
    P = problem(n, 7)
    s = Solver(P, "ising", "float32")
    out = {}
    for form in ("plain", "blocked"):
        np.random.seed(8)
        stats = {}
        out[form] = s.metropolis_update(num_sweeps=S, num_reads=reads, num_replicas=R,
                                        update_strategy="sequential", seed=9, form=form,
                                        engine_stats=stats)
        assert stats["form"] == form
    for col in ("states", "energies", "best_state", "best_energy", "temps", "min_sweeps"):
        for r in range(reads):
            assert np.array_equal(out["plain"][col][r], out["blocked"][col][r]), col
    stats = {}
    s.metropolis_update(num_sweeps=2, num_reads=1, num_replicas=R, update_strategy="sequential",
                        engine_stats=stats)
    assert stats["form"] == "blocked"  # auto picks the MFMA form where it applies
