"""Temperature-ladder heuristics and a TTS-driven parameter search
(tutorials/hpo_demo.ipynb, cells 8-22).

* ``get_delta_e_ising`` / ``get_delta_e_qubo`` -- the largest and smallest
  single-flip energy changes of a problem (mean-field bounds), cells 8-10.
* ``temperature_range`` -- the ladder ends that give a hot-end acceptance of
  `phot` % for the largest change and a cold-end acceptance of `pcold` %
  for the smallest, cell 12.
* ``tts_objective`` -- TTS(s) of a batch of best energies at an optimality
  gap, cell 14 (after arXiv:1905.10876).
* ``run_solver`` -- one Solver run scored by that objective, cell 16.
* ``search`` -- the hyperparameter search of cell 20: hyperopt's TPE when
  hyperopt is importable, else a seeded random search over the same space
  (log-uniform sweeps, uniform replicas), returning the best parameters and
  every trial.
"""

import math

import numpy as np

from .sa import Solver


def get_delta_e_ising(J, h, scaling_correction=True):
    """(delta_e_hot, delta_e_cold, count_min_i) of an Ising problem with
    symmetric couplings J (zero diagonal) and fields h.

    delta_e_hot = 2 max_i (sum_j |J_ij| + |h_i|); delta_e_cold = 2 x the
    smallest nonzero |J_ij| or |h_i|; count_min_i counts the variables whose
    smallest nonzero coefficient is that minimum (1 without the correction)."""
    J = np.asarray(J, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    delta_e_hot = 2 * np.max(np.sum(np.abs(J), axis=1) + np.abs(h))
    min_field = np.minimum(np.min(np.abs(J[J != 0]), initial=np.inf),
                           np.min(np.abs(h[h != 0]), initial=np.inf))
    min_ji = np.where(np.abs(J) != 0, np.abs(J), np.inf).min(axis=1)
    min_hi = np.where(np.abs(h) != 0, np.abs(h), min_ji)
    count_min_i = int(np.sum(np.minimum(min_ji, min_hi) == min_field)) if scaling_correction else 1
    return delta_e_hot, 2 * min_field, count_min_i


def get_delta_e_qubo(Q):
    """The same bounds for a symmetric QUBO matrix Q, through its Ising form
    J = (Q - diag Q) / 4, h = 1^T Q / 2."""
    Q = np.asarray(Q, dtype=np.float64)
    J = (Q - np.diag(np.diag(Q))) / 4.0
    h = np.ones(Q.shape[0]) @ Q / 2.0
    return get_delta_e_ising(J, h)


def temperature_range(problem, problem_type="ising", phot=50.0, pcold=1.0):
    """(min_temp, max_temp): exp(-dE_hot / T_max) = phot / 100 and
    exp(-dE_cold / T_min) = pcold / 100 / count_min_i."""
    P = np.asarray(problem, dtype=np.float64)
    if problem_type == "ising":
        J = P.copy()
        np.fill_diagonal(J, 0)
        d_hot, d_cold, cnt = get_delta_e_ising(J, np.diag(P))
    elif problem_type == "qubo":
        d_hot, d_cold, cnt = get_delta_e_qubo(P)
    else:
        raise ValueError(f"problem_type='{problem_type}' not supported.")
    max_temp = -d_hot / np.log(phot / 100.0)
    min_temp = -d_cold / np.log(pcold / 100.0 / cnt)
    return float(min_temp), float(max_temp)


def tts_objective(energies, mean_runtime, gs_energy, s=0.99, opt_gap=0.05, fail_value=1e10):
    """TTS(s): the runtime times log(1 - s) / log(1 - p), p the fraction of
    energies at or below gs_energy (1 - opt_gap); mean_runtime when p >= s,
    fail_value when p = 0."""
    p = float(np.mean(np.asarray(energies) <= gs_energy * (1.0 - opt_gap)))
    if p == 0.0:
        return fail_value
    if p >= s:
        return float(mean_runtime)
    return float(mean_runtime * np.log(1 - s) / np.log(1 - p))


def run_solver(problem, fixed_params, tuned_params, gs_energy, problem_type="ising",
               float_type="float32", energy_scale=1.0, opt_gap=0.05, s=0.99, **engine):
    """One metropolis_update run with fixed + tuned parameters, scored by
    tts_objective on energy_scale x best_energy (the tutorial doubles the
    energies of its (W + W^T) / 2 instance to match its ground-state file).
    The first read is dropped as in the tutorial (it includes compilation
    there).  Returns a dict with 'loss' (the TTS) and the run's parameters."""
    params = dict(fixed_params)
    params.update(tuned_params)
    params["num_sweeps"] = int(params["num_sweeps"])
    params["num_replicas"] = int(params["num_replicas"])
    solver = Solver(problem=np.asarray(problem), problem_type=problem_type,
                    float_type=float_type)
    result = solver.metropolis_update(**params, **engine)
    energies = energy_scale * np.asarray(result["best_energy"][1:], dtype=np.float64)
    mean_runtime = 1e-6 * float(np.mean(result["runtime (us)"][1:]))
    tts = tts_objective(energies, mean_runtime, gs_energy, s=s, opt_gap=opt_gap)
    return {"loss": tts, "status": "ok", "num_sweeps": params["num_sweeps"],
            "num_replicas": params["num_replicas"], "result": result}


def search(objective, max_evals=50, sweeps_log_range=(0.0, 4.0), replicas_range=(1, 16),
           seed=0, use_hyperopt=None):
    """Minimise objective({'num_sweeps', 'num_replicas'}) over the tutorial's
    space: num_sweeps = round(exp(U(a, b))), num_replicas = round(U(lo, hi)).

    Uses hyperopt's TPE when importable (or when use_hyperopt is True), else a
    seeded random search.  Returns (best_params, best_loss, trials), trials a
    list of (params, loss) in evaluation order."""
    if use_hyperopt is None:
        try:
            import hyperopt  # noqa: F401
            use_hyperopt = True
        except ImportError:
            use_hyperopt = False
    trials = []
    if use_hyperopt:
        from hyperopt import STATUS_OK, Trials, fmin, hp, tpe

        space = {"num_sweeps": hp.qloguniform("num_sweeps", sweeps_log_range[0],
                                              sweeps_log_range[1], 1),
                 "num_replicas": hp.quniform("num_replicas", replicas_range[0],
                                             replicas_range[1], 1)}

        def fn(params):
            p = {k: int(v) for k, v in params.items()}
            loss = float(objective(p))
            trials.append((p, loss))
            return {"loss": loss, "status": STATUS_OK}

        fmin(fn=fn, space=space, algo=tpe.suggest, max_evals=max_evals, trials=Trials(),
             rstate=np.random.default_rng(seed), show_progressbar=False)
    else:
        rng = np.random.default_rng(seed)
        for _ in range(max_evals):
            p = {"num_sweeps": max(1, int(round(math.exp(rng.uniform(*sweeps_log_range))))),
                 "num_replicas": max(1, int(round(rng.uniform(*replicas_range))))}
            trials.append((p, float(objective(p))))
    best = min(range(len(trials)), key=lambda i: trials[i][1])
    return trials[best][0], trials[best][1], trials
