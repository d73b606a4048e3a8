"""Drop-in ``Solver`` for PySA's Metropolis / parallel-tempering path.

Same constructor, methods, keyword arguments, validation messages and
DataFrame as pysa/sa.py (Solver :35-553).  The sweeps, replica exchange,
best tracking and log-omega of all reads run in one call into the HIP engine
(libpysa_amd.so); initial states still come from the global NumPy RNG, so
``np.random.seed`` gives the same initial states as PySA.

Engine options are keyword-only extras, all defaulted so that PySA calls are
unchanged:
  form        'auto' (blocked MFMA form where eligible, else plain),
              'plain', 'sparse' (the plain form updating only the nonzeros
              of a row: symmetric J, at most min(n/4, 64) nonzeros per row),
              'blocked' or 'blocked_exact'
  block_size  MFMA-form block size B: 32 (the MFMA tile) with the library's
              launch shape, or 0 (default) for the launch shape the
              autotuner picks (pysa_amd/tune.py, cached per batch shape)
  seed        Philox key (default: drawn from os.urandom, which leaves the
              global NumPy RNG untouched)
  sweep_offset  Philox sweep-counter offset (resume a stream)
  replay      dict(log_r=(reads,S,R,N), perm=(reads,S,R,N) or None,
              pt_logu=(reads,S,R-1) or None): host-supplied randomness
  engine_stats  dict to receive accepted flips, device time, form, updates/s
"""

import math
import secrets
from concurrent.futures import ThreadPoolExecutor
from time import time
from typing import Any, List

import numpy as np
import pandas as pd

from . import _lib
from . import ais as _ais
from . import tune as _tune

Vector = List[float]
Matrix = List[List[float]]
State = List[float]

_BLOCKED_MIN_N = 32
_BLOCKED_MAX_N = 2048
_BLOCKED_RANDOM_MIN_N = 64


def _tqdm(it, disable):
    if disable:
        return it
    try:
        from tqdm.auto import tqdm
        return tqdm(it)
    except ImportError:
        return it


class Solver(object):
    """Dense Ising/QUBO solver (pysa/sa.py:35-70)."""

    def __init__(self, problem: Matrix, problem_type: str, float_type: Any = 'float32'):
        if problem_type == 'qubo':
            self._module = 'qubo'
        elif problem_type == 'ising':
            self._module = 'ising'
        else:
            raise ValueError(f"problem_type='{problem_type}' not supported.")
        self.problem_type = problem_type
        self.n_vars = len(problem)
        self.float_type = float_type
        if problem.shape != (self.n_vars, self.n_vars):
            raise ValueError("Problem must be a square matrix.")
        if not np.allclose(problem, problem.T):
            raise ValueError("Problem must be a symmetric matrix.")
        self.local_fields = np.copy(np.diag(problem))
        self.couplings = np.copy(problem)
        np.fill_diagonal(self.couplings, 0)
        self._device = {}  # float dtype -> _lib.Problem

    # -- device problems ----------------------------------------------------
    def _problem(self, dtype):
        dt = np.dtype(dtype)
        p = self._device.get(dt)
        if p is None:
            p = _lib.Problem(self.couplings.astype(dt), self.local_fields.astype(dt),
                             self.problem_type, dt)
            self._device[dt] = p
        return p

    def _spins(self, states):
        """Engine spin encoding (int8 +1/-1 or 1/0) of float states."""
        s = np.asarray(states)
        if self.problem_type == 'ising':
            return np.where(s > 0, 1, -1).astype(np.int8)
        return np.where(s != 0, 1, 0).astype(np.int8)

    def get_energy(self, conf: Vector, dtype: Any = None):
        """Energy of one configuration, E = s.(J s / 2 + h) (pysa/sa.py:72-83).

        Accumulated in f64 on the GPU over the couplings cast to `dtype`, then
        rounded to `dtype`."""
        if dtype is None:
            dtype = self.couplings.dtype
        dt = np.dtype(dtype)
        e = self._problem(dt).energy(self._spins(np.asarray(conf).astype(dt))[None, :])[0]
        return dt.type(e)

    def _energies(self, states, dtype):
        dt = np.dtype(dtype)
        return self._problem(dt).energy(self._spins(states)).astype(dt)

    def metropolis_update_async(self, **kwargs):
        kwargs['send_background'] = False
        return self.metropolis_update(**kwargs)

    def metropolis_update(self,
                          num_sweeps: int,
                          num_reads: int = 1,
                          num_replicas: int = None,
                          temps: np.ndarray = None,
                          min_temp: float = 0.3,
                          max_temp: float = 1.5,
                          update_strategy: str = 'random',
                          initialize_strategy: str = 'random',
                          init_energies: List[float] = None,
                          recompute_energy: bool = False,
                          sort_output_temps: bool = False,
                          return_dataframe: bool = True,
                          parallel: bool = True,
                          use_pt: bool = True,
                          send_background: bool = False,
                          verbose: bool = False,
                          get_part_fun: bool = False,
                          beta0: bool = False,
                          *,
                          form: str = 'auto',
                          block_size: int = 0,
                          seed: int = None,
                          sweep_offset: int = 0,
                          replay: dict = None,
                          engine_stats: dict = None) -> pd.DataFrame:
        """Metropolis sweeps with parallel tempering over all reads in one
        engine call (pysa/sa.py:92-430).  Output columns and their order follow
        pysa/sa.py:395-422."""

        if send_background:  # pysa/sa.py:140-142
            return ThreadPoolExecutor(max_workers=1).submit(Solver.metropolis_update_async,
                                                             **locals())

        # Number of replicas (pysa/sa.py:144-161)
        if temps is not None:
            if num_replicas and len(temps) != num_replicas:
                raise ValueError(f"len('temps') != 'num_replicas'.")
            else:
                num_replicas = len(temps)
        if type(initialize_strategy) != str:
            if num_replicas and len(initialize_strategy) != num_replicas:
                raise ValueError(f"len('initialize_strategy') != 'num_replicas'.")
            else:
                num_replicas = len(initialize_strategy)
            if init_energies is not None and len(init_energies) != num_replicas:
                raise ValueError(f"len('init_energies') != 'num_replicas'.")

        if get_part_fun:  # pysa/sa.py:164-165
            beta0 = True
        betas = self._ladder(num_replicas, temps, min_temp, max_temp, beta0)
        num_replicas = len(betas)
        ft = np.dtype(self.float_type)
        couplings = self.couplings.astype(self.float_type)
        local_fields = self.local_fields.astype(self.float_type)

        init = self._init_strategy(initialize_strategy, init_energies, num_replicas, couplings,
                                   local_fields)
        if update_strategy not in ('random', 'sequential', 'block_random'):
            raise ValueError(f"update_strategy='{update_strategy}' not recognized.")
        if int(num_sweeps) < 1:
            raise ValueError("num_sweeps must be >= 1.")
        num_sweeps = int(num_sweeps)
        num_reads = int(num_reads)
        if num_reads < 1:
            return pd.DataFrame([]) if return_dataframe else []

        beta_sort = np.argsort(betas, kind='quicksort').astype(np.int32)
        if get_part_fun and betas[beta_sort[0]] != 0:  # pysa/ais.py:94-97
            raise ValueError(
                '''Calculating the Partition Function requires a zero beta.  The most likely solution is that when you are calling metrpolis_update you need to include np.inf in your temperature list (if using user defined temperatures) or use the beta0 = True option (if using min_temp and max_temp).'''
            )

        # Initial states of every read, drawn in read order from the global
        # RNG exactly as the per-read loop of pysa/sa.py:425-427 would.
        R, N = num_replicas, self.n_vars
        t0 = time()
        states = np.empty((num_reads, R, N), dtype=np.int8)
        energies = np.empty((num_reads, R), dtype=ft)
        for r in range(num_reads):
            s, e = init()
            states[r] = self._spins(s)
            energies[r] = e
        init_time = (time() - t0) / num_reads

        form_code, order = self._resolve_form(form, update_strategy, ft, R, replay)
        rep_lr = rep_perm = rep_pt = None
        if replay is not None:
            rep_lr = np.ascontiguousarray(replay['log_r'], dtype=np.float64)
            if rep_lr.shape != (num_reads, num_sweeps, R, N):
                raise ValueError("replay['log_r'] must have shape (reads, sweeps, replicas, n).")
            if replay.get('perm') is not None:
                rep_perm = np.ascontiguousarray(replay['perm'], dtype=np.int32)
                if rep_perm.shape != (num_reads, num_sweeps, R, N):
                    raise ValueError("replay['perm'] must have shape (reads, sweeps, replicas, n).")
                if not np.array_equal(np.sort(rep_perm, axis=-1),
                                      np.broadcast_to(np.arange(N, dtype=np.int32), rep_perm.shape)):
                    raise ValueError("replay['perm'] rows must be permutations of range(n).")
            elif update_strategy != 'sequential':
                raise ValueError(f"replay of update_strategy='{update_strategy}' needs replay['perm'].")
            if use_pt and R > 1:
                if replay.get('pt_logu') is None:
                    raise ValueError("replay with use_pt needs replay['pt_logu'].")
                rep_pt = np.ascontiguousarray(replay['pt_logu'], dtype=np.float64)
                if rep_pt.shape != (num_reads, num_sweeps, R - 1):
                    raise ValueError("replay['pt_logu'] must have shape (reads, sweeps, replicas-1).")
        if seed is None:
            seed = secrets.randbits(64)

        prob = self._problem(ft)
        if block_size not in (0, 32):
            raise ValueError("block_size must be 0 (tuned launch shape) or 32 (the MFMA tile).")
        if form_code in (_lib.FORM_BLOCKED, _lib.FORM_BLOCKED_EXACT):
            if block_size == 0:
                _tune.apply(prob, R, num_reads, order)
            else:
                prob.set_shape(0, 0)
        beta_idx = np.empty((num_reads, R), dtype=np.int64)
        best_state = np.empty((num_reads, N), dtype=np.int8)
        best_energy = np.empty(num_reads, dtype=ft)
        best_sweeps = np.empty(num_reads, dtype=np.int32)
        log_omegas = np.zeros((num_reads, num_sweeps), dtype=np.float64) if get_part_fun else None
        accepted = np.zeros(num_reads, dtype=np.uint64)
        kms = _lib._c_dbl(0.0)
        betas_c = np.ascontiguousarray(betas, dtype=ft)
        tw0 = time()
        lib = _lib.load()
        _lib.check(
            lib.pysa_amd_metropolis(prob.handle, num_reads, R, num_sweeps, order,
                                    int(bool(use_pt)), int(bool(get_part_fun)),
                                    0 if parallel else 1, _lib.ptr(betas_c), _lib.ptr(beta_sort),
                                    int(seed) & (2**64 - 1), int(sweep_offset), form_code,
                                    int(block_size), _lib.ptr(rep_lr), _lib.ptr(rep_perm),
                                    _lib.ptr(rep_pt), _lib.ptr(states), _lib.ptr(energies),
                                    _lib.ptr(beta_idx), _lib.ptr(best_state),
                                    _lib.ptr(best_energy), _lib.ptr(best_sweeps),
                                    _lib.ptr(log_omegas), _lib.ptr(accepted),
                                    _lib.ctypes.byref(kms)))
        wall = time() - tw0
        runtime = wall / num_reads
        if engine_stats is not None:
            upd = float(N) * num_sweeps * R * num_reads
            engine_stats.update({
                'form': {_lib.FORM_BLOCKED: 'blocked', _lib.FORM_BLOCKED_EXACT: 'blocked_exact',
                         _lib.FORM_SPARSE: 'sparse'}.get(form_code, 'plain'),
                'order': order, 'seed': int(seed), 'kernel_ms': kms.value, 'wall_s': wall,
                'accepted': accepted.copy(), 'updates': upd,
                'updates_per_s': upd / (kms.value * 1e-3) if kms.value > 0 else float('nan'),
            })

        def _read_result(r):
            out_states = states[r].astype(ft)
            out_energies = energies[r].copy()
            bstate = best_state[r].astype(ft)
            benergy = best_energy[r]
            out_beta_idx = beta_idx[r]
            if recompute_energy:  # pysa/sa.py:372-377, original problem dtype
                benergy = self.get_energy(bstate)
                out_energies = np.array([self.get_energy(s) for s in out_states])
            if sort_output_temps:  # pysa/sa.py:380-387
                out_betas = betas
                out_states = out_states[out_beta_idx]
                out_energies = out_energies[out_beta_idx]
            else:
                out_betas = betas[out_beta_idx]
            with np.errstate(divide='ignore'):
                out_temps = np.divide(1, out_betas)
            res = {
                'states': out_states,
                'energies': out_energies,
                'best_state': bstate,
                'best_energy': benergy,
                'temps': out_temps,
            }
            if get_part_fun:  # pysa/sa.py:392-394
                res['log_Zf'] = _ais.omegas_to_partition(log_omegas[r], self.n_vars * np.log(2))
            res.update({
                'num_sweeps': num_sweeps,
                'min_sweeps': best_sweeps[r],
                'init_time (us)': int(init_time * 1e6),
                'runtime (us)': int(runtime * 1e6),
                'problem_type': self.problem_type,
                'float_type': self.float_type,
            })
            return res

        _res = [_read_result(r) for r in _tqdm(range(num_reads), disable=not verbose)]
        return pd.DataFrame(_res) if return_dataframe else _res

    # -- helpers --------------------------------------------------------------
    def _ladder(self, num_replicas, temps, min_temp, max_temp, beta0):
        """Inverse temperatures in slot order (pysa/sa.py:167-201)."""
        if temps is None:
            if not num_replicas:
                num_replicas = 4
            if num_replicas == 1:
                betas = np.array([1 / min_temp], dtype=self.float_type)
            elif num_replicas == 2:
                betas = np.array([1 / min_temp, 1 / max_temp], dtype=self.float_type)
            elif num_replicas > 2:
                if beta0:
                    _ratio = np.exp(1 / (num_replicas - 2) * np.log(max_temp / min_temp))
                    betas = np.array([1 / (min_temp * _ratio**k) for k in range(num_replicas - 1)],
                                     dtype=self.float_type)
                    betas = np.append(betas, 0.0)
                else:
                    _ratio = np.exp(1 / (num_replicas - 1) * np.log(max_temp / min_temp))
                    betas = np.array([1 / (min_temp * _ratio**k) for k in range(num_replicas)],
                                     dtype=self.float_type)
            else:
                raise ValueError("num_replicas must be >= 1.")
        else:
            if beta0:
                if np.inf not in temps:
                    raise ValueError("get_part_fun = True or beta0 = True \
                    require the temps array to include np.inf")
            betas = 1 / np.array(temps, dtype=self.float_type)
        return np.asarray(betas).astype(self.float_type)

    def _init_strategy(self, initialize_strategy, init_energies, num_replicas, couplings,
                       local_fields):
        """Per-read initial states and energies (pysa/sa.py:213-313)."""
        ft = self.float_type
        n = self.n_vars
        if type(initialize_strategy) == str:
            if initialize_strategy == 'random':
                if self.problem_type == 'qubo':
                    def _init():
                        states = np.random.randint(2, size=(num_replicas, n)).astype(ft)
                        return states, self._energies(states, ft)
                elif self.problem_type == 'ising':
                    def _init():
                        states = 2 * np.random.randint(2, size=(num_replicas, n)).astype(ft) - 1
                        return states, self._energies(states, ft)
                else:
                    raise ValueError(f"self.problem_type=='{self.problem_type}' not supported.")
            elif initialize_strategy == 'zeros' and self.problem_type == 'qubo':
                def _init():
                    return (np.zeros((num_replicas, n), dtype=ft),
                            np.zeros(num_replicas, dtype=ft))
            elif initialize_strategy == 'ones':
                def _init():
                    states = np.ones((num_replicas, n), dtype=ft)
                    energies = np.ones(num_replicas, dtype=ft) * (np.sum(couplings) / 2 +
                                                                  np.sum(local_fields))
                    return states, energies
            else:
                raise ValueError(f"initialize_strategy='{initialize_strategy}' not recognized.")
        else:
            try:
                if init_energies is None:
                    init_energies = self._energies(np.asarray(initialize_strategy), ft)

                def _init():
                    states = np.copy(initialize_strategy).astype(ft)
                    energies = np.copy(init_energies).astype(ft)
                    return states, energies
            except Exception:
                raise ValueError("Cannot initialize system.")
        return _init

    def _sparse(self):
        """Whether the couplings qualify for the sparse form (symmetric, at
        most min(n/4, 64) nonzeros in every row)."""
        if getattr(self, '_is_sparse', None) is None:
            J = self.couplings
            n = self.n_vars
            d = int(np.count_nonzero(J, axis=1).max()) if n > 0 else 0
            self._is_sparse = bool(n > 0 and np.array_equal(J, J.T) and 0 < d <= 64 and
                                   4 * d <= n)
        return self._is_sparse

    def _resolve_form(self, form, update_strategy, ft, R, replay):
        """Engine form and visit order.

        'plain' runs every request: 'sequential' in position order, 'random'
        with a fresh permutation per replica and sweep (pysa/simulation.py:53),
        'block_random' in the block-random order.  The MFMA form
        ('blocked' / 'blocked_exact') needs float32, Philox randomness, at
        most 64 replicas and n <= 2048; it runs 'sequential' in position
        order and 'random' / 'block_random' in the block-random order (a
        random order of the 32-spin blocks per sweep with a random XOR
        shuffle inside each block, shared by all replicas; n > 64).  'auto'
        takes the MFMA form when it applies."""
        if form not in ('auto', 'plain', 'sparse', 'blocked', 'blocked_exact'):
            raise ValueError(f"form='{form}' not recognized.")
        blocked_order = 'sequential' if update_strategy == 'sequential' else 'block_random'
        ok = (ft == np.float32 and _BLOCKED_MIN_N <= self.n_vars <= _BLOCKED_MAX_N and
              R <= 64 and replay is None and _lib.blocked_available() and
              (blocked_order == 'sequential' or self.n_vars > _BLOCKED_RANDOM_MIN_N))
        if form == 'auto':
            form = 'blocked' if ok else 'plain'
        if form == 'sparse':
            if not self._sparse():
                raise ValueError("form='sparse' needs a symmetric problem with at most "
                                 "min(n/4, 64) nonzero couplings per row.")
            return _lib.FORM_SPARSE, _lib.ORDER[update_strategy]
        if form in ('blocked', 'blocked_exact'):
            if not ok:
                raise ValueError(
                    f"form='{form}' needs float_type='float32', {_BLOCKED_MIN_N} <= n <= "
                    f"{_BLOCKED_MAX_N} (n > {_BLOCKED_RANDOM_MIN_N} for a random order), "
                    "num_replicas <= 64 and no replay.")
            code = _lib.FORM_BLOCKED if form == 'blocked' else _lib.FORM_BLOCKED_EXACT
            return code, _lib.ORDER[blocked_order]
        return _lib.FORM_PLAIN, _lib.ORDER[update_strategy]

    def annealed_importance(self,
                            num_temps: int,
                            num_samps: int,
                            reference_prob: "str",
                            update_strategy: str = 'random',
                            reference_args=None,
                            beta: float = 1.0,
                            parallel: bool = True,
                            *,
                            seed: int = None,
                            ordering: str = 'reference',
                            replay: dict = None,
                            engine_stats: dict = None,
                            form: str = 'auto'):
        """Annealed importance sampling from a uniform or Bernoulli reference
        to the Boltzmann distribution at `beta` (pysa/sa.py:432-553).

        Returns (logZf, samples, omegas).  `ordering='reference'` takes each
        weight increment after the sweep at beta_t (pysa/simulation.py:305-316);
        'neal' takes it before (the unbiased AIS ordering).  `form`: 'plain',
        'blocked', 'blocked_exact' or 'auto'; 'auto' runs the MFMA form for
        float32 problems of 65 to 2048 variables with the uniform reference
        and no replay, where update_strategy='random' is the block-random
        order (as in metropolis_update)."""
        if reference_prob == "uniform":
            logZ0 = _ais.uniform_partition_fun(self.n_vars)
            initialize_prob = _ais.uniform_prob_initialization
            initial_args = None
            ref_code, p = 0, 0.5
        elif reference_prob == "bernoulli":
            logZ0 = _ais.bernoulli_partition_fun(self.n_vars)
            initialize_prob = _ais.bernoulli_prob_initialization
            initial_args = reference_args
            ref_code, p = 1, float(initial_args[0])
        else:
            raise ValueError("Reference Probability not recognized.")
        if update_strategy == 'random':
            order = _lib.ORDER['random']
        elif update_strategy == 'sequential':
            order = _lib.ORDER['sequential']
        else:
            raise ValueError(f"update_strategy='{update_strategy}' not recognized.")
        if self.problem_type not in ('qubo', 'ising'):
            raise ValueError(f"'{self.problem_type}' not supported for AIS.")
        if ordering not in ('reference', 'neal'):
            raise ValueError(f"ordering='{ordering}' not recognized.")

        ft = np.dtype(self.float_type)
        temps = np.array([(i + 1) / (num_temps) for i in range(num_temps)], dtype=self.float_type)
        couplings = beta * self.couplings.astype(self.float_type)
        local_fields = beta * self.local_fields.astype(self.float_type)

        initial_states = []
        for _ in range(num_samps):
            initial_states += [
                initialize_prob(self.n_vars, self.problem_type, initial_args).astype(self.float_type)
            ]
        initial_states = np.array(initial_states)
        spins = np.ascontiguousarray(self._spins(initial_states))
        rep_lr = rep_perm = None
        if replay is not None:
            rep_lr = np.ascontiguousarray(replay['log_r'], dtype=np.float64)
            if rep_lr.shape != (num_samps, num_temps, self.n_vars):
                raise ValueError("replay['log_r'] must have shape (samples, temps, n).")
            if replay.get('perm') is not None:
                rep_perm = np.ascontiguousarray(replay['perm'], dtype=np.int32)
                if rep_perm.shape != rep_lr.shape or not np.array_equal(
                        np.sort(rep_perm, axis=-1),
                        np.broadcast_to(np.arange(self.n_vars, dtype=np.int32), rep_perm.shape)):
                    raise ValueError("replay['perm'] rows must be permutations of range(n).")
            elif order != _lib.ORDER['sequential']:
                raise ValueError("replay of update_strategy='random' needs replay['perm'].")
        if seed is None:
            seed = secrets.randbits(64)
        codes = {'plain': _lib.FORM_PLAIN, 'blocked': _lib.FORM_BLOCKED,
                 'blocked_exact': _lib.FORM_BLOCKED_EXACT, 'auto': _lib.FORM_AUTO}
        if form not in codes:
            raise ValueError(f"form='{form}' not recognized.")
        mfma = (form != 'plain' and ft == np.float32 and ref_code == 0 and replay is None
                and 64 < self.n_vars <= 2048)
        if mfma and order == _lib.ORDER['random']:
            order = _lib.ORDER['block_random']

        prob = _lib.Problem(np.asarray(couplings, dtype=ft), np.asarray(local_fields, dtype=ft),
                            self.problem_type, ft)
        log_omegas = np.empty(num_samps, dtype=np.float64)
        kms = _lib._c_dbl(0.0)
        try:
            _lib.check(_lib.load().pysa_amd_ais(
                prob.handle, _lib.ptr(np.ascontiguousarray(temps)), int(num_temps),
                int(num_samps), ref_code, p, order, 0 if ordering == 'reference' else 1,
                int(seed) & (2**64 - 1), _lib.ptr(rep_lr), _lib.ptr(rep_perm), _lib.ptr(spins),
                _lib.ptr(log_omegas), None, codes[form], _lib.ctypes.byref(kms)))
        finally:
            prob.close()
        if engine_stats is not None:
            upd = float(self.n_vars) * num_temps * num_samps
            engine_stats.update({'kernel_ms': kms.value, 'updates': upd, 'seed': int(seed),
                                 'updates_per_s': upd / (kms.value * 1e-3) if kms.value else 0.0})
        samples = spins.astype(ft)
        logZf = _ais.omegas_to_partition(np.array(log_omegas), logZ0)
        lomax = log_omegas.max()
        normalization = np.sum(np.exp(log_omegas - lomax))
        omegas = [np.exp(log_omegas[i] - lomax) / normalization for i in range(num_samps)]
        return logZf, samples, omegas
