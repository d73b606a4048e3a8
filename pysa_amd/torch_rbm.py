"""Torch-native negative phase for RBM training.

PySA's RBM tutorial samples the negative phase by rebuilding a `Solver` from
host copies of the weights every training step and copying the chains back
(tutorials/rbm_pysa.py:111-140: the augmented QUBO
-[[diag(b_v), W], [W^T, diag(b_h)]], one read of `state_size` replicas,
sequential order, PT between min_temp and max_temp, the current chains as
the initial states).  `update_states_with_pysa` runs the same sampling on
the device tensors themselves: the problem is built with torch on the GPU,
handed to the engine by device pointer (pysa_amd_problem_create_device),
and the sweeps run on torch's current stream into device tensors (the
initial energies in f64 on the device, as the Solver's energy kernel
computes them).  Only the R-entry temperature ladder is made on the host.

Requires a ROCm build of torch with the tensors on a GPU; everything else
of pysa_amd works without torch.  torch and the engine must share one HIP
runtime: import torch before the engine library is first loaded (torch
loads its bundled runtime, which the library then reuses); a process holding
two runtimes is refused with an error.
"""

import secrets

import numpy as np
import torch

from . import _lib


def _hip_runtimes():
    """Distinct HIP runtime files mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return {line.split()[-1] for line in f if "libamdhip64" in line}
    except OSError:
        return set()


def _check_one_runtime():
    rts = _hip_runtimes()
    if len(rts) > 1:
        raise RuntimeError("torch and pysa_amd loaded different HIP runtimes (" +
                           ", ".join(sorted(rts)) + "); import torch before the first "
                           "pysa_amd call")


def rbm_qubo(W, v_bias, h_bias):
    """The augmented QUBO of rbm_pysa.py:115-122 (negated, as it is passed to
    the Solver there), float32 on W's device."""
    V, H = W.shape
    n = V + H
    P = torch.zeros((n, n), dtype=torch.float32, device=W.device)
    P[:V, V:] = W
    P[V:, :V] = W.t()
    P += torch.diag(torch.cat([v_bias.reshape(-1), h_bias.reshape(-1)]).to(torch.float32))
    return -P


def ladder(R, min_temp, max_temp):
    """The Solver's geometric ladder without a zero beta (pysa/sa.py:178-199)."""
    if R == 1:
        return np.array([1 / min_temp], dtype=np.float32)
    if R == 2:
        return np.array([1 / min_temp, 1 / max_temp], dtype=np.float32)
    ratio = np.exp(1 / (R - 1) * np.log(max_temp / min_temp))
    return np.array([1 / (min_temp * ratio**k) for k in range(R)], dtype=np.float32)


def update_states_with_pysa(W, v_bias, h_bias, v_state, h_state, steps, min_temp=1.0,
                            max_temp=3.5, parallel=True, seed=None, form="plain"):
    """New (v_state, h_state) after `steps` PT sweeps from the given chains.

    W (V x H), v_bias (V or V x 1), h_bias (H or H x 1), v_state (R x V) and
    h_state (R x H) with 0/1 entries, all on one GPU.  Returns float32
    tensors on that device, the chains in replica order like
    `results['states'].values[0]` of the tutorial.  `form` is the engine form
    ('plain', the drop-in; 'auto' or 'blocked' use the MFMA form when it
    applies)."""
    dev = W.device
    if dev.type != "cuda":
        raise ValueError("update_states_with_pysa needs tensors on a GPU.")
    _lib.load()
    _check_one_runtime()
    R, V = v_state.shape
    H = h_state.shape[1]
    n = V + H
    with torch.no_grad():
        P = rbm_qubo(W.detach(), v_bias.detach(), h_bias.detach())
        h = torch.diagonal(P).contiguous()
        J = (P - torch.diag(h)).contiguous()
        x = torch.cat([v_state, h_state], dim=1).detach()
        states = x.to(torch.int8).reshape(1, R, n).contiguous()
        xd = x.to(torch.float64)
        Jd, hd = J.to(torch.float64), h.to(torch.float64)
        energies = (0.5 * ((xd @ Jd) * xd).sum(1) + xd @ hd).to(torch.float32).reshape(1, R)
        betas_h = ladder(R, min_temp, max_temp)
        betas = torch.from_numpy(betas_h).to(dev)
        bsort = torch.from_numpy(np.argsort(betas_h, kind="quicksort").astype(np.int32)).to(dev)
        beta_idx = torch.empty((1, R), dtype=torch.int64, device=dev)
        best_state = torch.empty((1, n), dtype=torch.int8, device=dev)
        best_energy = torch.empty(1, dtype=torch.float32, device=dev)
        best_sweeps = torch.empty(1, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev)
        stream.synchronize()  # J and h are written before the engine copies them
        handle = _lib._c_i64(0)
        lib = _lib.load()
        _lib.check(lib.pysa_amd_problem_create_device(n, 1, 0, J.data_ptr(), h.data_ptr(),
                                                      _lib.ctypes.byref(handle)))
        try:
            code = {"plain": _lib.FORM_PLAIN, "auto": _lib.FORM_AUTO,
                    "blocked": _lib.FORM_BLOCKED}[form]
            seed = secrets.randbits(64) if seed is None else int(seed)
            _lib.check(lib.pysa_amd_metropolis_device(
                handle.value, 1, R, int(steps), _lib.ORDER["sequential"], 1, 0,
                0 if parallel else 1, betas.data_ptr(), bsort.data_ptr(), seed & (2**64 - 1), 0,
                code, 0, None, None, None, states.data_ptr(), energies.data_ptr(),
                beta_idx.data_ptr(), best_state.data_ptr(), best_energy.data_ptr(),
                best_sweeps.data_ptr(), None, None, None, stream.cuda_stream))
        finally:
            lib.pysa_amd_problem_destroy(handle.value)
        out = states.reshape(R, n).to(torch.float32)
    return out[:, :V], out[:, V:]
