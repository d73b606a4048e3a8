"""Torch-native RBM negative phase (tutorials/rbm_pysa.py:111-140) against
the Solver on host arrays: same chains, ladder, seed and form give the same
states bit for bit.  Runs in a child process that imports torch first, so
that torch and the engine share one HIP runtime."""

import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent("""
    import torch  # first: the engine reuses torch's HIP runtime
    import numpy as np
    from pysa_amd import torch_rbm
    from pysa_amd.sa import Solver

    g = torch.Generator().manual_seed(3)
    V, H, R, steps = 20, 12, 16, 10
    W = (0.3 * torch.randn(V, H, generator=g)).cuda()
    vb = (0.2 * torch.randn(V, 1, generator=g)).cuda()
    hb = (0.2 * torch.randn(H, 1, generator=g)).cuda()
    v = torch.randint(0, 2, (R, V), generator=g).float().cuda()
    h = torch.randint(0, 2, (R, H), generator=g).float().cuda()
    for parallel in (True, False):
        nv, nh = torch_rbm.update_states_with_pysa(W, vb, hb, v, h, steps, parallel=parallel,
                                                   seed=77)
        assert nv.device.type == "cuda" and nv.shape == (R, V) and nh.shape == (R, H)
        P = torch_rbm.rbm_qubo(W, vb, hb).cpu().numpy()
        x0 = torch.cat([v, h], 1).cpu().numpy()
        res = Solver(problem=P, problem_type="qubo", float_type="float32").metropolis_update(
            num_sweeps=steps, num_reads=1, num_replicas=R, update_strategy="sequential",
            min_temp=1.0, max_temp=3.5, initialize_strategy=x0, parallel=parallel,
            verbose=False, seed=77, form="plain")
        want = np.stack(res["states"].values[0])
        got = torch.cat([nv, nh], 1).cpu().numpy()
        assert np.array_equal(got, want), parallel
    print("torch rbm ok")
""")


def test_torch_negative_phase_equals_solver():
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "torch rbm ok" in r.stdout
