"""pysa_amd: MI355X-native Metropolis / parallel-tempering engine for dense
Ising and QUBO problems, with the public API of nasa/PySA's ``pysa.sa.Solver``.

    from pysa_amd.sa import Solver
    solver = Solver(problem, 'ising', 'float32')
    df = solver.metropolis_update(num_sweeps=100, num_reads=10, num_replicas=16)

The compute path is hand-written HIP for gfx950 behind a C ABI
(include/pysa_amd.h), loaded with ctypes from the in-tree libpysa_amd.so.
"""

__version__ = "0.1.0"

from . import ais  # noqa: F401
