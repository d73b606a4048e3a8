"""Batched uneven instances in one launch (pysa_amd_batched_device) against
the oracle's batched restatement, bit for bit: the plain form in every order,
the MFMA form (instances of 65 spins or more, the rest on the plain kernel in
the same call) in sequential and block-random order."""

import numpy as np
import pytest

import oracle
from pysa_amd import _lib, batched

pytestmark = pytest.mark.gpu


def make_instances(rng, sizes, kind):
    out = []
    for n in sizes:
        if kind == "sk":
            A = np.triu(2 * rng.integers(0, 2, (n, n)) - 1, 1)
            scale = np.float32(1 / np.sqrt(n))
        else:
            A = np.triu(rng.integers(-3, 4, (n, n)), 1)
            scale = np.float32(rng.choice([1.0, 0.37, 2.5]))
        out.append(((A + A.T).astype(np.int8), scale))
    return out


def run_both(inst, ptype, reads, R, S, order, best_sem, seed, with_h, rng, use_pt=True,
             form="plain"):
    h = [rng.normal(0, 0.5, S_.shape[0]).astype(np.float32) for S_, _ in inst] if with_h else None
    pk = batched.pack(inst, h)
    total = int(reads * R * pk["n"].astype(np.int64).sum())
    b = rng.integers(0, 2, total)
    st = (b if ptype else 2 * b - 1).astype(np.int8)
    betas = np.stack([np.geomspace(rng.uniform(1.5, 4.0), 0.2, R) for _ in inst]).astype(np.float32)
    en = batched.initial_energies(pk, inst, st, reads, R, h)
    want = oracle.batched(pk, ptype, reads, R, S, betas, st, en, order=order, use_pt=use_pt,
                          best_sem=best_sem, seed=seed, nthreads=8)
    got = batched.solve_batch(inst, problem_type="qubo" if ptype else "ising", reads=reads,
                              num_replicas=R, num_sweeps=S, betas=betas,
                              update_strategy={0: "sequential", 1: "random",
                                               2: "random_shared", 3: "block_random"}[order],
                              h=h, use_pt=use_pt, parallel=(best_sem == 0), seed=seed, states=st,
                              form=form)
    return got, want, pk


def compare(got, want, pk, reads, R):
    off = 0
    for k, g in enumerate(got):
        n = int(pk["n"][k])
        sl = slice(k * reads, (k + 1) * reads)
        assert np.array_equal(g["states"].reshape(-1), want["states"][off:off + reads * R * n]), k
        assert np.array_equal(g["best_state"].reshape(-1),
                              want["best_state"][off // R:off // R + reads * n]), k
        off += reads * R * n
        assert np.array_equal(g["energies"], want["energies"][sl]), k
        assert np.array_equal(g["beta_idx"], want["beta_idx"][sl]), k
        assert np.array_equal(g["best_energy"], want["best_energy"][sl]), k
        assert np.array_equal(g["best_sweeps"], want["best_sweeps"][sl]), k
        assert np.array_equal(g["accepted"], want["accepted"][sl]), k


@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("best_sem", [0, 1])
def test_batched_matches_oracle(ptype, order, best_sem):
    rng = np.random.default_rng(100 + 6 * ptype + 2 * order + best_sem)
    inst = make_instances(rng, [5, 33, 64, 100, 257, 31], "sk") + \
        make_instances(rng, [48, 130], "int")
    reads, R, S = 3, 5, 25
    got, want, pk = run_both(inst, ptype, reads, R, S, order, best_sem, 7, ptype == 0, rng)
    compare(got, want, pk, reads, R)


@pytest.mark.parametrize("R", [1, 8, 16])
def test_batched_replica_counts(R):
    rng = np.random.default_rng(200 + R)
    inst = make_instances(rng, [40, 90, 200], "sk")
    got, want, pk = run_both(inst, 0, 2, R, 15, 1, 0, 3, False, rng, use_pt=(R > 1))
    compare(got, want, pk, 2, R)


def test_batched_large_instances():
    """Instances above 1024 spins take the 32-slot register layout."""
    rng = np.random.default_rng(300)
    inst = make_instances(rng, [1100, 700, 2048], "sk")
    got, want, pk = run_both(inst, 0, 2, 4, 4, 0, 0, 11, False, rng)
    compare(got, want, pk, 2, 4)


def test_batched_rejects_bad_sizes():
    rng = np.random.default_rng(1)
    inst = make_instances(rng, [40], "sk")
    with pytest.raises(RuntimeError, match="R <= 16"):
        batched.solve_batch(inst, num_replicas=17, num_sweeps=2)


# ---- MFMA form

@pytest.mark.parametrize("ptype", [0, 1])
@pytest.mark.parametrize("order", [0, 3])
@pytest.mark.parametrize("best_sem", [0, 1])
def test_batched_mfma_matches_oracle(ptype, order, best_sem):
    """Mixed sizes around the MFMA threshold (65) and the block edges, SK and
    integer couplings at scales 1, 0.37 and 2.5, real local fields (ising)."""
    rng = np.random.default_rng(500 + 4 * ptype + 2 * (order == 3) + best_sem)
    inst = make_instances(rng, [5, 64, 65, 100, 257, 31, 96, 128, 161], "sk") + \
        make_instances(rng, [48, 130, 200], "int")
    reads, R, S = 3, 5, 25
    got, want, pk = run_both(inst, ptype, reads, R, S, order, best_sem, 7, ptype == 0, rng,
                             form="blocked")
    compare(got, want, pk, reads, R)


@pytest.mark.parametrize("reads,R", [(4, 8), (5, 16), (3, 1), (9, 4)])
def test_batched_mfma_tiles(reads, R):
    """One tile of 32 chains per workgroup (reads * R <= 32) and two (more),
    with several workgroups per instance."""
    rng = np.random.default_rng(600 + 17 * reads + R)
    inst = make_instances(rng, [70, 300, 129], "sk")
    got, want, pk = run_both(inst, 0, reads, R, 20, 3, 0, 5, True, rng, use_pt=(R > 1),
                             form="auto")
    compare(got, want, pk, reads, R)


@pytest.mark.parametrize("ptype", [0, 1])
def test_batched_mfma_paired_instances(ptype):
    """Workgroups pair tiles of different instances with the same number of
    32-spin blocks (sizes within one block range differ)."""
    rng = np.random.default_rng(650 + ptype)
    inst = make_instances(rng, [65, 70, 80, 96, 97, 100, 127, 128, 700, 704, 768], "sk") + \
        make_instances(rng, [90, 120], "int")
    got, want, pk = run_both(inst, ptype, 4, 8, 30, 3, ptype, 13, True, rng, form="blocked")
    compare(got, want, pk, 4, 8)


def test_batched_mfma_large_instances():
    """Up to 2048 spins: 16 spin words per matrix wave."""
    rng = np.random.default_rng(700)
    inst = make_instances(rng, [2048, 700, 1100], "sk")
    got, want, pk = run_both(inst, 0, 2, 4, 4, 3, 1, 11, True, rng, form="blocked")
    compare(got, want, pk, 2, 4)


def test_batched_forms_agree_at_full_length():
    """Plain and MFMA forms of the same block-random run, 300 sweeps of the
    C4 shape (4 reads x 8 replicas)."""
    rng = np.random.default_rng(800)
    inst = make_instances(rng, [768, 400, 97, 40], "sk")
    kw = dict(reads=4, num_replicas=8, num_sweeps=300, update_strategy="block_random", seed=9)
    b = rng.integers(0, 2, int(4 * 8 * sum(S_.shape[0] for S_, _ in inst)))
    st = (2 * b - 1).astype(np.int8)
    a = batched.solve_batch(inst, states=st, form="plain", **kw)
    m = batched.solve_batch(inst, states=st, form="blocked", **kw)
    for x, y in zip(a, m):
        for key in ("states", "energies", "beta_idx", "best_state", "best_energy",
                    "best_sweeps", "accepted"):
            assert np.array_equal(x[key], y[key]), key


def test_batched_mfma_needs_sequential_or_block_random():
    rng = np.random.default_rng(2)
    inst = make_instances(rng, [100], "sk")
    with pytest.raises(RuntimeError, match="sequential or block-random"):
        batched.solve_batch(inst, num_sweeps=2, update_strategy="random_shared", form="blocked")
