"""The oracle's random-number pieces, and the product's header against them.

* Philox4x32-10 against the known-answer vectors published with Random123
  (Salmon et al., SC'11; kat_vectors, philox4x32 with 10 rounds).
* det_log against numpy's log (<= 2 ulp) and exactly at powers of two.
* The Feistel visit order is a bijection of [0, n).
* pysa_amd/csrc/rng.hpp (the kernels' copy), compiled for the host with g++,
  gives the same words, logs and permutations as the oracle's own code.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KAT = [  # (counter, key, output)
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


def test_det_log_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        (rng.integers(0, 2**32, 200000) + 0.5) * 2.0**-32,  # the uniforms the engine draws
        np.exp(rng.uniform(-700, 700, 50000)),
        1 + rng.uniform(-1e-6, 1e-6, 1000),
        np.array([2.0**-32 * 0.5, 1 - 2.0**-33, 1.0, 0.5, 2.0, 1e-300, 1e300]),
    ])
    got = oracle.det_log(x)
    ref = np.log(x)
    ulp = np.spacing(np.abs(ref))
    err = np.abs(got - ref) / np.maximum(ulp, np.spacing(1e-300))
    assert err.max() <= 2.0, (err.max(), x[np.argmax(err)])


def test_det_log_powers_of_two():
    e = np.arange(-1000, 1000)
    got = oracle.det_log(np.ldexp(1.0, e))
    assert np.array_equal(got[e == 0], [0.0])
    ref = np.log(np.ldexp(1.0, e))
    assert np.all(np.abs(got - ref) <= 2 * np.spacing(np.abs(ref)))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 16, 17, 64, 130, 255, 1000, 1024, 4093])
def test_feistel_is_bijection(n):
    for keys in ([1, 2, 3, 4], [0xDEADBEEF, 7, 0, 0xFFFFFFFF]):
        p = oracle.perm(n, keys)
        assert np.array_equal(np.sort(p), np.arange(n))


def test_feistel_orders_differ_between_keys():
    a = oracle.perm(1024, [1, 2, 3, 4])
    b = oracle.perm(1024, [1, 2, 3, 5])
    assert not np.array_equal(a, b)
    assert not np.array_equal(a, np.arange(1024))


def test_uniform_stream_is_uniform():
    """Kolmogorov-Smirnov on the sweep stream's uniforms (exp of log_r)."""
    from scipy import stats
    lr, _ = oracle.sweep_stream(1234, 7, 3, 100000, 0)
    assert stats.kstest(np.exp(lr), "uniform").pvalue > 1e-4


_SHIM = r"""
#include "rng.hpp"
using namespace pysa;
extern "C" {
void shim_philox(const uint32_t* c, const uint32_t* k, uint32_t* o) {
  u32x4 r = philox4x32_10({c[0], c[1], c[2], c[3]}, k[0], k[1]);
  o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
}
void shim_log(const double* x, double* y, long n) { for (long i = 0; i < n; ++i) y[i] = det_log(x[i]); }
void shim_stream(uint64_t seed, uint64_t chain, uint64_t sweep, int n, int order, double* lr, int* pm) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int p = 0; p < n; ++p) lr[p] = log_u32(pick_word(sweep_block(chain, sweep, p, k0, k1), p & 3));
  if (order == 3) {
    FeistelKey f = block_order_key((n + 31) / 32, sweep, k0, k1);
    int o = 0;
    for (int i = 0; i < 32 * ((n + 31) / 32); ++i) {
      const int p = block_random_pos(f, sweep, i, n, k0, k1);
      if (p >= 0) pm[o++] = p;
    }
    return;
  }
  FeistelKey f = feistel_make(n, perm_keys(order == 2 ? 0 : chain, sweep,
                                           order == 2 ? kStreamShared : kStreamPerm, k0, k1));
  for (int i = 0; i < n; ++i) pm[i] = order == 0 ? i : (int)feistel_perm(f, i);
}
}
"""


@pytest.fixture(scope="module")
def product_rng(tmp_path_factory):
    d = tmp_path_factory.mktemp("rng")
    src = d / "shim.cpp"
    src.write_text(_SHIM)
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-fPIC", "-I", os.path.join(ROOT, "pysa_amd", "csrc"), str(src), "-o",
                    str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    vp = ctypes.c_void_p
    lib.shim_philox.argtypes = [vp, vp, vp]
    lib.shim_log.argtypes = [vp, vp, ctypes.c_long]
    lib.shim_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_int, vp, vp]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_product_header_matches_oracle(product_rng):
    for ctr, key, out in KAT:
        o = np.zeros(4, np.uint32)
        product_rng.shim_philox(_p(np.array(ctr, np.uint32)), _p(np.array(key, np.uint32)), _p(o))
        assert list(o) == out
    x = (np.random.default_rng(1).integers(0, 2**32, 100000) + 0.5) * 2.0**-32
    y = np.empty_like(x)
    product_rng.shim_log(_p(x), _p(y), x.size)
    assert np.array_equal(y, oracle.det_log(x))
    for order in (0, 1, 2, 3):
        for n, chain, sweep in ((5, 0, 0), (130, 77, 12), (1024, 2**30 + 5, 2**33 + 1)):
            lr = np.empty(n)
            pm = np.empty(n, np.int32)
            product_rng.shim_stream(99, chain, sweep, n, order, _p(lr), _p(pm))
            olr, opm = oracle.sweep_stream(99, chain, sweep, n, order)
            assert np.array_equal(lr, olr)
            assert np.array_equal(pm, opm)


@pytest.mark.parametrize("n", [1, 31, 32, 33, 65, 100, 1024, 2048])
def test_block_random_order_structure(n):
    """The block-random order (order 3) from its definition: every sweep
    visits each position once; blocks of 32 consecutive positions are visited
    whole, one after another, in a keyed order; inside block b slot k visits
    32 b + (k ^ m_b); the order is the same for every chain and changes with
    the sweep."""
    nblk = (n + 31) // 32
    seen_orders = set()
    for sweep in range(6):
        _, pm = oracle.sweep_stream(7, 0, sweep, n, 3)
        assert sorted(pm.tolist()) == list(range(n))
        _, pm2 = oracle.sweep_stream(7, 12345, sweep, n, 3)
        assert np.array_equal(pm, pm2)  # shared by every chain
        blocks = pm // 32
        # blocks appear as contiguous runs, each block once
        runs = [blocks[0]] + [b for a, b in zip(blocks[:-1], blocks[1:]) if a != b]
        assert sorted(runs) == list(range(nblk))
        for b in range(nblk):
            got = pm[blocks == b] - 32 * b
            nb = min(32, n - 32 * b)
            # slot k visits k ^ m for some mask m; slots past n are skipped
            masks = [m for m in range(32)
                     if [k ^ m for k in range(32) if (k ^ m) < nb] == got.tolist()]
            assert masks, (b, got)
        seen_orders.add(tuple(pm.tolist()))
    if n > 1:
        assert len(seen_orders) > 1
