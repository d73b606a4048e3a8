"""Benchmark: spin updates/s and TTS(99%) of the engine at N=1024.

Workload (one "step"): the planted Wishart Ising instance, N=1024, alpha=0.5
(M=512), float32 couplings, h=0; 512 reads x 32 PT replicas x 5000 sweeps,
temperatures geometric 0.3 -> 1.5 and update_strategy='random' (PySA's
defaults).  The MFMA form runs 'random' in the block-random order (a random
order of the 32-spin blocks per sweep, a random XOR shuffle inside each
block); the plain form runs PySA's per-replica permutation.  All inputs
(couplings, initial states, ladders) are resident in HBM before the timed
region; a step re-copies the initial states device-to-device and runs every
read of the batch.

Prints one JSON line.  `value` is the headline engine's whole-job spin
updates/s (the MFMA-blocked form when this build carries it, else the plain
form).  Riding along in the same line: both forms with their energy and
field errors against f64 recomputation after the full run, the sequential
order, TTS(99%) with a bootstrap interval on a planted instance (alpha=1.25)
whose ground state the runs reach, the QUBO (N=512), batched-instance and AIS
configurations, the roofline of the dominant kernel, and a CPU baseline (the
oracle's C restatement of the reference algorithm on the host cores).
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

HBM_PEAK_GBS = 8000.0

# Bytes per launch of the blocked kernel past L2 (Infinity Cache or HBM) on
# the default workload, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
# this script (FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000, the factors
# a 256 MiB device-to-device copy gives); summary in
# profiles/wishart_n1024_kernels.md.
MEASURED_TRAFFIC = {("blocked", "block_random", 512, 32, 5000): 184.0e9}
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ladder(R, tmin=0.3, tmax=1.5, dtype=np.float32):
    """pysa/sa.py:189-194 (geometric temperatures, slot 0 coldest)."""
    ratio = np.exp(1 / (R - 1) * np.log(tmax / tmin))
    return np.array([1 / (tmin * ratio**k) for k in range(R)], dtype=dtype)


def nts(success, ptgt=0.99):
    """Restarts to reach the target confidence from the success fraction on
    the last axis (the reference's tests/test_tts.py:171-194): log(1 - ptgt)
    / log(1 - p), 1 when p >= ptgt, +inf when p = 0."""
    p = np.mean(np.asarray(success, dtype=bool), axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(p < ptgt, np.log(1.0 - ptgt) / np.log(1.0 - p), 1.0)
    return np.where(p > 0.0, out, np.inf)


def tts_bootstrap(success, t_read, nboots=200, ptgt=0.99, seed=0):
    """TTS(ptgt) = t_read * nts and its bootstrap over reads (one instance:
    the reference's tests/test_tts.py:257-277 resampling of repetitions)."""
    success = np.asarray(success, dtype=bool)
    rng = np.random.default_rng(seed)
    k = len(success)
    boots = np.array([float(nts(success[rng.integers(0, k, k)], ptgt)) for _ in range(nboots)])
    fin = boots[np.isfinite(boots)]
    point = float(nts(success, ptgt)) * t_read
    out = {"tts_s": point, "p_success": float(success.mean()), "reads": int(k),
           "t_read_s": t_read, "ptgt": ptgt, "boot_finite_frac": float(len(fin) / nboots)}
    if len(fin):
        out["boot_mean_s"] = float(fin.mean() * t_read)
        out["boot_std_s"] = float(fin.std() * t_read)
        out["boot_ci95_s"] = [float(np.quantile(fin, 0.025) * t_read),
                              float(np.quantile(fin, 0.975) * t_read)]
    return out


def bf16_pieces_sum(J, npieces):
    """The couplings the bf16 MFMA products multiply with: the sum of the
    first `npieces` round-to-nearest-even bf16 pieces of each float32."""
    v = np.asarray(J, dtype=np.float32).copy()
    acc = np.zeros_like(v)
    for _ in range(npieces):
        u = v.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        b = u.astype(np.uint32).view(np.float32)
        acc = (acc + b).astype(np.float32)
        v = (v - b).astype(np.float32)
    return acc


class Workload:
    """One benchmark batch resident in HBM (library-owned device buffers):
    problem P (couplings off the diagonal, fields on it), `reads` x R
    replicas x S sweeps, geometric ladder tmin -> tmax."""

    def __init__(self, P, ptype, reads, R, S, tmin, tmax, seed, e_gs=None):
        from pysa_amd import _lib

        self.lib = _lib
        D = _lib.DeviceArray
        self.n = P.shape[0]
        self.ptype = ptype
        self.reads, self.R, self.S = reads, R, S
        self.e_gs = e_gs
        self.J = P.astype(np.float32).copy()
        np.fill_diagonal(self.J, 0)
        self.h = np.diag(P).astype(np.float32).copy()
        self.prob = _lib.Problem(self.J, self.h, "ising" if ptype == 0 else "qubo", np.float32)
        rng = np.random.default_rng(seed)
        b = rng.integers(0, 2, (reads, R, self.n))
        st0 = (b if ptype == 1 else 2 * b - 1).astype(np.int8)
        en0 = self.prob.energy(st0.reshape(-1, self.n)).astype(np.float32).reshape(reads, R)
        self.betas = ladder(R, tmin, tmax)
        bsort = np.argsort(self.betas, kind="quicksort").astype(np.int32)
        self.d_st0, self.d_en0 = D.from_numpy(st0), D.from_numpy(en0)
        self.d_betas, self.d_bsort = D.from_numpy(self.betas), D.from_numpy(bsort)
        self.d_st = D(st0.shape, np.int8)
        self.d_en = D(en0.shape, np.float32)
        self.d_bi = D((reads, R), np.int64)
        self.d_bs = D((reads, self.n), np.int8)
        self.d_be = D(reads, np.float32)
        self.d_bw = D(reads, np.int32)
        self.d_acc = D(reads, np.uint64)
        self.updates = float(self.n) * R * reads * S

    def step(self, form, order, seed, block_b=0):
        """One pass over the batch (synchronous); returns the device time (ms)
        of the run, measured with HIP events on the launch stream."""
        L = self.lib
        self.d_st.copy_from(self.d_st0)
        self.d_en.copy_from(self.d_en0)
        kms = L._c_dbl(0.0)
        L.check(L.load().pysa_amd_metropolis_device(
            self.prob.handle, self.reads, self.R, self.S, order, 1, 0, 0,
            self.d_betas.ptr, self.d_bsort.ptr, seed, 0, form, block_b, None, None, None,
            self.d_st.ptr, self.d_en.ptr, self.d_bi.ptr, self.d_bs.ptr, self.d_be.ptr,
            self.d_bw.ptr, None, self.d_acc.ptr, L.ctypes.byref(kms), None))
        return kms.value

    def best_energies(self):
        """Per-read f64 energy of best_state."""
        return self.prob.energy(self.d_bs.numpy())

    def errors(self, form_name, sample=256):
        """Energy and field errors of the last run against f64 recomputation.

        energy: max |E_tracked - E_f64(state)| / max(1, |E_f64|) over every
        chain, and the same for best_energy / best_state.  field: over the
        first `sample` chains, max |F - J s| with J s in f64, where F is the
        plain form's resident f32 field after the run, or for the MFMA forms
        the product of the couplings they multiply with (bf16 piece sums) with
        the spins, in f64 (the f32 accumulation of the products adds at most
        a few ulp of the partial sums on top)."""
        st = self.d_st.numpy().reshape(-1, self.n)
        en = self.d_en.numpy().reshape(-1).astype(np.float64)
        e64 = self.prob.energy(st)
        rel = np.abs(en - e64) / np.maximum(1.0, np.abs(e64))
        bs = self.d_bs.numpy()
        be = self.d_be.numpy().astype(np.float64)
        b64 = self.prob.energy(bs)
        brel = np.abs(be - b64) / np.maximum(1.0, np.abs(b64))
        k = min(sample, len(st))
        s = st[:k].astype(np.float64)
        f64 = s @ self.J.astype(np.float64).T
        if form_name == "plain":
            f = self.prob.plain_fields(k).astype(np.float64)
        else:
            npieces = 3 if form_name == "blocked_exact" else 2
            f = s @ bf16_pieces_sum(self.J, npieces).astype(np.float64).T
        ferr = np.abs(f - f64)
        return {"energy_err_max_rel": float(rel.max()), "energy_err_mean_rel": float(rel.mean()),
                "best_energy_err_max_rel": float(brel.max()),
                "field_err_max_abs": float(ferr.max()),
                "field_err_max_rel": float(ferr.max() / max(1e-30, np.abs(f64).max())),
                "field_sample_chains": int(k)}


def copy_calibration(nbytes):
    """Device-to-device copies of a buffer of known size (the HIP runtime's
    copy kernel), so that profiler byte counters can be scaled: run under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE beside the engine."""
    from pysa_amd import _lib
    a = _lib.DeviceArray(nbytes, np.uint8)
    b = _lib.DeviceArray(nbytes, np.uint8)
    for _ in range(3):
        b.copy_from(a)
    log(f"[bench] copied {nbytes} bytes device-to-device 3 times")
    a.free()
    b.free()


FORM_CODES = {"plain": 0, "blocked": 1, "blocked_exact": 3}


def run_form(wl, form_name, order_name, steps, warmup, seed, errors=True):
    """Time `steps` passes of one form over the workload (after `warmup`
    untimed ones); device time from HIP events on the launch stream."""
    L = wl.lib
    form = FORM_CODES[form_name]
    order = L.ORDER[order_name]
    shape = None
    if form_name != "plain":
        # the autotuner's launch shape for this batch (timed afresh, untimed
        # here; results do not depend on it)
        import tempfile

        from pysa_amd import tune
        path = os.path.join(tempfile.gettempdir(), f"pysa_amd_bench_tune_{os.getpid()}.json")
        shape = tune.tune(wl.prob, wl.R, wl.reads, order, force=True, path=path)
        log(f"[bench] {form_name}/{order_name}: tuned launch shape kernel {shape['kernel']}, "
            f"tiles {shape['tiles']} ({shape['ms']})")
    for w in range(warmup):
        wl.step(form, order, seed + 7919 * (w + 1))
        log(f"[bench] {form_name}/{order_name}: warmup step {w + 1}/{warmup} done")
    # Each step call returns only after its device work has completed, so the
    # wall clock below brackets exactly `steps` runs.
    t0 = time.perf_counter()
    kms, acc, bests = [], 0, []
    for k in range(steps):
        kms.append(wl.step(form, order, seed + k))
        acc += int(wl.d_acc.numpy().sum())
        bests.append(wl.best_energies())
        log(f"[bench] {form_name}/{order_name}: step {k + 1}/{steps}: {kms[-1]:.1f} ms device")
    wall = time.perf_counter() - t0
    dev_s = float(np.sum(kms)) * 1e-3
    out = {
        "order": order_name,
        # inputs resident in HBM: device time of the K runs (HIP events)
        "value": wl.updates * steps / dev_s,
        "ms_per_step": 1e3 * dev_s / steps,
        "wall_ms_per_step": 1e3 * wall / steps,
        "kernel_ms_mean": float(np.mean(kms)),
        "acceptance": acc / (wl.updates * steps),
        "t_read_s": float(np.mean(kms)) * 1e-3 / wl.reads,
    }
    if shape is not None:
        out["launch_shape"] = {"kernel": "pipelined" if shape["kernel"] == 1 else "split-K",
                               "tiles_of_32_replicas": shape["tiles"], "block_size": 32,
                               "tuning_ms": shape["ms"], "tuning_sweeps": shape["sweeps"]}
    if wl.e_gs is not None:
        gaps = (np.concatenate(bests) - wl.e_gs) / abs(wl.e_gs)
        out["best_gap_median"] = float(np.median(gaps))
        out["p_ground_state"] = float(np.mean(gaps <= 1e-6))
    if errors:
        out.update(wl.errors(form_name))
    return out


def tts_c3(forms, reads, sweeps, tmin, tmax, alpha, R=32, steps=2, seed=1003):
    """TTS(99%) on a planted Wishart instance (N=1024) whose planted ground
    state PT reaches within the run: success = best_state's f64 energy within
    1e-6 |E_gs| of E_gs; t_read = device time per read, amortised over the
    batch; bootstrap over reads.  At alpha = 0.5 (the throughput instance)
    no read of 5000 sweeps reaches the planted state, so TTS is infinite
    there; alpha = 1.25 with temperatures scaled by sqrt(alpha / 0.5) is
    the hardest ratio at which both forms succeed in a few hundred sweeps."""
    from pysa_amd import _lib, instances

    P, e_gs = instances.planted_wishart(1024, alpha, seed)
    wl = Workload(P, 0, reads, R, sweeps, tmin, tmax, seed, e_gs=e_gs)
    res = {"workload": f"planted Wishart Ising N=1024 alpha={alpha} float32, {reads} reads x "
                       f"{R} replicas x {sweeps} sweeps, T {tmin}->{tmax}, "
                       f"update_strategy='random'",
           "e_gs": e_gs}
    for f in forms:
        order = "random" if f == "plain" else "block_random"
        kms, succ = [], []
        for k in range(steps):
            kms.append(wl.step(FORM_CODES[f], _lib.ORDER[order], seed + 101 * k))
            e = wl.best_energies()
            succ.append(e <= e_gs + 1e-6 * abs(e_gs))
        t_read = float(np.mean(kms)) * 1e-3 / reads
        res[f] = tts_bootstrap(np.concatenate(succ), t_read)
        res[f]["order"] = order
        log(f"[bench] TTS {f}: p = {res[f]['p_success']:.3f}, TTS99 = {res[f]['tts_s']:.4g} s")
    return res


def sparse_lattice(L=32, reads=512, R=32, sweeps=100, seed=1006):
    """The sparse form against the dense plain form on a periodic L x L
    lattice with +-1 couplings (4 nonzeros per row), random order, PT
    (T 0.5 -> 3.0): same results bit for bit, device time of one pass each.
    The sparse form reads 4 (column, value) pairs, 32 B, per accepted flip
    against n x 4 B for the dense row, but runs half the waves (its update
    needs more registers), and the per-visit work, not the row, bounds both.
    The MFMA form (int8 products of the dense +-1 matrix) is timed as well."""
    from pysa_amd import _lib

    rng = np.random.default_rng(seed)
    n = L * L
    P = np.zeros((n, n), np.float32)
    for i in range(L):
        for j in range(L):
            a = i * L + j
            for b in (((i + 1) % L) * L + j, i * L + (j + 1) % L):
                P[a, b] = P[b, a] = 2 * rng.integers(0, 2) - 1
    wl = Workload(P, 0, reads, R, sweeps, 0.5, 3.0, seed)
    res = {"workload": f"periodic {L}x{L} lattice Ising (n={n}), +-1 couplings, {reads} reads x "
                       f"{R} replicas x {sweeps} sweeps, T 0.5->3.0, update_strategy='random'"}
    states = {}
    for name, form in (("plain", _lib.FORM_PLAIN), ("sparse", _lib.FORM_SPARSE)):
        wl.step(form, _lib.ORDER["random"], seed)  # warm-up
        ms = wl.step(form, _lib.ORDER["random"], seed + 1)
        states[name] = wl.d_st.numpy()
        res[name] = {"value": wl.updates / (ms * 1e-3), "ms": ms}
        log(f"[bench] lattice {name}: {ms:.1f} ms device")
    res["sparse_over_plain"] = res["sparse"]["value"] / res["plain"]["value"]
    res["same_states"] = bool(np.array_equal(states["plain"], states["sparse"]))
    # the MFMA form on the same lattice (int8 couplings, block-random order):
    # a +-1 lattice is an exact int8 matrix, so the dense products win
    wl.step(_lib.FORM_BLOCKED, _lib.ORDER["block_random"], seed)
    ms = wl.step(_lib.FORM_BLOCKED, _lib.ORDER["block_random"], seed + 1)
    res["blocked"] = {"value": wl.updates / (ms * 1e-3), "ms": ms, "order": "block_random"}
    log(f"[bench] lattice blocked: {ms:.1f} ms device")
    return res


def qubo_c2(forms, steps=2, warmup=1, seed=1002):
    """Dense QUBO N=512 (Q upper triangular N(0,1) off the diagonal and on
    it, passed as P = Q + Q^T - diag Q), 256 reads x 32 replicas x 2000
    sweeps, update_strategy='random', T 0.1 -> 5.0."""
    from pysa_amd import instances

    P = instances.gaussian_qubo(512, seed)
    wl = Workload(P, 1, 256, 32, 2000, 0.1, 5.0, seed)
    res = {"workload": "dense QUBO N=512 float32, Gaussian Q, 256 reads x 32 replicas x 2000 "
                       "sweeps, T 0.1->5.0, update_strategy='random'",
           "updates": wl.updates}
    for f in forms:
        res[f] = run_form(wl, f, "random" if f == "plain" else "block_random", steps, warmup,
                          seed)
        if f == "plain":  # the same order as the MFMA form, for a like-for-like ratio
            res["plain_block_random"] = run_form(wl, f, "block_random", steps, warmup, seed)
    if "plain" in res and "blocked" in res:
        res["blocked_over_plain"] = res["blocked"]["value"] / res["plain"]["value"]
    return res


def batched_c4(sweeps=2000, count=2048, loop_instances=64, seed=1004):
    """Sherrington-Kirkpatrick batch: `count` instances, n uniform in
    [32, 768], J = +-1/sqrt(n) (int8 signs + scale), h = 0, 4 reads x 8
    replicas each, T 0.3 -> 1.5, random order.  One batched call with inputs
    resident in HBM, in the MFMA form (block-random order: one workgroup per
    instance, int8 block products; instances below 65 spins on the plain
    kernel in the same call) and in the plain form (a fresh permutation per
    replica and sweep), against a per-instance loop (one plain-form engine
    call per instance, as a loop over Solver objects would make), timed by
    device time (HIP events) over the first `loop_instances` instances."""
    from pysa_amd import _lib, batched, instances

    reads, R = 4, 8
    inst = instances.sk_batch(count, 32, 768, seed)
    pk = batched.pack(inst)
    ns = pk["n"].astype(np.int64)
    total = int(reads * R * ns.sum())
    rng = np.random.default_rng(seed)
    st = (2 * rng.integers(0, 2, total) - 1).astype(np.int8)
    en = batched.initial_energies(pk, inst, st, reads, R)
    betas = np.tile(ladder(R), (count, 1))
    s_off = np.concatenate([[0], np.cumsum(reads * R * ns)[:-1]]).astype(np.int64)
    D = _lib.DeviceArray
    dev = {k: D.from_numpy(v) for k, v in pk.items()}
    d_betas, d_soff = D.from_numpy(betas), D.from_numpy(s_off)
    d_st0, d_en0 = D.from_numpy(st), D.from_numpy(en)
    d_st, d_en = D(st.shape, np.int8), D(en.shape, np.float32)
    d_bi = D((count * reads, R), np.int64)
    d_bs = D(total // R, np.int8)
    d_be, d_bw = D(count * reads, np.float32), D(count * reads, np.int32)
    d_acc = D(count * reads, np.uint64)
    L = _lib

    def launch(sd, form, order):
        d_st.copy_from(d_st0)
        d_en.copy_from(d_en0)
        kms = L._c_dbl(0.0)
        L.check(L.load().pysa_amd_batched_device(
            count, dev["n"].ptr, dev["ld"].ptr, dev["j_off"].ptr, dev["signs"].ptr,
            dev["scale"].ptr, None, None, 0, reads, R, sweeps, order, 1, 0, d_betas.ptr, sd, 0,
            d_soff.ptr, d_st.ptr, d_en.ptr, d_bi.ptr, d_bs.ptr, d_be.ptr, d_bw.ptr, d_acc.ptr,
            form, L.ctypes.byref(kms), None))
        return kms.value

    updates = float(total) * sweeps
    out = {}
    for name, form, order in (("blocked", L.FORM_BLOCKED, 3), ("plain", L.FORM_PLAIN, 1)):
        launch(1, form, order)
        log(f"[bench] batched C4 {name}: warmup launch done")
        kms = launch(2, form, order)
        accs = d_acc.numpy().astype(np.float64)
        r = {"order": {1: "random", 3: "block_random"}[order], "kernel_ms": kms,
             "value": updates / (kms * 1e-3), "acceptance": float(accs.sum()) / updates}
        if name == "plain":
            # one n_k-byte int8 coupling row per accepted flip of instance k,
            # L2-resident (every matrix is at most 768 x 768 bytes)
            gbs = float(np.sum(accs * np.repeat(ns, reads))) / (kms * 1e-3) / 1e9
            r["roofline"] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                             "note": "int8 coupling row per accepted flip; served from L2"}
        else:
            # int8 block products: 2 n ops per spin update of the instances
            # the MFMA form takes (n >= 65)
            mf = ns >= 65
            tops = float(np.sum(2.0 * ns[mf] * ns[mf] * reads * R * sweeps)) / (kms * 1e-3) / 1e12
            r["roofline"] = {"bound": "mfma", "achieved": tops, "peak": 5000.0,
                             "unit": "TOP/s", "frac": tops / 5000.0, "traffic": None,
                             "note": "int8 MFMA, 2n ops per spin update, instances n >= 65 "
                                     f"({int(mf.sum())} of {count})"}
        out[name] = r
        log(f"[bench] batched C4 {name}: {kms:.1f} ms device")
    # Per-instance loop: one plain-form engine call per instance (the batch
    # as a caller's loop would run it), first `loop_instances` instances,
    # device time of each call.
    loop_ms, loop_upd = 0.0, 0.0
    for k in range(loop_instances):
        S_k, sc = inst[k]
        n = S_k.shape[0]
        prob = L.Problem((S_k.astype(np.float32) * sc).astype(np.float32),
                         np.zeros(n, np.float32), "ising", np.float32)
        o = int(s_off[k])
        st_k = D.from_numpy(st[o:o + reads * R * n])
        en_k = D.from_numpy(en[k * reads:(k + 1) * reads])
        bi_k, bs_k = D((reads, R), np.int64), D((reads, n), np.int8)
        be_k, bw_k = D(reads, np.float32), D(reads, np.int32)
        b_k = D.from_numpy(ladder(R))
        srt = D.from_numpy(np.argsort(ladder(R)).astype(np.int32))
        km = L._c_dbl(0.0)
        L.check(L.load().pysa_amd_metropolis_device(
            prob.handle, reads, R, sweeps, 1, 1, 0, 0, b_k.ptr, srt.ptr, 3, 0, L.FORM_PLAIN, 0,
            None, None, None, st_k.ptr, en_k.ptr, bi_k.ptr, bs_k.ptr, be_k.ptr, bw_k.ptr, None,
            None, L.ctypes.byref(km), None))
        loop_ms += km.value
        loop_upd += float(n) * R * reads * sweeps
        for a in (st_k, en_k, bi_k, bs_k, be_k, bw_k, b_k, srt):
            a.free()
        prob.close()
    log(f"[bench] per-instance loop over {loop_instances} instances: {loop_ms:.1f} ms device")
    for a in list(dev.values()) + [d_betas, d_soff, d_st0, d_en0, d_st, d_en, d_bi, d_bs, d_be,
                                   d_bw, d_acc]:
        a.free()
    loop_rate = loop_upd / (loop_ms * 1e-3)
    best = out["blocked"]
    return {
        "workload": f"{count} SK instances, n uniform in [32, 768] (sum {int(ns.sum())}), "
                    f"J = +-1/sqrt(n) int8 signs x scale, h = 0, {reads} reads x {R} replicas x "
                    f"{sweeps} sweeps, T 0.3->1.5, random order, PT",
        "updates": updates,
        "kernel_ms": best["kernel_ms"],
        "value": best["value"],
        "unit": "spin_updates/s",
        "form": "blocked (MFMA, block-random order)",
        "roofline": best["roofline"],
        "forms": out,
        "blocked_over_plain": best["value"] / out["plain"]["value"],
        "per_instance_loop": {"instances": loop_instances, "updates_per_s": loop_rate,
                              "note": "device time (HIP events) of one plain-form engine call "
                                      "per instance, first instances of the batch"},
        "batched_over_loop": best["value"] / loop_rate,
    }


def ais_c5(nt=10000, chains=4096, seed=1005):
    """Annealed importance sampling on the RBM-shaped Ising model: 784 visible
    + 500 hidden spins, W ~ N(0, 0.05^2), biases ~ N(0, 0.1^2), float32,
    `chains` chains over `nt` linear betas to 1, uniform reference, log-weights
    in f64 (pysa/sa.py:432-553): the MFMA form (block-random order, two bf16
    pieces) and the plain form (a fresh permutation per chain and step)."""
    from pysa_amd import _lib, instances

    P = instances.rbm_ising(784, 500, seed=seed)
    n = P.shape[0]
    J = P.copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float32).copy()
    prob = _lib.Problem(J.astype(np.float32), h, "ising", np.float32)
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    rng = np.random.default_rng(seed)
    st0 = (2 * rng.integers(0, 2, (chains, n)) - 1).astype(np.int8)
    updates = float(n) * nt * chains
    out = {}
    for name, form, order in (("blocked", _lib.FORM_BLOCKED, 3), ("plain", _lib.FORM_PLAIN, 1)):
        st = st0.copy()
        lw = np.empty(chains, np.float64)
        en = np.empty(chains, np.float64)
        kms = _lib._c_dbl(0.0)
        _lib.check(_lib.load().pysa_amd_ais(prob.handle, _lib.ptr(temps), nt, chains, 0, 0.5,
                                            order, 0, seed, None, None, _lib.ptr(st),
                                            _lib.ptr(lw), _lib.ptr(en), form,
                                            _lib.ctypes.byref(kms)))
        # tracked final energies against the f64 energies of the returned states
        e64 = prob.energy(st)
        err = np.abs(en - e64)
        m = lw.max()
        logz = m + math.log(np.exp(lw - m).sum()) - math.log(chains) + n * math.log(2)
        log(f"[bench] AIS C5 {name}: {kms.value:.1f} ms device")
        out[name] = {"order": {1: "random", 3: "block_random"}[order], "kernel_ms": kms.value,
                     "value": updates / (kms.value * 1e-3), "log_z": logz,
                     "energy_err_max_abs": float(err.max()),
                     "energy_err_max_rel": float((err / np.maximum(1.0, np.abs(e64))).max())}
    prob.close()
    best = out["blocked"]
    return {"workload": f"AIS, RBM-shaped Ising 784+500 float32, {chains} chains x {nt} betas, "
                        "uniform reference, random order",
            "updates": updates, "kernel_ms": best["kernel_ms"], "value": best["value"],
            "unit": "spin_updates/s", "log_z": best["log_z"],
            "form": "blocked (MFMA, block-random order)",
            "energy_err_max_abs": best["energy_err_max_abs"],
            "energy_err_max_rel": best["energy_err_max_rel"], "forms": out,
            "blocked_over_plain": best["value"] / out["plain"]["value"]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the CPU baseline: the affinity cores, capped at the
    process's thread budget (OMP_NUM_THREADS when set)."""
    aff = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", aff) or aff)
    return max(1, min(aff, cap)), aff


def cpu_baseline(target_s=12.0, threads=None, one_thread_s=4.0):
    """The oracle's C restatement of the reference algorithm (an O(N) coupling
    row dot per proposal, replicas in parallel as numba.prange does, serial
    PT), built with -ffast-math as numba's fastmath=True, on the host cores,
    over a bounded sample of the headline workload; and the same on one
    thread."""
    import oracle
    from pysa_amd import instances

    n, R = 1024, 32
    P, _ = instances.planted_wishart(n, 0.5, 1003)
    J = P.copy()
    np.fill_diagonal(J, 0)
    h = np.zeros(n, np.float32)
    rng = np.random.default_rng(1003)
    st0 = (2 * rng.integers(0, 2, (1, R, n)) - 1).astype(np.int8)
    Jd = J.astype(np.float64)
    en0 = np.array([[s @ Jd @ s / 2 for s in st0[0].astype(np.float64)]], dtype=np.float32)
    betas = ladder(R)
    aff = None
    if threads is None:
        threads, aff = cpu_threads()
    oracle.load(baseline=True)

    def timed(nthr, budget):
        st, en = st0, en0
        oracle.metropolis(J, h, 0, betas, st, en, 2, order=1, seed=4, nthreads=nthr,
                          baseline=True)  # warm
        # chunks of sweeps (continuing the same chain) until the budget has passed
        sweeps, dt, chunk = 0, 0.0, 10
        while dt < budget:
            t0 = time.perf_counter()
            out = oracle.metropolis(J, h, 0, betas, st, en, chunk, order=1, seed=6,
                                    sweep_offset=sweeps, nthreads=nthr, baseline=True)
            dt += time.perf_counter() - t0
            sweeps += chunk
            st, en = out["states"], out["energies"]
            chunk = max(10, int(chunk * min(4.0, 0.5 * budget / max(dt, 1e-3))))
        return n * R * sweeps / dt, sweeps, dt

    v, sw, dt = timed(threads, target_s)
    v1, sw1, dt1 = timed(1, one_thread_s)
    return {"value": v, "unit": "spin_updates/s", "cores": threads, "kind": "port",
            "affinity_cores": aff, "cpu_model": cpu_model(),
            "one_thread": {"value": v1, "sample": f"{sw1} sweeps in {dt1:.1f} s"},
            "build": "oracle/oracle.cpp, g++ -O3 -march=x86-64-v3 -ffast-math -fopenmp",
            "sample": f"1 read x {R} replicas x {sw} sweeps of the same N=1024 Wishart "
                      f"instance ({dt:.1f} s, random order, PT on)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=512)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=5000)
    ap.add_argument("--strategy", default="random", choices=["random", "sequential"],
                    help="update_strategy of the headline workload")
    ap.add_argument("--forms", default="auto", help="auto | plain | blocked | both")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline forms (no exact-pieces, other-order, QUBO, TTS lines)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-instances line")
    ap.add_argument("--no-ais", action="store_true", help="skip the AIS line")
    ap.add_argument("--no-tts", action="store_true", help="skip the TTS line")
    ap.add_argument("--tts-reads", type=int, default=512)
    ap.add_argument("--tts-sweeps", type=int, default=500)
    ap.add_argument("--tts-alpha", type=float, default=1.25)
    ap.add_argument("--tts-tmin", type=float, default=0.158)
    ap.add_argument("--tts-tmax", type=float, default=2.372)
    ap.add_argument("--seed", type=int, default=1003)
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed steps of the plain form when it is the comparison, not the headline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run device-to-device copies of BYTES (scales profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine shards "
              "reads across GPUs with no collective, so a multi-GPU figure is N x this one's.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    from pysa_amd import _lib, instances
    forms = {"auto": ["plain", "blocked"] if _lib.blocked_available() else ["plain"],
             "both": ["plain", "blocked"], "plain": ["plain"], "blocked": ["blocked"]}[args.forms]
    P, e_gs = instances.planted_wishart(1024, 0.5, args.seed)
    wl = Workload(P, 0, args.reads, args.replicas, args.sweeps, 0.3, 1.5, args.seed, e_gs=e_gs)

    def order_of(form, strategy):
        if strategy == "sequential":
            return "sequential"
        return "random" if form == "plain" else "block_random"

    res = {}
    head = "blocked" if "blocked" in forms else "plain"
    for f in forms:
        k = args.steps if f == head else min(args.steps, args.plain_steps)
        w = args.warmup if f == head else 0
        res[f] = run_form(wl, f, order_of(f, args.strategy), k, w, args.seed)
    if head == "blocked" and not args.no_extra:
        # three exact bf16 pieces, and the other visit order, one step each
        res["blocked_exact"] = run_form(wl, "blocked_exact", order_of("blocked", args.strategy),
                                        1, 0, args.seed)
        other = "sequential" if args.strategy == "random" else "random"
        res[f"blocked_{other}"] = run_form(wl, "blocked", order_of("blocked", other), 1, 0,
                                           args.seed)
    H = res[head]
    n = wl.n
    if head == "plain":
        # one coupling row (n x 4 B) per accepted flip
        bytes_per_ms = H["acceptance"] * wl.updates * n * 4 / H["kernel_ms_mean"]
        achieved = bytes_per_ms / 1e6  # GB/s
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
        dtype = "f32"
    else:
        flop = wl.updates * 2 * n * 2  # split-bf16: two products per coupling
        achieved = flop / (H["kernel_ms_mean"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                "useful_frac": achieved / 2 / BF16_PEAK_TFLOPS,
                "traffic": MEASURED_TRAFFIC.get(("blocked", H["order"], args.reads,
                                                 args.replicas, args.sweeps))}
        dtype = "bf16x2/f32"
    line = {
        "metric": "spin updates/s + TTS(99%) at N=1024: MFMA-blocked vs plain kernel, same GPU",
        "value": H["value"],
        "unit": "spin_updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": H["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"planted Wishart Ising N={n} alpha=0.5 float32, "
                               f"{args.reads} reads x {args.replicas} replicas x {args.sweeps} "
                               f"sweeps, T 0.3->1.5, update_strategy='{args.strategy}', PT",
                   "form": head, "order": H["order"],
                   "parallelism": "single GPU (reads across workgroups)"},
        "roofline": roof,
        "forms": res,
    }
    if "plain" in res and "blocked" in res:
        line["blocked_over_plain"] = res["blocked"]["value"] / res["plain"]["value"]
    if not args.no_extra:
        line["qubo_c2"] = qubo_c2(forms)
        line["sparse_lattice"] = sparse_lattice()
    if not args.no_tts:
        line["tts_c3"] = tts_c3(forms, args.tts_reads, args.tts_sweeps, args.tts_tmin,
                                args.tts_tmax, args.tts_alpha)
    if not args.no_batched:
        line["batched_c4"] = batched_c4()
    if not args.no_ais:
        line["ais_c5"] = ais_c5()
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
