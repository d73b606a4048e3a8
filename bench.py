"""Benchmark: spin updates/s and TTS(99%) of the engine at N=1024.

Workload (one "step"): the planted Wishart Ising instance, N=1024, alpha=0.5
(M=512), float32 couplings, h=0; 512 reads x 32 PT replicas x 5000 sweeps,
temperatures geometric 0.3 -> 1.5 (PySA's defaults), sequential update
order (the order both engine forms run with the same semantics).
All inputs (couplings, initial states, ladders) are resident in HBM before
the timed region; a step re-copies the initial states device-to-device and
runs every read of the batch.

Prints one JSON line.  `value` is the headline engine's whole-job spin
updates/s (the MFMA-blocked form when this build carries it, else the plain
form); both forms, their ratio, TTS(99%), the roofline of the dominant kernel
and a CPU baseline (the oracle's C restatement of the reference algorithm on
the host cores) ride along in the same line.
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
# a 1 GiB device-to-device copy gives); summary in
# profiles/wishart_n1024_kernels.md.
MEASURED_TRAFFIC = {("blocked", 512, 32, 5000): 152.2e9}
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ladder(R, tmin=0.3, tmax=1.5, dtype=np.float32):
    """pysa/sa.py:189-194 (geometric temperatures, slot 0 coldest)."""
    ratio = np.exp(1 / (R - 1) * np.log(tmax / tmin))
    return np.array([1 / (tmin * ratio**k) for k in range(R)], dtype=dtype)


def tts(p, t_read, target=0.99):
    if p <= 0:
        return math.inf
    if p >= target:
        return t_read
    return t_read * math.log(1 - target) / math.log(1 - p)


class Workload:
    """The benchmark batch, resident in HBM (library-owned device buffers)."""

    def __init__(self, reads, R, S, seed=1003):
        from pysa_amd import _lib, instances

        self.lib = _lib
        D = _lib.DeviceArray
        self.n = 1024
        self.reads, self.R, self.S = reads, R, S
        P, self.e_gs = instances.planted_wishart(self.n, 0.5, seed)
        self.J = P.copy()
        np.fill_diagonal(self.J, 0)
        self.h = np.diag(P).astype(np.float32).copy()
        self.prob = _lib.Problem(self.J, self.h, "ising", np.float32)
        rng = np.random.default_rng(seed)
        st0 = (2 * rng.integers(0, 2, (reads, R, self.n)) - 1).astype(np.int8)
        en0 = self.prob.energy(st0.reshape(-1, self.n)).astype(np.float32).reshape(reads, R)
        self.betas = ladder(R)
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

    def success(self):
        """Per-read f64 energy of best_state and relative gap to E_gs."""
        bs = self.d_bs.numpy()
        e = self.prob.energy(bs)
        gap = (e - self.e_gs) / abs(self.e_gs)
        return e, gap


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


def run_form(wl, form_name, steps, warmup, seed):
    L = wl.lib
    form = {"plain": L.FORM_PLAIN, "blocked": L.FORM_BLOCKED}[form_name]
    order = L.ORDER["sequential"]
    for w in range(warmup):
        wl.step(form, order, seed + 7919 * (w + 1))
        log(f"[bench] {form_name}: warmup step {w + 1}/{warmup} done")
    # Each step call returns only after its device work has completed, so the
    # wall clock below brackets exactly `steps` runs.
    t0 = time.perf_counter()
    kms, acc, gaps = [], 0, []
    for k in range(steps):
        kms.append(wl.step(form, order, seed + k))
        acc += int(wl.d_acc.numpy().sum())
        _, gap = wl.success()
        gaps.append(gap)
        log(f"[bench] {form_name}: step {k + 1}/{steps}: {kms[-1]:.1f} ms device")
    wall = time.perf_counter() - t0
    gaps = np.concatenate(gaps)
    ms_step = 1e3 * wall / steps
    t_read = np.mean(kms) * 1e-3 / wl.reads
    p0 = float(np.mean(gaps <= 1e-6))
    dev_s = float(np.sum(kms)) * 1e-3
    out = {
        # inputs resident in HBM: device time of the K runs (HIP events)
        "value": wl.updates * steps / dev_s,
        "ms_per_step": 1e3 * dev_s / steps,
        "wall_ms_per_step": ms_step,
        "kernel_ms_mean": float(np.mean(kms)),
        "kernel_updates_per_s": wl.updates / (np.mean(kms) * 1e-3),
        "acceptance": acc / (wl.updates * steps),
        "p_success": p0,
        "tts99_s": tts(p0, t_read),
        "gap_tts99_s": {f"{g:g}": tts(float(np.mean(gaps <= g)), t_read)
                        for g in (0.01, 0.05, 0.1, 0.15, 0.2)},
        "best_gap_median": float(np.median(gaps)),
        "t_read_s": t_read,
    }
    return out


def batched_c4(sweeps=2000, count=2048, loop_instances=8, seed=1004):
    """Sherrington-Kirkpatrick batch: `count` instances, n uniform in
    [32, 768], J = +-1/sqrt(n) (int8 signs + scale), h = 0, 4 reads x 8
    replicas each, T 0.3 -> 1.5, random order.  One batched launch with inputs
    resident in HBM, against a per-instance launch loop (one engine call per
    instance, as a loop over Solver objects would make) timed on the first
    `loop_instances` instances."""
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

    def launch(sd):
        d_st.copy_from(d_st0)
        d_en.copy_from(d_en0)
        kms = L._c_dbl(0.0)
        L.check(L.load().pysa_amd_batched_device(
            count, dev["n"].ptr, dev["ld"].ptr, dev["j_off"].ptr, dev["signs"].ptr,
            dev["scale"].ptr, None, None, 0, reads, R, sweeps, 1, 1, 0, d_betas.ptr, sd, 0,
            d_soff.ptr, d_st.ptr, d_en.ptr, d_bi.ptr, d_bs.ptr, d_be.ptr, d_bw.ptr, d_acc.ptr,
            L.ctypes.byref(kms), None))
        return kms.value

    launch(1)
    log("[bench] batched C4: warmup launch done")
    kms = launch(2)
    updates = float(total) * sweeps
    acc = int(d_acc.numpy().sum())
    log(f"[bench] batched C4: {kms:.1f} ms device")
    # Per-instance loop: one plain-form engine call per instance (the batch
    # as a caller's loop would run it), first `loop_instances` instances.
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
        t0 = time.perf_counter()
        L.check(L.load().pysa_amd_metropolis_device(
            prob.handle, reads, R, sweeps, 1, 1, 0, 0, b_k.ptr, srt.ptr, 3, 0, L.FORM_PLAIN, 0,
            None, None, None, st_k.ptr, en_k.ptr, bi_k.ptr, bs_k.ptr, be_k.ptr, bw_k.ptr, None,
            None, L.ctypes.byref(km), None))
        loop_ms += 1e3 * (time.perf_counter() - t0)
        loop_upd += float(n) * R * reads * sweeps
        for a in (st_k, en_k, bi_k, bs_k, be_k, bw_k, b_k, srt):
            a.free()
        prob.close()
    log(f"[bench] per-instance loop over {loop_instances} instances: {loop_ms:.1f} ms wall")
    for a in list(dev.values()) + [d_betas, d_soff, d_st0, d_en0, d_st, d_en, d_bi, d_bs, d_be,
                                   d_bw, d_acc]:
        a.free()
    return {
        "workload": f"{count} SK instances, n uniform in [32, 768] (sum {int(ns.sum())}), "
                    f"J = +-1/sqrt(n) int8 signs x scale, h = 0, {reads} reads x {R} replicas x "
                    f"{sweeps} sweeps, T 0.3->1.5, random order, PT",
        "updates": updates,
        "kernel_ms": kms,
        "value": updates / (kms * 1e-3),
        "unit": "spin_updates/s",
        "acceptance": acc / updates,
        "per_instance_loop": {"instances": loop_instances, "updates_per_s": loop_upd / (loop_ms * 1e-3),
                              "note": "wall time of one plain-form engine call per instance"},
        "batched_over_loop": (updates / (kms * 1e-3)) / (loop_upd / (loop_ms * 1e-3)),
    }


def ais_c5(nt=10000, chains=4096, seed=1005):
    """Annealed importance sampling on the RBM-shaped Ising model: 784 visible
    + 500 hidden spins, W ~ N(0, 0.05^2), biases ~ N(0, 0.1^2), float32,
    `chains` chains over `nt` linear betas to 1, uniform reference, random
    order, log-weights in f64 (pysa/sa.py:432-553)."""
    from pysa_amd import _lib, instances

    P = instances.rbm_ising(784, 500, seed=seed)
    n = P.shape[0]
    J = P.copy()
    np.fill_diagonal(J, 0)
    h = np.diag(P).astype(np.float32).copy()
    prob = _lib.Problem(J.astype(np.float32), h, "ising", np.float32)
    temps = np.array([(t + 1) / nt for t in range(nt)], dtype=np.float32)
    rng = np.random.default_rng(seed)
    st = (2 * rng.integers(0, 2, (chains, n)) - 1).astype(np.int8)
    lw = np.empty(chains, np.float64)
    kms = _lib._c_dbl(0.0)
    _lib.check(_lib.load().pysa_amd_ais(prob.handle, _lib.ptr(temps), nt, chains, 0, 0.5, 1, 0,
                                        seed, None, None, _lib.ptr(st), _lib.ptr(lw),
                                        _lib.ctypes.byref(kms)))
    prob.close()
    updates = float(n) * nt * chains
    m = lw.max()
    logz = m + math.log(np.exp(lw - m).sum()) - math.log(chains) + n * math.log(2)
    log(f"[bench] AIS C5: {kms.value:.1f} ms device")
    return {"workload": f"AIS, RBM-shaped Ising 784+500 float32, {chains} chains x {nt} betas, "
                        "uniform reference, random order",
            "updates": updates, "kernel_ms": kms.value, "value": updates / (kms.value * 1e-3),
            "unit": "spin_updates/s", "log_z": logz, "form": "plain"}


def cpu_baseline(target_s=12.0, threads=None):
    """The oracle's C restatement of the reference algorithm (O(N) coupling
    row dot per proposal, replicas in parallel as numba.prange does) on the
    host cores, over a bounded sample of the same workload."""
    import oracle
    from pysa_amd import instances

    n, R = 1024, 32
    P, _ = instances.planted_wishart(n, 0.5, 1003)
    J = P.copy()
    np.fill_diagonal(J, 0)
    h = np.zeros(n, np.float32)
    rng = np.random.default_rng(1003)
    st = (2 * rng.integers(0, 2, (1, R, n)) - 1).astype(np.int8)
    Jd = J.astype(np.float64)
    en = np.array([[s @ Jd @ s / 2 for s in st[0].astype(np.float64)]], dtype=np.float32)
    betas = ladder(R)
    if threads is None:
        threads = min(16, len(os.sched_getaffinity(0)))
    oracle.load()
    oracle.metropolis(J, h, 0, betas, st, en, 5, order=0, seed=4, nthreads=threads)  # warm
    # Chunks of sweeps (continuing the same chain) until target_s has passed.
    sweeps, dt, chunk = 0, 0.0, 40
    while dt < target_s:
        t0 = time.perf_counter()
        out = oracle.metropolis(J, h, 0, betas, st, en, chunk, order=0, seed=6,
                                sweep_offset=sweeps, nthreads=threads)
        dt += time.perf_counter() - t0
        sweeps += chunk
        st, en = out["states"], out["energies"]
        chunk = max(40, int(chunk * min(4.0, 0.5 * target_s / max(dt, 1e-3))))
    return {"value": n * R * sweeps / dt, "unit": "spin_updates/s", "cores": threads,
            "kind": "port",
            "sample": f"1 read x {R} replicas x {sweeps} sweeps of the same N=1024 Wishart "
                      f"instance ({dt:.1f} s, sequential order, PT on)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--reads", type=int, default=512)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=5000)
    ap.add_argument("--forms", default="auto", help="auto | plain | blocked | both")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-instances line")
    ap.add_argument("--no-ais", action="store_true", help="skip the AIS line")
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

    from pysa_amd import _lib
    forms = {"auto": ["plain", "blocked"] if _lib.blocked_available() else ["plain"],
             "both": ["plain", "blocked"], "plain": ["plain"], "blocked": ["blocked"]}[args.forms]
    wl = Workload(args.reads, args.replicas, args.sweeps, args.seed)
    res = {}
    head = "blocked" if "blocked" in forms else "plain"
    for f in forms:
        k = args.steps if f == head else min(args.steps, args.plain_steps)
        res[f] = run_form(wl, f, k, args.warmup, args.seed)
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
                "traffic": MEASURED_TRAFFIC.get(("blocked", args.reads, args.replicas,
                                                 args.sweeps))}
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
                               "sweeps, T 0.3->1.5, sequential order, PT",
                   "form": head, "parallelism": "single GPU (reads across workgroups)"},
        "roofline": roof,
        "forms": res,
    }
    if "plain" in res and "blocked" in res:
        line["blocked_over_plain"] = res["blocked"]["value"] / res["plain"]["value"]
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
