"""Many independent instances of different sizes in one engine launch.

PySA solves a batch of problems as a Python loop over ``Solver`` objects
(pysa/sa.py:92-430 per instance, reads one after another at
pysa/sa.py:425-427).  ``solve_batch`` runs every read of every instance in
one call of ``pysa_amd_batched_device``, largest instances first: with
``form='plain'`` one workgroup per (instance, read) keeps each replica's fields
in registers; with ``form='blocked'`` (or ``'auto'``, the default) instances of
65 spins or more run the MFMA form, one workgroup per instance whose fields
come from int8 products of 32-spin blocks.  Both give the same results for the
same visit order.

Instances are given as integer coupling matrices S_k (int8, symmetric, zero
diagonal) with a float32 scale each, J_k = scale_k * S_k -- e.g. the
Sherrington-Kirkpatrick +-1/sqrt(n) ensemble -- and optional float32 local
fields.  The local field of a spin is the float32 value scale * (S_k s)_i + h_i
with the integer count S_k s kept exactly on the device.
"""

import numpy as np

from . import _lib

__all__ = ["pack", "initial_energies", "solve_batch"]


def pack(instances, h=None):
    """Pack [(S int8 (n, n), scale)] into the device layout: rows padded to a
    multiple of 16 bytes, instance k's matrix at j_off[k]."""
    ns = np.array([np.asarray(S).shape[0] for S, _ in instances], dtype=np.int32)
    if np.any(ns < 1) or np.any(ns > 2048):
        raise ValueError("instance sizes must be in [1, 2048]")
    ld = ((ns + 15) // 16 * 16).astype(np.int32)
    sizes = ns.astype(np.int64) * ld
    j_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    signs = np.zeros(int(sizes.sum()), dtype=np.int8)
    for k, (S, _) in enumerate(instances):
        S = np.asarray(S)
        n = ns[k]
        if S.shape != (n, n) or not np.array_equal(S, S.T) or np.any(np.diag(S) != 0):
            raise ValueError(f"instance {k}: couplings must be square, symmetric, zero-diagonal")
        if S.min() < -128 or S.max() > 127 or not np.array_equal(S, np.round(S)):
            raise ValueError(f"instance {k}: couplings must be int8 integers")
        block = signs[j_off[k]:j_off[k] + sizes[k]].reshape(n, ld[k])
        block[:, :n] = S.astype(np.int8)
    out = dict(n=ns, ld=ld, j_off=j_off, signs=signs,
               scale=np.array([s for _, s in instances], dtype=np.float32))
    if h is not None:
        hs = [np.asarray(x, dtype=np.float32) for x in h]
        if [len(x) for x in hs] != ns.tolist():
            raise ValueError("one local-field vector per instance, of its size")
        out["h"] = np.concatenate(hs).astype(np.float32)
        out["h_off"] = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    return out


def initial_energies(packed, instances, states, reads, R, h=None):
    """float32 energies E = s.(J s/2 + h) of the packed initial states,
    computed in f64 with J = scale * S (pysa/ising.py:37)."""
    out = []
    off = 0
    for k, (S, scale) in enumerate(instances):
        n = int(packed["n"][k])
        st = states[off:off + reads * R * n].reshape(reads * R, n).astype(np.float64)
        off += reads * R * n
        J = np.asarray(S, dtype=np.float64) * float(np.float32(scale))
        e = 0.5 * np.sum((st @ J) * st, axis=1)
        if h is not None:
            e = e + st @ np.asarray(h[k], dtype=np.float32).astype(np.float64)
        out.append(e.astype(np.float32).reshape(reads, R))
    return np.concatenate(out).reshape(-1, R)


def solve_batch(instances, *, problem_type="ising", reads=4, num_replicas=8, num_sweeps=1000,
                min_temp=0.3, max_temp=1.5, betas=None, update_strategy="random", h=None,
                use_pt=True, parallel=True, seed=0, sweep_offset=0, states=None,
                engine_stats=None, form="auto"):
    """Run every read of every instance; returns one dict per instance with
    PySA's per-read columns (states, energies, best_state, best_energy, temps,
    num_sweeps, min_sweeps).

    form: 'plain', 'blocked' or 'auto'.  With 'auto' and 'blocked',
    update_strategy='random' is the block-random order of the MFMA form (as in
    Solver.metropolis_update) when some instance has 65 spins or more; pass
    form='plain' for a fresh permutation per replica and sweep."""
    if problem_type not in ("ising", "qubo"):
        raise ValueError(f"problem_type='{problem_type}' not supported.")
    order = {"sequential": 0, "random": 1, "random_shared": 2,
             "block_random": 3}.get(update_strategy)
    if order is None:
        raise ValueError(f"update_strategy='{update_strategy}' not recognized.")
    codes = {"plain": _lib.FORM_PLAIN, "blocked": _lib.FORM_BLOCKED, "auto": _lib.FORM_AUTO}
    if form not in codes:
        raise ValueError(f"form='{form}' not recognized.")
    R = int(num_replicas)
    count = len(instances)
    pk = pack(instances, h)
    if form != "plain" and order == 1 and int(pk["n"].max()) >= 65:
        order = 3
    if betas is None:  # pysa/sa.py:167-194 (R >= 3) for every instance
        ratio = np.exp(1 / (R - 1) * np.log(max_temp / min_temp)) if R > 1 else 1.0
        lad = np.array([1 / (min_temp * ratio**k) for k in range(R)], dtype=np.float32)
        betas = np.tile(lad, (count, 1))
    betas = np.ascontiguousarray(np.broadcast_to(np.asarray(betas, np.float32), (count, R)))
    ptype = 0 if problem_type == "ising" else 1
    total = int(reads * R * pk["n"].astype(np.int64).sum())
    s_off = np.concatenate([[0], np.cumsum(reads * R * pk["n"].astype(np.int64))[:-1]])
    if states is None:
        b = np.random.randint(2, size=total)
        states = (b if ptype else 2 * b - 1).astype(np.int8)
    states = np.ascontiguousarray(states, dtype=np.int8)
    en = initial_energies(pk, instances, states, reads, R, h)
    D = _lib.DeviceArray
    dev = {k: D.from_numpy(v) for k, v in pk.items()}
    d_betas, d_soff = D.from_numpy(betas), D.from_numpy(s_off.astype(np.int64))
    d_st, d_en = D.from_numpy(states), D.from_numpy(en)
    d_bi = D((count * reads, R), np.int64)
    d_bs = D(total // R, np.int8)
    d_be = D(count * reads, np.float32)
    d_bw = D(count * reads, np.int32)
    d_acc = D(count * reads, np.uint64)
    kms = _lib._c_dbl(0.0)
    try:
        _lib.check(_lib.load().pysa_amd_batched_device(
            count, dev["n"].ptr, dev["ld"].ptr, dev["j_off"].ptr, dev["signs"].ptr,
            dev["scale"].ptr, dev["h"].ptr if "h" in dev else None,
            dev["h_off"].ptr if "h_off" in dev else None, ptype, reads, R, int(num_sweeps), order,
            int(bool(use_pt)), 0 if parallel else 1, d_betas.ptr, int(seed) & (2**64 - 1),
            int(sweep_offset), d_soff.ptr, d_st.ptr, d_en.ptr, d_bi.ptr, d_bs.ptr, d_be.ptr,
            d_bw.ptr, d_acc.ptr, codes[form], _lib.ctypes.byref(kms), None))
        st, ene, bi = d_st.numpy(), d_en.numpy(), d_bi.numpy()
        bs, be, bw, acc = d_bs.numpy(), d_be.numpy(), d_bw.numpy(), d_acc.numpy()
    finally:
        for a in list(dev.values()) + [d_betas, d_soff, d_st, d_en, d_bi, d_bs, d_be, d_bw, d_acc]:
            a.free()
    if engine_stats is not None:
        upd = float(total) * num_sweeps
        engine_stats.update(kernel_ms=kms.value, updates=upd, accepted=acc,
                            updates_per_s=upd / (kms.value * 1e-3) if kms.value else 0.0)
    out = []
    off = 0
    for k in range(count):
        n = int(pk["n"][k])
        S_k = st[off:off + reads * R * n].reshape(reads, R, n)
        B_k = bs[off // R:off // R + reads * n].reshape(reads, n)
        off += reads * R * n
        g = slice(k * reads, (k + 1) * reads)
        with np.errstate(divide="ignore"):
            temps = 1 / betas[k][bi[g]]
        out.append(dict(states=S_k, energies=ene[g], beta_idx=bi[g], best_state=B_k,
                        best_energy=be[g], best_sweeps=bw[g], temps=temps,
                        num_sweeps=int(num_sweeps), accepted=acc[g]))
    return out
