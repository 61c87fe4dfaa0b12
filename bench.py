#!/usr/bin/env python3
"""Benchmark of the regular-force path (gpunb_regf) on one MI355X.

Workload (one "step"): one regular-force call of the largest block-step i-set
of the N = 1,048,576 precision-study configuration -- a King W0 = 6 cluster
with 50 % of the stars in primordial binaries (a log-uniform in 0.005-50 AU,
thermal e), Kroupa masses, all N as j-particles, ni = 131,072 active
particles drawn at random, per-i RS giving log-normal neighbour counts (median
80, cap 500), dtr log-uniform in [2^-10, 2^-4].  Inputs are resident in HBM
(nb6_send_device / nb6_regf_device); the timed region is K such calls.

Output: one JSON line with the pair-interaction rate (ni * N per call, the
reference's own count, gpunb.velocity.cu:748), the roofline of the dominant
kernel (regf_tile_kernel, 60 flops per pair against the FP32 vector peak), the
accuracy of that step against the FP64 oracle on a sample of i-particles, and
the reference CPU back end (src/Main/reg.avx.cpp, built by oracle/Makefile)
timed on a bounded sample of the same step on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "nbody6ppgpu-beijing_amd"), str(ROOT / "oracle")]

from nb6 import ics, lib  # noqa: E402

N_J = 1 << 20
N_I = 1 << 17
LMAX, NNBMAX = 600, 550          # configure.ac LMAX for --with-par >= 128k; NNBMAX = LMAX-50
FP32_PEAK_TFLOPS = 157.3         # MI355X FP32 vector peak
FLOPS_PER_PAIR = 60              # gpunb.velocity.cu:895 convention
SEED = 20261202 + 2
PLAIN_REFINE = (0.5, 0.5, 2.0, 0.0)   # default FP64-pass weights of a plain call
ENS_REFINE = (0.5, 0.5, 0.5, 0.0)     # and of the batched ensemble step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(nj=N_J, ni=N_I, seed=SEED):
    t = time.time()
    w = ics.precision_study(nj, ni, seed)
    log(f"workload built in {time.time() - t:.1f} s")
    return w


def traffic_from_profiles():
    """HBM bytes per regf_tile launch, from the committed rocprofv3 summary."""
    p = ROOT / "profiles" / "regf_tile_king1m_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(m, x, v, idx, h2, dtr, target_s=15.0):
    """The reference's reg.avx.cpp (oracle/_ref) on a bounded sample of the step."""
    import oracle as O
    if not O.ref_available():
        return None
    threads = O.num_threads()
    R = O.RefReg(m.size)
    R.send(m, x, v)
    probe = 512
    t = time.time()
    R.regf(h2[:probe], dtr[:probe], x[idx[:probe]], v[idx[:probe]], LMAX, NNBMAX)
    rate = probe * m.size / max(time.time() - t, 1e-6)
    ns = int(min(idx.size, max(probe, rate * target_s / m.size)))
    t = time.time()
    res = R.regf(h2[:ns], dtr[:ns], x[idx[:ns]], v[idx[:ns]], LMAX, NNBMAX)
    dt = time.time() - t
    R.close()
    return {"value": ns * m.size / dt, "unit": "pair-interactions/s", "cores": threads,
            "kind": "reference",
            "sample": f"src/Main/reg.avx.cpp (g++ -O3 -fopenmp -mavx2 -mfma), {ns} of the "
                      f"{idx.size} i-particles of one step against all {m.size} j, "
                      f"{dt:.1f} s, OMP_NUM_THREADS={threads}"}, res


def accuracy(m, x, v, idx, h2, dtr, out, lst, k=256):
    """Per-particle relative error of the first k i-particles against FP64."""
    import oracle as O
    racc, rjrk, rpot, rlst = O.regf(m, x, v, h2[:k], dtr[:k], x[idx[:k]], v[idx[:k]], LMAX, NNBMAX)

    def rel(a, b):
        return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    ea = rel(out[0:3, :k].T, racc)
    ej = rel(out[3:6, :k].T, rjrk)
    ep = np.abs(out[6, :k] - rpot) / rpot
    same = all(lst[i, 0] == rlst[i, 0] and
               np.array_equal(lst[i, 1:1 + max(lst[i, 0], 0)], rlst[i, 1:1 + max(rlst[i, 0], 0)])
               for i in range(k))
    return {"sample_i": k,
            "acc_max": float(ea.max()), "acc_p99": float(np.percentile(ea, 99)),
            "jrk_max": float(ej.max()), "jrk_p99": float(np.percentile(ej, 99)),
            "pot_max": float(ep.max()), "lists_identical": bool(same)}


def bench_ensemble(steps, warmup):
    """Batched ensemble (256 Plummer clusters, N log-uniform in 1024-32768, Kroupa
    masses): one regular step of every cluster in one persistent hipGraph-replayed
    launch, against looping the plain gpunb path (send + regf) over the clusters."""
    import ctypes
    rng = np.random.Generator(np.random.PCG64(SEED + 40))
    nc = 256
    sizes = np.exp(rng.uniform(np.log(1024), np.log(32768), nc)).astype(int)
    t = time.time()
    ms, xs, vs, h2s, dts = [], [], [], [], []
    for c, n in enumerate(sizes):
        m, x, v = ics.plummer(int(n), SEED + 1000 + c, masses=ics.kroupa_masses(int(n), SEED + 2000 + c))
        rs = ics.rs_for_counts(x, np.arange(n), np.full(n, 50))
        r2 = (x * x).sum(1)
        ms.append(m); xs.append(x); vs.append(v); h2s.append(rs * rs)
        dts.append(np.minimum(0.125 * np.sqrt(1 + r2), 0.125) * 0.05)
    log(f"ensemble built in {time.time() - t:.1f} s, {sizes.sum()} particles")
    off = np.r_[0, np.cumsum(sizes)].astype(np.int32)
    ntot = int(off[-1])
    lmax, nnbmax = 400, 350
    m, x, v = np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
    h2, dtr = np.concatenate(h2s), np.concatenate(dts)
    L = lib.load()
    dj = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]])))
    di = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, h2[None], dtr[None]])))
    do = lib.DeviceArray(7 * ntot * 8)
    dl = lib.DeviceArray(ntot * lmax * 4)
    pairs = float((sizes.astype(float) ** 2).sum())

    # batched: send + one graph-replayed regular step per ensemble step, in the
    # configuration the autotuner picks on this ensemble
    L.nb6_ens_open(nc, ntot)
    L.nb6_ens_items.restype = ctypes.c_int
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    if L.nb6_ens_send_device(nc, offp, dj.ptr, ntot, None) != 0:
        lib.check_error()
    best = (ctypes.c_int * 4)()
    t_tune = time.perf_counter()
    if L.nb6_ens_autotune(di.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, best, None) != 0:
        lib.check_error()
    t_tune = time.perf_counter() - t_tune
    tuned = {"threads": best[0], "i_per_lane": best[1], "workgroups_per_cu": best[2],
             "step_us": best[3], "tuning_s": round(t_tune, 2)}
    log(f"ensemble autotune: {tuned}")

    def ens_step():
        if L.nb6_ens_send_device(nc, offp, dj.ptr, ntot, None) != 0:
            lib.check_error()
        if L.nb6_ens_regf_device(di.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None) != 0:
            lib.check_error()
    for _ in range(warmup):
        ens_step()
    L.nb6_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ens_step()
    L.nb6_sync()
    t_ens = (time.perf_counter() - t0) / steps
    # regular step alone (graph replays, send excluded)
    t0 = time.perf_counter()
    for _ in range(steps):
        L.nb6_ens_regf_device(di.ptr, ntot, lmax, nnbmax, 0, do.ptr, dl.ptr, 1, None)
    L.nb6_sync()
    t_ens_regf = (time.perf_counter() - t0) / steps
    refined_b = L.nb6_ens_last_refined()
    out_b = do.to_host(np.float64, (7, ntot))
    lst_b = dl.to_host(np.int32, (ntot, lmax))
    L.nb6_ens_close()
    L.nb6_set_regf_config(0, 1, 4096)      # the looped path runs the default configuration,
    L.nb6_set_regf_refine_weights(*ENS_REFINE)   # with the batched step's FP64-pass criterion

    # looped: the plain path, one cluster after the other
    g = lib.Gpunb(int(sizes.max()))

    def loop_step():
        for c in range(nc):
            a, n = int(off[c]), int(sizes[c])
            if L.nb6_send_device(n, ntot, dj.ptr + 8 * a, None) != 0:
                lib.check_error()
            if L.nb6_regf_device(n, ntot, di.ptr + 8 * a, lmax, nnbmax, 0, do.ptr + 8 * a,
                                 dl.ptr + 4 * lmax * a, None) != 0:
                lib.check_error()
    for _ in range(warmup):
        loop_step()
    L.nb6_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        loop_step()
    L.nb6_sync()
    t_loop = (time.perf_counter() - t0) / steps
    L.nb6_set_regf_refine_weights(*PLAIN_REFINE)
    out_l = do.to_host(np.float64, (7, ntot))
    lst_l = dl.to_host(np.int32, (ntot, lmax))
    # every particle of every cluster against the FP64 kernel (after the timed regions)
    L.nb6_set_regf_precision(2)
    loop_step()
    L.nb6_sync()
    L.nb6_set_regf_precision(0)
    ref = do.to_host(np.float64, (7, ntot))
    rl = dl.to_host(np.int32, (ntot, lmax))
    vs64 = {"particles": ntot, "acc": err_stats(rel_cols(out_b[0:3], ref[0:3])),
            "jerk": err_stats(rel_cols(out_b[3:6], ref[3:6])),
            "pot": err_stats(rel_cols(out_b[6], ref[6])), "lists_identical": same_rows(lst_b, rl)}
    g.close()
    for d in (dj, di, do, dl):
        d.free()
    # the two paths sum the FP32 partials in different orders: compare per particle
    da = np.linalg.norm(out_b[0:3] - out_l[0:3], axis=0) / np.linalg.norm(out_l[0:3], axis=0)
    dp = np.abs(out_b[6] - out_l[6]) / np.abs(out_l[6])
    same_lists = bool(np.array_equal(lst_b[:, 0], lst_l[:, 0]))
    for i in range(0, ntot, 997):
        c = lst_l[i, 0]
        if c > 0 and not np.array_equal(lst_b[i, 1:1 + c], lst_l[i, 1:1 + c]):
            same_lists = False
    return {"metric": "ensemble clusters/s (256 Plummer clusters, one regular step each)",
            "value": nc / t_ens, "unit": "clusters/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * t_ens, "higher_is_better": True,
            "dtype": "fp32", "data": "synthetic (seeded Plummer clusters, Kroupa masses)",
            "config": {"workload": "ensemble", "clusters": nc, "particles": ntot,
                       "pairs_per_step": pairs, "lmax": lmax, "nnbmax": nnbmax},
            "batched": {"clusters_per_s": nc / t_ens, "ms_per_step": 1e3 * t_ens,
                        "regf_only_ms": 1e3 * t_ens_regf, "pairs_per_s": pairs / t_ens,
                        "work_items": L.nb6_ens_items(), "autotuned": tuned},
            "looped": {"clusters_per_s": nc / t_loop, "ms_per_step": 1e3 * t_loop,
                       "pairs_per_s": pairs / t_loop},
            "speedup": t_loop / t_ens,
            "batched_vs_looped": {"acc_max_rel": float(da.max()), "pot_max_rel": float(dp.max()),
                                  "lists_identical": same_lists},
            "batched_vs_fp64_kernel": vs64,
            "fp64_pass": {"batched_rows": refined_b,
                          "weights_batched_and_looped": list(ENS_REFINE)}}


def bench_learned_rs(train_samples, eval_steps, no_cpu, lmax=600):
    """Learned neighbour radius: King W0=7, N=262144, NNBOPT=100, LMAX=600 (or `lmax`;
    NNBMAX = LMAX - 50).
    Pairs (features, radius giving NNBOPT neighbours) come from regular block steps
    of one realisation under the reference rule; the network trains on the GPU;
    then a second realisation runs the same number of block steps under the rule
    and under the network (rule as fallback), counting the util_gpu.F overflow
    re-runs of gpunb_regf."""
    from nb6.rsdrive import RSNet, Sim
    n, nnbopt = 262144, 100
    t = time.time()
    m, x, v = ics.king(n, SEED + 50, 7.0, masses=ics.kroupa_masses(n, SEED + 51))
    rec = []
    s = Sim(m, x, v, nnbopt=nnbopt, lmax=lmax)
    s.start()
    nsteps = 0
    while sum(r[1].size for r in rec) < train_samples:
        s.block_step("rule", record=rec)
        nsteps += 1
        if nsteps % 50 == 0:
            log(f"training data: {nsteps} block steps, {sum(r[1].size for r in rec)} pairs, "
                f"t = {s.t:.4f}, {time.time() - t:.0f} s")
    s.close()
    feat = np.concatenate([r[0] for r in rec])
    tgt = np.concatenate([r[1] for r in rec])
    t_data = time.time() - t
    t = time.time()
    net = RSNet(seed=SEED)
    hist = net.train(feat, tgt, epochs=4, batch=8192, lr=2e-3, log=log)
    t_train = time.time() - t
    m2, x2, v2 = ics.king(n, SEED + 60, 7.0, masses=ics.kroupa_masses(n, SEED + 61))
    res = {}
    for pol in ("rule", "net"):
        t = time.time()
        s = Sim(m2, x2, v2, nnbopt=nnbopt, lmax=lmax)
        s.start(pol, net if pol == "net" else None)
        s.run(eval_steps, pol, net=net if pol == "net" else None)
        st = dict(s.stats)
        st["wall_s"] = time.time() - t
        st["mean_abs_ln_count_ratio"] = st["absdev"] / max(st["samples"], 1)
        st["miss30_frac"] = st["miss30"] / max(st["samples"], 1)
        st["sim_time"] = s.t
        if not no_cpu:
            st["force_accuracy_vs_fp64"] = _rs_accuracy(s)
        s.close()
        res[pol] = st
        log(f"{pol}: {st}")
    net.close()
    r, q = res["rule"], res["net"]
    ratio = r["overflow_events"] / max(q["overflow_events"], 1)

    def split(st):
        # a particle's first regular step after start-up (radius from start-up, the same
        # for both policies' inputs) against its later steps (radius from the policy)
        return {"overflow_first_step": st["overflow_first"],
                "overflow_later_steps": st["overflow_later"],
                "later_samples": st["samples_later"],
                "later_miss30_frac": st["miss30_later"] / max(st["samples_later"], 1),
                "later_mean_abs_ln_count_ratio": st["absdev_later"] / max(st["samples_later"], 1)}
    steady = {"rule": split(r), "learned": split(q),
              "overflow_ratio_later_steps": (r["overflow_later"] / q["overflow_later"]
                                             if q["overflow_later"] else None),
              "miss30_ratio_later_steps": (split(r)["later_miss30_frac"] /
                                           max(split(q)["later_miss30_frac"], 1e-12))}
    return {"metric": "neighbour-list overflow retries: reference RS rule / learned RS "
                      f"(King W0=7 N=262144, NNBOPT=100, LMAX={lmax})",
            "value": ratio, "unit": "x fewer gpunb_regf re-runs", "n_gpus": 1,
            "start_up_vs_steady_state": steady,
            "higher_is_better": True, "dtype": "bf16 weights / fp32 accumulate",
            "data": "synthetic (seeded King W0=7, Kroupa masses)",
            "config": {"workload": "learned-rs", "n": n, "nnbopt": nnbopt, "lmax": lmax,
                       "nnbmax": lmax - 50, "eval_block_steps": eval_steps},
            "training": {"pairs": int(tgt.size), "block_steps": nsteps, "data_s": t_data,
                         "train_s": t_train, "device_epochs_s": net.train_s,
                         "samples_per_s": 4 * tgt.size / max(net.train_s, 1e-9),
                         "loss": hist, "target_var": float(tgt.var())},
            "rule": r, "learned": q}


def bench_precision(no_cpu):
    """Precision study at N=1M (King W0=6, 50% binaries, Kroupa): block i-sets with
    sizes drawn from p(n) ~ 1/n on [1, 2^17]; each block through the FP32 production
    kernel, the float-float kernel and the FP64 kernel.  Errors of acc, jerk, pot
    against the FP64 kernel on every particle of the block (the FP64 kernel is checked
    against the CPU oracle on a sample), list agreement, and throughput per mode;
    plus gpupot over all N against the FP64 kernel's all-pairs potential."""
    m, x, v, idx_all, h2_all, dtr_all = make_workload()
    L = lib.load()
    nj = m.size
    rng = np.random.Generator(np.random.PCG64(SEED + 70))
    sizes = np.unique(np.exp(rng.uniform(0.0, np.log(2.0 ** 17), 6)).astype(int).clip(1, 2 ** 17))
    sizes = np.r_[sizes, 2 ** 17]
    g = lib.Gpunb(nj)
    dj = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]])))
    if L.nb6_send_device(nj, nj, dj.ptr, None) != 0:
        lib.check_error()
    nmax = int(sizes.max())
    di = lib.DeviceArray(8 * nmax * 8)
    do = lib.DeviceArray(7 * nmax * 8)
    dl = lib.DeviceArray(nmax * LMAX * 4)
    names = {0: "fp32_fp64acc", 1: "float_float", 2: "fp64"}
    blocks = []
    for ni in sizes:
        sel = np.sort(rng.choice(idx_all.size, int(ni), replace=False))
        ib = np.ascontiguousarray(np.concatenate([x[idx_all[sel]].T, v[idx_all[sel]].T,
                                                  h2_all[sel][None], dtr_all[sel][None]]))
        L.nb6_memcpy_htod(di.ptr, ib.ctypes.data, ib.nbytes)
        res = {}
        for mode in (2, 0, 1):
            L.nb6_set_regf_precision(mode)
            if L.nb6_regf_device(int(ni), int(ni), di.ptr, LMAX, NNBMAX, 0, do.ptr, dl.ptr, None) != 0:
                lib.check_error()
            L.nb6_sync()
            ms = L.nb6_kernel_ms(b"regf_tile") + L.nb6_kernel_ms(b"regf_finalize")
            res[mode] = (do.to_host(np.float64, (7, int(ni))), dl.to_host(np.int32, (int(ni), LMAX)), ms)
        L.nb6_set_regf_precision(0)
        ref_out, ref_lst, _ = res[2]
        row = {"ni": int(ni)}
        for mode in (0, 1, 2):
            out, lst, ms = res[mode]
            ea = np.linalg.norm(out[0:3] - ref_out[0:3], axis=0) / np.linalg.norm(ref_out[0:3], axis=0)
            ej = np.linalg.norm(out[3:6] - ref_out[3:6], axis=0) / np.linalg.norm(ref_out[3:6], axis=0)
            ep = np.abs(out[6] - ref_out[6]) / ref_out[6]
            row[names[mode]] = {"pairs_per_s": float(ni) * nj / (ms * 1e-3), "ms": ms,
                                "acc_max": float(ea.max()), "jrk_p99": float(np.percentile(ej, 99)),
                                "jrk_max": float(ej.max()), "pot_max": float(ep.max()),
                                "lists_equal_fp64": bool(np.array_equal(lst[:, 0], ref_lst[:, 0]) and
                                                         all(np.array_equal(lst[i, 1:1 + max(lst[i, 0], 0)],
                                                                            ref_lst[i, 1:1 + max(lst[i, 0], 0)])
                                                             for i in range(0, int(ni), max(1, int(ni) // 512))))}
        if not no_cpu and ni == sizes[0]:
            # checker: the FP64 kernel against the CPU oracle on this small block
            import oracle as O
            k = min(int(ni), 64)
            racc, rjrk, rpot, _ = O.regf(m, x, v, h2_all[sel][:k], dtr_all[sel][:k], x[idx_all[sel][:k]],
                                         v[idx_all[sel][:k]], LMAX, NNBMAX)
            e = np.linalg.norm(ref_out[0:3, :k].T - racc, axis=1) / np.linalg.norm(racc, axis=1)
            row["fp64_kernel_vs_cpu_oracle_acc_max"] = float(e.max())
        blocks.append(row)
        log(f"precision block ni={ni}: " + ", ".join(
            f"{names[md]} {row[names[md]]['pairs_per_s']:.3e} pairs/s acc {row[names[md]]['acc_max']:.1e} "
            f"jrk99 {row[names[md]]['jrk_p99']:.1e}" for md in (0, 1)))
    # gpupot over all N against the FP64 kernel's all-pairs potential (h2 = 0)
    dx = lib.DeviceArray.from_host(np.ascontiguousarray(np.concatenate([x.T, m[None]])))
    dp = lib.DeviceArray(nj * 8)
    if L.nb6_pot_device(nj, nj, dx.ptr, 0, nj, dp.ptr, None) != 0:
        lib.check_error()
    L.nb6_sync()
    pot_ms = L.nb6_kernel_ms(b"pot")
    pot = dp.to_host(np.float64, nj)
    ks = np.arange(0, nj, nj // 4096)
    ib = np.ascontiguousarray(np.concatenate([x[ks].T, v[ks].T, np.zeros((2, ks.size))]))
    L.nb6_memcpy_htod(di.ptr, ib.ctypes.data, ib.nbytes)
    L.nb6_set_regf_precision(2)
    if L.nb6_regf_device(ks.size, ks.size, di.ptr, 2, 1, 0, do.ptr, dl.ptr, None) != 0:
        lib.check_error()
    L.nb6_set_regf_precision(0)
    ref_pot = do.to_host(np.float64, (7, ks.size))[6]
    ep = np.abs(pot[ks] - ref_pot) / ref_pot
    for d in (dj, di, do, dl, dx, dp):
        d.free()
    g.close()
    return {"metric": "precision study: regf pair-interactions/s and error vs the FP64 kernel "
                      "(King W0=6 N=1M, 50% binaries)",
            "value": blocks[-1]["fp32_fp64acc"]["pairs_per_s"], "unit": "pair-interactions/s",
            "n_gpus": 1, "higher_is_better": True, "dtype": "fp32 / float-float / fp64",
            "data": "synthetic (seeded King W0=6, 50% binaries, Kroupa masses)",
            "config": {"workload": "precision", "n_j": nj, "block_sizes": [int(s) for s in sizes]},
            "blocks": blocks,
            "gpupot_all_n": {"n": nj, "ms": pot_ms, "pairs_per_s": float(nj) * nj / (pot_ms * 1e-3),
                             "sample": int(ks.size), "max_rel_err_vs_fp64": float(ep.max()),
                             "median_rel_err_vs_fp64": float(np.median(ep))}}


def bench_irr(steps, warmup, no_cpu):
    """Irregular (neighbour-list) force gather at N=1M (the regf workload): lists of
    the ni=131072 i-set come straight from the device regf rows, predictor records
    from the device predictor store; one step = irr_simd_firr for all ni.
    Algorithmic bytes per active i: nnb (4 B index + 72 B record) + 72 B + 52 B."""
    m, x, v, idx, h2, dtr = make_workload()
    L = lib.load()
    n, ni = m.size, idx.size
    rng = np.random.Generator(np.random.PCG64(SEED + 80))
    f = rng.normal(0.0, 0.1, (n, 3))
    fd = rng.normal(0.0, 0.1, (n, 3))
    t0 = np.zeros(n)
    ti = 2.0 ** -8
    g = lib.Gpunb(n)
    jp = lib.JPred(n)
    irr = lib.IrrSimd(n, LMAX)
    addr = np.arange(1, n + 1, dtype=np.int32)
    jp.set(addr, x, v, f, fd, m, t0)
    jp.predict_send(ti, n)
    L.nb6_irr_sync_jp_(lib._ri(n))
    lib.check_error()
    # lists: the regf call of the i-set at the predicted state (device rows)
    xp = x + ti * (v + ti * (f + ti * fd))
    vp = v + 2 * ti * (f + 1.5 * ti * fd)
    ib = np.ascontiguousarray(np.concatenate([xp[idx].T, vp[idx].T, h2[None], dtr[None]]))
    di = lib.DeviceArray.from_host(ib)
    do = lib.DeviceArray(7 * ni * 8)
    dl = lib.DeviceArray(ni * LMAX * 4)
    if L.nb6_regf_device(ni, ni, di.ptr, LMAX, NNBMAX, 0, do.ptr, dl.ptr, None) != 0:
        lib.check_error()
    da = lib.DeviceArray.from_host((idx + 1).astype(np.int32))
    if L.nb6_irr_set_lists_device(ni, da.ptr, dl.ptr, LMAX, 1, None) != 0:
        lib.check_error()
    lst = dl.to_host(np.int32, (ni, LMAX))
    nnb_tot = float(np.maximum(lst[:, 0] - 1, 0).sum())
    dout = lib.DeviceArray(ni * 8 * 8)
    for _ in range(warmup):
        L.nb6_irr_firr_device(ti, ni, da.ptr, dout.ptr, None)
    L.nb6_sync()
    kms = []
    t_0 = time.perf_counter()
    for _ in range(steps):
        if L.nb6_irr_firr_device(ti, ni, da.ptr, dout.ptr, None) != 0:
            lib.check_error()
        kms.append(L.nb6_kernel_ms(b"irr"))
    L.nb6_sync()
    wall = (time.perf_counter() - t_0) / steps
    out = dout.to_host(np.float64, (ni, 8))
    kt = float(np.mean(kms)) * 1e-3
    alg = nnb_tot * (4 + 72) + ni * (72 + 52)
    # call time by block size: the first nb of the same active particles, on the device
    # (addresses and results in HBM) and through irr_simd_firr_vec_ (host arrays)
    sweep = []
    for nb in (1, 64, 1024, 16384):
        for _ in range(warmup):
            L.nb6_irr_firr_device(ti, nb, da.ptr, dout.ptr, None)
            irr.firr_vec(ti, idx[:nb] + 1)
        L.nb6_sync()
        t_0 = time.perf_counter()
        for _ in range(steps):
            if L.nb6_irr_firr_device(ti, nb, da.ptr, dout.ptr, None) != 0:
                lib.check_error()
        L.nb6_sync()
        dev_ms = (time.perf_counter() - t_0) * 1e3 / steps
        t_0 = time.perf_counter()
        for _ in range(steps):
            irr.firr_vec(ti, idx[:nb] + 1)
        abi_ms = (time.perf_counter() - t_0) * 1e3 / steps
        sweep.append({"ni": nb, "device_call_ms": dev_ms, "host_abi_call_ms": abi_ms})
        log(f"irr ni={nb}: device {dev_ms:.3f} ms, host ABI {abi_ms:.3f} ms")
    check = None
    if not no_cpu:
        import oracle as O
        k = 128
        lists = np.zeros((k, LMAX), np.int32)
        for q in range(k):
            c = lst[q, 0]
            row = [j + 1 for j in lst[q, 1:1 + c] if j != idx[q]]
            lists[q, 0] = len(row)
            lists[q, 1:1 + len(row)] = row
        full = np.zeros((n, LMAX), np.int32)
        full[idx[:k]] = lists
        racc, rjrk, rnn = O.firr(ti, idx[:k] + 1, x, v, f, fd, m, t0, full, LMAX)
        ea = np.linalg.norm(out[:k, 0:3] - racc, axis=1) / np.linalg.norm(racc, axis=1)
        ej = np.linalg.norm(out[:k, 3:6] - rjrk, axis=1) / np.linalg.norm(rjrk, axis=1)
        check = {"sample_i": k, "acc_max": float(ea.max()), "jrk_max": float(ej.max()),
                 "nnbid_equal": bool(np.array_equal(out[:k, 6].astype(int), rnn))}
    for d in (di, do, dl, da, dout):
        d.free()
    irr.close()
    jp.close()
    g.close()
    return {"metric": "irregular-force gather: algorithmic HBM GB/s (King W0=6 N=1M, ni=131072)",
            "value": alg / kt / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
            "dtype": "fp32 pair math, fp64 positions and sums",
            "data": "synthetic (seeded King W0=6, 50% binaries, Kroupa masses)",
            "config": {"workload": "irr", "n": n, "n_i": ni, "mean_nnb": nnb_tot / ni, "lmax": LMAX},
            "roofline": {"bound": "hbm", "achieved": round(alg / kt / 1e9, 1), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(alg / kt / 8e12, 4),
                         "algorithmic_bytes": alg, "kernel": "irr_predict_all + irr_force_kernel",
                         "avg_launch_ms": kt * 1e3},
            "check_vs_fp64": check, "call_time_by_block_size": sweep}


def bench_regcor(steps, warmup, no_cpu):
    """Regular corrector (nb6_regcor_device, REGCOR_GPU + FPCORR) of the ni = 131072
    block of the N = 1M regf workload: new rows and force from the device regf at RS,
    old lists from a device regf at radii and steps of the previous block; KZ(38) = 1.
    Each step restores the mutated state by device copies outside the timed kernels."""
    m, x, v, idx, h2, dtr = make_workload()
    L = lib.load()
    n, ni = m.size, idx.size
    rng = np.random.Generator(np.random.PCG64(SEED + 90))
    time_ = 1.0
    dtr_q = 2.0 ** np.round(np.log2(dtr))
    rs = np.sqrt(h2)
    counts0 = ics.lognormal_counts(ni, SEED + 4, median=80.0)
    fac = np.minimum(rng.uniform(0.85, 1.15, ni), (400.0 / counts0) ** (1.0 / 3.0))
    g = lib.Gpunb(n)
    g.send(m, x, v)

    def regf(rad, dt):
        ib = np.ascontiguousarray(np.concatenate([x[idx].T, v[idx].T, (rad * rad)[None], dt[None]]))
        di = lib.DeviceArray.from_host(ib)
        do = lib.DeviceArray(7 * ni * 8)
        dl = lib.DeviceArray(ni * LMAX * 4)
        if L.nb6_regf_device(ni, ni, di.ptr, LMAX, NNBMAX, 0, do.ptr, dl.ptr, None) != 0:
            lib.check_error()
        di.free()
        return do, dl

    do_old, dl_old = regf(rs * fac, dtr_q * 2.0 ** rng.integers(-1, 2, ni))
    old_rows = dl_old.to_host(np.int32, (ni, LMAX))
    do_old.free()
    dl_old.free()
    d_force, d_rows0 = regf(rs, dtr_q)
    rows = d_rows0.to_host(np.int32, (ni, LMAX))
    g.close()
    lst = np.zeros((n, LMAX), np.int32)
    nnb0 = np.zeros(ni, np.int64)
    for k, i in enumerate(idx):
        c = old_rows[k, 0]
        js = old_rows[k, 1:1 + max(c, 0)]
        js = js[js != i] + 1
        lst[i, 0] = len(js)
        lst[i, 1:1 + len(js)] = js
        nnb0[k] = len(js)
    dtr_all = np.full(n, 2.0 ** -6)
    dtr_all[idx] = dtr_q
    step = dtr_all / 2.0 ** rng.integers(0, 4, n)
    step[rng.random(n) < 0.05] = 2.0 ** -14
    z3 = np.zeros((3, n))
    S = dict(x=x.T.copy(), xdot=v.T.copy(), body=m, x0=(x - 1e-3 * v).T.copy(), x0dot=v.T.copy(),
             f=rng.normal(0, 0.1, (3, n)), fdot=rng.normal(0, 0.1, (3, n)), t0=time_ - step,
             h=None, step=step, timenw=time_ + step, fi=rng.normal(0, 0.1, (3, n)),
             fidot=rng.normal(0, 0.1, (3, n)), fr=rng.normal(0, 0.1, (3, n)),
             frdot=rng.normal(0, 0.1, (3, n)), d2=z3.copy(), d3=z3.copy(), t0r=time_ - dtr_all,
             stepr=dtr_all.copy(), rs=np.full(n, 0.01), list=lst, xn=z3.copy(), xndot=z3.copy(),
             d0=z3.copy(), d0r=z3.copy(), d1r=z3.copy(), d2r=z3.copy(), d3r=z3.copy())
    S["rs"][idx] = rs
    P = dict(time=time_, dtmin=2.0 ** -14, smin=2.0 ** -13, smax=0.125, etar=0.02, rscale=0.5,
             rc2=0.01, rsfac=12.0, znbmin=16.0, bodym=1.0 / n, tidal4=0.0, rdens=[0.0, 0.0, 0.0],
             n=n, ifirst=1, nnbopt=80, nnbmax=NNBMAX, nb_flag=1, m_flag=0, kz38=1)
    rc = lib.Regcor(S)
    del lst
    # pristine copies of everything the corrector reads and writes
    keep = ("fi", "fidot", "fr", "frdot", "d2", "d3", "t0r", "stepr", "rs", "step", "timenw",
            "list")
    saved = {}
    for k in keep:
        d = rc.dev[k]
        saved[k] = lib.DeviceArray(d.nbytes)
        L.nb6_device_copy(saved[k].ptr, d.ptr, d.nbytes, None)
    d_rows = lib.DeviceArray(d_rows0.nbytes)
    d_addr = lib.DeviceArray.from_host((idx + 1).astype(np.int32))
    d_cnt = lib.DeviceArray(lib.NCOUNT * 4)

    def restore():
        for k in keep:
            L.nb6_device_copy(rc.dev[k].ptr, saved[k].ptr, saved[k].nbytes, None)
        L.nb6_device_copy(d_rows.ptr, d_rows0.ptr, d_rows0.nbytes, None)
        L.nb6_memcpy_htod(d_cnt.ptr, np.zeros(lib.NCOUNT, np.int32).ctypes.data, 4 * lib.NCOUNT)

    def timed(mode):
        # mode 0: the default (packed neighbour records for a block this large), 2: the
        # neighbours' state read from the state arrays directly
        L.nb6_set_regcor_pack(mode)
        for _ in range(warmup):
            restore()
            rc.correct(P, ni, d_addr, d_force, ni, d_rows, d_cnt)
        L.nb6_sync()
        kms = []
        for _ in range(steps):
            restore()
            rc.correct(P, ni, d_addr, d_force, ni, d_rows, d_cnt)
            kms.append(L.nb6_kernel_ms(b"regcor"))
        L.nb6_sync()
        L.nb6_set_regcor_pack(0)
        return float(np.mean(kms)) * 1e-3
    kt_direct = timed(2)
    kt = timed(0)
    cnt = d_cnt.to_host(np.int32, (lib.NCOUNT,))
    out = rc.download()
    rows_g = d_rows.to_host(np.int32, (ni, LMAX))
    # algorithmic bytes of one call: per particle 276 B read (address, FREG/FDR, X/XDOT,
    # X0/X0DOT, FI/FIDOT/FR/FRDOT, T0R/STEPR/RS/STEP) + 288 B written (XN/XNDOT,
    # FI/FIDOT/FR/FRDOT, D0/D0R/D1R/D2R/D3R, T0R/STEPR/RS); list rows: regf row read,
    # converted row written, old list read, new list written where it changed; per
    # changed neighbour 136 B (X/XDOT/BODY, and T0/F/FDOT/X0DOT for FPCORR) + 8 B STEP;
    # per FPCORR call 144 B (F/FDOT of I, D2/D3 read and written)
    c = np.maximum(rows[:, 0], 0).astype(np.int64)
    nnb = rows_g[:, 0].astype(np.int64)
    changed = np.array([rows_g[k, 0] != nnb0[k] or not np.array_equal(
        rows_g[k, 1:1 + rows_g[k, 0]], S["list"][idx[k], 1:1 + nnb0[k]]) for k in range(ni)])
    alg = (ni * (276 + 288) + 4 * ((1 + c).sum() + (1 + nnb).sum() + (1 + nnb0).sum()
                                    + (1 + nnb[changed]).sum())
           + (int(cnt[1]) + int(cnt[2])) * 144 + int(cnt[3]) * 144)
    check = None
    cpu = None
    if not no_cpu:
        import regcor as RC
        sub = np.sort(rng.choice(ni, 256, replace=False))
        So = {k: (np.array(v_) if isinstance(v_, np.ndarray) else v_) for k, v_ in S.items()}
        RC.regcor_block(So, P, idx[sub] + 1, rc_force_host(d_force, ni)[:, sub], rows[sub].copy())
        ia = idx[sub]
        same = all(np.array_equal(out[k][..., ia], So[k][..., ia]) for k in lib.REG3 + lib.REG1)
        same = same and np.array_equal(out["list"][ia], So["list"][ia])
        check = {"sample_i": int(sub.size), "bit_identical": bool(same), "counters": cnt.tolist()}
        # the oracle (pure Python, one core) on a bounded sample of the same block
        k_cpu = 16384
        sub = np.sort(rng.choice(ni, k_cpu, replace=False))
        So = {k: (np.array(v_) if isinstance(v_, np.ndarray) else v_) for k, v_ in S.items()}
        t0 = time.perf_counter()
        RC.regcor_block(So, P, idx[sub] + 1, rc_force_host(d_force, ni)[:, sub], rows[sub].copy())
        tc = time.perf_counter() - t0
        cpu = {"value": k_cpu / tc, "unit": "particles/s", "cores": 1, "kind": "port",
               "sample": f"oracle/regcor.py (Python restatement of regcor_gpu.F + fpcorr.F), "
                         f"{k_cpu} of the {ni} particles of the block, {tc:.1f} s"}
    for d in list(saved.values()) + [d_rows, d_rows0, d_addr, d_cnt, d_force]:
        d.free()
    rc.free()
    return {"metric": "regular corrector: particles/s (King W0=6 N=1M, block ni=131072)",
            "value": ni / kt, "unit": "particles/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": kt * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp64 (REAL*4 where the reference is)",
            "data": "synthetic (seeded King W0=6, 50% binaries, Kroupa masses)",
            "config": {"workload": "regcor", "n": n, "n_i": ni, "lmax": LMAX, "kz38": 1,
                       "mean_nnb": float(nnb.mean()), "changed_fraction": float(changed.mean())},
            "roofline": {"bound": "hbm", "achieved": round(alg / kt / 1e9, 1), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(alg / kt / 8e12, 4),
                         "algorithmic_bytes": int(alg), "kernel": "regcor_lists + finish",
                         "avg_launch_ms": kt * 1e3, "unpacked_ms": kt_direct * 1e3},
            "cpu_baseline": cpu, "check_vs_oracle": check}


def bench_c2(steps, warmup, no_cpu):
    """Single regular-force calls at N = 65,536 (Plummer, Kroupa masses, all N as j) for
    i-sets of 1, 256, 4096 and 65,536 particles with log-normal neighbour counts
    (median 80, cap 500), lmax 500 and nnbmax 450 (configure.ac / input.F for 64k).
    Per size: the device-resident call (nb6_regf_device, inputs in HBM) and the host ABI
    call gpunb_regf_ (host arrays in and out), each timed over `steps` calls after
    `warmup`; the kernel times; every particle against the FP64 kernel; and the
    reference's reg.avx.cpp on the host cores for the same call."""
    n, lmax, nnbmax = 65536, 500, 450
    seed = SEED + 200
    mass = ics.kroupa_masses(n, seed)
    m, x, v = ics.plummer(n, seed + 1, masses=mass)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    L = lib.load()
    g = lib.Gpunb(n)
    g.send(m, x, v)
    ref, cores = None, None
    if not no_cpu:
        import oracle as O
        if O.ref_available():
            ref = O.RefReg(n)
            ref.send(m, x, v)
            cores = O.num_threads()
    rows = []
    for ni in (1, 256, 4096, 65536):
        idx = np.sort(rng.choice(n, size=ni, replace=False))
        counts = ics.lognormal_counts(ni, seed + 3 + ni, median=80.0)
        rs = ics.rs_for_counts(x, idx, counts)
        h2 = rs * rs
        dtr = np.exp(rng.uniform(np.log(2.0 ** -10), np.log(2.0 ** -4), ni))
        ib = np.ascontiguousarray(np.concatenate([x[idx].T, v[idx].T, h2[None], dtr[None]]))
        di = lib.DeviceArray.from_host(ib)
        do = lib.DeviceArray(7 * ni * 8)
        dl = lib.DeviceArray(ni * lmax * 4)

        def dev_call():
            if L.nb6_regf_device(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr, None) != 0:
                lib.check_error()

        for _ in range(warmup):
            dev_call()
        L.nb6_sync()
        # the call's latency: each call waited for, as a caller that needs the results
        # before it goes on does (no kernel timers, as in production)
        L.nb6_set_kernel_timing(0)
        t0 = time.perf_counter()
        for _ in range(steps):
            dev_call()
            L.nb6_sync()
        dev_ms = (time.perf_counter() - t0) * 1e3 / steps
        L.nb6_set_kernel_timing(1)
        kt, kf, kx = [], [], []
        for _ in range(steps):
            dev_call()
            kt.append(L.nb6_kernel_ms(b"regf_tile"))
            kf.append(L.nb6_kernel_ms(b"regf_finalize"))
            kx.append(max(L.nb6_kernel_ms(b"regf_chunkfix"), 0.0) +
                      max(L.nb6_kernel_ms(b"regf_fix"), 0.0))
        L.nb6_sync()
        chunks = L.nb6_last_regf_chunks()
        refined = L.nb6_last_regf_refined()
        out = do.to_host(np.float64, (7, ni))
        lst = dl.to_host(np.int32, (ni, lmax))
        full, ref64, _ = all_i_vs_fp64_kernel(L, ni, di, do, dl, out, lst, lmax, nnbmax)
        xi, vi = x[idx], v[idx]
        L.nb6_set_kernel_timing(0)
        for _ in range(warmup):
            g.regf(h2, dtr, xi, vi, lmax, nnbmax)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.regf(h2, dtr, xi, vi, lmax, nnbmax)
        abi_ms = (time.perf_counter() - t0) * 1e3 / steps
        L.nb6_set_kernel_timing(1)
        row = {"ni": ni, "pairs": ni * n, "device_call_ms": dev_ms,
               "tile_kernel_ms": float(np.median(kt)), "finalize_kernel_ms": float(np.median(kf)),
               "fp64_pass_ms": float(np.median(kx)), "fp64_pass_rows": refined,
               "j_chunks": chunks, "host_abi_call_ms": abi_ms,
               "pairs_per_s_device": ni * n / (dev_ms * 1e-3),
               "pairs_per_s_host_abi": ni * n / (abi_ms * 1e-3),
               "acc_max_vs_fp64": full["acc"]["max"], "jerk_p99_vs_fp64": full["jerk"]["p99"],
               "pot_max_vs_fp64": full["pot"]["max"], "lists_identical": full["lists_identical"]}
        if ref is not None:
            reps = max(1, min(steps, int(2e9 // (ni * n))))
            t0 = time.perf_counter()
            for _ in range(reps):
                res = ref.regf(h2, dtr, xi, vi, lmax, nnbmax)
            row["reference_cpu_call_ms"] = (time.perf_counter() - t0) * 1e3 / reps
            row["reference_cpu_acc_max_vs_fp64"] = float(rel_cols(res[0].T, ref64[0:3]).max())
            row["reference_cpu_pot_max_vs_fp64"] = float(rel_cols(res[2], ref64[6]).max())
        log(f"c2 ni={ni}: device {dev_ms:.3f} ms (tile {row['tile_kernel_ms']:.3f}, finalize "
            f"{row['finalize_kernel_ms']:.3f}, FP64 pass {row['fp64_pass_ms']:.3f} over "
            f"{refined} rows, {row['j_chunks']} chunks), host ABI "
            f"{abi_ms:.3f} ms, acc {row['acc_max_vs_fp64']:.1e}")
        rows.append(row)
        for d in (di, do, dl):
            d.free()
    g.close()
    if ref is not None:
        ref.close()
    return {"metric": "gpunb_regf call time and pair-interactions/s at N=65536 by i-set size",
            "value": rows[-1]["pairs_per_s_device"], "unit": "pair-interactions/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "fp32",
            "data": "synthetic (seeded Plummer, Kroupa masses)",
            "config": {"workload": "c2: one gpunb_regf call, N=65536, ni in 1/256/4096/65536",
                       "n_j": n, "lmax": lmax, "nnbmax": nnbmax,
                       "cpu_cores": cores},
            "calls": rows}


def bench_energy(steps, warmup, no_cpu):
    """Energy path of the GPU build (energy.F:28-62: GPUPOT over all bodies, PHICOR,
    POT/ZKIN sums) through nb6_energy_device on the N = 1M King W0 = 6 system with its
    binaries as single bodies.  The dominant kernel is pot_tile_kernel: N^2 pair terms,
    counted at 20 flops per pair (float-float difference 9, r^2 5, rsqrt 1, m rinv 1,
    compensated add 4) against the FP32 vector peak."""
    import ctypes
    m, x, v, idx, h2, dtr = make_workload()
    n = m.size
    L = lib.load()
    d = {"x": lib.DeviceArray.from_host(np.ascontiguousarray(x)),
         "xdot": lib.DeviceArray.from_host(np.ascontiguousarray(v)),
         "body": lib.DeviceArray.from_host(np.ascontiguousarray(m)),
         "list": lib.DeviceArray.from_host(np.zeros(n, np.int32)),
         "phidbl": lib.DeviceArray(n * 8)}
    e = lib.EnergyIn(nmax=n, lmax=1, n=n, ntot=n, npairs=0, cmsep2_rmin2=0.0)
    for k, a in d.items():
        setattr(e, k, a.ptr)
    out = (ctypes.c_double * 3)()

    def call():
        if L.nb6_energy_device(ctypes.byref(e), ctypes.cast(out, ctypes.c_void_p), None) != 0:
            lib.check_error()
    for _ in range(warmup):
        call()
    kms, pms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
        kms.append(L.nb6_kernel_ms(b"energy"))
        pms.append(L.nb6_kernel_ms(b"pot"))
    wall = (time.perf_counter() - t0) / steps
    te, tp = float(np.mean(kms)) * 1e-3, float(np.mean(pms)) * 1e-3
    pairs = float(n) * n
    phi = d["phidbl"].to_host(np.float64, (n,))
    zkin, pot = out[0], out[1]
    for a in d.values():
        a.free()
    check = cpu = None
    if not no_cpu:
        import oracle as O
        k = 256
        ref = O.pot(m, x, 1, k)
        err = np.abs(phi[:k] - ref) / ref
        check = {"sample_i": k, "phi_max_rel": float(err.max()),
                 "pot_vs_sample_mean": float(np.mean(m[:k] * (phi[:k] - ref)) / np.mean(m[:k] * ref)),
                 "zkin": zkin, "pot": pot}
        if O.ref_available():
            R = O._ref("pot")
            ns = 262144                      # about 5 s on 16 host cores
            n8 = (n + 7) // 8 * 8 + 8
            mp = np.zeros(n8)
            xp = np.zeros((n8, 3))
            mp[:n], xp[:n] = m, x
            po = np.zeros(n8)
            tc = time.perf_counter()
            R.gpupot_(ctypes.byref(ctypes.c_int(0)), ctypes.byref(ctypes.c_int(1)),
                      ctypes.byref(ctypes.c_int(ns)), ctypes.byref(ctypes.c_int(n)),
                      mp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      xp.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      po.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            tc = time.perf_counter() - tc
            cpu = {"value": ns * float(n) / tc, "unit": "pair-interactions/s",
                   "cores": O.num_threads(), "kind": "reference",
                   "sample": f"src/Main/pot.avx.cpp gpupot_ for {ns} of the {n} bodies against all "
                             f"{n}, {tc:.1f} s, OMP_NUM_THREADS={O.num_threads()}"}
    achieved = 20.0 * pairs / tp / 1e12
    return {"metric": "energy path (GPUPOT + PHICOR + sums): pair-interactions/s at N=1M",
            "value": pairs / te, "unit": "pair-interactions/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 pair math, fp64 separations and sums",
            "data": "synthetic (seeded King W0=6, 50% binaries, Kroupa masses)",
            "config": {"workload": "energy", "n": n, "npairs": 0},
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                         "kernel": "pot_tile_kernel", "avg_launch_ms": tp * 1e3,
                         "energy_call_ms": te * 1e3, "flops_per_pair": 20},
            "cpu_baseline": cpu, "check_vs_fp64": check}


def rc_force_host(d_force, ni):
    return d_force.to_host(np.float64, (7, ni))


def _rs_accuracy(s, k=128):
    """Regular force of k particles with the policy's current RS against FP64 (checker)."""
    import oracle as O
    idx = np.arange(0, s.n, s.n // k)[:k]
    xp, vp = s.predict(s.t)
    h2 = s.rs[idx] ** 2
    dtr = s.dt[idx]
    s._send(xp, vp)
    out, rows = s._regf(xp[idx], vp[idx], h2, dtr, s.d.lst, s.lmax, s.nnbmax)
    racc, rjrk, rpot, rlst = O.regf(s.m, xp, vp, h2, dtr, xp[idx], vp[idx], s.lmax, s.nnbmax)
    ea = np.linalg.norm(out[0:3].T - racc, axis=1) / np.linalg.norm(racc, axis=1)
    ep = np.abs(out[6] - rpot) / rpot
    return {"acc_max": float(ea.max()), "pot_max": float(ep.max()),
            "counts_identical": bool(np.array_equal(rows[:, 0], rlst[:, 0]))}


def err_stats(e):
    return {"max": float(e.max()), "p99_9": float(np.percentile(e, 99.9)),
            "p99": float(np.percentile(e, 99)), "median": float(np.median(e)),
            "n_above_1e-6": int((e > 1e-6).sum())}


def rel_cols(a, b):
    """Per-column |a - b| / |b| of SoA arrays [k][n]."""
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


def same_rows(a, b):
    """Every row's count and, for non-overflowing rows, its neighbours agree."""
    if not np.array_equal(a[:, 0], b[:, 0]):
        return False
    keep = np.arange(a.shape[1])[None, :] <= np.maximum(a[:, :1], 0)
    return bool(np.array_equal(np.where(keep, a, 0), np.where(keep, b, 0)))


def reference_vs_fp64(res, ref, rl):
    """The reference CPU back end's output on its sample against the FP64 kernel."""
    racc, rjrk, rpot, rlst = res
    k = len(rpot)
    return {"n_i": int(k), "acc": err_stats(rel_cols(racc.T, ref[0:3, :k])),
            "jerk": err_stats(rel_cols(rjrk.T, ref[3:6, :k])),
            "pot": err_stats(rel_cols(rpot, ref[6, :k])),
            "lists_identical": same_rows(rlst, rl[:k])}


def parity_vs_reference(res, out, lst, ref64, m, x, v, idx, h2, dtr):
    """This library's step against the reference's own reg.avx.cpp output on the CPU
    sample (checker, oracle/refband.py): list entries that differ, outside one FP32
    ulp of the threshold and outside the reference's derived FP32 rounding band, and
    the per-particle rule |ours - ref| <= 1e-6 |ref| + |ref - FP64|."""
    import refband as RB
    racc, rjrk, rpot, rlst = res
    k = len(rpot)
    c = RB.classify_rows(lst[:k], rlst, x, v, m, x[idx[:k]], v[idx[:k]], h2[:k], dtr[:k], 0)
    rule = {}
    for name, ours, ref, f64 in (("acc", out[0:3, :k].T, racc, ref64[0:3, :k].T),
                                 ("jerk", out[3:6, :k].T, rjrk, ref64[3:6, :k].T),
                                 ("pot", out[6, :k, None], rpot[:, None], ref64[6, :k, None])):
        nb, worst = RB.force_rule(ours, ref, f64)
        rule[name] = {"particles_breaking": nb, "worst_ratio": worst}
    return {"n_i": int(k), "lists": c, "force_rule_1e-6_plus_ref_error": rule}


def all_i_vs_fp64_kernel(L, ni, di, do, dl, out, lst, lmax=LMAX, nnbmax=NNBMAX):
    """Every i of the step against the FP64 kernel (nb6_set_regf_precision(2), which
    matches the CPU oracle to 1e-14): per-particle vector relative errors of acc, jerk
    and pot, and whether every row (count and neighbours) is identical.  Runs after
    the timed region."""
    L.nb6_set_regf_precision(2)
    try:
        if L.nb6_regf_device(ni, ni, di.ptr, lmax, nnbmax, 0, do.ptr, dl.ptr, None) != 0:
            lib.check_error()
        L.nb6_sync()
    finally:
        L.nb6_set_regf_precision(0)
    ref = do.to_host(np.float64, (7, ni))
    rl = dl.to_host(np.int32, (ni, lmax))
    ea = rel_cols(out[0:3], ref[0:3])
    ovf = lst[:, 0] < 0
    return {"n_i": int(ni), "acc": err_stats(ea), "jerk": err_stats(rel_cols(out[3:6], ref[3:6])),
            "pot": err_stats(rel_cols(out[6], ref[6])), "lists_identical": same_rows(lst, rl),
            "overflow_rows": int(ovf.sum()),
            "acc_max_overflow_rows": float(ea[ovf].max()) if ovf.any() else None}, ref, rl


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="regf",
                    choices=["regf", "ensemble", "learned-rs", "precision", "irr", "regcor",
                             "energy", "c2"],
                    help="regf: the headline N=1M step; ensemble: 256-cluster batch vs loop; "
                         "learned-rs: overflow retries of the RS network vs the reference rule")
    ap.add_argument("--rs-train-samples", type=int, default=10_000_000)
    ap.add_argument("--rs-eval-steps", type=int, default=300)
    ap.add_argument("--rs-lmax", type=int, default=600,
                    help="learned-rs: list length (NNBMAX = LMAX - 50); 200 makes the rule "
                         "overflow after start-up too")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one plain copy kernel of MIB MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        L = lib.load()
        nb = args.copy_calibration << 20
        a = lib.DeviceArray(nb)
        b = lib.DeviceArray(nb)
        L.nb6_device_copy(b.ptr, a.ptr, nb, None)
        L.nb6_sync()
        print(json.dumps({"copy_bytes": nb}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              f"(sharding across GPUs is not implemented)", file=sys.stderr)
        return 2
    # kernel times come from the library's HIP event pairs, recorded only when asked for
    lib.load().nb6_set_kernel_timing(1)

    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "irr":
        print(json.dumps(bench_irr(args.steps, args.warmup, args.no_cpu)), flush=True)
        return 0
    if args.workload == "regcor":
        print(json.dumps(bench_regcor(args.steps, args.warmup, args.no_cpu)), flush=True)
        return 0
    if args.workload == "c2":
        print(json.dumps(bench_c2(args.steps, args.warmup, args.no_cpu)), flush=True)
        return 0
    if args.workload == "energy":
        print(json.dumps(bench_energy(args.steps, args.warmup, args.no_cpu)), flush=True)
        return 0
    if args.workload == "precision":
        print(json.dumps(bench_precision(args.no_cpu)), flush=True)
        return 0
    if args.workload == "learned-rs":
        print(json.dumps(bench_learned_rs(args.rs_train_samples, args.rs_eval_steps,
                                          args.no_cpu, args.rs_lmax)), flush=True)
        return 0

    m, x, v, idx, h2, dtr = make_workload()
    L = lib.load()
    if L.nb6_device_count() < 1:
        raise RuntimeError("no GPU visible")
    ni, nj = idx.size, m.size
    g = lib.Gpunb(nj)
    js = np.ascontiguousarray(np.concatenate([x.T, v.T, m[None]]))      # SoA [7][nj]
    ib = np.ascontiguousarray(np.concatenate([x[idx].T, v[idx].T, h2[None], dtr[None]]))
    dj = lib.DeviceArray.from_host(js)
    di = lib.DeviceArray.from_host(ib)
    do = lib.DeviceArray(7 * ni * 8)
    dl = lib.DeviceArray(ni * LMAX * 4)
    if L.nb6_send_device(nj, nj, dj.ptr, None) != 0:
        lib.check_error()

    def step():
        rc = L.nb6_regf_device(ni, ni, di.ptr, LMAX, NNBMAX, 0, do.ptr, dl.ptr, None)
        if rc != 0:
            lib.check_error()
        # FP64 pass: segment recomputation (chunkfix) plus whole-row fixes (fix)
        return (L.nb6_kernel_ms(b"regf_tile"), L.nb6_kernel_ms(b"regf_finalize"),
                max(L.nb6_kernel_ms(b"regf_chunkfix"), 0.0) + max(L.nb6_kernel_ms(b"regf_fix"), 0.0))

    for _ in range(args.warmup):
        step()
    L.nb6_sync()
    tile_ms, fin_ms, fix_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, b, c = step()
        tile_ms.append(a)
        fin_ms.append(b)
        fix_ms.append(c)
    L.nb6_sync()
    wall = time.perf_counter() - t0
    chunks = L.nb6_last_regf_chunks()
    refined = L.nb6_last_regf_refined()
    refined_items = L.nb6_last_regf_refined_chunks()

    out = do.to_host(np.float64, (7, ni))
    lst = dl.to_host(np.int32, (ni, LMAX))
    full, ref64, rl64 = all_i_vs_fp64_kernel(L, ni, di, do, dl, out, lst)
    for d in (dj, di, do, dl):
        d.free()
    g.close()

    pairs = float(ni) * nj
    value = args.steps * pairs / wall
    t_tile = float(np.mean(tile_ms)) * 1e-3
    achieved = FLOPS_PER_PAIR * pairs / t_tile / 1e12
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic_from_profiles(),
            "kernel": "regf_tile_kernel", "avg_launch_ms": round(t_tile * 1e3, 3),
            "finalize_avg_ms": round(float(np.mean(fin_ms)), 3),
            "fp64_pass_avg_ms": round(float(np.mean(fix_ms)), 3), "fp64_pass_rows": refined,
            "fp64_pass_segments": refined_items, "j_chunks": chunks}
    # CPU leg: the reference's AVX back end timed on a sample, and the FP64 oracle
    # as the checker of the GPU result on the first 256 i-particles
    cpu, acc = None, None
    if not args.no_cpu:
        base = cpu_baseline(m, x, v, idx, h2, dtr)
        if base is not None:
            cpu, res = base
            full["reference_cpu_vs_fp64_kernel"] = reference_vs_fp64(res, ref64, rl64)
            full["parity_vs_reference"] = parity_vs_reference(res, out, lst, ref64, m, x, v,
                                                              idx, h2, dtr)
        acc = accuracy(m, x, v, idx, h2, dtr, out, lst)
    rec = {"metric": "regf pair-interactions/s at N=1M (%FP32 peak) at <=1e-6 rel err vs FP64",
           "value": value, "unit": "pair-interactions/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded King W0=6, 50% binaries, Kroupa masses)",
           "config": {"workload": "gpunb_regf, King W0=6 N=1048576 with 50% binaries, "
                                  "ni=131072 i-set, FP32 pair math, FP64 relative "
                                  "coordinates and accumulation",
                      "n_j": nj, "n_i": ni, "lmax": LMAX, "nnbmax": NNBMAX,
                      "pct_fp32_peak": round(100.0 * value * FLOPS_PER_PAIR / 1e12 /
                                             FP32_PEAK_TFLOPS, 2)},
           "roofline": roof, "cpu_baseline": cpu, "accuracy_vs_fp64": acc,
           "all_i_vs_fp64_kernel": full}
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
