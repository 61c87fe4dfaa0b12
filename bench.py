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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(nj=N_J, ni=N_I, seed=SEED):
    t = time.time()
    mass = ics.kroupa_masses(nj, seed)
    m, x, v = ics.king(nj, seed + 1, w0=6.0, masses=mass)
    m, x, v = ics.add_binaries(m, x, v, 0.5, seed + 2)
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    idx = np.sort(rng.choice(nj, size=ni, replace=False))
    counts = ics.lognormal_counts(ni, seed + 4, median=80.0)
    rs = ics.rs_for_counts(x, idx, counts)
    h2 = rs * rs
    dtr = np.exp(rng.uniform(np.log(2.0 ** -10), np.log(2.0 ** -4), ni))
    log(f"workload built in {time.time() - t:.1f} s")
    return m, x, v, idx, h2, dtr


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
    R.regf(h2[:ns], dtr[:ns], x[idx[:ns]], v[idx[:ns]], LMAX, NNBMAX)
    dt = time.time() - t
    R.close()
    return {"value": ns * m.size / dt, "unit": "pair-interactions/s", "cores": threads,
            "kind": "reference",
            "sample": f"src/Main/reg.avx.cpp (g++ -O3 -fopenmp -mavx2 -mfma), {ns} of the "
                      f"{idx.size} i-particles of one step against all {m.size} j, "
                      f"{dt:.1f} s, OMP_NUM_THREADS={threads}"}


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

    # batched: send + one graph-replayed regular step per ensemble step
    L.nb6_ens_open(nc, ntot)
    L.nb6_ens_items.restype = ctypes.c_int
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

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
    out_b = do.to_host(np.float64, (7, ntot))
    lst_b = dl.to_host(np.int32, (ntot, lmax))
    L.nb6_ens_close()

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
    out_l = do.to_host(np.float64, (7, ntot))
    lst_l = dl.to_host(np.int32, (ntot, lmax))
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
                        "work_items": L.nb6_ens_items()},
            "looped": {"clusters_per_s": nc / t_loop, "ms_per_step": 1e3 * t_loop,
                       "pairs_per_s": pairs / t_loop},
            "speedup": t_loop / t_ens,
            "batched_vs_looped": {"acc_max_rel": float(da.max()), "pot_max_rel": float(dp.max()),
                                  "lists_identical": same_lists}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="regf", choices=["regf", "ensemble"],
                    help="regf: the headline N=1M step; ensemble: 256-cluster batch vs loop")
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

    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args.steps, args.warmup)), flush=True)
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
        return L.nb6_kernel_ms(b"regf_tile"), L.nb6_kernel_ms(b"regf_finalize")

    for _ in range(args.warmup):
        step()
    L.nb6_sync()
    tile_ms, fin_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, b = step()
        tile_ms.append(a)
        fin_ms.append(b)
    L.nb6_sync()
    wall = time.perf_counter() - t0
    chunks = L.nb6_last_regf_chunks()

    out = do.to_host(np.float64, (7, ni))
    lst = dl.to_host(np.int32, (ni, LMAX))
    acc = accuracy(m, x, v, idx, h2, dtr, out, lst)
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
            "finalize_avg_ms": round(float(np.mean(fin_ms)), 3), "j_chunks": chunks}
    cpu = None if args.no_cpu else cpu_baseline(m, x, v, idx, h2, dtr)
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
           "roofline": roof, "cpu_baseline": cpu, "accuracy_vs_fp64": acc}
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
