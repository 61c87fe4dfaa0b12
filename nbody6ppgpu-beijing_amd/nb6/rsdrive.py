"""Regular block steps on the library: a host-side mirror of the caller flow
around gpunb_regf that the neighbour-radius policies are measured in.

Follows the reference's regular cycle:
  * start-up (fpoly0.F:53-58, :110-181): RS = RS0 sqrt(1 + r^2), STEPR =
    min(SMAX sqrt(1 + r^2), SMAX); lists of every particle with block re-runs
    while any count is < 2 (RS x 1.5) or overflows (util_gpu.F:80-89 rescale);
  * each block step (intgrt.F:875-896): predict every particle to the block time
    (xbpredall.F:55-59 polynomial), send the j-set, call gpunb_regf for the
    active block with h2 = RS^2 and dtr = STEPR, and on an overflow rescale that
    particle's RS and re-run the chunk immediately, one extra call per overflow
    (util_gpu.F:65-100);
  * the RS for the particle's next regular step comes from the reference rule
    (regcor_gpu.F:95-238, on the device) or from the RS network (rsnet.hip) with
    the rule as fallback.  The network aims at the count the rule aims at: NNBOPT,
    or 0.9 NBAVE where that is smaller (short lists, see `net_aim`).

The trajectory is integrated with individual power-of-two regular steps and a
fourth-order Hermite predictor-corrector on the full force (a gpunb_regf call
with h2 = 0), i.e. without the irregular/regular split: the neighbour radius
and lists follow the regular-step cadence, which is what the policies are
judged on.  All force, list, rule, network and target work runs on the GPU
through the library; the host only does the O(N) bookkeeping.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import lib

KTH_LMAX = 2049          # list width of the target calls (the largest the library sorts)


class Dev:
    """Device scratch of one simulation (sized for all N particles)."""

    def __init__(self, n: int, lmax: int):
        D = lib.DeviceArray
        self.j = D(7 * n * 8)          # x y z vx vy vz m, SoA [7][n]
        self.i = D(8 * n * 8)          # x y z vx vy vz h2 dtr, SoA [8][n]
        self.out = D(7 * n * 8)
        self.lst = D(n * lmax * 4)
        self.tmp = D(n * lmax * 4)
        self.big = D(n * KTH_LMAX * 4)
        self.blk = D(6 * n * 8)        # block x, v for the rule
        self.kth = D(n * 8)
        self.state = D(10 * n * 8)     # x[3] a[3] j[3] mass for the features
        self.rs = D(n * 8)
        self.dtr = D(n * 8)
        self.nnb = D(n * 4)
        self.nnb0 = D(n * 4)
        self.feat = D(8 * n * 4)
        self.xj = D(3 * n * 8)
        self.vj = D(3 * n * 8)

    def free(self):
        for v in self.__dict__.values():
            v.free()


class Sim:
    """One N-body system advanced by regular block steps."""

    def __init__(self, m, x, v, *, nnbopt=100, lmax=600, smax=2.0 ** -5, dtmin=2.0 ** -12,
                 eta=0.05, maxthr=1024, rs0=None):
        self.L = lib.load()
        self.n = n = m.size
        self.m, self.x, self.v = (np.array(a, np.float64) for a in (m, x, v))
        self.nnbopt, self.lmax = nnbopt, lmax
        self.nnbmax = min(n // 2, lmax - 50)                  # input.F:132
        self.znbmin = max(0.01 * self.nnbmax, 1.0)            # input.F:133
        self.aim = net_aim(nnbopt, self.nnbmax, self.znbmin)
        self.smax, self.dtmin, self.eta, self.maxthr = smax, dtmin, eta, maxthr
        self.g = lib.Gpunb(n)
        self.d = Dev(n, lmax)
        self.t = 0.0
        self.t0 = np.zeros(n)
        self.stats = dict(calls=0, overflow_events=0, extra_pairs=0, steps=0, samples=0,
                          miss30=0, absdev=0.0, fallback=0,
                          # split by the particle's regular step: its first after start-up
                          # (radius from start-up) or a later one (radius from the policy)
                          overflow_first=0, overflow_later=0, samples_later=0,
                          miss30_later=0, absdev_later=0.0)
        self.nreg = np.zeros(n, np.int64)    # regular steps taken by each particle
        r = np.linalg.norm(self.x, axis=1)
        if rs0 is None:
            # radius holding NNBOPT particles at the mean density inside the half-mass radius
            rh = np.median(r)
            rs0 = (3.0 * nnbopt / (4.0 * math.pi * (n / 2) / (4.0 / 3.0 * math.pi * rh ** 3))) ** (1 / 3)
        self.rs = rs0 * np.sqrt(1.0 + r * r)                      # fpoly0.F:53-55
        self.dt = np.minimum(smax * np.sqrt(1.0 + r * r), smax)   # fpoly0.F:57-58
        self.dt = self._pow2(self.dt)
        self.nnb = np.zeros(n, np.int32)
        self.feat_prev = np.zeros((n, 8), np.float32)
        self.rs_used_prev = np.ones(n)
        self.has_prev = np.zeros(n, bool)
        self.used_net = np.zeros(n, bool)
        self.hook = None

    # ------------------------------------------------------------ device helpers
    def close(self):
        self.d.free()
        self.g.close()

    def _step(self, a, j, cap, t):
        """eta |a| / |adot| as a power of two in [dtmin, min(SMAX, cap)] that divides t."""
        an = np.linalg.norm(a, axis=1)
        jn = np.linalg.norm(j, axis=1) + 1e-300
        d = np.minimum(self._pow2(self.eta * an / jn), cap)
        while True:
            bad = np.fmod(t, d) != 0.0
            if not bad.any():
                return d
            d[bad] *= 0.5

    def _pow2(self, dt):
        dt = np.clip(dt, self.dtmin, self.smax)
        return 2.0 ** np.floor(np.log2(dt))

    def _send(self, xp, vp):
        self._jx, self._jv = xp, vp
        soa = np.ascontiguousarray(np.concatenate([xp.T, vp.T, self.m[None]]))
        self.L.nb6_memcpy_htod(self.d.j.ptr, soa.ctypes.data, soa.nbytes)
        if self.L.nb6_send_device(self.n, self.n, self.d.j.ptr, None) != 0:
            lib.check_error()

    def _regf(self, xi, vi, h2, dtr, lst, lmax, nnbmax):
        """One gpunb_regf call on the device; returns (out [7][ni], counts [ni])."""
        ni = h2.size
        ib = np.ascontiguousarray(np.concatenate([xi.T, vi.T, h2[None], dtr[None]]))
        self.L.nb6_memcpy_htod(self.d.i.ptr, ib.ctypes.data, ib.nbytes)
        if self.L.nb6_regf_device(ni, ni, self.d.i.ptr, lmax, nnbmax, 0, self.d.out.ptr, lst.ptr,
                                  None) != 0:
            lib.check_error()
        self.stats["calls"] += 1
        out = self.d.out.to_host(np.float64, (7, ni))
        rows = np.empty((ni, lmax), np.int32)
        self.L.nb6_memcpy_dtoh(rows.ctypes.data, lst.ptr, rows.nbytes)
        return out, rows

    def predict(self, t):
        s = (t - self.t0)[:, None]
        xp = self.x + s * (self.v + s * (0.5 * self.a + s * self.j / 6.0))
        vp = self.v + s * (self.a + 0.5 * s * self.j)
        return xp, vp

    # ------------------------------------------------------------ start-up
    def start(self, policy="rule", net=None):
        """fpoly0.F: full force, then lists of everybody with block re-runs.  The
        start-up lists give every particle its first features; with policy "net"
        the network also sets the radius of the first regular step."""
        self._send(self.x, self.v)
        out, _ = self._regf(self.x, self.v, np.zeros(self.n), np.zeros(self.n), self.d.tmp, 2, 1)
        self.a, self.j = out[0:3].T.copy(), out[3:6].T.copy()
        self.pot = out[6].copy()
        # first regular steps from the force criterion, at most the fpoly0.F estimate
        self.dt = np.minimum(self.dt, self._step(self.a, self.j, np.inf, 0.0))
        reruns = 0
        for b in range(0, self.n, self.maxthr):
            idx = np.arange(b, min(self.n, b + self.maxthr))
            while True:
                _, rows = self._regf(self.x[idx], self.v[idx], self.rs[idx] ** 2, self.dt[idx],
                                     self.d.lst, self.lmax, self.nnbmax)
                c = rows[:, 0]
                bad = c < 2
                if not bad.any():
                    break
                reruns += 1
                ovf = c < 0
                self.rs[idx[ovf]] = _overflow(self.rs[idx[ovf]], c[ovf], self.nnbopt, self.nnbmax)
                zero = bad & ~ovf
                self.rs[idx[zero]] *= 1.5
            self.nnb[idx] = c - 1                                  # without self
        self.stats["startup_reruns"] = reruns
        self._globals()
        self.t = 0.0
        allp = np.arange(self.n)
        feats = np.empty((self.n, 8), np.float32)
        for b in range(0, self.n, 65536):
            sl = allp[b:b + 65536]
            feats[sl] = self.features(sl, self.x[sl], self.a[sl], self.j[sl], self.rs[sl],
                                      self.dt[sl], self.nnb[sl])
        self.feat_prev[:] = feats
        self.rs_used_prev[:] = self.rs
        self.has_prev[:] = True
        if policy == "net":
            self.rs = net.predict(feats, self.rs * self.aim, self.rs, np.zeros(self.n, np.int32))
            self.used_net[:] = True
        self.stats["calls"] = 0

    def _globals(self):
        """Density centre, core radius, RSCALE, RSFAC from the current lists."""
        rho = (self.nnb + 1.0) / (4.0 / 3.0 * math.pi * self.rs ** 3)
        w = rho / rho.sum()
        self.rdens = (w[:, None] * self.x).sum(0)
        d2 = ((self.x - self.rdens) ** 2).sum(1)
        self.rc2 = float((rho * rho * d2).sum() / (rho * rho).sum())
        wpot = 0.5 * (self.m * self.pot).sum()
        mt = self.m.sum()
        self.rscale = 0.5 * mt * mt / wpot                             # adjust.F:136
        ekin = 0.5 * (self.m[:, None] * self.v ** 2).sum()
        etot = ekin - wpot
        tcr = mt ** 2.5 / (2.0 * abs(etot)) ** 1.5
        vc = math.sqrt(2.0 * ekin / mt)
        self.rsfac = max(25.0 / tcr, 3.0 * vc / (2.0 * self.rs.min()))  # adjust.F:199
        self.glob = np.array([*self.rdens, self.rscale, self.rc2, self.rsfac, self.znbmin])

    # ------------------------------------------------------------ one block step
    def _rs_call(self, xi, vi, rs, dtr):
        """The block's regular call with its current RS.  util_gpu.F:65-100 re-runs
        the whole chunk (maxthr particles) once per overflowing particle after
        rescaling its RS; gpunb_regf is independent per i, so the same counts and
        lists come from one call for the block plus calls on the overflowed subset,
        while the re-runs of the reference are counted (events, pairs)."""
        ni = rs.size
        out, rows = self._regf(xi, vi, rs ** 2, dtr, self.d.lst, self.lmax, self.nnbmax)
        if self.hook is not None:   # e.g. a check against an FP64 reference at every call
            self.hook(self._jx, self._jv, xi, vi, rs ** 2, dtr, out, rows)
        cnt = rows[:, 0].copy()
        todo = np.nonzero(cnt < 0)[0]
        while todo.size:
            chunk = np.minimum(self.maxthr, ni - (todo // self.maxthr) * self.maxthr)
            self.stats["overflow_events"] += int(todo.size)
            first = getattr(self, "_first", None)
            if first is not None and first.size == ni:
                nf = int(first[todo].sum())
                self.stats["overflow_first"] += nf
                self.stats["overflow_later"] += int(todo.size) - nf
            self.stats["extra_pairs"] += int(chunk.sum()) * self.n
            rs[todo] = _overflow(rs[todo], cnt[todo], self.nnbopt, self.nnbmax)
            _, sub = self._regf(xi[todo], vi[todo], rs[todo] ** 2, dtr[todo], self.d.tmp,
                                self.lmax, self.nnbmax)
            for q, r in enumerate(todo):          # the block's rows stay on the device
                self.L.nb6_memcpy_htod(self.d.lst.ptr + 4 * self.lmax * int(r), sub[q].ctypes.data,
                                       4 * self.lmax)
            cnt[todo] = sub[:, 0]
            todo = todo[sub[:, 0] < 0]
        return cnt - 1                              # without self

    def block_step(self, policy="rule", net=None, record=None):
        """Advance the next block of particles by their regular step.

        policy: "rule" (regcor_gpu.F) or "net" (RS network; the rule where the
        network's previous count missed NNBOPT by more than 30 %).
        record: list that receives (features at the previous step, target) pairs."""
        L, d, n = self.L, self.d, self.n
        t = float((self.t0 + self.dt).min())
        act = np.nonzero(self.t0 + self.dt == t)[0]
        ni = act.size
        xp, vp = self.predict(t)
        self._send(xp, vp)
        xi, vi, dtr = xp[act], vp[act], self.dt[act]

        rs = self.rs[act].copy()
        self._first = self.nreg[act] == 0
        nnb_new = self._rs_call(xi, vi, rs, dtr)
        dev = np.abs(nnb_new - self.nnbopt) / self.nnbopt
        lad = np.abs(np.log((nnb_new + 1.0) / (self.nnbopt + 1.0)))
        self.stats["miss30"] += int((dev > 0.3).sum())
        self.stats["absdev"] += float(lad.sum())
        self.stats["samples"] += ni
        later = ~self._first
        self.stats["miss30_later"] += int((dev[later] > 0.3).sum())
        self.stats["absdev_later"] += float(lad[later].sum())
        self.stats["samples_later"] += int(later.sum())
        self.nreg[act] += 1
        if len(self.stats.setdefault("overflow_by_step", [])) < 8:
            self.stats["overflow_by_step"].append(self.stats["overflow_events"])

        # the rule's inputs: block x, v, dtr, counts, the lists just built and every
        # particle's predicted position and velocity
        nnb0 = self.nnb[act].astype(np.int32)
        blk = np.ascontiguousarray(np.concatenate([xi.T, vi.T]))
        L.nb6_memcpy_htod(d.blk.ptr, blk.ctypes.data, blk.nbytes)
        L.nb6_memcpy_htod(d.xj.ptr, np.ascontiguousarray(xp.T).ctypes.data, 3 * n * 8)
        L.nb6_memcpy_htod(d.vj.ptr, np.ascontiguousarray(vp.T).ctypes.data, 3 * n * 8)
        L.nb6_memcpy_htod(d.dtr.ptr, np.ascontiguousarray(dtr).ctypes.data, ni * 8)
        L.nb6_memcpy_htod(d.nnb.ptr, np.ascontiguousarray(nnb_new, np.int32).ctypes.data, ni * 4)
        L.nb6_memcpy_htod(d.nnb0.ptr, np.ascontiguousarray(nnb0).ctypes.data, ni * 4)
        L.nb6_memcpy_htod(d.rs.ptr, np.ascontiguousarray(rs).ctypes.data, ni * 8)
        if L.nb6_rs_rule_device(ni, ni, d.blk.ptr, d.blk.ptr + 3 * ni * 8, d.dtr.ptr, d.nnb.ptr,
                                d.nnb0.ptr, d.lst.ptr, self.lmax, d.xj.ptr, d.vj.ptr, n,
                                _host_d(self.glob), self.nnbopt, self.nnbmax, d.rs.ptr, None) != 0:
            lib.check_error()
        rs_rule = d.rs.to_host(np.float64, ni)

        # training target at this state, paired with the features of the previous step
        if record is not None:
            tgt = self.target_radius(xi, vi, dtr, rs, nnb_new)
            ok = (tgt > 0) & self.has_prev[act]
            record.append((self.feat_prev[act][ok],
                           np.log(tgt[ok] / self.rs_used_prev[act][ok]).astype(np.float32)))

        # full force and Hermite correction of the block
        out, _ = self._regf(xi, vi, np.zeros(ni), dtr, d.tmp, 2, 1)
        a1, j1 = out[0:3].T, out[3:6].T
        a0, j0 = self.a[act], self.j[act]
        h = dtr[:, None]
        a2 = (-6.0 * (a0 - a1) - h * (4.0 * j0 + 2.0 * j1)) / h ** 2
        a3 = (12.0 * (a0 - a1) + 6.0 * h * (j0 + j1)) / h ** 3
        self.x[act] = xi + h ** 4 / 24.0 * a2 + h ** 5 / 120.0 * a3
        self.v[act] = vi + h ** 3 / 6.0 * a2 + h ** 4 / 24.0 * a3
        self.a[act], self.j[act] = a1, j1
        self.pot[act] = out[6]
        self.t0[act] = t
        dnew = self._step(a1, j1, 2.0 * dtr, t)

        # RS of the block's next regular step
        feats = self.features(act, xi, a1, j1, rs, dtr, nnb_new)
        if policy == "net":
            use_rule = (self.used_net[act] & (dev > 0.3)).astype(np.int32)
            self.stats["fallback"] += int(use_rule.sum())
            rs_next = net.predict(feats, rs * self.aim, rs_rule, use_rule)
            self.used_net[act] = use_rule == 0
        else:
            rs_next = rs_rule
        self.feat_prev[act] = feats
        self.rs_used_prev[act] = rs
        self.has_prev[act] = True
        self.rs[act] = rs_next
        self.nnb[act] = nnb_new
        self.dt[act] = dnew
        self.t = t
        self.stats["steps"] += 1
        return ni

    def features(self, act, xi, a1, j1, rs, dtr, nnb):
        """The eight network inputs of the block (device kernel)."""
        d, L = self.d, self.L
        ni = act.size
        st = np.ascontiguousarray(np.concatenate([xi.T, a1.T, j1.T, self.m[act][None]]))
        L.nb6_memcpy_htod(d.state.ptr, st.ctypes.data, st.nbytes)
        L.nb6_memcpy_htod(d.rs.ptr, np.ascontiguousarray(rs).ctypes.data, ni * 8)
        L.nb6_memcpy_htod(d.dtr.ptr, np.ascontiguousarray(dtr).ctypes.data, ni * 8)
        L.nb6_memcpy_htod(d.nnb.ptr, np.ascontiguousarray(nnb, np.int32).ctypes.data, ni * 4)
        base = d.state.ptr
        if L.nb6_rs_features_device(ni, ni, base, base + 3 * ni * 8, base + 6 * ni * 8,
                                    base + 9 * ni * 8, d.rs.ptr, d.dtr.ptr, d.nnb.ptr,
                                    _host_d(self.rdens), d.feat.ptr, None) != 0:
            lib.check_error()
        return d.feat.to_host(np.float32, (8, ni)).T.copy()

    def target_radius(self, xi, vi, dtr, rs, nnb):
        """Radius giving exactly NNBOPT neighbours besides self for each particle of
        the block: a gpunb_regf list with a generous radius, then the k-th distance
        (k = NNBOPT + 1, self included) on the device; radii are widened or narrowed
        for rows whose wide list held too few entries or overflowed."""
        k = self.nnbopt + 1
        ni = xi.shape[0]
        guess = rs * ((k + 0.5) / np.maximum(nnb + 1.0, 1.0)) ** (1.0 / 3.0) * 1.3
        out = np.full(ni, -1.0)
        todo = np.arange(ni)
        for _ in range(16):
            if todo.size == 0:
                break
            for b in range(0, todo.size, 65536):
                sub = todo[b:b + 65536]
                h2 = guess[sub] ** 2
                ib = np.ascontiguousarray(np.concatenate([xi[sub].T, vi[sub].T, h2[None],
                                                          dtr[sub][None]]))
                self.L.nb6_memcpy_htod(self.d.i.ptr, ib.ctypes.data, ib.nbytes)
                if self.L.nb6_regf_device(sub.size, sub.size, self.d.i.ptr, KTH_LMAX, KTH_LMAX - 1,
                                          0, self.d.out.ptr, self.d.big.ptr, None) != 0:
                    lib.check_error()
                if self.L.nb6_rs_kth_device(sub.size, sub.size, self.d.i.ptr, self.d.big.ptr,
                                            KTH_LMAX, k, self.d.kth.ptr, None) != 0:
                    lib.check_error()
                out[sub] = self.d.kth.to_host(np.float64, sub.size)
            bad = out[todo] < 0
            found = -out[todo][bad] - 1.0
            todo = todo[bad]
            few = found < k
            guess[todo[few]] *= np.maximum(((k + 1.0) / np.maximum(found[few], 1.0)) ** (1 / 3) * 1.2, 1.3)
            guess[todo[~few]] *= 0.7
        return out

    def run(self, nsteps, policy="rule", net=None, record=None):
        for _ in range(nsteps):
            self.block_step(policy, net, record)
        return self.stats

    def energy(self):
        """ZKIN - POT at the common time of the last full block, through the energy path
        of the reference's GPU build (energy.F:28-62: GPUPOT, PHICOR, sums), here
        nb6_energy_device with single bodies only."""
        xp, vp = self.predict(self.t)
        zkin, pot, _, _ = lib.energy(xp, vp, self.m, np.zeros((self.n, 1), np.int32), self.n, 0, 0.0)
        return zkin - pot


def net_aim(nnbopt, nnbmax, znbmin):
    """Radius factor that moves the network's aim from NNBOPT to the rule's own target.

    The rule caps the count it aims at to 0.9 NBAVE, NBAVE = min(NNBOPT + 150,
    max(NNBMAX - 50, ZNBMIN)) (regcor_gpu.F:144-145): below NNBOPT when the list is
    short (NNBMAX < NNBOPT / 0.9 + 50), e.g. 90 at NNBOPT = 100, LMAX = 200.  The
    network predicts the radius holding NNBOPT neighbours; at a locally uniform
    density the radius holding k is that times (k / NNBOPT)^(1/3)."""
    nbave = min(nnbopt + 150, max(nnbmax - 50, int(znbmin)))
    return (min(float(nnbopt), 0.9 * nbave) / nnbopt) ** (1.0 / 3.0)


def _overflow(rs, nnb, nnbopt, nnbmax):
    """util_gpu.F:80-89 (NB_FLAG = 1)."""
    nnb = np.asarray(nnb, float)
    return np.where(-nnb > nnbopt, (-float(nnbopt) / nnb) ** 0.333, (float(nnbopt) / nnbmax) ** 0.4) * rs


_keep = []


def _host_d(a):
    a = np.ascontiguousarray(a, np.float64)
    _keep.append(a)
    if len(_keep) > 64:
        del _keep[:32]
    return a.ctypes.data


class RSNet:
    """The RS network on the device: training and fused prediction."""

    def __init__(self, seed=20261206):
        self.L = lib.load()
        if self.L.nb6_rsnet_open(seed) != 0:
            lib.check_error()
        self.nparam = self.L.nb6_rsnet_nparam()

    def set_norm(self, mu, sigma):
        mu = np.ascontiguousarray(mu, np.float32)
        sigma = np.ascontiguousarray(sigma, np.float32)
        if self.L.nb6_rsnet_set_norm(mu.ctypes.data, sigma.ctypes.data) != 0:
            lib.check_error()

    def params(self):
        p = np.empty(self.nparam, np.float32)
        self.L.nb6_rsnet_get_params(p.ctypes.data)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if self.L.nb6_rsnet_set_params(p.ctypes.data) != 0:
            lib.check_error()

    def train(self, feat, tgt, epochs=4, batch=4096, lr=2e-3, seed=1, log=None):
        """Adam on minibatches; features [n][8], targets [n].  The samples go to the
        device once; each epoch is shuffled and run on the device
        (nb6_rsnet_train_epoch_device).  Returns the loss history."""
        n = tgt.size
        mu = feat.mean(0)
        sigma = feat.std(0) + 1e-6
        self.set_norm(mu, sigma)
        fd = lib.DeviceArray.from_host(np.ascontiguousarray(feat.T, np.float32))
        td = lib.DeviceArray.from_host(np.ascontiguousarray(tgt, np.float32))
        hist = []
        loss = ctypes.c_float()
        t0 = time.time()
        self.train_s = 0.0
        for ep in range(epochs):
            t1 = time.time()
            if self.L.nb6_rsnet_train_epoch_device(n, fd.ptr, td.ptr, batch, ctypes.c_float(lr),
                                                   seed * 1000003 + ep, ctypes.byref(loss),
                                                   None) != 0:
                lib.check_error()
            self.train_s += time.time() - t1
            hist.append(loss.value)
            if log:
                log(f"rsnet epoch {ep}: loss {hist[-1]:.4f} ({time.time() - t0:.1f} s)")
        fd.free()
        td.free()
        return hist

    def predict(self, feats, rs_prev, rs_rule, use_rule):
        """Fused prediction of RS for a block (rule value where use_rule)."""
        ni = rs_prev.size
        fb = lib.DeviceArray.from_host(np.ascontiguousarray(feats.T, np.float32))
        rp = lib.DeviceArray.from_host(np.ascontiguousarray(rs_prev, np.float64))
        rr = lib.DeviceArray.from_host(np.ascontiguousarray(rs_rule, np.float64))
        ur = lib.DeviceArray.from_host(np.ascontiguousarray(use_rule, np.int32))
        ro = lib.DeviceArray(ni * 8)
        if self.L.nb6_rs_predict_device(ni, fb.ptr, rp.ptr, rr.ptr, ur.ptr, ro.ptr, None, ni, 0,
                                        ctypes.c_double(1.0), None) != 0:
            lib.check_error()
        out = ro.to_host(np.float64, ni)
        for a in (fb, rp, rr, ur, ro):
            a.free()
        return out

    def close(self):
        self.L.nb6_rsnet_close()


def input_setup(src, dat10=None):
    """Masses, positions, velocities and driver settings from a NAMELIST parameter file
    (input.F:18-60; path or text).  Initial conditions: unit 10 in NBODY format for
    KZ(22) = 2, 6, 9, 10 or TREE format for KZ(22) = 3, 7 (data.F:36-66, `dat10` path or
    text); otherwise an equal-mass Plummer sphere from this library's generator seeded
    with NRAND (the reference's own generators are not restated).  The regular-step
    parameter is sqrt(ETAR): regcor_gpu.F:641-643 takes the square root of ETAR times a
    squared time scale.  LMAX follows configure.ac (400 up to N = 16k, 500 up to 64k,
    600 above)."""
    from pathlib import Path

    from . import ics, io
    text = Path(src).read_text() if isinstance(src, (str, Path)) and Path(str(src)).exists() \
        else str(src)
    p = io.run_parameters(io.read_namelist(text))
    n = p["n"]
    if p["kz22"] in (2, 6, 9, 10):
        if dat10 is None:
            raise ValueError(f"KZ(22) = {p['kz22']} reads the initial conditions from unit 10")
        m, x, v = io.read_dat10(dat10, n)
    elif p["kz22"] in (3, 7):
        if dat10 is None:
            raise ValueError(f"KZ(22) = {p['kz22']} reads the TREE format from unit 10")
        m, x, v = io.read_tree(dat10)
    else:
        m, x, v = ics.plummer(n, p["nrand"])
    lmax = 400 if n <= 16384 else 500 if n <= 65536 else 600
    kw = dict(nnbopt=p["nnbopt"], lmax=lmax, smax=p["smax"], dtmin=p["dtmin"],
              eta=math.sqrt(p["etar"]), rs0=p["rs0"])
    return m, x, v, kw, p


def sim_from_input(src, dat10=None, **override):
    """A Sim set up from a parameter file (and unit 10), see input_setup."""
    m, x, v, kw, _ = input_setup(src, dat10)
    kw.update(override)
    return Sim(m, x, v, **kw)
