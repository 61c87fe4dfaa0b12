"""CPU restatement of the neighbour-radius rules and of the RS network
(TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, never by the
product path).

Neighbour-radius rules of the reference, for a single star (I <= N), M_FLAG = 0:

* ``initial_rs``      fpoly0.F:53-58   RS = RS0 sqrt(1 + r^2), STEPR = min(SMAX sqrt(1 + r^2), SMAX)
* ``overflow_rs``     util_gpu.F:80-89 (same as fpoly0.F:135-145) on NNB < 0
* ``regcor_rs``       regcor_gpu.F:20-33 (RI2, RH2), :95-127 (NNB = 0),
                      :138-220 (volume-ratio update), :222-238 (NNB <= 3)
  Left out: the c.m.-body branch (:166-181, needs KS pairs, none exist here) and
  the optional CHECKL call (:241-245, KZ(37)).

RS network: 8 -> 64 -> 64 -> 1, ReLU, weights rounded to bfloat16 (round to
nearest even) for the forward pass, FP64 arithmetic here.  Parameter vector
layout (4801 floats): W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3.
Loss: mean over the batch of (y - t)^2.
"""
from __future__ import annotations

import numpy as np

NF, NH = 8, 64
NPARAM = NH * NF + NH + NH * NH + NH + NH + 1
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH


# ----------------------------------------------------------------- RS rules
def initial_rs(x, rs0, smax):
    """fpoly0.F:53-58."""
    ri2 = (np.asarray(x) ** 2).sum(-1)
    return rs0 * np.sqrt(1.0 + ri2), np.minimum(smax * np.sqrt(1.0 + ri2), smax)


def overflow_rs(rs, nnb, nnbopt, nnbmax, nb_flag=1):
    """New RS after an overflow (nnb < 0 holds -count), util_gpu.F:80-89."""
    rs = np.asarray(rs, float)
    nnb = np.asarray(nnb)
    if not nb_flag:
        return 0.9 * rs
    big = -nnb > nnbopt
    return np.where(big, (-float(nnbopt) / nnb) ** 0.333 * rs,
                    (float(nnbopt) / float(nnbmax)) ** 0.4 * rs)


def regcor_rs(rs, nnb, nnb0, xi, vi, dtr, *, rdens, rscale, rc2, rsfac, nnbopt, nnbmax,
              znbmin, nb_positions=None):
    """Neighbour radius after a regular step, regcor_gpu.F (NB_FLAG = 1, M_FLAG = 0).

    rs, nnb (new count without self), nnb0 (previous count), dtr: per particle;
    xi, vi: [n][3].  nb_positions(i) -> (xj [k][3], vj [k][3]) of i's neighbours,
    needed only for particles with nnb <= 3 (regcor_gpu.F:222-238).
    """
    rs = np.array(rs, float)
    out = rs.copy()
    xi = np.asarray(xi, float)
    vi = np.asarray(vi, float)
    dri = xi - np.asarray(rdens, float)
    ri2_all = (dri * dri).sum(1)
    rh2 = rscale * rscale
    nbave = min(nnbopt + 150, max(nnbmax - 50, int(znbmin)))
    for i in range(rs.size):
        n, n0, r = int(nnb[i]), int(nnb0[i]), rs[i]
        ri2 = ri2_all[i]
        if n == 0:                                   # :95-127
            r2 = float(xi[i] @ xi[i])
            rdot = 3.0 * float(xi[i] @ vi[i]) / r2
            out[i] = max(0.75 * r, 0.1 * rscale) if rdot > 0.0 else 1.1 * r
            continue
        fac = 1.0 + 0.1 * (nnbopt - n) / n          # :141
        a3 = n * fac
        a3 = min(a3, 0.9 * nbave)
        nbp = int(a3)
        if ri2 > rh2:
            a3 = a3 * max(rscale / np.sqrt(ri2), 0.1)
        else:
            a3 = max(a3, znbmin)
        a4 = a3 / n
        if (a3 - n0) * (a3 - n) < 0.0:              # :158
            a4 = np.sqrt(a4)
        ridot = float(dri[i] @ vi[i])
        if ri2 > rc2 and ri2 < 100.0 * rh2:          # :164-166
            a4 = a4 * (1.0 + ridot * dtr[i] / ri2)
        a5 = min(rsfac * dtr[i], 0.25)               # :184
        a4 = a4 - 1.0
        a4 = a5 if a4 > a5 else max(a4, -a5)
        a3 = a4 / 3.0                                # :195-197
        a1 = 1.0 + a3 - a3 * a3
        if n < znbmin and nbp > znbmin and a1 < 1.0:
            a1 = 1.05
        if abs(a1 - 1.0) > 0.003:
            if r * r > ri2 and r * r < 2.0 * ri2 and n > 5:
                a1 = np.sqrt(a1)
            r = a1 * r
        if n > nbave and ridot < 0.0:               # :211-216
            r = r * (1.0 - (n - nbave) / (3.0 * n))
        if n <= 3 and ri2 < 100.0 * rh2 and nb_positions is not None:   # :222-238
            xj, vj = nb_positions(i)
            a1 = 2.0 * r
            for k in range(len(xj)):
                d = xi[i] - xj[k]
                rij = np.sqrt(d @ d)
                rsdot = float(d @ (vi[i] - vj[k])) / rij
                a1 = min(a1, rij + rsdot * dtr[i])
            r = max(a1, 1.1 * r)
        out[i] = r
    return out


def kth_radius(xi, vi, dtr, xj, vj, cand, k):
    """Radius that gives exactly k neighbours under min(r^2, |dx + dtr dv|^2):
    the k-th smallest such distance over the candidate j indices (0-based)."""
    out = np.empty(len(xi))
    for i in range(len(xi)):
        c = np.asarray(cand[i])
        dx = xj[c] - xi[i]
        dv = vj[c] - vi[i]
        r2 = (dx * dx).sum(1)
        e = dx + dtr[i] * dv
        d2 = np.minimum(r2, (e * e).sum(1))
        out[i] = np.sqrt(np.sort(d2)[k - 1])
    return out


# ----------------------------------------------------------------- network
def bf16(a):
    """Round FP32 values to bfloat16 (nearest, ties to even); returns float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NF), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def normalise(feat, mu, sigma):
    """feat [n][8] -> (feat - mu) / sigma in FP32, as the device does."""
    f = np.asarray(feat, np.float32)
    return ((f - np.asarray(mu, np.float32)) / np.asarray(sigma, np.float32)).astype(np.float32)


def forward(p, fn, round_weights=True):
    """Outputs y [n] for normalised features fn [n][8]."""
    q = bf16(p).astype(np.float64) if round_weights else np.asarray(p, np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(q)
    h1 = np.maximum(fn.astype(np.float64) @ w1.T + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss_and_grad(p, fn, t, round_weights=True):
    """Mean squared error and its gradient with respect to the (rounded) weights."""
    q = bf16(p).astype(np.float64) if round_weights else np.asarray(p, np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(q)
    x = fn.astype(np.float64)
    z1 = x @ w1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    n = len(t)
    r = y - np.asarray(t, np.float64)
    loss = float((r * r).mean())
    dy = 2.0 * r / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = dy @ h2
    g[OFF_B3] = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ w2) * (z1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g


def init_params(seed):
    """He-uniform initial weights (seeded), zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, NH * NF) * np.sqrt(6.0 / NF)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, NH * NH) * np.sqrt(6.0 / NH)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, NH) * np.sqrt(6.0 / NH) * 0.1
    return p
