"""CPU restatement of the regular corrector, regcor_gpu.F with fpcorr.F (TEST
INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU corrector).

One call handles one block of active particles after a regular force evaluation
(util_gpu.F:101-110 and :150-175): for each particle I

  * the regf list row is converted as util_gpu.F:101-110 does (0-based j-set index +
    IFIRST, self dropped);
  * no new neighbours (regcor_gpu.F:95-127): a small mass BODYM/100 at the origin
    gives DFIRR, DFD and RS is grown or shrunk;
  * otherwise RS follows the volume-ratio rule (regcor_gpu.F:139-238);
  * old and new lists are compared (losses, gains; regcor_gpu.F:252-296), lost
    neighbours with STEP <= SMIN inside 2 RS are kept (regcor_gpu.F:298-350) and the
    force of every lost (gained) neighbour is moved from the irregular to the regular
    side (regcor_gpu.F:355-415);
  * the Hermite corrector for the regular force (regcor_gpu.F:417-530), the optional
    derivative corrections of fpcorr.F (KZ(38) = 1, 2, 3; regcor_gpu.F:600-612), the
    list copy (:614-622) and the new regular step (:624-740).

Arithmetic kinds follow gfortran without -fdefault-real-8: literals such as 0.1, 0.6,
1.05, 1.1, 0.003 and every FLOAT() expression are REAL*4 (numpy float32 here, which
rounds every operation to single), everything else REAL*8 (Python floats, one IEEE
rounding per operation in Fortran's left-to-right order).  ONE3 = 1/3 and ONE12 = 1/12
in double (setup.F:486, zero.f:116).

CHECKL (regcor_gpu.F:240-244) is restated for KZ(37) = 1, the setting of the
reference's own example input: the high-velocity intruder test of checkl.F:243-340
(at most one body of LISTV joins the list per particle, with its force moved off the
regular side; a c.m. particle takes the resolved interactions of checkl.F:282-336
inside the c.m. approximation).  KZ(37) > 1 (close encounters and disrupted pairs,
checkl.F:17-242) is counted as unsupported.  A KS component J < IFIRST in FPCORR
uses its c.m. body's F and FDOT (fpcorr.F:46-61).

Not restated: tides other than the TIDAL(4) rotation of FIDOT (XTRNLF, ETIDE), cloud
forces (FCLOUD), and the print-only diagnostic block (regcor_gpu.F:700-722).  NB_FLAG
and M_FLAG are inputs (regcor_gpu.F sets them from KZ(39) and leaves them unset
otherwise).

Order within a block: the reference corrects the particles one after another, so the
retention test of particle I reads STEP(J) already divided by 8 when J precedes I in
the block and J's previous list was empty (regcor_gpu.F:736-739).  Here every
retention test reads the STEP values of the start of the block (the GPU corrects all
particles of a block at once); the two agree unless such a J is also a lost neighbour
of I with SMIN < STEP(J) <= 8 SMIN.

State layout (all numpy, index = particle number - 1): x, xdot, x0, x0dot, f, fdot, fi,
fidot, fr, frdot, d2, d3, xn, xndot, d0, d0r, d1r, d2r, d3r are [3][nmax]; body, t0,
step, timenw, t0r, stepr, rs are [nmax]; list is int32 [nmax][lmax] (count, then
1-based ascending indices); h is [npairs] (KS binding energies, used for I > N);
with KZ(37) = 1 also listv (int, [0] = count, then 1-based LISTV) and r [npairs] (KS
separations R, for c.m. particles), and P["cmsep2"].
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32
ONE3 = 1.0 / 3.0
ONE12 = 1.0 / 12.0


def d4(v):
    """A REAL*4 literal or value taken to REAL*8 (exact widening)."""
    return float(f32(v))


def _div(a, b):
    """IEEE division (inf / NaN instead of Python's exception, as in the Fortran)."""
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def _pair(xi, vi, xj, vj, mj):
    """Force and first derivative of body J on I, regcor_gpu.F:334-353 / :364-378."""
    a = [xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]]
    dv = [vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]]
    rij2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2]
    dr2i = 1.0 / rij2
    dr3i = mj * dr2i * math.sqrt(dr2i)
    drdv = a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2]
    drdp = 3.0 * drdv * dr2i
    return ([a[k] * dr3i for k in range(3)],
            [(dv[k] - a[k] * drdp) * dr3i for k in range(3)])


def merge_lists(old, new):
    """Losses (old not in new) and gains (new not in old), both ascending, as the
    two-pointer walk of regcor_gpu.F:262-296 produces them (strictly ascending input)."""
    losses, gains = [], []
    a = b = 0
    while a < len(old) or b < len(new):
        if b >= len(new) or (a < len(old) and old[a] < new[b]):
            losses.append(old[a])
            a += 1
        elif a >= len(old) or new[b] < old[a]:
            gains.append(new[b])
            b += 1
        else:
            a += 1
            b += 1
    return losses, gains


def fpcorr_terms(S, P, i, xi, vi, jlist, nloss, C):
    """fpcorr.F:11-98: summed F2DOT, F3DOT over the lost (sign -) then gained bodies."""
    save2 = [0.0, 0.0, 0.0]
    save3 = [0.0, 0.0, 0.0]
    for L, j1 in enumerate(jlist):
        j = j1 - 1
        A = [0.0] * 12
        if j1 < P["ifirst"]:          # KS component: its c.m. values (fpcorr.F:46-61)
            jcm = P["n"] + (j1 + 1) // 2 - 1          # N + KVEC(J), KVEC(J) = (J + 1) / 2
            for k in range(3):
                A[k] = S["x"][k][j] - xi[k]
                A[k + 3] = S["xdot"][k][j] - vi[k]
                A[k + 6] = 2.0 * (S["f"][k][jcm] - S["f"][k][i])
                A[k + 9] = 6.0 * (S["fdot"][k][jcm] - S["fdot"][k][i])
            C["ncmder"] += 1
        else:
            s = P["time"] - S["t0"][j]
            s3 = 3.0 * s
            for k in range(3):
                fj, fdj = S["f"][k][j], S["fdot"][k][j]
                A[k] = S["x"][k][j] - xi[k]
                A[k + 3] = (fdj * s3 + 2.0 * fj) * s + S["x0dot"][k][j] - vi[k]
                A[k + 6] = 2.0 * (fdj * s3 + fj - S["f"][k][i])
                A[k + 9] = 6.0 * (fdj - S["fdot"][k][i])
        a13 = 1.0 / (A[0] * A[0] + A[1] * A[1] + A[2] * A[2])
        a14 = S["body"][j] * a13 * math.sqrt(a13)
        a15 = (A[0] * A[3] + A[1] * A[4] + A[2] * A[5]) * a13
        a16 = a15 * a15
        a17 = 3.0 * a15
        a18 = 6.0 * a15
        a19 = 9.0 * a15
        a20 = (A[3] * A[3] + A[4] * A[4] + A[5] * A[5] + A[0] * A[6] + A[1] * A[7]
               + A[2] * A[8]) * a13 + a16
        a21 = 9.0 * a20
        a20 = 3.0 * a20
        a22 = ((9.0 * (A[3] * A[6] + A[4] * A[7] + A[5] * A[8])
                + 3.0 * (A[0] * A[9] + A[1] * A[10] + A[2] * A[11])) * a13
               + a17 * (a20 - 4.0 * a16))
        f2 = [0.0] * 3
        f3 = [0.0] * 3
        for k in range(3):
            f1 = A[k + 3] - a17 * A[k]
            f2[k] = (A[k + 6] - a18 * f1 - a20 * A[k]) * a14
            f3[k] = (A[k + 9] - a21 * f1 - a22 * A[k]) * a14 - a19 * f2[k]
        if L < nloss:
            f2 = [-v for v in f2]
            f3 = [-v for v in f3]
        for k in range(3):
            save2[k] = save2[k] + f2[k]
            save3[k] = save3[k] + f3[k]
    return save2, save3


def _pair_mw(xk, vk, xl, vl, mk, ml, mi):
    """checkl.F:300-329: mass-weighted term of KS component K on component L of I."""
    a = [xk[0] - xl[0], xk[1] - xl[1], xk[2] - xl[2]]
    rij2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2]
    dv = [vk[0] - vl[0], vk[1] - vl[1], vk[2] - vl[2]]
    dr2i = 1.0 / rij2
    dr3i = mk * ml * dr2i * math.sqrt(dr2i) / mi
    drdv = 3.0 * (a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2]) * dr2i
    return a, dv, dr3i, drdv


def _checkl_fast(S, P, I, xi, vi, rs2, freg, fdr, nlist, nnb, C):
    """checkl.F:243-340 (KZ(37) = 1): the first body J of LISTV inside 2 RS, approaching,
    with impact parameter below RS and not yet a neighbour joins the list of I; its
    force and derivative leave FREG and FDR (the irregular side is done later by the
    gain loop).  Returns the new NNB; nlist is updated in place."""
    i = I - 1
    X, XD, B = S["x"], S["xdot"], S["body"]
    lv = S["listv"]
    for L in range(1, int(lv[0]) + 1):
        J = int(lv[L])
        if J == I:
            continue
        j = J - 1
        a1, a2, a3 = X[0][j] - xi[0], X[1][j] - xi[1], X[2][j] - xi[2]
        rij2 = a1 * a1 + a2 * a2 + a3 * a3
        if rij2 > 4.0 * rs2:
            continue
        a4, a5, a6 = XD[0][j] - XD[0][i], XD[1][j] - XD[1][i], XD[2][j] - XD[2][i]
        a7 = a1 * a4 + a2 * a5 + a3 * a6
        if a7 > 0.0:
            continue
        p2 = rij2 - a7 * a7 / (a4 * a4 + a5 * a5 + a6 * a6)
        if p2 > rs2:
            continue
        if J in nlist:
            continue
        C["nbfast"] += 1
        if nnb > P["nnbmax"]:
            return nnb
        resolved = False
        if I > P["n"]:
            d = [xi[0] - X[0][j], xi[1] - X[1][j], xi[2] - X[2][j]]
            rij2c = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            rp = float(S["r"][I - P["n"] - 1])
            resolved = not (rij2c > P["cmsep2"] * (rp * rp) and J <= P["n"])
        if not resolved:                              # checkl.F:191-215
            dx = [X[k][j] - xi[k] for k in range(3)]
            dv = [XD[k][j] - vi[k] for k in range(3)]
            dr2 = 0.0
            drdv = 0.0
            for k in range(3):
                dr2 = dr2 + dx[k] * dx[k]
                drdv = drdv + dx[k] * dv[k]
            dr2i = 1.0 / dr2
            dr3i = float(B[j]) * dr2i * math.sqrt(dr2i)
            drdv = 3.0 * drdv * dr2i
            for k in range(3):
                freg[k] = freg[k] - dx[k] * dr3i
                fdr[k] = fdr[k] - (dv[k] - dx[k] * drdv) * dr3i
        else:                                         # checkl.F:282-336
            i1 = 2 * (I - P["n"]) - 1
            js = [J]
            if J > P["n"]:
                rq = float(S["r"][J - P["n"] - 1])
                if rij2c < P["cmsep2"] * (rq * rq):
                    js = [2 * (J - P["n"]) - 1, 2 * (J - P["n"])]
            for K in js:
                for Lc in (i1, i1 + 1):
                    k0, l0 = K - 1, Lc - 1
                    a, dv, dr3i, drdv = _pair_mw([X[q][k0] for q in range(3)],
                                                 [XD[q][k0] for q in range(3)],
                                                 [X[q][l0] for q in range(3)],
                                                 [XD[q][l0] for q in range(3)],
                                                 float(B[k0]), float(B[l0]), float(B[i]))
                    for q in range(3):
                        freg[q] = freg[q] - a[q] * dr3i
                    for q in range(3):
                        fdr[q] = fdr[q] - (dv[q] - a[q] * drdv) * dr3i
        pos = 0
        while pos < len(nlist) and nlist[pos] < J:
            pos += 1
        nlist.insert(pos, J)
        return nnb + 1
    return nnb


def regcor_block(S, P, addr, force, nlist):
    """Corrects the block in place.  addr: 1-based particle numbers; force: [7][ni]
    regf output (acc, jerk, pot); nlist: int32 [ni][lmax] regf rows (count, 0-based
    j-set indices with self), converted in place.  Returns the counters
    {nbvoid, nbflux, nbsmin, nbcorr, unsupported, nbfast, ncmder}."""
    C = dict(nbvoid=0, nbflux=0, nbsmin=0, nbcorr=0, unsupported=0, nbfast=0, ncmder=0)
    uns = [0]
    step0 = np.array(S["step"], float)        # STEP(J) of the retention test: block start
    for k_i, I in enumerate(addr):
        _one(S, P, int(I), [float(force[q][k_i]) for q in range(6)], nlist[k_i], C, uns, step0)
    C["unsupported"] = uns[0]
    return C


def _one(S, P, I, fg, row, C, uns, step0):
    i = I - 1
    time = P["time"]
    # regf row -> 1-based list without self (util_gpu.F:101-110)
    cnt = int(row[0])
    if cnt < 0:
        uns[0] += 1
        return
    L1 = 0
    for l in range(1, cnt + 1):
        t = int(row[l]) + P["ifirst"]
        if t != I:
            L1 += 1
            row[L1] = t
    row[0] = L1
    nnb = L1
    nlist = [int(row[l]) for l in range(1, nnb + 1)]

    xi = [float(S["x"][k][i]) for k in range(3)]
    vi = [float(S["xdot"][k][i]) for k in range(3)]
    freg = fg[0:3]
    fdr = fg[3:6]
    old_row = S["list"][i]
    nnb0 = int(old_row[0])
    old = [int(old_row[l]) for l in range(1, nnb0 + 1)]
    dtr = time - float(S["t0r"][i])
    rd = P["rdens"]
    rscale = P["rscale"]
    if P["nb_flag"] == 1:
        e = [xi[0] - rd[0], xi[1] - rd[1], xi[2] - rd[2]]
        ri2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
        rh2 = rscale * rscale
    else:
        ri2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]
        rh2 = 9.0 * (rscale * rscale)
    rs = float(S["rs"][i])
    rs2 = rs * rs
    dfirr = [0.0, 0.0, 0.0]
    dfd = [0.0, 0.0, 0.0]

    if nnb == 0:                                             # regcor_gpu.F:95-127
        r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]
        fij = d4(0.01) * P["bodym"] / (r2 * math.sqrt(r2))
        rdot = 3.0 * (xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]) / r2
        for k in range(3):
            dfirr[k] = -(fij * xi[k])
            dfd[k] = -((vi[k] - rdot * xi[k]) * fij)
        if rdot > 0.0:
            rs = max(0.75 * rs, d4(0.1) * rscale)
        else:
            rs = d4(1.1) * rs
        C["nbvoid"] += 1
    else:                                                    # regcor_gpu.F:139-238
        nnbopt, nnbmax, znbmin = P["nnbopt"], P["nnbmax"], P["znbmin"]
        fnb = f32(nnb)
        fac = float(f32(1.0) + f32(f32(f32(0.1) * (f32(nnbopt) - fnb)) / fnb))
        a3 = float(fnb) * fac
        nbave = min(nnbopt + 150, max(nnbmax - 50, int(znbmin)))
        a3 = min(a3, float(f32(0.9) * f32(nbave)))
        nbp = int(a3)
        if ri2 > rh2 and P["nb_flag"] == 1:
            a3 = a3 * max(rscale / math.sqrt(ri2), d4(0.1))
        else:
            a3 = max(a3, znbmin)
        a4 = a3 / float(fnb)
        if (a3 - float(nnb0)) * (a3 - float(nnb)) < 0.0:
            a4 = math.sqrt(a4)
        ridot = (xi[0] - rd[0]) * vi[0] + (xi[1] - rd[1]) * vi[1] + (xi[2] - rd[2]) * vi[2]
        if ri2 > P["rc2"] and P["nb_flag"] == 1 and ri2 < 100.0 * rh2:
            a4 = a4 * (1.0 + ridot * dtr / ri2)
        if I > P["n"] and ri2 <= rh2:                        # c.m. body, :169-181
            a2 = 100.0 * float(S["body"][i]) / abs(float(S["h"][I - P["n"] - 1]))
            if a2 > rs:
                a3 = max(float(f32(1.0) - f32(fnb / f32(nbave))), 0.0)
                a4 = 1.0 + 0.5 * a3
        a5 = min(P["rsfac"] * dtr, 0.25)
        a4 = a4 - 1.0
        a4 = a5 if a4 > a5 else max(a4, -a5)
        a3 = ONE3 * a4
        a1 = 1.0 + a3 - a3 * a3
        if nnb < znbmin and nbp > znbmin and a1 < 1.0:
            a1 = d4(1.05)
        if abs(a1 - 1.0) > d4(0.003):
            if rs * rs > ri2 and rs * rs < 2.0 * ri2 and nnb > 5:
                a1 = math.sqrt(a1)
            rs = a1 * rs
        if nnb > nbave and ridot < 0.0:
            drs = float(f32(fnb - f32(nbave)) / f32(f32(3.0) * fnb))
            rs = rs * (1.0 - drs)
        if P["m_flag"] == 0 and nnb <= 3 and ri2 < 100.0 * rh2:   # :222-238
            a1 = 2.0 * rs
            for j1 in nlist:
                j = j1 - 1
                d = [xi[k] - float(S["x"][k][j]) for k in range(3)]
                rij = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
                rsdot = (d[0] * (vi[0] - float(S["xdot"][0][j]))
                         + d[1] * (vi[1] - float(S["xdot"][1][j]))
                         + d[2] * (vi[2] - float(S["xdot"][2][j]))) / rij
                a1 = min(a1, rij + rsdot * dtr)
            rs = max(a1, d4(1.1) * rs)
        # optional neighbour additions, regcor_gpu.F:240-244
        kz37 = P.get("kz37", 0)
        if kz37 == 1 and S.get("listv") is not None and int(S["listv"][0]) > 0:
            nnb = _checkl_fast(S, P, I, xi, vi, rs2, freg, fdr, nlist, nnb, C)
        elif kz37 > 1:
            uns[0] += 1
    S["rs"][i] = rs

    # losses and gains, regcor_gpu.F:252-350
    nbloss = nbgain = 0
    if nnb0 == 0:
        losses, gains = [], list(nlist)
        df2, fr2 = 1.0, 0.0
    else:
        df2 = fr2 = 0.0
        for k in range(3):
            e = freg[k] - float(S["fr"][k][i])
            df2 = df2 + e * e
            fr2 = fr2 + freg[k] * freg[k]
        losses, gains = merge_lists(old, nlist)
        step = step0
        if any(float(step[j1 - 1]) < P["smin"] for j1 in losses):
            kk = 0
            while kk < len(losses):
                if nnb > P["nnbmax"] or I > P["n"]:
                    break
                j1 = losses[kk]
                j = j1 - 1
                if float(step[j]) > P["smin"] or j1 < P["ifirst"] or j1 > P["n"]:
                    kk += 1
                    continue
                e = [xi[k] - float(S["x"][k][j]) for k in range(3)]
                rij2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
                if rij2 > 4.0 * rs2:
                    kk += 1
                    continue
                pos = 0
                while pos < len(nlist) and nlist[pos] < j1:
                    pos += 1
                nlist.insert(pos, j1)
                nnb += 1
                C["nbsmin"] += 1
                fa, fj = _pair(xi, vi, [float(S["x"][k][j]) for k in range(3)],
                               [float(S["xdot"][k][j]) for k in range(3)], float(S["body"][j]))
                for k in range(3):
                    freg[k] = freg[k] - fa[k]
                    fdr[k] = fdr[k] - fj[k]
                del losses[kk]
    nbloss, nbgain = len(losses), len(gains)

    # force differences of the lost and gained neighbours, regcor_gpu.F:355-415
    for sign, js in ((-1.0, losses), (1.0, gains)):
        for j1 in js:
            j = j1 - 1
            fa, fj = _pair(xi, vi, [float(S["x"][k][j]) for k in range(3)],
                           [float(S["xdot"][k][j]) for k in range(3)], float(S["body"][j]))
            for k in range(3):
                if sign < 0:
                    dfirr[k] = dfirr[k] - fa[k]
                    dfd[k] = dfd[k] - fj[k]
                else:
                    dfirr[k] = dfirr[k] + fa[k]
                    dfd[k] = dfd[k] + fj[k]

    # corrector, regcor_gpu.F:417-530
    dtsq = dtr * dtr
    dt6 = _div(6.0, dtr * dtsq)
    dt2 = _div(2.0, dtsq)
    dtsq12 = ONE12 * dtsq
    dtr13 = ONE3 * dtr
    S["t0r"][i] = time
    if float(S["step"][i]) < 5.0 * P["dtmin"] and dtr > 50.0 * float(S["step"][i]):
        dtr13 = 0.0
        dtsq12 = 0.0
    tid4 = P["tidal4"]
    for k in range(3):
        dfr = float(S["fr"][k][i]) - freg[k] - dfirr[k]
        fdr0 = fdr[k] + dfd[k]
        frd = float(S["frdot"][k][i])
        sm = frd + fdr0
        at3 = 2.0 * dfr + dtr * sm
        bt2 = -3.0 * dfr - dtr * (sm + frd)
        S["xn"][k][i] = float(S["x0"][k][i]) + (d4(0.6) * at3 + bt2) * dtsq12
        S["xndot"][k][i] = float(S["x0dot"][k][i]) + (0.75 * at3 + bt2) * dtr13
        S["fi"][k][i] = float(S["fi"][k][i]) + dfirr[k]
        S["fr"][k][i] = freg[k]
        S["fidot"][k][i] = float(S["fidot"][k][i]) + dfd[k]
        S["frdot"][k][i] = fdr[k]
        if k == 1:
            S["fidot"][0][i] = float(S["fidot"][0][i]) - tid4 * dfr
        if k == 0:
            S["fidot"][1][i] = float(S["fidot"][1][i]) + tid4 * dfr
        S["d0"][k][i] = S["fi"][k][i]
        S["d0r"][k][i] = freg[k]
        S["d1r"][k][i] = fdr[k]
        S["d3r"][k][i] = at3 * dt6
        S["d2r"][k][i] = (3.0 * at3 + bt2) * dt2

    # derivative corrections, regcor_gpu.F:600-612 and fpcorr.F
    kz38 = P["kz38"]
    call = False
    if kz38 == 1:
        call = True
    elif kz38 == 2 and df2 > d4(0.0001) * fr2:
        call = True
    elif kz38 == 3 and nbloss > 0:
        nbgain = 0
        gains = []
        call = True
    if call and nbloss + nbgain > 0:
        s2, s3 = fpcorr_terms(S, P, i, xi, vi, losses + gains, nbloss, C)
        C["nbcorr"] += 1
        for k in range(3):
            S["d2"][k][i] = float(S["d2"][k][i]) + s2[k]
            S["d3"][k][i] = float(S["d3"][k][i]) + s3[k]
            S["d2r"][k][i] = float(S["d2r"][k][i]) - s2[k]
            S["d3r"][k][i] = float(S["d3r"][k][i]) - s3[k]

    # list copy, regcor_gpu.F:614-622
    if nbloss + nbgain > 0:
        old_row[0] = nnb
        old_row[1:nnb + 1] = nlist
        C["nbflux"] += nbloss + nbgain
    row[0] = nnb
    row[1:nnb + 1] = nlist

    # new regular step, regcor_gpu.F:624-740
    fr2 = freg[0] * freg[0] + freg[1] * freg[1] + freg[2] * freg[2]
    w1 = fdr[0] * fdr[0] + fdr[1] * fdr[1] + fdr[2] * fdr[2]
    d2r = [float(S["d2r"][k][i]) for k in range(3)]
    d3r = [float(S["d3r"][k][i]) for k in range(3)]
    w2 = d2r[0] * d2r[0] + d2r[1] * d2r[1] + d2r[2] * d2r[2]
    w3 = d3r[0] * d3r[0] + d3r[1] * d3r[1] + d3r[2] * d3r[2]
    w0 = _div(math.sqrt(fr2 * w2) + w1, math.sqrt(w1 * w3) + w2)
    w0 = P["etar"] * w0
    ttmp = math.sqrt(w0)
    dt0 = ttmp
    stepr = float(S["stepr"][i])
    smax = P["smax"]
    if ttmp > 2.0 * stepr:
        if float(np.fmod(time, 2.0 * dtr)) == 0.0:
            ttmp = min(2.0 * dtr, smax)
            while dt0 > 2.0 * ttmp and 2.0 * ttmp < smax and math.fmod(time, 2.0 * ttmp) == 0.0:
                ttmp = 2.0 * ttmp
        else:
            ttmp = stepr
    elif ttmp < dtr:
        ttmp = 0.5 * dtr
        if ttmp > dt0:
            ttmp = 0.5 * ttmp
    else:
        ttmp = stepr
    if ttmp > d4(0.1) and ttmp > dtr:
        dv2 = 0.0
        for k in range(3):
            e = vi[k] - float(S["xndot"][k][i])
            dv2 = dv2 + e * e
        q = _div(1.0e-4 * (dtr * dtr) * fr2, dv2)
        dtj = dtr * q ** d4(0.1)
        if dtj < ttmp:
            ttmp = dtr
    ttmp = min(ttmp, smax)
    S["stepr"][i] = ttmp
    if nnb0 == 0:
        S["step"][i] = 0.125 * float(S["step"][i])
        S["timenw"][i] = float(S["t0"][i]) + float(S["step"][i])
