// Regular corrector of a block on the GPU: REGCOR_GPU (regcor_gpu.F:1-740) with
// FPCORR (fpcorr.F:1-148), called by the reference once per particle after each
// regular force evaluation (util_gpu.F:150-175).
//
// One thread per active particle I.  Its work is O(old + new list length): the regf
// row is turned into the 1-based list without self (util_gpu.F:101-110), RS is
// updated (regcor_gpu.F:95-238), the old and new lists are walked in ascending
// order to find losses and gains (regcor_gpu.F:252-296), lost neighbours with
// STEP <= SMIN inside 2 RS are kept (regcor_gpu.F:298-350), the force of each lost
// (gained) neighbour moves between the irregular and regular parts
// (regcor_gpu.F:355-415), then the Hermite corrector (:417-530), FPCORR, the list
// copy and the new regular step (:600-740).
//
// Arithmetic follows the reference operation by operation: FP64 without FMA
// contraction in Fortran's left-to-right order, and the REAL*4 sub-expressions of the
// reference (FLOAT() of counts, single-precision literals such as 0.1, 0.6, 1.05,
// 1.1, 0.003) in FP32.  The only library function whose last bit may differ from the
// host's is pow() in the Makino test, which only enters a comparison.
#include "nb6_internal.h"
#include "../../include/nb6gpu.h"

#include <climits>
#include <cmath>

namespace nb6 {

namespace {

constexpr double ONE3 = 1.0 / 3.0;
constexpr double ONE12 = 1.0 / 12.0;

enum { C_VOID, C_FLUX, C_SMIN, C_CORR, C_UNS, NCOUNT };

// force and first derivative of body j (0-based) on i, regcor_gpu.F:334-353
__device__ __forceinline__ void pair_fj(const nb6_regstate &s, int j, const double *xi,
                                        const double *vi, double *fa, double *fj) {
#pragma clang fp contract(off)
    const size_t M = s.nmax;
    double a[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a[k] = s.x[k * M + j] - xi[k];
        dv[k] = s.xdot[k * M + j] - vi[k];
    }
    const double rij2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double dr2i = 1.0 / rij2;
    const double dr3i = s.body[j] * dr2i * sqrt(dr2i);
    const double drdv = a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2];
    const double drdp = 3.0 * drdv * dr2i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        fa[k] = a[k] * dr3i;
        fj[k] = (dv[k] - a[k] * drdp) * dr3i;
    }
}

// second and third derivative of the force of body j on i, fpcorr.F:24-74
__device__ __forceinline__ void fp_terms(const nb6_regstate &s, double time, int i, int j,
                                         const double *xi, const double *vi, double *f2,
                                         double *f3) {
#pragma clang fp contract(off)
    const size_t M = s.nmax;
    const double sj = time - s.t0[j];
    const double s3 = 3.0 * sj;
    double A[12];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double fj = s.f[k * M + j], fdj = s.fdot[k * M + j];
        A[k] = s.x[k * M + j] - xi[k];
        A[k + 3] = (fdj * s3 + 2.0 * fj) * sj + s.x0dot[k * M + j] - vi[k];
        A[k + 6] = 2.0 * (fdj * s3 + fj - s.f[k * M + i]);
        A[k + 9] = 6.0 * (fdj - s.fdot[k * M + i]);
    }
    const double a13 = 1.0 / (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
    const double a14 = s.body[j] * a13 * sqrt(a13);
    const double a15 = (A[0] * A[3] + A[1] * A[4] + A[2] * A[5]) * a13;
    const double a16 = a15 * a15;
    const double a17 = 3.0 * a15;
    const double a18 = 6.0 * a15;
    const double a19 = 9.0 * a15;
    double a20 = (A[3] * A[3] + A[4] * A[4] + A[5] * A[5] + A[0] * A[6] + A[1] * A[7] +
                  A[2] * A[8]) * a13 + a16;
    const double a21 = 9.0 * a20;
    a20 = 3.0 * a20;
    const double a22 = (9.0 * (A[3] * A[6] + A[4] * A[7] + A[5] * A[8]) +
                        3.0 * (A[0] * A[9] + A[1] * A[10] + A[2] * A[11])) * a13 +
                       a17 * (a20 - 4.0 * a16);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double f1 = A[k + 3] - a17 * A[k];
        f2[k] = (A[k + 6] - a18 * f1 - a20 * A[k]) * a14;
        f3[k] = (A[k + 9] - a21 * f1 - a22 * A[k]) * a14 - a19 * f2[k];
    }
}

__device__ __forceinline__ int at_or_max(const int *r, int l, int n) { return l <= n ? r[l] : INT_MAX; }

__device__ void regcor_one(const nb6_regcor_params &P, const nb6_regstate &s, int I,
                           const double *fg, int fstride, int *row, int *cnt,
                           unsigned char *shrink) {
#pragma clang fp contract(off)
    const size_t M = s.nmax;
    const int lmax = s.lmax;
    const double time = P.time;
    if (I < 1 || I > s.nmax) { cnt[C_UNS]++; return; }
    const int i = I - 1;

    // regf row -> 1-based list without self, util_gpu.F:101-110
    const int c = row[0];
    if (c < 0 || c > lmax - 1) { cnt[C_UNS]++; return; }
    int nnb = 0;
    for (int l = 1; l <= c; l++) {
        const int t = row[l] + P.ifirst;
        if (t < 1 || t > s.nmax) { cnt[C_UNS]++; return; }
        if (t != I) row[++nnb] = t;
    }
    row[0] = nnb;

    int *orow = s.list + (size_t)i * lmax;
    const int nnb0 = orow[0];
    if (nnb0 < 0 || nnb0 > lmax - 1) { cnt[C_UNS]++; return; }
    for (int l = 1; l <= nnb0; l++)
        if (orow[l] < 1 || orow[l] > s.nmax) { cnt[C_UNS]++; return; }

    double xi[3], vi[3], freg[3], fdr[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        xi[k] = s.x[k * M + i];
        vi[k] = s.xdot[k * M + i];
        freg[k] = fg[(size_t)k * fstride];
        fdr[k] = fg[(size_t)(3 + k) * fstride];
    }
    const double dtr = time - s.t0r[i];
    const double rscale = P.rscale;
    double ri2, rh2;
    if (P.nb_flag == 1) {
        const double e0 = xi[0] - P.rdens[0], e1 = xi[1] - P.rdens[1], e2 = xi[2] - P.rdens[2];
        ri2 = e0 * e0 + e1 * e1 + e2 * e2;
        rh2 = rscale * rscale;
    } else {
        ri2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        rh2 = 9.0 * (rscale * rscale);
    }
    double rs = s.rs[i];
    const double rs2 = rs * rs;
    double dfirr[3] = {0.0, 0.0, 0.0}, dfd[3] = {0.0, 0.0, 0.0};

    if (nnb == 0) {                                   // regcor_gpu.F:95-127
        const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        const double fij = (double)0.01f * P.bodym / (r2 * sqrt(r2));
        const double rdot = 3.0 * (xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]) / r2;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            dfirr[k] = -(fij * xi[k]);
            dfd[k] = -((vi[k] - rdot * xi[k]) * fij);
        }
        rs = rdot > 0.0 ? fmax(0.75 * rs, (double)0.1f * rscale) : (double)1.1f * rs;
        cnt[C_VOID]++;
    } else {                                          // regcor_gpu.F:139-238
        const float fnb = (float)nnb;
        const double fac = (double)(1.0f + (0.1f * ((float)P.nnbopt - fnb)) / fnb);
        double a3 = (double)fnb * fac;
        const int nbave = min(P.nnbopt + 150, max(P.nnbmax - 50, (int)P.znbmin));
        a3 = fmin(a3, (double)(0.9f * (float)nbave));
        const int nbp = (int)a3;
        if (ri2 > rh2 && P.nb_flag == 1)
            a3 = a3 * fmax(rscale / sqrt(ri2), (double)0.1f);
        else
            a3 = fmax(a3, P.znbmin);
        double a4 = a3 / (double)fnb;
        if ((a3 - (double)nnb0) * (a3 - (double)nnb) < 0.0) a4 = sqrt(a4);
        const double ridot = (xi[0] - P.rdens[0]) * vi[0] + (xi[1] - P.rdens[1]) * vi[1] +
                             (xi[2] - P.rdens[2]) * vi[2];
        if (ri2 > P.rc2 && P.nb_flag == 1 && ri2 < 100.0 * rh2) a4 = a4 * (1.0 + ridot * dtr / ri2);
        if (I > P.n && ri2 <= rh2) {                  // c.m. body, regcor_gpu.F:169-181
            if (!s.h) { cnt[C_UNS]++; return; }
            const double a2 = 100.0 * s.body[i] / fabs(s.h[I - P.n - 1]);
            if (a2 > rs) {
                a3 = fmax((double)(1.0f - fnb / (float)nbave), 0.0);
                a4 = 1.0 + 0.5 * a3;
            }
        }
        const double a5 = fmin(P.rsfac * dtr, 0.25);
        a4 = a4 - 1.0;
        a4 = a4 > a5 ? a5 : fmax(a4, -a5);
        a3 = ONE3 * a4;
        double a1 = 1.0 + a3 - a3 * a3;
        if (nnb < P.znbmin && nbp > P.znbmin && a1 < 1.0) a1 = (double)1.05f;
        if (fabs(a1 - 1.0) > (double)0.003f) {
            if (rs * rs > ri2 && rs * rs < 2.0 * ri2 && nnb > 5) a1 = sqrt(a1);
            rs = a1 * rs;
        }
        if (nnb > nbave && ridot < 0.0) {
            const double drs = (double)((fnb - (float)nbave) / (3.0f * fnb));
            rs = rs * (1.0 - drs);
        }
        if (P.m_flag == 0 && nnb <= 3 && ri2 < 100.0 * rh2) {   // regcor_gpu.F:222-238
            a1 = 2.0 * rs;
            for (int l = 1; l <= nnb; l++) {
                const int j = row[l] - 1;
                const double d0 = xi[0] - s.x[j], d1 = xi[1] - s.x[M + j], d2 = xi[2] - s.x[2 * M + j];
                const double rij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                const double rsdot = (d0 * (vi[0] - s.xdot[j]) + d1 * (vi[1] - s.xdot[M + j]) +
                                      d2 * (vi[2] - s.xdot[2 * M + j])) / rij;
                a1 = fmin(a1, rij + rsdot * dtr);
            }
            rs = fmax(a1, (double)1.1f * rs);
        }
    }
    s.rs[i] = rs;

    // losses and gains, regcor_gpu.F:252-350
    int nbloss = 0, nbgain = 0;
    double df2, fr2;
    if (nnb0 == 0) {
        nbgain = nnb;
        df2 = 1.0;
        fr2 = 0.0;
    } else {
        df2 = 0.0;
        fr2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double e = freg[k] - s.fr[k * M + i];
            df2 = df2 + e * e;
            fr2 = fr2 + freg[k] * freg[k];
        }
        bool jmin = false;
        for (int a = 1, b = 1; a <= nnb0 || b <= nnb;) {
            const int vo = at_or_max(orow, a, nnb0), vn = at_or_max(row, b, nnb);
            if (vo < vn) {
                nbloss++;
                if (s.step[vo - 1] < P.smin) jmin = true;
                a++;
            } else if (vn < vo) {
                nbgain++;
                b++;
            } else {
                a++;
                b++;
            }
        }
        // keep lost neighbours with small steps inside 2 RS, in ascending order, while
        // NNB <= NNBMAX (regcor_gpu.F:298-350)
        if (jmin) {
            for (int a = 1, b = 1; a <= nnb0 || b <= nnb;) {
                if (nnb > P.nnbmax || I > P.n) break;
                const int vo = at_or_max(orow, a, nnb0), vn = at_or_max(row, b, nnb);
                if (vo < vn) {
                    const int j = vo - 1;
                    bool keep = !(s.step[j] > P.smin || vo < P.ifirst || vo > P.n);
                    if (keep) {
                        const double e0 = xi[0] - s.x[j], e1 = xi[1] - s.x[M + j],
                                     e2 = xi[2] - s.x[2 * M + j];
                        keep = !(e0 * e0 + e1 * e1 + e2 * e2 > 4.0 * rs2);
                    }
                    if (keep) {
                        for (int l = nnb; l >= b; l--) row[l + 1] = row[l];
                        row[b] = vo;
                        nnb++;
                        nbloss--;
                        cnt[C_SMIN]++;
                        double fa[3], fj[3];
                        pair_fj(s, j, xi, vi, fa, fj);
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            freg[k] = freg[k] - fa[k];
                            fdr[k] = fdr[k] - fj[k];
                        }
                        a++;
                        b++;
                    } else {
                        a++;
                    }
                } else if (vn < vo) {
                    b++;
                } else {
                    a++;
                    b++;
                }
            }
        }
    }

    // FPCORR decision (regcor_gpu.F:600-612); KZ(38) = 3 drops the gains
    bool fp = false;
    if (P.kz38 == 1)
        fp = true;
    else if (P.kz38 == 2 && df2 > (double)0.0001f * fr2)
        fp = true;
    else if (P.kz38 == 3 && nbloss > 0)
        fp = true;
    const bool fp_gains = fp && P.kz38 != 3;
    fp = fp && (nbloss + (fp_gains ? nbgain : 0)) > 0;

    // force differences of lost then gained neighbours (regcor_gpu.F:355-415) and the
    // FPCORR sums in the same order (fpcorr.F:19-92)
    double sv2[3] = {0.0, 0.0, 0.0}, sv3[3] = {0.0, 0.0, 0.0};
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 0 && nbloss == 0) continue;
        if (pass == 1 && nbgain == 0) continue;
        const bool fps = fp && (pass == 0 || fp_gains);
        for (int a = 1, b = 1; a <= nnb0 || b <= nnb;) {
            const int vo = at_or_max(orow, a, nnb0), vn = at_or_max(row, b, nnb);
            int j1 = 0;
            if (vo < vn) {
                if (pass == 0) j1 = vo;
                a++;
            } else if (vn < vo) {
                if (pass == 1) j1 = vn;
                b++;
            } else {
                a++;
                b++;
            }
            if (!j1) continue;
            const int j = j1 - 1;
            double fa[3], fj[3];
            pair_fj(s, j, xi, vi, fa, fj);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (pass == 0) {
                    dfirr[k] = dfirr[k] - fa[k];
                    dfd[k] = dfd[k] - fj[k];
                } else {
                    dfirr[k] = dfirr[k] + fa[k];
                    dfd[k] = dfd[k] + fj[k];
                }
            }
            if (fps) {
                if (j1 < P.ifirst) {
                    cnt[C_UNS]++;
                } else {
                    double f2[3], f3[3];
                    fp_terms(s, time, i, j, xi, vi, f2, f3);
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        if (pass == 0) {
                            f2[k] = -f2[k];
                            f3[k] = -f3[k];
                        }
                        sv2[k] = sv2[k] + f2[k];
                        sv3[k] = sv3[k] + f3[k];
                    }
                }
            }
        }
    }
    if (P.kz38 == 3 && nbloss > 0) nbgain = 0;

    // corrector, regcor_gpu.F:417-530
    const double dtsq = dtr * dtr;
    const double dt6 = 6.0 / (dtr * dtsq);
    const double dt2 = 2.0 / dtsq;
    double dtsq12 = ONE12 * dtsq;
    double dtr13 = ONE3 * dtr;
    s.t0r[i] = time;
    const double stp = s.step[i];
    if (stp < 5.0 * P.dtmin && dtr > 50.0 * stp) {
        dtr13 = 0.0;
        dtsq12 = 0.0;
    }
    double xnd[3], d2r[3], d3r[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double dfr = s.fr[k * M + i] - freg[k] - dfirr[k];
        const double fdr0 = fdr[k] + dfd[k];
        const double frd = s.frdot[k * M + i];
        const double sm = frd + fdr0;
        const double at3 = 2.0 * dfr + dtr * sm;
        const double bt2 = -3.0 * dfr - dtr * (sm + frd);
        s.xn[k * M + i] = s.x0[k * M + i] + ((double)0.6f * at3 + bt2) * dtsq12;
        xnd[k] = s.x0dot[k * M + i] + (0.75 * at3 + bt2) * dtr13;
        s.xndot[k * M + i] = xnd[k];
        const double fik = s.fi[k * M + i] + dfirr[k];
        s.fi[k * M + i] = fik;
        s.fr[k * M + i] = freg[k];
        s.fidot[k * M + i] = s.fidot[k * M + i] + dfd[k];
        s.frdot[k * M + i] = fdr[k];
        if (k == 1) s.fidot[i] = s.fidot[i] - P.tidal4 * dfr;
        if (k == 0) s.fidot[M + i] = s.fidot[M + i] + P.tidal4 * dfr;
        s.d0[k * M + i] = fik;
        s.d0r[k * M + i] = freg[k];
        s.d1r[k * M + i] = fdr[k];
        d3r[k] = at3 * dt6;
        d2r[k] = (3.0 * at3 + bt2) * dt2;
    }
    if (fp) {                                         // fpcorr.F:93-97
        cnt[C_CORR]++;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s.d2[k * M + i] = s.d2[k * M + i] + sv2[k];
            s.d3[k * M + i] = s.d3[k * M + i] + sv3[k];
            d2r[k] = d2r[k] - sv2[k];
            d3r[k] = d3r[k] - sv3[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s.d2r[k * M + i] = d2r[k];
        s.d3r[k * M + i] = d3r[k];
    }

    // list copy, regcor_gpu.F:614-622
    row[0] = nnb;
    if (nbloss + nbgain > 0) {
        for (int l = 0; l <= nnb; l++) orow[l] = row[l];
        cnt[C_FLUX] += nbloss + nbgain;
    }

    // new regular step, regcor_gpu.F:624-740
    fr2 = freg[0] * freg[0] + freg[1] * freg[1] + freg[2] * freg[2];
    const double w1 = fdr[0] * fdr[0] + fdr[1] * fdr[1] + fdr[2] * fdr[2];
    const double w2 = d2r[0] * d2r[0] + d2r[1] * d2r[1] + d2r[2] * d2r[2];
    const double w3 = d3r[0] * d3r[0] + d3r[1] * d3r[1] + d3r[2] * d3r[2];
    double w0 = (sqrt(fr2 * w2) + w1) / (sqrt(w1 * w3) + w2);
    w0 = P.etar * w0;
    double ttmp = sqrt(w0);
    const double dt0 = ttmp;
    const double stepr = s.stepr[i];
    const double smax = P.smax;
    if (ttmp > 2.0 * stepr) {
        if (fmod(time, 2.0 * dtr) == 0.0) {
            ttmp = fmin(2.0 * dtr, smax);
            while (dt0 > 2.0 * ttmp && 2.0 * ttmp < smax && fmod(time, 2.0 * ttmp) == 0.0)
                ttmp = 2.0 * ttmp;
        } else {
            ttmp = stepr;
        }
    } else if (ttmp < dtr) {
        ttmp = 0.5 * dtr;
        if (ttmp > dt0) ttmp = 0.5 * ttmp;
    } else {
        ttmp = stepr;
    }
    if (ttmp > (double)0.1f && ttmp > dtr) {          // Makino criterion
        double dv2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double e = vi[k] - xnd[k];
            dv2 = dv2 + e * e;
        }
        const double q = dv2 != 0.0 ? 1.0e-4 * (dtr * dtr) * fr2 / dv2 : INFINITY;
        const double dtj = dtr * pow(q, (double)0.1f);
        if (dtj < ttmp) ttmp = dtr;
    }
    ttmp = fmin(ttmp, smax);
    s.stepr[i] = ttmp;
    if (nnb0 == 0) {        // STEP itself is scaled by regcor_step_kernel after the block
        s.timenw[i] = s.t0[i] + 0.125 * stp;
        *shrink = 1;
    }
}

__global__ __launch_bounds__(256) void regcor_kernel(nb6_regcor_params P, nb6_regstate s, int ni,
                                                     const int *__restrict__ addr,
                                                     const double *__restrict__ force, int fstride,
                                                     int *__restrict__ nlist, int *__restrict__ counters,
                                                     unsigned char *__restrict__ shrink) {
    __shared__ int red[NCOUNT];
    if (threadIdx.x < NCOUNT) red[threadIdx.x] = 0;
    __syncthreads();
    const int k = blockIdx.x * 256 + threadIdx.x;
    int cnt[NCOUNT] = {0, 0, 0, 0, 0};
    if (k < ni) {
        shrink[k] = 0;
        regcor_one(P, s, addr[k], force + k, fstride, nlist + (size_t)k * s.lmax, cnt, shrink + k);
    }
#pragma unroll
    for (int q = 0; q < NCOUNT; q++)
        if (cnt[q]) atomicAdd(&red[q], cnt[q]);
    __syncthreads();
    if (threadIdx.x < NCOUNT && red[threadIdx.x]) atomicAdd(&counters[threadIdx.x], red[threadIdx.x]);
}

// STEP(I) = STEP(I) / 8 for particles whose old list was empty (regcor_gpu.F:736-739),
// applied after the whole block so that every retention test of the block reads the
// steps the block started with
__global__ __launch_bounds__(256) void regcor_step_kernel(int ni, const int *__restrict__ addr,
                                                          const unsigned char *__restrict__ shrink,
                                                          double *__restrict__ step) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < ni && shrink[k]) step[addr[k] - 1] = 0.125 * step[addr[k] - 1];
}

DBuf g_shrink;

}  // namespace

int regcor_device(const nb6_regcor_params *p, const nb6_regstate *s, int ni, const int *d_addr,
                  const double *d_force, int force_stride, int *d_nlist, int *d_counters,
                  hipStream_t stream) {
    if (!p || !s) { set_error(E_ARG, "nb6_regcor_device: null parameter block"); return E_ARG; }
    if (ni < 0) { set_error(E_ARG, "nb6_regcor_device: ni < 0"); return E_ARG; }
    if (ni == 0) return OK;
    if (s->nmax <= 0 || s->lmax < 2 || p->nnbmax + 1 >= s->lmax) {
        set_error(E_ARG, "nb6_regcor_device: need nmax > 0 and lmax > nnbmax + 1");
        return E_ARG;
    }
    if (force_stride < ni) { set_error(E_ARG, "nb6_regcor_device: force_stride < ni"); return E_ARG; }
    if (p->ifirst < 1 || p->n < 0 || p->kz38 < 0 || p->kz38 > 3 || !(p->time == p->time)) {
        set_error(E_ARG, "nb6_regcor_device: bad IFIRST, N, KZ(38) or TIME");
        return E_ARG;
    }
    const void *need[] = {s->x, s->xdot, s->body, s->x0, s->x0dot, s->f, s->fdot, s->t0,
                          s->step, s->timenw, s->fi, s->fidot, s->fr, s->frdot, s->d2, s->d3,
                          s->t0r, s->stepr, s->rs, s->list, s->xn, s->xndot, s->d0, s->d0r,
                          s->d1r, s->d2r, s->d3r, d_addr, d_force, d_nlist, d_counters};
    for (const void *q : need)
        if (!q) { set_error(E_ARG, "nb6_regcor_device: null device array"); return E_ARG; }
    Device &d = device();
    if (!d.ready) return E_NODEV;
    KTimer &kt = ktimer("regcor");
    kt.start(stream);
    auto *shrink = static_cast<unsigned char *>(g_shrink.get((size_t)ni));
    hipLaunchKernelGGL(regcor_kernel, dim3((ni + 255) / 256), dim3(256), 0, stream, *p, *s, ni,
                       d_addr, d_force, force_stride, d_nlist, d_counters, shrink);
    NB6_HIP(hipGetLastError());
    hipLaunchKernelGGL(regcor_step_kernel, dim3((ni + 255) / 256), dim3(256), 0, stream, ni,
                       d_addr, shrink, s->step);
    NB6_HIP(hipGetLastError());
    kt.stop(stream);
    return OK;
}

}  // namespace nb6
