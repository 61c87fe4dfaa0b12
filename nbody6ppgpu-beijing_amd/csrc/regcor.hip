// Regular corrector of a block on the GPU: REGCOR_GPU (regcor_gpu.F:1-740) with
// FPCORR (fpcorr.F:1-148), called by the reference once per particle after each
// regular force evaluation (util_gpu.F:150-175).
//
// One wavefront per active particle I.  Its work is O(old + new list length): the regf
// row is turned into the 1-based list without self (util_gpu.F:101-110), RS is
// updated (regcor_gpu.F:95-238), the old and new lists are walked in ascending
// order to find losses and gains (regcor_gpu.F:252-296), lost neighbours with
// STEP <= SMIN inside 2 RS are kept (regcor_gpu.F:298-350), the force of each lost
// (gained) neighbour moves between the irregular and regular parts
// (regcor_gpu.F:355-415), then the Hermite corrector (:417-530), FPCORR, the list
// copy and the new regular step (:600-740).  Lists are read coalesced into LDS, losses
// and gains are found by binary search there, and the moved force terms are formed 64
// at a time by the lanes and summed in list order by twelve lanes, so every sum has the
// reference's order and rounding.
//
// Arithmetic follows the reference operation by operation: FP64 without FMA
// contraction in Fortran's left-to-right order, and the REAL*4 sub-expressions of the
// reference (FLOAT() of counts, single-precision literals such as 0.1, 0.6, 1.05,
// 1.1, 0.003) in FP32.  The only library function whose last bit may differ from the
// host's is pow() in the Makino test, which only enters a comparison.
#include "nb6_internal.h"
#include "../../include/nb6gpu.h"

#include <algorithm>
#include <cmath>

namespace nb6 {

namespace {

constexpr double ONE3 = 1.0 / 3.0;
constexpr double ONE12 = 1.0 / 12.0;

enum { C_VOID, C_FLUX, C_SMIN, C_CORR, C_UNS, C_FAST, C_CMDER, NCOUNT };

// Packed neighbour records: what the list kernel reads of a neighbour J, in one record
// of PK_STRIDE doubles per particle (192 bytes: at most two 128-byte lines) instead of
// eight arrays (eight lines).  Written at the start of a call from the state arrays,
// which the call does not change before the list kernel has read them, so the values
// are the same.
enum { PK_X = 0, PK_XDOT = 3, PK_BODY = 6, PK_STEP = 7, PK_T0 = 8, PK_F = 9, PK_FDOT = 12,
       PK_X0DOT = 15, PK_STRIDE = 24 };

struct NbSrc {
    const nb6_regstate &s;
    const double *pk;   // packed records or nullptr (read the state arrays)
    __device__ double x(int j, int k) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_X + k] : s.x[3 * (size_t)j + k]; }
    __device__ double xdot(int j, int k) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_XDOT + k] : s.xdot[3 * (size_t)j + k]; }
    __device__ double body(int j) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_BODY] : s.body[j]; }
    __device__ double step(int j) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_STEP] : s.step[j]; }
    __device__ double t0(int j) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_T0] : s.t0[j]; }
    __device__ double f(int j, int k) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_F + k] : s.f[3 * (size_t)j + k]; }
    __device__ double fdot(int j, int k) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_FDOT + k] : s.fdot[3 * (size_t)j + k]; }
    __device__ double x0dot(int j, int k) const { return pk ? pk[(size_t)j * PK_STRIDE + PK_X0DOT + k] : s.x0dot[3 * (size_t)j + k]; }
};

__global__ __launch_bounds__(256) void regcor_pack_kernel(nb6_regstate s, double *__restrict__ pk) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= s.nmax) return;
    double *o = pk + (size_t)j * PK_STRIDE;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o[PK_X + k] = s.x[3 * (size_t)j + k];
        o[PK_XDOT + k] = s.xdot[3 * (size_t)j + k];
        o[PK_F + k] = s.f[3 * (size_t)j + k];
        o[PK_FDOT + k] = s.fdot[3 * (size_t)j + k];
        o[PK_X0DOT + k] = s.x0dot[3 * (size_t)j + k];
    }
    o[PK_BODY] = s.body[j];
    o[PK_STEP] = s.step[j];
    o[PK_T0] = s.t0[j];
}

// force and first derivative of body j (0-based) on i, regcor_gpu.F:334-353
__device__ __forceinline__ void pair_fj(const NbSrc &s, int j, const double *xi,
                                        const double *vi, double *fa, double *fj) {
#pragma clang fp contract(off)
    double a[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a[k] = s.x(j, k) - xi[k];
        dv[k] = s.xdot(j, k) - vi[k];
    }
    const double rij2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double dr2i = 1.0 / rij2;
    const double dr3i = s.body(j) * dr2i * sqrt(dr2i);
    const double drdv = a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2];
    const double drdp = 3.0 * drdv * dr2i;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        fa[k] = a[k] * dr3i;
        fj[k] = (dv[k] - a[k] * drdp) * dr3i;
    }
}

// second and third derivative of the force of body j on i, fpcorr.F:24-74; a KS
// component j (jcm >= 0, 0-based c.m. body N + KVEC(J)) takes its own position and
// velocity with the c.m. body's F and FDOT (fpcorr.F:46-61)
__device__ __forceinline__ void fp_terms(const NbSrc &s, double time, int i, int j,
                                         int jcm, const double *xi, const double *vi,
                                         double *f2, double *f3) {
#pragma clang fp contract(off)
    double A[12];
    if (jcm >= 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            A[k] = s.x(j, k) - xi[k];
            A[k + 3] = s.xdot(j, k) - vi[k];
            A[k + 6] = 2.0 * (s.f(jcm, k) - s.f(i, k));
            A[k + 9] = 6.0 * (s.fdot(jcm, k) - s.fdot(i, k));
        }
    } else {
        const double sj = time - s.t0(j);
        const double s3 = 3.0 * sj;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double fj = s.f(j, k), fdj = s.fdot(j, k);
            A[k] = s.x(j, k) - xi[k];
            A[k + 3] = (fdj * s3 + 2.0 * fj) * sj + s.x0dot(j, k) - vi[k];
            A[k + 6] = 2.0 * (fdj * s3 + fj - s.f(i, k));
            A[k + 9] = 6.0 * (fdj - s.fdot(i, k));
        }
    }
    const double a13 = 1.0 / (A[0] * A[0] + A[1] * A[1] + A[2] * A[2]);
    const double a14 = s.body(j) * a13 * sqrt(a13);
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

// ------------------------------------------------------------------ one wavefront per particle
constexpr int WAVE = 64;
constexpr int NTERM = 12;   // moved force terms: DFIRR[3], DFD[3], F2DOT[3], F3DOT[3]
constexpr int TCH = 32;     // terms held in LDS at a time

// LDS of one wavefront: half a chunk of terms (double [NTERM][32]), the new and old
// lists (int [lmax + 1] each) and two bits per old-list entry: lost, and candidate for
// the SMIN retention
__host__ __device__ constexpr int loss_words(int lmax) { return 2 * ((lmax + 64) / 64); }
__host__ __device__ constexpr size_t wave_lds_bytes(int lmax) {
    return ((size_t)NTERM * TCH * 8 + (size_t)2 * (lmax + 1) * 4 +
            (size_t)2 * loss_words(lmax) * 4 + 15) & ~(size_t)15;
}

__device__ __forceinline__ bool lds_contains(const int *a, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && a[lo] == v;
}

// number of set bits of m in the lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ double pick3(const double *a, int q) {
    return q == 0 ? a[0] : q == 1 ? a[1] : a[2];
}

// NNB = 0: force of a mass BODYM/100 at the origin, regcor_gpu.F:108-116
__device__ __forceinline__ void void_terms(const nb6_regcor_params &P, const double *xi,
                                           const double *vi, double *dfirr, double *dfd) {
#pragma clang fp contract(off)
    const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    const double fij = (double)0.01f * P.bodym / (r2 * sqrt(r2));
    const double rdot = 3.0 * (xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]) / r2;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        dfirr[k] = -(fij * xi[k]);
        dfd[k] = -((vi[k] - rdot * xi[k]) * fij);
    }
}

// RS after the step, regcor_gpu.F:95-238.  nl: the new list before SMIN retention
// (1-based, ascending; read for NNB <= 3 only).
__device__ double rs_update(const nb6_regcor_params &P, const nb6_regstate &s, int I, int i,
                            const double *xi, const double *vi, int nnb, int nnb0, double dtr,
                            double ri2, double rh2, double rs, const int *nl) {
#pragma clang fp contract(off)
    const double rscale = P.rscale;
    if (nnb == 0) {
        const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        const double rdot = 3.0 * (xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]) / r2;
        return rdot > 0.0 ? fmax(0.75 * rs, (double)0.1f * rscale) : (double)1.1f * rs;
    }
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
    if (I > P.n && ri2 <= rh2) {                      // c.m. body, regcor_gpu.F:169-181
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
        for (int l = 0; l < nnb; l++) {
            const int j = nl[l] - 1;
            const double d0 = xi[0] - s.x[3 * (size_t)j], d1 = xi[1] - s.x[3 * (size_t)j + 1], d2 = xi[2] - s.x[3 * (size_t)j + 2];
            const double rij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double rsdot = (d0 * (vi[0] - s.xdot[3 * (size_t)j]) + d1 * (vi[1] - s.xdot[3 * (size_t)j + 1]) +
                                  d2 * (vi[2] - s.xdot[3 * (size_t)j + 2])) / rij;
            a1 = fmin(a1, rij + rsdot * dtr);
        }
        rs = fmax(a1, (double)1.1f * rs);
    }
    return rs;
}

// Insert v into the ascending list nl[0..n) at its place (all lanes of the wavefront:
// the tail moves up by one, 64 entries at a time from the top).
__device__ __forceinline__ void list_insert(int *nl, int n, int v, int lane) {
    int lo = 0, hi_ = n;
    while (lo < hi_) {
        const int mid = (lo + hi_) >> 1;
        if (nl[mid] < v)
            lo = mid + 1;
        else
            hi_ = mid;
    }
    for (int c0 = n - 1; c0 >= lo; c0 -= WAVE) {
        const int idx = c0 - lane;
        const bool act = idx >= lo;
        int w = 0;
        if (act) w = nl[idx];
        __builtin_amdgcn_wave_barrier();
        if (act) nl[idx + 1] = w;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) nl[lo] = v;
    __builtin_amdgcn_wave_barrier();
}

// CHECKL for KZ(37) = 1 (checkl.F:243-340): the first body J of LISTV (in LISTV order)
// with |x_J - x_I|^2 <= 4 RS^2, approaching (dx.dv <= 0), impact parameter^2 <= RS^2
// and not yet in the list joins it; its force and derivative leave FREG and FDR (the
// gain loop moves them to the irregular side).  NNB > NNBMAX stops the search.  A c.m.
// particle I takes the terms between its KS components and J (or J's components, when
// J is a c.m. body inside its c.m. approximation), mass-weighted, unless J is a single
// body outside CMSEP2 R^2.  The lanes test 64 LISTV entries at a time; every lane
// then runs the chosen body with the same values.  Returns the new NNB.
__device__ int checkl_fast(const nb6_regcor_params &P, const nb6_regstate &s, int I, int i,
                           const double *xi, const double *vi, double rs2, double *freg,
                           double *fdr, int *nl, int nnb, int *cnt, int lane) {
#pragma clang fp contract(off)
    const int nv = s.listv[0];
    for (int b = 0; b < nv; b += WAVE) {
        const int L = b + lane;
        bool q = false;
        int J = 0;
        if (L < nv) {
            J = s.listv[1 + L];
            if (J != I && J >= 1 && J <= s.nmax) {
                const size_t j3 = 3 * (size_t)(J - 1), i3 = 3 * (size_t)i;
                const double a1 = s.x[j3] - xi[0], a2 = s.x[j3 + 1] - xi[1], a3 = s.x[j3 + 2] - xi[2];
                const double rij2 = a1 * a1 + a2 * a2 + a3 * a3;
                const double a4 = s.xdot[j3] - s.xdot[i3], a5 = s.xdot[j3 + 1] - s.xdot[i3 + 1],
                             a6 = s.xdot[j3 + 2] - s.xdot[i3 + 2];
                const double a7 = a1 * a4 + a2 * a5 + a3 * a6;
                const double p2 = rij2 - a7 * a7 / (a4 * a4 + a5 * a5 + a6 * a6);
                q = !(rij2 > 4.0 * rs2) && !(a7 > 0.0) && !(p2 > rs2) && !lds_contains(nl, nnb, J);
            }
        }
        const unsigned long long m = __ballot(q);
        if (!m) continue;
        J = __shfl(J, __builtin_ctzll(m));
        cnt[C_FAST]++;
        if (nnb > P.nnbmax) return nnb;
        const int j = J - 1;
        bool resolved = false;
        double rij2c = 0.0;
        if (I > P.n) {
            const double d0 = xi[0] - s.x[3 * (size_t)j], d1 = xi[1] - s.x[3 * (size_t)j + 1],
                         d2 = xi[2] - s.x[3 * (size_t)j + 2];
            rij2c = d0 * d0 + d1 * d1 + d2 * d2;
            const double rp = s.r[I - P.n - 1];
            resolved = !(rij2c > P.cmsep2 * (rp * rp) && J <= P.n);
        }
        if (!resolved) {                              // checkl.F:191-215
            double dx[3], dv[3], dr2 = 0.0, drdv = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                dx[k] = s.x[3 * (size_t)j + k] - xi[k];
                dv[k] = s.xdot[3 * (size_t)j + k] - vi[k];
                dr2 = dr2 + dx[k] * dx[k];
                drdv = drdv + dx[k] * dv[k];
            }
            const double dr2i = 1.0 / dr2;
            const double dr3i = s.body[j] * dr2i * sqrt(dr2i);
            drdv = 3.0 * drdv * dr2i;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                freg[k] = freg[k] - dx[k] * dr3i;
                fdr[k] = fdr[k] - (dv[k] - dx[k] * drdv) * dr3i;
            }
        } else {                                      // checkl.F:282-336
            const int i1 = 2 * (I - P.n) - 1;
            int k1 = J, k2 = 0;
            if (J > P.n) {
                const double rq = s.r[J - P.n - 1];
                if (rij2c < P.cmsep2 * (rq * rq)) {
                    k1 = 2 * (J - P.n) - 1;
                    k2 = k1 + 1;
                }
            }
            for (int K = k1; K != 0; K = (K == k1 && k2) ? k2 : 0) {
                for (int Lc = i1; Lc <= i1 + 1; Lc++) {
                    const size_t k3 = 3 * (size_t)(K - 1), l3 = 3 * (size_t)(Lc - 1);
                    const double a1 = s.x[k3] - s.x[l3], a2 = s.x[k3 + 1] - s.x[l3 + 1],
                                 a3 = s.x[k3 + 2] - s.x[l3 + 2];
                    const double rij2 = a1 * a1 + a2 * a2 + a3 * a3;
                    const double dv0 = s.xdot[k3] - s.xdot[l3], dv1 = s.xdot[k3 + 1] - s.xdot[l3 + 1],
                                 dv2 = s.xdot[k3 + 2] - s.xdot[l3 + 2];
                    const double dr2i = 1.0 / rij2;
                    const double dr3i = s.body[K - 1] * s.body[Lc - 1] * dr2i * sqrt(dr2i) / s.body[i];
                    const double drdv = 3.0 * (a1 * dv0 + a2 * dv1 + a3 * dv2) * dr2i;
                    freg[0] = freg[0] - a1 * dr3i;
                    freg[1] = freg[1] - a2 * dr3i;
                    freg[2] = freg[2] - a3 * dr3i;
                    fdr[0] = fdr[0] - (dv0 - a1 * drdv) * dr3i;
                    fdr[1] = fdr[1] - (dv1 - a2 * drdv) * dr3i;
                    fdr[2] = fdr[2] - (dv2 - a3 * drdv) * dr3i;
                }
            }
        }
        list_insert(nl, nnb, J, lane);
        return nnb + 1;
    }
    return nnb;
}

// Hand-over from the list kernel to the finishing kernel, SoA over the block:
// doubles DFIRR, DFD, F2DOT and F3DOT sums, FREG, FDR (after retention); ints flags
// (1: corrected, 2: FPCORR applies), NNB before retention, NNB0 and, for NNB <= 3,
// the new list.
enum { H_DFIRR = 0, H_DFD = 3, H_SV2 = 6, H_SV3 = 9, H_FREG = 12, H_FDR = 15, NHD = 18 };
enum { HI_FLAGS = 0, HI_NNB = 1, HI_NNB0 = 2, HI_J3 = 3, NHI = 6 };

// Lists of particle I by one wavefront (regcor_gpu.F:252-415 and the list copy): values
// every lane computes from the same inputs are uniform, list work is spread over the
// lanes, lane 0 alone runs the rare SMIN retention.  hd, hi: this particle's column of
// the hand-over (stride ni).
template <bool CHECKL>
__device__ void regcor_wave(const nb6_regcor_params &P, const nb6_regstate &s, const double *pk,
                            int I, const double *fg, int fstride, int *row, int *cnt, double *T,
                            int *nl, int *ol, unsigned *lb, unsigned *cb, int lane, double *hd,
                            int *hi, int ni) {
#pragma clang fp contract(off)
    const NbSrc nb{s, pk};
    const int lmax = s.lmax;
    const double time = P.time;
    if (I < 1 || I > s.nmax) { cnt[C_UNS]++; return; }
    const int i = I - 1;
    if (I > P.n && !s.h) { cnt[C_UNS]++; return; }

    // every load that depends only on I is issued before any is used: the regf row's
    // count and first 128 entries, the same of LIST(.,I), and the particle's own state
    int *orow = s.list + (size_t)i * lmax;
    const int c = row[0];
    const int nnb0 = orow[0];
    int rpre0 = 1 + lane < lmax ? row[1 + lane] : 0;
    int rpre1 = 1 + WAVE + lane < lmax ? row[1 + WAVE + lane] : 0;
    int opre0 = 1 + lane < lmax ? orow[1 + lane] : 0;
    int opre1 = 1 + WAVE + lane < lmax ? orow[1 + WAVE + lane] : 0;
    double xi[3], vi[3], freg[3], fdr[3], fr_old[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        xi[k] = s.x[3 * (size_t)i + k];
        vi[k] = s.xdot[3 * (size_t)i + k];
        freg[k] = fg[(size_t)k * fstride];
        fdr[k] = fg[(size_t)(3 + k) * fstride];
        fr_old[k] = s.fr[3 * (size_t)i + k];
    }
    const double rs_old = s.rs[i];
    if (c < 0 || c > lmax - 1 || nnb0 < 0 || nnb0 > lmax - 1) { cnt[C_UNS]++; return; }

    // regf row -> 1-based list without self, util_gpu.F:101-110 (ballot compaction)
    int nnb = 0;
    bool bad = false;
    for (int b = 0; b < c; b += WAVE) {
        const int l = b + lane;
        int t = 0;
        bool keep = false;
        if (l < c) {
            t = (b == 0 ? rpre0 : b == WAVE ? rpre1 : row[1 + l]) + P.ifirst;
            bad = bad || t < 1 || t > s.nmax;
            keep = t != I;
        }
        const unsigned long long m = __ballot(keep);
        if (keep) nl[nnb + lanes_below(m)] = t;
        nnb += __popcll(m);
    }
    for (int b = 0; b < nnb0; b += WAVE) {
        const int l = b + lane;
        if (l < nnb0) {
            const int t = b == 0 ? opre0 : b == WAVE ? opre1 : orow[1 + l];
            bad = bad || t < 1 || t > s.nmax;
            ol[l] = t;
        }
    }
    if (__ballot(bad)) { cnt[C_UNS]++; return; }
    __builtin_amdgcn_wave_barrier();

    const double rs2 = rs_old * rs_old;
    double dfirr[3] = {0.0, 0.0, 0.0}, dfd[3] = {0.0, 0.0, 0.0};
    if (nnb == 0) {                                   // regcor_gpu.F:108-116
        void_terms(P, xi, vi, dfirr, dfd);
        cnt[C_VOID]++;
    }
    const int nnb_pre = nnb;
    if (lane < 3 && lane < nnb) hi[(HI_J3 + lane) * ni] = nl[lane];

    // CHECKL, regcor_gpu.F:240-244 (after the RS update, which the finishing kernel does
    // with the NNB and list from before this point)
    if (nnb > 0 && P.kz37 > 1) cnt[C_UNS]++;
    if (CHECKL && nnb > 0 && P.kz37 == 1 && s.listv && s.listv[0] > 0)
        nnb = checkl_fast(P, s, I, i, xi, vi, rs2, freg, fdr, nl, nnb, cnt, lane);

    // losses (ascending), gains and JMIN, regcor_gpu.F:252-296
    int nloss = 0, nbgain = 0;
    bool jmin = false;
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
            const double e = freg[k] - fr_old[k];
            df2 = df2 + e * e;
            fr2 = fr2 + freg[k] * freg[k];
        }
        // a lost J with STEP(J) < SMIN enables the retention (JMIN); a candidate has
        // STEP(J) <= SMIN, IFIRST <= J <= N and |x_J - x_I| <= 2 RS (regcor_gpu.F:305-313)
        bool jm = false;
        for (int b = 0; b < nnb0; b += WAVE) {
            const int l = b + lane;
            bool lost = false, cand = false;
            if (l < nnb0) {
                const int t = ol[l];
                lost = !lds_contains(nl, nnb, t);
                if (lost) {
                    const double st = nb.step(t - 1);
                    jm = jm || st < P.smin;
                    if (!(st > P.smin || t < P.ifirst || t > P.n)) {
                        const double e0 = xi[0] - nb.x(t - 1, 0), e1 = xi[1] - nb.x(t - 1, 1),
                                     e2 = xi[2] - nb.x(t - 1, 2);
                        cand = !(e0 * e0 + e1 * e1 + e2 * e2 > 4.0 * rs2);
                    }
                }
            }
            const unsigned long long m = __ballot(lost), mc = __ballot(cand);
            if (lane == 0) {
                lb[b / 32] = (unsigned)m;
                lb[b / 32 + 1] = (unsigned)(m >> 32);
                cb[b / 32] = (unsigned)mc;
                cb[b / 32 + 1] = (unsigned)(mc >> 32);
            }
            nloss += __popcll(m);
        }
        jmin = __ballot(jm) != 0;
        for (int b = 0; b < nnb; b += WAVE) {
            const int l = b + lane;
            const bool gain = l < nnb && !lds_contains(ol, nnb0, nl[l]);
            nbgain += __popcll(__ballot(gain));
        }
    }
    int nbloss = nloss;
    __builtin_amdgcn_wave_barrier();

    // keep the candidates, ascending, while NNB <= NNBMAX (regcor_gpu.F:298-350): NNB
    // only grows at a kept candidate, so testing the bound before each candidate is the
    // reference's test before each lost neighbour.  All lanes run this loop with the same
    // values; each insertion shifts the tail of the new list by one, 64 entries at a time
    // from the top.  A kept neighbour loses its loss bit.
    if (jmin && I <= P.n) {
        bool full = false;
        for (int wd = 0; wd < 2 * ((nnb0 + WAVE - 1) / WAVE) && !full; wd++) {
            unsigned bits = lb[wd] & cb[wd];
            while (bits) {
                if (nnb > P.nnbmax) {
                    full = true;
                    break;
                }
                const int q = wd * 32 + __builtin_ctz(bits);
                bits &= bits - 1;
                const int vo = ol[q];
                int lo = 0, hi_ = nnb;
                while (lo < hi_) {
                    const int mid = (lo + hi_) >> 1;
                    if (nl[mid] < vo)
                        lo = mid + 1;
                    else
                        hi_ = mid;
                }
                for (int c0 = nnb - 1; c0 >= lo; c0 -= WAVE) {
                    const int idx = c0 - lane;
                    const bool act = idx >= lo;
                    int v = 0;
                    if (act) v = nl[idx];
                    __builtin_amdgcn_wave_barrier();
                    if (act) nl[idx + 1] = v;
                    __builtin_amdgcn_wave_barrier();
                }
                if (lane == 0) {
                    nl[lo] = vo;
                    lb[wd] &= ~(1u << (q & 31));
                }
                __builtin_amdgcn_wave_barrier();
                nnb++;
                nbloss--;
                cnt[C_SMIN]++;
                double fa[3], fj[3];
                pair_fj(nb, vo - 1, xi, vi, fa, fj);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    freg[k] = freg[k] - fa[k];
                    fdr[k] = fdr[k] - fj[k];
                }
            }
        }
    }

    // FPCORR decision (regcor_gpu.F:600-612); KZ(38) = 3 drops the gains
    bool fp = P.kz38 == 1 || (P.kz38 == 2 && df2 > (double)0.0001f * fr2) ||
              (P.kz38 == 3 && nbloss > 0);
    const bool fp_gains = fp && P.kz38 != 3;
    fp = fp && (nbloss + (fp_gains ? nbgain : 0)) > 0;

    // force moved by lost then gained neighbours (regcor_gpu.F:355-415) and the FPCORR
    // sums (fpcorr.F:19-92): up to 64 terms at a time, compacted in list order into LDS,
    // then lane q < 12 adds term component q over them in that order, as the
    // reference's serial loop does
    double acc = lane < 3 ? pick3(dfirr, lane) : lane < 6 ? pick3(dfd, lane - 3) : 0.0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 0 ? nbloss == 0 : nbgain == 0) continue;
        const int np = pass == 0 ? nnb0 : nnb;
        const bool fps = fp && (pass == 0 || fp_gains);
        for (int b = 0; b < np; b += WAVE) {
            const int l = b + lane;
            int j1 = 0;
            if (l < np) {
                if (pass == 0) {
                    if ((lb[l >> 5] >> (l & 31)) & 1u) j1 = ol[l];
                } else {
                    const int t = nl[l];
                    if (nnb0 == 0 || !lds_contains(ol, nnb0, t)) j1 = t;
                }
            }
            const unsigned long long m = __ballot(j1 != 0);
            if (!m) continue;
            bool ks = false;
            double tv[NTERM];
            const int pos = lanes_below(m);
            if (j1) {
                pair_fj(nb, j1 - 1, xi, vi, tv, tv + 3);
                if (fps) {
                    // a KS component J < IFIRST: c.m. body N + KVEC(J), KVEC(J) = (J + 1) / 2
                    ks = j1 < P.ifirst;
                    const int jcm = ks ? P.n + (j1 + 1) / 2 - 1 : -1;
                    fp_terms(nb, time, i, j1 - 1, jcm, xi, vi, tv + 6, tv + 9);
                }
            }
            cnt[C_CMDER] += __popcll(__ballot(ks));
            // the compacted terms pass through LDS 32 at a time, in list order
            const int nt = __popcll(m);
            for (int h = 0; h < nt; h += TCH) {
                if (j1 && pos >= h && pos < h + TCH) {
#pragma unroll
                    for (int k = 0; k < NTERM; k++)
                        if (k < 6 || fps) T[k * TCH + (pos - h)] = tv[k];
                }
                __builtin_amdgcn_wave_barrier();
                const int nh = min(TCH, nt - h);
                if (lane < (fps ? NTERM : 6)) {
                    const double *Tq = T + lane * TCH;
                    if (pass == 0)
                        for (int t = 0; t < nh; t++) acc = acc - Tq[t];
                    else
                        for (int t = 0; t < nh; t++) acc = acc + Tq[t];
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    double sv2[3], sv3[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        dfirr[k] = __shfl(acc, k);
        dfd[k] = __shfl(acc, 3 + k);
        sv2[k] = __shfl(acc, 6 + k);
        sv3[k] = __shfl(acc, 9 + k);
    }
    if (P.kz38 == 3 && nbloss > 0) nbgain = 0;
    const bool copy = nbloss + nbgain > 0;
    if (copy) cnt[C_FLUX] += nbloss + nbgain;
    if (fp) cnt[C_CORR]++;

    // final list to the regf row and, when it changed, to LIST (regcor_gpu.F:614-622)
    for (int b = 0; b <= nnb; b += WAVE) {
        const int l = b + lane;
        if (l <= nnb) {
            const int v = l == 0 ? nnb : nl[l - 1];
            row[l] = v;
            if (copy) orow[l] = v;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            hd[(H_DFIRR + k) * ni] = dfirr[k];
            hd[(H_DFD + k) * ni] = dfd[k];
            hd[(H_SV2 + k) * ni] = sv2[k];
            hd[(H_SV3 + k) * ni] = sv3[k];
            hd[(H_FREG + k) * ni] = freg[k];
            hd[(H_FDR + k) * ni] = fdr[k];
        }
        hi[HI_NNB * ni] = nnb_pre;
        hi[HI_NNB0 * ni] = nnb0;
        hi[HI_FLAGS * ni] = 1 | (fp ? 2 : 0);
    }
}

// RS, corrector, FPCORR and step of particle I (regcor_gpu.F:95-238, :417-740) from
// the list kernel's hand-over, one thread per particle.
__device__ void regcor_finish(const nb6_regcor_params &P, const nb6_regstate &s, int I,
                              const double *hd, const int *hi, int ni) {
#pragma clang fp contract(off)
    const double time = P.time;
    const int i = I - 1;
    const int flags = hi[HI_FLAGS * ni];
    const int nnb = hi[HI_NNB * ni], nnb0 = hi[HI_NNB0 * ni];
    const bool fp = (flags & 2) != 0;
    double xi[3], vi[3], freg[3], fdr[3], dfirr[3], dfd[3], sv2[3], sv3[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        xi[k] = s.x[3 * (size_t)i + k];
        vi[k] = s.xdot[3 * (size_t)i + k];
        dfirr[k] = hd[(H_DFIRR + k) * ni];
        dfd[k] = hd[(H_DFD + k) * ni];
        sv2[k] = hd[(H_SV2 + k) * ni];
        sv3[k] = hd[(H_SV3 + k) * ni];
        freg[k] = hd[(H_FREG + k) * ni];
        fdr[k] = hd[(H_FDR + k) * ni];
    }
    const double dtr = time - s.t0r[i];
    double ri2, rh2;
    if (P.nb_flag == 1) {
        const double e0 = xi[0] - P.rdens[0], e1 = xi[1] - P.rdens[1], e2 = xi[2] - P.rdens[2];
        ri2 = e0 * e0 + e1 * e1 + e2 * e2;
        rh2 = P.rscale * P.rscale;
    } else {
        ri2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        rh2 = 9.0 * (P.rscale * P.rscale);
    }
    int j3[3] = {0, 0, 0};
    for (int q = 0; q < 3 && q < nnb; q++) j3[q] = hi[(HI_J3 + q) * ni];
    s.rs[i] = rs_update(P, s, I, i, xi, vi, nnb, nnb0, dtr, ri2, rh2, s.rs[i], j3);

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
        const double dfr = s.fr[3 * (size_t)i + k] - freg[k] - dfirr[k];
        const double fdr0 = fdr[k] + dfd[k];
        const double frd = s.frdot[3 * (size_t)i + k];
        const double sm = frd + fdr0;
        const double at3 = 2.0 * dfr + dtr * sm;
        const double bt2 = -3.0 * dfr - dtr * (sm + frd);
        s.xn[3 * (size_t)i + k] = s.x0[3 * (size_t)i + k] + ((double)0.6f * at3 + bt2) * dtsq12;
        xnd[k] = s.x0dot[3 * (size_t)i + k] + (0.75 * at3 + bt2) * dtr13;
        s.xndot[3 * (size_t)i + k] = xnd[k];
        const double fik = s.fi[3 * (size_t)i + k] + dfirr[k];
        s.fi[3 * (size_t)i + k] = fik;
        s.fr[3 * (size_t)i + k] = freg[k];
        s.fidot[3 * (size_t)i + k] = s.fidot[3 * (size_t)i + k] + dfd[k];
        s.frdot[3 * (size_t)i + k] = fdr[k];
        if (k == 1) s.fidot[3 * (size_t)i] = s.fidot[3 * (size_t)i] - P.tidal4 * dfr;
        if (k == 0) s.fidot[3 * (size_t)i + 1] = s.fidot[3 * (size_t)i + 1] + P.tidal4 * dfr;
        s.d0[3 * (size_t)i + k] = fik;
        s.d0r[3 * (size_t)i + k] = freg[k];
        s.d1r[3 * (size_t)i + k] = fdr[k];
        d3r[k] = at3 * dt6;
        d2r[k] = (3.0 * at3 + bt2) * dt2;
    }
    if (fp) {                                         // fpcorr.F:93-97
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s.d2[3 * (size_t)i + k] = s.d2[3 * (size_t)i + k] + sv2[k];
            s.d3[3 * (size_t)i + k] = s.d3[3 * (size_t)i + k] + sv3[k];
            d2r[k] = d2r[k] - sv2[k];
            d3r[k] = d3r[k] - sv3[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s.d2r[3 * (size_t)i + k] = d2r[k];
        s.d3r[3 * (size_t)i + k] = d3r[k];
    }

    // new regular step, regcor_gpu.F:624-740
    const double fr2 = freg[0] * freg[0] + freg[1] * freg[1] + freg[2] * freg[2];
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
        const double q = 1.0e-4 * (dtr * dtr) * fr2 / dv2;
        const double dtj = dtr * pow(q, (double)0.1f);
        if (dtj < ttmp) ttmp = dtr;
    }
    ttmp = fmin(ttmp, smax);
    s.stepr[i] = ttmp;
    if (nnb0 == 0) {
        s.step[i] = 0.125 * stp;
        s.timenw[i] = s.t0[i] + s.step[i];
    }
}

// CHECKL = false (KZ(37) != 1, or no LISTV): the kernel without the high-velocity test,
// whose registers would halve its occupancy
template <bool CHECKL>
__global__ __launch_bounds__(256) void regcor_lists_kernel(nb6_regcor_params P, nb6_regstate s,
                                                           const double *__restrict__ pk,
                                                           int ni, const int *__restrict__ addr,
                                                           const double *__restrict__ force,
                                                           int fstride, int *__restrict__ nlist,
                                                           int *__restrict__ counters,
                                                           double *__restrict__ hd,
                                                           int *__restrict__ hi) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int red[NCOUNT];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    if (threadIdx.x < NCOUNT) red[threadIdx.x] = 0;
    __syncthreads();
    const int k = blockIdx.x * (blockDim.x / WAVE) + w;
    int cnt[NCOUNT] = {};
    if (k < ni) {
        unsigned char *base = lds + (size_t)w * wave_lds_bytes(s.lmax);
        double *T = reinterpret_cast<double *>(base);
        int *nl = reinterpret_cast<int *>(base + (size_t)NTERM * TCH * 8);
        int *ol = nl + (s.lmax + 1);
        unsigned *lb = reinterpret_cast<unsigned *>(ol + (s.lmax + 1));
        unsigned *cb = lb + loss_words(s.lmax);
        if (lane == 0) hi[HI_FLAGS * ni + k] = 0;
        regcor_wave<CHECKL>(P, s, pk, addr[k], force + k, fstride, nlist + (size_t)k * s.lmax, cnt, T, nl,
                    ol, lb, cb, lane, hd + k, hi + k, ni);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NCOUNT; q++)
            if (cnt[q]) atomicAdd(&red[q], cnt[q]);
    }
    __syncthreads();
    if (threadIdx.x < NCOUNT && red[threadIdx.x]) atomicAdd(&counters[threadIdx.x], red[threadIdx.x]);
}

__global__ __launch_bounds__(256) void regcor_finish_kernel(nb6_regcor_params P, nb6_regstate s,
                                                            int ni, const int *__restrict__ addr,
                                                            const double *__restrict__ hd,
                                                            const int *__restrict__ hi) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < ni && (hi[HI_FLAGS * ni + k] & 1)) regcor_finish(P, s, addr[k], hd + k, hi + k, ni);
}

DBuf g_hand, g_pack;
constexpr size_t PACK_RATIO = 32;   // pack when ni >= nmax / 32
int g_pack_mode = 0;                // 0: by block size, 1: always, 2: never

}  // namespace

int regcor_set_pack(int mode) {
    if (mode < 0 || mode > 2) { set_error(E_ARG, "nb6_set_regcor_pack: mode must be 0, 1 or 2"); return E_ARG; }
    g_pack_mode = mode;
    return OK;
}

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
    if (p->ifirst < 1 || p->n < 0 || p->kz38 < 0 || p->kz38 > 3 || p->kz37 < 0 ||
        !(p->time == p->time)) {
        set_error(E_ARG, "nb6_regcor_device: bad IFIRST, N, KZ(37), KZ(38) or TIME");
        return E_ARG;
    }
    if (p->kz37 == 1 && s->listv && p->n < s->nmax && !s->r) {
        set_error(E_ARG, "nb6_regcor_device: KZ(37) = 1 with c.m. bodies needs the KS separations r");
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
    if (!stream) stream = d.stream;
    KTimer &kt = ktimer("regcor");
    kt.start(stream);
    // one wavefront per particle, up to four per work-group within the LDS of a CU
    const size_t wl = wave_lds_bytes(s->lmax);
    const int wpb = (int)std::min<size_t>(4, (160 * 1024) / wl);
    if (wpb < 1) { set_error(E_ARG, "nb6_regcor_device: lmax too large for LDS"); return E_ARG; }
    const size_t lds = wl * wpb;
    static size_t lds_set = 0;
    if (lds > lds_set) {
        NB6_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(regcor_lists_kernel<false>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        NB6_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(regcor_lists_kernel<true>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        lds_set = lds;
    }
    char *h = static_cast<char *>(g_hand.get((size_t)ni * (NHD * 8 + NHI * 4)));
    double *hd = reinterpret_cast<double *>(h);
    int *hi = reinterpret_cast<int *>(h + (size_t)ni * NHD * 8);
    // large blocks: one packed record per particle first (a pass over all nmax
    // particles, about 300 bytes each), so each gathered neighbour costs two cache
    // lines instead of eight; small blocks gather from the state arrays directly
    double *pk = nullptr;
    if (g_pack_mode == 1 || (g_pack_mode == 0 && (size_t)ni * PACK_RATIO >= (size_t)s->nmax)) {
        pk = static_cast<double *>(g_pack.get((size_t)s->nmax * PK_STRIDE * sizeof(double)));
        hipLaunchKernelGGL(regcor_pack_kernel, dim3((s->nmax + 255) / 256), dim3(256), 0, stream,
                           *s, pk);
        NB6_HIP(hipGetLastError());
    }
    if (p->kz37 == 1 && s->listv)
        hipLaunchKernelGGL(regcor_lists_kernel<true>, dim3((ni + wpb - 1) / wpb), dim3(WAVE * wpb),
                           lds, stream, *p, *s, pk, ni, d_addr, d_force, force_stride, d_nlist,
                           d_counters, hd, hi);
    else
        hipLaunchKernelGGL(regcor_lists_kernel<false>, dim3((ni + wpb - 1) / wpb), dim3(WAVE * wpb),
                           lds, stream, *p, *s, pk, ni, d_addr, d_force, force_stride, d_nlist,
                           d_counters, hd, hi);
    NB6_HIP(hipGetLastError());
    hipLaunchKernelGGL(regcor_finish_kernel, dim3((ni + 255) / 256), dim3(256), 0, stream, *p, *s,
                       ni, d_addr, hd, hi);
    NB6_HIP(hipGetLastError());
    kt.stop(stream);
    return OK;
}

}  // namespace nb6
