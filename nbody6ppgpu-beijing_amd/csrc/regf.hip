// Regular force + neighbour list for gfx950.
//
// Semantics (what is computed) follow the reference:
//   pair terms      gpunb.velocity.cu:163-209 / :212-258, reg.avx.cpp:250-305
//   neighbour test  min(r2, |dx + dtr dv|^2) < h2 (or < m_j h2)   gpunb.velocity.cu:188,236
//   neighbours add nothing to acc / jerk / pot                     gpunb.velocity.cu:192
//   list rows       [0] = count, ascending 0-based j, -count on overflow  reg.avx.cpp:316-327
//
// Method (how it is computed) is this library's own:
//
//  * The j-set is Morton-ordered at send time (jset.hip) and streamed as FP32
//    records relative to the origin of their 256-particle tile.  regf_tile_kernel
//    (one-dimensional grid of (j-chunk, i-block) items, chunk-major, each XCD taking a
//    contiguous run; T = 128, 256 or 512 threads, IPL i-particles per lane; chunks are
//    added until a call has min_groups work-groups, down to one 256- or 512-record
//    unit per chunk) stages T records per step in LDS by direct global-to-LDS loads
//    (double-buffered), forms each lane's i-particle relative to the same tile origin
//    in FP64, and evaluates every pair in FP32: v_rsq_f32 refined by one
//    Newton-Raphson step, FP32 partial sums flushed into FP64 every 64 j.
//
//  * Candidates.  A pair whose FP32 velocity criterion passes the neighbour
//    threshold widened by its rounding-error band is a "hard" candidate; a pair
//    inside twice the neighbour radius is a "soft" (near-field) candidate.  Both
//    are left out of the FP32 sums and their sorted j index goes to per-(i, chunk)
//    slot arrays: hard candidates always (they contain every FP64 neighbour),
//    soft ones while slots last (a soft pair without a slot simply stays in the
//    FP32 sum).  The near field is where the 1/r^3-weighted jerk terms amplify
//    FP32 rounding.  The test is two-stage: a cheap, conservative FP32 bound
//    (r2p >= r2 + 2 dtr min(rv, 0)) guards a rarely-taken branch with the full test.
//    A row whose hard slots fill up overflows; past that point a hard candidate
//    below the band's lower edge is a neighbour for certain (no force, counted),
//    and one inside the band takes a soft slot marked ~j so that the FP64 test
//    still decides it.  Overflowing rows thus keep their exact count and sums, as
//    the reference computes them; if even the soft slots run out the row's
//    forces are NaN.
//
//  * regf_finalize_kernel: one wavefront per i (regf_finalize_wide_kernel: eight for at
//    most 16384 i in 16 chunks or more, four for 32 chunks or more).  Sums the FP64 chunk partials,
//    re-tests every hard candidate in FP64 with the reference formula (no FMA
//    contraction), adds the FP64 force of every non-neighbour candidate, maps
//    neighbours back to the caller's indices and writes each row in ascending
//    order (rank sort in LDS).
#include "nb6_internal.h"

#include <algorithm>
#include <cmath>

namespace nb6 {

namespace {

constexpr int RF_FLUSH = 64;          // j per FP32 partial sum
constexpr int NACC = 7;               // FP32 partials: acc x y z, pot, jerk x y z
// FP64 partial rows per (i, chunk): acc x y z, jerk x y z, pot, then the four error
// scales of the FP32 sums (see tile_work)
constexpr int NERR = 4;
constexpr int NPART = 7 + NERR;
constexpr int CAP_SOFT = 1024;        // soft (near-field) slots per (i, chunk)
constexpr int BAND_SLOTS = 16;        // of which kept for band pairs of overflowing rows
constexpr int NB_SORT_MAX = 2048;     // largest neighbour row the finalize sorts
constexpr int FIN_WIDE_NI = 16384;    // i-set size below which 16 chunks take the wide finalize

// rounding-error margins (u = 2^-24)
constexpr double U24 = 5.9604644775390625e-08;
constexpr float KAPPA_P = 1.9073486328125e-06f;   // 2^-19 = 32u : full test
constexpr float KAPPA_Q = 3.814697265625e-06f;    // 2^-18 = 64u : cheap test
constexpr double H_REL = 64.0 * U24;              // relative slack on h2
constexpr float D_ABS = (float)(64.0 * U24);      // absolute slack factor (x2 spare)
constexpr float Q_REL = (float)(192.0 * U24);     // cheap threshold over the full one
                                                  // (includes 1/(1-KAPPA_Q))
constexpr float TINY = 1.1754944e-38f;            // keeps r2 = 0 a candidate

struct TileParams {
    const double *__restrict__ ib;   // SoA [8][ib_stride]: x y z vx vy vz h2 dtr
    int ni, ib_stride;               // ni = slot rows of the work arrays
    const int *__restrict__ imap;    // slot -> row of ib (nullptr: identity; < 0: empty slot)
    const float4 *__restrict__ rec32;
    const double *__restrict__ torg;
    const float *__restrict__ tbox;  // [ntile][TBOX] per-tile bounds
    int far_test;                    // 0: always run the pair test
    const double *__restrict__ rec64;   // sorted AoS [nj][8] (precision variants)
    int nj;
    int chunk_len;                   // multiple of the block size T and of JTILE
    int nib;                         // i-blocks of the grid (regf_tile_kernel)
    double mmax;
    double *__restrict__ part;       // [nchunk][NPART][ni]
    int *__restrict__ cnt;           // [nchunk][2][ni]: hard, soft
    int *__restrict__ hslot;         // [(chunk*ni + i) * caph + k]
    int caph;
    int *__restrict__ sslot;         // [(chunk*ni + i) * caps + k]
    int caps;
    float fnear;                     // near field: r < fnear RS (soft candidates)
    int nseg_c;                      // partial-sum segments per chunk (part rows chunk * nseg_c
    int seg_tiles;                   // + s), each seg_tiles T-tiles long (0: the whole chunk)
};

// per-lane constants that do not depend on the tile
struct ILane {
    double x, y, z, u, v, w;         // FP64 position / velocity
    float dt1, dt2;                  // dtr, 2 dtr
    float h2H, h2S;                  // (1+slack) h2 and fnear^2 (1+slack) h2 (MFLAG: slopes)
    float h2L;                       // (1-slack) h2 (MFLAG: slope)
    float rsH, rsS;                  // radii used in the absolute error bands
    bool ok;
};

// per-lane constants relative to one tile origin
struct IC {
    float x, y, z, u, v, w;          // i relative to the tile origin
    float dt1, dt2;
    float hH, dH, hS, dS, hQ, dQ;    // thresholds (MFLAG: slopes h*, offsets d*)
    float hL;                        // below it a pair is an FP64 neighbour for certain
                                     // (MFLAG: slope, offset -2 dH)
    float eP, eV;                    // bounds of |x_i - P| + |x_j - P| and of
                                     // |v_i - V| + |v_j - V| over the tile's records
};

template <bool MFLAG>
__device__ __forceinline__ ILane make_lane(const double *__restrict__ ib, int is, int i, bool ok,
                                           double mmax, float fnear) {
    ILane L;
    const int ii = ok ? i : 0;
    L.x = ib[0 * is + ii]; L.y = ib[1 * is + ii]; L.z = ib[2 * is + ii];
    L.u = ib[3 * is + ii]; L.v = ib[4 * is + ii]; L.w = ib[5 * is + ii];
    const double h2 = ib[6 * is + ii], dtr = ib[7 * is + ii];
    L.dt1 = (float)dtr;
    // r2 + 2 dtr min(rv,0) <= (r2 - kq r2 + 2 dtr rv)/(1 - kq): the cheap test uses
    // 2 dtr/(1-kq) here and thresholds divided by (1-kq) (folded into Q_REL)
    L.dt2 = (float)(2.0 * dtr / (1.0 - (double)KAPPA_Q));
    const double hh = h2 > 0.0 ? h2 : 0.0;
    const double hk = hh * (1.0 + H_REL) * (1.0 + 4.0 * U24);
    L.h2H = (float)hk;
    L.h2S = (float)(hk * (double)(fnear * fnear));
    L.h2L = (float)(hh * (1.0 - H_REL) * (1.0 - 12.0 * U24));
    const double rs = MFLAG ? sqrt(hh * mmax) : sqrt(hh);
    L.rsH = (float)(rs * (1.0 + 4.0 * U24));
    L.rsS = fnear * L.rsH;
    L.ok = ok;
    return L;
}

template <bool MFLAG>
__device__ __forceinline__ IC tile_ic(const ILane &L, const double *__restrict__ T,
                                      const float *__restrict__ TB) {
    IC c;
    c.x = (float)(L.x - T[0]); c.y = (float)(L.y - T[1]); c.z = (float)(L.z - T[2]);
    c.u = (float)(L.u - T[3]); c.v = (float)(L.v - T[4]); c.w = (float)(L.w - T[5]);
    c.dt1 = L.dt1;
    c.dt2 = L.dt2;   // 2 dtr / (1 - kappa_q), see make_lane
    // |x_i - P|, |v_i - V| bound the rounding error of the FP32 separations
    const float A = sqrtf(fmaf(c.z, c.z, fmaf(c.y, c.y, c.x * c.x))) * 1.0001f;
    const float B = sqrtf(fmaf(c.w, c.w, fmaf(c.v, c.v, c.u * c.u))) * 1.0001f;
    const float g = A + fabsf(c.dt1) * B;
    const float dH = D_ABS * L.rsH * (g + L.rsH) + TINY;
    const float dS = D_ABS * L.rsS * (g + L.rsS) + TINY;
    // the tile's records: |x_j - P| and |v_j - V| are at most the largest corner of
    // its position and velocity boxes (per component)
    float xm = 0.f, vm = 0.f;
    if (TB[13] > 0.f) {                  // a tile without records has no bounds
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float xq = fmaxf(fabsf(TB[q]), fabsf(TB[3 + q]));
            const float vq = fmaxf(fabsf(TB[6 + q]), fabsf(TB[9 + q]));
            xm = fmaf(xq, xq, xm);
            vm = fmaf(vq, vq, vm);
        }
    }
    c.eP = A + sqrtf(xm) * 1.0001f;
    c.eV = B + sqrtf(vm) * 1.0001f;
    if (MFLAG) {
        c.hH = L.h2H; c.dH = dH;
        c.hS = L.h2S; c.dS = dS;
        c.hQ = L.h2S * (1.0f + Q_REL); c.dQ = dS * (1.0f + Q_REL);
        c.hL = L.h2L;
    } else {
        c.hH = L.h2H + dH; c.dH = 0.f;
        c.hS = L.h2S + dS; c.dS = 0.f;
        c.hQ = c.hS * (1.0f + Q_REL); c.dQ = 0.f;
        c.hL = L.h2L - 2.0f * dH;
    }
    if (!L.ok) {
        c.hH = c.hS = c.hQ = c.hL = -INFINITY;
        c.dH = c.dS = c.dQ = -INFINITY;
    }
    return c;
}

// True when no pair of this lane's i with the 256-j tile whose bounds are B can be a
// hard or soft candidate, so the per-pair test can be skipped for the whole tile.
// The loop records a candidate only if min(r2, r2p - KAPPA_P r2) < ths, with
// r2 = |dx|^2 and r2p = |dx + dtr dv|^2 formed from FP32 tile-relative values.
// Every such dx lies in the position box and dx + dtr dv in the box of predicted
// positions (lo + dtr vlo, hi + dtr vhi), per component, so |dx| >= D and
// |dx + dtr dv| >= Dp (FP32 rounding is monotonic).  The test asks
// min(D, Dp)^2 - 2^-16 X^2 >= (1 + 2^-8) ths with X = Dmax + dtr Wmax >= |dx|, |dtr dv|;
// its margins exceed KAPPA_P X^2 plus the FP32 rounding of r2, r2p and of the bound.
template <bool MFLAG>
__device__ __forceinline__ bool lane_far(const IC &c, bool ok, const float *__restrict__ B) {
    if (!ok || B[13] == 0.f) return true;          // empty slot or tile without records
    const float dt = c.dt1;
    if (!(dt >= 0.f)) return false;
    const float ci[3] = {c.x, c.y, c.z};
    const float vi[3] = {c.u, c.v, c.w};
    float dd = 0.f, dp = 0.f, dm = 0.f, ww = 0.f;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float lo = B[q], hi = B[3 + q], vlo = B[6 + q], vhi = B[9 + q];
        const float e = fmaxf(fmaxf(lo - ci[q], ci[q] - hi), 0.f);
        dd = fmaf(e, e, dd);
        const float cp = fmaf(dt, vi[q], ci[q]);
        const float ep = fmaxf(fmaxf(fmaf(dt, vlo, lo) - cp, cp - fmaf(dt, vhi, hi)), 0.f);
        dp = fmaf(ep, ep, dp);
        const float f = fmaxf(fabsf(lo - ci[q]), fabsf(hi - ci[q]));
        dm = fmaf(f, f, dm);
        const float g = fmaxf(fabsf(vlo - vi[q]), fabsf(vhi - vi[q]));
        ww = fmaf(g, g, ww);
    }
    const float X = sqrtf(dm) + dt * sqrtf(ww);
    const float ths = MFLAG ? fmaf(B[12], c.hS, c.dS) : c.hS;
    return fmaf(-0x1p-16f * X, X, fminf(dd, dp)) >= ths * (1.0f + 0x1p-8f);
}

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void glb_void;

// j record beyond the end of a range: far and massless, never a candidate, adds zero
__device__ const float4 g_far_rec[2] = {{1.0e17f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};

// wait for this wavefront's outstanding global loads, LDS DMA included
__device__ __forceinline__ void lds_dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

struct PairBase {
    float dx, dy, dz, du, dv, dw, r2, rv;
};

__device__ __forceinline__ PairBase pair_base(const IC &c, const float4 &a, const float4 &b) {
    PairBase q;
    q.dx = a.x - c.x; q.dy = a.y - c.y; q.dz = a.z - c.z;
    q.du = b.x - c.u; q.dv = b.y - c.v; q.dw = b.z - c.w;
    q.r2 = fmaf(q.dz, q.dz, fmaf(q.dy, q.dy, q.dx * q.dx));
    q.rv = fmaf(q.dz, q.dw, fmaf(q.dy, q.dv, q.dx * q.du));
    return q;
}

// One j against the IPL i-particles of this lane.  TEST=false leaves out the
// candidate test: only used where lane_far holds for every lane of the wavefront,
// where the test could not have recorded anything, so both forms give the same sums.
template <bool MFLAG, bool TEST>
__device__ __forceinline__ void pair_step(const TileParams &p, const IC &ck, const float4 a,
                                          const float4 b, int jg, int j1, int &nhk, int &nsk,
                                          int *__restrict__ hpk, int *__restrict__ spk,
                                          float *__restrict__ acc, float &s3) {
    {
        const PairBase q = pair_base(ck, a, b);
        // v_rsq_f32 plus one Newton-Raphson step
        float rinv = __builtin_amdgcn_rsqf(q.r2);
        {
            const float h = rinv * rinv;
            const float e = fmaf(-q.r2, h, 1.0f);
            rinv = fmaf(rinv, 0.5f * e, rinv);
        }
        if (TEST) {
            // cheap conservative bound on the soft test:
            // r2p >= r2 + 2 dtr min(rv, 0)  (dt2 carries the 1/(1-kappa) slack)
            const float t1 = fmaf(ck.dt2, fminf(q.rv, 0.f), q.r2);
            const float thq = MFLAG ? fmaf(a.w, ck.hQ, ck.dQ) : ck.hQ;
            if (t1 < thq) {
                // keep the full test inside the rare branch: the empty asm
                // makes dtr opaque here, so the compiler cannot hoist it
                float dtl = ck.dt1;
                asm volatile("" : "+v"(dtl));
                const float ex = fmaf(dtl, q.du, q.dx);
                const float ey = fmaf(dtl, q.dv, q.dy);
                const float ez = fmaf(dtl, q.dw, q.dz);
                const float r2p = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                const float mn = fminf(q.r2, fmaf(-KAPPA_P, q.r2, r2p));
                const float thh = MFLAG ? fmaf(a.w, ck.hH, ck.dH) : ck.hH;
                const float ths = MFLAG ? fmaf(a.w, ck.hS, ck.dS) : ck.hS;
                if (mn < thh && jg < j1) {
                    if (nhk < p.caph) {
                        hpk[nhk] = jg;
                    } else {
                        // hard slots full (the row overflows): a pair below the lower
                        // band is a neighbour for certain and stays out of every sum;
                        // one inside the band goes to a soft slot marked ~j, where the
                        // FP64 test decides it (the last BAND_SLOTS slots are kept for
                        // these; the count runs on past caps if even they fill up)
                        const float thl = MFLAG ? fmaf(a.w, ck.hL, -2.0f * ck.dH) : ck.hL;
                        if (!(fmaf(KAPPA_P, q.r2, mn) < thl)) {
                            if (nsk < p.caps) spk[nsk] = ~jg;
                            nsk++;
                        }
                    }
                    nhk++;
                    rinv = 0.f;
                } else if (mn < ths && jg < j1 && nsk < p.caps - BAND_SLOTS) {
                    spk[nsk] = jg;
                    nsk++;
                    rinv = 0.f;
                }
            }
        }
        // jerk term m/r^3 (dv - 3 (dx.dv / r^2) dx) summed whole: its two parts partly
        // cancel within a pair, so the FP32 partial sum stays smaller than the sums of
        // m dv / r^3 and of 3 m (dx.dv) dx / r^5 would be apart
        const float rinv2 = rinv * rinv;
        const float mr1 = a.w * rinv;
        const float mr3 = mr1 * rinv2;
        const float al = q.rv * (-3.0f * rinv2);
        acc[0] = fmaf(mr3, q.dx, acc[0]);
        acc[1] = fmaf(mr3, q.dy, acc[1]);
        acc[2] = fmaf(mr3, q.dz, acc[2]);
        acc[3] += mr1;
        acc[4] = fmaf(mr3, fmaf(al, q.dx, q.du), acc[4]);
        acc[5] = fmaf(mr3, fmaf(al, q.dy, q.dv), acc[5]);
        acc[6] = fmaf(mr3, fmaf(al, q.dz, q.dw), acc[6]);
        // sum of m/r^3: times the tile's bounds of the FP32 rounding of positions
        // and velocities it bounds the rounding error of the pair terms
        s3 += mr3;
    }
}

// FP64 partial sums and error scales of this lane's i-particles to part row `row`
// (of the chunk's partial-sum segments), then zeroed for the next segment
template <int T, int IPL>
__device__ __forceinline__ void write_part(const TileParams &p, int ibase, int iend, int row,
                                           double (*a64)[7], float (*err)[NERR]) {
#pragma unroll
    for (int k = 0; k < IPL; k++) {
        const int i = ibase + k * T + (int)threadIdx.x;
        if (i < iend) {
            double *pp = p.part + (size_t)row * NPART * p.ni + i;
#pragma unroll
            for (int f = 0; f < 7; f++) pp[(size_t)f * p.ni] = a64[k][f];
#pragma unroll
            for (int f = 0; f < NERR; f++) pp[(size_t)(7 + f) * p.ni] = (double)err[k][f];
        }
#pragma unroll
        for (int f = 0; f < 7; f++) a64[k][f] = 0.0;
#pragma unroll
        for (int f = 0; f < NERR; f++) err[k][f] = 0.f;
    }
}

// One work item: i in [ibase, min(ibase + T*IPL, iend)) against sorted j in
// [j0, j1) (a whole number of tiles), partial-sum row = chunk * ni + i.
template <int T, int IPL, bool MFLAG>
__device__ __forceinline__ void tile_work(const TileParams &p, int ibase, int iend, int j0,
                                          int j1, int chunk, float4 (*sA)[T], float4 (*sB)[T]) {
    const int tid = threadIdx.x;

    ILane L[IPL];
    int nh[IPL], ns[IPL];
    int *hp[IPL], *sp[IPL];
    float a32[IPL][NACC];
    double a64[IPL][7];
    float err[IPL][NERR];            // error scales of the FP32 sums (FP32 is enough)
    float s3[IPL];                   // sum of m/r^3 over the current 64 j

#pragma unroll
    for (int k = 0; k < IPL; k++) {
        const int i = ibase + k * T + tid;   // slot
        const bool in = i < iend;
        int r = i;
        if (in && p.imap) r = p.imap[i];
        const bool ok = in && r >= 0;
        L[k] = make_lane<MFLAG>(p.ib, p.ib_stride, r, ok, p.mmax, p.fnear);
        nh[k] = 0;
        ns[k] = 0;
        const size_t row = (size_t)chunk * p.ni + (in ? i : 0);
        hp[k] = p.hslot + row * p.caph;
        sp[k] = p.sslot + row * p.caps;
#pragma unroll
        for (int f = 0; f < NACC; f++) a32[k][f] = 0.f;
#pragma unroll
        for (int f = 0; f < 7; f++) a64[k][f] = 0.0;
#pragma unroll
        for (int f = 0; f < NERR; f++) err[k][f] = 0.f;
        s3[k] = 0.f;
    }

    const int ntile = (j1 - j0 + T - 1) / T;
    int segw = 0;                    // partial-sum segment being summed

    // Tile t goes straight from HBM into LDS buffer t & 1 (global_load_lds_dwordx4:
    // lane l of a wavefront writes 16 bytes at the wavefront's LDS base + 16 l), so the
    // next tile's records occupy no registers while the current one is evaluated.
    // Lanes past j1 read the far, massless dummy record.
    const int w0 = tid & ~63;
    auto gload = [&](int t, int b) {
        const int j = j0 + t * T + tid;
        const float4 *src = j < j1 ? p.rec32 + 2 * (size_t)j : g_far_rec;
        __builtin_amdgcn_global_load_lds((glb_void *)src, (lds_void *)&sA[b][w0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((glb_void *)(src + 1), (lds_void *)&sB[b][w0], 16, 0, 0);
    };

    gload(0, 0);
    lds_dma_wait();
    __syncthreads();

    for (int t = 0; t < ntile; t++) {
        const int buf = t & 1;
        if (t + 1 < ntile) gload(t + 1, buf ^ 1);
        const int jt = j0 + t * T;
        IC c[IPL];
        bool allfar = false;
        for (int js0 = 0; js0 < T; js0 += RF_FLUSH) {
            if (js0 == 0 || (T > JTILE && js0 % JTILE == 0)) {
                // i relative to the origin of the 256-j tile this segment lies in,
                // and whether the whole wavefront is outside its near field
                const size_t tt = (size_t)((jt + js0) / JTILE);
                bool far = p.far_test != 0;
#pragma unroll
                for (int k = 0; k < IPL; k++) {
                    c[k] = tile_ic<MFLAG>(L[k], p.torg + tt * 8, p.tbox + tt * TBOX);
                    if (p.far_test) far = far && lane_far<MFLAG>(c[k], L[k].ok, p.tbox + tt * TBOX);
                }
                allfar = __ballot(!far) == 0ull;
            }
            if (allfar) {
#pragma unroll 4
                for (int jj = js0; jj < js0 + RF_FLUSH; jj++) {
                    const float4 a = sA[buf][jj], b = sB[buf][jj];
#pragma unroll
                    for (int k = 0; k < IPL; k++)
                        pair_step<MFLAG, false>(p, c[k], a, b, jt + jj, j1, nh[k], ns[k], hp[k],
                                                sp[k], a32[k], s3[k]);
                }
            } else {
#pragma unroll 2
                for (int jj = js0; jj < js0 + RF_FLUSH; jj++) {
                    const float4 a = sA[buf][jj], b = sB[buf][jj];
#pragma unroll
                    for (int k = 0; k < IPL; k++)
                        pair_step<MFLAG, true>(p, c[k], a, b, jt + jj, j1, nh[k], ns[k], hp[k],
                                               sp[k], a32[k], s3[k]);
                }
            }
            // flush FP32 partials into FP64
#pragma unroll
            for (int k = 0; k < IPL; k++) {
                a64[k][0] += (double)a32[k][0];
                a64[k][1] += (double)a32[k][1];
                a64[k][2] += (double)a32[k][2];
                a64[k][3] += (double)a32[k][4];
                a64[k][4] += (double)a32[k][5];
                a64[k][5] += (double)a32[k][6];
                a64[k][6] += (double)a32[k][3];
                // L1 norms of the flushed FP32 partials: the scale of their rounding;
                // sum of m/r^3 times the position and velocity rounding bounds
                err[k][0] += fabsf(a32[k][0]) + fabsf(a32[k][1]) + fabsf(a32[k][2]);
                err[k][1] += fabsf(a32[k][4]) + fabsf(a32[k][5]) + fabsf(a32[k][6]);
                err[k][2] = fmaf(s3[k], c[k].eP, err[k][2]);
                err[k][3] = fmaf(s3[k], c[k].eV, err[k][3]);
                s3[k] = 0.f;
#pragma unroll
                for (int f = 0; f < NACC; f++) a32[k][f] = 0.f;
            }
        }

        lds_dma_wait();   // this wavefront's part of tile t + 1 is in LDS
        __syncthreads();
        if (p.seg_tiles > 0 && (t + 1) % p.seg_tiles == 0 && t + 1 < ntile) {
            // end of a partial-sum segment: its sums to their own row, then afresh
            write_part<T, IPL>(p, ibase, iend, chunk * p.nseg_c + segw, a64, err);
            segw++;
        }
    }
    // the last segment, and empty segments past the end of a short chunk
    write_part<T, IPL>(p, ibase, iend, chunk * p.nseg_c + segw, a64, err);
    for (segw++; segw < p.nseg_c; segw++)
        write_part<T, IPL>(p, ibase, iend, chunk * p.nseg_c + segw, a64, err);

#pragma unroll
    for (int k = 0; k < IPL; k++) {
        const int i = ibase + k * T + tid;
        if (i < iend) {
            p.cnt[((size_t)chunk * 2 + 0) * p.ni + i] = nh[k];
            p.cnt[((size_t)chunk * 2 + 1) * p.ni + i] = ns[k];
        }
    }
}

// waves per SIMD the register budget is sized for, at one i per lane: 6 (<= 80 VGPRs;
// the few spills are set-up values reloaded once per tile) for 512-wide work-groups,
// measured 5 % faster on the N = 2^20 step than 5; 5 (<= 96 VGPRs, no spills) for
// narrower ones, where 6 measured slower
template <int T, int IPL>
constexpr int regf_waves() { return IPL == 1 ? (T >= 512 ? 6 : 5) : 2; }

template <int T, int IPL, bool MFLAG>
__global__ __launch_bounds__(T, (regf_waves<T, IPL>())) void regf_tile_kernel(TileParams p) {
    __shared__ float4 sA[2][T];   // x', y', z', m
    __shared__ float4 sB[2][T];   // vx', vy', vz', 0
    // (chunk, i-block) items chunk-major, each XCD taking a contiguous run: the
    // work-groups resident on one XCD stream the same j-chunk through its L2
    const int w = xcd_order(blockIdx.x, gridDim.x);
    const int chunk = w / p.nib;
    const int ib = w - chunk * p.nib;
    const int j0 = chunk * p.chunk_len;
    tile_work<T, IPL, MFLAG>(p, ib * (T * IPL), p.ni, j0, min(j0 + p.chunk_len, p.nj), chunk, sA,
                             sB);
}

// Persistent form for a batch of independent clusters: each work-group takes the
// next work item from a queue (items are ordered longest first) until none is left.
template <int T, int IPL, bool MFLAG>
__global__ __launch_bounds__(T, (regf_waves<T, IPL>())) void regf_items_kernel(TileParams p,
                                                       const int4 *__restrict__ items,
                                                       int nitems, int *__restrict__ queue) {
    __shared__ float4 sA[2][T];
    __shared__ float4 sB[2][T];
    __shared__ int it;
    for (;;) {
        if (threadIdx.x == 0) it = atomicAdd(queue, 1);
        __syncthreads();
        const int k = __builtin_amdgcn_readfirstlane(it);
        __syncthreads();
        if (k >= nitems) break;
        const int4 w = items[k];   // ibase, iend, j0, j1 (chunk 0)
        tile_work<T, IPL, MFLAG>(p, w.x, w.y, w.z, w.w, 0, sA, sB);
    }
}

// ------------------------------------------------------------------ precision variants
// The same pipeline (chunks, slots, finalize) with the pair arithmetic in float-float
// (PREC = 1: every FP64 input split into an FP32 hi/lo pair, error-free sums and
// products, one float-float Newton-Raphson step on v_rsq_f32) or in FP64 (PREC = 2:
// the oracle's formulas).  They exist for the precision study: each pair's neighbour
// test runs in the variant's arithmetic, every pair near the threshold is recorded as a
// candidate and left out of the sums, and the finalize stage decides and adds those in
// FP64 exactly as for the FP32 kernel, so the lists are the exact FP64 lists.
struct ff {
    float h, l;
};

__device__ __forceinline__ ff ff_two_sum(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    const float bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ ff ff_fast(float a, float b) {
#pragma clang fp contract(off)
    const float s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ ff ff_add(ff x, ff y) {
#pragma clang fp contract(off)
    const ff s = ff_two_sum(x.h, y.h);
    return ff_fast(s.h, s.l + (x.l + y.l));
}
__device__ __forceinline__ ff ff_neg(ff x) { return {-x.h, -x.l}; }
__device__ __forceinline__ ff ff_mul(ff x, ff y) {
#pragma clang fp contract(off)
    const float p = x.h * y.h;
    const float e = fmaf(x.h, y.h, -p);
    return ff_fast(p, e + (x.h * y.l + x.l * y.h));
}
__device__ __forceinline__ ff ff_of(double d) {
    const float h = (float)d;
    return {h, (float)(d - (double)h)};
}
__device__ __forceinline__ double ff_d(ff x) { return (double)x.h + (double)x.l; }

template <int PREC, bool MFLAG>
__global__ __launch_bounds__(256) void regf_var_kernel(TileParams p) {
    __shared__ double4 sA[256], sB[256];
    const int tid = threadIdx.x;
    const int chunk = blockIdx.y;
    const int j0 = chunk * p.chunk_len;
    const int j1 = min(j0 + p.chunk_len, p.nj);
    const int i = blockIdx.x * 256 + tid;   // slot
    const bool in = i < p.ni;
    int r = i;
    if (in && p.imap) r = p.imap[i];
    const bool ok = in && r >= 0;
    const int rr = ok ? r : 0;
    const size_t is = p.ib_stride;
    const double xi = p.ib[rr], yi = p.ib[is + rr], zi = p.ib[2 * is + rr];
    const double ui = p.ib[3 * is + rr], vi = p.ib[4 * is + rr], wi = p.ib[5 * is + rr];
    const double h2 = p.ib[6 * is + rr], dtr = p.ib[7 * is + rr];
    const size_t row = (size_t)chunk * p.ni + (in ? i : 0);
    int *hp = p.hslot + row * p.caph;
    int *sp = p.sslot + row * p.caps;
    int nh = 0, ns = 0;
    double a64[7] = {0, 0, 0, 0, 0, 0, 0};
    ff f[10];
#pragma unroll
    for (int q = 0; q < 10; q++) f[q] = {0.f, 0.f};
    const ff fxi = ff_of(xi), fyi = ff_of(yi), fzi = ff_of(zi);
    const ff fui = ff_of(ui), fvi = ff_of(vi), fwi = ff_of(wi);
    const ff fdt = ff_of(dtr);
    for (int jt = j0; jt < j1; jt += 256) {
        __syncthreads();
        const int j = jt + tid;
        if (j < j1) {
            sA[tid] = reinterpret_cast<const double4 *>(p.rec64)[2 * (size_t)j];
            sB[tid] = reinterpret_cast<const double4 *>(p.rec64)[2 * (size_t)j + 1];
        } else {
            sA[tid] = make_double4(1.0e17, 0.0, 0.0, 0.0);
            sB[tid] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        __syncthreads();
        if (!ok) continue;
        const int jn = min(256, j1 - jt);
        for (int jj = 0; jj < jn; jj++) {
            const double4 a = sA[jj], b = sB[jj];
            const int jg = jt + jj;
            if (PREC == 2) {
#pragma clang fp contract(off)
                const double dx = a.x - xi, dy = a.y - yi, dz = a.z - zi;
                const double du = b.x - ui, dv = b.y - vi, dw = b.z - wi;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double ex = dx + dtr * du, ey = dy + dtr * dv, ez = dz + dtr * dw;
                const double r2p = ex * ex + ey * ey + ez * ez;
                const double rmin = r2 < r2p ? r2 : r2p;
                const double thr = MFLAG ? a.w * h2 : h2;
                if (rmin < thr) {
                    if (nh < p.caph) hp[nh] = jg;
                    nh++;
                } else if (r2 > 0.0) {
                    const double rinv2 = 1.0 / r2;
                    const double rinv = sqrt(rinv2);
                    const double mr1 = a.w * rinv;
                    const double mr3 = mr1 * rinv2;
                    const double rv = dx * du + dy * dv + dz * dw;
                    const double al = -3.0 * rv * rinv2;
                    a64[0] += mr3 * dx;
                    a64[1] += mr3 * dy;
                    a64[2] += mr3 * dz;
                    a64[3] += mr3 * (du + al * dx);
                    a64[4] += mr3 * (dv + al * dy);
                    a64[5] += mr3 * (dw + al * dz);
                    a64[6] += mr1;
                }
            } else {
                const ff dx = ff_add(ff_of(a.x), ff_neg(fxi));
                const ff dy = ff_add(ff_of(a.y), ff_neg(fyi));
                const ff dz = ff_add(ff_of(a.z), ff_neg(fzi));
                const ff du = ff_add(ff_of(b.x), ff_neg(fui));
                const ff dv = ff_add(ff_of(b.y), ff_neg(fvi));
                const ff dw = ff_add(ff_of(b.z), ff_neg(fwi));
                const ff r2 = ff_add(ff_add(ff_mul(dx, dx), ff_mul(dy, dy)), ff_mul(dz, dz));
                const ff ex = ff_add(dx, ff_mul(fdt, du));
                const ff ey = ff_add(dy, ff_mul(fdt, dv));
                const ff ez = ff_add(dz, ff_mul(fdt, dw));
                const ff r2p = ff_add(ff_add(ff_mul(ex, ex), ff_mul(ey, ey)), ff_mul(ez, ez));
                const double d2 = ff_d(r2), d2p = ff_d(r2p);
                const double rmin = d2 < d2p ? d2 : d2p;
                const double thr = MFLAG ? a.w * h2 : h2;
                // float-float keeps ~2^-44 relative: 2^-30 covers the decision
                if (rmin < thr * (1.0 + 0x1p-30) + 0x1p-120 && jg < j1) {
                    if (nh < p.caph) {
                        hp[nh] = jg;
                    } else if (!(rmin < thr * (1.0 - 0x1p-30) - 0x1p-120)) {
                        // overflowing row, pair inside the band: the FP64 test decides
                        if (ns < p.caps) sp[ns] = ~jg;
                        ns++;
                    }
                    nh++;
                } else if (d2 > 0.0) {
                    // rinv: v_rsq_f32 refined by one float-float Newton-Raphson step
                    const float y0 = __builtin_amdgcn_rsqf(r2.h);
                    const ff y = {y0, 0.f};
                    const ff e = ff_add({1.f, 0.f}, ff_neg(ff_mul(r2, ff_mul(y, y))));
                    const ff rinv = ff_add(y, ff_mul(y, {0.5f * e.h, 0.5f * e.l}));
                    const ff rinv2 = ff_mul(rinv, rinv);
                    const ff mr1 = ff_mul(ff_of(a.w), rinv);
                    const ff mr3 = ff_mul(mr1, rinv2);
                    const ff rv = ff_add(ff_add(ff_mul(dx, du), ff_mul(dy, dv)), ff_mul(dz, dw));
                    const ff be = ff_mul(ff_mul(rv, rinv2), mr3);
                    f[0] = ff_add(f[0], ff_mul(mr3, dx));
                    f[1] = ff_add(f[1], ff_mul(mr3, dy));
                    f[2] = ff_add(f[2], ff_mul(mr3, dz));
                    f[3] = ff_add(f[3], mr1);
                    f[4] = ff_add(f[4], ff_mul(mr3, du));
                    f[5] = ff_add(f[5], ff_mul(mr3, dv));
                    f[6] = ff_add(f[6], ff_mul(mr3, dw));
                    f[7] = ff_add(f[7], ff_mul(be, dx));
                    f[8] = ff_add(f[8], ff_mul(be, dy));
                    f[9] = ff_add(f[9], ff_mul(be, dz));
                }
            }
        }
        if (PREC == 1) {   // float-float partials into FP64 after every tile
            a64[0] += ff_d(f[0]);
            a64[1] += ff_d(f[1]);
            a64[2] += ff_d(f[2]);
            a64[3] += ff_d(f[4]) - 3.0 * ff_d(f[7]);
            a64[4] += ff_d(f[5]) - 3.0 * ff_d(f[8]);
            a64[5] += ff_d(f[6]) - 3.0 * ff_d(f[9]);
            a64[6] += ff_d(f[3]);
#pragma unroll
            for (int q = 0; q < 10; q++) f[q] = {0.f, 0.f};
        }
    }
    if (in) {
        double *pp = p.part + (size_t)chunk * NPART * p.ni + i;
#pragma unroll
        for (int q = 0; q < 7; q++) pp[(size_t)q * p.ni] = a64[q];
#pragma unroll
        for (int q = 7; q < NPART; q++) pp[(size_t)q * p.ni] = 0.0;   // no FP32 sums
        p.cnt[((size_t)chunk * 2 + 0) * p.ni + i] = nh;
        p.cnt[((size_t)chunk * 2 + 1) * p.ni + i] = ns;
    }
}

struct FinParams {
    const double *__restrict__ ib;   // SoA [8][ib_stride]
    int ni, ib_stride;               // ni = slot rows of the work arrays
    const int *__restrict__ imap;    // slot -> caller row (nullptr: identity; < 0: empty)
    const double *__restrict__ rec64;   // sorted AoS [nj][8]
    const int *__restrict__ perm;       // sorted -> caller index
    const int *__restrict__ jbase;      // per i: index of its cluster's first j (or null)
    int nj;
    int nchunk;
    int nprow;                       // partial-sum rows per i: nchunk * segments per chunk
    const double *__restrict__ part;
    const int *__restrict__ cnt;
    const int *__restrict__ hslot;
    int caph;
    const int *__restrict__ sslot;
    int caps;
    int lmax, nnbmax;                // nnbmax already min(nnbmax, lmax-1)
    double *__restrict__ out;        // SoA [7][out_stride]
    int out_stride;
    int *__restrict__ list;          // [ni][lmax]
    int refine;                      // flag rows whose FP32 error scale is too large
    int row_base;                    // added to r for fix_rows (batches without imap)
    int fix_slot;                    // fix_rows holds slots i instead of caller rows
    float kf[NERR];                  // weights of the error scales (row_bounds)
    int *__restrict__ fix_rows;      // rows for the whole-row FP64 pass: caller rows, or
    int *__restrict__ nfix;          // slots (fix_slot); their count (device)
    // chunk-selective FP64 pass (plain calls): per flagged row the chunks to recompute,
    // as items k = (slot, chunk) stored contiguously per row in chunk order, each also
    // listed in its chunk's bucket (seg_items[c * ni + pos] = k, seg_cnt[c] entries), and
    // one entry {caller row, first item, items} per row
    int2 *__restrict__ items;
    int *__restrict__ nitems;
    int item_cap;
    int row_items;                   // > 0: slot i's items at i * row_items (at most that
                                     // many; a row needing more goes to the whole-row pass),
                                     // so the choice depends on the row alone; -1: no room,
                                     // every flagged row whole; 0: appended (batched form)
    int *__restrict__ seg_items;
    int *__restrict__ seg_cnt;
    int4 *__restrict__ rowsel;
    int *__restrict__ nrows;
    const int *__restrict__ padoff;  // batched form: segments are the clusters [nseg + 1]
    int nseg;
    double *__restrict__ dbg;        // diagnostic: the error scales of each row [NERR][out_stride]
};

// Error scales of a row's FP32 sums -> does the row need the FP64 pass?  u = 2^-24.
// e[0], e[1]: sums over the 64-j flushes of the L1 norms of the FP32 acc and jerk
// partials (the rounding of the sums themselves); e[2], e[3]: sums over the 64-j
// segments of sum(m/r^3) times the bound of |x_i - P| + |x_j - P| and of
// |v_i - V| + |v_j - V| in the segment's tile (the rounding of the FP32 separations
// and velocity differences: a relative error u of each moves an acc term by at most
// 2 m/r^3 u |x| and a jerk term by m/r^3 u |v| plus its position part).  A row is
// recomputed in FP64 when k0 u e[0] + k2 u e[2] exceeds REFINE_REL |acc| or
// k1 u e[1] + k3 u e[3] + k2 u e[2] |jerk| / |acc| exceeds REFINE_REL |jerk|.  Rows
// whose sums are NaN (band pairs left without a slot) are recomputed too.
constexpr double REFINE_REL = 1.0e-6;
struct RowRef {
    double ta, tj;       // allowed acc and jerk error: REFINE_REL |acc|, REFINE_REL |jerk|
    double jpa;          // |jerk| / |acc|
};
__device__ __forceinline__ RowRef row_ref(const double *s) {
    RowRef q;
    const double a = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double j = sqrt(s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    q.ta = REFINE_REL * a;
    q.tj = REFINE_REL * j;
    q.jpa = j / (a + 1e-300);
    return q;
}
// bound of the acc and jerk errors from the error scales e[NERR] (of a row or a chunk)
__device__ __forceinline__ void err_bound(const double *e, const float *k, const RowRef &q,
                                          double &ea, double &ej) {
    const double ep = U24 * (double)k[2] * e[2];
    ea = U24 * (double)k[0] * e[0] + ep;
    // the position part of a jerk term's error relative to the jerk: as for acc
    ej = U24 * ((double)k[1] * e[1] + (double)k[3] * e[3]) + ep * q.jpa;
}
__device__ __forceinline__ bool needs_fp64(const double *s, const double *e, bool nan_row,
                                           const float *k) {
    if (nan_row) return true;
    const RowRef q = row_ref(s);
    double ea, ej;
    err_bound(e, k, q, ea, ej);
    return !(ea <= q.ta) || !(ej <= q.tj);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int lane_prefix(unsigned long long b) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

struct IF64 {
    double x, y, z, u, v, w, h2, dtr;
};

// FP64 neighbour test of sorted j against i (formula order = the oracle's) and,
// for a non-neighbour, its exact force added to s[].  test=false: force only.
template <bool MFLAG>
__device__ __forceinline__ bool pair64(const double4 &r0, const double4 &r1, const IF64 &I,
                                       bool test, double *s) {
    double dx, dy, dz, du, dv, dw, r2, rmin;
    {
#pragma clang fp contract(off)
        dx = r0.x - I.x; dy = r0.y - I.y; dz = r0.z - I.z;
        du = r1.x - I.u; dv = r1.y - I.v; dw = r1.z - I.w;
        r2 = dx * dx + dy * dy + dz * dz;
        const double ex = dx + I.dtr * du;
        const double ey = dy + I.dtr * dv;
        const double ez = dz + I.dtr * dw;
        const double r2p = ex * ex + ey * ey + ez * ez;
        rmin = r2 < r2p ? r2 : r2p;
    }
    const double thr = MFLAG ? r0.w * I.h2 : I.h2;
    const bool isnb = test && rmin < thr;
    if (!isnb && r2 > 0.0) {
        // 1/sqrt(r2): v_rsq_f64 (~23 bits) and two Newton steps, y (3/2 - r2 y^2 / 2),
        // which leave a few ulp, in place of a correctly rounded divide and square root
        const double hr = 0.5 * r2;
        double rinv = __builtin_amdgcn_rsq(r2);
        rinv = rinv * fma(-hr * rinv, rinv, 1.5);
        rinv = rinv * fma(-hr * rinv, rinv, 1.5);
        const double rinv2 = rinv * rinv;
        const double mr1 = r0.w * rinv;
        const double mr3 = mr1 * rinv2;
        const double rv = dx * du + dy * dv + dz * dw;
        const double al = -3.0 * rv * rinv2;
        s[0] += mr3 * dx;
        s[1] += mr3 * dy;
        s[2] += mr3 * dz;
        s[3] += mr3 * (du + al * dx);
        s[4] += mr3 * (dv + al * dy);
        s[5] += mr3 * (dw + al * dz);
        s[6] += mr1;
    }
    return isnb;
}

// pair64's force terms (test on) without a branch, for the FP64 pass's long j loops:
// the same operations on the pairs that count and exact zeros added for the others,
// so that unrolled iterations interleave (one pair's chain of dependent FP64
// operations alone leaves the VALU idle)
template <bool MFLAG>
__device__ __forceinline__ void pair64_sum(const double4 &r0, const double4 &r1, const IF64 &I,
                                           double *s) {
    double dx, dy, dz, du, dv, dw, r2, rmin;
    {
#pragma clang fp contract(off)
        dx = r0.x - I.x; dy = r0.y - I.y; dz = r0.z - I.z;
        du = r1.x - I.u; dv = r1.y - I.v; dw = r1.z - I.w;
        r2 = dx * dx + dy * dy + dz * dz;
        const double ex = dx + I.dtr * du;
        const double ey = dy + I.dtr * dv;
        const double ez = dz + I.dtr * dw;
        const double r2p = ex * ex + ey * ey + ez * ez;
        rmin = r2 < r2p ? r2 : r2p;
    }
    const double thr = MFLAG ? r0.w * I.h2 : I.h2;
    const bool use = !(rmin < thr) && r2 > 0.0;
    const double r2s = use ? r2 : 1.0;
    const double hr = 0.5 * r2s;
    double rinv = __builtin_amdgcn_rsq(r2s);
    rinv = rinv * fma(-hr * rinv, rinv, 1.5);
    rinv = rinv * fma(-hr * rinv, rinv, 1.5);
    const double rinv2 = rinv * rinv;
    const double mr1 = use ? r0.w * rinv : 0.0;
    const double mr3 = mr1 * rinv2;
    const double rv = dx * du + dy * dv + dz * dw;
    const double al = -3.0 * rv * rinv2;
    s[0] += mr3 * dx;
    s[1] += mr3 * dy;
    s[2] += mr3 * dz;
    s[3] += mr3 * (du + al * dx);
    s[4] += mr3 * (dv + al * dy);
    s[5] += mr3 * (dw + al * dz);
    s[6] += mr1;
}

// pair64_sum over jn records of an LDS tile, four at a time, in j order
template <bool MFLAG>
__device__ __forceinline__ void tile_sum64(const double4 *sA, const double4 *sB, int jn,
                                           const IF64 &I, double *s) {
    int jj = 0;
    for (; jj + 4 <= jn; jj += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) pair64_sum<MFLAG>(sA[jj + u], sB[jj + u], I, s);
    }
    for (; jj < jn; jj++) pair64_sum<MFLAG>(sA[jj], sB[jj], I, s);
}

template <bool MFLAG>
__device__ __forceinline__ bool classify(const double *__restrict__ rec, const IF64 &I, bool test,
                                         double *s) {
    const double4 r0 = reinterpret_cast<const double4 *>(rec)[0];   // x y z m
    const double4 r1 = reinterpret_cast<const double4 *>(rec)[1];   // vx vy vz 0
    return pair64<MFLAG>(r0, r1, I, test, s);
}

// The FP64 pass over one i's candidates, in the chunks c = c0 + cstep (64 p + lane),
// p = 0, 1, ...: every hard candidate is re-tested (neighbours into buf), every
// non-neighbour candidate's force is added to s.  SHARED: several wavefronts fill one
// row, taking buffer positions from *nshared; otherwise this wavefront's nnb counts them.
template <bool MFLAG, bool SHARED>
__device__ __forceinline__ void fin_candidates(const FinParams &f, const IF64 &I, int i, int jb,
                                               int c0, int cstep, int lane, double *s, int *buf,
                                               int *nshared, int &nnb, int &dropped, int &unres,
                                               bool &dropped_any) {
    for (int cg = 0; c0 + cstep * cg < f.nchunk; cg += 64) {
        const int c = c0 + cstep * (cg + lane);
        const int nh = c < f.nchunk ? f.cnt[((size_t)c * 2 + 0) * f.ni + i] : 0;
        const int ns = c < f.nchunk ? f.cnt[((size_t)c * 2 + 1) * f.ni + i] : 0;
        const int over = wave_sum_i(nh > f.caph ? nh - f.caph : 0);
        dropped += over;
        dropped_any = dropped_any || over > 0;
        unsigned long long mask = __ballot(nh > 0 || ns > 0);
        while (mask) {
            const int cl = __builtin_ctzll(mask);
            mask &= mask - 1;
            const size_t rowc = (size_t)(c0 + cstep * (cg + cl)) * f.ni + i;
            const int nhc = min(__shfl(nh, cl, 64), f.caph);
            const int nsc = __shfl(ns, cl, 64);
            const int *hs = f.hslot + rowc * f.caph;
            for (int kb = 0; kb < nhc; kb += 64) {
                const int k = kb + lane;
                bool isnb = false;
                int jo = 0;
                if (k < nhc) {
                    const int sj = hs[k];
                    if ((unsigned)sj < (unsigned)f.nj) {
                        isnb = classify<MFLAG>(f.rec64 + (size_t)sj * 8, I, true, s);
                        jo = f.perm[sj] - jb;
                    }
                }
                const unsigned long long b = __ballot(isnb);
                int base = nnb;
                if (SHARED) {
                    int t = 0;
                    if (lane == 0 && b) t = atomicAdd(nshared, __popcll(b));
                    base = __shfl(t, 0, 64);
                }
                const int pos = base + lane_prefix(b);
                if (isnb && pos < NB_SORT_MAX) buf[pos] = jo;
                nnb += __popcll(b);
            }
            // soft slots: near-field pairs (force only) and, marked ~j, band pairs
            // beyond the hard slots of an overflowing row; those were counted in
            // dropped and are re-counted by the FP64 test here
            const int *ss = f.sslot + rowc * f.caps;
            const int nsr = min(nsc, f.caps);
            unres += nsc - nsr;
            for (int kb = 0; kb < nsr; kb += 64) {
                const int k = kb + lane;
                bool band = false, isnb = false;
                if (k < nsr) {
                    const int e = ss[k];
                    band = e < 0;
                    const int sj = band ? ~e : e;
                    if ((unsigned)sj < (unsigned)f.nj)
                        isnb = classify<MFLAG>(f.rec64 + (size_t)sj * 8, I, band, s);
                }
                dropped += __popcll(__ballot(isnb)) - __popcll(__ballot(band));
            }
        }
    }

}

// A flagged row whose sums are finite: the chunks to recompute in FP64, chosen so that
// the bounds of the chunks left in FP32 add up to at most the row's allowed acc and jerk
// errors: greedily by largest share of the allowance (at most 64 chunks, one per lane),
// or, with more chunks, every chunk above 1/nchunk of it.  Run by one whole wavefront
// (s: the row's sums and error scales, the same in every lane).  The row's items are
// contiguous and in chunk order; a row whose items do not fit the item buffer takes the
// whole-row pass instead.
constexpr int SHARE_LDS = 2048;   // segments whose shares the wide finalize keeps in LDS

__device__ void emit_chunks(const FinParams &f, int i, int r, const double *s, int lane,
                            float2 *lds_share = nullptr) {
    // batched form: the row's cluster cl (slot i in [padoff[cl], padoff[cl + 1])); its
    // segments are the cluster's, bucketed per (cluster, segment)
    int cl = 0;
    if (f.padoff) {
        int hi = f.nseg;
        while (hi - cl > 1) {
            const int mid = (cl + hi) >> 1;
            if (f.padoff[mid] <= i) cl = mid; else hi = mid;
        }
    }
    const int fix_id = f.fix_slot ? i : r + f.row_base;
    const RowRef q = row_ref(s);
    auto bound = [&](int c, double &ea, double &ej) {
        ea = ej = 0.0;
        if (c >= f.nprow) return;
        const double *pp = f.part + (size_t)c * NPART * f.ni + i;
        double e[NERR];
#pragma unroll
        for (int k = 0; k < NERR; k++) e[k] = pp[(size_t)(7 + k) * f.ni];
        err_bound(e, f.kf, q, ea, ej);
    };
    bool mine = false;                    // (greedy form) this lane's chunk is selected
    if (f.nprow <= 64) {
        double ea, ej;
        bound(lane, ea, ej);
        const double w = fmax(ea / (q.ta + 1e-300), ej / (q.tj + 1e-300));
        for (int it = 0; it < f.nprow; it++) {
            const double ra = wave_sum(mine ? 0.0 : ea), rj = wave_sum(mine ? 0.0 : ej);
            if (ra <= q.ta && rj <= q.tj) break;
            // the unselected chunk with the largest share (lowest lane on ties)
            double best = mine || lane >= f.nprow ? -1.0 : w;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) best = fmax(best, __shfl_xor(best, o, 64));
            const unsigned long long m = __ballot(!mine && lane < f.nprow && w == best);
            if (!m) break;
            mine = mine || lane == __builtin_ctzll(m);
        }
    }
    // more than 64 (many short chunks, small i-sets): the segments whose share exceeds
    // a threshold t, t found by bisection as the largest for which the rest fits the
    // allowed error (selecting every segment with a share above 0 always fits)
    // shares relative to the allowed errors (from LDS where the caller gave room: the
    // bisection below reads them twenty times)
    const bool cached = lds_share && f.nprow > 64 && f.nprow <= SHARE_LDS;
    if (cached) {
        for (int c = lane; c < f.nprow; c += 64) {
            double ea, ej;
            bound(c, ea, ej);
            lds_share[c] = make_float2((float)(ea / (q.ta + 1e-300)), (float)(ej / (q.tj + 1e-300)));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // share of segment c; ea, ej: its bounds relative to the allowed errors
    auto share = [&](int c, double &ea, double &ej) {
        if (cached) {
            const float2 w = c < f.nprow ? lds_share[c] : make_float2(0.f, 0.f);
            ea = w.x;
            ej = w.y;
        } else {
            bound(c, ea, ej);
            ea = ea / (q.ta + 1e-300);
            ej = ej / (q.tj + 1e-300);
        }
        return fmax(ea, ej);
    };
    double tsel = 0.0;
    if (f.nprow > 64) {
        double wmax = 0.0;
        for (int cb = 0; cb < f.nprow; cb += 64) {
            double ea, ej;
            wmax = fmax(wmax, share(cb + lane, ea, ej));
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) wmax = fmax(wmax, __shfl_xor(wmax, o, 64));
        if (!(wmax < 1e300)) {
            tsel = 0.0;                      // not finite: every segment with a share
        } else {
            double lo = 0.0, hi = wmax;      // fits at lo, not at hi (the row is flagged)
            for (int it = 0; it < 20; it++) {
                const double t = 0.5 * (lo + hi);
                double ra = 0.0, rj = 0.0;
                for (int cb = 0; cb < f.nprow; cb += 64) {
                    double ea, ej;
                    if (!(share(cb + lane, ea, ej) > t)) {
                        ra += ea;
                        rj += ej;
                    }
                }
                ra = wave_sum(ra);
                rj = wave_sum(rj);
                if (ra <= 1.0 && rj <= 1.0) lo = t; else hi = t;
            }
            tsel = lo;
        }
    }
    auto sel = [&](int c) {
        if (c >= f.nprow) return false;
        if (f.nprow <= 64) return mine;
        double ea, ej;
        return share(c, ea, ej) > tsel;
    };
    int cnt = 0;
    for (int cb = 0; cb < f.nprow; cb += 64) cnt += __popcll(__ballot(sel(cb + lane)));
    if (cnt == 0) return;    // within the bound with every chunk in FP32
    int start = 0;
    if (f.row_items != 0) {
        if (cnt > f.row_items) {
            if (lane == 0) f.fix_rows[atomicAdd(f.nfix, 1)] = fix_id;
            return;
        }
        if (lane == 0) atomicAdd(f.nitems, cnt);   // (a count only)
        start = i * f.row_items;
    } else {
        if (lane == 0) start = atomicAdd(f.nitems, cnt);
        start = __shfl(start, 0, 64);
        if (start + cnt > f.item_cap) {
            if (lane == 0) f.fix_rows[atomicAdd(f.nfix, 1)] = fix_id;
            return;
        }
    }
    int k = start;
    for (int cb = 0; cb < f.nprow; cb += 64) {
        const int c = cb + lane;
        const bool b = sel(c);
        const unsigned long long m = __ballot(b);
        if (b) {
            const int kk = k + lane_prefix(m);
            if (f.padoff) {
                // (cluster, segment) g; bucket g holds at most the cluster's slots
                const int g = cl * f.nprow + c;
                f.items[kk] = make_int2(i, g);
                const int pos = atomicAdd(&f.seg_cnt[g], 1);
                f.seg_items[(size_t)f.padoff[cl] * f.nprow +
                            (size_t)c * (f.padoff[cl + 1] - f.padoff[cl]) + pos] = kk;
            } else {
                f.items[kk] = make_int2(i, c);
                const int pos = atomicAdd(&f.seg_cnt[c], 1);
                f.seg_items[(size_t)c * f.ni + pos] = kk;
            }
        }
        k += __popcll(m);
    }
    if (lane == 0) f.rowsel[atomicAdd(f.nrows, 1)] = make_int4(r, start, cnt, 0);
}

template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_finalize_kernel(FinParams f) {
    __shared__ int nbuf[4][NB_SORT_MAX];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wid;   // slot
    if (i >= f.ni) return;   // whole wavefront leaves together
    const int r = f.imap ? f.imap[i] : i;   // caller row
    if (r < 0) return;
    int *buf = nbuf[wid];

    // 1. FP64 sum of the chunk partials (and of the error scales)
    double s[7], e[NERR] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 7; q++) s[q] = 0.0;
    for (int c = lane; c < f.nprow; c += 64) {
        const double *pp = f.part + (size_t)c * NPART * f.ni + i;
#pragma unroll
        for (int q = 0; q < 7; q++) s[q] += pp[(size_t)q * f.ni];
        if (f.refine || f.dbg) {
#pragma unroll
            for (int q = 0; q < NERR; q++) e[q] += pp[(size_t)(7 + q) * f.ni];
        }
    }

    // 2. FP64 work on the candidates
    const int is = f.ib_stride;
    IF64 I;
    I.x = f.ib[0 * is + r]; I.y = f.ib[1 * is + r]; I.z = f.ib[2 * is + r];
    I.u = f.ib[3 * is + r]; I.v = f.ib[4 * is + r]; I.w = f.ib[5 * is + r];
    I.h2 = f.ib[6 * is + r]; I.dtr = f.ib[7 * is + r];
    const int jb = f.jbase ? f.jbase[r] : 0;

    int nnb = 0;          // exact neighbours found (wave-uniform)
    int dropped = 0;      // hard candidates beyond the slot capacity
    int unres = 0;        // band pairs of an overflowing row left without a slot
    bool dropped_any = false;
    fin_candidates<MFLAG, false>(f, I, i, jb, 0, 1, lane, s, buf, nullptr, nnb, dropped, unres,
                                 dropped_any);

#pragma unroll
    for (int q = 0; q < 7; q++) s[q] = wave_sum(s[q]);
    if (f.refine || f.dbg) {
#pragma unroll
        for (int q = 0; q < NERR; q++) e[q] = wave_sum(e[q]);
    }
    int *row = f.list + (size_t)r * f.lmax;
    const int tot = nnb + dropped;
    // a row whose hard slots overflowed reports -count, as any overflow does, even in
    // the (never seen) case where band pairs brought the exact count back under nnbmax:
    // its neighbours past the slots were not kept.  Band pairs left without a slot
    // would leave the sums incomplete: such a row's forces are NaN.
    const bool ovf = tot > f.nnbmax || dropped_any;
    if (unres > 0) {
#pragma unroll
        for (int q = 0; q < 7; q++) s[q] = __builtin_nan("");
    }
    if (lane == 0) {
        double *o = f.out + r;
        const int os = f.out_stride;
        o[0 * os] = s[0]; o[1 * os] = s[1]; o[2 * os] = s[2];
        o[3 * os] = s[3]; o[4 * os] = s[4]; o[5 * os] = s[5];
        o[6 * os] = s[6];
        row[0] = ovf ? -tot : tot;
    }
    if (f.dbg && lane == 0) {
#pragma unroll
        for (int q = 0; q < NERR; q++) f.dbg[(size_t)q * f.out_stride + r] = e[q];
    }
    // s and e are wave sums: every lane takes the same decision
    if (f.refine && needs_fp64(s, e, unres > 0, f.kf)) {
        if (unres > 0 || !f.items) {
            if (lane == 0) f.fix_rows[atomicAdd(f.nfix, 1)] = f.fix_slot ? i : r + f.row_base;
        } else {
            double se[7 + NERR];
#pragma unroll
            for (int q = 0; q < 7; q++) se[q] = s[q];
#pragma unroll
            for (int q = 0; q < NERR; q++) se[7 + q] = e[q];
            emit_chunks(f, i, r, se, lane);
        }
    }
    if (!ovf && nnb > 0) {
        // ascending row: the rank of each (distinct) index among the row
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int k = lane; k < nnb; k += 64) {
            const int e = buf[k];
            int rank = 0;
            for (int m = 0; m < nnb; m++) rank += buf[m] < e;
            row[1 + rank] = e;
        }
    }
}

// One work-group of W wavefronts per i, for small i-sets, where one wavefront per i
// leaves most of the GPU idle while it waits on its gathers: wavefront w takes the
// chunks c = w (mod W), so the few chunks that hold an i's candidates (they lie close
// together in the Morton order) are spread over the W.  Forces are combined in a fixed
// order; the row is rank-sorted as in regf_finalize_kernel.
template <bool MFLAG, int W>
__global__ __launch_bounds__(64 * W) void regf_finalize_wide_kernel(FinParams f) {
    __shared__ int buf[NB_SORT_MAX];
    __shared__ double red[W][NPART];
    __shared__ int cnts[W][3];
    __shared__ int nshared;
    __shared__ int flag;             // 1: chunk items, 2: whole-row pass
    __shared__ double vtot[NPART];
    __shared__ float2 shr[SHARE_LDS];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int i = blockIdx.x;   // slot
    const int r = f.imap ? f.imap[i] : i;
    if (r < 0) return;   // the whole work-group leaves together
    if (threadIdx.x == 0) {
        nshared = 0;
        flag = 0;
    }
    __syncthreads();

    double s[NPART];
#pragma unroll
    for (int q = 0; q < NPART; q++) s[q] = 0.0;
    for (int c = w + W * lane; c < f.nprow; c += 64 * W) {
        const double *pp = f.part + (size_t)c * NPART * f.ni + i;
#pragma unroll
        for (int q = 0; q < 7; q++) s[q] += pp[(size_t)q * f.ni];
        if (f.refine || f.dbg) {
#pragma unroll
            for (int q = 7; q < NPART; q++) s[q] += pp[(size_t)q * f.ni];
        }
    }
    const int is = f.ib_stride;
    IF64 I;
    I.x = f.ib[0 * is + r]; I.y = f.ib[1 * is + r]; I.z = f.ib[2 * is + r];
    I.u = f.ib[3 * is + r]; I.v = f.ib[4 * is + r]; I.w = f.ib[5 * is + r];
    I.h2 = f.ib[6 * is + r]; I.dtr = f.ib[7 * is + r];
    const int jb = f.jbase ? f.jbase[r] : 0;
    int nnb = 0, dropped = 0, unres = 0;
    bool dropped_any = false;
    fin_candidates<MFLAG, true>(f, I, i, jb, w, W, lane, s, buf, &nshared, nnb, dropped, unres,
                                dropped_any);
#pragma unroll
    for (int q = 0; q < NPART; q++) s[q] = wave_sum(s[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NPART; q++) red[w][q] = s[q];
        cnts[w][0] = dropped;
        cnts[w][1] = unres;
        cnts[w][2] = dropped_any ? 1 : 0;
    }
    __syncthreads();
    const int nall = nshared;
    int tot = nall, unr = 0, dq = 0;
#pragma unroll
    for (int q = 0; q < W; q++) {
        tot += cnts[q][0];
        unr += cnts[q][1];
        dq |= cnts[q][2];
    }
    const bool dany = dq != 0;
    // as in regf_finalize_kernel: a row whose hard slots overflowed reports -count;
    // band pairs left without a slot make the row's forces NaN
    const bool ovf = tot > f.nnbmax || dany;
    int *row = f.list + (size_t)r * f.lmax;
    if (threadIdx.x == 0) {
        double *o = f.out + r;
        const int os = f.out_stride;
        double v[NPART];
#pragma unroll
        for (int q = 0; q < NPART; q++) {
            v[q] = 0.0;   // a fixed order
#pragma unroll
            for (int u = 0; u < W; u++) v[q] += red[u][q];
        }
#pragma unroll
        for (int q = 0; q < 7; q++) o[q * os] = unr > 0 ? __builtin_nan("") : v[q];
        row[0] = ovf ? -tot : tot;
        flag = f.refine && needs_fp64(v, v + 7, unr > 0, f.kf) ? (unr > 0 || !f.items ? 2 : 1) : 0;
        if (f.dbg) {
#pragma unroll
            for (int q = 0; q < NERR; q++) f.dbg[(size_t)q * f.out_stride + r] = v[7 + q];
        }
        if (flag == 2) f.fix_rows[atomicAdd(f.nfix, 1)] = f.fix_slot ? i : r + f.row_base;
#pragma unroll
        for (int q = 0; q < NPART; q++) vtot[q] = v[q];
    }
    __syncthreads();
    if (flag == 1 && w == 0) {
        double se[NPART];
#pragma unroll
        for (int q = 0; q < NPART; q++) se[q] = vtot[q];
        emit_chunks(f, i, r, se, lane, shr);
    }
    if (!ovf && nall > 0) {
        for (int k = threadIdx.x; k < nall; k += 64 * W) {
            const int e = buf[k];
            int rank = 0;
            for (int m = 0; m < nall; m++) rank += buf[m] < e;
            row[1 + rank] = e;
        }
    }
}

// ------------------------------------------------------------------ FP64 pass
// Rows flagged by the finalize are recomputed over every j in FP64: the same pair
// function as the finalize's (pair64: the oracle's neighbour test, so exactly the
// row's neighbours are left out), j tiles staged in LDS and shared by the 256 rows of
// a work-group.  A persistent grid walks (row block, j-chunk) items; the number of
// chunks follows from the number of flagged rows alone, and each row's chunk sums are
// added in chunk order, so a row's result does not depend on which slot it got.
constexpr int FIX_ITEMS = 2048;       // work items aimed at per launch

struct FixParams {
    const double *__restrict__ ib;   // SoA [8][ib_stride] (caller rows)
    int ib_stride;
    const int *__restrict__ rows;    // flagged caller rows
    const int *__restrict__ nfix;
    const double *__restrict__ rec64;
    int j0, j1;                      // sorted j range
    double *__restrict__ part;       // [nch][7][nfix]
    double *__restrict__ out;        // SoA [7][out_stride]
    int out_stride;
    int items;                       // work items aimed at (the launch's grid size)
};

__device__ __forceinline__ void fix_chunks(int nfix, int nj, int items, int &nch, int &tpc) {
    const int nib = (nfix + 255) / 256;
    const int ntile = (nj + JTILE - 1) / JTILE;
    nch = max(1, min(items / nib, ntile));
    tpc = (ntile + nch - 1) / nch;          // tiles per chunk
    nch = (ntile + tpc - 1) / tpc;
}

template <bool MFLAG>
__device__ __forceinline__ void fix_body(const FixParams &f, int bid, int nblk, double4 *sA,
                                         double4 *sB) {
    const int nfix = *f.nfix;
    if (nfix == 0) return;
    const int nj = f.j1 - f.j0;
    int nch, tpc;
    fix_chunks(nfix, nj, f.items, nch, tpc);
    const int nib = (nfix + 255) / 256;
    const int items = nib * nch;
    for (int it = bid; it < items; it += nblk) {   // uniform per work-group
        const int b = it % nib, ch = it / nib;
        const int k = b * 256 + threadIdx.x;
        const bool ok = k < nfix;
        const int r = ok ? f.rows[k] : 0;
        const int is = f.ib_stride;
        IF64 I;
        I.x = f.ib[0 * is + r]; I.y = f.ib[1 * is + r]; I.z = f.ib[2 * is + r];
        I.u = f.ib[3 * is + r]; I.v = f.ib[4 * is + r]; I.w = f.ib[5 * is + r];
        I.h2 = f.ib[6 * is + r]; I.dtr = f.ib[7 * is + r];
        double sum[7] = {0, 0, 0, 0, 0, 0, 0};
        const int ja = f.j0 + ch * tpc * JTILE;
        const int jb = min(f.j1, ja + tpc * JTILE);
        for (int jt = ja; jt < jb; jt += JTILE) {
            __syncthreads();
            const int j = jt + threadIdx.x;
            if (j < jb) {
                sA[threadIdx.x] = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j];
                sB[threadIdx.x] = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j + 1];
            }
            __syncthreads();
            const int jn = min(JTILE, jb - jt);
            if (ok) tile_sum64<MFLAG>(sA, sB, jn, I, sum);
        }
        if (ok) {
#pragma unroll
            for (int q = 0; q < 7; q++) f.part[((size_t)ch * 7 + q) * nfix + k] = sum[q];
        }
    }
}

template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_fix_kernel(FixParams f) {
    __shared__ double4 sA[JTILE], sB[JTILE];
    fix_body<MFLAG>(f, blockIdx.x, gridDim.x, sA, sB);
}

// element q of v[7] without indexing by a lane-varying value
__device__ __forceinline__ double pick7(const double *v, int q) {
    double r = v[0];
#pragma unroll
    for (int k = 1; k < 7; k++) r = q == k ? v[k] : r;
    return r;
}

// one wavefront per whole row: lane l sums chunks l, l + 64, ... in order, then a
// fixed tree over the lanes
__device__ __forceinline__ void fix_reduce_body(const FixParams &f, int bid) {
    const int nfix = *f.nfix;
    const int lane = threadIdx.x & 63;
    const int k = bid * 4 + (threadIdx.x >> 6);
    if (k >= nfix) return;   // whole wavefront
    int nch, tpc;
    fix_chunks(nfix, f.j1 - f.j0, f.items, nch, tpc);
    double sum[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int ch = lane; ch < nch; ch += 64) {
#pragma unroll
        for (int q = 0; q < 7; q++) sum[q] += f.part[((size_t)ch * 7 + q) * nfix + k];
    }
#pragma unroll
    for (int q = 0; q < 7; q++) sum[q] = wave_sum(sum[q]);
    const int r = f.rows[k];
    if (lane < 7) f.out[(size_t)lane * f.out_stride + r] = pick7(sum, lane);
}

__global__ __launch_bounds__(256) void regf_fix_reduce_kernel(FixParams f) {
    fix_reduce_body(f, blockIdx.x);
}

// Chunk-selective FP64 pass.  Item k = (slot i, chunk c) replaces the FP32 sums of
// chunk c in row i by FP64: all of the chunk's pairs that are not FP64 neighbours
// (pair64, the finalize's test), minus the candidates the finalize already added in
// FP64 (the same slots and tests as fin_candidates), minus the chunk's FP32 partials.
//  * regf_cand_kernel: one wavefront per item, the candidate terms -> cand[k];
//  * regf_segfix_kernel: work units (chunk c, up to 256 of its items, sub-range s of
//    FIX_SUB j): the sub-range's records pass through LDS once for all the unit's items,
//    one thread per item -> part64[k][s];
//  * regf_segapply_kernel: each row adds, in item and sub-range order,
//    sum_s part64 - cand - FP32 partial of every one of its items.
// Every value depends only on (i, c, s), so results do not depend on the order in which
// items were emitted.
constexpr int FIX_SUB = 1024;          // j per sub-range
constexpr int SEG_MIN = 8192;          // shortest partial-sum segment
constexpr int SEG_ENS = 4096;          // partial-sum segment of a cluster (batched form)
constexpr int SEGCNT_LDS = 4096;       // bucket sizes the segment-sum kernel keeps in LDS
constexpr size_t ITEM_BUDGET = 4ull << 30;   // FP64-pass item memory aimed at per call
constexpr int MIN_ROW_ITEMS = 8;
constexpr int PASS_MERGE_NI = 16384;   // batches up to this many i: the FP64 pass in two launches

struct ChunkFixParams {
    const double *__restrict__ ib;   // SoA [8][ib_stride] (as the finalize of the batch)
    int ib_stride;
    const int *__restrict__ imap;    // slot -> caller row (nullptr: identity)
    const double *__restrict__ rec64;
    int nj, clen, nchunk, ni, nsub;  // clen, nchunk: the partial-sum segments' length, count
    int nseg_c;                      // segments per work-array chunk
    const double *__restrict__ part;
    const int *__restrict__ cnt;
    const int *__restrict__ hslot;
    int caph;
    const int *__restrict__ sslot;
    int caps;
    const int2 *__restrict__ items;
    const int *__restrict__ nitems;
    int item_cap;
    int row_items;                   // most items of one row entry (1 in the batched form)
    const int *__restrict__ seg_items;
    const int *__restrict__ seg_cnt;
    double *__restrict__ cand;       // [item_cap][7]
    double *__restrict__ part64;     // [item_cap][nsub][7]
    const int4 *__restrict__ rowsel;
    const int *__restrict__ nrows;
    double *__restrict__ out;        // SoA [7][out_stride] (of the batch)
    int out_stride;
    const int *__restrict__ padoff;  // batched form: segment c = cluster [padoff[c],
    int maxseg;                      // padoff[c + 1]), work-array chunk 0; maxseg its
};                                   // longest segment

__device__ __forceinline__ IF64 load_if64(const double *__restrict__ ib, int is, int r) {
    IF64 I;
    I.x = ib[0 * is + r]; I.y = ib[1 * is + r]; I.z = ib[2 * is + r];
    I.u = ib[3 * is + r]; I.v = ib[4 * is + r]; I.w = ib[5 * is + r];
    I.h2 = ib[6 * is + r]; I.dtr = ib[7 * is + r];
    return I;
}

// one wavefront per item of the row entries (rows whose items were all stored)
template <bool MFLAG>
__device__ __forceinline__ void cand_body(const ChunkFixParams &f, int bid, int nblk) {
    const int lane = threadIdx.x & 63;
    const int nr = *f.nrows;
    const int ri = max(f.row_items, 1);
    // (entry, item) pairs in contiguous runs per wavefront, a row's items one after
    // another; past a row's last item the run goes on at the next row
    const long long nw = (long long)nr * ri;
    const long long nwave = (long long)nblk * 4;
    const long long per = (nw + nwave - 1) / nwave;
    const long long w0 = (long long)(bid * 4 + (threadIdx.x >> 6)) * per;
    const long long w1 = min(nw, w0 + per);
    int ecur = -1;
    int4 rs = make_int4(0, 0, 0, 0);
    for (long long w = w0; w < w1; w++) {
        const int e = (int)(w / ri), m = (int)(w - (long long)e * ri);
        if (e != ecur) {
            rs = f.rowsel[e];
            ecur = e;
        }
        if (m >= rs.z) {
            w = (long long)(e + 1) * ri - 1;   // the next row
            continue;
        }
        const int k = rs.y + m;
        const int2 it = f.items[k];
        const int i = it.x, c = f.padoff ? 0 : it.y / f.nseg_c;   // work-array chunk
        // the segment's j range (the batched form: segment it.y % nseg_c of cluster
        // it.y / nseg_c)
        unsigned jlo, jhi;
        if (f.padoff) {
            const int cl = it.y / f.nseg_c, sg = it.y - cl * f.nseg_c;
            jlo = (unsigned)(f.padoff[cl] + sg * f.clen);
            jhi = (unsigned)min(f.padoff[cl + 1], f.padoff[cl] + (sg + 1) * f.clen);
        } else {
            jlo = (unsigned)(it.y * f.clen);
            jhi = (unsigned)min(f.nj, (it.y + 1) * f.clen);
        }
        const IF64 I = load_if64(f.ib, f.ib_stride, f.imap ? f.imap[i] : i);
        double cd[7] = {0, 0, 0, 0, 0, 0, 0};
        const size_t rowc = (size_t)c * f.ni + i;
        const int nh = min(f.cnt[((size_t)c * 2 + 0) * f.ni + i], f.caph);
        const int nsr = min(f.cnt[((size_t)c * 2 + 1) * f.ni + i], f.caps);
        for (int q = lane; q < nh; q += 64) {
            const unsigned sj = (unsigned)f.hslot[rowc * f.caph + q];
            if (sj >= jlo && sj < jhi) (void)classify<MFLAG>(f.rec64 + (size_t)sj * 8, I, true, cd);
        }
        for (int q = lane; q < nsr; q += 64) {
            const int e = f.sslot[rowc * f.caps + q];
            const bool band = e < 0;
            const unsigned sj = (unsigned)(band ? ~e : e);
            if (sj >= jlo && sj < jhi) (void)classify<MFLAG>(f.rec64 + (size_t)sj * 8, I, band, cd);
        }
#pragma unroll
        for (int q = 0; q < 7; q++) cd[q] = wave_sum(cd[q]);
        if (lane < 7) f.cand[(size_t)k * 7 + lane] = pick7(cd, lane);
    }
}

template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_cand_kernel(ChunkFixParams f) {
    cand_body<MFLAG>(f, blockIdx.x, gridDim.x);
}

template <bool MFLAG>
__device__ __forceinline__ void segfix_body(const ChunkFixParams &f, int bid, int nblk,
                                            double4 *sA, double4 *sB, int *scnt) {
    const int t = threadIdx.x;
    int gmax = ((f.padoff ? f.maxseg : f.ni) + 255) / 256;
    // bucket sizes into LDS (scnt[SEGCNT_LDS]: their maximum): the unit loop then reads
    // them there and ends after the last group that any bucket fills
    const bool cached = f.nchunk <= SEGCNT_LDS;
    if (cached) {
        if (t == 0) scnt[SEGCNT_LDS] = 0;
        __syncthreads();
        int mx = 0;
        for (int c = t; c < f.nchunk; c += 256) {
            const int v = f.seg_cnt[c];
            scnt[c] = v;
            mx = max(mx, v);
        }
        atomicMax(&scnt[SEGCNT_LDS], mx);
        __syncthreads();
        gmax = min(gmax, (scnt[SEGCNT_LDS] + 255) / 256);
    }
    const long long units = (long long)f.nchunk * gmax * f.nsub;
    // units ordered (item group, chunk, sub-range): the few busy groups of every chunk
    // come first and spread over the whole grid
    const int per_g = f.nchunk * f.nsub;
    for (long long u = bid; u < units; u += nblk) {   // uniform per work-group
        const int g = (int)(u / per_g);
        const int rem = (int)(u - (long long)g * per_g);
        const int c = rem / f.nsub, sub = rem - c * f.nsub;
        const int nc = cached ? scnt[c] : f.seg_cnt[c];
        if (g * 256 >= nc) continue;
        // (the last chunk or a short cluster: trailing sub-ranges are empty and write 0)
        int sb, se;
        size_t bucket;
        if (f.padoff) {
            // segment sg of cluster cl
            const int cl = c / f.nseg_c, sg = c - cl * f.nseg_c;
            const int p0 = f.padoff[cl], p1 = f.padoff[cl + 1];
            sb = p0 + sg * f.clen;
            se = min(p1, sb + f.clen);
            bucket = (size_t)p0 * f.nseg_c + (size_t)sg * (p1 - p0);
        } else {
            sb = c * f.clen;
            se = min(f.nj, (c + 1) * f.clen);
            bucket = (size_t)c * f.ni;
        }
        const int j0 = sb + sub * FIX_SUB;
        const int j1 = min(se, j0 + FIX_SUB);
        const bool ok = g * 256 + t < nc;
        const int k = ok ? f.seg_items[bucket + g * 256 + t] : 0;
        IF64 I;
        if (ok) {
            const int i = f.items[k].x;
            I = load_if64(f.ib, f.ib_stride, f.imap ? f.imap[i] : i);
        }
        double sum[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int jt = j0; jt < j1; jt += JTILE) {
            __syncthreads();
            const int j = jt + t;
            if (j < j1) {
                sA[t] = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j];
                sB[t] = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j + 1];
            }
            __syncthreads();
            const int jn = min(JTILE, j1 - jt);
            if (ok) tile_sum64<MFLAG>(sA, sB, jn, I, sum);
        }
        if (ok) {
            double *o = f.part64 + ((size_t)k * f.nsub + sub) * 7;
#pragma unroll
            for (int q = 0; q < 7; q++) o[q] = sum[q];
        }
    }
}

template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_segfix_kernel(ChunkFixParams f) {
    __shared__ double4 sA[JTILE], sB[JTILE];
    __shared__ int scnt[SEGCNT_LDS + 1];
    segfix_body<MFLAG>(f, blockIdx.x, gridDim.x, sA, sB, scnt);
}

// each flagged row, one wavefront: lane l sums the corrections of items l, l + 64, ...
// (sub-ranges in order), then a fixed tree over the lanes
__device__ __forceinline__ void segapply_body(const ChunkFixParams &f, int bid, int nblk) {
    const int lane = threadIdx.x & 63;
    const int nr = *f.nrows;
    for (int k = bid * 4 + (threadIdx.x >> 6); k < nr; k += nblk * 4) {   // whole wavefronts
        const int4 e = f.rowsel[k];
        double d[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int m = lane; m < e.z; m += 64) {
            const int kk = e.y + m;
            const int2 it = f.items[kk];
            // the partial-sum row: the segment (the batched form: within its cluster)
            const int prow = f.padoff ? it.y % f.nseg_c : it.y;
            const double *pp = f.part + (size_t)prow * NPART * f.ni + it.x;
            double v[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int sb = 0; sb < f.nsub; sb++) {
#pragma unroll
                for (int q = 0; q < 7; q++) v[q] += f.part64[((size_t)kk * f.nsub + sb) * 7 + q];
            }
#pragma unroll
            for (int q = 0; q < 7; q++) d[q] += (v[q] - f.cand[(size_t)kk * 7 + q]) - pp[(size_t)q * f.ni];
        }
#pragma unroll
        for (int q = 0; q < 7; q++) d[q] = wave_sum(d[q]);
        if (lane < 7) f.out[(size_t)lane * f.out_stride + e.x] += pick7(d, lane);
    }
}

__global__ __launch_bounds__(256) void regf_segapply_kernel(ChunkFixParams f) {
    segapply_body(f, blockIdx.x, gridDim.x);
}

// A batch's FP64 pass in two launches (small calls feel every launch): the segment
// sums, the candidates' FP64 terms and, with the last batch, the whole-row partial
// sums, in disjoint ranges of work-groups (none reads what another writes; the heavy
// segment sums first); then the segment corrections and the whole-row results
// (disjoint rows).
template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_pass1_kernel(ChunkFixParams f, FixParams x, int gs,
                                                         int gc) {
    __shared__ double4 sA[JTILE], sB[JTILE];
    __shared__ int scnt[SEGCNT_LDS + 1];
    const int b = blockIdx.x;   // uniform per work-group
    if (b < gs)
        segfix_body<MFLAG>(f, b, gs, sA, sB, scnt);
    else if (b < gs + gc)
        cand_body<MFLAG>(f, b - gs, gc);
    else
        fix_body<MFLAG>(x, b - gs - gc, (int)gridDim.x - gs - gc, sA, sB);
}
__global__ __launch_bounds__(256) void regf_pass2_kernel(ChunkFixParams f, FixParams x, int ga) {
    const int b = blockIdx.x;
    if (b < ga)
        segapply_body(f, b, ga);
    else
        fix_reduce_body(x, b - ga);
}

// Ensemble rows: one work-group per flagged slot, over the padded j range of the
// slot's cluster (padoff: cluster starts in slot order, found by binary search).
// Padding records are massless and far, so they add nothing.  Thread t sums
// j = lo + t + 256 k; the 256 sums are added by a fixed tree, so a row's result does
// not depend on the work-group that takes it.
struct FixRowParams {
    const double *__restrict__ ib;   // SoA [8][ib_stride] (caller rows)
    int ib_stride;
    const int *__restrict__ imap;    // slot -> caller row
    const int *__restrict__ slots;   // flagged slots
    const int *__restrict__ nfix;
    const int *__restrict__ padoff;  // [nc + 1]
    int nc;
    const double *__restrict__ rec64;
    double *__restrict__ out;        // SoA [7][out_stride]
    int out_stride;
};

template <bool MFLAG>
__global__ __launch_bounds__(256) void regf_fix_rows_kernel(FixRowParams f) {
    __shared__ double red[7][256];
    const int nfix = *f.nfix;
    const int t = threadIdx.x;
    for (int k = blockIdx.x; k < nfix; k += gridDim.x) {   // uniform per work-group
        const int slot = f.slots[k];
        const int r = f.imap[slot];
        int lo = 0, hi = f.nc;   // the cluster c with padoff[c] <= slot < padoff[c + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (f.padoff[mid] <= slot) lo = mid; else hi = mid;
        }
        const int j0 = f.padoff[lo], j1 = f.padoff[lo + 1];
        const int is = f.ib_stride;
        IF64 I;
        I.x = f.ib[0 * is + r]; I.y = f.ib[1 * is + r]; I.z = f.ib[2 * is + r];
        I.u = f.ib[3 * is + r]; I.v = f.ib[4 * is + r]; I.w = f.ib[5 * is + r];
        I.h2 = f.ib[6 * is + r]; I.dtr = f.ib[7 * is + r];
        double sum[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int j = j0 + t; j < j1; j += 256) {
            const double4 r0 = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j];
            const double4 r1 = reinterpret_cast<const double4 *>(f.rec64)[2 * (size_t)j + 1];
            (void)pair64<MFLAG>(r0, r1, I, true, sum);
        }
#pragma unroll
        for (int q = 0; q < 7; q++) red[q][t] = sum[q];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (t < o) {
#pragma unroll
                for (int q = 0; q < 7; q++) red[q][t] += red[q][t + o];
            }
            __syncthreads();
        }
        if (t < 7) f.out[(size_t)t * f.out_stride + r] = red[t][0];
        __syncthreads();   // red is reused by the next row
    }
}

__global__ void zero_int_kernel(int *p) { *p = 0; }
__global__ void zero_ints_kernel(int *p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}
// between batches: rows and items of the batch so far into [3], [4]; per-batch
// counters (and the nchunk bucket counts after them) cleared
__global__ void next_batch_kernel(int *c, int nchunk) {
    __syncthreads();
    if (threadIdx.x == 0) {
        c[3] += c[2];
        c[4] += c[1];
        c[1] = 0;
        c[2] = 0;
    }
    for (int k = threadIdx.x; k < nchunk; k += blockDim.x) c[6 + k] = 0;
}
constexpr int CHUNKFIX_GROUPS = 2048;

RegfWork g_work;
int g_last_chunks = 0;
int g_last_batches = 0;
int *g_nfix = nullptr;       // device count of the rows of the last call's FP64 pass
int *g_ens_nfix = nullptr;   // the same for the last batched (ensemble) step
double *g_dbg = nullptr;     // diagnostic output of the rows' error scales (or null)
bool g_last_refine = false;

template <int T, int IPL>
void launch_tile_t(bool mflag, dim3 grid, const TileParams &tp, hipStream_t s) {
    if (mflag)
        hipLaunchKernelGGL((regf_tile_kernel<T, IPL, true>), grid, dim3(T), 0, s, tp);
    else
        hipLaunchKernelGGL((regf_tile_kernel<T, IPL, false>), grid, dim3(T), 0, s, tp);
}

template <int T, int IPL>
void launch_items_t(bool mflag, int grid, const TileParams &tp, const int4 *items, int nitems,
                    int *queue, hipStream_t s) {
    if (mflag)
        hipLaunchKernelGGL((regf_items_kernel<T, IPL, true>), dim3(grid), dim3(T), 0, s, tp, items,
                           nitems, queue);
    else
        hipLaunchKernelGGL((regf_items_kernel<T, IPL, false>), dim3(grid), dim3(T), 0, s, tp, items,
                           nitems, queue);
}

// dispatch on the tile configuration (block size T in {128, 256, 512}, IPL in {1, 2})
void launch_tile(const RegfConfig &cfg, bool mflag, dim3 grid, const TileParams &tp, hipStream_t s) {
    if (cfg.prec == 1 || cfg.prec == 2) {
        grid = dim3(tp.nib, grid.x / tp.nib);   // the variant kernels take (i-block, chunk)
        if (cfg.prec == 1) {
            if (mflag) hipLaunchKernelGGL((regf_var_kernel<1, true>), grid, dim3(256), 0, s, tp);
            else hipLaunchKernelGGL((regf_var_kernel<1, false>), grid, dim3(256), 0, s, tp);
        } else {
            if (mflag) hipLaunchKernelGGL((regf_var_kernel<2, true>), grid, dim3(256), 0, s, tp);
            else hipLaunchKernelGGL((regf_var_kernel<2, false>), grid, dim3(256), 0, s, tp);
        }
        return;
    }
    const int key = cfg.threads * 10 + cfg.ipl;
    switch (key) {
        case 1281: launch_tile_t<128, 1>(mflag, grid, tp, s); break;
        case 1282: launch_tile_t<128, 2>(mflag, grid, tp, s); break;
        case 2562: launch_tile_t<256, 2>(mflag, grid, tp, s); break;
        case 5121: launch_tile_t<512, 1>(mflag, grid, tp, s); break;
        case 5122: launch_tile_t<512, 2>(mflag, grid, tp, s); break;
        default: launch_tile_t<256, 1>(mflag, grid, tp, s); break;
    }
}

void launch_items(const RegfConfig &cfg, bool mflag, int grid, const TileParams &tp,
                  const int4 *items, int nitems, int *queue, hipStream_t s) {
    const int key = cfg.threads * 10 + cfg.ipl;
    switch (key) {
        case 1281: launch_items_t<128, 1>(mflag, grid, tp, items, nitems, queue, s); break;
        case 1282: launch_items_t<128, 2>(mflag, grid, tp, items, nitems, queue, s); break;
        case 2562: launch_items_t<256, 2>(mflag, grid, tp, items, nitems, queue, s); break;
        case 5121: launch_items_t<512, 1>(mflag, grid, tp, items, nitems, queue, s); break;
        case 5122: launch_items_t<512, 2>(mflag, grid, tp, items, nitems, queue, s); break;
        default: launch_items_t<256, 1>(mflag, grid, tp, items, nitems, queue, s); break;
    }
}

}  // namespace

RegfConfig &regf_config() {
    static RegfConfig c;
    return c;
}

// 512-wide work-groups amortise the per-tile set-up best once there are enough
// i-blocks to fill the machine; smaller i-sets keep more j-chunks with 256.
RegfConfig regf_effective(int ni) {
    RegfConfig c = regf_config();
    if (c.threads == 0) c.threads = ni >= 65536 ? 512 : 256;
    if (c.prec != 0) {   // the precision variants run 256-wide, one i per lane
        c.threads = 256;
        c.ipl = 1;
    }
    return c;
}

// Carves the caller's work memory (no allocation): regions are 256-byte aligned, so a
// set of regions needs up to 255 bytes each beyond their sizes.
constexpr size_t ARENA_SLACK = 8 * 256;
struct Arena {
    unsigned char *p;
    size_t left;
    void *take(size_t n) {
        n = (n + 255) & ~(size_t)255;
        if (n > left) return nullptr;
        void *r = p;
        p += n;
        left -= n;
        return r;
    }
};

// na == nullptr: the library's own, growing work buffers.  With na, every work array
// comes from na (at most rows_bytes of it for the per-(i, chunk) rows), the i-set is
// not Morton-ordered (an ordering only, results do not depend on it), the number of
// j-chunks is capped so one i-block of rows fits, no kernel timers are recorded, and
// nothing is allocated.
struct NoAlloc {
    Arena *arena;
    size_t rows_bytes;
    int fix_items;
};

int regf_launch_impl(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                     int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s,
                     const NoAlloc *na);

int regf_launch(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s) {
    return regf_launch_impl(js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out, d_list, s,
                            nullptr);
}

int regf_launch_impl(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                     int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s,
                     const NoAlloc *na) {
    if (ni <= 0) return OK;
    if (js.n <= 0) {
        set_error(E_STATE, "gpunb_regf: no j-particles sent");
        return E_STATE;
    }
    if (lmax < 1 || lmax > NB_SORT_MAX + 1) {
        set_error(E_ARG, "gpunb_regf: lmax outside [1, 2049]");
        return E_ARG;
    }
    const RegfConfig cfg = regf_effective(ni);
    const int nbi = cfg.threads * cfg.ipl;
    const int nib = (ni + nbi - 1) / nbi;
    const int nj = js.n;
    const int unit = std::max(cfg.threads, JTILE);       // chunk granularity (j)
    const int nunit = (nj + unit - 1) / unit;
    const int nnbeff = std::max(0, std::min(nnbmax, lmax - 1));
    // small i-sets get chunks down to one unit each, so that even ni = 1 spreads over
    // the whole GPU
    int want = (cfg.min_groups + nib - 1) / nib;
    if (na) {
        // one i-block of rows (largest row size, FP64-pass items included) must fit the
        // caller's work memory
        const size_t rb = NPART * sizeof(double) + 2 * sizeof(int) +
                          ((size_t)nnbeff + 8 + (size_t)CAP_SOFT) * sizeof(int) + sizeof(int);
        const size_t cmax = (na->rows_bytes - ARENA_SLACK) / ((size_t)nbi * rb);
        if (cmax < 1) {
            set_error(E_ARG, "regf: work memory too small for one i-block");
            return E_ARG;
        }
        want = (int)std::min<size_t>((size_t)want, cmax);
    }
    int nchunk = std::max(1, std::min(want, nunit));
    const int upc = (nunit + nchunk - 1) / nchunk;        // units per chunk
    const int clen = upc * unit;
    nchunk = (nunit + upc - 1) / upc;
    // a chunk cannot hold more candidates than records
    const int caph = std::min(nnbeff + 8, clen);
    const int caps = std::min(CAP_SOFT, clen);
    // the finalize form depends on the whole call, never on the batching below, so a
    // batched call gives the same bits as an unbatched one
    const int fin_form = (ni <= FIN_WIDE_NI && nchunk >= 16) ? 8 : (nchunk >= 32 ? 4 : 1);

    // Work arrays hold one row per (i, chunk): FP64 partials, counts and the hard and
    // soft candidate slots.  Calls whose rows would exceed the slot budget run in
    // batches of i (whole i-blocks); chunks and results do not change.
    // FP32 partial sums per segment of a chunk: with long chunks, P segments of at
    // least SEG_MIN j (whole tiles), so that the FP64 pass recomputes a segment, not the
    // chunk.  FP64 pass: per (i, segment) a bucket entry
    const bool refine = cfg.refine != 0 && cfg.prec == 0;
    const int T = cfg.threads;
    auto row_size = [&](int nsegc) {
        return (size_t)nsegc * NPART * sizeof(double) + 2 * sizeof(int) +
               ((size_t)caph + (size_t)caps) * sizeof(int) + (refine ? nsegc * sizeof(int) : 0);
    };
    const size_t budget = na ? na->rows_bytes - ARENA_SLACK
                             : (size_t)std::max(cfg.slot_budget_mb, 1) << 20;
    int P = 1;
    // (at most 64 per i, the width of the finalize's greedy selection; with caller
    // work memory, one i-block of rows must still fit)
    while (refine && P < 8 && nchunk * 2 * P <= 64 && clen / (2 * P) >= SEG_MIN &&
           (clen / (2 * P)) % T == 0 && clen % (2 * P) == 0 &&
           (!na || (size_t)nchunk * nbi * row_size(2 * P) <= budget))
        P *= 2;
    const int seglen = clen / P;
    const size_t row_bytes = row_size(P);
    int nbatch_i = ni;
    if ((size_t)nchunk * (size_t)ni * row_bytes > budget) {
        const size_t rows_max = budget / row_bytes;
        nbatch_i = (int)std::max<size_t>((size_t)nbi, rows_max / (size_t)nchunk / nbi * nbi);
        nbatch_i = std::min(nbatch_i, ni);
    }
    const size_t rows = (size_t)nchunk * nbatch_i;
    // rows the finalize flags for the FP64 pass (FP32 production kernel only)
    const int fix_items = na ? na->fix_items : FIX_ITEMS;
    const size_t fix_cap = std::max((size_t)fix_items * 256, (size_t)ni);
    // items of the segment-selective pass: (slot, segment), candidate terms, FP64 sums
    // per sub-range
    const int nsub = (seglen + FIX_SUB - 1) / FIX_SUB;
    const size_t per_item = sizeof(int2) + 7 * sizeof(double) * (1 + (size_t)nsub);
    const int nseg = nchunk * P;
    // items per row: every segment where the whole call's rows allow it within
    // ITEM_BUDGET, at least MIN_ROW_ITEMS (from ni, not the batch: a batched call makes
    // the same choices)
    int row_items = (int)std::min<size_t>(
        nseg, std::max<size_t>(MIN_ROW_ITEMS, ITEM_BUDGET / ((size_t)ni * per_item)));
    size_t item_cap = (size_t)nbatch_i * row_items;
    double *part, *fixpart = nullptr, *cand = nullptr, *part64 = nullptr;
    int *cnt, *hslot, *sslot, *fix_rows = nullptr, *nfix = nullptr, *ctr = nullptr;
    int *seg_items = nullptr;
    int2 *items = nullptr;
    int4 *rowsel = nullptr;
    if (na) {
        Arena rows_ar{na->arena->p, std::min(na->rows_bytes, na->arena->left)};
        part = static_cast<double *>(rows_ar.take(rows * P * NPART * sizeof(double)));
        cnt = static_cast<int *>(rows_ar.take(rows * 2 * sizeof(int)));
        hslot = static_cast<int *>(rows_ar.take(rows * caph * sizeof(int)));
        sslot = static_cast<int *>(rows_ar.take(rows * caps * sizeof(int)));
        if (refine) seg_items = static_cast<int *>(rows_ar.take(rows * P * sizeof(int)));
        na->arena->take(na->rows_bytes);
        if (refine) {
            fix_rows = static_cast<int *>(na->arena->take((size_t)ni * sizeof(int)));
            ctr = static_cast<int *>(na->arena->take((6 + (size_t)nseg) * sizeof(int)));
            fixpart = static_cast<double *>(na->arena->take(fix_cap * 7 * sizeof(double)));
            rowsel = static_cast<int4 *>(na->arena->take((size_t)nbatch_i * sizeof(int4)));
            // the items take what is left
            const size_t left = na->arena->left > ARENA_SLACK ? na->arena->left - ARENA_SLACK : 0;
            row_items = (int)std::min<size_t>(row_items, left / ((size_t)nbatch_i * per_item));
            item_cap = std::max<size_t>((size_t)nbatch_i * row_items, 1);
            items = static_cast<int2 *>(na->arena->take(item_cap * sizeof(int2)));
            cand = static_cast<double *>(na->arena->take(item_cap * 7 * sizeof(double)));
            part64 = static_cast<double *>(na->arena->take(item_cap * nsub * 7 * sizeof(double)));
        }
        if (!part || !cnt || !hslot || !sslot ||
            (refine && (!fix_rows || !ctr || !fixpart || !rowsel || !seg_items || !items ||
                        !cand || !part64))) {
            set_error(E_ARG, "regf: the call does not fit the work memory");
            return E_ARG;
        }
    } else {
        // the work arrays grow with the call (14 GB at ni = 131,072 against 2^20 j): a
        // device that cannot hold them gets an error, not an abort
        part = static_cast<double *>(g_work.part.try_get(rows * P * NPART * sizeof(double)));
        cnt = static_cast<int *>(g_work.cnt.try_get(rows * 2 * sizeof(int)));
        hslot = static_cast<int *>(g_work.hslot.try_get(rows * caph * sizeof(int)));
        sslot = static_cast<int *>(g_work.sslot.try_get(rows * caps * sizeof(int)));
        if (refine) {
            fix_rows = static_cast<int *>(g_work.fixrows.try_get((size_t)ni * sizeof(int)));
            ctr = static_cast<int *>(g_work.nfix.try_get((6 + (size_t)nseg) * sizeof(int)));
            fixpart = static_cast<double *>(g_work.fixpart.try_get(fix_cap * 7 * sizeof(double)));
            seg_items = static_cast<int *>(g_work.segitems.try_get(rows * P * sizeof(int)));
            items = static_cast<int2 *>(g_work.items.try_get(item_cap * sizeof(int2)));
            cand = static_cast<double *>(g_work.delta.try_get(item_cap * 7 * sizeof(double)));
            part64 = static_cast<double *>(g_work.part64.try_get(item_cap * nsub * 7 * sizeof(double)));
            rowsel = static_cast<int4 *>(g_work.rowsel.try_get((size_t)nbatch_i * sizeof(int4)));
        }
        if (!part || !cnt || !hslot || !sslot ||
            (refine && (!fix_rows || !ctr || !fixpart || !seg_items || !items || !cand ||
                        !part64 || !rowsel))) {
            set_error(E_STATE, "gpunb_regf: device memory for the work arrays is not available "
                               "(lower nb6_set_regf_slot_budget)");
            return E_STATE;
        }
    }
    int *seg_cnt = ctr ? ctr + 6 : nullptr;   // items per chunk of the current batch
    // counters: [0] whole-row passes, [1] chunk items, [2] rows with items (this batch),
    // [3] rows with items and [4] items of the earlier batches
    int *nitems = ctr ? ctr + 1 : nullptr, *nrows = ctr ? ctr + 2 : nullptr;
    g_last_chunks = nchunk;
    g_last_batches = (ni + nbatch_i - 1) / nbatch_i;
    if (refine) {
        nfix = ctr;
        g_nfix = ctr;
        hipLaunchKernelGGL(zero_ints_kernel, dim3((6 + nseg + 255) / 256), dim3(256), 0, s, ctr,
                           6 + nseg);
    }
    g_last_refine = refine;

    // Morton-ordered i: a wavefront's particles are neighbours in space, so their
    // candidates fall on the same j and most tiles are far from all of them
    const int *imap = !na && cfg.isort_min > 0 && ni >= cfg.isort_min
                          ? jset_sort_i(js, ni, ni_stride, d_i, g_work, s)
                          : nullptr;

    KTimer dummy;   // never started: with caller work memory no events are recorded
    KTimer &t1 = na ? dummy : ktimer("regf_tile");
    KTimer &t2 = na ? dummy : ktimer("regf_finalize");
    KTimer &t4 = na ? dummy : ktimer("regf_chunkfix");
    const bool timed = !na;
    if (timed) {   // the FP64-pass timers report this call only (or -1: no such pass)
        t4.armed = false;
        ktimer("regf_fix").armed = false;
    }
    for (int b0 = 0; b0 < ni; b0 += nbatch_i) {
        const int nb = std::min(nbatch_i, ni - b0);
        // slots [b0, b0 + nb): through imap the rows are absolute; without it the
        // SoA inputs and outputs are offset by b0
        const int *bmap = imap ? imap + b0 : nullptr;
        const double *bi = imap ? d_i : d_i + b0;
        double *bout = imap ? d_out : d_out + b0;
        int *blist = imap ? d_list : d_list + (size_t)b0 * lmax;
        const int bnib = (nb + nbi - 1) / nbi;

        TileParams tp{};
        tp.ib = bi; tp.ni = nb; tp.ib_stride = ni_stride; tp.imap = bmap;
        tp.rec32 = js.rec32.as<float4>(); tp.torg = js.torg.as<double>(); tp.nj = nj;
        tp.tbox = js.tbox.as<float>();
        tp.far_test = cfg.far_test;
        tp.rec64 = js.rec64.as<double>();
        tp.chunk_len = clen;
        tp.nib = bnib;
        tp.mmax = js.mmax;
        tp.part = part; tp.cnt = cnt;
        tp.hslot = hslot; tp.caph = caph; tp.sslot = sslot; tp.caps = caps;
        tp.fnear = cfg.near;
        tp.nseg_c = P;
        tp.seg_tiles = P > 1 ? seglen / T : 0;
        dim3 grid(bnib * nchunk);
        if (timed && b0 == 0) t1.start(s);
        launch_tile(cfg, m_flag != 0, grid, tp, s);
        NB6_HIP(hipGetLastError());
        if (timed && b0 + nb >= ni) t1.stop(s);

        FinParams fp{};
        fp.ib = bi; fp.ni = nb; fp.ib_stride = ni_stride; fp.imap = bmap;
        fp.rec64 = js.rec64.as<double>(); fp.perm = js.perm.as<int>(); fp.jbase = nullptr;
        fp.nj = nj;
        fp.nchunk = nchunk;
        fp.nprow = nseg;
        fp.part = part; fp.cnt = cnt;
        fp.hslot = hslot; fp.caph = caph; fp.sslot = sslot; fp.caps = caps;
        fp.lmax = lmax; fp.nnbmax = nnbeff;
        fp.out = bout; fp.out_stride = ni_stride; fp.list = blist;
        fp.refine = refine ? 1 : 0;
        fp.row_base = imap ? 0 : b0;
        fp.fix_slot = 0;
        for (int q = 0; q < NERR; q++) fp.kf[q] = nj <= cfg.small_nj ? cfg.kref_small[q] : cfg.kref[q];
        fp.fix_rows = fix_rows;
        fp.nfix = nfix;
        fp.items = items; fp.nitems = nitems;
        fp.item_cap = (int)std::min<size_t>(item_cap, INT32_MAX);
        fp.row_items = row_items > 0 ? row_items : -1;
        fp.seg_items = seg_items; fp.seg_cnt = seg_cnt;
        fp.padoff = nullptr; fp.nseg = 0;
        fp.dbg = g_dbg ? (imap ? g_dbg : g_dbg + b0) : nullptr;
        fp.rowsel = rowsel; fp.nrows = nrows;
        if (refine && b0 > 0)   // item counters per batch (the rows' total is kept in [3])
            hipLaunchKernelGGL(next_batch_kernel, dim3(1), dim3(256), 0, s, ctr, nseg);
        if (timed && b0 == 0) t2.start(s);
        if (fin_form == 8) {
            // few i: eight wavefronts per i
            if (m_flag)
                hipLaunchKernelGGL((regf_finalize_wide_kernel<true, 8>), dim3(nb), dim3(512), 0, s, fp);
            else
                hipLaunchKernelGGL((regf_finalize_wide_kernel<false, 8>), dim3(nb), dim3(512), 0, s, fp);
        } else if (fin_form == 4) {
            // many chunks per i: four wavefronts per i
            if (m_flag)
                hipLaunchKernelGGL((regf_finalize_wide_kernel<true, 4>), dim3(nb), dim3(256), 0, s, fp);
            else
                hipLaunchKernelGGL((regf_finalize_wide_kernel<false, 4>), dim3(nb), dim3(256), 0, s, fp);
        } else {
            const dim3 fgrid((nb + 3) / 4);
            if (m_flag)
                hipLaunchKernelGGL(regf_finalize_kernel<true>, fgrid, dim3(256), 0, s, fp);
            else
                hipLaunchKernelGGL(regf_finalize_kernel<false>, fgrid, dim3(256), 0, s, fp);
        }
        NB6_HIP(hipGetLastError());
        if (timed && b0 + nb >= ni) t2.stop(s);
        if (refine) {
            // chunk-selective FP64 pass of this batch's flagged rows, while its work
            // arrays (partials, slots) are still there
            ChunkFixParams cp{};
            cp.ib = bi; cp.ib_stride = ni_stride; cp.imap = bmap;
            cp.rec64 = js.rec64.as<double>(); cp.nj = nj; cp.clen = seglen; cp.nchunk = nseg;
            cp.ni = nb; cp.nsub = nsub; cp.nseg_c = P;
            cp.part = part; cp.cnt = cnt; cp.hslot = hslot; cp.caph = caph;
            cp.sslot = sslot; cp.caps = caps;
            cp.items = items; cp.nitems = nitems; cp.item_cap = fp.item_cap;
            cp.row_items = std::max(row_items, 1);
            cp.seg_items = seg_items; cp.seg_cnt = seg_cnt;
            cp.cand = cand; cp.part64 = part64;
            cp.rowsel = rowsel; cp.nrows = nrows;
            cp.out = bout; cp.out_stride = ni_stride;
            cp.padoff = nullptr; cp.maxseg = 0;
            if (timed && b0 == 0) t4.start(s);
            // whole-row FP64 pass (rows with NaN sums, rows with more segments than
            // they may store; their count stays on the device) with the last batch,
            // when every batch has listed its rows
            const bool last = b0 + nb >= ni;
            FixParams xp{};
            xp.ib = d_i; xp.ib_stride = ni_stride;
            xp.rows = fix_rows; xp.nfix = nfix;
            xp.rec64 = js.rec64.as<double>(); xp.j0 = 0; xp.j1 = nj;
            xp.part = fixpart;
            xp.out = d_out; xp.out_stride = ni_stride;
            xp.items = fix_items;
            // grids no larger than the work there can be (small calls launch small
            // grids); whole rows: at most (ni / 256 row groups) x (j tiles) items
            const int gc = (int)std::min<size_t>(CHUNKFIX_GROUPS,
                                                 ((size_t)nb * cp.row_items + 3) / 4);
            const int gs = (int)std::min<size_t>(
                CHUNKFIX_GROUPS, (size_t)nseg * (((size_t)nb + 255) / 256) * (size_t)nsub);
            const int gf = last ? (int)std::min<size_t>(fix_items, (((size_t)ni + 255) / 256) *
                                                      (((size_t)nj + JTILE - 1) / JTILE))
                                : 0;
            const int ga = (nb + 3) / 4;
            const int gr = last ? (ni + 3) / 4 : 0;
            if (nb <= PASS_MERGE_NI) {
                // small batches: two launches
                if (m_flag)
                    hipLaunchKernelGGL(regf_pass1_kernel<true>, dim3(gs + gc + gf), dim3(256), 0,
                                       s, cp, xp, gs, gc);
                else
                    hipLaunchKernelGGL(regf_pass1_kernel<false>, dim3(gs + gc + gf), dim3(256), 0,
                                       s, cp, xp, gs, gc);
                NB6_HIP(hipGetLastError());
                hipLaunchKernelGGL(regf_pass2_kernel, dim3(ga + gr), dim3(256), 0, s, cp, xp, ga);
            } else {
                // large ones: each part at its own occupancy (the same work, the same bits)
                if (m_flag) {
                    hipLaunchKernelGGL(regf_segfix_kernel<true>, dim3(gs), dim3(256), 0, s, cp);
                    hipLaunchKernelGGL(regf_cand_kernel<true>, dim3(gc), dim3(256), 0, s, cp);
                    if (gf) hipLaunchKernelGGL(regf_fix_kernel<true>, dim3(gf), dim3(256), 0, s, xp);
                } else {
                    hipLaunchKernelGGL(regf_segfix_kernel<false>, dim3(gs), dim3(256), 0, s, cp);
                    hipLaunchKernelGGL(regf_cand_kernel<false>, dim3(gc), dim3(256), 0, s, cp);
                    if (gf) hipLaunchKernelGGL(regf_fix_kernel<false>, dim3(gf), dim3(256), 0, s, xp);
                }
                NB6_HIP(hipGetLastError());
                hipLaunchKernelGGL(regf_segapply_kernel, dim3(ga), dim3(256), 0, s, cp);
                if (gr) hipLaunchKernelGGL(regf_fix_reduce_kernel, dim3(gr), dim3(256), 0, s, xp);
            }
            NB6_HIP(hipGetLastError());
            if (timed && last) t4.stop(s);
        }
    }
    return OK;
}

// ------------------------------------------------------------------ no-alloc form
// Work memory of nb6_regf_device_noalloc: one array per GPU, held until the process
// ends and shared by every call on that GPU (calls that overlap on one GPU race).
// Host code never passes `scratch` itself to a kernel: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr size_t SCRATCH_BYTES = 256ull << 20;
constexpr size_t SCRATCH_ROWS = 232ull << 20;   // of which the per-(i, chunk) rows
constexpr int FIX_ITEMS_NA = 512;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
static unsigned char *g_scratch_ptr[256] = {};

int regf_noalloc_setup() {
    int dev = -1;
    NB6_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) {
        set_error(E_NODEV, "nb6_regf_device_noalloc: device index out of range");
        return E_NODEV;
    }
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(E_STATE, std::string("nb6_regf_device_noalloc: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
            return E_STATE;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    return OK;
}

int regf_launch_noalloc(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                        int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s) {
    if (ni > REGF_NOALLOC_NI_MAX) {
        set_error(E_ARG, "nb6_regf_device_noalloc: ni above 65536");
        return E_ARG;
    }
    if (regf_config().prec != 0) {
        set_error(E_STATE, "nb6_regf_device_noalloc: only the FP32 production kernel");
        return E_STATE;
    }
    const int rc = regf_noalloc_setup();
    if (rc != OK) return rc;
    int dev = 0;
    NB6_HIP(hipGetDevice(&dev));
    Arena ar{g_scratch_ptr[dev], SCRATCH_BYTES};
    const NoAlloc na{&ar, SCRATCH_ROWS, FIX_ITEMS_NA};
    return regf_launch_impl(js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out, d_list, s, &na);
}

// Batched form (ensemble.hip): one persistent launch over work items, nchunk = 1,
// work rows indexed by slot, caller rows by imap[slot].
int regf_items_launch(const JSet &js, int nslots, const int *imap, int stride, const double *d_i,
                      const int4 *items, int nitems, int *queue, const int *jbase,
                      const int *padoff, int nc, int maxseg, int lmax, int nnbmax, int m_flag,
                      double *d_out, int *d_list, int grid, RegfWork &w, hipStream_t s) {
    if (lmax < 1 || lmax > NB_SORT_MAX + 1) {
        set_error(E_ARG, "nb6_ens_regf_device: lmax outside [1, 2049]");
        return E_ARG;
    }
    const RegfConfig cfg = regf_effective(0);   // clusters are small: 256 unless set
    const int nnbeff = std::max(0, std::min(nnbmax, lmax - 1));
    const int caph = nnbeff + 8;
    const int caps = CAP_SOFT;
    const size_t rows = (size_t)nslots;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    NB6_HIP(hipStreamIsCapturing(s, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    const bool refine = cfg.refine != 0;
    // FP32 partial sums per segment of SEG_ENS j of a cluster (whole tiles of every
    // work-group width; at most 64 per cluster); FP64 pass: items (row, segment of its
    // cluster), bucketed per (cluster, segment)
    int sege = SEG_ENS;
    while ((maxseg + sege - 1) / sege > 64) sege *= 2;
    const int nsegc = refine ? std::max(1, (maxseg + sege - 1) / sege) : 1;
    const int nsub = (sege + FIX_SUB - 1) / FIX_SUB;
    const size_t item_cap = std::max<size_t>(rows / 2, 1);
    const size_t nbk = (size_t)nc * nsegc;
    const size_t ctr_bytes = (6 + nbk) * sizeof(int);
    if (capturing && (w.part.bytes < rows * nsegc * NPART * sizeof(double) ||
                      w.hslot.bytes < rows * caph * sizeof(int) ||
                      w.sslot.bytes < rows * caps * sizeof(int) || w.cnt.bytes < rows * 2 * sizeof(int) ||
                      (refine && (w.fixrows.bytes < rows * sizeof(int) || w.nfix.bytes < ctr_bytes ||
                                  w.segitems.bytes < rows * nsegc * sizeof(int) ||
                                  w.items.bytes < item_cap * sizeof(int2) ||
                                  w.delta.bytes < item_cap * 7 * sizeof(double) ||
                                  w.part64.bytes < item_cap * nsub * 7 * sizeof(double) ||
                                  w.rowsel.bytes < rows * sizeof(int4))))) {
        set_error(E_STATE, "regf_items_launch: work buffers must be sized before stream capture");
        return E_STATE;
    }
    double *part = static_cast<double *>(w.part.get(rows * nsegc * NPART * sizeof(double)));
    int *cnt = static_cast<int *>(w.cnt.get(rows * 2 * sizeof(int)));
    int *hslot = static_cast<int *>(w.hslot.get(rows * caph * sizeof(int)));
    int *sslot = static_cast<int *>(w.sslot.get(rows * caps * sizeof(int)));
    int *fix_slots = nullptr, *nfix = nullptr, *ctr = nullptr, *seg_items = nullptr;
    int2 *fitems = nullptr;
    double *cand = nullptr, *part64 = nullptr;
    int4 *rowsel = nullptr;
    if (refine) {
        fix_slots = static_cast<int *>(w.fixrows.get(rows * sizeof(int)));
        ctr = static_cast<int *>(w.nfix.get(ctr_bytes));
        seg_items = static_cast<int *>(w.segitems.get(rows * nsegc * sizeof(int)));
        fitems = static_cast<int2 *>(w.items.get(item_cap * sizeof(int2)));
        cand = static_cast<double *>(w.delta.get(item_cap * 7 * sizeof(double)));
        part64 = static_cast<double *>(w.part64.get(item_cap * nsub * 7 * sizeof(double)));
        rowsel = static_cast<int4 *>(w.rowsel.get(rows * sizeof(int4)));
        nfix = ctr;   // [0] whole rows, [1] items, [2] rows with items, [6 ..] per bucket
        hipLaunchKernelGGL(zero_ints_kernel, dim3((unsigned)((6 + nbk + 255) / 256)), dim3(256), 0,
                           s, ctr, (int)(6 + nbk));
    }
    TileParams tp{};
    tp.ib = d_i; tp.ni = nslots; tp.ib_stride = stride; tp.imap = imap;
    tp.rec32 = js.rec32.as<float4>(); tp.torg = js.torg.as<double>(); tp.nj = js.n;
    tp.tbox = js.tbox.as<float>();
    tp.far_test = cfg.far_test;
    tp.rec64 = js.rec64.as<double>();
    tp.chunk_len = js.n;
    tp.mmax = js.mmax;
    tp.part = part; tp.cnt = cnt;
    tp.hslot = hslot; tp.caph = caph; tp.sslot = sslot; tp.caps = caps;
    tp.fnear = cfg.near;
    tp.nseg_c = nsegc;
    tp.seg_tiles = nsegc > 1 ? sege / cfg.threads : 0;
    KTimer &t1 = ktimer("regf_items");
    if (!capturing) t1.start(s);
    launch_items(cfg, m_flag != 0, grid, tp, items, nitems, queue, s);
    NB6_HIP(hipGetLastError());
    if (!capturing) t1.stop(s);
    FinParams fp{};
    fp.ib = d_i; fp.ni = nslots; fp.ib_stride = stride; fp.imap = imap;
    fp.rec64 = js.rec64.as<double>(); fp.perm = js.perm.as<int>(); fp.jbase = jbase; fp.nj = js.n;
    fp.nchunk = 1;
    fp.nprow = nsegc;
    fp.part = part; fp.cnt = cnt;
    fp.hslot = hslot; fp.caph = caph; fp.sslot = sslot; fp.caps = caps;
    fp.lmax = lmax; fp.nnbmax = nnbeff;
    fp.out = d_out; fp.out_stride = stride; fp.list = d_list;
    fp.refine = refine ? 1 : 0; fp.row_base = 0; fp.fix_slot = 1;
    fp.items = fitems; fp.nitems = ctr ? ctr + 1 : nullptr;
    fp.item_cap = (int)std::min<size_t>(item_cap, INT32_MAX);
    fp.seg_items = seg_items; fp.seg_cnt = ctr ? ctr + 6 : nullptr;
    fp.dbg = g_dbg;
    fp.rowsel = rowsel; fp.nrows = ctr ? ctr + 2 : nullptr;
    fp.padoff = padoff; fp.nseg = nc;
    fp.row_items = 0;                    // items appended (up to item_cap)
    for (int q = 0; q < NERR; q++) fp.kf[q] = cfg.kref_ens[q];
    fp.fix_rows = fix_slots; fp.nfix = nfix;
    const dim3 fgrid((nslots + 3) / 4);
    if (m_flag)
        hipLaunchKernelGGL(regf_finalize_kernel<true>, fgrid, dim3(256), 0, s, fp);
    else
        hipLaunchKernelGGL(regf_finalize_kernel<false>, fgrid, dim3(256), 0, s, fp);
    NB6_HIP(hipGetLastError());
    if (refine) {
        // FP64 pass of the flagged rows: items bucketed by cluster, each cluster's records
        // through LDS once per 256 of its rows
        ChunkFixParams cp{};
        cp.ib = d_i; cp.ib_stride = stride; cp.imap = imap;
        cp.rec64 = js.rec64.as<double>(); cp.nj = js.n; cp.clen = sege; cp.nchunk = (int)nbk;
        cp.ni = nslots; cp.nsub = nsub; cp.nseg_c = nsegc;
        cp.part = part; cp.cnt = cnt; cp.hslot = hslot; cp.caph = caph;
        cp.sslot = sslot; cp.caps = caps;
        cp.items = fitems; cp.nitems = ctr + 1; cp.item_cap = fp.item_cap;
        cp.row_items = nsegc;            // a row's items: at most one per segment
        cp.seg_items = seg_items; cp.seg_cnt = ctr + 6;
        cp.cand = cand; cp.part64 = part64;
        cp.rowsel = rowsel; cp.nrows = ctr + 2;
        cp.out = d_out; cp.out_stride = stride;
        cp.padoff = padoff; cp.maxseg = maxseg;
        if (m_flag) {
            hipLaunchKernelGGL(regf_cand_kernel<true>, dim3(CHUNKFIX_GROUPS), dim3(256), 0, s, cp);
            hipLaunchKernelGGL(regf_segfix_kernel<true>, dim3(CHUNKFIX_GROUPS), dim3(256), 0, s, cp);
        } else {
            hipLaunchKernelGGL(regf_cand_kernel<false>, dim3(CHUNKFIX_GROUPS), dim3(256), 0, s, cp);
            hipLaunchKernelGGL(regf_segfix_kernel<false>, dim3(CHUNKFIX_GROUPS), dim3(256), 0, s, cp);
        }
        NB6_HIP(hipGetLastError());
        hipLaunchKernelGGL(regf_segapply_kernel, dim3(std::min((nslots + 3) / 4, 16384)), dim3(256),
                           0, s, cp);
        NB6_HIP(hipGetLastError());
        // rows whose items did not fit and rows with NaN sums: whole rows
        FixRowParams xp{};
        xp.ib = d_i; xp.ib_stride = stride; xp.imap = imap;
        xp.slots = fix_slots; xp.nfix = nfix;
        xp.padoff = padoff; xp.nc = nc;
        xp.rec64 = js.rec64.as<double>();
        xp.out = d_out; xp.out_stride = stride;
        const dim3 xgrid(1024);
        if (m_flag)
            hipLaunchKernelGGL(regf_fix_rows_kernel<true>, xgrid, dim3(256), 0, s, xp);
        else
            hipLaunchKernelGGL(regf_fix_rows_kernel<false>, xgrid, dim3(256), 0, s, xp);
        NB6_HIP(hipGetLastError());
        g_ens_nfix = nfix;
    }
    return OK;
}

void regf_items_forget() { g_ens_nfix = nullptr; }
void regf_set_debug_bounds(double *d) { g_dbg = d; }

int regf_items_refined() {
    if (!g_ens_nfix) return 0;
    int h[3] = {0, 0, 0};
    NB6_HIP(hipDeviceSynchronize());
    NB6_HIP(hipMemcpy(h, g_ens_nfix, sizeof(h), hipMemcpyDeviceToHost));
    return h[0] + h[2];
}

int regf_last_chunks() { return g_last_chunks; }
int regf_last_batches() { return g_last_batches; }
int regf_last_refined(int *chunk_items) {
    if (chunk_items) *chunk_items = 0;
    if (!g_last_refine || !g_nfix) return 0;
    int h[6] = {0, 0, 0, 0, 0, 0};
    NB6_HIP(hipDeviceSynchronize());   // the call may have run on any stream
    NB6_HIP(hipMemcpy(h, g_nfix, sizeof(h), hipMemcpyDeviceToHost));
    if (chunk_items) *chunk_items = h[1] + h[4];
    return h[0] + h[2] + h[3];
}

// Sweep work-group size (= LDS j-tile) 128/256/512, i-particles per lane 1/2 and the
// minimum number of work-groups (occupancy through j-chunking) on the caller's own
// inputs; the fastest becomes the configuration.  best = {threads, ipl, min_groups, us}.
int regf_autotune(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                  int m_flag, double *d_out, int *d_list, int *best, hipStream_t s) {
    RegfConfig &cfg = regf_config();
    const RegfConfig keep = cfg;
    if (cfg.prec != 0) { set_error(E_STATE, "nb6_regf_autotune: set the FP32 kernel first"); return E_STATE; }
    float best_ms = 1e30f;
    int bt = 256, bipl = 1, bmg = 4096;
    const int threads[3] = {128, 256, 512};
    const bool timing = kernel_timing();
    set_kernel_timing(true);   // the tuner times with the kernel timers
    for (int ti = 0; ti < 3; ti++)
        for (int ipl = 1; ipl <= 2; ipl++)
            for (int mg : {2048, 4096, 8192}) {
                cfg.threads = threads[ti];
                cfg.ipl = ipl;
                cfg.min_groups = mg;
                float ms = 0.f;
                for (int rep = 0; rep < 2; rep++) {   // the first run warms up
                    const int rc = regf_launch(js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out,
                                               d_list, s);
                    if (rc != OK) { cfg = keep; set_kernel_timing(timing); return rc; }
                    ms = (float)(ktimer("regf_tile").ms() + ktimer("regf_finalize").ms());
                }
                if (ms < best_ms) { best_ms = ms; bt = threads[ti]; bipl = ipl; bmg = mg; }
            }
    set_kernel_timing(timing);
    cfg.threads = bt;
    cfg.ipl = bipl;
    cfg.min_groups = bmg;
    if (best) { best[0] = bt; best[1] = bipl; best[2] = bmg; best[3] = (int)(best_ms * 1000.f); }
    return OK;
}

void regf_release() {
    g_work.release();
    g_nfix = nullptr;
    g_last_refine = false;
}

}  // namespace nb6
