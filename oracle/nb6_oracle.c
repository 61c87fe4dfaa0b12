/*
 * nb6_oracle.c -- CPU oracle for the regular-force / neighbour-list / potential /
 * irregular-force path.  TEST INFRASTRUCTURE ONLY: this file is the checker used by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  It is
 * never linked into the product library and the product never calls it.
 *
 * Everything here is a plain FP64 restatement of the reference algorithm, written
 * from the reference's semantics (not its code), with the formula order fixed so that
 * the GPU library's FP64 neighbour decision reproduces it bit for bit:
 *
 *   orc_regf  : gpunb_regf_ semantics
 *               - pair terms        gpunb.velocity.cu:163-209 (dev_gravity),
 *                                   gpunb.velocity.cu:212-258 (dev_gravity_m),
 *                                   reg.avx.cpp:250-305
 *               - neighbour test    min(r2, |dx + dtr*dv|^2) < h2   (or < m_j*h2),
 *                                   gpunb.velocity.cu:188 / :236, reg.avx.cpp:265-273
 *               - neighbours add nothing to acc, jerk or pot (rinv = 0),
 *                                   gpunb.velocity.cu:192, reg.avx.cpp:292
 *               - lists ascending, 0-based, row layout list[i*lmax] = count,
 *                 overflow count > nnbmax -> -count, entries not written,
 *                                   reg.avx.cpp:316-327
 *   orc_pot   : gpupot_ semantics, phi_i = sum_{j, r2>0} m_j / r_ij (positive),
 *                                   gpupot.gpu.cu:32-57, pot.avx.cpp:112-136
 *   orc_firr  : irr_simd_firr_vec_ semantics: neighbour-only force and jerk with
 *               third-order prediction of i and of every j to time ti, nearest
 *               neighbour as the 1-based index of min r2,
 *                                   irr.avx.cpp:142-156 (prediction),
 *                                   irr.avx.cpp:303-332 (pair terms),
 *                                   irr.avx.cpp:271-301 (nnbid),
 *                                   nbint.F:19-58 (FP64 form of the same sum)
 *
 * Deliberate, documented differences from the reference:
 *   - a pair with r2 == 0 that is not a neighbour (only possible when h2 <= 0) is
 *     skipped here; the reference divides by zero and returns inf/NaN.
 *   - orc_firr returns nnbid = 0 for an empty list (the reference returns an
 *     unspecified integer, irr.avx.cpp:246-251,293).
 *
 * Compile with -ffp-contract=off: the neighbour decision must be the plain
 * multiply/add sequence written below.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------
 * Regular force, FP64.
 * ------------------------------------------------------------------------- */
void orc_regf(int nj, const double *mj, const double *xj /*[nj][3]*/,
              const double *vj /*[nj][3]*/,
              int ni, const double *h2, const double *dtr,
              const double *xi /*[ni][3]*/, const double *vi /*[ni][3]*/,
              double *acc /*[ni][3]*/, double *jrk /*[ni][3]*/, double *pot,
              int lmax, int nnbmax, int *list /*[ni][lmax]*/, int m_flag)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < ni; i++) {
        const double x0 = xi[3 * i + 0], y0 = xi[3 * i + 1], z0 = xi[3 * i + 2];
        const double u0 = vi[3 * i + 0], v0 = vi[3 * i + 1], w0 = vi[3 * i + 2];
        const double hh = h2[i], dt = dtr[i];
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, ph = 0;
        int *row = list + (size_t)i * lmax;
        int nnb = 0;
        for (int j = 0; j < nj; j++) {
            const double dx = xj[3 * j + 0] - x0;
            const double dy = xj[3 * j + 1] - y0;
            const double dz = xj[3 * j + 2] - z0;
            const double dvx = vj[3 * j + 0] - u0;
            const double dvy = vj[3 * j + 1] - v0;
            const double dvz = vj[3 * j + 2] - w0;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double dxp = dx + dt * dvx;
            const double dyp = dy + dt * dvy;
            const double dzp = dz + dt * dvz;
            const double r2p = dxp * dxp + dyp * dyp + dzp * dzp;
            const double rmin = r2 < r2p ? r2 : r2p;
            const double thr = m_flag ? mj[j] * hh : hh;
            if (rmin < thr) {
                if (nnb < nnbmax) row[1 + nnb] = j;
                nnb++;
                continue;
            }
            if (r2 == 0.0) continue;
            const double rinv2 = 1.0 / r2;
            const double rinv = sqrt(rinv2);
            const double mr1 = mj[j] * rinv;
            const double mr3 = mr1 * rinv2;
            const double rv = dx * dvx + dy * dvy + dz * dvz;
            const double al = -3.0 * rv * rinv2;
            ph += mr1;
            ax += mr3 * dx;
            ay += mr3 * dy;
            az += mr3 * dz;
            jx += mr3 * (dvx + al * dx);
            jy += mr3 * (dvy + al * dy);
            jz += mr3 * (dvz + al * dz);
        }
        acc[3 * i + 0] = ax; acc[3 * i + 1] = ay; acc[3 * i + 2] = az;
        jrk[3 * i + 0] = jx; jrk[3 * i + 1] = jy; jrk[3 * i + 2] = jz;
        pot[i] = ph;
        row[0] = (nnb > nnbmax) ? -nnb : nnb;
    }
}

/* Neighbour decision only (same formula as above), for one pair.  Used by tests to
 * classify pairs near the threshold band. Returns 1 if j is a neighbour of i. */
int orc_is_neighbour(const double *xj, const double *vj, double mj,
                     const double *xi, const double *vi, double h2, double dtr,
                     int m_flag)
{
    const double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
    const double dvx = vj[0] - vi[0], dvy = vj[1] - vi[1], dvz = vj[2] - vi[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double dxp = dx + dtr * dvx, dyp = dy + dtr * dvy, dzp = dz + dtr * dvz;
    const double r2p = dxp * dxp + dyp * dyp + dzp * dzp;
    const double rmin = r2 < r2p ? r2 : r2p;
    const double thr = m_flag ? mj * h2 : h2;
    return rmin < thr;
}

/* ---------------------------------------------------------------------------
 * Potential, FP64:  pot[k] = sum_{j, r2>0} m_j / r(istart-1+k, j), k < ni.
 * istart is 1-based like gpupot_ (gpupot.gpu.cu:35, :101-105).
 * ------------------------------------------------------------------------- */
void orc_pot(int n, const double *m, const double *x /*[n][3]*/,
             int istart, int ni, double *pot)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int k = 0; k < ni; k++) {
        const int i = istart - 1 + k;
        const double x0 = x[3 * i + 0], y0 = x[3 * i + 1], z0 = x[3 * i + 2];
        double ph = 0.0;
        for (int j = 0; j < n; j++) {
            const double dx = x[3 * j + 0] - x0;
            const double dy = x[3 * j + 1] - y0;
            const double dz = x[3 * j + 2] - z0;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0.0) ph += m[j] / sqrt(r2);
        }
        pot[k] = ph;
    }
}

/* ---------------------------------------------------------------------------
 * Irregular (neighbour-list) force with prediction, FP64.
 *   state arrays are 0-based over particles [0, nmax);
 *   list is [nmax][lmax] with list[a*lmax] = nnb and entries 1-based particle
 *   addresses (the form irr_simd_set_list_ receives, irr.avx.cpp:590-595);
 *   addr[] are 1-based active particle addresses.
 * Prediction (irr.avx.cpp:142-156, xbpredall.F:55-59):
 *   s = ti - t0;  x = x0 + s*(v0 + s*(a2 + s*j6));  v = v0 + 2s*(a2 + 1.5s*j6)
 * ------------------------------------------------------------------------- */
static inline void orc_predict(double ti, const double *x0, const double *v0,
                               const double *a2, const double *j6, double t0,
                               double *xp, double *vp)
{
    const double s = ti - t0;
    const double s1 = s + s, s2 = 1.5 * s;
    for (int k = 0; k < 3; k++) {
        xp[k] = x0[k] + s * (v0[k] + s * (a2[k] + s * j6[k]));
        vp[k] = v0[k] + s1 * (a2[k] + s2 * j6[k]);
    }
}

void orc_firr(double ti, int ni, const int *addr,
              const double *x0, const double *v0, const double *a2,
              const double *j6, const double *m, const double *t0,
              const int *list, int lmax,
              double *acc /*[ni][3]*/, double *jrk /*[ni][3]*/, int *nnbid)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < ni; k++) {
        const int i = addr[k] - 1;
        double xi[3], vi[3];
        orc_predict(ti, x0 + 3 * i, v0 + 3 * i, a2 + 3 * i, j6 + 3 * i, t0[i], xi, vi);
        const int *row = list + (size_t)i * lmax;
        const int nnb = row[0];
        double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0;
        double best = INFINITY;
        int bestj = -1;
        for (int l = 0; l < nnb; l++) {
            const int j = row[1 + l] - 1;
            double xj[3], vj[3];
            orc_predict(ti, x0 + 3 * j, v0 + 3 * j, a2 + 3 * j, j6 + 3 * j, t0[j], xj, vj);
            const double dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
            const double dvx = vj[0] - vi[0], dvy = vj[1] - vi[1], dvz = vj[2] - vi[2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < best || (r2 == best && j < bestj)) { best = r2; bestj = j; }
            const double rinv2 = 1.0 / r2;
            const double rinv = sqrt(rinv2);
            const double mr3 = m[j] * rinv * rinv2;
            const double rv = dx * dvx + dy * dvy + dz * dvz;
            const double al = -3.0 * rv * rinv2;
            ax += mr3 * dx;
            ay += mr3 * dy;
            az += mr3 * dz;
            jx += mr3 * (dvx + al * dx);
            jy += mr3 * (dvy + al * dy);
            jz += mr3 * (dvz + al * dz);
        }
        acc[3 * k + 0] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;
        jrk[3 * k + 0] = jx; jrk[3 * k + 1] = jy; jrk[3 * k + 2] = jz;
        nnbid[k] = bestj + 1;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
