// Total-energy terms on the GPU: the GPU branch of ENERGY (energy.F:28-62, :119-124) with
// the differential potential corrections of PHICOR (phicor.f:1-76).
//
//   * GPUPOT over the bodies IFIRST..NTOT (singles and KS centres of mass), here the
//     pot_tile_kernel of pot.hip on a packed copy of their positions and masses;
//   * each KS component takes the potential of its centre of mass (energy.F:36-44);
//   * PHICOR per body: a single I corrects, for every c.m. neighbour J closer than
//     sqrt(CMSEP2 RMIN2), the c.m. term by the two component terms; a c.m. body I
//     corrects the potentials of its two components over the perturbers of LIST(.,I1)
//     (phicor.f:30-71);
//   * POT = 1/2 sum m PHIDBL over singles and components, ZKIN = 1/2 sum m v^2 over
//     IFIRST..NTOT, EBIN = sum m1 m2 HT / m_cm over the pairs with a positive c.m. mass.
// One thread per body for PHICOR (its neighbour loop is O(list length)); the three sums
// are FP64 with a fixed reduction tree, so repeated calls give the same bits.
#include "nb6_internal.h"
#include "../../include/nb6gpu.h"

#include <algorithm>

namespace nb6 {

namespace {

constexpr int EB = 256;

__global__ __launch_bounds__(EB) void energy_pack_kernel(const double *__restrict__ x,
                                                         const double *__restrict__ body,
                                                         int ifirst, int nnt,
                                                         double *__restrict__ xs) {
    const int k = blockIdx.x * EB + threadIdx.x;
    if (k >= nnt) return;
    const size_t i = (size_t)(ifirst - 1 + k);
    xs[k] = x[3 * i];
    xs[(size_t)nnt + k] = x[3 * i + 1];
    xs[2 * (size_t)nnt + k] = x[3 * i + 2];
    xs[3 * (size_t)nnt + k] = body[i];
}

__device__ __forceinline__ double dist2(const double *__restrict__ x, int a, int b) {
#pragma clang fp contract(off)
    double r2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double d = x[3 * (size_t)a + k] - x[3 * (size_t)b + k];
        r2 = r2 + d * d;
    }
    return r2;
}

// PHICOR and the per-body terms of energy.F:47-58 and :119-124 for body I = IFIRST + k
__global__ __launch_bounds__(EB) void phicor_kernel(nb6_energy_in e, const double *__restrict__ phi,
                                                    double *__restrict__ pterm,
                                                    double *__restrict__ kterm,
                                                    int *__restrict__ bad) {
#pragma clang fp contract(off)
    const int nnt = e.ntot - (2 * e.npairs + 1) + 1;
    const int k = blockIdx.x * EB + threadIdx.x;
    if (k >= nnt) return;
    const int I = 2 * e.npairs + 1 + k;               // 1-based
    const int i = I - 1;
    const double *x = e.x, *m = e.body;
    const size_t lm = e.lmax;
    const double thr = e.cmsep2_rmin2;
    double term;
    if (I <= e.n) {
        double d1 = 0.0;
        const int *row = e.list + (size_t)i * lm;
        const int nnb = row[0];
        if (nnb < 0 || nnb > (int)lm - 1) { *bad = 1; return; }
        // backwards over the list until the last c.m. neighbour (phicor.f:14-29)
        for (int l = nnb; l >= 1; l--) {
            const int J = row[l];
            if (J <= e.n) break;
            if (J > e.ntot) { *bad = 1; return; }
            const double rij2 = dist2(x, i, J - 1);
            if (rij2 < thr) {
                const int j1 = 2 * (J - e.n) - 1, j2 = j1 + 1;
                const double rj1 = dist2(x, i, j1 - 1), rj2 = dist2(x, i, j2 - 1);
                d1 = d1 - m[J - 1] / sqrt(rij2) + m[j1 - 1] / sqrt(rj1) + m[j2 - 1] / sqrt(rj2);
            }
        }
        const double p = phi[k] + d1;
        e.phidbl[i] = p;
        term = m[i] * p;
    } else {
        const int i1 = 2 * (I - e.n) - 1, i2 = i1 + 1;   // 1-based components
        const double pcm = phi[k];
        double d1 = 0.0, d2 = 0.0;
        const int *row = e.list + (size_t)(i1 - 1) * lm;
        const int np = row[0];
        if (np < 0 || np > (int)lm - 1) { *bad = 1; return; }
        for (int l = 1; l <= np; l++) {                 // phicor.f:36-71
            const int J = row[l];
            if (J < 1 || J > e.ntot) { *bad = 1; return; }
            const double rij2 = dist2(x, i, J - 1);
            if (J <= e.n || rij2 > thr) {
                const double ri1 = dist2(x, i1 - 1, J - 1), ri2 = dist2(x, i2 - 1, J - 1);
                d1 = d1 - m[J - 1] / sqrt(rij2) + m[J - 1] / sqrt(ri1);
                d2 = d2 - m[J - 1] / sqrt(rij2) + m[J - 1] / sqrt(ri2);
            } else {
                d1 = d1 - m[J - 1] / sqrt(rij2);
                d2 = d2 - m[J - 1] / sqrt(rij2);
                const int j1 = 2 * (J - e.n) - 1;
                for (int jj = j1; jj <= j1 + 1; jj++) {
                    const double rj1 = dist2(x, i1 - 1, jj - 1), rj2 = dist2(x, i2 - 1, jj - 1);
                    d1 = d1 + m[jj - 1] / sqrt(rj1);
                    d2 = d2 + m[jj - 1] / sqrt(rj2);
                }
            }
        }
        const double p1 = pcm + d1, p2 = pcm + d2;
        e.phidbl[i] = pcm;
        e.phidbl[i1 - 1] = p1;
        e.phidbl[i2 - 1] = p2;
        term = m[i1 - 1] * p1 + m[i2 - 1] * p2;
    }
    pterm[k] = term;
    const double v0 = e.xdot[3 * (size_t)i], v1 = e.xdot[3 * (size_t)i + 1],
                 v2 = e.xdot[3 * (size_t)i + 2];
    kterm[k] = m[i] * (v0 * v0 + v1 * v1 + v2 * v2);
}

// EBIN terms (energy.F:13-24): m1 m2 HT / m_cm for pairs with m_cm > 0
__global__ __launch_bounds__(EB) void ebin_kernel(nb6_energy_in e, double *__restrict__ bterm) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * EB + threadIdx.x;
    if (p >= e.npairs) return;
    const double mcm = e.body[e.n + p];
    bterm[p] = mcm > 0.0 ? e.body[2 * p] * e.body[2 * p + 1] * e.ht[p] / mcm : 0.0;
}

// fixed-tree FP64 sum of v[0..n) per work-group into out[blockIdx.x]
__global__ __launch_bounds__(EB) void sum_kernel(const double *__restrict__ v, int n,
                                                 double *__restrict__ out) {
    __shared__ double s[EB];
    double a = 0.0;
    for (int k = blockIdx.x * EB + threadIdx.x; k < n; k += gridDim.x * EB) a += v[k];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = EB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void clear_kernel(int *flag) { *flag = 0; }

DBuf g_pack, g_phi, g_terms, g_part;
HBuf g_hout;

}  // namespace

int energy_device(const nb6_energy_in *e, double *h_out, hipStream_t s) {
    if (!e || !h_out) { set_error(E_ARG, "nb6_energy_device: null argument"); return E_ARG; }
    const int ifirst = 2 * e->npairs + 1;
    const int nnt = e->ntot - ifirst + 1;
    if (e->npairs < 0 || e->n < ifirst - 1 || e->ntot < e->n || e->ntot > e->nmax ||
        e->ntot - e->n != e->npairs || nnt < 1 || e->lmax < 1) {
        set_error(E_ARG, "nb6_energy_device: need 0 <= 2 NPAIRS <= N <= NTOT = N + NPAIRS <= NMAX");
        return E_ARG;
    }
    if (!e->x || !e->xdot || !e->body || !e->list || !e->phidbl || (e->npairs && !e->ht)) {
        set_error(E_ARG, "nb6_energy_device: null device array");
        return E_ARG;
    }
    Device &d = device();
    if (!d.ready) return E_NODEV;
    if (!s) s = d.stream;
    double *xs = static_cast<double *>(g_pack.get((size_t)4 * nnt * sizeof(double)));
    double *phi = static_cast<double *>(g_phi.get((size_t)nnt * sizeof(double)));
    const int nb = std::max(nnt, e->npairs);
    double *terms = static_cast<double *>(g_terms.get((size_t)3 * nb * sizeof(double)));
    constexpr int NPART = 64;
    double *part = static_cast<double *>(g_part.get((size_t)(3 * NPART + 4) * sizeof(double)));
    int *bad = reinterpret_cast<int *>(part + 3 * NPART + 3);
    hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(1), 0, s, bad);
    KTimer &kt = ktimer("energy");
    kt.start(s);
    hipLaunchKernelGGL(energy_pack_kernel, dim3((nnt + EB - 1) / EB), dim3(EB), 0, s, e->x, e->body,
                       ifirst, nnt, xs);
    NB6_HIP(hipGetLastError());
    if (pot_launch(nnt, nnt, xs, 0, nnt, phi, s) != OK) return E_ARG;
    hipLaunchKernelGGL(phicor_kernel, dim3((nnt + EB - 1) / EB), dim3(EB), 0, s, *e, phi, terms,
                       terms + nb, bad);
    NB6_HIP(hipGetLastError());
    if (e->npairs > 0) {
        hipLaunchKernelGGL(ebin_kernel, dim3((e->npairs + EB - 1) / EB), dim3(EB), 0, s, *e,
                           terms + 2 * (size_t)nb);
        NB6_HIP(hipGetLastError());
    }
    const int cnt[3] = {nnt, nnt, e->npairs};
    for (int q = 0; q < 3; q++) {
        if (cnt[q] == 0) {
            hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(EB), 0, s, terms, 0, part + 3 * NPART + q);
        } else {
            hipLaunchKernelGGL(sum_kernel, dim3(NPART), dim3(EB), 0, s, terms + (size_t)q * nb, cnt[q],
                               part + (size_t)q * NPART);
            hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(EB), 0, s, part + (size_t)q * NPART, NPART,
                               part + 3 * NPART + q);
        }
        NB6_HIP(hipGetLastError());
    }
    kt.stop(s);
    double *ho = static_cast<double *>(g_hout.get(4 * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(ho, part + 3 * NPART, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    NB6_HIP(hipStreamSynchronize(s));
    if (*reinterpret_cast<int *>(ho + 3)) {
        set_error(E_ARG, "nb6_energy_device: a LIST row holds a count or index out of range");
        return E_ARG;
    }
    h_out[0] = 0.5 * ho[1];    // ZKIN  (sum over the kinetic terms, energy.F:119-124)
    h_out[1] = 0.5 * ho[0];    // POT   (energy.F:61)
    h_out[2] = ho[2];          // EBIN
    return OK;
}

}  // namespace nb6
