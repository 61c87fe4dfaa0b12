// Device-resident j-particles with a GPU predictor.
//
// The reference predicts every particle on the host at each regular block
// (xbpredall.F:52-60) and ships all of them to the GPU (gpunb_send_,
// intgrt.F:875-896).  Here the predictor state lives in HBM: the caller uploads
// only the particles it has just corrected (nb6_jp_set_), the prediction to the
// block time runs on the GPU and writes the j records the regular-force kernel
// reads, and a regular block can take its i-particles' predicted positions
// straight from the same arrays (nb6_jp_regf_).
//
// State per particle (SoA, FP64): x0[3], x0dot[3], F[3] = a/2, FDOT[3] = adot/6,
// t0, m.  Prediction (xbpredall.F:55-59):
//   s = t - t0;  x = ((FDOT s + F) s + x0dot) s + x0;  xdot = (1.5 FDOT s + F) 2s + x0dot
#include "nb6_internal.h"

namespace nb6 {

namespace {

constexpr int NST = 14;   // state rows

__global__ __launch_bounds__(256) void jp_scatter_kernel(int n, const int *__restrict__ addr,
                                                         const double *__restrict__ upd,
                                                         double *__restrict__ st, int nmax) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int j = addr[k];
    if ((unsigned)j >= (unsigned)nmax) return;
    const double *u = upd + (size_t)k * NST;
#pragma unroll
    for (int q = 0; q < NST; q++) st[(size_t)q * nmax + j] = u[q];
}

__global__ __launch_bounds__(256) void jp_predict_kernel(int nj, const double *__restrict__ st,
                                                         int nmax, double time,
                                                         double *__restrict__ aos,
                                                         double *__restrict__ pred) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= nj) return;
    const size_t M = nmax;
    const double s = time - st[12 * M + j];
    const double s1 = 1.5 * s, s2 = 2.0 * s;
    double x[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double x0 = st[d * M + j], v0 = st[(3 + d) * M + j];
        const double f = st[(6 + d) * M + j], fd = st[(9 + d) * M + j];
        x[d] = ((fd * s + f) * s + v0) * s + x0;
        v[d] = (fd * s1 + f) * s2 + v0;
        pred[d * M + j] = x[d];
        pred[(3 + d) * M + j] = v[d];
    }
    double *o = aos + (size_t)j * 8;
    o[0] = x[0]; o[1] = x[1]; o[2] = x[2]; o[3] = st[13 * M + j];
    o[4] = v[0]; o[5] = v[1]; o[6] = v[2]; o[7] = 0.0;
}

// i-block of ni particles given by 0-based address: predicted x, v and the caller's h2, dtr
__global__ __launch_bounds__(256) void jp_gather_kernel(int ni, const int *__restrict__ addr,
                                                        const double *__restrict__ pred, int nmax,
                                                        const double *__restrict__ h2dtr,
                                                        double *__restrict__ ib) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ni) return;
    const int j = addr[i];
    const bool ok = (unsigned)j < (unsigned)nmax;
    const size_t M = nmax, N = ni;
#pragma unroll
    for (int q = 0; q < 6; q++) ib[q * N + i] = ok ? pred[q * M + j] : 0.0;
    ib[6 * N + i] = h2dtr[i];
    ib[7 * N + i] = h2dtr[N + i];
}

struct JPred {
    bool open = false;
    int nmax = 0;
    DBuf st, pred, upd, addr, hd;
    HBuf hupd, haddr, hh;
    int npred = 0;            // particles covered by the last prediction
} g_jp;

}  // namespace

int jp_open(int nmax) {
    JPred &p = g_jp;
    if (p.open) {
        fprintf(stderr, "nb6 jp: already open\n");
        return OK;
    }
    if (nmax <= 0) { set_error(E_ARG, "nb6_jp_open: nmax must be positive"); return E_ARG; }
    Device &d = device();
    if (!d.ready) return E_NODEV;
    p.nmax = nmax;
    p.st.get((size_t)NST * nmax * sizeof(double));
    p.pred.get((size_t)6 * nmax * sizeof(double));
    NB6_HIP(hipMemsetAsync(p.st.p, 0, (size_t)NST * nmax * sizeof(double), d.stream));
    NB6_HIP(hipStreamSynchronize(d.stream));
    p.npred = 0;
    p.open = true;
    return OK;
}

int jp_close() {
    JPred &p = g_jp;
    if (!p.open) return OK;
    p.st.release(); p.pred.release(); p.upd.release(); p.addr.release(); p.hd.release();
    p.hupd.release(); p.haddr.release(); p.hh.release();
    p.open = false;
    return OK;
}

// addresses are 1-based (Fortran particle index)
int jp_set(int n, const int *addr, const double (*x0)[3], const double (*v0)[3],
           const double (*f)[3], const double (*fdot)[3], const double *m, const double *t0) {
    JPred &p = g_jp;
    if (!p.open) { set_error(E_STATE, "nb6_jp_set: not open"); return E_STATE; }
    if (n <= 0) return OK;
    for (int k = 0; k < n; k++)
        if (addr[k] < 1 || addr[k] > p.nmax) {
            set_error(E_ARG, "nb6_jp_set: address outside [1, nmax]");
            return E_ARG;
        }
    Device &d = device();
    double *h = static_cast<double *>(p.hupd.get((size_t)n * NST * sizeof(double)));
    int *ha = static_cast<int *>(p.haddr.get((size_t)n * sizeof(int)));
    for (int k = 0; k < n; k++) {
        double *u = h + (size_t)k * NST;
        for (int q = 0; q < 3; q++) {
            u[q] = x0[k][q];
            u[3 + q] = v0[k][q];
            u[6 + q] = f[k][q];
            u[9 + q] = fdot[k][q];
        }
        u[12] = t0[k];
        u[13] = m[k];
        ha[k] = addr[k] - 1;
    }
    double *du = static_cast<double *>(p.upd.get((size_t)n * NST * sizeof(double)));
    int *da = static_cast<int *>(p.addr.get((size_t)n * sizeof(int)));
    NB6_HIP(hipMemcpyAsync(du, h, (size_t)n * NST * sizeof(double), hipMemcpyHostToDevice, d.stream));
    NB6_HIP(hipMemcpyAsync(da, ha, (size_t)n * sizeof(int), hipMemcpyHostToDevice, d.stream));
    hipLaunchKernelGGL(jp_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, d.stream, n, da, du,
                       p.st.as<double>(), p.nmax);
    NB6_HIP(hipGetLastError());
    NB6_HIP(hipStreamSynchronize(d.stream));   // the pinned staging is reused by the next call
    return OK;
}

// predict particles [0, nj) to `time` into the AoS records aos[nj][8] (a j-set's in64)
int jp_predict(double time, int nj, double *aos, hipStream_t s) {
    JPred &p = g_jp;
    if (!p.open) { set_error(E_STATE, "nb6_jp_predict_send: not open"); return E_STATE; }
    if (nj < 0 || nj > p.nmax) { set_error(E_ARG, "nb6_jp_predict_send: nj outside [0, nmax]"); return E_ARG; }
    if (nj > 0) {
        hipLaunchKernelGGL(jp_predict_kernel, dim3((nj + 255) / 256), dim3(256), 0, s, nj,
                           p.st.as<double>(), p.nmax, time, aos, p.pred.as<double>());
        NB6_HIP(hipGetLastError());
    }
    p.npred = nj;
    return OK;
}

const double *jp_state(int *nmax) {
    if (!g_jp.open) return nullptr;
    *nmax = g_jp.nmax;
    return g_jp.st.as<double>();
}

// i-block SoA [8][ni] on the device from 1-based addresses and host h2, dtr
int jp_iblock(int ni, const int *addr, const double *h2, const double *dtr, double *d_ib,
              hipStream_t s) {
    JPred &p = g_jp;
    if (!p.open) { set_error(E_STATE, "nb6_jp_regf: not open"); return E_STATE; }
    for (int k = 0; k < ni; k++)
        if (addr[k] < 1 || addr[k] > p.npred) {
            set_error(E_ARG, "nb6_jp_regf: address outside the predicted range [1, nj]");
            return E_ARG;
        }
    int *ha = static_cast<int *>(p.haddr.get((size_t)ni * sizeof(int)));
    double *hh = static_cast<double *>(p.hh.get((size_t)2 * ni * sizeof(double)));
    for (int k = 0; k < ni; k++) {
        ha[k] = addr[k] - 1;
        hh[k] = h2[k];
        hh[ni + k] = dtr[k];
    }
    int *da = static_cast<int *>(p.addr.get((size_t)ni * sizeof(int)));
    double *dh = static_cast<double *>(p.hd.get((size_t)2 * ni * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(da, ha, (size_t)ni * sizeof(int), hipMemcpyHostToDevice, s));
    NB6_HIP(hipMemcpyAsync(dh, hh, (size_t)2 * ni * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(jp_gather_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, ni, da,
                       p.pred.as<double>(), p.nmax, dh, d_ib);
    NB6_HIP(hipGetLastError());
    NB6_HIP(hipStreamSynchronize(s));           // the pinned staging is reused by the next call
    return OK;
}

}  // namespace nb6
