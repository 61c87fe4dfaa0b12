// All-pairs potential for gfx950.
//
// Semantics follow gpupot (gpupot.gpu.cu:32-57, pot.avx.cpp:112-136):
//   phi_i = sum_{j, r_ij^2 > 0} m_j / r_ij   (positive; the self pair and any
//   coincident particle are skipped).
//
// Method (own): grid (i-blocks, j-chunks), 256 threads, 4 i-particles per lane.
// j-tiles of 256 particles are staged in LDS as FP64 positions + FP32 mass.  The
// separation is formed in FP64 and rounded once to FP32 (v_add_f64 + v_cvt), so
// the relative error of dx is one FP32 rounding even for the tightest binaries
// (the reference reaches the same with float-float positions, cuda_share.h:45-54).
// r^2, v_rsq_f32 and m*rinv run in FP32 and are summed per tile with Kahan
// compensation (a tight binary can dominate a tile); tile sums are flushed into
// FP64, chunk partials are summed in FP64 by a second kernel.
#include "nb6_internal.h"

#include <algorithm>

namespace nb6 {

namespace {

constexpr int PT_THREADS = 256;
constexpr int PT_TJ = 256;
constexpr int PT_IPL = 4;

struct PotParams {
    const double *__restrict__ xs;   // SoA [4][stride] x, y, z, m
    int n, stride;
    int i0, ni;
    int chunk_len;
    double *__restrict__ part;       // [nchunk][ni]
};

__global__ __launch_bounds__(PT_THREADS) void pot_tile_kernel(PotParams p) {
    __shared__ double4 sJ[PT_TJ];
    const int tid = threadIdx.x;
    const int chunk = blockIdx.y;
    const int kbase = blockIdx.x * (PT_THREADS * PT_IPL);
    const int st = p.stride;

    double xi[PT_IPL], yi[PT_IPL], zi[PT_IPL];
    float a32[PT_IPL], c32[PT_IPL];   // Kahan sum and compensation per tile
    double a64[PT_IPL];
#pragma unroll
    for (int k = 0; k < PT_IPL; k++) {
        const int kk = kbase + k * PT_THREADS + tid;
        const int i = p.i0 + (kk < p.ni ? kk : 0);
        xi[k] = p.xs[0 * st + i]; yi[k] = p.xs[1 * st + i]; zi[k] = p.xs[2 * st + i];
        a32[k] = 0.f;
        c32[k] = 0.f;
        a64[k] = 0.0;
    }
    const int j0 = chunk * p.chunk_len;
    const int j1 = min(j0 + p.chunk_len, p.n);
    for (int jt = j0; jt < j1; jt += PT_TJ) {
        const int j = jt + tid;
        if (j < j1)
            sJ[tid] = make_double4(p.xs[0 * st + j], p.xs[1 * st + j], p.xs[2 * st + j],
                                   p.xs[3 * st + j]);
        else
            sJ[tid] = make_double4(1.0e17, 0.0, 0.0, 0.0);
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < PT_TJ; jj++) {
            const double4 b = sJ[jj];
            const float mj = (float)b.w;
#pragma unroll
            for (int k = 0; k < PT_IPL; k++) {
                const float dx = (float)(b.x - xi[k]);
                const float dy = (float)(b.y - yi[k]);
                const float dz = (float)(b.z - zi[k]);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float ri = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
                // compensated (Kahan) accumulation: one close pair (a tight binary)
                // can dominate the tile sum
                const float y = fmaf(mj, ri, -c32[k]);
                const float t = a32[k] + y;
                c32[k] = (t - a32[k]) - y;
                a32[k] = t;
            }
        }
#pragma unroll
        for (int k = 0; k < PT_IPL; k++) {
            a64[k] += (double)a32[k] - (double)c32[k];
            a32[k] = 0.f;
            c32[k] = 0.f;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < PT_IPL; k++) {
        const int kk = kbase + k * PT_THREADS + tid;
        if (kk < p.ni) p.part[(size_t)chunk * p.ni + kk] = a64[k];
    }
}

__global__ __launch_bounds__(256) void pot_sum_kernel(const double *__restrict__ part,
                                                      int ni, int nchunk,
                                                      double *__restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= ni) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(size_t)c * ni + k];
    out[k] = s;
}

DBuf g_part;

}  // namespace

int pot_launch(int n, int n_stride, const double *d_x, int i0, int ni, double *d_pot,
               hipStream_t s) {
    if (ni <= 0) return OK;
    if (n <= 0 || i0 < 0 || i0 + ni > n) {
        set_error(E_ARG, "gpupot: particle range outside [0, n)");
        return E_ARG;
    }
    const int nbi = PT_THREADS * PT_IPL;
    const int nib = (ni + nbi - 1) / nbi;
    const int want = (2048 + nib - 1) / nib;
    const int maxc = std::max(1, n / (4 * PT_TJ));
    int nchunk = std::max(1, std::min(want, maxc));
    int clen = (n + nchunk - 1) / nchunk;
    clen = (clen + PT_TJ - 1) / PT_TJ * PT_TJ;
    nchunk = (n + clen - 1) / clen;
    double *part = static_cast<double *>(g_part.get((size_t)nchunk * ni * sizeof(double)));
    PotParams p{d_x, n, n_stride, i0, ni, clen, part};
    KTimer &t = ktimer("pot");
    t.start(s);
    hipLaunchKernelGGL(pot_tile_kernel, dim3(nib, nchunk), dim3(PT_THREADS), 0, s, p);
    NB6_HIP(hipGetLastError());
    t.stop(s);
    hipLaunchKernelGGL(pot_sum_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, part, ni,
                       nchunk, d_pot);
    NB6_HIP(hipGetLastError());
    return OK;
}

void pot_release() { g_part.release(); }

}  // namespace nb6
