// Learned neighbour radius for gfx950: a small network predicts each particle's
// next neighbour radius RS from eight per-particle features, so that the count of
// the next gpunb_regf call lands near NNBOPT; the reference's RS rule stays as the
// fallback and the exact count of the kernel judges every prediction.
//
// Pieces (all device-resident):
//  * rs_rule_kernel      the reference rule after a regular step (regcor_gpu.F:95-238,
//                        single stars, NB_FLAG = 1, M_FLAG = 0) in FP64.
//  * rs_kth_kernel       training target: the radius that yields exactly k neighbours
//                        under min(r^2, |dx + dtr dv|^2), from a regf list built with a
//                        generous radius (one wavefront per i, bitonic sort in LDS).
//  * rs_feat_kernel      the eight features (log density from the previous list,
//                        previous count, |a|, |jerk|, distance from the density centre,
//                        previous RS, regular step, mass).
//  * rs_predict_kernel   inference fused ahead of gpunb_regf: normalise, 8->64->64->1
//                        (ReLU, bfloat16 weights, FP32 accumulation), RS = RS_prev e^y,
//                        rule value where the fallback flag is set, h2 written into the
//                        regf i-block.
//  * rsnet_grad_kernel / rsnet_sum_kernel / rsnet_adam_kernel   training: per-block
//                        forward + backward of 64 samples with activations in LDS,
//                        a fixed-order sum of the block partials, Adam on FP32 master
//                        weights with a bfloat16 copy for the forward pass.
#include "nb6_internal.h"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace nb6 {

namespace {

constexpr int NF = 8, NH = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NF;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH;
constexpr int NPARAM = OFF_B3 + 1;   // 4801
constexpr int GB = 64;               // samples per training work-group
constexpr int KTH_MAX = 2048;        // largest list the target kernel sorts

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float((uint32_t)h << 16);
}

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    u = u + 0x7FFFu + ((u >> 16) & 1u);    // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

struct NetState {
    bool open = false;
    DBuf master, m1, m2, wbf, grad, part, loss, norm;   // norm: mu[8], 1/sigma[8]
    DBuf lacc, keys, skeys, vals, perm, sort_tmp;          // epoch: loss sum, shuffle
    int step = 0;
} g_net;

// ------------------------------------------------------------------ weights in LDS
struct SW {
    float w1[NH * NF], b1[NH], w2[NH * (NH + 1)], b2[NH], w3[NH], b3;
};

__device__ __forceinline__ void load_weights(SW &s, const uint16_t *__restrict__ w) {
    for (int k = threadIdx.x; k < NPARAM; k += blockDim.x) {
        const float v = bf16_to_f32(w[k]);
        if (k < OFF_B1) s.w1[k] = v;
        else if (k < OFF_W2) s.b1[k - OFF_B1] = v;
        else if (k < OFF_B2) { const int q = k - OFF_W2; s.w2[(q / NH) * (NH + 1) + q % NH] = v; }
        else if (k < OFF_W3) s.b2[k - OFF_B2] = v;
        else if (k < OFF_B3) s.w3[k - OFF_W3] = v;
        else s.b3 = v;
    }
}

__device__ __forceinline__ void load_features(const float *__restrict__ feat, int n, int s,
                                              const float *__restrict__ norm, float *x) {
#pragma unroll
    for (int f = 0; f < NF; f++) x[f] = (feat[(size_t)f * n + s] - norm[f]) * norm[NF + f];
}

// forward pass of one sample; h1 kept for the caller
__device__ __forceinline__ float net_forward(const SW &w, const float *x, float *h1) {
#pragma unroll
    for (int k = 0; k < NH; k++) {
        float z = w.b1[k];
#pragma unroll
        for (int f = 0; f < NF; f++) z = fmaf(w.w1[k * NF + f], x[f], z);
        h1[k] = fmaxf(z, 0.f);
    }
    float y = w.b3;
    for (int o = 0; o < NH; o++) {
        float z = w.b2[o];
#pragma unroll
        for (int k = 0; k < NH; k++) z = fmaf(w.w2[o * (NH + 1) + k], h1[k], z);
        y = fmaf(w.w3[o], fmaxf(z, 0.f), y);
    }
    return y;
}

// ------------------------------------------------------------------ training
struct GradLds {
    SW w;
    float x[GB][NF];
    float h1[GB][NH + 1];
    float d2[GB][NH + 1];
    float dy[GB];
};

// 64 x 64 x 64 products of the hidden layer on the matrix cores, one wavefront:
// C[m][n] = sum_k A(m, k) B(k, n), A and B read through the callables (LDS), each
// element of C handed to out(m, n, c).  v_mfma_f32_16x16x4f32 (FP32 in, FP32
// accumulate): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] of a 16 x 4 and
// a 4 x 16 block, and holds C[4 (l / 16) + i][l % 16], i = 0..3, of the 16 x 16 block.
typedef float v4f __attribute__((ext_vector_type(4)));
template <class FA, class FB, class FO>
__device__ __forceinline__ void gemm64(FA a, FB b, FO out, int lane) {
    const int r = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < 4; mt++) {
        v4f acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[nt] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < NH; kb += 4) {
            const float av = a(mt * 16 + r, kb + q);
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b(kb + q, nt * 16 + r), acc[nt],
                                                               0, 0, 0);
        }
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) out(mt * 16 + q * 4 + i, nt * 16 + r, acc[nt][i]);
        }
    }
}

// Minibatch of n samples: sample s is row idx[base + s] (idx null: base + s) of the
// SoA features float[8][fstride] and of the targets.  One wavefront per 64 samples:
// the 8 -> 64 layer and the gradients of W1, b1, W3, b3 on the VALU, the three 64 x 64
// products of the hidden layer (forward, dW2, back-propagated delta) on the matrix
// cores.
__global__ __launch_bounds__(GB) void rsnet_grad_kernel(int n, const float *__restrict__ feat,
                                                        const float *__restrict__ tgt,
                                                        const int *__restrict__ idx, int base,
                                                        int fstride,
                                                        const uint16_t *__restrict__ wbf,
                                                        const float *__restrict__ norm,
                                                        float *__restrict__ part,
                                                        float *__restrict__ lpart) {
    __shared__ GradLds L;
    load_weights(L.w, wbf);
    const int t = threadIdx.x;
    const int s = blockIdx.x * GB + t;
    const bool ok = s < n;
    const int g = ok ? (idx ? idx[base + s] : base + s) : 0;
    float x[NF];
    if (ok) load_features(feat, fstride, g, norm, x);
    else
        for (int f = 0; f < NF; f++) x[f] = 0.f;
    __syncthreads();
    // layer 1 (VALU): h1 of this lane's sample into LDS
#pragma unroll
    for (int k = 0; k < NH; k++) {
        float z = L.w.b1[k];
#pragma unroll
        for (int f = 0; f < NF; f++) z = fmaf(L.w.w1[k * NF + f], x[f], z);
        L.h1[t][k] = fmaxf(z, 0.f);
    }
#pragma unroll
    for (int f = 0; f < NF; f++) L.x[t][f] = x[f];
    __syncthreads();
    // layer 2 (matrix cores): h2[s][o] = relu(b2[o] + sum_k h1[s][k] W2[o][k]) into d2
    gemm64([&](int m, int k) { return L.h1[m][k]; },
           [&](int k, int o) { return L.w.w2[o * (NH + 1) + k]; },
           [&](int m, int o, float z) { L.d2[m][o] = fmaxf(z + L.w.b2[o], 0.f); }, t);
    __syncthreads();
    float y = L.w.b3;
    for (int o = 0; o < NH; o++) y = fmaf(L.w.w3[o], L.d2[t][o], y);
    const float r = ok ? y - tgt[g] : 0.f;
    const float dy = 2.0f * r / (float)n;
    L.dy[t] = dy;
    __syncthreads();

    float *P = part + (size_t)blockIdx.x * NPARAM;
    // w3, b3 from h2 (still in d2)
    {
        float gw = 0.f;
        for (int q = 0; q < GB; q++) gw = fmaf(L.dy[q], L.d2[q][t], gw);
        P[OFF_W3 + t] = gw;
        if (t == 0) {
            float gb = 0.f;
            for (int q = 0; q < GB; q++) gb += L.dy[q];
            P[OFF_B3] = gb;
        }
    }
    __syncthreads();
    // delta2 = dy w3 [h2 > 0], in place of h2
    for (int o = 0; o < NH; o++) {
        const float h2 = L.d2[t][o];
        L.d2[t][o] = h2 > 0.f ? dy * L.w.w3[o] : 0.f;
    }
    __syncthreads();
    // dW2[o][k] = sum_s delta2[s][o] h1[s][k] (matrix cores); b2 = sum_s delta2[s][o]
    gemm64([&](int o, int q) { return L.d2[q][o]; },
           [&](int q, int k) { return L.h1[q][k]; },
           [&](int o, int k, float v) { P[OFF_W2 + o * NH + k] = v; }, t);
    {
        float gb = 0.f;
        for (int q = 0; q < GB; q++) gb += L.d2[q][t];
        P[OFF_B2 + t] = gb;
    }
    __syncthreads();
    // delta1[s][k] = (sum_o delta2[s][o] W2[o][k]) [h1 > 0] (matrix cores), in place of h1:
    // each element is read (for the mask) and written by the same lane only
    gemm64([&](int m, int o) { return L.d2[m][o]; },
           [&](int o, int k) { return L.w.w2[o * (NH + 1) + k]; },
           [&](int m, int k, float v) { L.h1[m][k] = L.h1[m][k] > 0.f ? v : 0.f; }, t);
    __syncthreads();
    // W1 (row k = t), b1
    {
        const int k = t;
        float acc[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) acc[f] = 0.f;
        float gb = 0.f;
        for (int q = 0; q < GB; q++) {
            const float d = L.h1[q][k];
            gb += d;
#pragma unroll
            for (int f = 0; f < NF; f++) acc[f] = fmaf(d, L.x[q][f], acc[f]);
        }
#pragma unroll
        for (int f = 0; f < NF; f++) P[OFF_W1 + k * NF + f] = acc[f];
        P[OFF_B1 + k] = gb;
    }
    // loss partial: sum of r^2 over the block
    {
        float v = r * r;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (t == 0) lpart[blockIdx.x] = v;
    }
}

// fixed-order sum over the block partials (deterministic)
__global__ __launch_bounds__(256) void rsnet_sum_kernel(const float *__restrict__ part, int nblk,
                                                        float *__restrict__ grad,
                                                        const float *__restrict__ lpart,
                                                        float *__restrict__ loss, int n,
                                                        double *__restrict__ lacc) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < NPARAM) {
        float g = 0.f;
        for (int b = 0; b < nblk; b++) g += part[(size_t)b * NPARAM + p];
        grad[p] = g;
    }
    if (p == NPARAM) {
        double l = 0.0;
        for (int b = 0; b < nblk; b++) l += lpart[b];
        *loss = (float)(l / n);
        if (lacc) *lacc += l;   // sum of squared residuals over an epoch
    }
}

// epoch shuffle: a random 32-bit key per sample (SplitMix64 of seed and index), sorted
__global__ __launch_bounds__(256) void rsnet_keys_kernel(int n, unsigned long long seed,
                                                         unsigned *__restrict__ key,
                                                         int *__restrict__ val) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(k + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    key[k] = (unsigned)(z >> 32);
    val[k] = k;
}

__global__ void rsnet_zero_kernel(double *p) { *p = 0.0; }

__global__ __launch_bounds__(256) void rsnet_adam_kernel(float *__restrict__ w, float *__restrict__ m1,
                                                         float *__restrict__ m2,
                                                         const float *__restrict__ g,
                                                         uint16_t *__restrict__ wbf, float lr,
                                                         float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gp = g[p];
    const float a = fmaf(b1, m1[p], (1.f - b1) * gp);
    const float v = fmaf(b2, m2[p], (1.f - b2) * gp * gp);
    m1[p] = a;
    m2[p] = v;
    const float wp = w[p] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[p] = wp;
    wbf[p] = f32_to_bf16(wp);
}

__global__ __launch_bounds__(256) void rsnet_round_kernel(const float *__restrict__ w,
                                                          uint16_t *__restrict__ wbf) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < NPARAM) wbf[p] = f32_to_bf16(w[p]);
}

// ------------------------------------------------------------------ inference
__global__ __launch_bounds__(256) void rsnet_forward_kernel(int n, const float *__restrict__ feat,
                                                            const uint16_t *__restrict__ wbf,
                                                            const float *__restrict__ norm,
                                                            float *__restrict__ y) {
    __shared__ SW w;
    load_weights(w, wbf);
    __syncthreads();
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float x[NF], h1[NH];
    load_features(feat, n, s, norm, x);
    y[s] = net_forward(w, x, h1);
}

// RS = RS_prev e^y (or the rule value where use_rule != 0); h2 = RS^2 (/ bodym if
// m_flag) written into row 6 of the regf i-block.
__global__ __launch_bounds__(256) void rs_predict_kernel(int n, const float *__restrict__ feat,
                                                         const uint16_t *__restrict__ wbf,
                                                         const float *__restrict__ norm,
                                                         const double *__restrict__ rs_prev,
                                                         const double *__restrict__ rs_rule,
                                                         const int *__restrict__ use_rule,
                                                         double *__restrict__ rs_out,
                                                         double *__restrict__ ib, int ib_stride,
                                                         int m_flag, double bodym) {
    __shared__ SW w;
    load_weights(w, wbf);
    __syncthreads();
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    double rs;
    if (use_rule && use_rule[s]) {
        rs = rs_rule[s];
    } else {
        float x[NF], h1[NH];
        load_features(feat, n, s, norm, x);
        const float y = net_forward(w, x, h1);
        rs = rs_prev[s] * exp((double)fminf(fmaxf(y, -4.f), 4.f));
    }
    rs_out[s] = rs;
    if (ib) ib[6 * (size_t)ib_stride + s] = m_flag ? rs * rs / bodym : rs * rs;
}

__global__ __launch_bounds__(256) void rs_feat_kernel(int n, int stride, const double *__restrict__ x,
                                                      const double *__restrict__ acc,
                                                      const double *__restrict__ jrk,
                                                      const double *__restrict__ mass,
                                                      const double *__restrict__ rs,
                                                      const double *__restrict__ dtr,
                                                      const int *__restrict__ nnb, double cx,
                                                      double cy, double cz,
                                                      float *__restrict__ feat) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const size_t st = stride;
    const double r = rs[s];
    const double c = (double)max(nnb[s], 0);
    const double a = sqrt(acc[s] * acc[s] + acc[st + s] * acc[st + s] + acc[2 * st + s] * acc[2 * st + s]);
    const double j = sqrt(jrk[s] * jrk[s] + jrk[st + s] * jrk[st + s] + jrk[2 * st + s] * jrk[2 * st + s]);
    const double dx = x[s] - cx, dy = x[st + s] - cy, dz = x[2 * st + s] - cz;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    float *f = feat + s;
    const size_t fn = n;
    f[0 * fn] = (float)log10((c + 1.0) / (4.18879020478639 * r * r * r));
    f[1 * fn] = (float)log10(c + 1.0);
    f[2 * fn] = (float)log10(a + 1e-30);
    f[3 * fn] = (float)log10(j + 1e-30);
    f[4 * fn] = (float)log10(d + 1e-6);
    f[5 * fn] = (float)log10(r);
    f[6 * fn] = (float)log10(dtr[s]);
    f[7 * fn] = (float)log10(mass[s]);
}

// ------------------------------------------------------------------ reference rule
struct RuleParams {
    int n, stride;
    const double *x, *v, *dtr;       // SoA [3][stride], [stride]
    const int *nnb, *nnb0;           // new count (without self), previous count
    const int *list;                 // [n][lmax] 0-based j (without self), may be null
    int lmax;
    const double *xj, *vj;           // SoA [3][jstride] of every particle (for NNB <= 3)
    int jstride;
    double rdx, rdy, rdz, rscale, rc2, rsfac, znbmin;
    int nnbopt, nnbmax;
    double *rs;                      // in / out
};

__global__ __launch_bounds__(256) void rs_rule_kernel(RuleParams q) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= q.n) return;
    const size_t st = q.stride;
    const double xi[3] = {q.x[i], q.x[st + i], q.x[2 * st + i]};
    const double vi[3] = {q.v[i], q.v[st + i], q.v[2 * st + i]};
    const double dri[3] = {xi[0] - q.rdx, xi[1] - q.rdy, xi[2] - q.rdz};
    const double ri2 = dri[0] * dri[0] + dri[1] * dri[1] + dri[2] * dri[2];
    const double rh2 = q.rscale * q.rscale;
    const int nnb = q.nnb[i], nnb0 = q.nnb0[i];
    const double dtr = q.dtr[i];
    double rs = q.rs[i];
    if (nnb == 0) {   // regcor_gpu.F:95-127
        const double r2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
        const double rdot = 3.0 * (xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2]) / r2;
        q.rs[i] = rdot > 0.0 ? fmax(0.75 * rs, 0.1 * q.rscale) : 1.1 * rs;
        return;
    }
    const int nbave = min(q.nnbopt + 150, max(q.nnbmax - 50, (int)q.znbmin));
    const double fn = (double)nnb;
    const double fac = 1.0 + 0.1 * ((double)q.nnbopt - fn) / fn;
    double a3 = fn * fac;
    a3 = fmin(a3, 0.9 * nbave);
    const int nbp = (int)a3;
    if (ri2 > rh2) a3 = a3 * fmax(q.rscale / sqrt(ri2), 0.1);
    else a3 = fmax(a3, q.znbmin);
    double a4 = a3 / fn;
    if ((a3 - (double)nnb0) * (a3 - fn) < 0.0) a4 = sqrt(a4);
    const double ridot = dri[0] * vi[0] + dri[1] * vi[1] + dri[2] * vi[2];
    if (ri2 > q.rc2 && ri2 < 100.0 * rh2) a4 = a4 * (1.0 + ridot * dtr / ri2);
    const double a5 = fmin(q.rsfac * dtr, 0.25);
    a4 = a4 - 1.0;
    a4 = a4 > a5 ? a5 : fmax(a4, -a5);
    a3 = a4 / 3.0;
    double a1 = 1.0 + a3 - a3 * a3;
    if (fn < q.znbmin && (double)nbp > q.znbmin && a1 < 1.0) a1 = 1.05;
    if (fabs(a1 - 1.0) > 0.003) {
        if (rs * rs > ri2 && rs * rs < 2.0 * ri2 && nnb > 5) a1 = sqrt(a1);
        rs = a1 * rs;
    }
    if (nnb > nbave && ridot < 0.0) rs = rs * (1.0 - (fn - nbave) / (3.0 * fn));
    if (nnb <= 3 && ri2 < 100.0 * rh2 && q.list) {   // regcor_gpu.F:222-238
        // rows as gpunb_regf returns them: the particle itself (zero distance) is
        // among the entries and is skipped
        double b = 2.0 * rs;
        const int *row = q.list + (size_t)i * q.lmax;
        const int c = min(row[0], q.lmax - 1);
        const size_t js = q.jstride;
        for (int l = 0; l < c; l++) {
            const int j = row[1 + l];
            if ((unsigned)j >= (unsigned)q.jstride) continue;
            const double d[3] = {xi[0] - q.xj[j], xi[1] - q.xj[js + j], xi[2] - q.xj[2 * js + j]};
            const double w[3] = {vi[0] - q.vj[j], vi[1] - q.vj[js + j], vi[2] - q.vj[2 * js + j]};
            const double rij = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (rij == 0.0) continue;
            const double rsdot = (d[0] * w[0] + d[1] * w[1] + d[2] * w[2]) / rij;
            b = fmin(b, rij + rsdot * dtr);
        }
        rs = fmax(b, 1.1 * rs);
    }
    q.rs[i] = rs;
}

// ------------------------------------------------------------------ target radius
// One wavefront per i: squared criterion distances of the listed j (caller
// order, FP64, the oracle's formula order) sorted in LDS; rs_out = sqrt of the
// k-th smallest, or -count when the list holds fewer than k entries or overflowed.
__global__ __launch_bounds__(256) void rs_kth_kernel(int ni, int ib_stride, const double *__restrict__ ib,
                                                     const int *__restrict__ list, int lmax, int k,
                                                     const double *__restrict__ jaos, int nj,
                                                     double *__restrict__ rs_out) {
    __shared__ double buf[4][KTH_MAX];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wid;
    if (i >= ni) return;
    double *b = buf[wid];
    const int *row = list + (size_t)i * lmax;
    const int c = row[0];
    if (c < k || c > KTH_MAX || c > lmax - 1) {
        if (lane == 0) rs_out[i] = -(double)(c < 0 ? -c : c) - 1.0;
        return;
    }
    const size_t is = ib_stride;
    const double xi = ib[i], yi = ib[is + i], zi = ib[2 * is + i];
    const double ui = ib[3 * is + i], vi = ib[4 * is + i], wi = ib[5 * is + i];
    const double dt = ib[7 * is + i];
    int p = 1;
    while (p < c) p <<= 1;
    for (int l = lane; l < p; l += 64) {
        double d2 = INFINITY;
        if (l < c) {
            const int j = row[1 + l];
            if ((unsigned)j < (unsigned)nj) {
#pragma clang fp contract(off)
                const double *r = jaos + (size_t)j * 8;   // x y z m vx vy vz 0
                const double dx = r[0] - xi, dy = r[1] - yi, dz = r[2] - zi;
                const double du = r[4] - ui, dv = r[5] - vi, dw = r[6] - wi;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double ex = dx + dt * du, ey = dy + dt * dv, ez = dz + dt * dw;
                const double r2p = ex * ex + ey * ey + ez * ez;
                d2 = r2 < r2p ? r2 : r2p;
            }
        }
        b[l] = d2;
    }
    // bitonic sort of p values (ascending); all lanes of the wavefront take part
    for (int kk = 2; kk <= p; kk <<= 1) {
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int t = lane; t < p / 2; t += 64) {
                const int lo = (t / jj) * 2 * jj + (t % jj);
                const int hi = lo + jj;
                const bool up = (lo & kk) == 0;
                const double a = b[lo], e = b[hi];
                if ((a > e) == up) { b[lo] = e; b[hi] = a; }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0) rs_out[i] = sqrt(b[k - 1]);
}

// He-uniform initial weights from a SplitMix64 stream; zero biases
void init_host(std::vector<float> &p, unsigned long long seed) {
    unsigned long long z = seed;
    auto next = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        unsigned long long r = z;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        return (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    p.assign(NPARAM, 0.f);
    for (int k = 0; k < NH * NF; k++) p[OFF_W1 + k] = (float)(next() * std::sqrt(6.0 / NF));
    for (int k = 0; k < NH * NH; k++) p[OFF_W2 + k] = (float)(next() * std::sqrt(6.0 / NH));
    for (int k = 0; k < NH; k++) p[OFF_W3 + k] = (float)(next() * std::sqrt(6.0 / NH) * 0.1);
}

bool net_ready() {
    if (!g_net.open) {
        set_error(E_STATE, "nb6_rsnet: network not open");
        return false;
    }
    return true;
}

}  // namespace

int rsnet_nparam() { return NPARAM; }

int rsnet_open(unsigned long long seed) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    NetState &s = g_net;
    if (!s.open) {
        s.master.get(NPARAM * sizeof(float));
        s.m1.get(NPARAM * sizeof(float));
        s.m2.get(NPARAM * sizeof(float));
        s.wbf.get(NPARAM * sizeof(uint16_t));
        s.grad.get((NPARAM + 1) * sizeof(float));
        s.loss.get(sizeof(float));
        s.norm.get(2 * NF * sizeof(float));
        s.open = true;
    }
    std::vector<float> p;
    init_host(p, seed);
    std::vector<float> zero(NPARAM, 0.f), nrm(2 * NF, 0.f);
    for (int f = 0; f < NF; f++) nrm[NF + f] = 1.f;
    NB6_HIP(hipMemcpyAsync(s.master.p, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, d.stream));
    NB6_HIP(hipMemcpyAsync(s.m1.p, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, d.stream));
    NB6_HIP(hipMemcpyAsync(s.m2.p, zero.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, d.stream));
    NB6_HIP(hipMemcpyAsync(s.norm.p, nrm.data(), 2 * NF * sizeof(float), hipMemcpyHostToDevice, d.stream));
    hipLaunchKernelGGL(rsnet_round_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, d.stream,
                       s.master.as<float>(), s.wbf.as<uint16_t>());
    NB6_HIP(hipGetLastError());
    NB6_HIP(hipStreamSynchronize(d.stream));
    s.step = 0;
    return OK;
}

int rsnet_close() {
    NetState &s = g_net;
    if (!s.open) return OK;
    s.master.release(); s.m1.release(); s.m2.release(); s.wbf.release();
    s.grad.release(); s.part.release(); s.loss.release(); s.norm.release();
    s.lacc.release(); s.keys.release(); s.skeys.release(); s.vals.release(); s.perm.release();
    s.sort_tmp.release();
    s.open = false;
    return OK;
}

int rsnet_set_params(const float *p) {
    if (!net_ready()) return E_STATE;
    Device &d = device();
    NetState &s = g_net;
    NB6_HIP(hipMemcpyAsync(s.master.p, p, NPARAM * sizeof(float), hipMemcpyHostToDevice, d.stream));
    hipLaunchKernelGGL(rsnet_round_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, d.stream,
                       s.master.as<float>(), s.wbf.as<uint16_t>());
    NB6_HIP(hipGetLastError());
    NB6_HIP(hipStreamSynchronize(d.stream));
    return OK;
}

int rsnet_get_params(float *p) {
    if (!net_ready()) return E_STATE;
    Device &d = device();
    NB6_HIP(hipMemcpyAsync(p, g_net.master.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, d.stream));
    NB6_HIP(hipStreamSynchronize(d.stream));
    return OK;
}

int rsnet_set_norm(const float *mu, const float *sigma) {
    if (!net_ready()) return E_STATE;
    float h[2 * NF];
    for (int f = 0; f < NF; f++) {
        if (!(sigma[f] > 0.f)) {
            set_error(E_ARG, "nb6_rsnet_set_norm: every sigma must be positive");
            return E_ARG;
        }
        h[f] = mu[f];
        h[NF + f] = 1.f / sigma[f];
    }
    Device &d = device();
    NB6_HIP(hipMemcpyAsync(g_net.norm.p, h, sizeof(h), hipMemcpyHostToDevice, d.stream));
    NB6_HIP(hipStreamSynchronize(d.stream));
    return OK;
}

int rsnet_forward(int n, const float *d_feat, float *d_y, hipStream_t s) {
    if (!net_ready()) return E_STATE;
    if (n <= 0) return OK;
    hipLaunchKernelGGL(rsnet_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, d_feat,
                       g_net.wbf.as<uint16_t>(), g_net.norm.as<float>(), d_y);
    NB6_HIP(hipGetLastError());
    return OK;
}

// gradient of the mean squared error over n samples into d_grad (NPARAM) and the
// loss into *d_loss (both device pointers; null: the library's own buffers)
int rsnet_grad(int n, const float *d_feat, const float *d_t, float *d_grad, float *d_loss,
               hipStream_t s) {
    if (!net_ready()) return E_STATE;
    if (n <= 0) { set_error(E_ARG, "nb6_rsnet_grad: n must be positive"); return E_ARG; }
    NetState &g = g_net;
    const int nblk = (n + GB - 1) / GB;
    float *part = static_cast<float *>(g.part.get(((size_t)nblk * NPARAM + nblk) * sizeof(float)));
    float *lpart = part + (size_t)nblk * NPARAM;
    hipLaunchKernelGGL(rsnet_grad_kernel, dim3(nblk), dim3(GB), 0, s, n, d_feat, d_t, nullptr, 0,
                       n, g.wbf.as<uint16_t>(), g.norm.as<float>(), part, lpart);
    NB6_HIP(hipGetLastError());
    hipLaunchKernelGGL(rsnet_sum_kernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, part, nblk,
                       d_grad ? d_grad : g.grad.as<float>(), lpart, d_loss ? d_loss : g.loss.as<float>(), n,
                       nullptr);
    NB6_HIP(hipGetLastError());
    return OK;
}

// One epoch on the device: the n samples (SoA float[8][n], targets[n]) shuffled by a
// sort of seeded random keys, then n / batch Adam steps on consecutive minibatches of
// the shuffled order with no host synchronisation until the end; *h_loss = mean
// squared residual over the epoch (each residual before its step).
int rsnet_train_epoch(int n, const float *d_feat, const float *d_t, int batch, float lr,
                      unsigned long long seed, float *h_loss, hipStream_t s) {
    if (!net_ready()) return E_STATE;
    if (n <= 0 || batch <= 0 || batch > n) {
        set_error(E_ARG, "nb6_rsnet_train_epoch_device: need 0 < batch <= n");
        return E_ARG;
    }
    NetState &g = g_net;
    unsigned *key = static_cast<unsigned *>(g.keys.get((size_t)n * sizeof(unsigned)));
    unsigned *skey = static_cast<unsigned *>(g.skeys.get((size_t)n * sizeof(unsigned)));
    int *val = static_cast<int *>(g.vals.get((size_t)n * sizeof(int)));
    int *perm = static_cast<int *>(g.perm.get((size_t)n * sizeof(int)));
    hipLaunchKernelGGL(rsnet_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, seed, key, val);
    NB6_HIP(hipGetLastError());
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, skey, val, perm, n, 0, 32, s));
    void *tmp = g.sort_tmp.get(tb);
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key, skey, val, perm, n, 0, 32, s));
    double *lacc = static_cast<double *>(g.lacc.get(sizeof(double)));
    hipLaunchKernelGGL(rsnet_zero_kernel, dim3(1), dim3(1), 0, s, lacc);
    const int nblk = (batch + GB - 1) / GB;
    float *part = static_cast<float *>(g.part.get(((size_t)nblk * NPARAM + nblk) * sizeof(float)));
    float *lpart = part + (size_t)nblk * NPARAM;
    const int nb = n / batch;
    for (int b = 0; b < nb; b++) {
        hipLaunchKernelGGL(rsnet_grad_kernel, dim3(nblk), dim3(GB), 0, s, batch, d_feat, d_t, perm,
                           b * batch, n, g.wbf.as<uint16_t>(), g.norm.as<float>(), part, lpart);
        hipLaunchKernelGGL(rsnet_sum_kernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, part,
                           nblk, g.grad.as<float>(), lpart, g.loss.as<float>(), batch, lacc);
        g.step++;
        const float c1 = 1.f - std::pow(0.9f, (float)g.step);
        const float c2 = 1.f - std::pow(0.999f, (float)g.step);
        hipLaunchKernelGGL(rsnet_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s,
                           g.master.as<float>(), g.m1.as<float>(), g.m2.as<float>(),
                           g.grad.as<float>(), g.wbf.as<uint16_t>(), lr, c1, c2);
    }
    NB6_HIP(hipGetLastError());
    if (h_loss) {
        double l = 0.0;
        NB6_HIP(hipMemcpyAsync(&l, lacc, sizeof(double), hipMemcpyDeviceToHost, s));
        NB6_HIP(hipStreamSynchronize(s));
        *h_loss = (float)(l / ((double)nb * batch));
    }
    return OK;
}

int rsnet_train_step(int n, const float *d_feat, const float *d_t, float lr, float *h_loss,
                     hipStream_t s) {
    int rc = rsnet_grad(n, d_feat, d_t, nullptr, nullptr, s);
    if (rc != OK) return rc;
    NetState &g = g_net;
    g.step++;
    const float c1 = 1.f - std::pow(0.9f, (float)g.step);
    const float c2 = 1.f - std::pow(0.999f, (float)g.step);
    hipLaunchKernelGGL(rsnet_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s,
                       g.master.as<float>(), g.m1.as<float>(), g.m2.as<float>(), g.grad.as<float>(),
                       g.wbf.as<uint16_t>(), lr, c1, c2);
    NB6_HIP(hipGetLastError());
    if (h_loss) {
        NB6_HIP(hipMemcpyAsync(h_loss, g.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        NB6_HIP(hipStreamSynchronize(s));
    }
    return OK;
}

int rs_features(int n, int stride, const double *x, const double *acc, const double *jrk,
                const double *mass, const double *rs, const double *dtr, const int *nnb,
                const double *centre, float *feat, hipStream_t s) {
    if (n <= 0) return OK;
    hipLaunchKernelGGL(rs_feat_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, stride, x, acc, jrk,
                       mass, rs, dtr, nnb, centre[0], centre[1], centre[2], feat);
    NB6_HIP(hipGetLastError());
    return OK;
}

int rs_predict(int n, const float *feat, const double *rs_prev, const double *rs_rule,
               const int *use_rule, double *rs_out, double *ib, int ib_stride, int m_flag,
               double bodym, hipStream_t s) {
    if (!net_ready()) return E_STATE;
    if (n <= 0) return OK;
    if (ib && ib_stride < n) { set_error(E_ARG, "nb6_rs_predict_device: stride below n"); return E_ARG; }
    if (m_flag && !(bodym > 0.0)) { set_error(E_ARG, "nb6_rs_predict_device: bodym must be positive"); return E_ARG; }
    hipLaunchKernelGGL(rs_predict_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, feat,
                       g_net.wbf.as<uint16_t>(), g_net.norm.as<float>(), rs_prev, rs_rule, use_rule,
                       rs_out, ib, ib_stride, m_flag, bodym);
    NB6_HIP(hipGetLastError());
    return OK;
}

int rs_rule(int n, int stride, const double *x, const double *v, const double *dtr, const int *nnb,
            const int *nnb0, const int *list, int lmax, const double *xj, const double *vj,
            int jstride, const double *glob, int nnbopt, int nnbmax, double *rs, hipStream_t s) {
    if (n <= 0) return OK;
    RuleParams q;
    q.n = n; q.stride = stride; q.x = x; q.v = v; q.dtr = dtr; q.nnb = nnb; q.nnb0 = nnb0;
    q.list = list; q.lmax = lmax; q.xj = xj; q.vj = vj; q.jstride = jstride;
    q.rdx = glob[0]; q.rdy = glob[1]; q.rdz = glob[2];
    q.rscale = glob[3]; q.rc2 = glob[4]; q.rsfac = glob[5]; q.znbmin = glob[6];
    q.nnbopt = nnbopt; q.nnbmax = nnbmax; q.rs = rs;
    hipLaunchKernelGGL(rs_rule_kernel, dim3((n + 255) / 256), dim3(256), 0, s, q);
    NB6_HIP(hipGetLastError());
    return OK;
}

int rs_kth(const JSet &js, int ni, int ib_stride, const double *ib, const int *list, int lmax, int k,
           double *rs_out, hipStream_t s) {
    if (ni <= 0) return OK;
    if (k < 1 || lmax < 2) { set_error(E_ARG, "nb6_rs_kth_device: need k >= 1 and lmax >= 2"); return E_ARG; }
    hipLaunchKernelGGL(rs_kth_kernel, dim3((ni + 3) / 4), dim3(256), 0, s, ni, ib_stride, ib, list,
                       lmax, k, js.in64.as<double>(), js.n, rs_out);
    NB6_HIP(hipGetLastError());
    return OK;
}

}  // namespace nb6
