// Device-side layout of the j-particles (what gpunb_send_ uploads).
//
// The reference packs j into FP32 absolute coordinates (Jparticle,
// gpunb.velocity.cu:58-93).  Here the j-set is kept in FP64 and re-ordered
// along a Morton (Z-order) curve on the device at every send, so that each tile
// of 256 consecutive j occupies a small region of space.  Every tile t has an
// origin (P_t, V_t): P_t the position of its middle record, V_t the mean velocity
// of its massive records (the middle record's own velocity would put the orbital
// velocity of a tight binary member into every relative velocity of the tile);
// the regular-force kernel streams
//   rec32[s] = { x - Px, y - Py, z - Pz, m,  vx - Vx, vy - Vy, vz - Vz, 0 }  (FP32)
// and forms the i side relative to the same origin in FP64, so the FP32
// separation carries the rounding error of the tile's extent, not of |x|.
// rec64 (sorted, AoS, 64 B) feeds the FP64 near-field / neighbour re-test, and
// perm[s] maps a sorted position back to the caller's 0-based j index.
#include "nb6_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>

namespace nb6 {

namespace {

// 21 bits per axis: a cluster with escapers far out still gets a fine order in its core
__device__ __forceinline__ unsigned long long spread21(unsigned long long a) {
    a &= 0x1fffffull;
    a = (a | (a << 32)) & 0x1f00000000ffffull;
    a = (a | (a << 16)) & 0x1f0000ff0000ffull;
    a = (a | (a << 8)) & 0x100f00f00f00f00full;
    a = (a | (a << 4)) & 0x10c30c30c30c30c3ull;
    a = (a | (a << 2)) & 0x1249249249249249ull;
    return a;
}

constexpr double MORTON_CELLS = 2097152.0;   // 2^21 per axis

__device__ __forceinline__ unsigned long long morton63(double x, double y, double z, double lx,
                                                       double ly, double lz, double sx, double sy,
                                                       double sz) {
    const double fx = (x - lx) * sx, fy = (y - ly) * sy, fz = (z - lz) * sz;
    // comparisons in FP64 first: NaN and far-outside values clamp without overflow
    const unsigned qx = fx > 0.0 ? (fx < MORTON_CELLS - 1.0 ? (unsigned)fx : 2097151u) : 0u;
    const unsigned qy = fy > 0.0 ? (fy < MORTON_CELLS - 1.0 ? (unsigned)fy : 2097151u) : 0u;
    const unsigned qz = fz > 0.0 ? (fz < MORTON_CELLS - 1.0 ? (unsigned)fz : 2097151u) : 0u;
    return spread21(qx) | (spread21(qy) << 1) | (spread21(qz) << 2);
}

__global__ __launch_bounds__(256) void morton_kernel(const double *__restrict__ in, int n,
                                                     double lx, double ly, double lz, double sx,
                                                     double sy, double sz,
                                                     unsigned long long *__restrict__ key,
                                                     int *__restrict__ val) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double *r = in + (size_t)j * 8;
    key[j] = morton63(r[0], r[1], r[2], lx, ly, lz, sx, sy, sz);
    val[j] = j;
}

// sorted position s <- original j = perm[s]
__global__ __launch_bounds__(256) void build_kernel(const double *__restrict__ in, int n,
                                                    const int *__restrict__ perm,
                                                    double *__restrict__ rec64) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const double *r = in + (size_t)perm[s] * 8;
    double *o = rec64 + (size_t)s * 8;
#pragma unroll
    for (int f = 0; f < 8; f++) o[f] = r[f];
}

// One work-group per tile: P = position of record `pos` of the tile (its middle one,
// or its first), V = mean velocity of the tile's records with m > 0 (that record's
// velocity if there are none).
__global__ __launch_bounds__(256) void origin_kernel(const double *__restrict__ rec64, int n,
                                                     int middle, double *__restrict__ torg) {
    __shared__ double sh[4][4];
    const int t = blockIdx.x;
    const int c = min(JTILE, n - t * JTILE);
    const int k = threadIdx.x;
    const double *r = rec64 + ((size_t)t * JTILE + k) * 8;
    const bool use = k < c && r[3] > 0.0;
    double q[4] = {use ? r[4] : 0.0, use ? r[5] : 0.0, use ? r[6] : 0.0, use ? 1.0 : 0.0};
#pragma unroll
    for (int f = 0; f < 4; f++)
        for (int o = 32; o > 0; o >>= 1) q[f] += __shfl_xor(q[f], o, 64);
    if ((k & 63) == 0)
        for (int f = 0; f < 4; f++) sh[k >> 6][f] = q[f];
    __syncthreads();
    if (k == 0) {
        double v[4];
        for (int f = 0; f < 4; f++) v[f] = (sh[0][f] + sh[1][f]) + (sh[2][f] + sh[3][f]);
        const double *p = rec64 + ((size_t)t * JTILE + (middle ? c / 2 : 0)) * 8;
        double *o = torg + (size_t)t * 8;
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        if (v[3] > 0.0) {
            o[3] = v[0] / v[3]; o[4] = v[1] / v[3]; o[5] = v[2] / v[3];
        } else {
            o[3] = p[4]; o[4] = p[5]; o[5] = p[6];
        }
        o[6] = 0.0; o[7] = 0.0;
    }
}

// FP32 records relative to their tile's origin
__global__ __launch_bounds__(256) void rec32_kernel(const double *__restrict__ rec64, int n,
                                                    const double *__restrict__ torg,
                                                    float4 *__restrict__ rec32) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const double *p = torg + (size_t)(s / JTILE) * 8;
    const double *r = rec64 + (size_t)s * 8;
    rec32[2 * (size_t)s] = make_float4((float)(r[0] - p[0]), (float)(r[1] - p[1]),
                                       (float)(r[2] - p[2]), (float)r[3]);
    rec32[2 * (size_t)s + 1] = make_float4((float)(r[4] - p[3]), (float)(r[5] - p[4]),
                                           (float)(r[6] - p[5]), 0.f);
}

// SoA [7][stride] (x y z vx vy vz m) -> AoS [n][8] (x y z m vx vy vz 0)
__global__ __launch_bounds__(256) void soa_to_aos_kernel(const double *__restrict__ soa,
                                                         int n, int stride,
                                                         double *__restrict__ aos) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double *o = aos + (size_t)j * 8;
    o[0] = soa[0 * (size_t)stride + j];
    o[1] = soa[1 * (size_t)stride + j];
    o[2] = soa[2 * (size_t)stride + j];
    o[3] = soa[6 * (size_t)stride + j];
    o[4] = soa[3 * (size_t)stride + j];
    o[5] = soa[4 * (size_t)stride + j];
    o[6] = soa[5 * (size_t)stride + j];
    o[7] = 0.0;
}

// bounding box of positions and maximum mass, pass 1: each work-group reduces a
// strided share of the records into one partial row of 7
constexpr int STATS_GROUPS = 256;

__global__ __launch_bounds__(256) void stats_kernel(const double *__restrict__ aos, int n,
                                                    double *__restrict__ part /* [groups][7] */) {
    __shared__ double sh[7][256];
    double v[7] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX, 0.0};
    for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) {
        const double *r = aos + (size_t)j * 8;
        for (int d = 0; d < 3; d++) {
            v[d] = fmin(v[d], r[d]);
            v[3 + d] = fmax(v[3 + d], r[d]);
        }
        v[6] = fmax(v[6], r[3]);
    }
    for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int d = 0; d < 3; d++) sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + o]);
            for (int d = 3; d < 7; d++) sh[d][threadIdx.x] = fmax(sh[d][threadIdx.x], sh[d][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 7 + threadIdx.x] = sh[threadIdx.x][0];
}

// pass 2: one work-group reduces the partial rows
__global__ __launch_bounds__(256) void stats_final_kernel(const double *__restrict__ part, int groups,
                                                          double *__restrict__ out /* 7 */) {
    __shared__ double sh[7][256];
    const int t = threadIdx.x;
    for (int q = 0; q < 7; q++) {
        const double id = q < 3 ? DBL_MAX : (q < 6 ? -DBL_MAX : 0.0);
        sh[q][t] = t < groups ? part[t * 7 + q] : id;
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) {
            for (int d = 0; d < 3; d++) sh[d][t] = fmin(sh[d][t], sh[d][t + o]);
            for (int d = 3; d < 7; d++) sh[d][t] = fmax(sh[d][t], sh[d][t + o]);
        }
        __syncthreads();
    }
    if (t < 7) out[t] = sh[t][0];
}

// Morton key of i-particle positions (SoA rows 0-2) in a given box (clamped)
__global__ __launch_bounds__(256) void imorton_kernel(const double *__restrict__ ib, int stride,
                                                      int ni, double lx, double ly, double lz,
                                                      double sx, double sy, double sz,
                                                      unsigned long long *__restrict__ key,
                                                      int *__restrict__ val) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ni) return;
    key[i] = morton63(ib[i], ib[(size_t)stride + i], ib[2 * (size_t)stride + i], lx, ly, lz, sx,
                      sy, sz);
    val[i] = i;
}

// FP32 bounds of one tile's tile-relative records (see JSet::tbox)
__global__ __launch_bounds__(256) void tile_box_kernel(const float4 *__restrict__ rec32, int n,
                                                       float *__restrict__ tbox) {
    __shared__ float sh[14][256];
    const int t = blockIdx.x;
    const int s = t * JTILE + threadIdx.x;
    float v[14];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        v[q] = FLT_MAX; v[3 + q] = -FLT_MAX; v[6 + q] = FLT_MAX; v[9 + q] = -FLT_MAX;
    }
    v[12] = 0.f;
    v[13] = 0.f;
    if (threadIdx.x < JTILE && s < n) {
        const float4 a = rec32[2 * (size_t)s];
        const float4 b = rec32[2 * (size_t)s + 1];
        if (fabsf(a.x) < 1.0e16f) {
            v[0] = v[3] = a.x; v[1] = v[4] = a.y; v[2] = v[5] = a.z;
            v[6] = v[9] = b.x; v[7] = v[10] = b.y; v[8] = v[11] = b.z;
            v[12] = a.w;
            v[13] = 1.f;
        }
    }
#pragma unroll
    for (int q = 0; q < 14; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            const int u = threadIdx.x, w = threadIdx.x + o;
            for (int q = 0; q < 3; q++) {
                sh[q][u] = fminf(sh[q][u], sh[q][w]);
                sh[3 + q][u] = fmaxf(sh[3 + q][u], sh[3 + q][w]);
                sh[6 + q][u] = fminf(sh[6 + q][u], sh[6 + q][w]);
                sh[9 + q][u] = fmaxf(sh[9 + q][u], sh[9 + q][w]);
            }
            sh[12][u] = fmaxf(sh[12][u], sh[12][w]);
            sh[13][u] += sh[13][w];
        }
        __syncthreads();
    }
    if (threadIdx.x < TBOX) tbox[(size_t)t * TBOX + threadIdx.x] = threadIdx.x < 14 ? sh[threadIdx.x][0] : 0.f;
}

// plain streaming copy (used to calibrate HBM byte counters against a known size)
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, size_t n16) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n16; k += (size_t)gridDim.x * 256)
        dst[k] = src[k];
}

}  // namespace

void device_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    if (n16 == 0) return;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    NB6_HIP(hipGetLastError());
}

void jset_reserve(JSet &js, int cap) {
    if (cap <= js.cap) return;
    js.cap = cap;
    const int ntile = (cap + JTILE - 1) / JTILE;
    js.in64.get((size_t)cap * 8 * sizeof(double));
    js.rec64.get((size_t)cap * 8 * sizeof(double));
    js.rec32.get((size_t)cap * 2 * sizeof(float4));
    js.torg.get((size_t)ntile * 8 * sizeof(double));
    js.perm.get((size_t)cap * sizeof(int));
    js.keys.get((size_t)cap * 2 * sizeof(unsigned long long));
    js.vals.get((size_t)cap * sizeof(int));
    js.stat.get((8 + 7 * STATS_GROUPS) * sizeof(double));
    js.tbox.get((size_t)ntile * TBOX * sizeof(float));
}

void jset_tile_boxes(JSet &js, int n, hipStream_t s) {
    const int nt = (n + JTILE - 1) / JTILE;
    if (nt <= 0) return;
    hipLaunchKernelGGL(tile_box_kernel, dim3(nt), dim3(256), 0, s, js.rec32.as<float4>(), n,
                       js.tbox.as<float>());
    NB6_HIP(hipGetLastError());
}

void jset_release(JSet &js) {
    js.in64.release(); js.rec64.release(); js.rec32.release(); js.torg.release();
    js.perm.release(); js.keys.release(); js.vals.release(); js.stat.release();
    js.sort_tmp.release(); js.tbox.release();
    js.cap = 0; js.n = 0; js.ntile = 0;
}

// in64 holds n AoS records in the caller's order; box = {lo[3], hi[3]} or nullptr
void jset_build(JSet &js, int n, const double *box, double mmax, hipStream_t s) {
    js.n = n;
    js.ntile = (n + JTILE - 1) / JTILE;
    if (n <= 0) return;
    double st[7];
    if (!box) {
        const int groups = std::min(STATS_GROUPS, (n + 255) / 256);
        double *part = js.stat.as<double>() + 8;
        hipLaunchKernelGGL(stats_kernel, dim3(groups), dim3(256), 0, s, js.in64.as<double>(), n, part);
        hipLaunchKernelGGL(stats_final_kernel, dim3(1), dim3(256), 0, s, part, groups,
                           js.stat.as<double>());
        NB6_HIP(hipGetLastError());
        NB6_HIP(hipMemcpyAsync(st, js.stat.p, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
        NB6_HIP(hipStreamSynchronize(s));
        box = st;
        mmax = st[6];
    }
    js.mmax = mmax;
    for (int d = 0; d < 6; d++) js.box[d] = box[d];
    double sc[3];
    for (int d = 0; d < 3; d++) {
        const double w = box[3 + d] - box[d];
        sc[d] = w > 0.0 ? MORTON_CELLS / (w * (1.0 + 1e-12)) : 0.0;
    }
    unsigned long long *k0 = js.keys.as<unsigned long long>();
    unsigned long long *k1 = k0 + js.cap;
    int *v0 = js.vals.as<int>();
    int *v1 = js.perm.as<int>();
    hipLaunchKernelGGL(morton_kernel, dim3((n + 255) / 256), dim3(256), 0, s, js.in64.as<double>(),
                       n, box[0], box[1], box[2], sc[0], sc[1], sc[2], k0, v0);
    NB6_HIP(hipGetLastError());
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, n, 0, 63, s));
    void *tmp = js.sort_tmp.get(tb);
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, n, 0, 63, s));
    hipLaunchKernelGGL(build_kernel, dim3((n + 255) / 256), dim3(256), 0, s, js.in64.as<double>(), n,
                       js.perm.as<int>(), js.rec64.as<double>());
    NB6_HIP(hipGetLastError());
    jset_tiles(js, n, 1, s);
}

void jset_tiles(JSet &js, int n, int middle, hipStream_t s) {
    const int nt = (n + JTILE - 1) / JTILE;
    hipLaunchKernelGGL(origin_kernel, dim3(nt), dim3(256), 0, s, js.rec64.as<double>(), n, middle,
                       js.torg.as<double>());
    hipLaunchKernelGGL(rec32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, js.rec64.as<double>(),
                       n, js.torg.as<double>(), js.rec32.as<float4>());
    NB6_HIP(hipGetLastError());
    jset_tile_boxes(js, n, s);
}

const int *jset_sort_i(const JSet &js, int ni, int stride, const double *d_i, RegfWork &w,
                       hipStream_t s) {
    double sc[3];
    for (int d = 0; d < 3; d++) {
        const double wd = js.box[3 + d] - js.box[d];
        sc[d] = wd > 0.0 ? MORTON_CELLS / (wd * (1.0 + 1e-12)) : 0.0;
    }
    unsigned long long *k0 =
        static_cast<unsigned long long *>(w.ikey.get((size_t)ni * 2 * sizeof(unsigned long long)));
    unsigned long long *k1 = k0 + ni;
    int *v0 = static_cast<int *>(w.ival.get((size_t)ni * 2 * sizeof(int)));
    int *v1 = v0 + ni;
    hipLaunchKernelGGL(imorton_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, d_i, stride, ni,
                       js.box[0], js.box[1], js.box[2], sc[0], sc[1], sc[2], k0, v0);
    NB6_HIP(hipGetLastError());
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, ni, 0, 63, s));
    void *tmp = w.itmp.get(tb);
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, ni, 0, 63, s));
    return v1;
}

namespace {
// row i of a [n][lmax] list array holds 1 + count ints (1 if the count is negative or
// above maxc); len[n] = 0 so that the exclusive scan's last entry is the total
__global__ __launch_bounds__(256) void row_len_kernel(const int *__restrict__ list, int n, int lmax,
                                                      int maxc, int *__restrict__ len) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i > n) return;
    int l = 0;
    if (i < n) {
        const int c = list[(size_t)i * lmax];
        l = 1 + (c > 0 && c <= maxc ? c : 0);
    }
    len[i] = l;
}

// one wavefront per row: copy its 1 + count ints to pack + off[i]
__global__ __launch_bounds__(256) void row_pack_kernel(const int *__restrict__ list, int n, int lmax,
                                                       const int *__restrict__ off,
                                                       int *__restrict__ pack) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int lane = threadIdx.x & 63;
    const int o = off[i], l = off[i + 1] - o;
    const int *src = list + (size_t)i * lmax;
    for (int k = lane; k < l; k += 64) pack[o + k] = src[k];
}
}  // namespace

void pack_rows(const int *list, int n, int lmax, int maxc, int *len, int *off, int *pack,
               DBuf &tmp, hipStream_t s) {
    hipLaunchKernelGGL(row_len_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, s, list, n, lmax,
                       maxc, len);
    NB6_HIP(hipGetLastError());
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len, off, n + 1, s));
    void *t = tmp.get(tb);
    NB6_HIP(hipcub::DeviceScan::ExclusiveSum(t, tb, len, off, n + 1, s));
    hipLaunchKernelGGL(row_pack_kernel, dim3((n + 3) / 4), dim3(256), 0, s, list, n, lmax, off, pack);
    NB6_HIP(hipGetLastError());
}

void jset_from_soa(JSet &js, int n, int stride, const double *d_soa, hipStream_t s) {
    if (n > 0) {
        hipLaunchKernelGGL(soa_to_aos_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_soa, n,
                           stride, js.in64.as<double>());
        NB6_HIP(hipGetLastError());
    }
    jset_build(js, n, nullptr, 0.0, s);
}

}  // namespace nb6
