// Batched regular force for an ensemble of independent clusters.
//
// The reference has one j-set per library instance (gpunb_open_/gpunb_send_,
// gpunb.velocity.cu:629-730), so many small clusters mean many small launches.
// Here all clusters live in one device j-set:
//   * send: one segmented Morton sort of all particles (key = cluster << 32 |
//     Morton code inside the cluster's own bounding box); each cluster's sorted
//     range is padded to whole 256-j tiles with far, massless dummies so no tile
//     spans two clusters, and every tile gets its FP64 origin (as in jset.hip);
//   * regular step: one persistent launch of regf_items_kernel walks a queue of
//     (cluster, i-block) work items ordered longest first, then one finalize over
//     all particles (per-i cluster base for the output indices).  The step (queue
//     reset, persistent kernel, finalize) is captured once into a hipGraph and
//     replayed while the buffers stay the same.
// Each cluster's results are exactly those of a gpunb_send_ + gpunb_regf_ of
// that cluster alone (same lists, same FP64 re-test).
#include "nb6_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <numeric>
#include <vector>

namespace nb6 {

namespace {

__device__ __forceinline__ unsigned spread10e(unsigned a) {
    a &= 0x3ffu;
    a = (a | (a << 16)) & 0x030000FFu;
    a = (a | (a << 8)) & 0x0300F00Fu;
    a = (a | (a << 4)) & 0x030C30C3u;
    a = (a | (a << 2)) & 0x09249249u;
    return a;
}

// per-cluster bounding box and maximum mass: one work-group per cluster
__global__ __launch_bounds__(256) void ens_bbox_kernel(const double *__restrict__ aos,
                                                       const int *__restrict__ off,
                                                       double *__restrict__ box) {
    __shared__ double sh[7][256];
    const int c = blockIdx.x;
    double v[7] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX, 0.0};
    for (int j = off[c] + threadIdx.x; j < off[c + 1]; j += 256) {
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
    if (threadIdx.x < 7) box[c * 7 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void ens_key_kernel(const double *__restrict__ aos, int n,
                                                      const int *__restrict__ cid,
                                                      const double *__restrict__ box,
                                                      unsigned long long *__restrict__ key,
                                                      int *__restrict__ val) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int c = cid[j];
    const double *r = aos + (size_t)j * 8;
    const double *b = box + c * 7;
    unsigned q[3];
    for (int d = 0; d < 3; d++) {
        const double w = b[3 + d] - b[d];
        const double sc = w > 0.0 ? 1024.0 / (w * (1.0 + 1e-12)) : 0.0;
        q[d] = (unsigned)min(1023, max(0, (int)((r[d] - b[d]) * sc)));
    }
    key[j] = ((unsigned long long)c << 32) |
             (spread10e(q[0]) | (spread10e(q[1]) << 1) | (spread10e(q[2]) << 2));
    val[j] = j;
}

// fill every padded slot with a far massless dummy (overwritten for real particles)
__global__ __launch_bounds__(256) void ens_dummy_kernel(double *__restrict__ rec64,
                                                        int *__restrict__ perm, int npad) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= npad) return;
    double *o = rec64 + (size_t)s * 8;
    o[0] = 1.0e17; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0;
    o[4] = 0.0; o[5] = 0.0; o[6] = 0.0; o[7] = 0.0;
    perm[s] = -1;
}

// compact sorted position s -> padded position (cluster start + rank in cluster)
__global__ __launch_bounds__(256) void ens_scatter_kernel(const double *__restrict__ in, int n,
                                                          const int *__restrict__ sorted,
                                                          const int *__restrict__ cid,
                                                          const int *__restrict__ off,
                                                          const int *__restrict__ padoff,
                                                          double *__restrict__ rec64,
                                                          int *__restrict__ perm) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int j = sorted[s];
    const int c = cid[j];
    const int ps = padoff[c] + (s - off[c]);   // clusters keep their order in the sort
    const double *r = in + (size_t)j * 8;
    double *o = rec64 + (size_t)ps * 8;
#pragma unroll
    for (int f = 0; f < 8; f++) o[f] = r[f];
    perm[ps] = j;
}

__global__ __launch_bounds__(256) void ens_soa_to_aos(const double *__restrict__ soa, int n,
                                                      int stride, double *__restrict__ aos) {
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

__global__ void ens_zero_kernel(int *q) { *q = 0; }

struct GraphKey {
    const void *ptr[4] = {nullptr, nullptr, nullptr, nullptr};   // d_i, d_out, d_list, items
    long long val[18] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool operator==(const GraphKey &o) const {
        return std::equal(ptr, ptr + 4, o.ptr) && std::equal(val, val + 18, o.val);
    }
};

struct Ens {
    bool open = false;
    int max_clusters = 0, max_particles = 0;
    int nc = 0, ntot = 0, npad = 0, maxseg = 0;   // maxseg: longest padded cluster
    std::vector<int> off, padoff;
    JSet js;                         // padded, sorted j-set of all clusters
    DBuf d_off, d_padoff, d_cid, d_box, d_keys, d_vals, d_skeys, d_svals, d_tmp, d_jbase;
    DBuf d_items, d_queue;
    RegfWork work;                   // own scratch: a captured graph holds its addresses
    int nitems = 0;
    int items_bi = 0;                // i-block the items were cut for
    bool layout = false;             // off/padoff, cluster ids and bases on the device match
                                     // the last send's offsets
    int per_cu = 0;                  // persistent work-groups per CU (0: 8 waves per SIMD)
    // graph of the last regular step and everything it captured by address or value
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey gkey;
} g_ens;

// work items: (cluster, block of Morton-ordered i slots), longest first; the real
// particles of cluster c fill slots [padoff[c], padoff[c] + n_c); the i-block is the
// current work-group size times i-particles per lane
int build_items(Ens &e, hipStream_t s) {
    const RegfConfig cfg = regf_effective(0);
    const int bi = cfg.threads * cfg.ipl;
    std::vector<int4> items;
    std::vector<long long> cost;
    for (int c = 0; c < e.nc; c++) {
        const int nci = e.off[c + 1] - e.off[c];
        const int plen = e.padoff[c + 1] - e.padoff[c];
        const int send = e.padoff[c] + nci;
        for (int ib = e.padoff[c]; ib < send; ib += bi) {
            items.push_back(make_int4(ib, send, e.padoff[c], e.padoff[c + 1]));
            cost.push_back((long long)std::min(bi, send - ib) * plen);
        }
    }
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int4> sorted(items.size());
    for (size_t k = 0; k < items.size(); k++) sorted[k] = items[order[k]];
    e.nitems = (int)sorted.size();
    e.items_bi = bi;
    e.d_items.get(sorted.size() * sizeof(int4));
    NB6_HIP(hipMemcpyAsync(e.d_items.p, sorted.data(), sorted.size() * sizeof(int4),
                           hipMemcpyHostToDevice, s));
    NB6_HIP(hipStreamSynchronize(s));
    return OK;
}

void drop_graph() {
    if (g_ens.exec) (void)hipGraphExecDestroy(g_ens.exec);
    if (g_ens.graph) (void)hipGraphDestroy(g_ens.graph);
    g_ens.exec = nullptr;
    g_ens.graph = nullptr;
}

}  // namespace

int ens_open(int max_clusters, int max_particles) {
    Ens &e = g_ens;
    if (e.open) {
        fprintf(stderr, "nb6 ensemble: it is already open\n");
        return OK;
    }
    if (max_clusters <= 0 || max_particles <= 0) {
        set_error(E_ARG, "nb6_ens_open: sizes must be positive");
        return E_ARG;
    }
    Device &d = device();
    if (!d.ready) return E_NODEV;
    e.max_clusters = max_clusters;
    e.max_particles = max_particles;
    const int padmax = max_particles + max_clusters * JTILE;
    jset_reserve(e.js, padmax);
    e.d_off.get((size_t)(max_clusters + 1) * sizeof(int));
    e.d_padoff.get((size_t)(max_clusters + 1) * sizeof(int));
    e.d_cid.get((size_t)max_particles * sizeof(int));
    e.d_box.get((size_t)max_clusters * 7 * sizeof(double));
    e.d_keys.get((size_t)max_particles * sizeof(unsigned long long));
    e.d_skeys.get((size_t)max_particles * sizeof(unsigned long long));
    e.d_vals.get((size_t)max_particles * sizeof(int));
    e.d_svals.get((size_t)max_particles * sizeof(int));
    e.d_jbase.get((size_t)max_particles * sizeof(int));
    e.d_queue.get(sizeof(int));
    e.open = true;
    return OK;
}

int ens_close() {
    Ens &e = g_ens;
    if (!e.open) return OK;
    drop_graph();
    jset_release(e.js);
    e.d_off.release(); e.d_padoff.release(); e.d_cid.release(); e.d_box.release();
    e.d_keys.release(); e.d_skeys.release(); e.d_vals.release(); e.d_svals.release();
    e.d_tmp.release(); e.d_jbase.release(); e.d_items.release(); e.d_queue.release();
    e.work.release();
    regf_items_forget();
    e.nc = 0;
    e.per_cu = 0;
    e.items_bi = 0;
    e.layout = false;
    e.open = false;
    return OK;
}

// offsets[nc+1] (host) into the concatenated SoA [7][stride] device array
int ens_send_device(int nc, const int *offsets, const double *d_soa, int stride, hipStream_t s) {
    Ens &e = g_ens;
    if (!e.open) { set_error(E_STATE, "nb6_ens_send_device: ensemble not open"); return E_STATE; }
    if (nc <= 0 || nc > e.max_clusters || offsets[0] != 0 || offsets[nc] > e.max_particles ||
        stride < offsets[nc]) {
        set_error(E_ARG, "nb6_ens_send_device: bad cluster count, offsets or stride");
        return E_ARG;
    }
    for (int c = 0; c < nc; c++)
        if (offsets[c + 1] <= offsets[c]) {
            set_error(E_ARG, "nb6_ens_send_device: every cluster needs at least one particle");
            return E_ARG;
        }
    // the cluster layout (padded offsets, per-particle cluster ids and bases, work items)
    // depends on the offsets alone: an ensemble sent again with the same offsets, as
    // every step of an integration is, keeps it
    const bool same_layout = e.layout && e.nc == nc && (int)e.off.size() == nc + 1 &&
                             std::equal(offsets, offsets + nc + 1, e.off.begin());
    if (!same_layout) {
        e.layout = false;
        e.nc = nc;
        e.ntot = offsets[nc];
        e.off.assign(offsets, offsets + nc + 1);
        e.padoff.assign(nc + 1, 0);
        for (int c = 0; c < nc; c++) {
            const int n = offsets[c + 1] - offsets[c];
            e.padoff[c + 1] = e.padoff[c] + (n + JTILE - 1) / JTILE * JTILE;
        }
        e.npad = e.padoff[nc];
        e.maxseg = 0;
        for (int c = 0; c < nc; c++) e.maxseg = std::max(e.maxseg, e.padoff[c + 1] - e.padoff[c]);
        std::vector<int> cid(e.ntot), jb(e.ntot);
        for (int c = 0; c < nc; c++)
            for (int j = offsets[c]; j < offsets[c + 1]; j++) {
                cid[j] = c;
                jb[j] = offsets[c];
            }
        NB6_HIP(hipMemcpyAsync(e.d_off.p, e.off.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice, s));
        NB6_HIP(hipMemcpyAsync(e.d_padoff.p, e.padoff.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice, s));
        NB6_HIP(hipMemcpyAsync(e.d_cid.p, cid.data(), e.ntot * sizeof(int), hipMemcpyHostToDevice, s));
        NB6_HIP(hipMemcpyAsync(e.d_jbase.p, jb.data(), e.ntot * sizeof(int), hipMemcpyHostToDevice, s));
        // (the copies read host vectors that end with this block)
        NB6_HIP(hipStreamSynchronize(s));
    }
    const int n = e.ntot;
    double *in64 = e.js.in64.as<double>();
    hipLaunchKernelGGL(ens_soa_to_aos, dim3((n + 255) / 256), dim3(256), 0, s, d_soa, n, stride, in64);
    hipLaunchKernelGGL(ens_bbox_kernel, dim3(nc), dim3(256), 0, s, in64, e.d_off.as<int>(),
                       e.d_box.as<double>());
    hipLaunchKernelGGL(ens_key_kernel, dim3((n + 255) / 256), dim3(256), 0, s, in64, n,
                       e.d_cid.as<int>(), e.d_box.as<double>(),
                       e.d_keys.as<unsigned long long>(), e.d_vals.as<int>());
    NB6_HIP(hipGetLastError());
    int hi = 32;
    while ((1 << (hi - 32)) <= nc && hi < 63) hi++;
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, e.d_keys.as<unsigned long long>(),
                                               e.d_skeys.as<unsigned long long>(), e.d_vals.as<int>(),
                                               e.d_svals.as<int>(), n, 0, hi, s));
    void *tmp = e.d_tmp.get(tb);
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, e.d_keys.as<unsigned long long>(),
                                               e.d_skeys.as<unsigned long long>(), e.d_vals.as<int>(),
                                               e.d_svals.as<int>(), n, 0, hi, s));
    const int np = e.npad;
    hipLaunchKernelGGL(ens_dummy_kernel, dim3((np + 255) / 256), dim3(256), 0, s,
                       e.js.rec64.as<double>(), e.js.perm.as<int>(), np);
    hipLaunchKernelGGL(ens_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, in64, n,
                       e.d_svals.as<int>(), e.d_cid.as<int>(), e.d_off.as<int>(),
                       e.d_padoff.as<int>(), e.js.rec64.as<double>(), e.js.perm.as<int>());
    NB6_HIP(hipGetLastError());
    // tile origins and FP32 records on the padded layout: P from a tile's first record
    // (padding sits at the end of a cluster's last tile)
    jset_tiles(e.js, np, 0, s);
    e.js.n = np;
    e.js.ntile = np / JTILE;
    // maximum mass, for the m_flag error band
    std::vector<double> box((size_t)nc * 7);
    NB6_HIP(hipMemcpyAsync(box.data(), e.d_box.p, box.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NB6_HIP(hipStreamSynchronize(s));
    double mmax = 0.0;
    for (int c = 0; c < nc; c++) mmax = std::max(mmax, box[(size_t)c * 7 + 6]);
    e.js.mmax = mmax;
    const RegfConfig cfg = regf_effective(0);
    if (same_layout && e.items_bi == cfg.threads * cfg.ipl) return OK;
    const int rc = build_items(e, s);
    if (rc == OK) e.layout = true;
    return rc;
}

int ens_regf_device(const double *d_i, int stride, int lmax, int nnbmax, int m_flag, double *d_out,
                    int *d_list, int use_graph, hipStream_t s) {
    Ens &e = g_ens;
    if (!e.open || e.nc == 0) { set_error(E_STATE, "nb6_ens_regf_device: nothing sent"); return E_STATE; }
    if (stride < e.ntot) { set_error(E_ARG, "nb6_ens_regf_device: stride below the particle count"); return E_ARG; }
    Device &d = device();
    const RegfConfig cfg = regf_effective(0);
    if (e.items_bi != cfg.threads * cfg.ipl) {   // configuration changed since the send
        drop_graph();
        const int rc = build_items(e, s);
        if (rc != OK) return rc;
    }
    const int per_cu = e.per_cu > 0 ? e.per_cu : std::max(1, (8 * 256) / cfg.threads);
    const int grid = std::min(e.nitems, d.num_cu * per_cu);
    auto step = [&](hipStream_t st) {
        hipLaunchKernelGGL(ens_zero_kernel, dim3(1), dim3(1), 0, st, e.d_queue.as<int>());
        return regf_items_launch(e.js, e.npad, e.js.perm.as<int>(), stride, d_i,
                                 e.d_items.as<int4>(), e.nitems, e.d_queue.as<int>(),
                                 e.d_jbase.as<int>(), e.d_padoff.as<int>(), e.nc, e.maxseg, lmax,
                                 nnbmax, m_flag, d_out, d_list, grid, e.work, st);
    };
    if (!use_graph) return step(s);
    // The graph holds kernel arguments by value: pointers, sizes, the launch shape and
    // (for m_flag) the maximum mass of the j-set.  Any change there needs a new capture.
    GraphKey key;
    key.ptr[0] = d_i; key.ptr[1] = d_out; key.ptr[2] = d_list; key.ptr[3] = e.d_items.p;
    long long mbits = 0;
    if (m_flag) std::memcpy(&mbits, &e.js.mmax, sizeof(mbits));
    int nbits = 0, kb[4] = {0, 0, 0, 0};
    std::memcpy(&nbits, &cfg.near, sizeof(nbits));
    std::memcpy(kb, cfg.kref_ens, sizeof(kb));
    const long long v[18] = {stride, lmax, nnbmax, m_flag, cfg.threads, cfg.ipl, e.nitems, grid,
                             e.ntot, e.npad, mbits, cfg.far_test, cfg.refine, nbits,
                             kb[0], kb[1], kb[2], kb[3]};
    std::copy(v, v + 18, key.val);
    const bool same = e.exec && key == e.gkey;
    if (!same) {
        drop_graph();
        // work buffers are sized (and allocated) by one eager step first
        int rc = step(s);
        if (rc != OK) return rc;
        NB6_HIP(hipStreamSynchronize(s));
        NB6_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = step(s);
        NB6_HIP(hipStreamEndCapture(s, &e.graph));
        if (rc != OK) return rc;
        NB6_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        e.gkey = key;
        return OK;   // the eager step already produced this call's results
    }
    NB6_HIP(hipGraphLaunch(e.exec, s));
    return OK;
}

int ens_items() { return g_ens.nitems; }

// Sweep work-group size (= LDS j-tile) 128/256/512, i-particles per lane 1/2 (i-block
// 128..1024) and persistent work-groups per CU on the sent ensemble; keep the fastest.
int ens_autotune(const double *d_i, int stride, int lmax, int nnbmax, int m_flag, double *d_out,
                 int *d_list, int *best, hipStream_t s) {
    Ens &e = g_ens;
    if (!e.open || e.nc == 0) { set_error(E_STATE, "nb6_ens_autotune: nothing sent"); return E_STATE; }
    RegfConfig &cfg = regf_config();
    const RegfConfig keep = cfg;
    hipEvent_t a, b;
    NB6_HIP(hipEventCreate(&a));
    NB6_HIP(hipEventCreate(&b));
    float best_ms = 1e30f;
    int bt = 256, bipl = 1, bcu = 0;
    const int threads[3] = {128, 256, 512};
    for (int ti = 0; ti < 3; ti++)
        for (int ipl = 1; ipl <= 2; ipl++) {
            const int waves_per_group = threads[ti] / 64;
            for (int wpc : {8, 16, 32}) {        // wavefronts per CU
                const int pcu = std::max(1, wpc / waves_per_group);
                cfg.threads = threads[ti];
                cfg.ipl = ipl;
                e.per_cu = pcu;
                int rc = ens_regf_device(d_i, stride, lmax, nnbmax, m_flag, d_out, d_list, 0, s);
                if (rc != OK) { cfg = keep; e.per_cu = 0; return rc; }
                NB6_HIP(hipEventRecord(a, s));
                rc = ens_regf_device(d_i, stride, lmax, nnbmax, m_flag, d_out, d_list, 0, s);
                NB6_HIP(hipEventRecord(b, s));
                NB6_HIP(hipEventSynchronize(b));
                if (rc != OK) { cfg = keep; e.per_cu = 0; return rc; }
                float ms = 0.f;
                NB6_HIP(hipEventElapsedTime(&ms, a, b));
                if (ms < best_ms) { best_ms = ms; bt = threads[ti]; bipl = ipl; bcu = pcu; }
            }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    cfg.threads = bt;
    cfg.ipl = bipl;
    e.per_cu = bcu;
    drop_graph();
    const int rc = build_items(e, s);
    if (best) { best[0] = bt; best[1] = bipl; best[2] = bcu; best[3] = (int)(best_ms * 1000.f); }
    return rc;
}

}  // namespace nb6
