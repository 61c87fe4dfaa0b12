// Irregular (neighbour-list) force for gfx950, behind irr_simd_*_ / gpuirr_*_.
//
// Semantics follow irr.avx.cpp:
//   predictor  s = ti - t0;  x += s(v + s(a/2 + s jdot/6));  v += 2s(a/2 + 1.5 s jdot/6)
//                                                   irr.avx.cpp:142-156, xbpredall.F:55-59
//   pair terms acc += m r^-3 dx;  jrk += m r^-3 (dv - 3 (dx.dv) r^-2 dx)
//                                                   irr.avx.cpp:303-332, nbint.F:36-57
//   nnbid      1-based index of the nearest list member (ties: lower index)
//                                                   irr.avx.cpp:271-283,293
//
// Method (own): device-resident particle records, 64 B each
//   { double x, y, z;  float m, vx, vy, vz, ax2, ay2, az2, jx6, jy6, jz6 }
// plus a separate FP64 time array, and device-resident lists [nmax][lmax]
// (row[0] = nnb, 0-based entries).  set_jp / set_list only append to host logs;
// the logs are uploaded and scattered by one kernel each at the next firr call.
// firr: one wavefront per active particle, lanes over its list members; the
// separation is formed in FP64 and the predictor increments in FP32; FP32 lane
// sums are reduced across the wavefront in FP64; nearest neighbour by a 64-bit
// (r^2 bits, index) minimum.
#include "nb6_internal.h"

#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cstring>
#include <vector>

namespace nb6 {

namespace {

struct alignas(16) IrrRec {
    double x, y, z;
    float m, vx, vy, vz;
    float ax, ay, az, jx;
    float jy, jz;
};
static_assert(sizeof(IrrRec) == 64, "record must be 64 bytes");

struct IrrState {
    bool open = false;
    int nmax = 0, lmax = 0;
    DBuf rec, tim, lists;          // device
    DBuf pre;                      // [nmax + 1] records predicted to the current call's time
    // host logs
    std::vector<int> jp_pos;       // addr -> slot in jp log (or -1)
    std::vector<int> ls_pos;       // addr -> slot in list log (or -1)
    HBuf jp_log;                   // records: IrrRec + double t0 (72 B) + int addr
    int jp_n = 0;
    HBuf ls_log;                   // [addr, nnb, entries...] variable length
    std::vector<size_t> ls_off;    // offsets (ints) of live entries in ls_log
    size_t ls_used = 0;            // ints used
    DBuf d_jp, d_ls, d_lsoff;
    DBuf d_addr, d_out, d_bad, d_okey, d_oval, d_otmp;
    HBuf h_addr, h_out;
    // profile
    double time_grav = 0.0;
    unsigned long long num_inter = 0, num_fcall = 0, num_steps = 0;
} g_irr;

struct JpEntry {
    IrrRec r;
    double t0;
    int addr;
    int pad;
};
static_assert(sizeof(JpEntry) == 80, "jp entry layout");

__global__ void irr_scatter_jp(const JpEntry *__restrict__ e, int n, IrrRec *__restrict__ rec,
                               double *__restrict__ tim) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = e[k].addr;
    rec[a] = e[k].r;
    tim[a] = e[k].t0;
}

// one wavefront per list entry: entry = [addr, nnb, e_1 .. e_nnb] (e 1-based)
__global__ void irr_scatter_list(const int *__restrict__ log, const long long *__restrict__ off,
                                 int n, int *__restrict__ lists, int lmax) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= n) return;
    const int *e = log + off[w];
    const int a = e[0], nnb = e[1];
    int *row = lists + (size_t)a * lmax;
    if (lane == 0) row[0] = nnb;
    for (int k = lane; k < nnb; k += 64) row[1 + k] = e[2 + k] - 1;
}

// regf rows -> irregular-force lists, one wavefront per row: row i holds the
// gpunb_regf list (count, 0-based j of the sent j-set, the particle itself
// included) of the particle at 1-based address addr[i]; j becomes the 0-based
// address j + jbase - 1 (util_gpu.F:101-110 adds IFIRST and drops self).  Rows
// with a negative (overflow) count leave the particle's list unchanged.
__global__ __launch_bounds__(256) void irr_lists_from_regf(int ni, const int *__restrict__ addr,
                                                           const int *__restrict__ rl, int lmax_r,
                                                           int jbase, int *__restrict__ lists,
                                                           int lmax, int nmax,
                                                           int *__restrict__ bad) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= ni) return;
    const int a = addr[w] - 1;
    if (a < 0 || a >= nmax) {
        if (lane == 0) atomicAdd(bad, 1);
        return;
    }
    const int *src = rl + (size_t)w * lmax_r;
    const int c = min(src[0], lmax_r - 1);
    if (c < 0) return;
    int *row = lists + (size_t)a * lmax;
    int n = 0;
    for (int kb = 0; kb < c; kb += 64) {
        const int k = kb + lane;
        int e = -1;
        if (k < c) e = src[1 + k] + jbase - 1;
        const bool keep = k < c && e != a && e >= 0 && e < nmax;
        const unsigned long long b = __ballot(keep);
        const int pos = n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
        if (keep && pos < lmax - 1) row[1 + pos] = e;
        n += __popcll(b);
    }
    if (lane == 0) {
        row[0] = min(n, lmax - 1);
        if (n > lmax - 1) atomicAdd(bad, 1);
    }
}

// predictor records from the device predictor store (SoA [14][jmax]: x0[3], x0dot[3],
// F = a/2 [3], FDOT = adot/6 [3], t0, m) for addresses [0, n)
__global__ __launch_bounds__(256) void irr_from_jp(int n, const double *__restrict__ st, int jmax,
                                                   IrrRec *__restrict__ rec,
                                                   double *__restrict__ tim) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= n) return;
    const size_t M = jmax;
    IrrRec r;
    r.x = st[a]; r.y = st[M + a]; r.z = st[2 * M + a];
    r.vx = (float)st[3 * M + a]; r.vy = (float)st[4 * M + a]; r.vz = (float)st[5 * M + a];
    r.ax = (float)st[6 * M + a]; r.ay = (float)st[7 * M + a]; r.az = (float)st[8 * M + a];
    r.jx = (float)st[9 * M + a]; r.jy = (float)st[10 * M + a]; r.jz = (float)st[11 * M + a];
    r.m = (float)st[13 * M + a];
    rec[a] = r;
    tim[a] = st[12 * M + a];
}

struct Pred {
    double x, y, z;
    float dx, dy, dz;     // FP32 predictor increments
    float vx, vy, vz;
    float m;
};

// a Pred stored in one 64-byte line
struct alignas(16) PredRec {
    double x, y, z;
    float dx, dy, dz, vx, vy, vz, m, pad;
};
static_assert(sizeof(PredRec) == 64, "predicted record must be 64 bytes");

__device__ __forceinline__ Pred predict(const IrrRec *__restrict__ rec,
                                        const double *__restrict__ tim, int a, double ti) {
    // explicit fmaf and no contraction: the same bits whether the prediction is
    // inlined into the gather or stored first (irr_predict_all)
#pragma clang fp contract(off)
    const uint4 *p = reinterpret_cast<const uint4 *>(rec + a);
    const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    IrrRec r;
    uint4 *rr = reinterpret_cast<uint4 *>(&r);
    rr[0] = q0; rr[1] = q1; rr[2] = q2; rr[3] = q3;
    const float s = (float)(ti - tim[a]);
    const float s1 = s + s, s2 = 1.5f * s;
    Pred o;
    o.x = r.x; o.y = r.y; o.z = r.z;
    o.dx = s * fmaf(s, fmaf(s, r.jx, r.ax), r.vx);
    o.dy = s * fmaf(s, fmaf(s, r.jy, r.ay), r.vy);
    o.dz = s * fmaf(s, fmaf(s, r.jz, r.az), r.vz);
    o.vx = fmaf(s1, fmaf(s2, r.jx, r.ax), r.vx);
    o.vy = fmaf(s1, fmaf(s2, r.jy, r.ay), r.vy);
    o.vz = fmaf(s1, fmaf(s2, r.jz, r.az), r.vz);
    o.m = r.m;
    return o;
}

__device__ __forceinline__ Pred load_pred(const PredRec *__restrict__ pre, int a) {
    const uint4 *p = reinterpret_cast<const uint4 *>(pre + a);
    const uint4 q0 = p[0], q1 = p[1], q2 = p[2], q3 = p[3];
    PredRec r;
    uint4 *rr = reinterpret_cast<uint4 *>(&r);
    rr[0] = q0; rr[1] = q1; rr[2] = q2; rr[3] = q3;
    Pred o;
    o.x = r.x; o.y = r.y; o.z = r.z;
    o.dx = r.dx; o.dy = r.dy; o.dz = r.dz;
    o.vx = r.vx; o.vy = r.vy; o.vz = r.vz;
    o.m = r.m;
    return o;
}

// every record predicted to ti once (the same predict() as the per-neighbour path), so
// that each neighbour costs one 64-byte line instead of a record and a time
__global__ __launch_bounds__(256) void irr_predict_all(double ti, int n, const IrrRec *__restrict__ rec,
                                                       const double *__restrict__ tim,
                                                       PredRec *__restrict__ pre) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= n) return;
    const Pred p = predict(rec, tim, a, ti);
    PredRec r;
    r.x = p.x; r.y = p.y; r.z = p.z;
    r.dx = p.dx; r.dy = p.dy; r.dz = p.dz;
    r.vx = p.vx; r.vy = p.vy; r.vz = p.vz;
    r.m = p.m;
    r.pad = 0.f;
    const uint4 *rr = reinterpret_cast<const uint4 *>(&r);
    uint4 *q = reinterpret_cast<uint4 *>(pre + a);
    q[0] = rr[0]; q[1] = rr[1]; q[2] = rr[2]; q[3] = rr[3];
}

// Launch shape of the irregular-force gather: LPI lanes per active i (16: 4 i per
// wavefront; 64: one i per wavefront), each lane keeps two neighbour gathers in
// flight; the i are taken in the order of `order` (Morton order of their positions
// when sorting is on) and consecutive work-groups go to the same XCD, so one XCD's
// L2 serves one compact region of the cluster.
struct IrrConfig {
    int lpi = 64;
    int sort_i = 1;
    int pre = 1;          // predict all records first: 0 never, 1 for large calls, 2 always
} g_irrcfg;

__device__ __forceinline__ void irr_pair(const Pred &pi, const Pred &pj, int j, double &ax,
                                         double &ay, double &az, double &jx, double &jy, double &jz,
                                         unsigned long long &best) {
    const float dx = (float)(pj.x - pi.x) + (pj.dx - pi.dx);
    const float dy = (float)(pj.y - pi.y) + (pj.dy - pi.dy);
    const float dz = (float)(pj.z - pi.z) + (pj.dz - pi.dz);
    const float du = pj.vx - pi.vx, dv = pj.vy - pi.vy, dw = pj.vz - pi.vz;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float rv = fmaf(dz, dw, fmaf(dy, dv, dx * du));
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    const float mr3 = pj.m * rinv * rinv2;
    const float al = -3.0f * rv * rinv2;
    // pair terms in FP32, the lane's running sums in FP64
    ax += (double)(mr3 * dx);
    ay += (double)(mr3 * dy);
    az += (double)(mr3 * dz);
    jx += (double)(mr3 * fmaf(al, dx, du));
    jy += (double)(mr3 * fmaf(al, dy, dv));
    jz += (double)(mr3 * fmaf(al, dz, dw));
    const unsigned long long key = ((unsigned long long)__float_as_uint(r2) << 32) | (unsigned)j;
    best = key < best ? key : best;
}

template <int LPI>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
    for (int o = LPI / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int LPI, bool PRE>
__global__ __launch_bounds__(256) void irr_force_kernel(
    double ti, int ni, const int *__restrict__ addr /* in processing order */,
    const int *__restrict__ order /* processing slot -> output row; nullptr: the same */,
    const IrrRec *__restrict__ rec, const double *__restrict__ tim,
    const PredRec *__restrict__ pre, const int *__restrict__ lists, int lmax,
    double *__restrict__ out /* [ni][8]: acc3, jrk3, nnbid, nnb */) {
    const int blk = xcd_order(blockIdx.x, gridDim.x);
    const int q = blk * (256 / LPI) + (threadIdx.x / LPI);
    const int lane = threadIdx.x % LPI;
    // the whole group leaves together; the shuffles below stay within a group
    if (q >= ni) return;
    const int w = order ? order[q] : q;
    const int ia = addr[q] - 1;
    auto get = [&](int a) { return PRE ? load_pred(pre, a) : predict(rec, tim, a, ti); };
    const Pred pi = get(ia);
    const int *row = lists + (size_t)ia * lmax;
    // the lane's first two entries are read with the count, and each loop trip reads the
    // next two while the current records are in flight (entries past the count are
    // never used)
    int k = lane;
    int j0 = 1 + k < lmax ? row[1 + k] : 0;
    int j1 = 1 + k + LPI < lmax ? row[1 + k + LPI] : 0;
    const int nnb = row[0];
    double ax = 0.0, ay = 0.0, az = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    unsigned long long best = ~0ull;
    for (; k + LPI < nnb; k += 2 * LPI) {
        const Pred p0 = get(j0);
        const Pred p1 = get(j1);
        const int kn = k + 2 * LPI;
        const int n0 = 1 + kn < lmax ? row[1 + kn] : 0;
        const int n1 = 1 + kn + LPI < lmax ? row[1 + kn + LPI] : 0;
        irr_pair(pi, p0, j0, ax, ay, az, jx, jy, jz, best);
        irr_pair(pi, p1, j1, ax, ay, az, jx, jy, jz, best);
        j0 = n0;
        j1 = n1;
    }
    if (k < nnb) irr_pair(pi, get(j0), j0, ax, ay, az, jx, jy, jz, best);
    const double sax = gsum<LPI>(ax), say = gsum<LPI>(ay), saz = gsum<LPI>(az);
    const double sjx = gsum<LPI>(jx), sjy = gsum<LPI>(jy), sjz = gsum<LPI>(jz);
#pragma unroll
    for (int o = LPI / 2; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(best, o, 64);
        best = other < best ? other : best;
    }
    if (lane == 0) {
        double *o = out + (size_t)w * 8;
        o[0] = sax; o[1] = say; o[2] = saz;
        o[3] = sjx; o[4] = sjy; o[5] = sjz;
        o[6] = nnb > 0 ? (double)((unsigned)(best & 0xffffffffu) + 1u) : 0.0;
        o[7] = (double)nnb;
    }
}

// Morton key of each active i's record position (for the processing order)
__global__ __launch_bounds__(256) void irr_key_kernel(int ni, const int *__restrict__ addr,
                                                      const IrrRec *__restrict__ rec, double lx,
                                                      double ly, double lz, double sx, double sy,
                                                      double sz, unsigned *__restrict__ key,
                                                      int *__restrict__ val) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= ni) return;
    const IrrRec &r = rec[addr[w] - 1];
    const double f[3] = {(r.x - lx) * sx, (r.y - ly) * sy, (r.z - lz) * sz};
    unsigned q[3];
    for (int d = 0; d < 3; d++) q[d] = f[d] > 0.0 ? (f[d] < 1023.0 ? (unsigned)f[d] : 1023u) : 0u;
    unsigned kk = 0;
    for (int bit = 9; bit >= 0; bit--)
        for (int d = 2; d >= 0; d--) kk = (kk << 1) | ((q[d] >> bit) & 1u);
    key[w] = kk;
    val[w] = w;
}

void launch_irr(double ti, int ni, const int *d_addr, const int *order, const IrrState &g,
                double *d_out, hipStream_t s) {
    const int lpi = g_irrcfg.lpi;
    const dim3 grid((ni + 256 / lpi - 1) / (256 / lpi));
    // one prediction of every record (a fixed cost of ~0.14 KB per particle) against a
    // gather that is 20 % faster with one line per neighbour: measured even at
    // ni = nmax / 8 with ~100 neighbours, a gain above
    const bool pre = g_irrcfg.pre == 2 || (g_irrcfg.pre == 1 && (size_t)ni * 8 >= (size_t)g.nmax);
    const PredRec *pr = nullptr;
    if (pre) {
        PredRec *p = static_cast<PredRec *>(const_cast<DBuf &>(g.pre).get((size_t)(g.nmax + 1) * sizeof(PredRec)));
        hipLaunchKernelGGL(irr_predict_all, dim3((g.nmax + 1 + 255) / 256), dim3(256), 0, s, ti,
                           g.nmax + 1, g.rec.as<IrrRec>(), g.tim.as<double>(), p);
        NB6_HIP(hipGetLastError());
        pr = p;
    }
    const IrrRec *rc = g.rec.as<IrrRec>();
    const double *tm = g.tim.as<double>();
    const int *ls = g.lists.as<int>();
#define NB6_IRR_LAUNCH(L, P) \
    hipLaunchKernelGGL((irr_force_kernel<L, P>), grid, dim3(256), 0, s, ti, ni, d_addr, order, rc, \
                       tm, pr, ls, g.lmax, d_out)
    if (lpi == 16) {
        if (pre) NB6_IRR_LAUNCH(16, true); else NB6_IRR_LAUNCH(16, false);
    } else {
        if (pre) NB6_IRR_LAUNCH(64, true); else NB6_IRR_LAUNCH(64, false);
    }
#undef NB6_IRR_LAUNCH
    NB6_HIP(hipGetLastError());
}

// processing order of the active i: Morton order of their record positions in the
// box of the gpunb j-set (nullptr: as given, when sorting is off or no box is known);
// *oaddr receives the addresses in processing order
const int *irr_order(int ni, const int *d_addr, const int **oaddr, hipStream_t s);

void flush_logs(hipStream_t s) {
    IrrState &g = g_irr;
    if (g.jp_n > 0) {
        const size_t bytes = (size_t)g.jp_n * sizeof(JpEntry);
        void *d = g.d_jp.get(bytes);
        NB6_HIP(hipMemcpyAsync(d, g.jp_log.p, bytes, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(irr_scatter_jp, dim3((g.jp_n + 255) / 256), dim3(256), 0, s,
                           static_cast<const JpEntry *>(d), g.jp_n, g.rec.as<IrrRec>(),
                           g.tim.as<double>());
        NB6_HIP(hipGetLastError());
        const JpEntry *e = g.jp_log.as<JpEntry>();
        for (int k = 0; k < g.jp_n; k++) g.jp_pos[e[k].addr] = -1;
        g.jp_n = 0;
    }
    if (!g.ls_off.empty()) {
        const size_t bytes = g.ls_used * sizeof(int);
        void *d = g.d_ls.get(bytes);
        NB6_HIP(hipMemcpyAsync(d, g.ls_log.p, bytes, hipMemcpyHostToDevice, s));
        // live entries only (a repeated set_list of one particle keeps the last)
        std::vector<long long> live;
        live.reserve(g.ls_off.size());
        const int *lg = g.ls_log.as<int>();
        for (size_t k = 0; k < g.ls_off.size(); k++) {
            const int a = lg[g.ls_off[k]];
            if (a >= 0 && g.ls_pos[a] == (int)k) live.push_back((long long)g.ls_off[k]);
        }
        const int n = (int)live.size();
        HBuf &hb = g.h_addr;   // reuse as staging for offsets
        long long *ho = static_cast<long long *>(hb.get(live.size() * sizeof(long long)));
        std::memcpy(ho, live.data(), live.size() * sizeof(long long));
        void *doff = g.d_lsoff.get(live.size() * sizeof(long long));
        NB6_HIP(hipMemcpyAsync(doff, ho, live.size() * sizeof(long long), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(irr_scatter_list, dim3((n + 3) / 4), dim3(256), 0, s,
                           static_cast<const int *>(d), static_cast<const long long *>(doff), n,
                           g.lists.as<int>(), g.lmax);
        NB6_HIP(hipGetLastError());
        NB6_HIP(hipStreamSynchronize(s));   // staging buffers are reused next
        for (size_t k = 0; k < g.ls_off.size(); k++) {
            const int a = lg[g.ls_off[k]];
            if (a >= 0) g.ls_pos[a] = -1;
        }
        g.ls_off.clear();
        g.ls_used = 0;
    }
    NB6_HIP(hipStreamSynchronize(s));   // the host logs are rewritten after this
}

}  // namespace

namespace {
__global__ __launch_bounds__(256) void irr_oaddr_kernel(int ni, const int *__restrict__ addr,
                                                        const int *__restrict__ order,
                                                        int *__restrict__ oaddr) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < ni) oaddr[q] = addr[order[q]];
}

const int *irr_order(int ni, const int *d_addr, const int **oaddr, hipStream_t s) {
    IrrState &g = g_irr;
    const JSet *js = gpunb_jset();
    *oaddr = d_addr;
    if (!g_irrcfg.sort_i || ni < 2048 || !js || js->n <= 0) return nullptr;
    double sc[3];
    for (int d = 0; d < 3; d++) {
        const double w = js->box[3 + d] - js->box[d];
        sc[d] = w > 0.0 ? 1024.0 / (w * (1.0 + 1e-12)) : 0.0;
    }
    unsigned *k0 = static_cast<unsigned *>(g.d_okey.get((size_t)ni * 2 * sizeof(unsigned)));
    unsigned *k1 = k0 + ni;
    int *v0 = static_cast<int *>(g.d_oval.get((size_t)ni * 2 * sizeof(int)));
    int *v1 = v0 + ni;
    hipLaunchKernelGGL(irr_key_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, ni, d_addr,
                       g.rec.as<IrrRec>(), js->box[0], js->box[1], js->box[2], sc[0], sc[1], sc[2],
                       k0, v0);
    NB6_HIP(hipGetLastError());
    size_t tb = 0;
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, ni, 0, 30, s));
    void *tmp = g.d_otmp.get(tb);
    NB6_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, ni, 0, 30, s));
    int *oa = reinterpret_cast<int *>(k0);     // the unsorted keys are no longer needed
    hipLaunchKernelGGL(irr_oaddr_kernel, dim3((ni + 255) / 256), dim3(256), 0, s, ni, d_addr, v1,
                       oa);
    NB6_HIP(hipGetLastError());
    *oaddr = oa;
    return v1;
}
}  // namespace

int irr_predict_mode(int mode) {
    if (mode < 0 || mode > 2) {
        set_error(E_ARG, "nb6_set_irr_predict: mode must be 0, 1 or 2");
        return E_ARG;
    }
    g_irrcfg.pre = mode;
    return OK;
}

int irr_config(int lanes_per_i, int sort_i) {
    if ((lanes_per_i != 16 && lanes_per_i != 64) || (sort_i != 0 && sort_i != 1)) {
        set_error(E_ARG, "nb6_set_irr_config: lanes_per_i must be 16 or 64 and sort_i 0 or 1");
        return E_ARG;
    }
    g_irrcfg.lpi = lanes_per_i;
    g_irrcfg.sort_i = sort_i;
    return OK;
}

void irr_release() {
    IrrState &g = g_irr;
    g.rec.release(); g.tim.release(); g.lists.release(); g.pre.release();
    g.d_jp.release(); g.d_ls.release(); g.d_lsoff.release();
    g.d_addr.release(); g.d_out.release(); g.d_bad.release();
    g.d_okey.release(); g.d_oval.release(); g.d_otmp.release();
    g.jp_log.release(); g.ls_log.release(); g.h_addr.release(); g.h_out.release();
    g.jp_pos.clear(); g.ls_pos.clear(); g.ls_off.clear();
    g.jp_n = 0; g.ls_used = 0;
    g.open = false;
}

// ------------------------------------------------------------------ host API
int irr_open(int nmax, int lmax, int rank) {
    IrrState &g = g_irr;
    if (g.open) {
        fprintf(stderr, "irr_simd: it is already open\n");
        return OK;
    }
    if (nmax <= 0 || lmax < 1) {
        set_error(E_ARG, "irr_simd_open: nmax and lmax must be positive");
        return E_ARG;
    }
    Device &dev = device(rank);
    if (!dev.ready) return E_NODEV;
    g.nmax = nmax; g.lmax = lmax;
    g.rec.get((size_t)(nmax + 1) * sizeof(IrrRec));
    g.tim.get((size_t)(nmax + 1) * sizeof(double));
    g.lists.get((size_t)(nmax + 1) * lmax * sizeof(int));
    NB6_HIP(hipMemsetAsync(g.rec.p, 0, (size_t)(nmax + 1) * sizeof(IrrRec), dev.stream));
    NB6_HIP(hipMemsetAsync(g.tim.p, 0, (size_t)(nmax + 1) * sizeof(double), dev.stream));
    NB6_HIP(hipMemsetAsync(g.lists.p, 0, (size_t)(nmax + 1) * lmax * sizeof(int), dev.stream));
    NB6_HIP(hipStreamSynchronize(dev.stream));
    g.jp_pos.assign(nmax + 1, -1);
    g.ls_pos.assign(nmax + 1, -1);
    g.jp_n = 0; g.ls_used = 0; g.ls_off.clear();
    g.time_grav = 0.0; g.num_inter = g.num_fcall = g.num_steps = 0;
    g.open = true;
    fprintf(stderr, "# Opening IRR lib. gfx950 ver. - rank: %d; nmax: %d, lmax: %d\n", rank, nmax,
            lmax);
    return OK;
}

int irr_close(int rank) {
    IrrState &g = g_irr;
    if (!g.open) {
        fprintf(stderr, "irr_simd: it is already close\n");
        return OK;
    }
    irr_release();
    fprintf(stderr, "Closing IRR lib. gfx950 ver. - rank: %d\n", rank);
    return OK;
}

int irr_set_jp(int addr1, const double *pos, const double *vel, const double *acc2,
               const double *jrk6, double mass, double time) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "irr_simd_set_jp: library not open"); return E_STATE; }
    const int a = addr1 - 1;
    if (a < 0 || a >= g.nmax) { set_error(E_ARG, "irr_simd_set_jp: address out of range"); return E_ARG; }
    int slot = g.jp_pos[a];
    if (slot < 0) {
        slot = g.jp_n++;
        JpEntry *base = static_cast<JpEntry *>(g.jp_log.p);
        if ((size_t)g.jp_n * sizeof(JpEntry) > g.jp_log.bytes) {
            // grow, keeping the entries already logged
            const size_t nb = std::max<size_t>((size_t)g.jp_n * sizeof(JpEntry) * 2, 4096);
            HBuf nbuf;
            nbuf.get(nb);
            if (base) std::memcpy(nbuf.p, base, (size_t)(g.jp_n - 1) * sizeof(JpEntry));
            g.jp_log.release();
            g.jp_log = nbuf;
            nbuf.p = nullptr;
        }
        g.jp_pos[a] = slot;
    }
    JpEntry &e = g.jp_log.as<JpEntry>()[slot];
    e.r.x = pos[0]; e.r.y = pos[1]; e.r.z = pos[2];
    e.r.m = (float)mass;
    e.r.vx = (float)vel[0]; e.r.vy = (float)vel[1]; e.r.vz = (float)vel[2];
    e.r.ax = (float)acc2[0]; e.r.ay = (float)acc2[1]; e.r.az = (float)acc2[2];
    e.r.jx = (float)jrk6[0]; e.r.jy = (float)jrk6[1]; e.r.jz = (float)jrk6[2];
    e.t0 = time;
    e.addr = a;
    e.pad = 0;
    return OK;
}

int irr_set_list(int addr1, const int *nblist) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "irr_simd_set_list: library not open"); return E_STATE; }
    const int a = addr1 - 1;
    const int nnb = nblist[0];
    if (a < 0 || a >= g.nmax) { set_error(E_ARG, "irr_simd_set_list: address out of range"); return E_ARG; }
    if (nnb < 0 || nnb > g.lmax - 1) {
        set_error(E_ARG, "irr_simd_set_list: neighbour count outside [0, lmax-1]");
        return E_ARG;
    }
    for (int k = 1; k <= nnb; k++) {
        if (nblist[k] < 1 || nblist[k] > g.nmax) {
            set_error(E_ARG, "irr_simd_set_list: neighbour address out of range");
            return E_ARG;
        }
    }
    const size_t need = g.ls_used + 2 + (size_t)nnb;
    if (need * sizeof(int) > g.ls_log.bytes) {
        HBuf nbuf;
        nbuf.get(std::max<size_t>(need * sizeof(int) * 2, 1 << 16));
        if (g.ls_log.p) std::memcpy(nbuf.p, g.ls_log.p, g.ls_used * sizeof(int));
        g.ls_log.release();
        g.ls_log = nbuf;
        nbuf.p = nullptr;
    }
    int *e = g.ls_log.as<int>() + g.ls_used;
    e[0] = a;
    e[1] = nnb;
    std::memcpy(e + 2, nblist + 1, (size_t)nnb * sizeof(int));
    g.ls_pos[a] = (int)g.ls_off.size();
    g.ls_off.push_back(g.ls_used);
    g.ls_used = need;
    return OK;
}

int irr_firr_vec(double ti, int ni, const int *addr, double (*acc)[3], double (*jrk)[3],
                 int *nnbid) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "irr_simd_firr_vec: library not open"); return E_STATE; }
    if (ni <= 0) return OK;
    for (int k = 0; k < ni; k++) {
        if (addr[k] < 1 || addr[k] > g.nmax) {
            set_error(E_ARG, "irr_simd_firr_vec: address out of range");
            return E_ARG;
        }
    }
    // wall time of the whole call, as the reference profile counts it (irr.avx.cpp:551,561-562)
    const auto t0 = std::chrono::steady_clock::now();
    Device &dev = device();
    hipStream_t s = dev.stream;
    flush_logs(s);
    int *ha = static_cast<int *>(g.h_addr.get((size_t)ni * sizeof(int)));
    std::memcpy(ha, addr, (size_t)ni * sizeof(int));
    int *da = static_cast<int *>(g.d_addr.get((size_t)ni * sizeof(int)));
    double *dout = static_cast<double *>(g.d_out.get((size_t)ni * 8 * sizeof(double)));
    double *hout = static_cast<double *>(g.h_out.get((size_t)ni * 8 * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(da, ha, (size_t)ni * sizeof(int), hipMemcpyHostToDevice, s));
    const int *oaddr = nullptr;
    const int *order = irr_order(ni, da, &oaddr, s);
    KTimer &t = ktimer("irr");
    t.start(s);
    launch_irr(ti, ni, oaddr, order, g, dout, s);
    t.stop(s);
    NB6_HIP(hipMemcpyAsync(hout, dout, (size_t)ni * 8 * sizeof(double), hipMemcpyDeviceToHost, s));
    NB6_HIP(hipStreamSynchronize(s));
    unsigned long long ninter = 0;
    for (int k = 0; k < ni; k++) {
        const double *o = hout + (size_t)k * 8;
        acc[k][0] = o[0]; acc[k][1] = o[1]; acc[k][2] = o[2];
        jrk[k][0] = o[3]; jrk[k][1] = o[4]; jrk[k][2] = o[5];
        nnbid[k] = (int)o[6];
        ninter += (unsigned long long)o[7];
    }
    g.num_inter += ninter;
    g.time_grav += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    g.num_fcall++;
    g.num_steps += ni;
    return OK;
}

// lists of ni particles straight from gpunb_regf rows on the device (no host round trip)
int irr_set_lists_device(int ni, const int *d_addr, const int *d_list, int lmax_r, int jbase,
                         hipStream_t s) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "nb6_irr_set_lists_device: irr library not open"); return E_STATE; }
    if (ni <= 0) return OK;
    if (lmax_r < 1) { set_error(E_ARG, "nb6_irr_set_lists_device: lmax must be >= 1"); return E_ARG; }
    flush_logs(s);   // earlier host set_list calls land first
    int *bad = static_cast<int *>(g.d_bad.get(sizeof(int)));
    NB6_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(irr_lists_from_regf, dim3((ni + 3) / 4), dim3(256), 0, s, ni, d_addr, d_list,
                       lmax_r, jbase, g.lists.as<int>(), g.lmax, g.nmax, bad);
    NB6_HIP(hipGetLastError());
    int hb = 0;
    NB6_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    NB6_HIP(hipStreamSynchronize(s));
    if (hb) {
        set_error(E_ARG, "nb6_irr_set_lists_device: address out of range or list longer than lmax-1");
        return E_ARG;
    }
    return OK;
}

// predictor records of addresses [0, n) from the device predictor store
int irr_sync_jp(int n, hipStream_t s) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "nb6_irr_sync_jp: irr library not open"); return E_STATE; }
    int jmax = 0;
    const double *st = jp_state(&jmax);
    if (!st) { set_error(E_STATE, "nb6_irr_sync_jp: predictor store not open"); return E_STATE; }
    if (n < 0 || n > g.nmax || n > jmax) { set_error(E_ARG, "nb6_irr_sync_jp: n outside both stores"); return E_ARG; }
    flush_logs(s);
    if (n > 0) {
        hipLaunchKernelGGL(irr_from_jp, dim3((n + 255) / 256), dim3(256), 0, s, n, st, jmax,
                           g.rec.as<IrrRec>(), g.tim.as<double>());
        NB6_HIP(hipGetLastError());
    }
    return OK;
}

// irregular force of ni particles (1-based device addresses), results on the device
int irr_firr_device(double ti, int ni, const int *d_addr, double *d_out, hipStream_t s) {
    IrrState &g = g_irr;
    if (!g.open) { set_error(E_STATE, "nb6_irr_firr_device: irr library not open"); return E_STATE; }
    if (ni <= 0) return OK;
    flush_logs(s);
    const int *oaddr = nullptr;
    const int *order = irr_order(ni, d_addr, &oaddr, s);
    KTimer &t = ktimer("irr");
    t.start(s);
    launch_irr(ti, ni, oaddr, order, g, d_out, s);
    t.stop(s);
    // not counted in the profile line: its pair count and time are on the device
    return OK;
}

int irr_profile(int rank) {
    IrrState &g = g_irr;
    if (g.open && g.num_fcall) {
        const double gf = 60.0 * double(g.num_inter) * 1e-9 / (g.time_grav > 0 ? g.time_grav : 1e-30);
        const double us = 1e6 * g.time_grav / double(g.num_fcall);
        const double nb = g.num_steps ? double(g.num_inter) / double(g.num_steps) : 0.0;
        // the reference's <NI> field is neighbour pairs per call, num_inter / num_fcall
        // (irr.avx.cpp:414); kept as it is, since the line is read by people and scripts
        const int niavr = (int)(g.num_inter / g.num_fcall);
        fprintf(stderr,
                "[R.%d GPU Irr.F ] Ncall: %llu <NI>: %d <NB>: %f grav: %f s, %f Gflops, %f usec\n",
                rank, g.num_fcall, niavr, nb, g.time_grav, gf, us);
    }
    g.time_grav = 0.0;
    g.num_inter = g.num_fcall = g.num_steps = 0;
    return OK;
}

}  // namespace nb6
