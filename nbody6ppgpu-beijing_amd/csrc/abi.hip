// Host side of libnb6gpu.so: device context, buffers, error state, and the
// Fortran-callable C ABI (gpunb_*_, gpupot_, irr_simd_*_, gpuirr_*_) plus the
// nb6_* extension entry points.  See include/nb6gpu.h for the contract of each
// symbol and the reference lines it replaces.
#include "nb6_internal.h"
#include "host_threads.h"
#include "../../include/nb6gpu.h"

#include <sys/time.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace nb6 {

// irr.hip
int irr_open(int nmax, int lmax, int rank);
int irr_close(int rank);
int irr_set_jp(int addr1, const double *pos, const double *vel, const double *acc2,
               const double *jrk6, double mass, double time);
int irr_set_list(int addr1, const int *nblist);
int irr_firr_vec(double ti, int ni, const int *addr, double (*acc)[3], double (*jrk)[3],
                 int *nnbid);
int irr_profile(int rank);

// ensemble.hip
int ens_open(int max_clusters, int max_particles);
int ens_close();
int ens_send_device(int nc, const int *offsets, const double *d_soa, int stride, hipStream_t s);
int ens_regf_device(const double *d_i, int stride, int lmax, int nnbmax, int m_flag, double *d_out,
                    int *d_list, int use_graph, hipStream_t s);
int ens_items();

// ------------------------------------------------------------------ errors
static int g_err = 0;
static std::string g_msg;
static int g_abort = -1;   // -1: not yet read from NB6_ABORT_ON_ERROR

static bool abort_enabled() {
    if (g_abort < 0) {
        const char *e = getenv("NB6_ABORT_ON_ERROR");
        g_abort = (e && e[0] == '0') ? 0 : 1;
    }
    return g_abort == 1;
}

void set_error(int code, const std::string &msg) {
    g_err = code;
    g_msg = msg;
    fprintf(stderr, "nb6gpu error %d: %s\n", code, msg.c_str());
    if (abort_enabled()) abort();
}

void hip_fail(hipError_t e, const char *what, const char *file, int line) {
    fprintf(stderr, "nb6gpu: HIP error '%s' in %s at %s:%d\n", hipGetErrorString(e), what, file,
            line);
    abort();
}

static double wtime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// ------------------------------------------------------------------ device
Device &device(int irank) {
    static Device d;
    if (d.ready) return d;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error(E_NODEV, "no HIP device available");
        return d;
    }
    std::vector<int> ids;
    if (const char *gl = getenv("GPU_LIST")) {   // same variable as gpunb.velocity.cu:583
        std::string s(gl);
        size_t p = 0;
        while (p < s.size()) {
            while (p < s.size() && s[p] == ' ') p++;
            size_t q = p;
            while (q < s.size() && s[q] != ' ') q++;
            if (q > p) ids.push_back(atoi(s.substr(p, q - p).c_str()));
            p = q;
        }
    }
    if (ids.empty())
        d.id = irank >= 0 ? irank % n : 0;
    else
        d.id = ids[(irank >= 0 ? irank : 0) % ids.size()];
    if (d.id < 0 || d.id >= n) {
        set_error(E_NODEV, "GPU_LIST names a device that does not exist");
        return d;
    }
    NB6_HIP(hipSetDevice(d.id));
    NB6_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    NB6_HIP(hipGetDeviceProperties(&prop, d.id));
    d.num_cu = prop.multiProcessorCount;
    char host[150] = {0};
    gethostname(host, sizeof(host) - 1);
    fprintf(stderr, "# GPU initialization - rank: %d; HOST %s; NGPU 1; device: %d %s (%s, %d CU)\n",
            irank, host, d.id, prop.name, prop.gcnArchName, d.num_cu);
    d.ready = true;
    return d;
}

void *DBuf::get(size_t need) {
    if (need == 0) need = 1;
    if (need > bytes) {
        if (p) NB6_HIP(hipFree(p));
        size_t nb = need + need / 8;
        NB6_HIP(hipMalloc(&p, nb));
        bytes = nb;
    }
    return p;
}
void *DBuf::try_get(size_t need) {
    if (need == 0) need = 1;
    if (need > bytes) {
        if (p) NB6_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        const size_t nb = need + need / 8;
        if (hipMalloc(&p, nb) != hipSuccess) {
            (void)hipGetLastError();   // clear the sticky error of the failed allocation
            p = nullptr;
            return nullptr;
        }
        bytes = nb;
    }
    return p;
}
void DBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
void *HBuf::get(size_t need) {
    if (need == 0) need = 1;
    if (need > bytes) {
        if (p) NB6_HIP(hipHostFree(p));
        size_t nb = need + need / 8;
        NB6_HIP(hipHostMalloc(&p, nb, hipHostMallocDefault));
        bytes = nb;
    }
    return p;
}
void HBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

// Event pairs around kernels cost the stream an idle gap of several microseconds each
// (measured about 10 us per pair on a small regular-force call), so they are recorded
// only when asked for (nb6_set_kernel_timing) and inside the autotuner.
static bool g_timing = false;
bool kernel_timing() { return g_timing; }
void set_kernel_timing(bool on) { g_timing = on; }
void KTimer::start(hipStream_t s) {
    if (!g_timing) {
        armed = false;
        return;
    }
    if (!a) {
        NB6_HIP(hipEventCreate(&a));
        NB6_HIP(hipEventCreate(&b));
    }
    NB6_HIP(hipEventRecord(a, s));
}
void KTimer::stop(hipStream_t s) {
    if (!g_timing || !a) return;
    NB6_HIP(hipEventRecord(b, s));
    armed = true;
}
double KTimer::ms() {
    if (!armed) return -1.0;
    NB6_HIP(hipEventSynchronize(b));
    float t = 0.f;
    NB6_HIP(hipEventElapsedTime(&t, a, b));
    return t;
}
KTimer &ktimer(const char *name) {
    static std::map<std::string, KTimer> m;
    return m[name];
}

// ------------------------------------------------------------------ gpunb state
struct Gpunb {
    bool open = false;
    int nbmax = 0;
    JSet js;
    HBuf hj, hi, hout, hlist, hoff;
    DBuf di, dout, dlist, dlen, doff, dpack, dtmp;
    double time_send = 0, time_grav = 0, time_reduce = 0, time_out = 0;
    long long num_inter = 0;
    int icall = 0, ini = 0, isend = 0;
} g_nb;

static bool check_finite(const double *a, size_t n) {
    for (size_t k = 0; k < n; k++)
        if (!(a[k] == a[k])) return false;
    return true;
}

const JSet *gpunb_jset() { return g_nb.open ? &g_nb.js : nullptr; }

}  // namespace nb6

using namespace nb6;

extern "C" {

// ================================================================== errors
int nb6_last_error(void) { return g_err; }
const char *nb6_last_error_message(void) { return g_msg.c_str(); }
void nb6_clear_error(void) { g_err = 0; g_msg.clear(); }
void nb6_set_abort_on_error(int on) { g_abort = on ? 1 : 0; }
int nb6_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ================================================================== gpunb
void gpunb_devinit_(int *irank) { (void)device(irank ? *irank : 0); }

void gpunb_open_(int *nbmax, int *irank) {
    Gpunb &g = g_nb;
    g.time_send = g.time_grav = g.time_reduce = g.time_out = 0.0;
    g.num_inter = 0;
    g.icall = g.ini = g.isend = 0;
    Device &d = device(*irank);
    if (!d.ready) return;
    if (g.open) {
        fprintf(stderr, "gpunb: it is already open\n");
        return;
    }
    if (*nbmax <= 0) {
        set_error(E_ARG, "gpunb_open: nbmax must be positive");
        return;
    }
    g.nbmax = *nbmax;
    jset_reserve(g.js, g.nbmax);
    g.js.n = 0;
    // device address of the no-alloc form's work array, with this device current
    if (regf_noalloc_setup() != OK) return;
    g.open = true;
    fprintf(stderr, "# Open GPU regular force - rank: %d; nbmax: %d\n", *irank, *nbmax);
}

void gpunb_close_(void) {
    Gpunb &g = g_nb;
    if (!g.open) {
        fprintf(stderr, "gpunb: it is already close\n");
        return;
    }
    g.open = false;
    jset_release(g.js);
    g.hj.release(); g.hi.release(); g.hout.release(); g.hlist.release(); g.hoff.release();
    g.di.release(); g.dout.release(); g.dlist.release();
    g.dlen.release(); g.doff.release(); g.dpack.release(); g.dtmp.release();
    regf_release();
    g.nbmax = 0;
}

void gpunb_send_(int *nj, double mj[], double xj[][3], double vj[][3]) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "gpunb_send: library not open"); return; }
    const int n = *nj;
    if (n < 0 || n > g.nbmax) { set_error(E_ARG, "gpunb_send: nj outside [0, nbmax]"); return; }
    const double tstart = wtime();
    Device &d = device();
    // AoS records (x y z m vx vy vz 0) in the caller's order, plus box, max mass and a
    // NaN check, in one pass spread over a few threads (one range each)
    double *h = static_cast<double *>(g.hj.get((size_t)8 * std::max(n, 1) * sizeof(double)));
    const int nt = n > 65536 ? std::max(1, std::min(8, (int)std::thread::hardware_concurrency())) : 1;
    double box[6];
    double mmax = 0.0;
    if (!pack_j_records(n, nt, mj, xj, vj, h, box, &mmax)) {
        set_error(E_ARG, "gpunb_send: NaN in j-particle data");
        return;
    }
    g.time_send -= tstart;   // the pack counts as send time
    g.isend++;
    if (n > 0)
        NB6_HIP(hipMemcpyAsync(g.js.in64.p, h, (size_t)8 * n * sizeof(double),
                               hipMemcpyHostToDevice, d.stream));
    jset_build(g.js, n, box, mmax, d.stream);
    NB6_HIP(hipStreamSynchronize(d.stream));
    g.time_send += wtime();
}

static void regf_finish(Gpunb &g, int ni, const double *di, int lmax, int nnbmax, int m_flag,
                        double acc[][3], double jrk[][3], double pot[], int *list, hipStream_t s,
                        double t0);

void gpunb_regf_(int *ni_, double h2[], double dtr[], double xi[][3], double vi[][3],
                 double acc[][3], double jrk[][3], double pot[], int *lmax_, int *nnbmax_,
                 int *list, int *m_flag) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "gpunb_regf: library not open"); return; }
    const int ni = *ni_, lmax = *lmax_, nnbmax = *nnbmax_;
    if (ni < 0) { set_error(E_ARG, "gpunb_regf: ni must not be negative"); return; }
    if (ni == 0) return;   // an empty block (the CUDA back end launches no work either)
    if (lmax < 1) { set_error(E_ARG, "gpunb_regf: lmax must be >= 1"); return; }
    if (!check_finite(h2, ni) || !check_finite(dtr, ni) || !check_finite(&xi[0][0], 3 * (size_t)ni) ||
        !check_finite(&vi[0][0], 3 * (size_t)ni)) {
        set_error(E_ARG, "gpunb_regf: NaN in i-particle data");
        return;
    }
    const double t0 = wtime();
    g.num_inter += (long long)ni * g.js.n;
    g.ini += ni;
    g.icall++;
    Device &d = device();
    hipStream_t s = d.stream;
    double *hi = static_cast<double *>(g.hi.get((size_t)8 * ni * sizeof(double)));
    par_for(ni, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            hi[0 * (size_t)ni + i] = xi[i][0];
            hi[1 * (size_t)ni + i] = xi[i][1];
            hi[2 * (size_t)ni + i] = xi[i][2];
            hi[3 * (size_t)ni + i] = vi[i][0];
            hi[4 * (size_t)ni + i] = vi[i][1];
            hi[5 * (size_t)ni + i] = vi[i][2];
            hi[6 * (size_t)ni + i] = h2[i];
            hi[7 * (size_t)ni + i] = dtr[i];
        }
    });
    double *di = static_cast<double *>(g.di.get((size_t)8 * ni * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(di, hi, (size_t)8 * ni * sizeof(double), hipMemcpyHostToDevice, s));
    regf_finish(g, ni, di, lmax, nnbmax, *m_flag, acc, jrk, pot, list, s, t0);
}

// regf on a staged device i-block, results copied into the caller's arrays
static void regf_finish(Gpunb &g, int ni, const double *di, int lmax, int nnbmax, int m_flag,
                        double acc[][3], double jrk[][3], double pot[], int *list, hipStream_t s,
                        double t0) {
    double *dout = static_cast<double *>(g.dout.get((size_t)7 * ni * sizeof(double)));
    int *dlist = static_cast<int *>(g.dlist.get((size_t)ni * lmax * sizeof(int)));
    if (regf_launch(g.js, ni, ni, di, lmax, nnbmax, m_flag, dout, dlist, s) != OK) return;
    double *ho = static_cast<double *>(g.hout.get((size_t)7 * ni * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(ho, dout, (size_t)7 * ni * sizeof(double), hipMemcpyDeviceToHost, s));
    const int nnbeff = std::max(0, std::min(nnbmax, lmax - 1));
    const size_t full = (size_t)ni * lmax * sizeof(int);
    if (full > ((size_t)2 << 20)) {
        // large blocks: only the rows' used part crosses PCIe (about nnb + 1 of lmax ints
        // per row), packed on the device and spread into the caller's rows here
        int *len = static_cast<int *>(g.dlen.get((size_t)(ni + 1) * sizeof(int)));
        int *off = static_cast<int *>(g.doff.get((size_t)(ni + 1) * sizeof(int)));
        int *pack = static_cast<int *>(g.dpack.get(full));
        pack_rows(dlist, ni, lmax, nnbeff, len, off, pack, g.dtmp, s);
        int *hoff = static_cast<int *>(g.hoff.get((size_t)(ni + 1) * sizeof(int)));
        NB6_HIP(hipMemcpyAsync(hoff, off, (size_t)(ni + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
        NB6_HIP(hipStreamSynchronize(s));
        const size_t total = (size_t)hoff[ni];
        int *hp = static_cast<int *>(g.hlist.get(total * sizeof(int)));
        NB6_HIP(hipMemcpyAsync(hp, pack, total * sizeof(int), hipMemcpyDeviceToHost, s));
        NB6_HIP(hipStreamSynchronize(s));
        const double t1 = wtime();
        par_for(ni, [&](int a, int b) {
            for (int i = a; i < b; i++) {
                acc[i][0] = ho[0 * (size_t)ni + i];
                acc[i][1] = ho[1 * (size_t)ni + i];
                acc[i][2] = ho[2 * (size_t)ni + i];
                jrk[i][0] = ho[3 * (size_t)ni + i];
                jrk[i][1] = ho[4 * (size_t)ni + i];
                jrk[i][2] = ho[5 * (size_t)ni + i];
                pot[i] = ho[6 * (size_t)ni + i];
                std::memcpy(list + (size_t)i * lmax, hp + hoff[i],
                            (size_t)(hoff[i + 1] - hoff[i]) * sizeof(int));
            }
        });
        const double t2 = wtime();
        g.time_grav += t1 - t0;
        g.time_reduce += t2 - t1;
        return;
    }
    int *hl = static_cast<int *>(g.hlist.get(full));
    NB6_HIP(hipMemcpyAsync(hl, dlist, full, hipMemcpyDeviceToHost, s));
    NB6_HIP(hipStreamSynchronize(s));
    const double t1 = wtime();
    par_for(ni, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            acc[i][0] = ho[0 * (size_t)ni + i];
            acc[i][1] = ho[1 * (size_t)ni + i];
            acc[i][2] = ho[2 * (size_t)ni + i];
            jrk[i][0] = ho[3 * (size_t)ni + i];
            jrk[i][1] = ho[4 * (size_t)ni + i];
            jrk[i][2] = ho[5 * (size_t)ni + i];
            pot[i] = ho[6 * (size_t)ni + i];
            const int *src = hl + (size_t)i * lmax;
            int *dst = list + (size_t)i * lmax;
            const int c = src[0];
            dst[0] = c;
            if (c > 0 && c <= nnbeff) std::memcpy(dst + 1, src + 1, (size_t)c * sizeof(int));
        }
    });
    const double t2 = wtime();
    g.time_grav += t1 - t0;
    g.time_reduce += t2 - t1;
}

void gpunb_profile_(int *irank) {
    Gpunb &g = g_nb;
    if (g.icall) {
        fprintf(stderr,
                "[R.%d GPU Reg.F ] Nsend %d  Ngrav %d  <Ni> %d   send(s) %f grav(s) %f  nb(s) %f  "
                "out(s) %f  Perf.(Gflops) %f\n",
                *irank, g.isend, g.icall, g.isend ? g.ini / g.isend : 0, g.time_send, g.time_grav,
                g.time_reduce, g.time_out,
                g.time_grav > 0 ? 60.e-9 * (double)g.num_inter / g.time_grav : 0.0);
    }
    g.time_send = g.time_grav = g.time_reduce = g.time_out = 0.0;
    g.num_inter = 0;
    g.icall = g.ini = g.isend = 0;
}

// ================================================================== gpupot
static HBuf g_pot_h, g_pot_ho;
static DBuf g_pot_d, g_pot_do;

void gpupot_(int *irank, int *istart, int *ni_, int *n_, double m[], double x[][3], double pot[]) {
    Device &d = device(*irank);
    if (!d.ready) return;
    const int n = *n_, ni = *ni_, i0 = *istart - 1;
    if (n <= 0 || ni < 0 || i0 < 0 || i0 + ni > n) {
        set_error(E_ARG, "gpupot: istart/ni outside the particle range");
        return;
    }
    if (!check_finite(m, n) || !check_finite(&x[0][0], 3 * (size_t)n)) {
        set_error(E_ARG, "gpupot: NaN in particle data");
        return;
    }
    const double t0 = wtime();
    double *h = static_cast<double *>(g_pot_h.get((size_t)4 * n * sizeof(double)));
    par_for(n, [&](int a, int b) {
        for (int j = a; j < b; j++) {
            h[0 * (size_t)n + j] = x[j][0];
            h[1 * (size_t)n + j] = x[j][1];
            h[2 * (size_t)n + j] = x[j][2];
            h[3 * (size_t)n + j] = m[j];
        }
    });
    double *dx = static_cast<double *>(g_pot_d.get((size_t)4 * n * sizeof(double)));
    double *dp = static_cast<double *>(g_pot_do.get((size_t)std::max(ni, 1) * sizeof(double)));
    double *hp = static_cast<double *>(g_pot_ho.get((size_t)std::max(ni, 1) * sizeof(double)));
    NB6_HIP(hipMemcpyAsync(dx, h, (size_t)4 * n * sizeof(double), hipMemcpyHostToDevice, d.stream));
    if (pot_launch(n, n, dx, i0, ni, dp, d.stream) != OK) return;
    NB6_HIP(hipMemcpyAsync(hp, dp, (size_t)ni * sizeof(double), hipMemcpyDeviceToHost, d.stream));
    NB6_HIP(hipStreamSynchronize(d.stream));
    std::memcpy(pot, hp, (size_t)ni * sizeof(double));
    const double t1 = wtime();
    fprintf(stderr, "[R.%d GPU Pot.A] Ni %d  NTOT %d  pot(s) %f\n", *irank, ni, n, t1 - t0);
}

// ================================================================== irr_simd
void irr_simd_open_(int *nmax, int *lmax, int *rank) { irr_open(*nmax, *lmax, *rank); }
void irr_simd_close_(int *rank) { irr_close(*rank); }
void irr_simd_profile_(int *rank) { irr_profile(*rank); }
void irr_simd_set_jp_(int *addr, double pos[3], double vel[3], double acc2[3], double jrk6[3],
                      double *mass, double *time) {
    irr_set_jp(*addr, pos, vel, acc2, jrk6, *mass, *time);
}
void irr_simd_set_list_(int *addr, int *nblist) { irr_set_list(*addr, nblist); }
void irr_simd_firr_vec_(double *ti, int *ni, int addr[], double acc[][3], double jrk[][3],
                        int nnbid[]) {
    irr_firr_vec(*ti, *ni, addr, acc, jrk, nnbid);
}

// gpuirr_* aliases (argument lists of the commented GPUIRR calls, fpoly0.F:42,70,202)
void gpuirr_open_(int *nmax, int *lmax) { irr_open(*nmax, *lmax, 0); }
void gpuirr_close_(void) { irr_close(0); }
void gpuirr_set_jp_(int *addr, double pos[3], double vel[3], double acc2[3], double jrk6[3],
                    double *mass, double *time) {
    irr_set_jp(*addr, pos, vel, acc2, jrk6, *mass, *time);
}
void gpuirr_set_list_(int *addr, int *nblist) { irr_set_list(*addr, nblist); }
void gpuirr_firr_vec_(double *ti, int *ni, int addr[], double acc[][3], double jrk[][3],
                      int nnbid[]) {
    irr_firr_vec(*ti, *ni, addr, acc, jrk, nnbid);
}

// ================================================================== extensions
int nb6_regf_device(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax, int m_flag,
                    double *d_out, int *d_list, void *stream) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_regf_device: library not open"); return E_STATE; }
    if (ni < 0 || ni_stride < ni) { set_error(E_ARG, "nb6_regf_device: bad ni / stride"); return E_ARG; }
    if (ni == 0) return OK;
    Device &d = device();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : d.stream;
    g.num_inter += (long long)ni * g.js.n;
    g.ini += ni;
    g.icall++;
    return regf_launch(g.js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out, d_list, s);
}

int nb6_regf_device_noalloc(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                            int m_flag, double *d_out, int *d_list, void *stream) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_regf_device_noalloc: library not open"); return E_STATE; }
    if (ni < 0 || ni_stride < ni) { set_error(E_ARG, "nb6_regf_device_noalloc: bad ni / stride"); return E_ARG; }
    if (ni == 0) return OK;
    Device &d = device();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : d.stream;
    g.num_inter += (long long)ni * g.js.n;
    g.ini += ni;
    g.icall++;
    return regf_launch_noalloc(g.js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out, d_list, s);
}

int nb6_send_device(int nj, int nj_stride, const double *d_j, void *stream) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_send_device: library not open"); return E_STATE; }
    if (nj < 0 || nj > g.nbmax || nj_stride < nj) {
        set_error(E_ARG, "nb6_send_device: nj outside [0, nbmax] or bad stride");
        return E_ARG;
    }
    Device &d = device();
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : d.stream;
    jset_from_soa(g.js, nj, nj_stride, d_j, s);
    g.isend++;
    return OK;
}

int nb6_pot_device(int n, int n_stride, const double *d_x, int i0, int ni, double *d_pot,
                   void *stream) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : d.stream;
    return pot_launch(n, n_stride, d_x, i0, ni, d_pot, s);
}

int nb6_set_regf_config(int threads, int i_per_lane, int min_groups) {
    if (threads != 0 && threads != 128 && threads != 256 && threads != 512) {
        set_error(E_ARG, "nb6_set_regf_config: threads must be 0 (by ni), 128, 256 or 512");
        return E_ARG;
    }
    if (i_per_lane != 1 && i_per_lane != 2) {
        set_error(E_ARG, "nb6_set_regf_config: i_per_lane must be 1 or 2");
        return E_ARG;
    }
    if (min_groups < 1) {
        set_error(E_ARG, "nb6_set_regf_config: min_groups must be positive");
        return E_ARG;
    }
    regf_config().threads = threads;
    regf_config().ipl = i_per_lane;
    regf_config().min_groups = min_groups;
    return OK;
}

int nb6_set_regf_tuning(int isort_min, int far_test) {
    if (isort_min < 0 || (far_test != 0 && far_test != 1)) {
        set_error(E_ARG, "nb6_set_regf_tuning: isort_min must be >= 0 and far_test 0 or 1");
        return E_ARG;
    }
    regf_config().isort_min = isort_min;
    regf_config().far_test = far_test;
    return OK;
}

int nb6_set_regf_precision(int mode) {
    if (mode < 0 || mode > 2) {
        set_error(E_ARG, "nb6_set_regf_precision: mode must be 0 (FP32), 1 (float-float) or 2 (FP64)");
        return E_ARG;
    }
    regf_config().prec = mode;
    return OK;
}

double nb6_kernel_ms(const char *name) { return ktimer(name).ms(); }
int nb6_set_kernel_timing(int on) {
    set_kernel_timing(on != 0);
    return OK;
}

int nb6_ens_open(int max_clusters, int max_particles) { return ens_open(max_clusters, max_particles); }
int nb6_ens_close(void) { return ens_close(); }
int nb6_ens_send_device(int nclusters, const int *offsets, const double *d_soa, int stride,
                        void *stream) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    return ens_send_device(nclusters, offsets, d_soa, stride,
                           stream ? static_cast<hipStream_t>(stream) : d.stream);
}
int nb6_ens_regf_device(const double *d_i, int stride, int lmax, int nnbmax, int m_flag,
                        double *d_out, int *d_list, int use_graph, void *stream) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    return ens_regf_device(d_i, stride, lmax, nnbmax, m_flag, d_out, d_list, use_graph,
                           stream ? static_cast<hipStream_t>(stream) : d.stream);
}
int nb6_ens_items(void) { return ens_items(); }
int nb6_ens_last_refined(void) { return regf_items_refined(); }
int nb6_ens_autotune(const double *d_i, int stride, int lmax, int nnbmax, int m_flag,
                     double *d_out, int *d_list, int *best, void *stream) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    return ens_autotune(d_i, stride, lmax, nnbmax, m_flag, d_out, d_list, best,
                        stream ? static_cast<hipStream_t>(stream) : d.stream);
}
int nb6_regf_autotune_device(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                             int m_flag, double *d_out, int *d_list, int *best, void *stream) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_regf_autotune_device: library not open"); return E_STATE; }
    if (ni <= 0 || ni_stride < ni) { set_error(E_ARG, "nb6_regf_autotune_device: bad ni / stride"); return E_ARG; }
    Device &d = device();
    return regf_autotune(g.js, ni, ni_stride, d_i, lmax, nnbmax, m_flag, d_out, d_list, best,
                         stream ? static_cast<hipStream_t>(stream) : d.stream);
}

int nb6_device_copy(void *dst, const void *src, unsigned long long bytes, void *stream) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    if (bytes % 16) { set_error(E_ARG, "nb6_device_copy: size must be a multiple of 16"); return E_ARG; }
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : d.stream;
    device_copy(dst, src, bytes, s);
    return OK;
}
int nb6_last_regf_chunks(void) { return regf_last_chunks(); }
int nb6_last_regf_batches(void) { return regf_last_batches(); }
int nb6_set_regf_refine(int on) {
    if (on != 0 && on != 1) { set_error(E_ARG, "nb6_set_regf_refine: on must be 0 or 1"); return E_ARG; }
    regf_config().refine = on;
    return OK;
}
int nb6_last_regf_refined(void) { return regf_last_refined(); }
int nb6_last_regf_refined_chunks(void) {
    int c = 0;
    (void)regf_last_refined(&c);
    return c;
}
int nb6_set_regf_refine_weights(double k0, double k1, double k2, double k3) {
    const double k[4] = {k0, k1, k2, k3};
    for (int q = 0; q < 4; q++)
        if (!(k[q] >= 0.0 && k[q] <= 1e6)) {
            set_error(E_ARG, "nb6_set_regf_refine_weights: weights outside [0, 1e6]");
            return E_ARG;
        }
    for (int q = 0; q < 4; q++) regf_config().kref[q] = regf_config().kref_small[q] = (float)k[q];
    return OK;
}
int nb6_set_regf_debug_bounds(double *d_bounds) {
    regf_set_debug_bounds(d_bounds);
    return OK;
}
int nb6_set_ens_refine_weights(double k0, double k1, double k2, double k3) {
    const double k[4] = {k0, k1, k2, k3};
    for (int q = 0; q < 4; q++)
        if (!(k[q] >= 0.0 && k[q] <= 1e6)) {
            set_error(E_ARG, "nb6_set_ens_refine_weights: weights outside [0, 1e6]");
            return E_ARG;
        }
    for (int q = 0; q < 4; q++) regf_config().kref_ens[q] = (float)k[q];
    return OK;
}
int nb6_set_regf_near(double f) {
    if (!(f >= 1.0 && f <= 4.0)) { set_error(E_ARG, "nb6_set_regf_near: factor outside [1, 4]"); return E_ARG; }
    regf_config().near = (float)f;
    return OK;
}
int nb6_set_regf_slot_budget(int mib) {
    if (mib < 1) { set_error(E_ARG, "nb6_set_regf_slot_budget: budget must be at least 1 MiB"); return E_ARG; }
    regf_config().slot_budget_mb = mib;
    return OK;
}

static hipStream_t pick(void *stream) {
    return stream ? static_cast<hipStream_t>(stream) : device().stream;
}

void nb6_jp_open_(int *nmax) { (void)jp_open(*nmax); }
void nb6_jp_close_(void) { (void)jp_close(); }
void nb6_jp_set_(int *n, int addr[], double x0[][3], double x0dot[][3], double f[][3],
                 double fdot[][3], double m[], double t0[]) {
    const int k = *n;
    if (k < 0) { set_error(E_ARG, "nb6_jp_set: n must not be negative"); return; }
    if (k == 0) return;
    const size_t k3 = 3 * (size_t)k;
    if (!check_finite(&x0[0][0], k3) || !check_finite(&x0dot[0][0], k3) ||
        !check_finite(&f[0][0], k3) || !check_finite(&fdot[0][0], k3) ||
        !check_finite(m, (size_t)k) || !check_finite(t0, (size_t)k)) {
        set_error(E_ARG, "nb6_jp_set: NaN in particle data");
        return;
    }
    (void)jp_set(k, addr, x0, x0dot, f, fdot, m, t0);
}
void nb6_jp_predict_send_(double *time, int *nj) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_jp_predict_send: gpunb library not open"); return; }
    const int n = *nj;
    if (n < 0 || n > g.nbmax) { set_error(E_ARG, "nb6_jp_predict_send: nj outside [0, nbmax]"); return; }
    const double t0 = wtime();
    Device &d = device();
    if (jp_predict(*time, n, g.js.in64.as<double>(), d.stream) != OK) return;
    jset_build(g.js, n, nullptr, 0.0, d.stream);
    g.isend++;
    g.time_send += wtime() - t0;
}
void nb6_jp_regf_(int *ni_, int addr[], double h2[], double dtr[], double acc[][3],
                  double jrk[][3], double pot[], int *lmax_, int *nnbmax_, int *list, int *m_flag) {
    Gpunb &g = g_nb;
    if (!g.open) { set_error(E_STATE, "nb6_jp_regf: gpunb library not open"); return; }
    const int ni = *ni_;
    if (ni <= 0 || *lmax_ < 1) { set_error(E_ARG, "nb6_jp_regf: ni and lmax must be positive"); return; }
    if (!check_finite(h2, ni) || !check_finite(dtr, ni)) {
        set_error(E_ARG, "nb6_jp_regf: NaN in h2 or dtr");
        return;
    }
    const double t0 = wtime();
    g.num_inter += (long long)ni * g.js.n;
    g.ini += ni;
    g.icall++;
    Device &d = device();
    double *di = static_cast<double *>(g.di.get((size_t)8 * ni * sizeof(double)));
    if (jp_iblock(ni, addr, h2, dtr, di, d.stream) != OK) return;
    regf_finish(g, ni, di, *lmax_, *nnbmax_, *m_flag, acc, jrk, pot, list, d.stream, t0);
}

void nb6_irr_sync_jp_(int *n) {
    Device &d = device();
    if (!d.ready) return;
    (void)irr_sync_jp(*n, d.stream);
}
int nb6_irr_set_lists_device(int ni, const int *d_addr, const int *d_list, int lmax, int jbase,
                             void *stream) {
    if (!device().ready) return E_NODEV;
    return irr_set_lists_device(ni, d_addr, d_list, lmax, jbase, pick(stream));
}
int nb6_set_irr_config(int lanes_per_i, int sort_i) { return irr_config(lanes_per_i, sort_i); }
int nb6_set_irr_predict(int mode) { return irr_predict_mode(mode); }
int nb6_irr_firr_device(double ti, int ni, const int *d_addr, double *d_out, void *stream) {
    if (!device().ready) return E_NODEV;
    return irr_firr_device(ti, ni, d_addr, d_out, pick(stream));
}

int nb6_energy_device(const nb6_energy_in *e, double *h_out, void *stream) {
    return energy_device(e, h_out, static_cast<hipStream_t>(stream));   // validates first
}

int nb6_set_regcor_pack(int mode) { return regcor_set_pack(mode); }
int nb6_regcor_device(const nb6_regcor_params *p, const nb6_regstate *s, int ni,
                      const int *d_addr, const double *d_force, int force_stride,
                      int *d_nlist, int *d_counters, void *stream) {
    return regcor_device(p, s, ni, d_addr, d_force, force_stride, d_nlist, d_counters,
                         static_cast<hipStream_t>(stream));   // validates first
}

int nb6_rsnet_nparam(void) { return rsnet_nparam(); }
int nb6_rsnet_open(unsigned long long seed) { return rsnet_open(seed); }
int nb6_rsnet_close(void) { return rsnet_close(); }
int nb6_rsnet_set_params(const float *params) { return rsnet_set_params(params); }
int nb6_rsnet_get_params(float *params) { return rsnet_get_params(params); }
int nb6_rsnet_set_norm(const float *mu, const float *sigma) { return rsnet_set_norm(mu, sigma); }
int nb6_rsnet_forward_device(int n, const float *d_feat, float *d_y, void *stream) {
    if (!device().ready) return E_NODEV;
    return rsnet_forward(n, d_feat, d_y, pick(stream));
}
int nb6_rsnet_grad_device(int n, const float *d_feat, const float *d_t, float *d_grad,
                          float *d_loss, void *stream) {
    if (!device().ready) return E_NODEV;
    return rsnet_grad(n, d_feat, d_t, d_grad, d_loss, pick(stream));
}
int nb6_rsnet_train_step_device(int n, const float *d_feat, const float *d_t, float lr,
                                float *h_loss, void *stream) {
    if (!device().ready) return E_NODEV;
    return rsnet_train_step(n, d_feat, d_t, lr, h_loss, pick(stream));
}
int nb6_rsnet_train_epoch_device(int n, const float *d_feat, const float *d_t, int batch,
                                 float lr, unsigned long long seed, float *h_loss, void *stream) {
    if (!device().ready) return E_NODEV;
    return rsnet_train_epoch(n, d_feat, d_t, batch, lr, seed, h_loss, pick(stream));
}
int nb6_rs_features_device(int n, int stride, const double *d_x, const double *d_acc,
                           const double *d_jrk, const double *d_mass, const double *d_rs,
                           const double *d_dtr, const int *d_nnb, const double *centre,
                           float *d_feat, void *stream) {
    if (!device().ready) return E_NODEV;
    if (stride < n) { set_error(E_ARG, "nb6_rs_features_device: stride below n"); return E_ARG; }
    return rs_features(n, stride, d_x, d_acc, d_jrk, d_mass, d_rs, d_dtr, d_nnb, centre, d_feat,
                       pick(stream));
}
int nb6_rs_predict_device(int n, const float *d_feat, const double *d_rs_prev,
                          const double *d_rs_rule, const int *d_use_rule, double *d_rs_out,
                          double *d_ib, int ib_stride, int m_flag, double bodym, void *stream) {
    if (!device().ready) return E_NODEV;
    if (d_use_rule && !d_rs_rule) {
        set_error(E_ARG, "nb6_rs_predict_device: use_rule given without rs_rule");
        return E_ARG;
    }
    return rs_predict(n, d_feat, d_rs_prev, d_rs_rule, d_use_rule, d_rs_out, d_ib, ib_stride,
                      m_flag, bodym, pick(stream));
}
int nb6_rs_rule_device(int n, int stride, const double *d_x, const double *d_v,
                       const double *d_dtr, const int *d_nnb, const int *d_nnb0,
                       const int *d_list, int lmax, const double *d_xj, const double *d_vj,
                       int jstride, const double *glob, int nnbopt, int nnbmax, double *d_rs,
                       void *stream) {
    if (!device().ready) return E_NODEV;
    if (stride < n || (d_list && (lmax < 2 || !d_xj || !d_vj))) {
        set_error(E_ARG, "nb6_rs_rule_device: bad stride or list arguments");
        return E_ARG;
    }
    return rs_rule(n, stride, d_x, d_v, d_dtr, d_nnb, d_nnb0, d_list, lmax, d_xj, d_vj, jstride, glob,
                   nnbopt, nnbmax, d_rs, pick(stream));
}
int nb6_rs_kth_device(int ni, int ib_stride, const double *d_ib, const int *d_list, int lmax,
                      int k, double *d_rs, void *stream) {
    const JSet *js = nb6::gpunb_jset();
    if (!js) { set_error(E_STATE, "nb6_rs_kth_device: library not open"); return E_STATE; }
    if (ib_stride < ni) { set_error(E_ARG, "nb6_rs_kth_device: stride below ni"); return E_ARG; }
    return rs_kth(*js, ni, ib_stride, d_ib, d_list, lmax, k, d_rs, pick(stream));
}

void *nb6_dev_alloc(unsigned long long bytes) {
    if (!device().ready) return nullptr;
    void *p = nullptr;
    NB6_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return p;
}
void nb6_dev_free(void *p) {
    if (p) NB6_HIP(hipFree(p));
}
int nb6_memcpy_htod(void *dst, const void *src, unsigned long long bytes) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    NB6_HIP(hipStreamSynchronize(d.stream));
    NB6_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return OK;
}
int nb6_memcpy_dtoh(void *dst, const void *src, unsigned long long bytes) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    NB6_HIP(hipStreamSynchronize(d.stream));
    NB6_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return OK;
}
int nb6_sync(void) {
    Device &d = device();
    if (!d.ready) return E_NODEV;
    NB6_HIP(hipStreamSynchronize(d.stream));
    NB6_HIP(hipDeviceSynchronize());
    return OK;
}

}  // extern "C"
