// Internal declarations shared by the translation units of libnb6gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace nb6 {

// ---------------------------------------------------------------- errors
enum Err : int {
    OK = 0,
    E_ARG = 1,        // invalid argument (sizes, NaN input, ...)
    E_STATE = 2,      // call in the wrong library state (not open, ...)
    E_HIP = 3,        // HIP runtime failure
    E_NODEV = 4,      // no usable device
};

void set_error(int code, const std::string &msg);   // records; aborts if enabled
void hip_fail(hipError_t e, const char *what, const char *file, int line);  // aborts

#define NB6_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) ::nb6::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

// ---------------------------------------------------------------- device
struct Device {
    bool ready = false;
    int id = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
};
Device &device(int irank = 0);          // initialises on first use

// Grow-only device buffer.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need);
    void *try_get(size_t need);   // nullptr (and no buffer) where the device has no room
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Grow-only pinned host buffer.
struct HBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t need);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Event timing of the last launch of a kernel family.
struct KTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool armed = false;
    void start(hipStream_t s);
    void stop(hipStream_t s);
    double ms();          // synchronises on the stop event
};
KTimer &ktimer(const char *name);
// kernel timers record events only while this is on (nb6_set_kernel_timing)
bool kernel_timing();
void set_kernel_timing(bool on);

// ---------------------------------------------------------------- regf
constexpr int JTILE = 256;   // j-particles per tile (one tile origin each)

// j-set on the device, Morton-ordered at every send (jset.hip):
//   in64  AoS [cap][8] caller order    (x y z m vx vy vz 0), upload staging
//   rec64 AoS [cap][8] sorted order    (x y z m vx vy vz 0)
//   rec32 [cap][2] float4 sorted       (x-P, y-P, z-P, m), (vx-V, vy-V, vz-V, 0)
//   torg  [ntile][8] FP64 tile origins (Px Py Pz Vx Vy Vz 0 0)
//   perm  [cap] int: sorted position -> caller's 0-based index
//   tbox  [ntile][16] FP32 bounds of each tile's records relative to its origin:
//         pos lo[3], pos hi[3], vel lo[3], vel hi[3], max mass, record count, 0, 0
//         (far massless padding records, |x| >= 1e16, are left out)
constexpr int TBOX = 16;
struct JSet {
    int n = 0;
    int cap = 0;
    int ntile = 0;
    double mmax = 0.0;
    double box[6] = {0, 0, 0, 0, 0, 0};   // lo[3], hi[3] of the positions (host copy)
    DBuf in64, rec64, rec32, torg, perm, keys, vals, stat, sort_tmp, tbox;
};
// per-tile bounds of rec32 (after rec32 is built): one work-group per tile
// Work-group b of G -> position in the work order such that XCD x (work-groups are
// dealt to the 8 XCDs round robin, b mod 8) receives a contiguous range of the work,
// and its L2 sees one region of it at a time.
__device__ __forceinline__ int xcd_order(int b, int G) {
    const int xcd = b & 7, k = b >> 3;
    const int per = G >> 3, rem = G & 7;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

void jset_tile_boxes(JSet &js, int n, hipStream_t s);
// tile origins (P from the middle record, or the first: middle = 0), FP32 records,
// tile boxes, from the sorted rec64 of n records
void jset_tiles(JSet &js, int n, int middle, hipStream_t s);
void jset_reserve(JSet &js, int cap);
void jset_release(JSet &js);
// in64 already holds n records; box = {lo x,y,z, hi x,y,z} or nullptr (computed)
void jset_build(JSet &js, int n, const double *box, double mmax, hipStream_t s);
void jset_from_soa(JSet &js, int n, int stride, const double *d_soa, hipStream_t s);
// Rows [count, entries...] of a [n][lmax] list array packed end to end on the device:
// off[i] (n + 1 entries, off[n] = total) and pack[off[i] .. off[i + 1]) = the row's
// 1 + count leading ints (only the count if it is negative or above maxc); len is
// n + 1 ints of scratch (jset.hip).
void pack_rows(const int *list, int n, int lmax, int maxc, int *len, int *off, int *pack,
               DBuf &tmp, hipStream_t s);

struct RegfConfig {
    int threads = 0;      // work-group size = LDS j-tile (128, 256 or 512; 0: by ni)
    int ipl = 1;          // i-particles per lane (1 or 2)
    int min_groups = 4096;
    int isort_min = 2048; // i-sets this large are Morton-ordered (0: never)
    int far_test = 1;     // skip the pair test on tiles far from the whole wavefront
    int prec = 0;         // pair arithmetic: 0 FP32, 1 float-float, 2 FP64 (study variants)
    int slot_budget_mb = 32768;   // largest work-array footprint of one call; larger
                                  // calls run in batches of i (same results)
    int refine = 1;       // FP64 pass over rows whose FP32 error scale is too large
    float near = 2.0f;    // near field (FP64 in the finalize): r < near RS
    float kref[4] = {0.5f, 0.5f, 2.0f, 0.0f};   // weights of the refine error scales
    // weights for j-sets of at most small_nj: the velocity-rounding scale (a linear
    // bound, loose by about the square root of the terms on large j-sets) counts too,
    // where recomputing a row costs at most small_nj FP64 pairs
    float kref_small[4] = {0.5f, 0.5f, 2.0f, 1.0f};
    int small_nj = 4096;
    float kref_ens[4] = {0.5f, 0.5f, 0.5f, 0.0f};   // the same for the batched form
};
RegfConfig &regf_config();
// the configuration a call with ni i-particles runs with (threads resolved)
RegfConfig regf_effective(int ni);

// Launch the regular-force pipeline on device data.  d_i SoA[8][ni_stride],
// d_out SoA[7][ni_stride], d_list [ni][lmax].
int regf_launch(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s);
// The same pipeline with no allocation: work memory from a fixed device array (see
// nb6_regf_device_noalloc); ni at most REGF_NOALLOC_NI_MAX.
constexpr int REGF_NOALLOC_NI_MAX = 65536;
int regf_noalloc_setup();
int regf_launch_noalloc(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax,
                        int nnbmax, int m_flag, double *d_out, int *d_list, hipStream_t s);
int regf_last_chunks();
int regf_last_batches();
// rows of the last call's FP64 pass, and its (row, chunk) items (synchronises)
int regf_last_refined(int *chunk_items = nullptr);
void regf_release();

// Scratch of one regular-force pipeline: FP32 partials, candidate counts and slots.
struct RegfWork {
    DBuf part, cnt, hslot, sslot;
    DBuf ikey, ival, itmp;           // Morton sort of the i-set
    DBuf fixrows, nfix, fixpart;     // FP64 pass: whole rows, counters, chunk sums
    DBuf items, delta, rowsel;       // FP64 pass: (row, chunk) items, candidate terms, rows
    DBuf segitems, part64;           // FP64 pass: items by chunk, FP64 sums per sub-range
    void release() {
        part.release(); cnt.release(); hslot.release(); sslot.release();
        ikey.release(); ival.release(); itmp.release();
        fixrows.release(); nfix.release(); fixpart.release();
        items.release(); delta.release(); rowsel.release();
        segitems.release(); part64.release();
    }
};

// Morton order of ni i-particles (SoA [8][stride] positions) in the j-set's box:
// returns the device array slot -> i row (jset.hip).
const int *jset_sort_i(const JSet &js, int ni, int stride, const double *d_i, RegfWork &w,
                       hipStream_t s);

// One persistent launch over work items (slot begin, slot end, jbeg, jend) of a padded
// j-set; nslots rows of work arrays, imap[slot] = row of d_i / d_out / d_list (< 0:
// empty slot); queue is a device counter the caller zeroes first; jbase[row] maps j
// back to the caller's indexing; padoff[nc + 1] are the clusters' slot ranges, maxseg the
// longest (the FP64 pass of flagged rows runs over the row's own cluster).  Safe under stream capture once
// w is sized.
int regf_items_launch(const JSet &js, int nslots, const int *imap, int stride, const double *d_i,
                      const int4 *items, int nitems, int *queue, const int *jbase,
                      const int *padoff, int nc, int maxseg, int lmax, int nnbmax, int m_flag,
                      double *d_out, int *d_list, int grid, RegfWork &w, hipStream_t s);
int regf_items_refined();     // rows of the last batched step's FP64 pass (synchronises)
void regf_items_forget();     // the ensemble's work buffers are gone
void regf_set_debug_bounds(double *d);   // rows' error scales to d[4][out_stride] (or null)

// streaming copy of bytes (multiple of 16) on stream s
void device_copy(void *dst, const void *src, size_t bytes, hipStream_t s);

// ---------------------------------------------------------------- learned RS (rsnet.hip)
int rsnet_nparam();
int rsnet_open(unsigned long long seed);
int rsnet_close();
int rsnet_set_params(const float *p);
int rsnet_get_params(float *p);
int rsnet_set_norm(const float *mu, const float *sigma);
int rsnet_forward(int n, const float *d_feat, float *d_y, hipStream_t s);
int rsnet_grad(int n, const float *d_feat, const float *d_t, float *d_grad, float *d_loss,
               hipStream_t s);
int rsnet_train_epoch(int n, const float *d_feat, const float *d_t, int batch, float lr,
                      unsigned long long seed, float *h_loss, hipStream_t s);
int rsnet_train_step(int n, const float *d_feat, const float *d_t, float lr, float *h_loss,
                     hipStream_t s);
int rs_features(int n, int stride, const double *x, const double *acc, const double *jrk,
                const double *mass, const double *rs, const double *dtr, const int *nnb,
                const double *centre, float *feat, hipStream_t s);
int rs_predict(int n, const float *feat, const double *rs_prev, const double *rs_rule,
               const int *use_rule, double *rs_out, double *ib, int ib_stride, int m_flag,
               double bodym, hipStream_t s);
int rs_rule(int n, int stride, const double *x, const double *v, const double *dtr, const int *nnb,
            const int *nnb0, const int *list, int lmax, const double *xj, const double *vj,
            int jstride, const double *glob, int nnbopt, int nnbmax, double *rs, hipStream_t s);
int rs_kth(const JSet &js, int ni, int ib_stride, const double *ib, const int *list, int lmax, int k,
           double *rs_out, hipStream_t s);
int regf_autotune(const JSet &js, int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                  int m_flag, double *d_out, int *d_list, int *best, hipStream_t s);
int ens_autotune(const double *d_i, int stride, int lmax, int nnbmax, int m_flag, double *d_out,
                 int *d_list, int *best, hipStream_t s);
// device-resident predictor (jpred.hip)
int jp_open(int nmax);
int jp_close();
int jp_set(int n, const int *addr, const double (*x0)[3], const double (*v0)[3],
           const double (*f)[3], const double (*fdot)[3], const double *m, const double *t0);
int jp_predict(double time, int nj, double *aos, hipStream_t s);
int jp_iblock(int ni, const int *addr, const double *h2, const double *dtr, double *d_ib,
              hipStream_t s);
const double *jp_state(int *nmax);   // predictor store SoA [14][nmax] (nullptr when closed)
// irregular force fed from the device (irr.hip)
int irr_set_lists_device(int ni, const int *d_addr, const int *d_list, int lmax_r, int jbase,
                         hipStream_t s);
int irr_sync_jp(int n, hipStream_t s);
int irr_firr_device(double ti, int ni, const int *d_addr, double *d_out, hipStream_t s);

// ---------------------------------------------------------------- regular corrector
}  // namespace nb6
struct nb6_regcor_params;
struct nb6_regstate;
struct nb6_energy_in;
namespace nb6 {
int energy_device(const ::nb6_energy_in *e, double *h_out, hipStream_t s);
int regcor_set_pack(int mode);
int regcor_device(const ::nb6_regcor_params *p, const ::nb6_regstate *s, int ni,
                  const int *d_addr, const double *d_force, int force_stride, int *d_nlist,
                  int *d_counters, hipStream_t stream);
int irr_config(int lanes_per_i, int sort_i);
int irr_predict_mode(int mode);
// the j-set of the open gpunb library (nullptr when closed)
const JSet *gpunb_jset();

// ---------------------------------------------------------------- pot
int pot_launch(int n, int n_stride, const double *d_x, int i0, int ni, double *d_pot,
               hipStream_t s);
void pot_release();

// ---------------------------------------------------------------- irr
void irr_release();

}  // namespace nb6
