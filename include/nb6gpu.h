/*
 * nb6gpu.h -- C ABI of libnb6gpu.so, the gfx950 (MI355X) regular-force,
 * neighbour-list, potential and irregular-force library for Nbody6++GPU.
 *
 * Part 1 is the drop-in boundary: the exact symbols the Fortran integrator
 * already calls, with gfortran/flang mangling (lower case, one trailing
 * underscore, every argument by reference).  Part 2 is this library's own
 * extension API (device-resident entry points, profiling, error status); the
 * Fortran code never needs it.
 *
 * Array conventions (Fortran column-major seen from C):
 *   X(3,N)           == double x[N][3]
 *   LISTGP(LMAX,NI)  == int list[NI][LMAX]; list[i][0] = count, list[i][1..] = entries
 */
#ifndef NB6GPU_H
#define NB6GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== *
 * Part 1 -- reference ABI                                                *
 * ===================================================================== */

/* Select and initialise the device.  Replaces gpunb_devinit_
 * (src/Main/gpunb.velocity.cu:906-908, GPUNB_devinit :575-627).  GPU_LIST is
 * honoured like the reference (:583-598); the first listed device is used. */
void gpunb_devinit_(int *irank);

/* Open the regular-force library for up to *nbmax j-particles.  Replaces
 * gpunb_open_ (gpunb.velocity.cu:909-911, GPUNB_open :629-667).  A second open
 * prints "gpunb: it is already open" and returns, like :637-640. */
void gpunb_open_(int *nbmax, int *irank);

/* Release every device buffer.  Replaces gpunb_close_ (gpunb.velocity.cu:912-914,
 * GPUNB_close :669-702). */
void gpunb_close_(void);

/* Upload the j-particles (mass, position, velocity, FP64).  Replaces
 * gpunb_send_ (gpunb.velocity.cu:915-921, GPUNB_send :704-730).  The data stay
 * FP64 on the device. */
void gpunb_send_(int *nj, double mj[], double xj[][3], double vj[][3]);

/* Regular force, jerk and potential over the non-neighbours of ni i-particles,
 * and their neighbour lists.  Replaces gpunb_regf_ (gpunb.velocity.cu:922-936,
 * GPUNB_regf :732-881; AVX reg.avx.cpp:367-381).
 *   h2[i]   : RS^2 (or RS^2/BODYM when *m_flag != 0)
 *   dtr[i]  : regular step used by the velocity criterion
 *   j is a neighbour of i iff min(|dx|^2, |dx + dtr*dv|^2) < h2 (or < m_j*h2)
 *   list[i][0] = count, list[i][1..count] = ascending 0-based j;  a count above
 *   min(*nnbmax, *lmax-1) is returned as -count with no entries (reg.avx.cpp:320-322).
 * Any ni >= 1 is accepted (the reference is limited to NIMAX = 2048). */
void gpunb_regf_(int *ni, double h2[], double dtr[], double xi[][3], double vi[][3],
                 double acc[][3], double jrk[][3], double pot[], int *lmax,
                 int *nnbmax, int *list, int *m_flag);

/* Print the profile line and reset the counters.  Replaces gpunb_profile_
 * (gpunb.velocity.cu:937-939, GPUNB_profile :883-903). */
void gpunb_profile_(int *irank);

/* All-pairs potential pot[k] = sum_{j, r>0} m_j / r(istart-1+k, j) over the n
 * particles, k < *ni, *istart 1-based.  Replaces gpupot_ (gpupot.gpu.cu:116-127,
 * gpupot :61-114; AVX pot.avx.cpp:148-159). */
void gpupot_(int *irank, int *istart, int *ni, int *n, double m[], double x[][3],
             double pot[]);

/* Irregular (neighbour-list) force.  Replace irr_simd_*_ (src/Main/irr.avx.cpp:568-606). */
void irr_simd_open_(int *nmax, int *lmax, int *rank);          /* irr.avx.cpp:569-571 */
void irr_simd_close_(int *rank);                               /* irr.avx.cpp:572-574 */
void irr_simd_profile_(int *rank);                             /* irr.avx.cpp:575-577 */
/* Store the predictor record of particle *addr (1-based): irr.avx.cpp:579-589. */
void irr_simd_set_jp_(int *addr, double pos[3], double vel[3], double acc2[3],
                      double jrk6[3], double *mass, double *time);
/* Store the neighbour list of *addr: nblist[0] = nnb, nblist[1..] 1-based
 * addresses (irr.avx.cpp:590-595). */
void irr_simd_set_list_(int *addr, int *nblist);
/* Neighbour force and jerk of the *ni particles addr[] (1-based) at time *ti,
 * every member predicted to *ti; nnbid[k] = 1-based nearest neighbour
 * (irr.avx.cpp:596-605). */
void irr_simd_firr_vec_(double *ti, int *ni, int addr[], double acc[][3],
                        double jrk[][3], int nnbid[]);

/* gpuirr_* names used by the commented-out GPUIRR calls (fpoly0.F:42,70,202,
 * intgrt.F:1329).  The reference has no implementation; these are aliases of the
 * irr_simd_* entry points with the argument lists of those calls. */
void gpuirr_open_(int *nmax, int *lmax);
void gpuirr_close_(void);
void gpuirr_set_jp_(int *addr, double pos[3], double vel[3], double acc2[3],
                    double jrk6[3], double *mass, double *time);
void gpuirr_set_list_(int *addr, int *nblist);
void gpuirr_firr_vec_(double *ti, int *ni, int addr[], double acc[][3],
                      double jrk[][3], int nnbid[]);

/* ===================================================================== *
 * Part 2 -- extension API (no counterpart in the reference)              *
 * ===================================================================== */

/* Last error code (0 = none) and message; cleared by nb6_clear_error(). */
int         nb6_last_error(void);
const char *nb6_last_error_message(void);
void        nb6_clear_error(void);

/* 1 (default): argument errors abort the process with a message, like the
 * reference's assert()s.  0: they are reported through nb6_last_error() and the
 * call returns without touching its outputs. */
void nb6_set_abort_on_error(int on);

/* Number of HIP devices visible (0 when none); never initialises a context
 * beyond hipGetDeviceCount. */
int nb6_device_count(void);

/* Device-resident regular force (no host copies).  All pointers are device
 * pointers on the library's device:
 *   d_i   : SoA double[8][ni_stride]  (x, y, z, vx, vy, vz, h2, dtr)
 *   d_out : SoA double[7][ni_stride]  (ax, ay, az, jx, jy, jz, pot)
 *   d_list: int[ni][lmax] in the gpunb_regf_ row format
 * The j-set is the one uploaded by gpunb_send_ / nb6_send_device.
 * stream: a hipStream_t (NULL = the library's stream).  Asynchronous.
 * Returns 0 or an error code. */
int nb6_regf_device(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                    int m_flag, double *d_out, int *d_list, void *stream);

/* nb6_regf_device that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments, lists and FP64
 * pass; the work memory is one fixed 256 MiB device array per GPU, held until the
 * process ends and shared by every call on that GPU, so calls that overlap on one
 * GPU race.  ni at most 65536 (larger: E_ARG).  Against nb6_regf_device the i-set
 * is not Morton-ordered and fewer j-chunks may be used, so sums can differ in the
 * last FP32 roundings; lists are the same.  No kernel timers are recorded.  Needs
 * a j-set already sent (gpunb_send_ / nb6_send_device allocate). */
int nb6_regf_device_noalloc(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                            int m_flag, double *d_out, int *d_list, void *stream);

/* Device-resident j upload: d_j is SoA double[7][nj_stride] (x,y,z,vx,vy,vz,m). */
int nb6_send_device(int nj, int nj_stride, const double *d_j, void *stream);

/* Device-resident potential: d_x SoA double[4][n_stride] (x,y,z,m) for all n
 * particles; d_pot[k] for particles [i0, i0+ni) (0-based). */
int nb6_pot_device(int n, int n_stride, const double *d_x, int i0, int ni,
                   double *d_pot, void *stream);

/* Tile configuration of the regular-force kernel: work-group size = LDS j-tile
 * (128, 256 or 512; 0 = chosen per call: 512 for ni >= 65536, else 256),
 * i-particles per lane (1 or 2), and the minimum number of work-groups a call is
 * split into (j-chunks are added until it is reached).  Default: 0, 1, 4096. */
int nb6_set_regf_config(int threads, int i_per_lane, int min_groups);

/* Ordering and pruning of the regular-force kernel: i-sets of at least isort_min
 * particles are processed in Morton order of their positions (0 = never), and
 * far_test = 1 lets a wavefront skip the per-pair neighbour test on a 256-j tile
 * that is provably outside every lane's near field.  Neither changes results.
 * Defaults: 2048, 1. */
int nb6_set_regf_tuning(int isort_min, int far_test);

/* Pair arithmetic of the regular-force kernel, for precision studies: 0 = FP32
 * with FP64 differencing and accumulation (default, the production kernel),
 * 1 = float-float (FP32 hi/lo pairs), 2 = FP64.  Lists are the exact FP64 lists
 * in every mode; modes 1 and 2 run 256-wide with one i per lane. */
int nb6_set_regf_precision(int mode);

/* Average device time (ms) of the last launch of the named kernel family
 * ("regf_tile", "regf_finalize", "regf_chunkfix", "regf_fix", "pot", "irr",
 * "regcor", "energy") measured with HIP events (-1: not recorded), and the number
 * of j-chunks the last regf call used.  Events are recorded only after
 * nb6_set_kernel_timing(1) (off by default: each pair leaves the stream idle for
 * several microseconds, which small calls feel). */
double nb6_kernel_ms(const char *name);
int    nb6_set_kernel_timing(int on);
int    nb6_last_regf_chunks(void);

/* Work-array budget of one regular-force call, in MiB (default 32768).  The work
 * arrays hold one row per (i-particle, j-chunk): FP64 partials, candidate counts and
 * candidate slots, (nnbmax + 8 + 1024) * 4 + 64 bytes each (about 6.3 KB at
 * nnbmax 550).  A call whose rows would exceed the budget runs in batches of whole
 * i-blocks; chunks, finalize form and results are those of the unbatched call.
 * nb6_last_regf_batches() returns the number of batches of the last call.  (No
 * counterpart in the reference, whose i-sets are capped at NIMAX = 2048,
 * gpunb.velocity.cu:24,751.) */
int nb6_set_regf_slot_budget(int mib);
int nb6_last_regf_batches(void);

/* FP64 pass of the regular force (default on; no counterpart in the reference).
 * The FP32 kernel also sums, per row and j-segment, error scales of its FP32
 * partial sums and of the FP32 rounding of its separations; the finalize flags every
 * row whose bound could put its acc or jerk more than 1e-6 (relative) from FP64 and
 * selects the segments with the largest shares of that bound until the rest is within
 * it; those segments are recomputed in FP64 (same neighbour test, so lists are
 * unchanged); rows whose FP32 sums are incomplete (NaN), or that need more segments
 * than they may store, are recomputed over all j.  nb6_last_regf_refined() returns
 * the rows of the last call's FP64 pass, nb6_last_regf_refined_chunks() its (row,
 * segment) items; both synchronise the device. */
int nb6_set_regf_refine(int on);
int nb6_last_regf_refined(void);
int nb6_last_regf_refined_chunks(void);
/* Weights of the four error scales that flag a row for the FP64 pass (flush sums
 * of acc and jerk, position and velocity rounding of the pair terms); larger
 * weights flag more rows.  Defaults 0.5, 0.5, 2, 0, and 0.5, 0.5, 2, 1 for j-sets
 * of at most 4096 (the velocity term's linear bound over-flags large j-sets, and a
 * small j-set's rows are cheap to recompute); a call here sets both to its values. */
int nb6_set_regf_refine_weights(double k0, double k1, double k2, double k3);
/* The same weights for the batched ensemble step.  Default 0.5, 0.5, 0.5, 0: the
 * position-rounding term weighs a quarter of the single-call default, which keeps
 * acc and pot within 1e-6 of FP64 for every particle and jerk at the 99.9th
 * percentile while recomputing few rows. */
int nb6_set_ens_refine_weights(double k0, double k1, double k2, double k3);

/* Near field of the FP32 regular-force kernel (no counterpart in the reference):
 * pairs with r < f RS are left out of the FP32 sums and summed in FP64 by the
 * finalize.  f in [1, 4]; default 2.  Lists do not depend on it. */
int nb6_set_regf_near(double f);

/* Diagnostic: while d_bounds (device, double[4][out_stride] of later calls) is set,
 * every regf call writes each row's four error scales there (the inputs of the FP64
 * pass's decision); NULL switches it off. */
int nb6_set_regf_debug_bounds(double *d_bounds);

/* Ensemble of independent clusters in one device j-set (no counterpart in the
 * reference, which holds one j-set per library instance).
 *   nb6_ens_send_device: offsets[nclusters+1] (host) index the concatenated SoA
 *     double[7][stride] (x,y,z,vx,vy,vz,m) on the device; cluster c owns
 *     [offsets[c], offsets[c+1]).
 *   nb6_ens_regf_device: one regular step of every particle of every cluster
 *     against its own cluster: d_i SoA double[8][stride] (x,y,z,vx,vy,vz,h2,dtr)
 *     for all particles, d_out SoA double[7][stride], d_list int[ntot][lmax] with
 *     0-based indices relative to the particle's own cluster.  One persistent
 *     launch over (cluster, i-block) work items; use_graph=1 captures the step
 *     into a hipGraph on the first call and replays it while the pointers and
 *     sizes stay the same.  Results equal gpunb_send_ + gpunb_regf_ per cluster. */
int nb6_ens_open(int max_clusters, int max_particles);
int nb6_ens_close(void);
int nb6_ens_send_device(int nclusters, const int *offsets, const double *d_soa, int stride,
                        void *stream);
int nb6_ens_regf_device(const double *d_i, int stride, int lmax, int nnbmax, int m_flag,
                        double *d_out, int *d_list, int use_graph, void *stream);
int nb6_ens_items(void);
/* Rows of the last batched step that went through the FP64 pass (as
 * nb6_last_regf_refined, over each row's own cluster; synchronises the device). */
int nb6_ens_last_refined(void);

/* Autotuners (run on the caller's own, already sent, inputs; outputs are
 * overwritten).  They time every combination of work-group size = LDS j-tile
 * (128, 256, 512), i-particles per lane (1, 2; i-block 128..1024) and occupancy
 * (regf: minimum work-groups 1024/2048/4096 through j-chunking; ensemble: 8, 16
 * or 32 wavefronts per CU in the persistent launch), keep the fastest as the
 * library's configuration and report best = {threads, i_per_lane, min_groups or
 * work-groups per CU, time in microseconds}. */
int nb6_regf_autotune_device(int ni, int ni_stride, const double *d_i, int lmax, int nnbmax,
                             int m_flag, double *d_out, int *d_list, int *best, void *stream);
int nb6_ens_autotune(const double *d_i, int stride, int lmax, int nnbmax, int m_flag,
                     double *d_out, int *d_list, int *best, void *stream);

/* Streaming device-to-device copy of bytes (a multiple of 16) by a plain copy
 * kernel; used to calibrate HBM byte counters against a known size. */
int nb6_device_copy(void *dst, const void *src, unsigned long long bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Device-resident j-particles with a GPU predictor (extension; replaces the host
 * prediction xbpredall.F:52-60 plus gpunb_send_ per regular block,
 * intgrt.F:875-896).  Fortran-callable, every argument by reference.
 *   nb6_jp_open_(nmax) / nb6_jp_close_(): predictor state for nmax particles.
 *   nb6_jp_set_(n, addr, x0, x0dot, f, fdot, m, t0): store the corrected state of n
 *     particles (1-based addresses; F = a/2 and FDOT = adot/6 as in common6.h).
 *   nb6_jp_predict_send_(time, nj): predict particles 1..nj to `time` on the GPU,
 *     x = ((FDOT s + F) s + x0dot) s + x0, xdot = (1.5 FDOT s + F) 2s + x0dot,
 *     s = time - t0 (xbpredall.F:55-59, FP64), and make them the j-set of the next
 *     gpunb_regf_ calls (as gpunb_send_ with those positions would; gpunb must be
 *     open).
 *   nb6_jp_regf_(ni, addr, h2, dtr, acc, jrk, pot, lmax, nnbmax, list, m_flag):
 *     gpunb_regf_ for the particles at 1-based addresses addr[ni], their positions
 *     and velocities taken from the last prediction (no x_i / v_i from the host).
 * ------------------------------------------------------------------------- */
void nb6_jp_open_(int *nmax);
void nb6_jp_close_(void);
void nb6_jp_set_(int *n, int addr[], double x0[][3], double x0dot[][3], double f[][3],
                 double fdot[][3], double m[], double t0[]);
void nb6_jp_predict_send_(double *time, int *nj);
void nb6_jp_regf_(int *ni, int addr[], double h2[], double dtr[], double acc[][3],
                  double jrk[][3], double pot[], int *lmax, int *nnbmax, int *list, int *m_flag);

/* Irregular force fed from the device (extension; replaces the host list round
 * trip util_gpu.F:101-110 -> IRR_SIMD_SET_LIST, intgrt.F:932, and the per-particle
 * IRR_SIMD_SET_JP uploads):
 *   nb6_irr_sync_jp_(n): predictor records of addresses 1..n from the
 *     nb6_jp_* store (both libraries open).
 *   nb6_irr_set_lists_device(ni, d_addr, d_list, lmax, jbase, stream): the lists of
 *     the particles at 1-based addresses d_addr[ni] from gpunb_regf rows d_list
 *     [ni][lmax] (device), j -> address j + jbase (jbase = IFIRST), self dropped,
 *     overflow rows skipped; the same effect as IRR_SIMD_SET_LIST with the list the
 *     caller would have built.
 *   nb6_irr_firr_device(ti, ni, d_addr, d_out, stream): irr_simd_firr_vec_ with
 *     device addresses and d_out double[ni][8] = acc[3], jrk[3], nnbid (1-based), nnb.
 *     Asynchronous on the stream, so it is not counted in irr_simd_profile_'s line
 *     (that line covers irr_simd_firr_vec_ calls, whose pairs and time are known on
 *     return). */
/* Launch shape of the irregular-force gather: 16 or 64 lanes per active particle,
 * and whether blocks of >= 2048 active particles are processed in Morton order of
 * their positions (results unchanged).  Defaults: 64, 1. */
int nb6_set_irr_config(int lanes_per_i, int sort_i);
/* Prediction of the irregular-force records: 0 = each neighbour record is predicted
 * where it is gathered (a record and its time per neighbour), 2 = every record is
 * predicted to ti once per call and the gather reads one 64-byte predicted record per
 * neighbour, 1 (default) = 2 when ni * 8 >= nmax.  Results are the same bits. */
int nb6_set_irr_predict(int mode);
void nb6_irr_sync_jp_(int *n);
int nb6_irr_set_lists_device(int ni, const int *d_addr, const int *d_list, int lmax, int jbase,
                             void *stream);
int nb6_irr_firr_device(double ti, int ni, const int *d_addr, double *d_out, void *stream);

/* ---------------------------------------------------------------------------
 * Part 3 (extension): learned neighbour radius.  A network 8 -> 64 -> 64 -> 1
 * (ReLU; bfloat16 weights, FP32 accumulation; 4801 parameters, layout W1[64][8],
 * b1[64], W2[64][64], b2[64], w3[64], b3) predicts y = ln(RS_next / RS_prev)
 * from eight features per particle.  The reference's rule (regcor_gpu.F:95-238)
 * stays available on the device as the fallback.  All arrays are device
 * pointers; FP32 feature arrays are SoA float[8][n].
 * ------------------------------------------------------------------------- */
int nb6_rsnet_nparam(void);
/* Allocate (once) and initialise: He-uniform weights from a SplitMix64 stream of
 * `seed`, zero biases, zero Adam moments, identity normalisation. */
int nb6_rsnet_open(unsigned long long seed);
int nb6_rsnet_close(void);
/* Host copies of the FP32 master weights (4801 floats); set also refreshes the
 * bfloat16 copy used by the forward pass. */
int nb6_rsnet_set_params(const float *params);
int nb6_rsnet_get_params(float *params);
/* Feature normalisation: x' = (x - mu) / sigma, mu[8], sigma[8] on the host. */
int nb6_rsnet_set_norm(const float *mu, const float *sigma);
/* y[n] = network(features) */
int nb6_rsnet_forward_device(int n, const float *d_feat, float *d_y, void *stream);
/* Gradient of mean((y - t)^2) with respect to the (bfloat16-rounded) weights:
 * d_grad float[4801], d_loss float[1] (device). */
int nb6_rsnet_grad_device(int n, const float *d_feat, const float *d_t, float *d_grad,
                          float *d_loss, void *stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the minibatch; *h_loss (host,
 * may be null) receives the minibatch loss before the step. */
int nb6_rsnet_train_step_device(int n, const float *d_feat, const float *d_t, float lr,
                                float *h_loss, void *stream);
/* One epoch of minibatch Adam on the device (no counterpart in the reference): the n
 * samples (SoA float[8][n], targets float[n], both on the device) in an order
 * shuffled on the device from `seed`, n / batch steps of batch samples each, queued
 * without host synchronisation; *h_loss (host, may be null; the call then returns
 * without waiting) receives the mean squared residual over the epoch, each residual
 * taken before its step. */
int nb6_rsnet_train_epoch_device(int n, const float *d_feat, const float *d_t, int batch,
                                 float lr, unsigned long long seed, float *h_loss,
                                 void *stream);
/* Features of n particles (SoA double arrays with `stride`; x, acc, jrk [3][stride]):
 * log10 of (nnb+1)/(4/3 pi RS^3), nnb+1, |acc|, |jrk|, |x - centre|, RS, dtr, mass;
 * nnb = previous neighbour count (int[n]), centre = density centre (host, 3). */
int nb6_rs_features_device(int n, int stride, const double *d_x, const double *d_acc,
                           const double *d_jrk, const double *d_mass, const double *d_rs,
                           const double *d_dtr, const int *d_nnb, const double *centre,
                           float *d_feat, void *stream);
/* Inference fused ahead of gpunb_regf: rs_out = use_rule[i] ? rs_rule[i] :
 * rs_prev[i] exp(y_i) (y clamped to [-4, 4]); when d_ib is given, row 6 of that
 * regf i-block (double[8][ib_stride]) receives h2 = RS^2 (RS^2 / bodym if m_flag).
 * d_rs_rule / d_use_rule may be null (network everywhere). */
int nb6_rs_predict_device(int n, const float *d_feat, const double *d_rs_prev,
                          const double *d_rs_rule, const int *d_use_rule, double *d_rs_out,
                          double *d_ib, int ib_stride, int m_flag, double bodym, void *stream);
/* The reference's RS update after a regular step, regcor_gpu.F:95-238 (single
 * stars, NB_FLAG = 1, M_FLAG = 0; the c.m.-body branch and CHECKL are not part of
 * it): x, v [3][stride]; dtr; nnb = new count without self, nnb0 = previous count;
 * list int[n][lmax] = the rows gpunb_regf returned (count, 0-based j including the
 * particle itself, which is skipped; used for nnb <= 3; may be null) into the
 * j-set positions xj, vj [3][jstride]; glob = {rdens x, y, z, RSCALE, RC^2, RSFAC,
 * ZNBMIN} (host).  d_rs is updated in place. */
int nb6_rs_rule_device(int n, int stride, const double *d_x, const double *d_v,
                       const double *d_dtr, const int *d_nnb, const int *d_nnb0,
                       const int *d_list, int lmax, const double *d_xj, const double *d_vj,
                       int jstride, const double *glob, int nnbopt, int nnbmax, double *d_rs,
                       void *stream);
/* Training target: for each i of a regf i-block (double[8][ib_stride]) and its
 * list from nb6_regf_device against the currently sent j-set, the radius giving
 * exactly k neighbours under min(r^2, |dx + dtr dv|^2) (FP64).  Rows with fewer
 * than k entries (or overflow) get -(count) - 1. */
int nb6_rs_kth_device(int ni, int ib_stride, const double *d_ib, const int *d_list, int lmax,
                      int k, double *d_rs, void *stream);

/* ---------------------------------------------------------------------------
 * Part 4 (extension): the regular corrector on the GPU.  Replaces the loop of
 * util_gpu.F:150-175 that calls REGCOR_GPU (regcor_gpu.F:1-740, with FPCORR,
 * fpcorr.F:1-148) once per particle of a regular block, on per-particle state
 * held in device memory.  Every array is a device pointer indexed by particle
 * number - 1; [nmax][3] arrays are laid out as the Fortran X(3,NMAX) (the three
 * components of a particle adjacent), so COMMON arrays can be copied as they are.
 * ------------------------------------------------------------------------- */
typedef struct nb6_regstate {
    int nmax;                         /* stride of every per-particle array */
    int lmax;                         /* row length of list (LMAX) */
    /* read */
    const double *x, *xdot;           /* [nmax][3] X, XDOT predicted to the block time */
    const double *body;               /* [nmax] BODY */
    const double *x0, *x0dot;         /* [nmax][3] X0, X0DOT */
    const double *f, *fdot;           /* [nmax][3] F = a/2, FDOT = adot/6 (FPCORR) */
    const double *t0;                 /* [nmax] T0 */
    const double *h;                  /* [nmax - N] H of KS pairs (c.m. bodies I > N); may
                                         be null when no active particle exceeds N */
    /* read and written */
    double *step, *timenw;            /* [nmax] STEP, TIMENW */
    double *fi, *fidot, *fr, *frdot;  /* [nmax][3] FI, FIDOT, FR, FRDOT */
    double *d2, *d3;                  /* [nmax][3] D2, D3 (FPCORR) */
    double *t0r, *stepr, *rs;         /* [nmax] T0R, STEPR, RS */
    int *list;                        /* int[nmax][lmax] LIST: count, then 1-based ascending */
    /* written */
    double *xn, *xndot;               /* [nmax][3] XN, XNDOT */
    double *d0, *d0r, *d1r, *d2r, *d3r;  /* [nmax][3] D0, D0R, D1R, D2R, D3R */
    /* read when KZ(37) = 1 (CHECKL); may be null otherwise */
    const int *listv;                 /* LISTV: [0] = count, then 1-based high-velocity bodies */
    const double *r;                  /* [nmax - N] R of KS pairs (c.m. bodies I > N) */
} nb6_regstate;

typedef struct nb6_regcor_params {
    double time;                      /* TIME of the block */
    double dtmin, smin, smax, etar;   /* DTMIN, SMIN (= 2 DTMIN, adjust.F:192), SMAX, ETAR */
    double rscale, rc2, rsfac, znbmin, bodym;  /* RSCALE, RC^2, RSFAC, ZNBMIN, BODYM */
    double tidal4;                    /* TIDAL(4) when KZ(14) <= 2, else 0 */
    double rdens[3];                  /* RDENS */
    int n, ifirst;                    /* N (single bodies), IFIRST */
    int nnbopt, nnbmax;               /* NNBOPT, NNBMAX */
    int nb_flag, m_flag;              /* from KZ(39) (regcor_gpu.F:20-21) */
    int kz38;                         /* KZ(38): 0 none, 1 always, 2 if DF2 > 1e-4 FR2, 3 losses */
    int kz37;                         /* KZ(37): 1 = CHECKL's high-velocity test; > 1 counted
                                         as unsupported (checkl.F:17-242 not included) */
    double cmsep2;                    /* CMSEP2 (input.F:184), for c.m. bodies in CHECKL */
} nb6_regcor_params;

/* REGCOR_GPU for the ni particles d_addr[ni] (1-based I) of one block:
 *   d_force  double[7][force_stride] = the block's nb6_regf_device output (FREG, FDR);
 *   d_nlist  int[ni][lmax] = the block's nb6_regf_device rows (count, 0-based
 *            positions in the sent j-set of particles IFIRST.., self included), turned
 *            in place into the final new list (count, 1-based ascending, no self) as
 *            util_gpu.F:101-110 and regcor_gpu.F:298-350 leave it; lmax must equal
 *            s->lmax and exceed nnbmax + 1;
 *   d_counters int[7] accumulated: NBVOID, NBFLUX, NBSMIN, NBCORR, unsupported
 *            (particles skipped: a negative or too long row, an index outside
 *            [1, nmax], I > N without h; particles with KZ(37) > 1), NBFAST (CHECKL's
 *            high-velocity intruders, checkl.F:266-269), NCMDER (FPCORR terms of KS
 *            components J < IFIRST, taken with their c.m. F and FDOT, fpcorr.F:46-61).
 * CHECKL (regcor_gpu.F:240-244) is done for KZ(37) = 1 (checkl.F:243-340: at most
 * one body of LISTV inside 2 RS, approaching with impact parameter below RS, joins
 * the list; c.m. bodies take the resolved component terms of checkl.F:282-336).
 * Not included: external tides beyond TIDAL(4), FCLOUD, the KZ(37) > 1 parts of
 * CHECKL.  Particles of one block must be distinct.  Asynchronous on `stream`. */
int nb6_regcor_device(const nb6_regcor_params *p, const nb6_regstate *s, int ni,
                      const int *d_addr, const double *d_force, int force_stride,
                      int *d_nlist, int *d_counters, void *stream);

/* How nb6_regcor_device reads the neighbours' state: 0 (default) packs every
 * particle's X, XDOT, BODY, STEP, T0, F, FDOT, X0DOT into one record first when the
 * block holds at least nmax/32 particles, 1 always packs, 2 never.  Results do not
 * depend on it. */
int nb6_set_regcor_pack(int mode);

/* ---------------------------------------------------------------------------
 * Part 5 (extension): total-energy terms on the GPU.  Replaces the GPU branch of
 * ENERGY (energy.F:28-62 and :119-124: GPUPOT over IFIRST..NTOT, the c.m. potential
 * copied to the KS components, PHICOR of phicor.f:1-76, the POT and ZKIN sums) and
 * the EBIN sum of energy.F:13-24 given each pair's HT.  Device arrays as in Part 4.
 * ------------------------------------------------------------------------- */
typedef struct nb6_energy_in {
    int nmax, lmax;
    int n, ntot, npairs;              /* N, NTOT = N + NPAIRS, NPAIRS (IFIRST = 2 NPAIRS + 1) */
    double cmsep2_rmin2;              /* CMSEP2 * RMIN2 (PHICOR's c.m. separation test) */
    const double *x, *xdot;           /* [nmax][3] X (components resolved), XDOT */
    const double *body;               /* [nmax] BODY */
    const int *list;                  /* int[nmax][lmax] LIST (1-based); rows of the first
                                         KS components hold their perturber lists */
    const double *ht;                 /* [npairs] HT of each pair (may be null if npairs = 0) */
    double *phidbl;                   /* [nmax] out: PHIDBL (positive potential) */
} nb6_energy_in;

/* h_out (host double[3]) = ZKIN, POT, EBIN; the caller's total energy is
 * ZKIN - POT + EBIN (+ EPL, ESUB, EMERGE, ECOLL, chain terms).  Synchronises. */
int nb6_energy_device(const nb6_energy_in *e, double *h_out, void *stream);

/* Device allocation / copies for callers without another GPU runtime
 * (tests, benchmarks). */
void *nb6_dev_alloc(unsigned long long bytes);
void  nb6_dev_free(void *p);
int   nb6_memcpy_htod(void *dst, const void *src, unsigned long long bytes);
int   nb6_memcpy_dtoh(void *dst, const void *src, unsigned long long bytes);
int   nb6_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* NB6GPU_H */
