// Host-side sanitizer driver for libnb6gpu.so: drives the library's threaded host
// paths through the Fortran-facing entry points with sizes that switch the
// threads on (the j pack of gpunb_send_ above 65,536 particles, the list assembly
// and unpacking of gpunb_regf_ above 4,096 i, the packing of gpupot_), plus the
// irregular-force logs (irr_simd_set_jp_ / set_list_ / firr_vec_).  Built with the
// library's host code instrumented by AddressSanitizer + UndefinedBehaviorSanitizer
// or by ThreadSanitizer (Makefile beside this file); a sanitizer report fails the
// run.  The results are checked for plausibility only (finite forces, counts in
// range): their accuracy is covered by the parity tests.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "nb6gpu.h"

static int fail(const char *what) {
    fprintf(stderr, "san_driver: FAILED: %s (%s)\n", what, nb6_last_error_message());
    return 1;
}

int main() {
    nb6_set_abort_on_error(0);
    const int n = 131072, ni = 8192, lmax = 128, nnbmax = 100, rank = 0;
    std::mt19937_64 rng(20261203);
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n);
    for (auto &a : x) a = g(rng);
    for (auto &a : v) a = 0.5 * g(rng);
    int nb = n, nj = n, r = rank;
    gpunb_open_(&nb, &r);
    if (nb6_last_error()) return fail("gpunb_open_");
    for (int rep = 0; rep < 2; rep++) {
        gpunb_send_(&nj, m.data(), reinterpret_cast<double(*)[3]>(x.data()),
                    reinterpret_cast<double(*)[3]>(v.data()));
        if (nb6_last_error()) return fail("gpunb_send_");
        std::vector<double> h2(ni, 0.01), dtr(ni, 1e-3), acc(3 * ni), jrk(3 * ni), pot(ni);
        std::vector<int> list((size_t)ni * lmax);
        int nii = ni, lm = lmax, nm = nnbmax, mf = 0;
        gpunb_regf_(&nii, h2.data(), dtr.data(), reinterpret_cast<double(*)[3]>(x.data()),
                    reinterpret_cast<double(*)[3]>(v.data()), reinterpret_cast<double(*)[3]>(acc.data()),
                    reinterpret_cast<double(*)[3]>(jrk.data()), pot.data(), &lm, &nm, list.data(), &mf);
        if (nb6_last_error()) return fail("gpunb_regf_");
        for (int i = 0; i < ni; i++) {
            const int c = list[(size_t)i * lmax];
            if (!std::isfinite(acc[3 * i]) || !std::isfinite(pot[i]) || c > lmax - 1 || c < -n)
                return fail("gpunb_regf_ result out of range");
        }
    }
    gpunb_profile_(&r);
    gpunb_close_();
    if (nb6_last_error()) return fail("gpunb_close_");

    int i0 = 1, np = 8192, nall = 8192;
    std::vector<double> pot(np);
    gpupot_(&r, &i0, &np, &nall, m.data(), reinterpret_cast<double(*)[3]>(x.data()), pot.data());
    if (nb6_last_error()) return fail("gpupot_");
    for (int k = 0; k < np; k++)
        if (!std::isfinite(pot[k]) || pot[k] <= 0.0) return fail("gpupot_ result out of range");

    const int nirr = 16384, lirr = 64;
    int nmax = nirr, lm2 = lirr;
    irr_simd_open_(&nmax, &lm2, &r);
    if (nb6_last_error()) return fail("irr_simd_open_");
    for (int a = 1; a <= nirr; a++) {
        double acc2[3] = {0.0, 0.0, 0.0}, jrk6[3] = {0.0, 0.0, 0.0}, mass = 1.0 / nirr, t = 0.0;
        irr_simd_set_jp_(&a, &x[3 * (size_t)(a - 1)], &v[3 * (size_t)(a - 1)], acc2, jrk6, &mass, &t);
    }
    std::vector<int> nbl(lirr);
    for (int a = 1; a <= nirr; a++) {
        nbl[0] = 20;
        for (int k = 1; k <= 20; k++) nbl[k] = 1 + (a - 1 + 37 * k) % nirr;
        irr_simd_set_list_(&a, nbl.data());
    }
    std::vector<int> addr(4096), nnbid(4096);
    std::vector<double> ai(3 * 4096), ji(3 * 4096);
    for (int k = 0; k < 4096; k++) addr[k] = 1 + 3 * k;
    int nq = 4096;
    double ti = 1e-3;
    for (int rep = 0; rep < 2; rep++) {
        irr_simd_firr_vec_(&ti, &nq, addr.data(), reinterpret_cast<double(*)[3]>(ai.data()),
                           reinterpret_cast<double(*)[3]>(ji.data()), nnbid.data());
        if (nb6_last_error()) return fail("irr_simd_firr_vec_");
        for (int k = 0; k < nq; k++)
            if (!std::isfinite(ai[3 * k]) || nnbid[k] < 1 || nnbid[k] > nirr)
                return fail("irr_simd_firr_vec_ result out of range");
        // half the records change between the calls (the logs are replayed again)
        for (int a = 1; a <= nirr; a += 2) {
            double acc2[3] = {0.0, 0.0, 0.0}, jrk6[3] = {0.0, 0.0, 0.0}, mass = 1.0 / nirr, t = 0.0;
            irr_simd_set_jp_(&a, &x[3 * (size_t)a], &v[3 * (size_t)a], acc2, jrk6, &mass, &t);
        }
    }
    irr_simd_profile_(&r);
    irr_simd_close_(&r);
    printf("san_driver: ok\n");
    fflush(stdout);
    // leave without the exit-time teardown of the HIP and HSA runtimes: under
    // AddressSanitizer that teardown (libhsa-runtime64 freeing through the sanitizer's
    // device allocator after the device runtime is marked unloaded) stops on an
    // internal CHECK of the sanitizer runtime, outside this library's code
    std::_Exit(0);
}
