// Checks the library's host fork-join helpers (csrc/host_threads.h) on the CPU,
// built under ThreadSanitizer or AddressSanitizer + UndefinedBehaviorSanitizer by
// tests/test_host_threads_cpu.py: the ranges cover [0, n) once, the threaded j pack
// equals the one-thread pack record for record, and a NaN in any range is reported.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host_threads.h"

static int bad(const char *what) {
    fprintf(stderr, "host_threads_check: FAILED: %s\n", what);
    return 1;
}

int main() {
    using namespace nb6;
    // every index written once, by exactly one range, for several n and thread counts
    for (int n : {0, 3, 4097, 100003}) {
        for (int nt : {1, 3, 8}) {
            std::vector<int> hit(n, 0);
            std::vector<long long> sum(nt > 0 ? nt : 1, 0);
            run_ranges(nt, n, [&](int t, int a, int b) {
                for (int i = a; i < b; i++) { hit[i] += 1; sum[t] += i; }
            });
            long long s = 0;
            for (long long x : sum) s += x;
            for (int i = 0; i < n; i++)
                if (hit[i] != 1) return bad("run_ranges: an index not written exactly once");
            if (s != (long long)n * (n - 1) / 2) return bad("run_ranges: wrong range sums");
        }
    }
    const int n = 1 << 18;
    std::vector<double> out(n, 0.0);
    par_for(n, [&](int a, int b) {
        for (int i = a; i < b; i++) out[i] = std::sqrt((double)i);
    });
    for (int i = 0; i < n; i++)
        if (out[i] != std::sqrt((double)i)) return bad("par_for: wrong value");

    // the j pack: threaded against one thread
    std::vector<double> m(n), x(3 * (size_t)n), v(3 * (size_t)n);
    for (int j = 0; j < n; j++) {
        m[j] = 1.0 + (j % 97) * 1e-3;
        for (int k = 0; k < 3; k++) {
            x[3 * (size_t)j + k] = std::sin(0.37 * j + k) * (1 + j % 13);
            v[3 * (size_t)j + k] = std::cos(0.11 * j - k);
        }
    }
    const auto *xj = reinterpret_cast<const double(*)[3]>(x.data());
    const auto *vj = reinterpret_cast<const double(*)[3]>(v.data());
    std::vector<double> h1(8 * (size_t)n), h8(8 * (size_t)n);
    double b1[6], b8[6], m1 = 0, m8 = 0;
    if (!pack_j_records(n, 1, m.data(), xj, vj, h1.data(), b1, &m1)) return bad("pack: NaN on one thread");
    if (!pack_j_records(n, 8, m.data(), xj, vj, h8.data(), b8, &m8)) return bad("pack: NaN on 8 threads");
    if (std::memcmp(h1.data(), h8.data(), h1.size() * sizeof(double)) != 0) return bad("pack: records differ");
    if (std::memcmp(b1, b8, sizeof b1) != 0 || m1 != m8) return bad("pack: box or mass differ");
    if (m1 != 1.0 + 96e-3) return bad("pack: largest mass");
    for (int j : {0, n / 2, n - 1}) {        // a NaN in the first, a middle and the last range
        std::vector<double> xn(x);
        xn[3 * (size_t)j + 1] = std::nan("");
        double bx[6], mm;
        if (pack_j_records(n, 8, m.data(), reinterpret_cast<const double(*)[3]>(xn.data()), vj,
                           h8.data(), bx, &mm))
            return bad("pack: NaN not reported");
    }
    printf("host_threads_check: ok\n");
    return 0;
}
