// Host-side fork-join helpers of the C-ABI layer: plain C++ (no HIP), so that the
// exact threading of the j pack and of the list assembly also builds and runs
// under ThreadSanitizer on a machine without a GPU.
#pragma once

#include <algorithm>
#include <thread>
#include <vector>

namespace nb6 {

// Run f(t, a, b) for t in [0, nt) over the contiguous ranges [a, b) that split
// [0, n), one std::thread each; nt <= 1 runs f(0, 0, n) on the calling thread.  A
// thread that cannot be created (std::system_error under a process or thread
// limit, or std::bad_alloc) leaves its range, and every later one, to the calling
// thread; every created thread is joined before this returns, so no joinable
// std::thread is destroyed and every range has run.
template <class F>
inline void run_ranges(int nt, int n, F f) {
    if (nt <= 1) {
        f(0, 0, n);
        return;
    }
    auto range = [&](int t) {
        f(t, (int)((long long)n * t / nt), (int)((long long)n * (t + 1) / nt));
    };
    std::vector<std::thread> th;
    int t = 0;
    try {
        th.reserve(nt);
        for (; t < nt; t++) th.emplace_back(range, t);
    } catch (...) {
    }
    for (int r = t; r < nt; r++) range(r);
    for (auto &x : th) x.join();
}

// A host loop over [0, n) in ranges f(a, b): up to 8 threads above 4,096 items.
template <class F>
inline void par_for(int n, F f) {
    const int nt = n > 4096 ? std::min(8, (int)std::thread::hardware_concurrency()) : 1;
    run_ranges(nt, n, [&](int, int a, int b) { f(a, b); });
}

// gpunb_send_'s host pass: AoS records (x y z m vx vy vz 0) of the caller's j
// arrays into h[8 n], the bounding box (lo[3], hi[3]) and the largest mass, in one
// pass over nt ranges.  Returns false when any value is NaN (box and mass are then
// unspecified).
inline bool pack_j_records(int n, int nt, const double *mj, const double (*xj)[3],
                           const double (*vj)[3], double *h, double box[6], double *mmax) {
    nt = std::max(1, nt);
    std::vector<double> part((size_t)nt * 7);
    std::vector<char> nan(nt, 0);
    run_ranges(nt, n, [&](int t, int a, int b) {
        double bx[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
        double mm = 0.0;
        bool bad = false;
        for (int j = a; j < b; j++) {
            double *r = h + (size_t)8 * j;
            r[0] = xj[j][0]; r[1] = xj[j][1]; r[2] = xj[j][2]; r[3] = mj[j];
            r[4] = vj[j][0]; r[5] = vj[j][1]; r[6] = vj[j][2]; r[7] = 0.0;
            for (int k = 0; k < 7; k++) bad = bad || !(r[k] == r[k]);
            for (int k = 0; k < 3; k++) {
                bx[k] = r[k] < bx[k] ? r[k] : bx[k];
                bx[3 + k] = r[k] > bx[3 + k] ? r[k] : bx[3 + k];
            }
            mm = mj[j] > mm ? mj[j] : mm;
        }
        for (int k = 0; k < 6; k++) part[(size_t)t * 7 + k] = bx[k];
        part[(size_t)t * 7 + 6] = mm;
        nan[t] = bad;
    });
    for (int k = 0; k < 3; k++) { box[k] = 1e300; box[3 + k] = -1e300; }
    *mmax = 0.0;
    for (int t = 0; t < nt; t++) {
        if (nan[t]) return false;
        for (int k = 0; k < 3; k++) {
            box[k] = std::min(box[k], part[(size_t)t * 7 + k]);
            box[3 + k] = std::max(box[3 + k], part[(size_t)t * 7 + 3 + k]);
        }
        *mmax = std::max(*mmax, part[(size_t)t * 7 + 6]);
    }
    return true;
}

}  // namespace nb6
