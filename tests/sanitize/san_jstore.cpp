// Thread-sanitizer harness (CPU only) for set_jp's host side
// (hermite_amd/csrc/hh_jstore.hpp, the code hh_set_jp runs): threads write
// distinct slots of one j-store concurrently and raise the shared dirty
// bound, as the reference driver's OpenMP loop does (SRC/nbodysystem.h:
// 753-756).  std::thread rather than OpenMP, whose runtime is not built with
// the sanitizer.  Checks the record layout and that the bound ends at the
// highest slot written.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "hh_jstore.hpp"

int main()
{
    int fails = 0;
    for (int order : {4, 6, 8, 9}) {
        const int ps = order == 4 ? 24 : order == 6 ? 28 : order == 8 ? 36 : 40;
        const int jw = order == 4 ? 14 : order == 6 ? 20 : order == 8 ? 26 : 29;
        const int n = 4000, nthreads = 8;
        std::vector<double> recs((size_t)n * ps), jst((size_t)n * jw, -1.0);
        for (size_t k = 0; k < recs.size(); k++) recs[k] = (double)k;
        std::atomic<int> dirty{0};
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; t++)
            th.emplace_back([&, t] {
                for (int i = t; i < n; i += nthreads) {  // distinct slots per thread
                    hh::jstore_write(order, jst.data(), i, &recs[(size_t)i * ps]);
                    hh::dirty_raise(dirty, i + 1);
                }
            });
        for (auto &x : th) x.join();
        if (dirty.load() != n) fails++;
        for (int i = 0; i < n; i++) {
            const double *p = &recs[(size_t)i * ps], *d = &jst[(size_t)i * jw];
            if (d[3] != p[1] || d[7] != p[2] || d[0] != p[4]) fails++;
            if (order == 9 ? (d[4] != p[10] || d[26] != p[7]) : d[4] != p[7]) fails++;
        }
    }
    std::printf("san_jstore: %d failed checks\n", fails);
    return fails ? 1 : 0;
}
