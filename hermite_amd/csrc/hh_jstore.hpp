// Host side of set_jp: the field copy of one corrected Particle record into
// the pinned j-store and the lock-free bound of the slots written since the
// last upload.  Pure host code (no HIP call), so OpenMP threads may call it
// concurrently for distinct slots, as the reference driver does
// (SRC/nbodysystem.h:753-756).  Shared by hh_gravity.hip and the
// thread-sanitizer harness of the tests.
#pragma once
#include <atomic>
#include <cstddef>

namespace hh {

// set_jp: SRC/hermite6-titan.h:70-97 (field copy into the j-store record
// pos[3] mass vel[3] tlast acc jrk [snp crk [d4a d5a]] of jw words).  The
// double-double Particle (SRC/hermite8dd.h:44-78) has posL after posH; its
// j-store record is the 8th-order one followed by posL.
inline void jstore_write(int order, double *jst, int addr, const double *p)
{
    if (order == 9) {
        double *d = jst + (size_t)addr * 29;
        d[0] = p[4];  d[1] = p[5];  d[2] = p[6];     // posH
        d[3] = p[1];                                 // mass
        d[4] = p[10]; d[5] = p[11]; d[6] = p[12];    // vel
        d[7] = p[2];                                 // tlast
        for (int w = 0; w < 18; w++) d[8 + w] = p[13 + w];  // acc jrk snp crk d4a d5a
        d[26] = p[7]; d[27] = p[8]; d[28] = p[9];    // posL
        return;
    }
    const int jw = order == 4 ? 14 : order == 6 ? 20 : 26;
    double *d = jst + (size_t)addr * jw;
    d[0] = p[4];  d[1] = p[5];  d[2] = p[6];   // pos
    d[3] = p[1];                               // mass
    d[4] = p[7];  d[5] = p[8];  d[6] = p[9];   // vel
    d[7] = p[2];                               // tlast
    for (int w = 0; w < 6; w++) d[8 + w] = p[10 + w];                     // acc, jrk
    if (order >= 6) for (int w = 0; w < 6; w++) d[14 + w] = p[16 + w];  // snp, crk
    if (order == 8) for (int w = 0; w < 6; w++) d[20 + w] = p[22 + w];  // d4a, d5a
}

// Raises the dirty bound to hi (a lock-free maximum).
inline void dirty_raise(std::atomic<int> &dirty_hi, int hi)
{
    int cur = dirty_hi.load(std::memory_order_relaxed);
    while (cur < hi && !dirty_hi.compare_exchange_weak(cur, hi, std::memory_order_relaxed)) {
    }
}

}  // namespace hh
