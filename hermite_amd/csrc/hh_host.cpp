// Host-side input generation for the Hermite back end: the Aarseth-Henon-
// Wielen Plummer sampler of the reference's RUN/gen-plum.c, its random
// number source, and the Salpeter mass function used by the large workloads.
//
// Compiled with -ffp-contract=off and no -march flags, so that with the
// reference's seed (19640916) and truncation radius (100) the sample is
// bit-identical to the snapshot gen-plum writes (checked by
// tests/test_golden.py::test_generator_reproduces_gen_plum against a
// gen-plum built from its source).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// glibc's default rand(): the additive feedback generator of random_r's
// TYPE_3 state (degree 31, separation 3).  Seeding fills r[0..30] with the
// Lehmer sequence r[i] = 16807 r[i-1] mod (2^31 - 1), then
// r[i] = r[i-31] + r[i-3] (mod 2^32) and the first 310 outputs are
// discarded; each output is r[i] >> 1.
struct GlibcRand {
    uint32_t ring[31];  // the last 31 values; f indexes r[i-3], b r[i-31]
    int f = 3, b = 0;
    explicit GlibcRand(uint32_t seed)
    {
        if (seed == 0) seed = 1;
        int32_t x = (int32_t)seed;
        ring[0] = (uint32_t)x;
        for (int i = 1; i < 31; i++) {
            // Schrage's method for 16807 * x mod (2^31 - 1) without overflow
            const int32_t hi = x / 127773, lo = x % 127773;
            int32_t w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            x = w;
            ring[i] = (uint32_t)x;
        }
        f = 3;
        b = 0;
        for (int i = 0; i < 310; i++) next_raw();
    }
    uint32_t next_raw()
    {
        ring[f] += ring[b];
        const uint32_t out = ring[f] >> 1;
        if (++f >= 31) f = 0;
        if (++b >= 31) b = 0;
        return out;
    }
    int next() { return (int)next_raw(); }
};

}  // namespace

extern "C" {

// First n outputs of glibc rand() after srand(seed).
void hh_glibc_rand(unsigned seed, int n, int *out)
{
    GlibcRand g(seed);
    for (int i = 0; i < n; i++) out[i] = g.next();
}

// Plummer sphere by the Aarseth-Henon-Wielen method, RUN/gen-plum.c:116-181:
// radius from the inverse cumulative mass X1^(-2/3) - 1 = R^-2, rejected
// unless R < rmax (gen-plum.c:123), isotropic direction, speed by von
// Neumann rejection on q^2 (1-q^2)^(7/2) (gen-plum.c:135-138), scaling by
// conv = 3 pi / 16 to Henon units (gen-plum.c:104,150-151), equal masses
// 1/n, then the centre-of-mass correction (gen-plum.c:187-226).
// pos and vel are n x 3, row-major.
void hh_plummer(int n, unsigned seed, double rmax, double *mass, double *pos, double *vel)
{
    const double pi = 3.141592653589793238462643;
    const double conv = 3.0 * pi / 16.0;
    const double M = 1.0;
    GlibcRand g(seed);
    auto u01 = [&g]() { return (double)(g.next() / (double)2147483647); };
    int i = 0;
    while (i < n) {
        const double X1 = u01(), X2 = u01(), X3 = u01();
        const double R = 1.0 / std::sqrt((std::pow(X1, -2.0 / 3.0) - 1.0));
        if (!(R < rmax)) continue;
        const double Z = (1.0 - 2.0 * X2) * R;
        const double X = std::sqrt(R * R - Z * Z) * std::cos(2.0 * M_PI * X3);
        const double Y = std::sqrt(R * R - Z * Z) * std::sin(2.0 * M_PI * X3);
        const double Ve = std::sqrt(2.0) * std::pow((1.0 + R * R), -0.25);
        double X4 = 0.0, X5 = 0.0;
        while (0.1 * X5 >= X4 * X4 * std::pow((1.0 - X4 * X4), 3.5)) {
            X4 = u01();
            X5 = u01();
        }
        const double V = Ve * X4;
        const double X6 = u01(), X7 = u01();
        const double Vz = (1.0 - 2.0 * X6) * V;
        const double Vx = std::sqrt(V * V - Vz * Vz) * std::cos(2.0 * M_PI * X7);
        const double Vy = std::sqrt(V * V - Vz * Vz) * std::sin(2.0 * M_PI * X7);
        mass[i] = M / n;
        pos[3 * i + 0] = X * conv;
        pos[3 * i + 1] = Y * conv;
        pos[3 * i + 2] = Z * conv;
        vel[3 * i + 0] = Vx / std::sqrt(conv);
        vel[3 * i + 1] = Vy / std::sqrt(conv);
        vel[3 * i + 2] = Vz / std::sqrt(conv);
        i++;
    }
    double mcm = 0.0, xcm[3] = {0, 0, 0}, vcm[3] = {0, 0, 0};
    for (int k = 0; k < n; k++) {
        mcm += mass[k];
        for (int d = 0; d < 3; d++) {
            xcm[d] += mass[k] * pos[3 * k + d];
            vcm[d] += mass[k] * vel[3 * k + d];
        }
    }
    for (int d = 0; d < 3; d++) {
        xcm[d] /= mcm;
        vcm[d] /= mcm;
    }
    for (int k = 0; k < n; k++)
        for (int d = 0; d < 3; d++) {
            pos[3 * k + d] -= xcm[d];
            vel[3 * k + d] -= vcm[d];
        }
}

// Salpeter masses: inverse CDF of dN/dm ~ m^-alpha on [mlo, mhi], drawn with
// the same generator from seed, then rescaled to total mass 1.
void hh_salpeter(int n, unsigned seed, double alpha, double mlo, double mhi, double *mass)
{
    GlibcRand g(seed);
    const double e = 1.0 - alpha;
    const double a = std::pow(mlo, e), b = std::pow(mhi, e);
    double tot = 0.0;
    for (int i = 0; i < n; i++) {
        const double u = g.next() / 2147483648.0;
        mass[i] = std::pow(a + u * (b - a), 1.0 / e);
        tot += mass[i];
    }
    for (int i = 0; i < n; i++) mass[i] /= tot;
}

}  // extern "C"
