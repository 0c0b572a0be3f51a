// Block-timestep driver on the gfx950 back end.
//
// The caller side of the force + predictor path: the reference's
// NbodySystem (SRC/nbodysystem.h:253-970, TITAN flow without FAST_OMP_SYNC)
// with its per-order Particle physics (SRC/hermite{4,6,8}.h: corrector,
// derivative reconstruction, quantised Aarseth step), driving the HIP
// back end through the C-ABI of include/hermite_hip.h.  Floating-point
// expressions keep the reference's operation order and this file is
// compiled without FMA contraction, so with the HH_EXACT variant a run is
// bit-identical to the reference's CPU build.
//
// Exposed to tests, the benchmark and hermite_run through extern "C" hhd_*.
#include "hermite_hip.h"

#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

// 3-vector with the operator semantics of the reference's dvec3
// (SRC/vector3.h:105-150): a*b is the dot product ((x*x' + y*y') + z*z'),
// v/s multiplies by the rounded reciprocal 1/s.
struct V3 {
    double x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit V3(const double *p) : x(p[0]), y(p[1]), z(p[2]) {}
    void store(double *p) const { p[0] = x; p[1] = y; p[2] = z; }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(double s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(double s) const { const double r = 1.0 / s; return *this * r; }
    double operator*(const V3 &o) const { return x * o.x + y * o.y + z * o.z; }
    double norm2() const { return (*this) * (*this); }
};
inline V3 operator*(double s, const V3 &v) { return v * s; }

// Particle record words (the reference Particle, SRC/hermite8.h:31-45)
enum { ID = 0, MASS = 1, TLAST = 2, DT = 3, POS = 4, VEL = 7, ACC = 10, JRK = 13, SNP = 16, CRK = 19,
       D4A = 22, D5A = 25, D6A = 28, D7A = 31 };
// Order code 9: the double-double 8th-order build (SRC/hermite8dd.h), whose
// Particle has posL after posH and is 40 words long (SRC/hermite8dd.h:44-78).
enum { DD_POSH = 4, DD_POSL = 7, DD_VEL = 10, DD_ACC = 13, DD_JRK = 16, DD_SNP = 19, DD_CRK = 22,
       DD_D4A = 25, DD_D5A = 28, DD_D6A = 31, DD_D7A = 34 };
constexpr int pwords(int o) { return o == 4 ? 24 : o == 6 ? 28 : o == 8 ? 36 : 40; }
constexpr int fwords(int o) { return o == 4 ? 6 : o == 6 ? 9 : 12; }
constexpr int vel_word(int o) { return o == 9 ? int(DD_VEL) : int(VEL); }
constexpr bool valid_order(int o) { return o == 4 || o == 6 || o == 8 || o == 9; }

// pow_one_nth_quant<N>: 2^floor(log2(x)/N) from the exponent bits
// (SRC/pownth.h:3-19).
template <int N>
double pow_nth_quant(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, 8);
    uint64_t e = b >> 52;
    e = (e - (uint64_t)(1023 % N)) / N + (uint64_t)(1023 - 1023 / N);
    b = e << 52;
    std::memcpy(&x, &b, 8);
    return x;
}

// h = dt/2 and 1/h for a power-of-two dt by exponent arithmetic
// (SRC/pownth.h:21-34).
inline void half_and_inverse(double dt, double &h, double &hinv)
{
    uint64_t b;
    std::memcpy(&b, &dt, 8);
    b -= (uint64_t)1 << 52;
    std::memcpy(&h, &b, 8);
    b = ((uint64_t)2046 << 52) - b;
    std::memcpy(&hinv, &b, 8);
}

template <int ORDER>
struct Phys;

// 4th order: SRC/hermite4.h:67-153
template <>
struct Phys<4> {
    static void assign(double *p, const double *f) { V3(f).store(p + ACC); V3(f + 3).store(p + JRK); }
    static void init_dt(double *p, double eta_s, double dtmax)
    {
        const double dtnat = eta_s * std::sqrt(V3(p + ACC).norm2() / V3(p + JRK).norm2());
        double dt = 0.25 * dtmax;
        while (dt > dtnat) dt *= 0.5;
        p[DT] = dt;
    }
    static double step(const double *p, double etapow)
    {
        const double s0 = V3(p + ACC).norm2(), s1 = V3(p + JRK).norm2();
        const double s2 = V3(p + SNP).norm2(), s3 = V3(p + CRK).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s1 * s3) + s2;
        return pow_nth_quant<2>(etapow * (u / l));
    }
    static void correct(double *p, const double *f, double etapow, double dtlim)
    {
        double h, hinv;
        half_and_inverse(p[DT], h, hinv);
        const V3 fa(f), fj(f + 3), acc(p + ACC), jrk(p + JRK), vel(p + VEL);
        const V3 Ap = fa + acc, Am = fa - acc;
        const V3 Jp = (fj + jrk) * h, Jm = (fj - jrk) * h;
        const V3 vel1 = vel + h * (Ap - (1. / 3.) * Jm);
        const V3 pos = V3(p + POS) + h * ((vel + vel1) + h * (-(1. / 3.) * Am));
        pos.store(p + POS);
        vel1.store(p + VEL);
        fa.store(p + ACC);
        fj.store(p + JRK);
        p[TLAST] += p[DT];
        V3 snp = (0.5 * hinv * hinv) * Jm;
        const V3 crk = (1.5 * hinv * hinv * hinv) * (Jp - Am);
        snp = snp + h * crk;
        snp.store(p + SNP);
        crk.store(p + CRK);
        const double dtq = step(p, etapow);
        p[DT] = dtq < dtlim ? dtq : dtlim;
    }
};

// 6th order: SRC/hermite6.h:75-184
template <>
struct Phys<6> {
    static void assign(double *p, const double *f)
    {
        V3(f).store(p + ACC); V3(f + 3).store(p + JRK); V3(f + 6).store(p + SNP);
    }
    static void init_dt(double *p, double eta_s, double dtmax)
    {
        const double dtnat = eta_s * std::sqrt(std::sqrt(V3(p + ACC).norm2() / V3(p + SNP).norm2()));
        double dt = 0.25 * dtmax;
        while (dt > dtnat) dt *= 0.5;
        p[DT] = dt;
    }
    static double step(const double *p, double etapow)
    {
        const double s0 = V3(p + ACC).norm2(), s1 = V3(p + JRK).norm2(), s2 = V3(p + SNP).norm2();
        const double s3 = V3(p + CRK).norm2(), s4 = V3(p + D4A).norm2(), s5 = V3(p + D5A).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s3 * s5) + s4;
        return pow_nth_quant<6>(etapow * (u / l));
    }
    static void correct(double *p, const double *f, double etapow, double dtlim)
    {
        double h, hinv;
        half_and_inverse(p[DT], h, hinv);
        const V3 fa(f), fj(f + 3), fs(f + 6), acc(p + ACC), jrk(p + JRK), snp(p + SNP), vel(p + VEL);
        const V3 Ap = fa + acc, Am = fa - acc;
        const V3 Jp = (fj + jrk) * h, Jm = (fj - jrk) * h;
        const V3 Sp = (fs + snp) * (h * h), Sm = (fs - snp) * (h * h);
        const V3 vel1 = vel + h * (Ap - 0.4 * Jm + (1. / 15) * Sp);
        const V3 pos = V3(p + POS) + h * ((vel + vel1) + h * (-0.4 * Am + (1. / 15) * Jp));
        pos.store(p + POS);
        vel1.store(p + VEL);
        fa.store(p + ACC);
        fj.store(p + JRK);
        fs.store(p + SNP);
        p[TLAST] += p[DT];
        const double hinv2 = hinv * hinv, hinv3 = hinv2 * hinv, hinv4 = hinv2 * hinv2, hinv5 = hinv2 * hinv3;
        V3 crk = (0.75 * hinv3) * (-5. * Am + 5. * Jp - Sm);
        V3 d4a = (1.5 * hinv4) * (-Jm + Sp);
        const V3 d5a = (7.5 * hinv5) * (3 * Am - 3 * Jp + Sm);
        crk = crk + h * (d4a + 0.5 * h * d5a);
        d4a = d4a + h * d5a;
        crk.store(p + CRK);
        d4a.store(p + D4A);
        d5a.store(p + D5A);
        const double dtq = step(p, etapow);
        p[DT] = dtq < dtlim ? dtq : dtlim;
    }
};

// 8th order: SRC/hermite8.h:91-222
template <>
struct Phys<8> {
    static void assign(double *p, const double *f)
    {
        V3(f).store(p + ACC); V3(f + 3).store(p + JRK); V3(f + 6).store(p + SNP); V3(f + 9).store(p + CRK);
    }
    static void init_dt(double *p, double eta_s, double dtmax)
    {
        const double s0 = V3(p + ACC).norm2(), s1 = V3(p + JRK).norm2();
        const double s2 = V3(p + SNP).norm2(), s3 = V3(p + CRK).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s1 * s3) + s2;
        const double dtnat = eta_s * std::sqrt(u / l);
        double dt = 0.25 * dtmax;
        while (dt > dtnat) dt *= 0.5;
        p[DT] = dt;
    }
    static double step(const double *p, double etapow)
    {
        const double s0 = V3(p + ACC).norm2(), s1 = V3(p + JRK).norm2(), s2 = V3(p + SNP).norm2();
        const double s5 = V3(p + D5A).norm2(), s6 = V3(p + D6A).norm2(), s7 = V3(p + D7A).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s5 * s7) + s6;
        return pow_nth_quant<10>(etapow * (u / l));
    }
    static void correct(double *p, const double *f, double etapow, double dtlim)
    {
        double h, hinv;
        half_and_inverse(p[DT], h, hinv);
        const V3 fa(f), fj(f + 3), fs(f + 6), fc(f + 9);
        const V3 acc(p + ACC), jrk(p + JRK), snp(p + SNP), crk(p + CRK), vel(p + VEL);
        const V3 Ap = fa + acc, Am = fa - acc;
        const V3 Jp = (fj + jrk) * h, Jm = (fj - jrk) * h;
        const V3 Sp = (fs + snp) * (h * h), Sm = (fs - snp) * (h * h);
        const V3 Cp = (fc + crk) * (h * h * h), Cm = (fc - crk) * (h * h * h);
        const V3 vel1 = vel + h * (Ap + (-3. / 7.) * Jm + (2. / 21) * Sp - (1. / 105.) * Cm);
        const V3 pos = V3(p + POS) + h * ((vel + vel1) + h * ((-3. / 7.) * Am + (2. / 21) * Jp - (1. / 105.) * Sm));
        pos.store(p + POS);
        vel1.store(p + VEL);
        fa.store(p + ACC);
        fj.store(p + JRK);
        fs.store(p + SNP);
        fc.store(p + CRK);
        p[TLAST] += p[DT];
        const double hinv2 = hinv * hinv, hinv3 = hinv2 * hinv, hinv4 = hinv2 * hinv2;
        const double hinv5 = hinv2 * hinv3, hinv6 = hinv3 * hinv3, hinv7 = hinv3 * hinv4;
        V3 d4a = (hinv4 * 24. / 32.) * (-5. * Jm + 5. * Sp - Cm);
        V3 d5a = (hinv5 * 120. / 32.) * (21. * Am - 21. * Jp + 8. * Sm - Cp);
        V3 d6a = (hinv6 * 720. / 32.) * (Jm - Sp + Cm / 3.);
        const V3 d7a = (hinv7 * 5040. / 32.) * (-5. * Am + 5. * Jp - 2. * Sm + Cp / 3.);
        const double h2 = 0.5 * h, h3 = (1. / 3.) * h;
        d4a = d4a + h * (d5a + h2 * (d6a + h3 * d7a));
        d5a = d5a + h * (d6a + h2 * d7a);
        d6a = d6a + h * d7a;
        d4a.store(p + D4A);
        d5a.store(p + D5A);
        d6a.store(p + D6A);
        d7a.store(p + D7A);
        const double dtq = step(p, etapow);
        p[DT] = dtq < dtlim ? dtq : dtlim;
    }
};

// Profile phases in the order of the reference's Profile (SRC/nbodysystem.h:58-71)
// Double-double 8th order: SRC/hermite8dd.h:119-239.  The 8th-order
// formulas on the shifted offsets; the corrector adds the position increment
// to posH + posL with dd_accumulate (SRC/hermite8dd.h:4-24).
template <>
struct Phys<9> {
    static void assign(double *p, const double *f)
    {
        V3(f).store(p + DD_ACC); V3(f + 3).store(p + DD_JRK); V3(f + 6).store(p + DD_SNP); V3(f + 9).store(p + DD_CRK);
    }
    static void init_dt(double *p, double eta_s, double dtmax)
    {
        const double s0 = V3(p + DD_ACC).norm2(), s1 = V3(p + DD_JRK).norm2();
        const double s2 = V3(p + DD_SNP).norm2(), s3 = V3(p + DD_CRK).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s1 * s3) + s2;
        const double dtnat = eta_s * std::sqrt(u / l);
        double dt = 0.25 * dtmax;
        while (dt > dtnat) dt *= 0.5;
        p[DT] = dt;
    }
    static double step(const double *p, double etapow)
    {
        const double s0 = V3(p + DD_ACC).norm2(), s1 = V3(p + DD_JRK).norm2(), s2 = V3(p + DD_SNP).norm2();
        const double s5 = V3(p + DD_D5A).norm2(), s6 = V3(p + DD_D6A).norm2(), s7 = V3(p + DD_D7A).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s5 * s7) + s6;
        return pow_nth_quant<10>(etapow * (u / l));
    }
    static void dd_accumulate(double &xh, double &xl, double dx)
    {
        const double yh = xh + dx;
        const double res = dx - (yh - xh);
        const double yl = xl + res;
        const double zh = yh + yl;
        const double zl = yl - (zh - yh);
        xh = zh;
        xl = zl;
    }
    static void correct(double *p, const double *f, double etapow, double dtlim)
    {
        const double dt = p[DT];
        const double h = 0.5 * dt;
        const double hinv = 2.0 / dt;
        const V3 fa(f), fj(f + 3), fs(f + 6), fc(f + 9);
        const V3 acc(p + DD_ACC), jrk(p + DD_JRK), snp(p + DD_SNP), crk(p + DD_CRK), vel(p + DD_VEL);
        const V3 Ap = fa + acc, Am = fa - acc;
        const V3 Jp = (fj + jrk) * h, Jm = (fj - jrk) * h;
        const V3 Sp = (fs + snp) * (h * h), Sm = (fs - snp) * (h * h);
        const V3 Cp = (fc + crk) * (h * h * h), Cm = (fc - crk) * (h * h * h);
        const V3 vel1 = vel + h * (Ap + (-3. / 7.) * Jm + (2. / 21) * Sp - (1. / 105.) * Cm);
        const V3 dpos = h * ((vel + vel1) + h * ((-3. / 7.) * Am + (2. / 21) * Jp - (1. / 105.) * Sm));
        dd_accumulate(p[DD_POSH], p[DD_POSL], dpos.x);
        dd_accumulate(p[DD_POSH + 1], p[DD_POSL + 1], dpos.y);
        dd_accumulate(p[DD_POSH + 2], p[DD_POSL + 2], dpos.z);
        vel1.store(p + DD_VEL);
        fa.store(p + DD_ACC);
        fj.store(p + DD_JRK);
        fs.store(p + DD_SNP);
        fc.store(p + DD_CRK);
        p[TLAST] += dt;
        const double hinv2 = hinv * hinv, hinv3 = hinv2 * hinv, hinv4 = hinv2 * hinv2;
        const double hinv5 = hinv2 * hinv3, hinv6 = hinv3 * hinv3, hinv7 = hinv3 * hinv4;
        V3 d4a = (hinv4 * 24. / 32.) * (-5. * Jm + 5. * Sp - Cm);
        V3 d5a = (hinv5 * 120. / 32.) * (21. * Am - 21. * Jp + 8. * Sm - Cp);
        V3 d6a = (hinv6 * 720. / 32.) * (Jm - Sp + Cm / 3.);
        const V3 d7a = (hinv7 * 5040. / 32.) * (-5. * Am + 5. * Jp - 2. * Sm + Cp / 3.);
        const double h2 = 0.5 * h, h3 = (1. / 3.) * h;
        d4a = d4a + h * (d5a + h2 * (d6a + h3 * d7a));
        d5a = d5a + h * (d6a + h2 * d7a);
        d6a = d6a + h * d7a;
        d4a.store(p + DD_D4A);
        d5a.store(p + DD_D5A);
        d6a.store(p + DD_D6A);
        d7a.store(p + DD_D7A);
        const double dtq = step(p, etapow);
        p[DT] = dtq < dtlim ? dtq : dtlim;
    }
};

enum Phase { P_FORCE = 0, P_POT, P_PREDICT, P_CORRECT, P_SORT, P_SCAN, P_SETJP, P_IO, P_TOTAL, P_NUM };

// Unquantised Aarseth step used when a run restarts from a dump
// (recalc_dt: SRC/hermite4.h:95-116, SRC/hermite6.h:110-135,
// SRC/hermite8.h:136-165): eta * (u/l)^(1/p), then dt = dt_max halved until
// it is not above that.
double aarseth_unquant(int order, const double *p, double eta)
{
    if (order == 9) {  // SRC/hermite8dd.h:164-186, the 8th-order step on shifted offsets
        const double s0 = V3(p + DD_ACC).norm2(), s1 = V3(p + DD_JRK).norm2(), s2 = V3(p + DD_SNP).norm2();
        const double s5 = V3(p + DD_D5A).norm2(), s6 = V3(p + DD_D6A).norm2(), s7 = V3(p + DD_D7A).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s5 * s7) + s6;
        return eta * std::pow(u / l, 1. / 10.);
    }
    const double s0 = V3(p + ACC).norm2(), s1 = V3(p + JRK).norm2(), s2 = V3(p + SNP).norm2();
    const double s3 = V3(p + CRK).norm2();
    if (order == 4) {
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s1 * s3) + s2;
        return eta * std::sqrt(u / l);
    }
    if (order == 6) {
        const double s4 = V3(p + D4A).norm2(), s5 = V3(p + D5A).norm2();
        const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s3 * s5) + s4;
        return eta * std::pow(u / l, 1. / 6.);
    }
    const double s5 = V3(p + D5A).norm2(), s6 = V3(p + D6A).norm2(), s7 = V3(p + D7A).norm2();
    const double u = std::sqrt(s0 * s2) + s1, l = std::sqrt(s5 * s7) + s6;
    return eta * std::pow(u / l, 1. / 10.);
}

// Vectors a dump holds after id and mass (Particle::dump,
// SRC/hermite4.h:45-54, SRC/hermite6.h:49-60, SRC/hermite8.h:60-73,
// SRC/hermite8dd.h:80-94): pos vel acc jrk snp crk [d4a d5a [d6a d7a]], and
// posH posL vel ... d7a for the double-double build, i.e. the record's
// 3-vectors from word 4 in order.
constexpr int dump_vectors(int o) { return o == 4 ? 6 : o == 6 ? 8 : o == 8 ? 10 : 11; }
inline double *dump_vec(double *p, int v) { return p + POS + 3 * v; }
inline const double *dump_vec(const double *p, int v) { return p + POS + 3 * v; }

std::string g_io_err;

int io_fail(const std::string &m)
{
    g_io_err = m;
    return HH_E_ARG;
}

}  // namespace

struct hhd_sys {
    int order = 0, n = 0, ps = 0, fs = 0;
    double eps2 = 0, eta = 0, etapow = 0, eta_s = 0, dtmax = 0, dtmin = 0, tsys = 0;
    double init_energy = 0, prev_energy = 0;
    long num_step = 0, num_bstep = 0, dtmin_hits = 0;
    // print_stat bookkeeping (SRC/nbodysystem.h:934-969): totals are
    // advanced only when a log line is made; marks are the step counters at
    // the previous line
    long num_step_tot = 0, num_bstep_tot = 0, mark_step = 0, mark_bstep = 0;
    std::vector<double> ptcl, force, pot, dtbuf, tmp;
    hh_ctx *g = nullptr;
    double prof[P_NUM] = {0};
    std::string err;
    int rc = HH_OK;
    bool verbose = false;  // print init_step's lines as the reference does

    double *P(int i) { return ptcl.data() + (size_t)i * ps; }

    bool ok(int r, const char *what)
    {
        if (r != HH_OK && rc == HH_OK) {
            rc = r;
            err = std::string(what) + ": " + hh_last_error(g);
        }
        return r == HH_OK;
    }

    void set_jp(int i) { ok(hh_set_jp(g, i, P(i)), "hh_set_jp"); }

    void assign(int i)
    {
        const double *f = force.data() + (size_t)i * fs;
        switch (order) {
        case 4: Phys<4>::assign(P(i), f); break;
        case 6: Phys<6>::assign(P(i), f); break;
        case 8: Phys<8>::assign(P(i), f); break;
        default: Phys<9>::assign(P(i), f); break;
        }
    }
    void init_dt(int i)
    {
        switch (order) {
        case 4: Phys<4>::init_dt(P(i), eta_s, dtmax); break;
        case 6: Phys<6>::init_dt(P(i), eta_s, dtmax); break;
        case 8: Phys<8>::init_dt(P(i), eta_s, dtmax); break;
        default: Phys<9>::init_dt(P(i), eta_s, dtmax); break;
        }
    }
    void correct(int i, double dtlim)
    {
        const double *f = force.data() + (size_t)i * fs;
        switch (order) {
        case 4: Phys<4>::correct(P(i), f, etapow, dtlim); break;
        case 6: Phys<6>::correct(P(i), f, etapow, dtlim); break;
        case 8: Phys<8>::correct(P(i), f, etapow, dtlim); break;
        default: Phys<9>::correct(P(i), f, etapow, dtlim); break;
        }
    }

    void swap(int i, int j)
    {
        std::memcpy(tmp.data(), P(i), 8 * ps);
        std::memcpy(P(i), P(j), 8 * ps);
        std::memcpy(P(j), tmp.data(), 8 * ps);
    }

    // sort_ptcl(nact, dtlim): groups by dt level with the two-ended swap
    // partition std::partition uses (SRC/nbodysystem.h:486-504).
    void sort_by_level(int nact, double dtlim)
    {
        int end = nact;
        double dt = dtlim;
        while (end != 0) {
            int first = 0, last = end;
            for (;;) {
                while (first != last && P(first)[DT] < dt) ++first;
                if (first == last) break;
                --last;
                while (first != last && !(P(last)[DT] < dt)) --last;
                if (first == last) break;
                swap(first, last);
                ++first;
            }
            end = first;
            dt *= 0.5;
        }
        for (int i = 0; i < nact; i++) dtbuf[i] = P(i)[DT] + P(i)[TLAST];
    }

    // sort_ptcl_dtcache: the same partition keyed on the dt bit pattern,
    // halving by decrementing the exponent field (SRC/nbodysystem.h:505-555).
    void sort_by_level_cached(int nact, double dtlim)
    {
        int64_t key;
        std::memcpy(&key, &dtlim, 8);
        int end = nact;
        auto bits = [this](int i) { int64_t b; std::memcpy(&b, &dtbuf[i], 8); return b; };
        while (end != 0) {
            int i = 0, j = end;
            for (;;) {
                while (i != j && bits(i) != key) ++i;
                if (i == j) break;
                --j;
                while (i != j && bits(j) == key) --j;
                if (i == j) break;
                std::swap(dtbuf[i], dtbuf[j]);
                swap(i, j);
                ++i;
            }
            end = i;
            key -= (int64_t)1 << 52;
        }
        for (int i = 0; i < nact; i++) dtbuf[i] = P(i)[DT] + P(i)[TLAST];
    }

    // Largest power of two <= dtmax dividing tnext (SRC/nbodysystem.h:652-661).
    // Returns 0 when tnext is not on the power-of-two grid (a collapsed step).
    double calc_dtlim(double tnext) const
    {
        double dtlim = dtmax, s = tnext / dtmax;
        for (int k = 0; k < 1100 && std::isfinite(s); k++) {
            if (s == std::floor(s)) return dtlim;
            s *= 2.0;
            dtlim *= 0.5;
        }
        return 0.0;
    }

    int count_nact(double tnext) const
    {
        int k = 0;
        while (k < n && dtbuf[k] == tnext) k++;
        return k;
    }

    double energy(bool print = false)
    {
        const double t0 = omp_get_wtime();
        ok(hh_potential(g, eps2, pot.data()), "hh_potential");
        prof[P_POT] += omp_get_wtime() - t0;
        double ke = 0.0, pe = 0.0;
        for (int i = 0; i < n; i++) {
            ke += P(i)[MASS] * V3(P(i) + vel_word(order)).norm2();
            pe += P(i)[MASS] * pot[i];
        }
        ke *= 0.5;
        pe *= 0.5;
        if (print && verbose) std::fprintf(stderr, "ke = %24.16e, pe = %24.16e, e = %24.16e\n", ke, pe, ke + pe);
        return ke + pe;
    }

    // init_step: SRC/nbodysystem.h:565-631
    void init_step()
    {
        for (int i = 0; i < n; i++) {
            P(i)[TLAST] = tsys;
            set_jp(i);
        }
        init_energy = prev_energy = energy(true);
        if (order > 4) {
            if (verbose) std::puts("INITIAL FORCE");
            ok(hh_predict_all(g, tsys), "hh_predict_all");
            ok(hh_force_first_nact(g, n, eps2, force.data(), nullptr), "hh_force_first_nact");
            for (int i = 0; i < n; i++) {
                assign(i);
                set_jp(i);
            }
        }
        double t0 = omp_get_wtime();
        ok(hh_predict_all(g, tsys), "hh_predict_all");
        double t1 = omp_get_wtime();
        ok(hh_force_first_nact(g, n, eps2, force.data(), nullptr), "hh_force_first_nact");
        double t2 = omp_get_wtime();
        prof[P_PREDICT] += t1 - t0;
        prof[P_FORCE] += t2 - t1;
        if (verbose) std::printf("## Init N square : %f Gflops\n", ((1.0e-9 * n) * n) * flops() / (t2 - t1));
        for (int i = 0; i < n; i++) {
            assign(i);
            init_dt(i);
        }
        sort_by_level(n, dtmax);
        for (int i = 0; i < n; i++) set_jp(i);
        init_energy = energy(true);
    }

    int flops() const { return order == 4 ? 60 : order == 6 ? 97 : 144; }

    // integrate_one_block: SRC/nbodysystem.h:692-765.  Returns nact.
    int integrate_one_block()
    {
        double t0 = omp_get_wtime();
        const double tnext = P(0)[TLAST] + P(0)[DT];
        const double dtlim = calc_dtlim(tnext);
        const int nact = count_nact(tnext);
        // A non-finite or non-advancing block time means some particle's step
        // collapsed (the quantised Aarseth step of noisy derivatives can run
        // away); stop with an error instead of looping.
        if (!(tnext > tsys) || !std::isfinite(tnext) || dtlim <= 0.0 || nact <= 0) {
            char msg[160];
            std::snprintf(msg, sizeof msg, "block time does not advance: tsys=%.17g tnext=%.17g nact=%d", tsys,
                          tnext, nact);
            if (rc == HH_OK) {
                rc = HH_E_STATE;
                err = msg;
            }
            return 0;
        }
        double t1 = omp_get_wtime();
        ok(hh_predict_all(g, tnext), "hh_predict_all");
        double t2 = omp_get_wtime();
        ok(hh_force_first_nact(g, nact, eps2, force.data(), nullptr), "hh_force_first_nact");
        double t3 = omp_get_wtime();
        long hits = 0, bad = 0;
#pragma omp parallel for reduction(+ : hits, bad) if (nact > 256)
        for (int i = 0; i < nact; i++) {
            correct(i, dtlim);
            if (dtmin > 0.0 && P(i)[DT] < dtmin) {
                P(i)[DT] = dtmin;
                hits++;
            }
            const int vw = vel_word(order);
            if (!std::isfinite(P(i)[POS] + P(i)[POS + 1] + P(i)[POS + 2] + P(i)[vw] + P(i)[vw + 1] + P(i)[vw + 2]))
                bad++;
            dtbuf[i] = P(i)[DT];
        }
        dtmin_hits += hits;
        if (bad && rc == HH_OK) {
            rc = HH_E_STATE;
            err = "a particle state became non-finite";
        }
        double t4 = omp_get_wtime();
        sort_by_level_cached(nact, dtlim);
        double t5 = omp_get_wtime();
#pragma omp parallel for if (nact > 256)
        for (int i = 0; i < nact; i++) set_jp(i);
        double t6 = omp_get_wtime();
        prof[P_SCAN] += t1 - t0;
        prof[P_PREDICT] += t2 - t1;
        prof[P_FORCE] += t3 - t2;
        prof[P_CORRECT] += t4 - t3;
        prof[P_SORT] += t5 - t4;
        prof[P_SETJP] += t6 - t5;
        num_step += nact;
        num_bstep++;
        tsys = tnext;
        return nact;
    }

    // reinitialize_steps (SRC/nbodysystem.h:447-458): unquantised steps from
    // the restored derivatives, dt-level partition, j-store refresh.
    void reinitialize_steps(double eta_, double dtmax_)
    {
        eta = eta_;
        dtmax = dtmax_;
        for (int i = 0; i < n; i++) {
            const double dta = aarseth_unquant(order, P(i), eta);
            double dt = dtmax;
            while (dt > dta) dt *= 0.5;
            P(i)[DT] = dt;
        }
        sort_by_level(n, dtmax);
        for (int i = 0; i < n; i++) set_jp(i);
    }

    void integrate_one_dtmax()
    {
        const double t0 = omp_get_wtime();
        const double tt = tsys + dtmax;
        while (tsys < tt && rc == HH_OK) integrate_one_block();
        prof[P_TOTAL] += omp_get_wtime() - t0;
    }
};

extern "C" {

// Creates the driver around a new back-end context.  recs: n reference
// Particle records (id, mass, pos, vel; other words ignored).
hhd_sys *hhd_create(int order, int n, const double *recs, double tsys, double eps2, double eta, double eta_s,
                    double dtmax, double dtmin, int variant, int device)
{
    if (!valid_order(order) || n <= 0 || !recs) return nullptr;
    hhd_sys *s = new hhd_sys();
    s->order = order;
    s->n = n;
    s->ps = pwords(order);
    s->fs = fwords(order);
    s->eps2 = eps2;
    s->eta = eta;
    s->eta_s = eta_s;
    s->dtmax = dtmax;
    s->dtmin = dtmin;
    s->tsys = tsys;
    s->etapow = std::pow(eta, double(2 * ((order == 9 ? 8 : order) - 3)));
    s->ptcl.assign(recs, recs + (size_t)n * s->ps);
    s->force.assign((size_t)n * s->fs, 0.0);
    s->pot.assign(n, 0.0);
    s->dtbuf.assign(n, 0.0);
    s->tmp.assign(s->ps, 0.0);
    if (hh_create(order, n, variant, device, &s->g) != HH_OK) {
        s->rc = HH_E_NODEV;
        s->err = std::string("hh_create: ") + hh_last_error(nullptr);
        s->g = nullptr;
    }
    return s;
}

void hhd_destroy(hhd_sys *s)
{
    if (!s) return;
    if (s->g) hh_destroy(s->g);
    delete s;
}

int hhd_status(const hhd_sys *s) { return s ? s->rc : HH_E_ARG; }
hh_ctx *hhd_gravity(hhd_sys *s) { return s ? s->g : nullptr; }
const char *hhd_error(const hhd_sys *s) { return s ? s->err.c_str() : "null driver"; }

int hhd_init_step(hhd_sys *s)
{
    if (!s || !s->g) return HH_E_STATE;
    s->init_step();
    return s->rc;
}

int hhd_integrate_one_block(hhd_sys *s)
{
    if (!s || !s->g) return HH_E_STATE;
    const int nact = s->integrate_one_block();
    return s->rc == HH_OK ? nact : s->rc;
}

int hhd_integrate_one_dtmax(hhd_sys *s)
{
    if (!s || !s->g) return HH_E_STATE;
    s->integrate_one_dtmax();
    return s->rc;
}

int hhd_integrate_until(hhd_sys *s, double tend)
{
    if (!s || !s->g) return HH_E_STATE;
    while (s->tsys < tend && s->rc == HH_OK) s->integrate_one_dtmax();
    return s->rc;
}

double hhd_energy(hhd_sys *s) { return s && s->g ? s->energy() : NAN; }
double hhd_init_energy(const hhd_sys *s) { return s ? s->init_energy : NAN; }
double hhd_tsys(const hhd_sys *s) { return s ? s->tsys : NAN; }
long hhd_num_step(const hhd_sys *s) { return s ? s->num_step : -1; }
long hhd_num_bstep(const hhd_sys *s) { return s ? s->num_bstep : -1; }
long hhd_dtmin_hits(const hhd_sys *s) { return s ? s->dtmin_hits : -1; }

void hhd_get_ptcl(const hhd_sys *s, double *out)
{
    if (s) std::memcpy(out, s->ptcl.data(), sizeof(double) * s->ptcl.size());
}

// Wall seconds per phase: force, pot, predict, correct, sort, scan, set_jp,
// I/O, total.
void hhd_profile(const hhd_sys *s, double *out)
{
    if (s) std::memcpy(out, s->prof, sizeof(s->prof));
}

void hhd_add_io_time(hhd_sys *s, double seconds)
{
    if (s) s->prof[P_IO] += seconds;
}

// With verbose on, init_step prints the reference's lines ("INITIAL FORCE",
// "## Init N square", the energy lines of calc_energy(true)).
void hhd_set_verbose(hhd_sys *s, int on)
{
    if (s) s->verbose = on != 0;
}

const char *hh_io_error(void) { return g_io_err.c_str(); }

int hhd_set_eps2(hhd_sys *s, double eps2)
{
    if (!s) return HH_E_ARG;
    s->eps2 = eps2;
    return HH_OK;
}
double hhd_eps2(const hhd_sys *s) { return s ? s->eps2 : NAN; }

// Masaki snapshot reader (read_snapshot_masaki, SRC/nbodysystem.h:379-403):
// "snapid nbody tsys" then per particle "id mass x y z vx vy vz pot" with
// hexadecimal floats.  Call with recs == NULL to read the header only.
// recs: n Particle records of the order (id, mass, pos, vel; the rest 0).
int hh_read_snapshot(const char *path, int order, int *snapid, long *n, double *tsys, double *recs)
{
    if (!path || !snapid || !n || !tsys || !valid_order(order)) return io_fail("bad argument");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return io_fail(std::string("cannot open ") + path);
    int sid;
    long nb;
    double ts;
    if (std::fscanf(fp, "%d %ld %lf", &sid, &nb, &ts) != 3 || nb <= 0) {
        std::fclose(fp);
        return io_fail(std::string("bad snapshot header in ") + path);
    }
    *snapid = sid;
    *n = nb;
    *tsys = ts;
    if (recs) {
        const int ps = pwords(order);
        for (long i = 0; i < nb; i++) {
            long id;
            double m, x, y, z, vx, vy, vz, pot;
            if (std::fscanf(fp, "%ld %lA %lA %lA %lA %lA %lA %lA %lA", &id, &m, &x, &y, &z, &vx, &vy, &vz, &pot) != 9) {
                std::fclose(fp);
                return io_fail(std::string("bad particle line in ") + path);
            }
            double *p = recs + (size_t)i * ps;
            std::memset(p, 0, sizeof(double) * ps);
            std::memcpy(p + ID, &id, 8);
            p[MASS] = m;
            p[POS] = x;
            p[POS + 1] = y;
            p[POS + 2] = z;
            const int vw = vel_word(order);  // posH = pos, posL = 0 for the double-double build
            p[vw] = vx;
            p[vw + 1] = vy;
            p[vw + 2] = vz;
        }
    }
    std::fclose(fp);
    return HH_OK;
}

// write_snapshot_masaki (SRC/nbodysystem.h:359-378) of n records with their
// potentials.
int hh_write_snapshot(const char *path, int order, int snapid, long n, double tsys, const double *recs,
                      const double *pot)
{
    if (!path || !recs || !pot || n <= 0 || !valid_order(order)) return io_fail("bad argument");
    FILE *fp = std::fopen(path, "w");
    if (!fp) return io_fail(std::string("cannot create ") + path);
    const int ps = pwords(order);
    std::fprintf(fp, "%d\n%ld\n%f\n", snapid, n, tsys);
    for (long i = 0; i < n; i++) {
        const double *p = recs + (size_t)i * ps;
        long id;
        std::memcpy(&id, p + ID, 8);
        // the double-double build writes posH + posL (the intent of
        // SRC/nbodysystem.h:366-370, whose misspelt EIGTHDD test makes the
        // reference build read a pos member it does not have)
        const int vw = vel_word(order);
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = order == 9 ? p[DD_POSH + k] + p[DD_POSL + k] : p[POS + k];
        std::fprintf(fp, "%6ld    %A    %A %A %A    %A %A %A    %A\n", id, p[MASS], x[0], x[1], x[2], p[vw],
                     p[vw + 1], p[vw + 2], pot[i]);
    }
    return std::fclose(fp) == 0 ? HH_OK : io_fail(std::string("write error on ") + path);
}

// dump (SRC/nbodysystem.h:405-427 with Particle::dump): header
// "order dumpid / nbody / tsys (%20.16f) / init_energy prev_energy (%A) /
// num_step_tot num_bstep_tot", then id, mass, pos, vel and the derivative
// vectors of the order in hexadecimal floats.
int hh_write_dump(const char *path, int order, int dumpid, long n, double tsys, double e0, double eprev,
                  long nstep_tot, long nbstep_tot, const double *recs)
{
    if (!path || !recs || n <= 0 || !valid_order(order)) return io_fail("bad argument");
    FILE *fp = std::fopen(path, "w");
    if (!fp) return io_fail(std::string("cannot create ") + path);
    const int ps = pwords(order);
    std::fprintf(fp, "%d %d\n%ld\n%20.16f\n", order == 9 ? 8 : order, dumpid, n, tsys);
    std::fprintf(fp, "%A %A\n", e0, eprev);
    std::fprintf(fp, "%ld %ld\n", nstep_tot, nbstep_tot);
    for (long i = 0; i < n; i++) {
        const double *p = recs + (size_t)i * ps;
        long id;
        std::memcpy(&id, p + ID, 8);
        std::fprintf(fp, "%8ld %A", id, p[MASS]);
        for (int v = 0; v < dump_vectors(order); v++) {
            const double *q = dump_vec(p, v);
            std::fprintf(fp, "   %A %A %A", q[0], q[1], q[2]);
        }
        std::fprintf(fp, "\n");
    }
    return std::fclose(fp) == 0 ? HH_OK : io_fail(std::string("write error on ") + path);
}

// restore (SRC/nbodysystem.h:428-446): the dump's header and records, with
// tlast = tsys and dt = 0 (reinitialize_steps sets it).  recs == NULL reads
// the header only.
int hh_read_dump(const char *path, int order, int *dumpid, long *n, double *tsys, double *e0, double *eprev,
                 long *nstep_tot, long *nbstep_tot, double *recs)
{
    if (!path || !dumpid || !n || !tsys || !e0 || !eprev || !nstep_tot || !nbstep_tot ||
        !valid_order(order))
        return io_fail("bad argument");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return io_fail(std::string("cannot open ") + path);
    int ord, did;
    long nb, a, b;
    double ts, x0, x1;
    if (std::fscanf(fp, "%d %d %ld %lf", &ord, &did, &nb, &ts) != 4 || std::fscanf(fp, "%lA %lA", &x0, &x1) != 2 ||
        std::fscanf(fp, "%ld %ld", &a, &b) != 2 || nb <= 0) {
        std::fclose(fp);
        return io_fail(std::string("bad dump header in ") + path);
    }
    if (ord != (order == 9 ? 8 : order)) {
        std::fclose(fp);
        return io_fail(std::string("dump ") + path + " is of order " + std::to_string(ord));
    }
    *dumpid = did;
    *n = nb;
    *tsys = ts;
    *e0 = x0;
    *eprev = x1;
    *nstep_tot = a;
    *nbstep_tot = b;
    if (recs) {
        const int ps = pwords(order);
        for (long i = 0; i < nb; i++) {
            double *p = recs + (size_t)i * ps;
            std::memset(p, 0, sizeof(double) * ps);
            long id;
            int got = std::fscanf(fp, "%ld %lA", &id, p + MASS);
            for (int v = 0; v < dump_vectors(order); v++) {
                double *q = dump_vec(p, v);
                got += std::fscanf(fp, "%lA %lA %lA", q, q + 1, q + 2);
            }
            if (got != 2 + 3 * dump_vectors(order)) {
                std::fclose(fp);
                return io_fail(std::string("bad particle record in ") + path);
            }
            std::memcpy(p + ID, &id, 8);
            p[TLAST] = ts;
        }
    }
    std::fclose(fp);
    return HH_OK;
}

// The driver's current records and the potentials of its last energy.
int hhd_write_snapshot(hhd_sys *s, const char *path, int snapid)
{
    if (!s) return HH_E_ARG;
    const int r = hh_write_snapshot(path, s->order, snapid, s->n, s->tsys, s->ptcl.data(), s->pot.data());
    if (r != HH_OK && s->rc == HH_OK) s->err = g_io_err;
    return r;
}

int hhd_dump(hhd_sys *s, const char *path, int dumpid)
{
    if (!s) return HH_E_ARG;
    const int r = hh_write_dump(path, s->order, dumpid, s->n, s->tsys, s->init_energy, s->prev_energy,
                                s->num_step_tot, s->num_bstep_tot, s->ptcl.data());
    if (r != HH_OK && s->rc == HH_OK) s->err = g_io_err;
    return r;
}

// State carried by a dump besides the records: energies and step totals.
int hhd_set_restart_state(hhd_sys *s, double e0, double eprev, long nstep_tot, long nbstep_tot)
{
    if (!s) return HH_E_ARG;
    s->init_energy = e0;
    s->prev_energy = eprev;
    s->num_step_tot = nstep_tot;
    s->num_bstep_tot = nbstep_tot;
    return HH_OK;
}

// Records at a dt_max boundary from another integrator: every tlast equals
// tsys, so grouping by dt level orders the next block times as the driver
// expects (the state integrate_one_dtmax leaves).
int hhd_load_state(hhd_sys *s, const double *recs, double tsys, long num_step, long num_bstep, long dtmin_hits)
{
    if (!s || !s->g || !recs) return HH_E_ARG;
    for (int i = 0; i < s->n; i++)
        if (recs[(size_t)i * s->ps + TLAST] != tsys) return HH_E_ARG;
    std::memcpy(s->ptcl.data(), recs, sizeof(double) * (size_t)s->n * s->ps);
    s->tsys = tsys;
    s->num_step = num_step;
    s->num_bstep = num_bstep;
    s->dtmin_hits = dtmin_hits;
    s->sort_by_level(s->n, s->dtmax);
    for (int i = 0; i < s->n; i++) s->set_jp(i);
    return s->rc;
}

int hhd_reinitialize_steps(hhd_sys *s, double eta, double dtmax)
{
    if (!s || !s->g) return HH_E_STATE;
    s->reinitialize_steps(eta, dtmax);
    return s->rc;
}

// print_stat (SRC/nbodysystem.h:934-969) without the printing: energy, the
// interval and total step counters, then the interval counters restart.
// out: tsys, nbstep, nstep, nbstep_tot, nstep_tot, nact_avr, de_loc, de_glo,
// energy.
int hhd_stat(hhd_sys *s, double *out)
{
    if (!s || !s->g || !out) return HH_E_ARG;
    const double e = s->energy();
    const long ns = s->num_step - s->mark_step, nbs = s->num_bstep - s->mark_bstep;
    s->num_step_tot += ns;
    s->num_bstep_tot += nbs;
    out[0] = s->tsys;
    out[1] = (double)nbs;
    out[2] = (double)ns;
    out[3] = (double)s->num_bstep_tot;
    out[4] = (double)s->num_step_tot;
    out[5] = double(ns) / double(nbs);
    out[6] = (s->prev_energy - e) / s->init_energy;
    out[7] = (s->init_energy - e) / s->init_energy;
    out[8] = e;
    s->prev_energy = e;
    s->mark_step = s->num_step;
    s->mark_bstep = s->num_bstep;
    return s->rc;
}

long hhd_num_step_tot(const hhd_sys *s) { return s ? s->num_step_tot : -1; }
long hhd_num_bstep_tot(const hhd_sys *s) { return s ? s->num_bstep_tot : -1; }

}  // extern "C"
