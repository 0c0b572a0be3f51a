// Address and undefined-behaviour sanitizer harness (CPU only) for the host
// code around the back end: the Plummer and Salpeter generators
// (hh_host.cpp), the snapshot and dump writers and readers (hh_driver.cpp)
// and the CPU oracle (hermite_oracle.c): a round trip of every record
// format and a short oracle integration of every order.  Exit status 0 and
// no sanitizer report mean the checked paths stay in bounds.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hermite_driver.h"

extern "C" {
int orc_pstride(int order);
int orc_predstride(int order);
int orc_fstride(int order);
void orc_predict(int order, int n, const double *ptcl, double tsys, double *pred);
void orc_force(int order, int nj, const double *pred, int nact, double eps2, double *force);
void orc_potential(int order, int n, const double *ptcl, double eps2, double *pot);
struct orc_sys;
orc_sys *orc_sys_new(int order, int n, const double *ptcl_in, double tsys, double eps2, double eta, double eta_s,
                     double dtmax);
void orc_sys_free(orc_sys *s);
void orc_sys_init_step(orc_sys *s);
void orc_sys_integrate_until_inplace(orc_sys *s, double tend);
void orc_sys_integrate_one_dtmax(orc_sys *s);
double orc_sys_energy(orc_sys *s);
double orc_sys_init_energy(const orc_sys *s);
}

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            fails++;                                               \
        }                                                          \
    } while (0)

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int n = 257;  // not a multiple of anything the code might assume
    std::vector<double> m(n), x(3 * n), v(3 * n);
    hh_plummer(n, 7, 22.8, m.data(), x.data(), v.data());
    std::vector<double> ms(n);
    hh_salpeter(n, 11, 2.35, 0.1, 100.0, ms.data());
    double tot = 0;
    for (double a : ms) tot += a;
    CHECK(std::fabs(tot - 1.0) < 1e-12);
    for (int order : {4, 6, 8, 9}) {
        const int ps = orc_pstride(order);
        std::vector<double> P((size_t)n * ps, 0.0), Q((size_t)n * ps), pot(n, 0.5);
        for (int i = 0; i < n; i++) {
            double *p = &P[(size_t)i * ps];
            long id = i;
            std::memcpy(p, &id, 8);
            p[1] = m[i];
            const int vw = order == 9 ? 10 : 7;
            for (int k = 0; k < 3; k++) {
                p[4 + k] = x[3 * i + k];
                p[vw + k] = v[3 * i + k];
            }
            for (int w = vw + 3; w < ps; w++) p[w] = 1e-3 * (w + i);  // derivatives
        }
        const std::string snap = dir + "/snap" + std::to_string(order), dump = dir + "/dump" + std::to_string(order);
        CHECK(hh_write_snapshot(snap.c_str(), order, 3, n, 0.25, P.data(), pot.data()) == 0);
        int sid = 0;
        long nn = 0;
        double ts = 0;
        CHECK(hh_read_snapshot(snap.c_str(), order, &sid, &nn, &ts, nullptr) == 0 && nn == n && sid == 3);
        CHECK(hh_read_snapshot(snap.c_str(), order, &sid, &nn, &ts, Q.data()) == 0);
        for (int i = 0; i < n; i++) CHECK(Q[(size_t)i * ps + 1] == P[(size_t)i * ps + 1]);
        CHECK(hh_write_dump(dump.c_str(), order, 5, n, 0.25, -0.25, -0.2501, 100, 10, P.data()) == 0);
        int did;
        double e0, ep;
        long a, b;
        CHECK(hh_read_dump(dump.c_str(), order, &did, &nn, &ts, &e0, &ep, &a, &b, Q.data()) == 0 && nn == n);
        CHECK(e0 == -0.25 && ep == -0.2501 && a == 100 && b == 10);
        CHECK(hh_read_dump(snap.c_str(), order, &did, &nn, &ts, &e0, &ep, &a, &b, Q.data()) != 0);  // wrong format
        CHECK(hh_read_snapshot((dir + "/missing").c_str(), order, &sid, &nn, &ts, Q.data()) != 0);
        // oracle: predict, force, potential, a short integration
        std::vector<double> pr((size_t)n * orc_predstride(order)), f((size_t)n * orc_fstride(order)), ph(n);
        orc_predict(order, n, P.data(), 0.0, pr.data());
        orc_force(order, n, pr.data(), n, 1.0 / 4096, f.data());
        orc_potential(order, n, P.data(), 1.0 / 4096, ph.data());
        if (order != 9) {
            std::vector<double> R((size_t)n * ps, 0.0);
            for (int i = 0; i < n; i++)
                for (int w = 0; w < 10; w++) R[(size_t)i * ps + w] = P[(size_t)i * ps + w];
            orc_sys *s = orc_sys_new(order, n, R.data(), 0.0, 1.0 / 4096, 0.4, 0.1, 1.0 / 16);
            orc_sys_init_step(s);
            orc_sys_integrate_one_dtmax(s);
            orc_sys_integrate_until_inplace(s, 2.0 / 16);
            const double de = (orc_sys_energy(s) - orc_sys_init_energy(s)) / orc_sys_init_energy(s);
            CHECK(std::isfinite(de) && std::fabs(de) < 1e-3);
            orc_sys_free(s);
        }
    }
    std::printf("san_io: %d failed checks\n", fails);
    return fails ? 1 : 0;
}
