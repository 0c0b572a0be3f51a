// Harness that links the reference's own CPU back ends (compiled from the
// sources where they lie, see oracle/Makefile) behind a flat extern "C"
// surface, so tests can compare the oracle restatement and the HIP path
// against the reference itself.
//
// TEST INFRASTRUCTURE ONLY.  The prelude below mirrors what the reference's
// own SRC/compile.cpp:1-61 does before including the per-order physics
// header: vector types, the aligned allocator, then hermite{4,6,8}.h (which
// pulls in the back end selected by -DAVX_GRAVITY or the generic Gravity)
// and SRC/nbodysystem.h for the block-step driver.
#include <cstdio>
#include <cassert>
#include <cstring>
#include <new>
#include "vector3.h"
#ifdef __AVX__
typedef double v2df __attribute__((vector_size(16)));
typedef double v4df __attribute__((vector_size(32)));
#endif
#include "allocate.h"

#if defined FOURTH
#include "hermite4.h"
#elif defined SIXTH
#include "hermite6.h"
#elif defined EIGHTH
#include "hermite8.h"
#elif defined EIGHTHDD
// The double-double build's physics and generic Gravity only: the
// reference's driver does not compile for it (SRC/nbodysystem.h:366 tests
// the misspelt macro EIGTHDD and then reads Particle::pos, which the
// double-double Particle does not have).
#include "hermite8dd.h"
#else
#error "define FOURTH, SIXTH, EIGHTH or EIGHTHDD"
#endif
#ifndef EIGHTHDD
#include "nbodysystem.h"
#endif

namespace {
Particle *as_particle(void *buf) { return reinterpret_cast<Particle *>(buf); }
}

extern "C" {

int ref_order() { return Particle::order; }
int ref_particle_bytes() { return (int)sizeof(Particle); }
int ref_force_bytes() { return (int)sizeof(Force); }

// ---- Gravity back end (SRC/hermite*.h generic or SRC/hermite*-avx.h) ----
void *ref_gravity_new(int n) { return new Gravity(n); }
void ref_gravity_free(void *g) { delete static_cast<Gravity *>(g); }

void ref_set_jp(void *g, int addr, const double *rec)
{
    alignas(64) unsigned char buf[sizeof(Particle)];
    std::memcpy(buf, rec, sizeof(Particle));
    static_cast<Gravity *>(g)->set_jp(addr, *as_particle(buf));
}

void ref_predict_all(void *g, double tsys) { static_cast<Gravity *>(g)->predict_all(tsys); }

void ref_force(void *g, int nact, double eps2, double *force_out)
{
    // Force has no padding on non-MIC builds (acc, jrk[, snp[, crk]]).
    static_cast<Gravity *>(g)->calc_force_on_first_nact(nact, eps2, reinterpret_cast<Force *>(force_out));
}

void ref_potential(void *g, double eps2, double *pot) { static_cast<Gravity *>(g)->calc_potential(eps2, pot); }

// ---- Particle physics (SRC/hermite*.h struct Particle) ----
void ref_correct(double *rec, const double *f, double eta, double etapow, double dtlim)
{
    alignas(64) unsigned char buf[sizeof(Particle)];
    std::memcpy(buf, rec, sizeof(Particle));
    Force ff;
    std::memcpy(&ff, f, sizeof(Force));
    as_particle(buf)->correct(ff, eta, etapow, dtlim);
    std::memcpy(rec, buf, sizeof(Particle));
}

void ref_init_dt(double *rec, double eta_s, double dtmax)
{
    alignas(64) unsigned char buf[sizeof(Particle)];
    std::memcpy(buf, rec, sizeof(Particle));
    as_particle(buf)->init_dt(eta_s, dtmax);
    std::memcpy(rec, buf, sizeof(Particle));
}

double ref_pow_one_nth_quant(int n, double x)
{
    switch (n) {
    case 2: return pow_one_nth_quant<2>(x);
    case 6: return pow_one_nth_quant<6>(x);
    case 10: return pow_one_nth_quant<10>(x);
    default: return -1.0;
    }
}

void ref_invfordt(double dt, double *h, double *hinv)
{
    InvForDt t(dt);
    *h = t.h;
    *hinv = t.hinv;
}

#ifndef EIGHTHDD
// ---- Block-step driver (SRC/nbodysystem.h struct NbodySystem) ----
void *ref_sys_new(int n, const double *recs, double tsys, double eps2, double eta, double eta_s,
                  double dtmax, double tend)
{
    NbodySystem *s = new NbodySystem();
    s->set_run_params(tend, dtmax, eps2, eta, eta_s);
    s->nbody = n;
    s->allocate_resources();
    std::memcpy((void *)s->ptcl, recs, sizeof(Particle) * (size_t)n);
    s->tsys = tsys;
    return s;
}
void ref_sys_free(void *s) { delete static_cast<NbodySystem *>(s); }
void ref_sys_init_step(void *s) { static_cast<NbodySystem *>(s)->init_step(); }
void ref_sys_integrate_one_dtmax(void *s) { static_cast<NbodySystem *>(s)->integrate_one_dtmax(); }
double ref_sys_energy(void *s) { return static_cast<NbodySystem *>(s)->calc_energy(false); }
double ref_sys_init_energy(void *s) { return static_cast<NbodySystem *>(s)->init_energy; }
double ref_sys_tsys(void *s) { return static_cast<NbodySystem *>(s)->tsys; }
long ref_sys_num_step(void *s) { return static_cast<NbodySystem *>(s)->num_step; }
long ref_sys_num_bstep(void *s) { return static_cast<NbodySystem *>(s)->num_bstep; }
void ref_sys_get_ptcl(void *s, double *recs)
{
    NbodySystem *sy = static_cast<NbodySystem *>(s);
    std::memcpy(recs, (const void *)sy->ptcl, sizeof(Particle) * (size_t)sy->nbody);
}
void ref_sys_get_pot(void *s, double *pot)
{
    NbodySystem *sy = static_cast<NbodySystem *>(s);
    std::memcpy(pot, sy->potbuf, sizeof(double) * (size_t)sy->nbody);
}
#endif
}

#ifdef REF_MAIN
// Stand-alone driver for the AVX2 back end.  Its per-thread partial-force
// buffers are 32-byte aligned thread-locals (SRC/hermite4-avx.h:282), which
// need the static TLS of an executable; loading it with dlopen is unsafe.
//
//   ref_avxN force <particles.bin> <n> <tsys> <nact> <eps2> <reps> <out.bin|->
//   ref_avxN pot   <particles.bin> <n> <eps2> <out.bin>
//   ref_avxN run   <particles.bin> <n> <eps2> <eta> <eta_s> <dtmax> <tend>
#include <cstdlib>
#include <string>
#include <vector>
#include <algorithm>
#include <omp.h>

static std::vector<double> read_recs(const char *path, int n)
{
    std::vector<double> v((size_t)n * (sizeof(Particle) / 8));
    FILE *fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(Particle), n, fp) != (size_t)n) {
        std::fprintf(stderr, "cannot read %d particles from %s\n", n, path);
        std::exit(2);
    }
    std::fclose(fp);
    return v;
}

// The AVX back end stores Force records and potentials with aligned 32-byte
// vector stores (SRC/hermite6-avx.h:95-104,543), as the reference driver's
// 64-byte aligned buffers allow (SRC/nbodysystem.h:301-304).
template <typename T>
struct Aligned {
    T *p;
    explicit Aligned(size_t n) : p(allocate<T, 64>(n)) { std::memset((void *)p, 0, n * sizeof(T)); }
    ~Aligned() { deallocate<T, 64>(p); }
};

static void write_out(const char *path, const void *p, size_t bytes)
{
    if (std::string(path) == "-") return;
    FILE *fp = std::fopen(path, "wb");
    if (!fp || std::fwrite(p, 1, bytes, fp) != bytes) std::exit(3);
    std::fclose(fp);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 1;
    const std::string cmd = argv[1];
    if (cmd == "force" && argc == 9) {
        const int n = std::atoi(argv[3]);
        const double tsys = std::atof(argv[4]);
        const int nact = std::atoi(argv[5]);
        const double eps2 = std::atof(argv[6]);
        const int reps = std::atoi(argv[7]);
        std::vector<double> recs = read_recs(argv[2], n);
        void *g = ref_gravity_new(n);
        for (int i = 0; i < n; i++) ref_set_jp(g, i, recs.data() + (size_t)i * (sizeof(Particle) / 8));
        Aligned<double> f((size_t)(nact + 4) * (sizeof(Force) / 8));
        ref_predict_all(g, tsys);
        ref_force(g, nact, eps2, f.p);  // warm-up
        const double t0 = omp_get_wtime();
        for (int r = 0; r < reps; r++) ref_force(g, nact, eps2, f.p);
        const double t1 = omp_get_wtime();
        std::printf("{\"seconds\": %.9e, \"reps\": %d, \"nact\": %d, \"n\": %d, \"threads\": %d}\n",
                    t1 - t0, reps, nact, n, omp_get_max_threads());
        write_out(argv[8], f.p, (size_t)nact * sizeof(Force));
        ref_gravity_free(g);
        return 0;
    }
    if (cmd == "pot" && argc == 6) {
        const int n = std::atoi(argv[3]);
        const double eps2 = std::atof(argv[4]);
        std::vector<double> recs = read_recs(argv[2], n);
        void *g = ref_gravity_new(n);
        for (int i = 0; i < n; i++) ref_set_jp(g, i, recs.data() + (size_t)i * (sizeof(Particle) / 8));
        Aligned<double> pot((size_t)n + 4);
        ref_potential(g, eps2, pot.p);
        write_out(argv[5], pot.p, sizeof(double) * n);
        ref_gravity_free(g);
        return 0;
    }
    if (cmd == "run" && (argc == 9 || argc == 10)) {
        // the same integration from the same initial state, reps times (only
        // the integrations are timed; steps and energies are one run's)
        const int n = std::atoi(argv[3]);
        const int reps = argc == 10 ? std::max(1, std::atoi(argv[9])) : 1;
        std::vector<double> recs = read_recs(argv[2], n);
        double secs = 0.0, e0 = 0.0, e1 = 0.0;
        long nstep = 0, nbstep = 0;
        for (int r = 0; r < reps; r++) {
            void *s = ref_sys_new(n, recs.data(), 0.0, std::atof(argv[4]), std::atof(argv[5]),
                                  std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]));
            ref_sys_init_step(s);
            e0 = ref_sys_init_energy(s);
            const double t0 = omp_get_wtime();
            while (ref_sys_tsys(s) < std::atof(argv[8])) ref_sys_integrate_one_dtmax(s);
            secs += omp_get_wtime() - t0;
            e1 = ref_sys_energy(s);
            nstep = ref_sys_num_step(s);
            nbstep = ref_sys_num_bstep(s);
            ref_sys_free(s);
        }
        std::printf("{\"e0\": %.17e, \"e1\": %.17e, \"de\": %.6e, \"num_step\": %ld, \"num_bstep\": %ld, "
                    "\"seconds\": %.6e, \"reps\": %d, \"threads\": %d}\n",
                    e0, e1, (e0 - e1) / e0, nstep, nbstep, secs, reps, omp_get_max_threads());
        return 0;
    }
    std::fprintf(stderr, "bad command\n");
    return 1;
}
#endif
