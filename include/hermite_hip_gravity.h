/*
 * hermite_hip_gravity.h - drop-in `struct Gravity` for the reference's
 * block-step driver, backed by the gfx950 C-ABI of hermite_hip.h.
 *
 * The reference selects its force back end inside hermite{4,6,8}.h after
 * `Particle` and `Force` are defined (SRC/hermite4.h:278-288,
 * SRC/hermite6.h:257-267, SRC/hermite8.h:304-314).  Adding one branch there
 *
 *     #elif defined HIP_GRAVITY
 *     #include "hermite_hip_gravity.h"
 *
 * and linking libhermite_hip.so binds SRC/nbodysystem.h to this back end.
 * The class keeps the interface of the TITAN back end
 * (SRC/hermite6-titan.h:4-133): GParticle / GPredictor types (their sizes
 * feed the driver's bandwidth print, SRC/nbodysystem.h:152-153), set_jp,
 * predict_all, calc_force_on_first_nact, calc_potential, and the
 * FAST_OMP_SYNC pair predict_all_fast_omp / calc_force_on_first_nact_fast_omp
 * with NACT_PARALLEL_THRESH.  Like the
 * reference's checkCudaErrors (SRC/cuda_pointer.h:8-9) it aborts on a device
 * error, after printing the library's message.
 *
 * Variant (HH_FP64 / HH_SPLIT / HH_EXACT): compile-time default
 * HERMITE_HIP_VARIANT, overridden at run time by the environment variable
 * HERMITE_HIP_VARIANT=fp64|split|exact.  Device: HERMITE_HIP_DEVICE (0).
 * Launch geometry: HERMITE_HIP_TUNING_FILE names a tuning table in the text
 * form of hh_load_tuning_file (written by hermite_amd/autotune.py); without
 * it the library's built-in choice is used.
 */
#ifndef HERMITE_HIP_GRAVITY_H
#define HERMITE_HIP_GRAVITY_H

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hermite_hip.h"

#ifndef HERMITE_HIP_VARIANT
#define HERMITE_HIP_VARIANT HH_FP64
#endif

struct Gravity {
    enum {
        // FAST_OMP_SYNC drivers correct nact particles with an omp for
        // above this count, on the master thread below it
        // (SRC/nbodysystem.h:800-858); it does not change results.
        NACT_PARALLEL_THRESH = 32,
        ORDER = Particle::order,
        JP_WORDS = ORDER == 4 ? 14 : ORDER == 6 ? 20 : 26,
        PRED_WORDS = ORDER == 4 ? 7 : ORDER == 6 ? 10 : 13,
    };
    // Device record types (sizes only; the records live in HBM).
    struct GParticle { double w[JP_WORDS]; };
    struct GPredictor { double w[PRED_WORDS]; };

    const int nbody;
    hh_ctx *ctx;

    static int variant_from_env()
    {
        const char *v = std::getenv("HERMITE_HIP_VARIANT");
        if (!v) return HERMITE_HIP_VARIANT;
        if (!std::strcmp(v, "fp64")) return HH_FP64;
        if (!std::strcmp(v, "split")) return HH_SPLIT;
        if (!std::strcmp(v, "exact")) return HH_EXACT;
        std::fprintf(stderr, "HERMITE_HIP_VARIANT=%s is not fp64, split or exact\n", v);
        std::abort();
    }

    void check(int rc, const char *what) const
    {
        if (rc != HH_OK) {
            std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, hh_last_error(ctx));
            std::abort();
        }
    }

    explicit Gravity(const int _nbody) : nbody(_nbody), ctx(nullptr)
    {
        static_assert(sizeof(Particle) == (ORDER == 4 ? 192 : ORDER == 6 ? 224 : 288),
                      "Particle layout differs from SRC/hermite{4,6,8}.h");
        static_assert(sizeof(Force) == 24 * (ORDER / 2), "Force layout differs from SRC/hermite{4,6,8}.h");
        const char *dev = std::getenv("HERMITE_HIP_DEVICE");
        const int rc = hh_create(ORDER, nbody, variant_from_env(), dev ? std::atoi(dev) : 0, &ctx);
        if (rc != HH_OK) {
            std::fprintf(stderr, "hh_create failed (%d): %s\n", rc, hh_last_error(nullptr));
            std::abort();
        }
        const char *tab = std::getenv("HERMITE_HIP_TUNING_FILE");
        if (tab && *tab) check(hh_load_tuning_file(ctx, tab), "hh_load_tuning_file");
    }
    ~Gravity() { hh_destroy(ctx); }
    Gravity(const Gravity &) = delete;
    Gravity &operator=(const Gravity &) = delete;

    // Pure host write: safe from concurrent OpenMP threads for distinct addr.
    void set_jp(const int addr, const Particle &p) { check(hh_set_jp(ctx, addr, (const double *)&p), "hh_set_jp"); }

    void predict_all(const double tsys) { check(hh_predict_all(ctx, tsys), "hh_predict_all"); }

    void calc_force_on_first_nact(const int nact, const double eps2, Force force[])
    {
        check(hh_force_first_nact(ctx, nact, eps2, (double *)force, nullptr), "hh_force_first_nact");
    }

    void calc_potential(const double eps2, double potbuf[]) { check(hh_potential(ctx, eps2, potbuf), "hh_potential"); }

    // FAST_OMP_SYNC entry points (SRC/nbodysystem.h:767-902; the AVX back
    // end's are at SRC/hermite6-avx.h:176,496): every thread of the driver's
    // omp parallel region calls them.  The master thread does the HIP work;
    // the barrier after the force makes force[] visible to all threads
    // before their corrector loop.  The driver's own barrier between the
    // two calls orders the prediction before the force.
    void predict_all_fast_omp(const double tsys)
    {
#pragma omp master
        predict_all(tsys);
    }

    void calc_force_on_first_nact_fast_omp(const int nact, const double eps2, Force force[])
    {
#pragma omp master
        calc_force_on_first_nact(nact, eps2, force);
#pragma omp barrier
    }
};

#endif /* HERMITE_HIP_GRAVITY_H */
