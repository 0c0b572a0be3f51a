// The reference's own block-step driver (SRC/nbodysystem.h) and per-order
// physics (SRC/hermite{4,6,8}.h), compiled from the reference sources, bound
// to the HIP back end through include/hermite_hip_gravity.h instead of one of
// the reference's back ends.  This is the drop-in check: the reference
// driver, unmodified, runs on the gfx950 Gravity.
//
// TEST INFRASTRUCTURE ONLY (it links reference code).  The generic Gravity
// that hermite{4,6,8}.h defines when no back-end macro is set is renamed out
// of the way; everything after it sees the HIP Gravity.
//
//   ref_hipN run <particles.bin> <n> <eps2> <eta> <eta_s> <dtmax> <tend>
// Variant from HERMITE_HIP_VARIANT (fp64 | split | exact).
#include <cstdio>
#include <cassert>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "vector3.h"
#include "allocate.h"

#define Gravity ReferenceGenericGravity
#if defined FOURTH
#include "hermite4.h"
#elif defined SIXTH
#include "hermite6.h"
#elif defined EIGHTH
#include "hermite8.h"
#else
#error "define FOURTH, SIXTH or EIGHTH"
#endif
#undef Gravity
#include "hermite_hip_gravity.h"
#include "nbodysystem.h"

int main(int argc, char **argv)
{
    if (argc != 9 || std::strcmp(argv[1], "run") != 0) {
        std::fprintf(stderr, "usage: %s run <particles.bin> <n> <eps2> <eta> <eta_s> <dtmax> <tend>\n", argv[0]);
        return 1;
    }
    const int n = std::atoi(argv[3]);
    const double tend = std::atof(argv[8]);
    NbodySystem sys;
    sys.set_run_params(tend, std::atof(argv[7]), std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]));
    sys.nbody = n;
    sys.allocate_resources();
    FILE *fp = std::fopen(argv[2], "rb");
    if (!fp || std::fread((void *)sys.ptcl, sizeof(Particle), n, fp) != (size_t)n) return 2;
    std::fclose(fp);
    sys.tsys = 0.0;
    sys.init_step();
    const double e0 = sys.init_energy;
    long ns1 = 0, nb1 = 0;
    while (sys.tsys < tend) {
        sys.integrate_one_dtmax();
        if (sys.tsys == 1.0) {
            ns1 = sys.num_step;
            nb1 = sys.num_bstep;
        }
    }
    const double e1 = sys.calc_energy(false);
    std::printf("{\"e0\": %.17e, \"e1\": %.17e, \"num_step\": %ld, \"num_bstep\": %ld, "
                "\"num_step_t1\": %ld, \"num_bstep_t1\": %ld}\n",
                e0, e1, sys.num_step, sys.num_bstep, ns1, nb1);
    return 0;
}
