/* Host-side block-step driver and initial conditions (libhermite_host.so).
 *
 * hhd_* is NbodySystem of SRC/nbodysystem.h in C++ over the hh_* back end:
 * the host keeps the reference Particle records, the device does predict and
 * force.  The sorted-by-timestep active list, corrector and timestep follow
 * the reference operation by operation, so with the HH_EXACT back end the
 * integration is bit-identical to the reference CPU build.
 *
 * Status codes are those of hermite_hip.h. */
#ifndef HERMITE_DRIVER_H
#define HERMITE_DRIVER_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hhd_sys hhd_sys;

/* NbodySystem + allocate_resources + Particle(id, mass, pos, vel) records
 * (SRC/nbodysystem.h:253-318, 379-402): n reference Particle records (id,
 * mass, pos, vel used), the run parameters of SRC/parameter.h:6-48 (tsys,
 * eps2, eta, eta_s, dt_max) plus a lower step clamp dt_min (0 = none; the
 * reference has no clamp) and the back end (variant, device).  Returns NULL
 * on bad arguments; a back end that cannot be created is reported by
 * hhd_status / hhd_error. */
hhd_sys *hhd_create(int order, int n, const double *recs, double tsys, double eps2, double eta, double eta_s,
                    double dtmax, double dtmin, int variant, int device);
void hhd_destroy(hhd_sys *s);
int hhd_status(const hhd_sys *s);
/* The driver's back-end context (for hh_set_tuning_table and the like);
 * owned by the driver.  hh_ctx is declared in hermite_hip.h. */
struct hh_ctx *hhd_gravity(hhd_sys *s);
const char *hhd_error(const hhd_sys *s);

/* init_step, SRC/nbodysystem.h:565-631. */
int hhd_init_step(hhd_sys *s);
/* integrate_one_block, SRC/nbodysystem.h:693-766; returns nact (>= 0) or a
 * negative status (HH_E_STATE when the step collapses or turns non-finite). */
int hhd_integrate_one_block(hhd_sys *s);
/* integrate_one_dtmax, SRC/nbodysystem.h:904-944. */
int hhd_integrate_one_dtmax(hhd_sys *s);
/* integrate_one_dtmax repeated until tsys >= tend. */
int hhd_integrate_until(hhd_sys *s, double tend);

/* calc_energy, SRC/nbodysystem.h:633-650. */
double hhd_energy(hhd_sys *s);
double hhd_init_energy(const hhd_sys *s);
double hhd_tsys(const hhd_sys *s);
long hhd_num_step(const hhd_sys *s);
long hhd_num_bstep(const hhd_sys *s);
long hhd_dtmin_hits(const hhd_sys *s);
/* The n Particle records in the driver's current (sorted) order. */
void hhd_get_ptcl(const hhd_sys *s, double *out);
/* Wall seconds per phase: force, pot, predict, correct, sort, scan, set_jp,
 * total (no reference counterpart). */
void hhd_profile(const hhd_sys *s, double *out);

/* ---- initial conditions ---- */

/* The first n values of glibc rand() after srand(seed) (the generator
 * RUN/gen-plum.c draws from). */
void hh_glibc_rand(unsigned seed, int n, int *out);
/* Plummer sphere of RUN/gen-plum.c:100-226 (Henon units, centre of mass at
 * rest); pos and vel n x 3 row-major. */
void hh_plummer(int n, unsigned seed, double rmax, double *mass, double *pos, double *vel);
/* Salpeter masses m^-alpha on [mlo, mhi] from the same generator, total 1
 * (no reference counterpart). */
void hh_salpeter(int n, unsigned seed, double alpha, double mlo, double mhi, double *mass);

#ifdef __cplusplus
}
#endif
#endif /* HERMITE_DRIVER_H */
