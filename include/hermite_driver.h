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
/* Wall seconds per phase in the order of the reference's Profile
 * (SRC/nbodysystem.h:58-71): force, pot, predict, correct, sort, scan,
 * set_jp, I/O, total.  hhd_add_io_time adds the caller's output time. */
void hhd_profile(const hhd_sys *s, double *out);
void hhd_add_io_time(hhd_sys *s, double seconds);
/* With on != 0, init_step prints the reference's lines: "INITIAL FORCE" and
 * "## Init N square : G Gflops" on stdout, the energy lines of
 * calc_energy(true) on stderr (SRC/nbodysystem.h:565-650). */
void hhd_set_verbose(hhd_sys *s, int on);
/* Softening used by later force and energy calls (set_thermal_eps,
 * SRC/nbodysystem.h:480-484, passes 4 / (N kT) squared). */
int hhd_set_eps2(hhd_sys *s, double eps2);
double hhd_eps2(const hhd_sys *s);

/* ---- run control and output of the reference's main loop
 * (SRC/compile.cpp:127-216) ---- */

/* reinitialize_steps (SRC/nbodysystem.h:447-458) after a restart: dt from the
 * unquantised Aarseth step of the restored derivatives, dt-level partition,
 * j-store refresh. */
int hhd_reinitialize_steps(hhd_sys *s, double eta, double dtmax);
/* Energies and step totals a dump carries (restore, SRC/nbodysystem.h:428-446). */
int hhd_set_restart_state(hhd_sys *s, double e0, double eprev, long nstep_tot, long nbstep_tot);
/* State integrated elsewhere, at a dt_max boundary (no reference counterpart:
 * for hermite_run's GPU-resident mode, which integrates with hhr_* and hands
 * the records back for energies, logs, snapshots and dumps).  recs: n
 * Particle records, every one with tlast == tsys; num_step, num_bstep and
 * dtmin_hits replace the counters since hhd_create.  The dt-level partition
 * and the j-store are rebuilt, so host integration may continue.  HH_E_ARG
 * when a record is not at tsys. */
int hhd_load_state(hhd_sys *s, const double *recs, double tsys, long num_step, long num_bstep, long dtmin_hits);
/* print_stat (SRC/nbodysystem.h:934-969) without the printing: out[9] =
 * tsys, block steps and particle steps since the previous call, their
 * totals, mean nact, dE local, dE global, energy. */
int hhd_stat(hhd_sys *s, double *out);
long hhd_num_step_tot(const hhd_sys *s);
long hhd_num_bstep_tot(const hhd_sys *s);
/* write_snapshot_masaki / dump of the driver's current state. */
int hhd_write_snapshot(hhd_sys *s, const char *path, int snapid);
int hhd_dump(hhd_sys *s, const char *path, int dumpid);

/* ---- file formats (host only, no device needed) ----
 * Masaki snapshot (SRC/nbodysystem.h:359-403): "snapid\nnbody\ntsys\n" then
 * "id mass x y z vx vy vz pot" per particle in %A hexadecimal floats.
 * Dump (SRC/nbodysystem.h:405-446, Particle::dump/restore of
 * SRC/hermite{4,6,8}.h): "order dumpid\nnbody\ntsys(%20.16f)\n",
 * "init_energy prev_energy" (%A), "num_step_tot num_bstep_tot", then id, mass,
 * pos, vel, acc, jrk, snp, crk [, d4a, d5a [, d6a, d7a]] per particle.
 * recs are reference Particle records of the order; the readers take
 * recs == NULL to read the header only.  Errors return HH_E_ARG with the
 * text in hh_io_error(). */
int hh_read_snapshot(const char *path, int order, int *snapid, long *n, double *tsys, double *recs);
int hh_write_snapshot(const char *path, int order, int snapid, long n, double tsys, const double *recs,
                      const double *pot);
int hh_read_dump(const char *path, int order, int *dumpid, long *n, double *tsys, double *e0, double *eprev,
                 long *nstep_tot, long *nbstep_tot, double *recs);
int hh_write_dump(const char *path, int order, int dumpid, long n, double tsys, double e0, double eprev,
                  long nstep_tot, long nbstep_tot, const double *recs);
const char *hh_io_error(void);

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
