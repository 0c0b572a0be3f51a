/*
 * hermite_hip.h - C-ABI of the MI355X (gfx950) Hermite force + predictor
 * back end.
 *
 * The entry points replace the methods of the reference's compile-time force
 * back end `struct Gravity` (one per Hermite order), which the block-step
 * driver calls through exactly five methods (SRC/nbodysystem.h:308,326,569,
 * 575-576,583,588,635,703,707,731,755):
 *
 *   Gravity(int nbody) / ~Gravity()      -> hh_create / hh_destroy
 *   set_jp(int addr, const Particle &p)  -> hh_set_jp
 *   predict_all(double tsys)             -> hh_predict_all
 *   calc_force_on_first_nact(nact, eps2, Force force[])
 *                                        -> hh_force_first_nact
 *   calc_potential(eps2, double potbuf[])-> hh_potential
 *
 * Reference interfaces replaced (TITAN back end): SRC/hermite4-titan.h:4-114,
 * SRC/hermite6-titan.h:4-133, SRC/hermite8-titan.h:4-148 and the host code of
 * SRC/hermite{4,6,8}-titan.cu and SRC/pot-titan.hu:131-171.
 *
 * Records crossing the boundary use the reference's own host layouts:
 *   Particle : SRC/hermite4.h:16-31 (24 doubles), SRC/hermite6.h:19-35 (28),
 *              SRC/hermite8.h:26-45 (36); word 0 is the int64 id.
 *   Force    : acc, jrk[, snp[, crk]] = 6 / 9 / 12 doubles
 *              (SRC/hermite4.h:4-14, SRC/hermite6.h:4-17, SRC/hermite8.h:8-24).
 *
 * Every function returns HH_OK (0) or a negative HH_E* code; the message of
 * the last failure on a context is available from hh_last_error().  No
 * function aborts the process (the reference's checkCudaErrors does,
 * SRC/cuda_pointer.h:8-9; the C++ Gravity shim restores that behaviour).
 *
 * Threading: hh_set_jp is a pure host-memory write plus an atomic
 * bookkeeping update, safe to call concurrently for distinct addr from
 * OpenMP threads (the driver does so, SRC/nbodysystem.h:753-756).  Every
 * other function must be called from one host thread at a time per context.
 */
#ifndef HERMITE_HIP_H
#define HERMITE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hh_ctx hh_ctx;

/* Kernel variants. */
enum {
    HH_FP64 = 0,  /* all-FP64 interaction, FMA + FP32-seeded rsqrt refined to
                     FP64; the drop-in for the TITAN back end */
    HH_SPLIT = 1, /* per-family precision: the acceleration (8th order:
                     and the jerk) and the potential in FP64 arithmetic;
                     the higher families' chain rule in FP32 on FP64
                     relative positions, FP32 partial sums of 16 j folded
                     into FP64 accumulators */
    HH_EXACT = 2  /* IEEE FP64 with the generic back end's exact operation
                     order (SRC/hermite{4,6,8}.h generic Gravity): results
                     bit-identical to the reference CPU build */
};

/* Order code of the double-double 8th-order build (EIGHTHDD,
 * SRC/compile.cpp:47-48, SRC/hermite8dd.h): positions held as hi + lo
 * doubles (Particle of 40 doubles: id mass tlast dt posH[3] posL[3] vel[3]
 * acc[3] ... d7a[3] pad[3], SRC/hermite8dd.h:44-78), relative positions
 * (posHj - posHi) + (posLj - posLi), the 8th-order Force.  Variants HH_FP64
 * and HH_EXACT; the GPU-resident driver (hhr_*) does not take it. */
#define HH_ORDER_8DD 9

/* Status codes. */
enum {
    HH_OK = 0,
    HH_E_ARG = -1,     /* invalid argument (order, size, null pointer) */
    HH_E_HIP = -2,     /* HIP runtime error */
    HH_E_NODEV = -3,   /* no usable gfx950 device */
    HH_E_STATE = -4,   /* call out of sequence */
    HH_E_LIMIT = -5    /* hhr_run stopped at its block-step limit */
};

/* Doubles per record at the boundary and on the device. */
int hh_particle_words(int order);  /* 24 / 28 / 36 / 40 (reference Particle) */
int hh_force_words(int order);     /* 6 / 9 / 12 / 12  (reference Force)      */
int hh_jp_words(int order);        /* 14 / 20 / 26 / 29 (TITAN GParticle,
                                      SRC/hermite6-titan.h:20-29; the
                                      double-double record appends posL) */
int hh_pred_words(int order);      /* 7 / 10 / 13 / 16 (TITAN GPredictor,
                                      SRC/hermite6-titan.h:31-36; posL last
                                      for the double-double build)        */

/* Gravity(int nbody): SRC/hermite6-titan.h:55-62.  order is 4, 6, 8 or
 * HH_ORDER_8DD; variant one of HH_FP64, HH_SPLIT, HH_EXACT (not HH_SPLIT
 * for HH_ORDER_8DD); device a HIP ordinal. */
int hh_create(int order, int nbody, int variant, int device, hh_ctx **out);

/* ~Gravity(): SRC/hermite6-titan.h:63-68. */
int hh_destroy(hh_ctx *ctx);

/* set_jp(addr, p): SRC/hermite6-titan.h:70-97.  Copies pos, mass, vel, tlast
 * and the stored derivatives of one corrected Particle record into the
 * pinned host j-store slot addr.  No HIP call. */
int hh_set_jp(hh_ctx *ctx, int addr, const double *particle);

/* Pointer to the pinned host j-store (nbody records of hh_jp_words doubles,
 * word order pos[3] mass vel[3] tlast acc[3] jrk[3] snp[3] crk[3] d4a[3]
 * d5a[3]); a caller that writes it directly must then call hh_mark_dirty. */
double *hh_jstore(hh_ctx *ctx);
int hh_mark_dirty(hh_ctx *ctx, int lo, int hi);

/* predict_all(tsys): SRC/hermite6-titan.cu:269-283.  Uploads the slots
 * written since the last upload, then predicts all nbody j-particles to
 * tsys on the device. */
int hh_predict_all(hh_ctx *ctx, double tsys);

/* calc_force_on_first_nact(nact, eps2, force): SRC/hermite6-titan.cu:495-538.
 * Forces (acc, jrk[, snp[, crk]]) on the predicted particles [0, nact) from
 * all nbody predicted particles, written as nact Force records.  If pot is
 * not NULL it also receives phi_i = -sum_{j != i} m_j / sqrt(r_ij^2 + eps2)
 * over the predicted positions (an output the reference Force lacks). */
int hh_force_first_nact(hh_ctx *ctx, int nact, double eps2, double *force, double *pot);

/* No counterpart in the reference.  hh_force_first_nact for callers that
 * cannot allocate device memory during the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * forces, but the j-slice partial sums live in a static device array of
 * hh_static_scratch_doubles() doubles (32 MiB, held by every GPU the
 * library runs on until the process ends) instead of a buffer that grows
 * with hipMalloc; the call allocates nothing.  The j-slice count of the
 * launch is capped to what the scratch holds (so for large nact the
 * summation order, and the last bits, can differ from hh_force_first_nact's;
 * with the same slice count, hh_set_tuning, the results are identical).
 * Calls that overlap on one GPU (contexts on different threads or streams)
 * share that GPU's scratch and race: serialise them.  Returns HH_E_LIMIT if
 * the partial sums would not fit. */
int hh_force_first_nact_static(hh_ctx *ctx, int nact, double eps2, double *force, double *pot);
long long hh_static_scratch_doubles(void);

/* calc_potential(eps2, potbuf): SRC/pot-titan.hu:131-171.  phi_i over the
 * corrected (set_jp) positions of all nbody particles, self excluded. */
int hh_potential(hh_ctx *ctx, double eps2, double *pot);

/* Message of the last failure on ctx (or of the last failed hh_create when
 * ctx is NULL). */
const char *hh_last_error(const hh_ctx *ctx);

/* ---- measurement and tuning (no reference counterpart) ---- */

/* Force launch geometry.  kernel: 0 = automatic, 1 = i-parallel (i in lanes,
 * j tiles in LDS), 2 = j-parallel (j in lanes, small i-groups);
 * jsplit: j-slices per i-block (0 = automatic). */
int hh_set_tuning(hh_ctx *ctx, int kernel, int jsplit);

/* Register blocking and split accumulation: i_per_thread 1 or 2 i-particles
 * per lane in the i-parallel kernel (0 = default 1); split_fold j per FP32
 * partial sum of the SPLIT body before it is added to FP64 (4, 8, 16, 32;
 * 0 = default). */
int hh_set_tuning_ex(hh_ctx *ctx, int i_per_thread, int split_fold);

/* Tuning table (written by hermite_amd/autotune.py): row k applies to active
 * counts up to nact_hi[k] (ascending): kernel (1 i-parallel, 2 j-parallel),
 * i per lane (1 or 2) and j-slices (0 = automatic).  n = 0 clears it.
 * hh_set_tuning overrides take precedence over the table. */
int hh_set_tuning_table(hh_ctx *ctx, int n, const int *nact_hi, const int *kernel, const int *ni,
                        const int *jsplit);
/* Same from a text file of "nact_hi kernel ni jsplit" lines ('#' comments). */
int hh_load_tuning_file(hh_ctx *ctx, const char *path);

/* Runs the force on [0, nact) reps times on the device without copying the
 * result back, and reports the mean device time per call in milliseconds,
 * measured with HIP events on the context's stream. */
int hh_time_force(hh_ctx *ctx, int nact, double eps2, int reps, double *ms_per_call);

/* Same for predict_all (no upload) and the potential. */
int hh_time_predict(hh_ctx *ctx, double tsys, int reps, double *ms_per_call);

/* Launch the force on [0, nact) and return without waiting; the result
 * stays in device memory until hh_synchronize / the next blocking call.
 * For timing the device-resident hot path (predict + force). */
int hh_force_async(hh_ctx *ctx, int nact, double eps2, int with_pot);
int hh_predict_async(hh_ctx *ctx, double tsys);
int hh_synchronize(hh_ctx *ctx);

/* y[i] = 1/sqrt(x[i]) as the FP64 kernels compute it (v_rsq_f32 seed and
 * a second-order refinement; valid for FLT_MIN < x < FLT_MAX), on the
 * context's device (no reference counterpart: the reference's rsqrtf-seeded
 * refinement, SRC/cuda-common.hu:72-79, has no entry point; this one lets
 * tests bound the error against a correctly rounded 1/sqrt).  Host arrays of
 * n doubles; blocking. */
int hh_eval_rsqrt(hh_ctx *ctx, int n, const double *x, double *y);

/* Geometry the last force call used: kernel id, j-slices, workgroups. */
int hh_last_geometry(const hh_ctx *ctx, int *kernel, int *jsplit, int *workgroups);

/* ---- GPU-resident block-step driver (hhr_*) ----
 * The whole NbodySystem loop of SRC/nbodysystem.h on the device for one
 * cluster or an ensemble of independent clusters: block selection by minimum
 * tlast+dt (integrate_one_block, SRC/nbodysystem.h:693-766, without the
 * particle permutation: actives are taken in index order), predict, force,
 * corrector and quantised timestep (SRC/hermite{4,6,8}.h), steps launched in
 * batches captured into a hipGraph.  The reference runs one system per
 * process; clusters here share every launch. */
typedef struct hhr_ctx hhr_ctx;

/* recs: the clusters' reference Particle records (hh_particle_words(order)
 * doubles each) concatenated, after init_step (SRC/nbodysystem.h:565-631:
 * derivatives and dt set).  ns[k]: cluster sizes.  params: 5 doubles per
 * cluster {eps2, eta, dt_max, dt_min, tsys}. */
int hhr_create(int order, int variant, int device, int nclu, const int *ns, const double *recs,
               const double *params, hhr_ctx **out);
int hhr_destroy(hhr_ctx *ctx);

/* Integrate every cluster to tend, a multiple of its dt_max (the loop of
 * integrate_one_dtmax, SRC/nbodysystem.h:904-944, repeated).  batch_steps
 * block steps per launch batch (captured once as a hipGraph when use_graph).
 * Returns HH_E_STATE when a cluster's step collapses (the next block time
 * does not advance) or its state turns non-finite. */
int hhr_run(hhr_ctx *ctx, double tend, int use_graph, int batch_steps);

/* Counters of cluster k: particle steps, block steps, system time and the
 * number of steps clamped to dt_min. */
int hhr_stats(hhr_ctx *ctx, int k, long long *num_step, long long *num_bstep, double *tsys,
              long long *dtmin_hits);
/* Wall milliseconds of the last hhr_run and block steps it launched. */
double hhr_last_run_ms(const hhr_ctx *ctx);
long long hhr_steps_launched(const hhr_ctx *ctx);
/* Block-step limit of hhr_run (no reference counterpart; 0, the default,
 * is none): a call that has launched max_block_steps or more steps without
 * reaching tend returns HH_E_LIMIT after the batch in flight.  The state is
 * that of a completed block step, so a later hhr_run continues the
 * integration; for bounded runs of integrations that may crawl (steps
 * pinned at dt_min). */
int hhr_set_step_limit(hhr_ctx *ctx, long long max_block_steps);
/* Work split of the persistent force kernel (no reference counterpart):
 * active counts up to jpar_max use j-parallel groups of 2 i; larger ones
 * i-parallel groups of 512 i (128 or 256 when tiles are split over waves,
 * hhr_set_split_waves) with at most smax j-slices; a block step aims at
 * task_target i-parallel (jpar_target j-parallel) tasks over all clusters
 * (both at most 8192; jpar_max at most 4096).  hhr_create starts from the measured plan of the
 * same order and kind (one cluster or an ensemble) nearest in total
 * particles, within a factor of 4, where the autotuner found one at least
 * 2 % faster than the built-in plan (hermite_amd/csrc/hh_plan_table.hpp),
 * otherwise from the built-in 64, 256, 8192, 256 (split waves 8192,
 * 16384), which measured within 1.3 % of the best plan found for 4th order
 * N = 1024, a 256-cluster ensemble of 550k particles, 6th order N = 2^16 and
 * 2^20 and 8th order N = 2^18. */
int hhr_set_plan(hhr_ctx *ctx, int jpar_max, int smax, int task_target, int jpar_target);
/* i-parallel tasks of clusters with at most wj4_max (wj2_max) active
 * particles split each j-tile over 4 (2) waves of the workgroup, so a task
 * covers 128 (256) instead of 512 i (no reference counterpart; defaults
 * 8192 and 16384; EXACT always uses whole tiles per wave). */
int hhr_set_split_waves(hhr_ctx *ctx, int wj4_max, int wj2_max);

/* Two-launch block steps (no reference counterpart): a single cluster of
 * at most max_n particles (at most 4096; 0 disables, the default is 4096)
 * integrated with HH_FP64 or HH_SPLIT runs each block step as two launches
 * instead of three.  The force launch has about `grid` workgroups (0:
 * 512), cut evenly over j-slices of 256 jpl slots (jpl: 1, 2 or 4 j per
 * lane; 0 picks by size) and predicts its i and j on the fly; up to
 * islice_max particles (0: 1024, at 4th order 2048; -1: never) each
 * workgroup takes its i from one slice of 256 slots and issues every load
 * in one round trip, above that every workgroup builds the whole active
 * list itself.  The corrector launch, one thread per slot, sums the slices
 * in order and its last workgroup does the step's bookkeeping.  HH_EXACT
 * keeps the three-launch step, whose j summation order is the
 * reference's.  A particle's force differs from the three-launch step's
 * in the j summation order only. */
int hhr_set_fused(hhr_ctx *ctx, int max_n, int grid, int jpl, int islice_max);

/* Block-step timing (no reference counterpart).  on = 1: hhr_run launches
 * its steps directly on the context's stream (no graph; the kernels are the
 * same) with a HIP event pair around every force launch.  on = 2: the first
 * workgroup of each kernel of a step writes the device's constant-rate clock
 * at kernel start, in replayed graphs too; a step's force time runs from the
 * force kernel's start to the corrector's.  hhr_force_timing reports the
 * summed force milliseconds and the number of launches since timing was
 * switched on; on = 0 switches timing off. */
int hhr_set_timing(hhr_ctx *ctx, int on);
int hhr_force_timing(const hhr_ctx *ctx, double *ms_total, long long *launches);
/* Per-step log of the timed steps since timing was switched on (no reference
 * counterpart; the reference's Profile, SRC/nbodysystem.h:57-194, keeps
 * phase totals only): active particles over all clusters, force-kernel ms
 * and ms of the step's other launches, for every step that had
 * active particles.  Copies min(cap, count) entries and returns count. */
long long hhr_step_log(const hhr_ctx *ctx, long long cap, int *nact, float *force_ms, float *other_ms);
/* Reference Particle records of cluster k (ns[k] * hh_particle_words). */
int hhr_get_particles(hhr_ctx *ctx, int k, double *recs);
/* calc_energy of cluster k (SRC/nbodysystem.h:633-650). */
int hhr_energy(hhr_ctx *ctx, int k, double *energy);
const char *hhr_last_error(const hhr_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* HERMITE_HIP_H */
