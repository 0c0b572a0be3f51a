// GPU-resident block-timestep integration of one or many independent
// clusters (ensembles): the whole block step of SRC/nbodysystem.h:692-765
// runs on the device, so a step costs three kernel launches (four for an
// ensemble of more than R_PF_MAX clusters; capturable in a hipGraph) and no
// host round trip.
//
// Per block step:
//   r_compact_kernel  active list of every cluster (wave-aggregated append),
//                     prediction of every slot of the active clusters to the
//                     block time, and the minimum block time of the slots
//                     that stay inactive
//   r_plan_kernel     per-cluster dtlim, step counters, kernel choice and
//                     geometry, task table of the persistent force kernel
//                     (for a single cluster the compaction's last workgroup
//                     does this instead, for an ensemble of at most
//                     R_PF_MAX clusters the force kernel)
//   r_force_kernel    persistent: a fixed grid walks the step's table of
//                     (cluster, i-group, j-slice) tasks, writing partial forces
//   r_reduce_correct_kernel
//                     sums each active particle's slices in a fixed order,
//                     applies the Hermite corrector (SRC/hermite{4,6,8}.h
//                     correct) to it in place and folds its new block time
//                     into the next step's minimum.
// The next block time of a cluster, min_i(tlast_i + dt_i), is therefore
// complete when a step ends: min over the slots left inactive (compact) and
// over the new times of the corrected ones (reduce_correct).  The minimum and
// the active count live in two parity buffers (RGlob::step & 1 selects the
// current one), so no kernel has to clear what another is still reading.
// Particles are not permuted (the reference sorts them by dt level,
// SRC/nbodysystem.h:505-555): the active set is taken by comparison, and j
// is always summed in the particles' original order.
//
// Resident record, SoA rp[w*npad + i]: words 0..25 are the j-store layout of
// hh_kernels.hpp (pos mass vel tlast acc jrk snp crk d4a d5a), then dt (26),
// id bits (27), d6a (28-30), d7a (31-33).
#pragma once

namespace hh {

constexpr int R_DT = 26, R_ID = 27, R_D6A = 28, R_D7A = 31;
__host__ __device__ constexpr int r_words(int o) { return o == 8 ? 34 : 28; }

constexpr int R_GI = 2;          // i per group of the j-parallel task (half the wave sums of 4: short steps 30 % faster)
// i per lane of the i-parallel task: two i share every LDS read of a j
// (halving LDS traffic per interaction); the EXACT variant keeps one
__host__ __device__ constexpr int r_ni(int mode) { return mode == EXACT ? 1 : 2; }
constexpr int R_TASK_MAX = 8192; // largest task target a plan may ask for
constexpr int R_JPAR_MAX = 4096; // largest active count a plan may give j-parallel groups
constexpr int R_SLOTS = 32;      // words a block-time minimum is spread over
// Steps with at most this many active particles sum each particle's slices
// on a whole wave (one round of loads); larger ones on 8 lanes
constexpr int R_WIDE_MAX = 4096;

// Plan parameters (hhr_set_plan): active counts up to jpar_max use
// j-parallel groups; otherwise i-parallel groups of NT with at most smax
// j-slices.  A block step aims at task_target i-parallel tasks
// (jpar_target j-parallel ones) over all clusters, shared in proportion to
// each cluster's nact * n.
// Active counts up to wj4_max (wj2_max) split each tile's j over 4 (2) of a
// workgroup's waves, so an i-parallel group is 64 (128) i per lane-slot
// row instead of 256: short lists waste fewer lanes.
struct RPlan {
    int jpar_max = 64, smax = 256, task_target = 8192, jpar_target = 256;
    int wj4_max = 8192, wj2_max = 16384;
};

struct RClu {
    double tsys, tend, dtmax, dtmin, eps2, etapow;
    double tnext, dtlim;
    // block-time minimum, parity buffered and spread over R_SLOTS words so
    // that the atomics of one step do not all hit one address
    unsigned long long tmin_slot[2][R_SLOTS];
    long long num_step, num_bstep, dtmin_hits;
    int acnt[2];         // active-list append counters, parity buffered
    int off, n, ntiles;  // particles [off, off + n), tiles from off / TJ
    int nact;
    int mode;            // 0 idle, 1 i-parallel tiles, 2 j-parallel groups
    int wj;              // mode 1: waves sharing each tile's j (1, 2 or 4)
    int ngrp, nslice, tps, jchunk;
    int task0, ntask, act0, ipad;
    long long pofs;      // partials part[pofs + (s*ipad + ii)*NC + c]: a slice's words of one particle together
};

struct RGlob {
    int total_tasks, nclu, nactive, exact;
    int error;     // a cluster's block time stopped advancing
    int total_active;
    int step;      // block steps planned; step & 1 selects the current buffers
    RPlan plan;
    unsigned int cdone[2];  // fused corrector workgroups done, parity buffered (one cluster)
    // one cluster's compaction: workgroups done (high 32 bits) and active
    // particles appended (low 32 bits) in one word, parity buffered
    unsigned long long cpack[2];
    // one cluster of more than R_PACK_MAX tiles: compaction workgroups done
    // per group of tiles (tile mod R_CGRP) and groups done, parity buffered
    unsigned int cgrp[2][64];
    unsigned int ctop[2];
    // timing mode 2: kernel start clocks of step s at stamp[(s % cap) * 4 + k]
    // (read by the timed kernel instances only, so the untimed ones keep
    // their kernel arguments)
    long long *stamp;
};
constexpr int R_CGRP = 64;
// Largest tile count whose compaction workgroups all count themselves on
// one packed atomic: they reach it together, right after their first round
// trip, and same-address atomics are serialised (4096 tiles, N = 2^20: the
// compaction took 130 us instead of 85).
constexpr int R_PACK_MAX = 1024;

struct RArgs {
    double *__restrict__ rp;     // resident records, SoA
    double *__restrict__ jpr;    // predictions, SoA
    double *__restrict__ ipr;    // predictions of the active particles, SoA at their active-list positions
    double *__restrict__ tn;     // tlast + dt per slot (+inf on padding)
    int *__restrict__ ilist;     // active list, at each cluster's offset
    const int *__restrict__ tile_clu;
    RClu *__restrict__ clu;
    RGlob *__restrict__ glob;
    double *__restrict__ part;
    int npad;                    // SoA stride = total padded slots
    int ntiles_total;
    int nclu;                    // clusters (fixed for the context)
    int *__restrict__ steplog;   // optional: active count of step s at [s % steplog_cap]
    int steplog_cap;
};

// Step timing inside replayed graphs (hhr_set_timing mode 2): the first
// thread of the first workgroup of each kernel of block step s writes the
// device's constant-rate clock to stamp[(s % cap) * 4 + k]: k = 0
// compaction (the fused force kernel: the step's first launch), 2 force,
// 3 corrector.  The clock is read as soon as the kernel's first loads are
// issued, not before them: the clock read and the stamp's store count as
// memory side effects, and ahead of the loads they make the compiler turn
// uniform scalar loads into vector loads that wait one after the other.
// The launches of a graph follow each other with no host in between, so
// the next step's first stamp ends a step.  Kernels that may advance the
// step counter themselves pass the value they read before their own
// atomics.  Each kernel has a timed instance (template flag STAMP) and an
// untimed one without this code, which would change the untimed kernels'
// register allocation and cost them 1-3 % (measured).

__device__ __forceinline__ void r_stamp(const RArgs &a, int k, int step, long long t)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) a.glob->stamp[(size_t)(step % a.steplog_cap) * 4 + k] = t;
}

// For kernels during which nothing changes the step counter (back = 1
// after the plan has advanced it): the clock is read by the one thread that
// stores it.
__device__ __forceinline__ void r_stamp_glob(const RArgs &a, int k, int back)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long long t = wall_clock64();
        long long *st = a.glob->stamp + (size_t)((a.glob->step - back) % a.steplog_cap) * 4;
        st[k] = t;
        if (k == 0) st[2] = t;
    }
}

__device__ __forceinline__ double r_inf() { return __longlong_as_double(0x7ff0000000000000LL); }

// ------------------------------------------------------------ corrector
// Device restatement of the per-order Particle::correct with its quantised
// Aarseth step (SRC/hermite4.h:118-153, SRC/hermite6.h:136-184,
// SRC/hermite8.h:166-222), same operation order, no contraction.
struct D3 {
    double x, y, z;
    __device__ D3() : x(0), y(0), z(0) {}
    __device__ D3(double a, double b, double c) : x(a), y(b), z(c) {}
    __device__ D3 operator+(const D3 &o) const { return D3(x + o.x, y + o.y, z + o.z); }
    __device__ D3 operator-(const D3 &o) const { return D3(x - o.x, y - o.y, z - o.z); }
    __device__ D3 operator-() const { return D3(-x, -y, -z); }
    __device__ D3 operator*(double s) const { return D3(x * s, y * s, z * s); }
    __device__ D3 operator/(double s) const { const double r = 1.0 / s; return *this * r; }
    __device__ double norm2() const { return x * x + y * y + z * z; }
};
__device__ __forceinline__ D3 operator*(double s, const D3 &v) { return v * s; }

template <int N>
__device__ __forceinline__ double r_pow_nth_quant(double x)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    unsigned long long e = b >> 52;
    e = (e - (unsigned long long)(1023 % N)) / N + (unsigned long long)(1023 - 1023 / N);
    return __longlong_as_double((long long)(e << 52));
}

// Record words the corrector of an order reads or writes (r_correct): 4th
// order pos vel tlast acc jrk snp crk and dt; 6th order also d4a d5a; 8th
// order every word but the mass and the id, which it never touches.
__host__ __device__ constexpr bool r_used(int o, int k)
{
    return k != 3 && k != R_ID && (o == 8 || k == R_DT || k < (o == 4 ? 20 : 26));
}

// One particle's SoA record words held in registers (loaded and stored by
// the caller).
template <int RW>
struct RRec {
    mutable double v[RW];
    __device__ double &w(int k) const { return v[k]; }
    __device__ D3 v3(int k) const { return D3(v[k], v[k + 1], v[k + 2]); }
    __device__ void set3(int k, const D3 &x) const { v[k] = x.x; v[k + 1] = x.y; v[k + 2] = x.z; }
};

// f: acc jrk [snp [crk]] of this particle.  Sets the new dt.  P: RRec
// (the record's words in registers).
template <int ORDER, class P>
__device__ void r_correct(const P &p, const double *f, double etapow, double dtlim)
{
    const double dt = p.w(R_DT);
    const unsigned long long b = (unsigned long long)__double_as_longlong(dt) - (1ULL << 52);
    const double h = __longlong_as_double((long long)b);
    const double hinv = __longlong_as_double((long long)((2046ULL << 52) - b));
    const D3 fa(f[0], f[1], f[2]), fj(f[3], f[4], f[5]);
    const D3 acc = p.v3(8), jrk = p.v3(11), vel = p.v3(4);
    const D3 Ap = fa + acc, Am = fa - acc;
    const D3 Jp = (fj + jrk) * h, Jm = (fj - jrk) * h;
    double u, l;
    if constexpr (ORDER == 4) {
        const D3 vel1 = vel + h * (Ap - (1. / 3.) * Jm);
        const D3 pos = p.v3(0) + h * ((vel + vel1) + h * (-(1. / 3.) * Am));
        p.set3(0, pos);
        p.set3(4, vel1);
        p.set3(8, fa);
        p.set3(11, fj);
        p.w(7) = p.w(7) + dt;
        D3 snp = (0.5 * hinv * hinv) * Jm;
        const D3 crk = (1.5 * hinv * hinv * hinv) * (Jp - Am);
        snp = snp + h * crk;
        p.set3(14, snp);
        p.set3(17, crk);
        const double s0 = fa.norm2(), s1 = fj.norm2(), s2 = snp.norm2(), s3 = crk.norm2();
        u = sqrt(s0 * s2) + s1;
        l = sqrt(s1 * s3) + s2;
        const double dtq = r_pow_nth_quant<2>(etapow * (u / l));
        p.w(R_DT) = dtq < dtlim ? dtq : dtlim;
    } else if constexpr (ORDER == 6) {
        const D3 fs(f[6], f[7], f[8]), snp = p.v3(14);
        const D3 Sp = (fs + snp) * (h * h), Sm = (fs - snp) * (h * h);
        const D3 vel1 = vel + h * (Ap - 0.4 * Jm + (1. / 15) * Sp);
        const D3 pos = p.v3(0) + h * ((vel + vel1) + h * (-0.4 * Am + (1. / 15) * Jp));
        p.set3(0, pos);
        p.set3(4, vel1);
        p.set3(8, fa);
        p.set3(11, fj);
        p.set3(14, fs);
        p.w(7) = p.w(7) + dt;
        const double hinv2 = hinv * hinv, hinv3 = hinv2 * hinv, hinv4 = hinv2 * hinv2, hinv5 = hinv2 * hinv3;
        D3 crk = (0.75 * hinv3) * (-5. * Am + 5. * Jp - Sm);
        D3 d4a = (1.5 * hinv4) * (-Jm + Sp);
        const D3 d5a = (7.5 * hinv5) * (3 * Am - 3 * Jp + Sm);
        crk = crk + h * (d4a + 0.5 * h * d5a);
        d4a = d4a + h * d5a;
        p.set3(17, crk);
        p.set3(20, d4a);
        p.set3(23, d5a);
        const double s0 = fa.norm2(), s1 = fj.norm2(), s2 = fs.norm2();
        const double s3 = crk.norm2(), s4 = d4a.norm2(), s5 = d5a.norm2();
        u = sqrt(s0 * s2) + s1;
        l = sqrt(s3 * s5) + s4;
        const double dtq = r_pow_nth_quant<6>(etapow * (u / l));
        p.w(R_DT) = dtq < dtlim ? dtq : dtlim;
    } else {
        const D3 fs(f[6], f[7], f[8]), fc(f[9], f[10], f[11]), snp = p.v3(14), crk = p.v3(17);
        const D3 Sp = (fs + snp) * (h * h), Sm = (fs - snp) * (h * h);
        const D3 Cp = (fc + crk) * (h * h * h), Cm = (fc - crk) * (h * h * h);
        const D3 vel1 = vel + h * (Ap + (-3. / 7.) * Jm + (2. / 21) * Sp - (1. / 105.) * Cm);
        const D3 pos = p.v3(0) + h * ((vel + vel1) + h * ((-3. / 7.) * Am + (2. / 21) * Jp - (1. / 105.) * Sm));
        p.set3(0, pos);
        p.set3(4, vel1);
        p.set3(8, fa);
        p.set3(11, fj);
        p.set3(14, fs);
        p.set3(17, fc);
        p.w(7) = p.w(7) + dt;
        const double hinv2 = hinv * hinv, hinv3 = hinv2 * hinv, hinv4 = hinv2 * hinv2;
        const double hinv5 = hinv2 * hinv3, hinv6 = hinv3 * hinv3, hinv7 = hinv3 * hinv4;
        D3 d4a = (hinv4 * 24. / 32.) * (-5. * Jm + 5. * Sp - Cm);
        D3 d5a = (hinv5 * 120. / 32.) * (21. * Am - 21. * Jp + 8. * Sm - Cp);
        D3 d6a = (hinv6 * 720. / 32.) * (Jm - Sp + Cm / 3.);
        const D3 d7a = (hinv7 * 5040. / 32.) * (-5. * Am + 5. * Jp - 2. * Sm + Cp / 3.);
        const double h2 = 0.5 * h, h3 = (1. / 3.) * h;
        d4a = d4a + h * (d5a + h2 * (d6a + h3 * d7a));
        d5a = d5a + h * (d6a + h2 * d7a);
        d6a = d6a + h * d7a;
        p.set3(20, d4a);
        p.set3(23, d5a);
        p.set3(R_D6A, d6a);
        p.set3(R_D7A, d7a);
        const double s0 = fa.norm2(), s1 = fj.norm2(), s2 = fs.norm2();
        const double s5 = d5a.norm2(), s6 = d6a.norm2(), s7 = d7a.norm2();
        u = sqrt(s0 * s2) + s1;
        l = sqrt(s5 * s7) + s6;
        const double dtq = r_pow_nth_quant<10>(etapow * (u / l));
        p.w(R_DT) = dtq < dtlim ? dtq : dtlim;
    }
}

// The fields of a cluster record that the reduction and corrector read.
struct RCluR {
    double etapow, dtlim, dtmin;
    long long pofs;
    int act0, nslice, ipad, off;
};

__device__ __forceinline__ RCluR r_clu_rfields(const RClu &c)
{
    RCluR f;
    f.etapow = c.etapow;
    f.dtlim = c.dtlim;
    f.dtmin = c.dtmin;
    f.pofs = c.pofs;
    f.act0 = c.act0;
    f.nslice = c.nslice;
    f.ipad = c.ipad;
    f.off = c.off;
    return f;
}

// Correct particle i, whose record the caller has loaded into rr, with
// force f; clamp dt to dtmin, store the record and refresh tn[i].
template <int ORDER>
__device__ double r_finish(const RArgs &a, const RCluR &c, int i, RRec<r_words(ORDER)> &rr, const double *f)
{
    constexpr int RW = r_words(ORDER);
    r_correct<ORDER>(rr, f, c.etapow, c.dtlim);
    double dt = rr.v[R_DT];
    if (c.dtmin > 0.0 && dt < c.dtmin) {
        dt = c.dtmin;
        rr.v[R_DT] = dt;
        atomicAdd((unsigned long long *)&a.clu[a.tile_clu[i / TJ]].dtmin_hits, 1ULL);
    }
    const double tn = rr.v[7] + dt;
    a.tn[i] = tn;
#pragma unroll
    for (int k = 0; k < RW; k++)
        if (r_used(ORDER, k)) a.rp[(size_t)k * a.npad + i] = rr.v[k];
    // a non-finite state (a step that ran away on noisy derivatives) stops
    // the run instead of integrating garbage
    if (!isfinite(rr.v[0] + rr.v[1] + rr.v[2] + rr.v[4] + rr.v[5] + rr.v[6])) atomicExch(&a.glob->error, 2);
    return tn;
}

// ------------------------------------------------------------ schedule
// Minimum over the 64 lanes of a wave, in every lane (the stages of
// wave_sum; rows left out of a broadcast keep their own value).
__device__ __forceinline__ double r_wave_min(double x)
{
    x = fmin(x, dpp_f64<DPP_QUAD_XOR1, 0xf>(x, x));
    x = fmin(x, dpp_f64<DPP_QUAD_XOR2, 0xf>(x, x));
    x = fmin(x, dpp_f64<DPP_ROW_HALF_MIRROR, 0xf>(x, x));
    x = fmin(x, dpp_f64<DPP_ROW_MIRROR, 0xf>(x, x));
    x = fmin(x, dpp_f64<DPP_ROW_BCAST15, 0xa>(x, x));
    x = fmin(x, dpp_f64<DPP_ROW_BCAST31, 0xc>(x, x));
    return lane63(x);
}

__device__ __forceinline__ unsigned long long r_bits(double t) { return (unsigned long long)__double_as_longlong(t); }

template <int ORDER, bool PF = false>
__device__ void r_plan_body(const RArgs &a);

// Fields of a single cluster that its plan reads, loaded by every workgroup
// of the compaction in its first round trip, so that the last workgroup
// plans from registers instead of reading them after its count.
struct RPlanIn {
    double tsys, dtmax;
    long long num_step, num_bstep;
    int ntiles, exact;
    RPlan plan;
};

__device__ __forceinline__ RPlanIn r_plan_in(const RArgs &a, const RClu &c)
{
    RPlanIn f;
    f.tsys = c.tsys;
    f.dtmax = c.dtmax;
    f.num_step = c.num_step;
    f.num_bstep = c.num_bstep;
    f.ntiles = c.ntiles;
    f.exact = a.glob->exact;
    f.plan = a.glob->plan;
    return f;
}

template <int ORDER>
__device__ void r_plan_one(const RArgs &a, RClu &c, const RPlanIn &in, int step, int nact, double tmin);

// One workgroup per 256-slot tile; a tile belongs to one cluster.  Times
// are positive, and positive doubles order like their bit patterns, so the
// minima are integer atomicMin on the bits.
//
// One cluster: every load of the kernel (both parities of the block-time
// minimum, the step counter, the slot's block time and record, the fields
// the plan reads) is issued in one round trip, and one 64-bit atomic both
// reserves the tile's list positions and counts the workgroup done; the
// last workgroup to finish plans the step from registers (r_plan_one).
// That saves the plan kernel's launch and every dependent read but the
// atomic.  Ensembles append with one atomic per tile to their cluster's
// counter and are planned in the force launch (up to R_PF_MAX clusters) or
// in r_plan_kernel (planning in the compaction's last workgroup, with the
// workgroups counting themselves done in two levels of 64 counters, made
// the 256-cluster ensemble 2.7 % slower on MI355X).
// CM: 0 ensemble, 1 one cluster of at most R_PACK_MAX tiles, 2 one larger
// cluster.
template <int ORDER, bool EX, int CM, bool STAMP>
__global__ __launch_bounds__(NT) void r_compact_kernel(RArgs a)
{
    constexpr bool ONE = CM != 0, packed = CM == 1;
    constexpr int W = jp_words(ORDER), Q = pr_words(ORDER);
    constexpr int NW = NT / 64;
    __shared__ double wmin[NW];
    __shared__ int wcnt[NW];
    __shared__ int s_base, s_last, s_nact;
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int k = ONE ? 0 : a.tile_clu[tile];
    RClu &c = a.clu[k];
    const int i = tile * TJ + tid;
    const int step = a.glob->step;
    const unsigned long long none = 0x7ff0000000000000ULL;
    const unsigned long long b0 = lane < R_SLOTS ? c.tmin_slot[0][lane] : none;
    const unsigned long long b1 = lane < R_SLOTS ? c.tmin_slot[1][lane] : none;
    const double tend = c.tend;
    const double t = a.tn[i];
    RPlanIn pin{};
    if constexpr (ONE) pin = r_plan_in(a, c);
    // One cluster of at most R_PACK_MAX tiles: the record is read before
    // the block time is known (slot i is inside the store, i < npad), but
    // only after the loads above have been issued: loads complete in issue
    // order for the wait counter, so the schedule's first step waits for
    // those few words alone while the records stream in (at N = 2^18 they
    // are 54 MB per step).  Larger clusters read them once the block time
    // is known: issued all at once, the 168 MB of records of N = 2^20 made
    // the compaction 139 us instead of 90.  Ensembles read only the records
    // of clusters still live: most are done during the last block steps of
    // a run.
    __builtin_amdgcn_sched_barrier(0);
    double r[W];
    if constexpr (packed) {
#pragma unroll
        for (int w = 0; w < W; w++) r[w] = a.rp[(size_t)w * a.npad + i];
    }
    // The timed instance reads the clock (workgroup 0) and stores the stamp
    // only once the loads above are issued: the clock read and the store
    // count as memory side effects, and placed before those loads they
    // turned the uniform ones into vector loads waited for one after the
    // other (C3 compaction 48-51 us instead of 21).
    if constexpr (STAMP) {
        if (blockIdx.x == 0) r_stamp(a, 0, step, wall_clock64());
    }
    const int p = step & 1;
    const double tmin = r_wave_min(__longlong_as_double((long long)(p ? b1 : b0)));
    const bool live = tmin <= tend;
    const bool act = live && t == tmin;
    const unsigned long long mask = __ballot(act);
    // minimum block time of the slots this step leaves alone
    double m = r_wave_min(act ? r_inf() : t);
    if (lane == 0) {
        wmin[wave] = m;
        wcnt[wave] = __popcll(mask);
    }
    __syncthreads();
    // one append per tile (same-address atomics are serialised at the
    // memory side, so fewer of them is what matters)
    // one cluster of at most R_PACK_MAX tiles: the packed atomic below;
    // more tiles: the append reserves (tiles with active particles only)
    // and the workgroups count themselves done in two levels at their end
    if (tid == 0) {
        int tot = 0;
        for (int w = 0; w < NW; w++) {
            m = fmin(m, wmin[w]);
            tot += wcnt[w];
        }
        if (m < r_inf()) atomicMin(&c.tmin_slot[p ^ 1][tile % R_SLOTS], r_bits(m));
        if constexpr (packed) {
            const unsigned long long old = atomicAdd(&a.glob->cpack[p], (1ULL << 32) | (unsigned long long)tot);
            s_base = (int)(old & 0xffffffffULL);
            s_nact = s_base + tot;
            s_last = (unsigned)(old >> 32) == gridDim.x - 1;
        } else {
            s_base = tot > 0 ? atomicAdd(&c.acnt[p], tot) : 0;
        }
    }
    double q[Q];
    if constexpr (!packed) {
        if (live) {
#pragma unroll
            for (int w = 0; w < W; w++) r[w] = a.rp[(size_t)w * a.npad + i];
        }
    }
    if (live) {
        predict_words<ORDER, EX>(r, tmin, q);
#pragma unroll
        for (int w = 0; w < Q; w++) a.jpr[(size_t)w * a.npad + i] = q[w];
    }
    __syncthreads();
    if (act) {
        int base = s_base;
        for (int w = 0; w < wave; w++) base += wcnt[w];
        const unsigned long long below = lane == 0 ? 0ULL : (mask & ((~0ULL) >> (64 - lane)));
        const int pos = c.off + base + __popcll(below);
        a.ilist[pos] = i;
        // the i-side copy the force kernel reads without the ilist indirection
#pragma unroll
        for (int w = 0; w < Q; w++) a.ipr[(size_t)w * a.npad + pos] = q[w];
    }
    // the last workgroup's count returned the total: every other workgroup
    // has appended; the plan reads nothing else they wrote
    if constexpr (ONE) {
        if constexpr (!packed) {
            if (tid == 0) {
                // this tile's append has been performed (its atomic
                // returned) before the tile counts itself done
                __builtin_amdgcn_s_waitcnt(0);
                const int g = tile % R_CGRP;
                const unsigned in_g = (gridDim.x - g + R_CGRP - 1) / R_CGRP;  // tiles of group g
                bool last = false;
                if (atomicAdd(&a.glob->cgrp[p][g], 1u) == in_g - 1) {
                    a.glob->cgrp[p][g] = 0;  // next used two steps on
                    last = atomicAdd(&a.glob->ctop[p], 1u) == (unsigned)R_CGRP - 1;
                }
                if (last) {
                    // every append has been performed: read the count as an atomic
                    s_nact = __hip_atomic_load(&c.acnt[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    c.acnt[p] = 0;
                    a.glob->ctop[p] = 0;
                }
                s_last = last;
            }
            __syncthreads();
        }
        if (s_last) r_plan_one<ORDER>(a, c, pin, step, s_nact, tmin);
    }
}

// calc_dtlim (SRC/nbodysystem.h:652-661) with floor() for the int cast;
// 0 when tnext is off the power-of-two grid.
__device__ double r_dtlim(double tnext, double dtmax)
{
    double dtlim = dtmax, s = tnext / dtmax;
    for (int k = 0; k < 1100 && isfinite(s); k++) {
        if (s == floor(s)) return dtlim;
        s *= 2.0;
        dtlim *= 0.5;
    }
    return 0.0;
}

template <int ORDER, bool POT>
__host__ __device__ constexpr int r_nc() { return n_acc(ORDER, POT); }

// Exclusive scan over the NT threads of a workgroup, in thread order
// (integer sums, so the grouping does not matter); returns this thread's
// prefix and leaves the total in *total.
template <typename T>
__device__ T r_block_exscan(T v, T *s_buf, T *total)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_buf[wave] = x;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        if (w < wave) before += s_buf[w];
        all += s_buf[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Task geometry of one cluster's block step from its active count and its
// share of the step's task budget: i-parallel groups (mode 1) or
// j-parallel groups (mode 2), j-slices, tiles per slice or j per chunk.
struct RGeom {
    int mode, wj, ngrp, tps, jchunk, nslice, ipad;
};

__device__ __forceinline__ RGeom r_geometry(const RPlan &pl, bool exact, int nact, int ntiles, double share)
{
    RGeom g{};
    if (exact || nact > pl.jpar_max) {
        const int nil = 64 * r_ni(exact ? EXACT : FP64);  // i per wave of a mode-1 task
        const int budget = max(1, (int)ceil(pl.task_target * share));
        g.mode = 1;
        // EXACT sums each tile's j in index order on one wave
        g.wj = exact ? 1 : nact <= pl.wj4_max ? 4 : nact <= pl.wj2_max ? 2 : 1;
        const int gs = nil * (NT / 64 / g.wj);
        g.ngrp = (nact + gs - 1) / gs;
        const int ns = exact ? 1 : min(min(ntiles, pl.smax), max(1, (budget + g.ngrp - 1) / g.ngrp));
        g.tps = (ntiles + ns - 1) / ns;
        g.nslice = (ntiles + g.tps - 1) / g.tps;
        g.ipad = g.ngrp * gs;
    } else {
        const int budget = max(1, (int)ceil(pl.jpar_target * share));
        g.mode = 2;
        g.ngrp = (nact + R_GI - 1) / R_GI;
        const int ns = min(ntiles, max(1, (budget + g.ngrp - 1) / g.ngrp));
        const int tpc = (ntiles + ns - 1) / ns;  // tiles per chunk
        g.jchunk = tpc * TJ;
        g.nslice = (ntiles + tpc - 1) / tpc;
        g.ipad = g.ngrp * R_GI;
    }
    return g;
}

// Total work of a step's clusters, w = nact * n of the calling thread's
// clusters: a DPP tree per wave, then the waves in order (fixed order, so
// every workgroup that computes it gets the same bits).
__device__ __forceinline__ double r_total_work(double w, double *s_wsum)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    w = wave_sum(w);
    if (lane == 0) s_wsum[wave] = w;
    __syncthreads();
    double total = 0.0;
#pragma unroll
    for (int q = 0; q < NT / 64; q++) total += s_wsum[q];
    return total;
}

// One workgroup; thread t owns the clusters [t*cpt, (t+1)*cpt).  Geometry
// per cluster from its active count; the task budget is shared in
// proportion to each cluster's nact * n work; task, active-list and partial
// offsets are block scans in cluster order.
//
// PF: the bookkeeping of a step whose force workgroups planned their tasks
// themselves (r_pf_plan), run by one workgroup of that force launch.  The
// active count is the sum of both parity counters (the one not in use is
// zero: the reduce kernel clears a step's counter once every force
// workgroup has read it), and a cluster whose step collapses keeps the
// place of its tasks and partials, which the force workgroups planned
// without that check.
template <int ORDER, bool PF>
__device__ void r_plan_body(const RArgs &a)
{
    constexpr int NC = f_words(ORDER);
    constexpr int NW = NT / 64;
    __shared__ double s_wsum[NW];
    __shared__ long long s_scan[NW];
    const int tid = threadIdx.x;
    const int nclu = a.nclu;
    const int p = a.glob->step & 1;
    const bool exact = a.glob->exact != 0;
    const RPlan pl = a.glob->plan;
    const int cpt = (nclu + NT - 1) / NT;
    const int k0 = min(tid * cpt, nclu), k1 = min(k0 + cpt, nclu);

    // active counts and block-time minima of the owned clusters; this
    // step's block-time slots (and, planned here, its counter) are cleared
    // for their next use
    double w = 0.0;
    for (int k = k0; k < k1; k++) {
        RClu &c = a.clu[k];
        if constexpr (PF) {
            c.nact = c.acnt[0] + c.acnt[1];
        } else {
            c.nact = c.acnt[p];
            c.acnt[p] = 0;
        }
        c.mode = 0;
        c.ntask = 0;
        double tmin = r_inf();
#pragma unroll
        for (int q = 0; q < R_SLOTS; q++) tmin = fmin(tmin, __longlong_as_double((long long)c.tmin_slot[p][q]));
#pragma unroll
        for (int q = 0; q < R_SLOTS; q++) c.tmin_slot[p][q] = r_bits(r_inf());
        c.tnext = tmin;
        if (c.nact > 0) w += (double)c.nact * c.n;
    }
    const double total = r_total_work(w, s_wsum);

    int my_tasks = 0, my_act = 0, my_live = 0;
    long long my_part = 0;
    for (int k = k0; k < k1; k++) {
        RClu &c = a.clu[k];
        if (c.nact == 0) continue;
        const double tmin = c.tnext;
        const double dtlim = r_dtlim(tmin, c.dtmax);
        const bool stop = !(tmin > c.tsys) || dtlim <= 0.0;  // collapsed step: stop this cluster
        const double share = ((double)c.nact * c.n) / total;
        const RGeom g = r_geometry(pl, exact, c.nact, c.ntiles, share);
        if (stop) {
            c.nact = 0;
            c.tend = c.tsys;
            atomicExch(&a.glob->error, 1);
            if constexpr (!PF) continue;
        } else {
            c.dtlim = dtlim;
            c.num_step += c.nact;
            c.num_bstep += 1;
            c.tsys = tmin;
            c.mode = g.mode;
            my_act += c.nact;
            my_live += 1;
        }
        c.wj = g.wj;
        c.ngrp = g.ngrp;
        c.tps = g.tps;
        c.jchunk = g.jchunk;
        c.nslice = g.nslice;
        c.ipad = g.ipad;
        c.ntask = g.ngrp * g.nslice;
        my_tasks += c.ntask;
        my_part += (long long)c.nslice * NC * c.ipad;
    }
    // offsets in cluster order
    long long tot_tasks, tot_act, tot_live, tot_part;
    long long t0 = r_block_exscan<long long>(my_tasks, s_scan, &tot_tasks);
    long long act = r_block_exscan<long long>(my_act, s_scan, &tot_act);
    (void)r_block_exscan<long long>(my_live, s_scan, &tot_live);
    long long p0 = r_block_exscan<long long>(my_part, s_scan, &tot_part);
    for (int k = k0; k < k1; k++) {
        RClu &c = a.clu[k];
        c.task0 = (int)t0;
        c.act0 = (int)act;
        c.pofs = p0;
        t0 += c.ntask;
        if (c.ntask > 0) p0 += (long long)c.nslice * NC * c.ipad;
        if (c.mode != 0) act += c.nact;
    }
    if (tid == 0) {
        a.glob->total_tasks = (int)tot_tasks;
        a.glob->nactive = (int)tot_live;
        a.glob->total_active = (int)tot_act;
        if (a.steplog) a.steplog[a.glob->step % a.steplog_cap] = (int)tot_act;
        a.glob->step += 1;
    }
}

// End of a batch of timed steps: the clock after its last step.
__global__ void r_stamp_end_kernel(RArgs a)
{
    a.glob->stamp[(size_t)a.steplog_cap * 4] = wall_clock64();
}

template <int ORDER>
__global__ __launch_bounds__(NT) void r_plan_kernel(RArgs a)
{
    r_plan_body<ORDER>(a);
}

// The plan of one cluster's block step (r_plan_body for nclu = 1, with the
// whole task budget), by the last workgroup of the compaction: the active
// count and the block time come from the compaction itself, the cluster's
// fields from its first round trip.  Clears this step's parity buffers for
// their next use and advances the step counter.
template <int ORDER>
__device__ void r_plan_one(const RArgs &a, RClu &c, const RPlanIn &in, int step, int nact, double tmin)
{
    const int p = step & 1, tid = threadIdx.x;
    if (tid < R_SLOTS) c.tmin_slot[p][tid] = r_bits(r_inf());
    if (tid != 0) return;
    a.glob->cpack[p] = 0ULL;  // next used two steps on
    c.tnext = tmin;
    int ntask = 0, act = 0;
    RGeom g{};
    if (nact > 0) {
        const double dtlim = r_dtlim(tmin, in.dtmax);
        if (!(tmin > in.tsys) || dtlim <= 0.0) {  // collapsed step: stop the cluster
            c.tend = in.tsys;
            atomicExch(&a.glob->error, 1);
        } else {
            c.dtlim = dtlim;
            c.num_step = in.num_step + nact;
            c.num_bstep = in.num_bstep + 1;
            c.tsys = tmin;
            g = r_geometry(in.plan, in.exact != 0, nact, in.ntiles, 1.0);
            c.wj = g.wj;
            c.ngrp = g.ngrp;
            c.tps = g.tps;
            c.jchunk = g.jchunk;
            c.nslice = g.nslice;
            c.ipad = g.ipad;
            ntask = g.ngrp * g.nslice;
            act = nact;
        }
    }
    c.nact = act;
    c.mode = g.mode;
    c.ntask = ntask;
    c.task0 = 0;
    c.act0 = 0;
    c.pofs = 0;
    a.glob->total_tasks = ntask;
    a.glob->nactive = act > 0 ? 1 : 0;
    a.glob->total_active = act;
    if (a.steplog) a.steplog[step % a.steplog_cap] = act;
    a.glob->step = step + 1;
}

// The fields of a cluster record that a force task reads.
struct RCluF {
    double eps2;
    long long pofs;
    int off, n, ntiles, nact, mode, wj, nslice, tps, jchunk, task0, ipad;
};

__device__ __forceinline__ RCluF r_clu_fields(const RClu &c)
{
    RCluF f;
    f.eps2 = c.eps2;
    f.pofs = c.pofs;
    f.off = c.off;
    f.n = c.n;
    f.ntiles = c.ntiles;
    f.nact = c.nact;
    f.mode = c.mode;
    f.wj = c.wj;
    f.nslice = c.nslice;
    f.tps = c.tps;
    f.jchunk = c.jchunk;
    f.task0 = c.task0;
    f.ipad = c.ipad;
    return f;
}

// Clusters whose first task (first active-list position) the force (reduce)
// kernel keeps in LDS for its lookups; larger ensembles search the cluster
// records in global memory.
constexpr int R_CLU_LDS = 2048;

// The last cluster whose key (task0 or act0) is <= x, from an LDS copy of
// the keys (keys[k], k < nclu) or from the records themselves.
__device__ __forceinline__ int r_find_lds(const int *keys, int nclu, int x)
{
    int lo = 0, hi = nclu - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (keys[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Cluster of a task: the last cluster whose first task is <= task.
__device__ int r_find_cluster(const RClu *clu, int nclu, int task)
{
    int lo = 0, hi = nclu - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (clu[mid].task0 <= task) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Ensembles of at most R_PF_MAX clusters skip the plan launch: every
// workgroup of the force kernel plans the step's tasks itself (the
// geometry of r_plan_body, one thread per cluster, the same bits in every
// workgroup) into LDS, and one workgroup then does the step's bookkeeping
// (r_plan_body<ORDER, true>) for the reduce kernel.  LDS rows of the plan
// (R_PF_MAX ints each, in the cluster-key array): task0, nact, mode | wj
// << 4, nslice, tps or jchunk, ipad, pofs low and high words.
constexpr int R_PF_MAX = NT;
static_assert(8 * R_PF_MAX <= R_CLU_LDS, "plan rows in the cluster-key array");

template <int ORDER>
__device__ int r_pf_plan(const RArgs &a, int *s_key)
{
    constexpr int NC = f_words(ORDER);
    __shared__ double s_wsum[NT / 64];
    __shared__ long long s_scan[NT / 64];
    const int tid = threadIdx.x, nclu = a.nclu;
    const bool exact = a.glob->exact != 0;
    const RPlan pl = a.glob->plan;
    int nact = 0, n = 0, ntiles = 0;
    if (tid < nclu) {
        const RClu &c = a.clu[tid];
        nact = c.acnt[0] + c.acnt[1];  // the counter not in use is zero
        n = c.n;
        ntiles = c.ntiles;
    }
    double w = 0.0;
    if (nact > 0) w += (double)nact * n;
    const double total = r_total_work(w, s_wsum);
    RGeom g{};
    // partial words << 24 | tasks: a step has fewer than 2^24 tasks (the
    // budgets plus one per group, with at most R_JPAR_MAX / 2 j-parallel
    // groups per cluster and 128 i or more per i-parallel group)
    long long mine = 0;
    if (nact > 0) {
        const double share = ((double)nact * n) / total;
        g = r_geometry(pl, exact, nact, ntiles, share);
        mine = ((long long)g.nslice * NC * g.ipad << 24) | (long long)(g.ngrp * g.nslice);
    }
    long long all;
    const long long before = r_block_exscan<long long>(mine, s_scan, &all);
    if (tid < nclu) {
        const long long pofs = before >> 24;
        s_key[tid] = (int)(before & 0xffffff);
        s_key[R_PF_MAX + tid] = nact;
        s_key[2 * R_PF_MAX + tid] = g.mode | g.wj << 4;
        s_key[3 * R_PF_MAX + tid] = g.nslice;
        s_key[4 * R_PF_MAX + tid] = g.mode == 1 ? g.tps : g.jchunk;
        s_key[5 * R_PF_MAX + tid] = g.ipad;
        s_key[6 * R_PF_MAX + tid] = (int)(pofs & 0xffffffffLL);
        s_key[7 * R_PF_MAX + tid] = (int)(pofs >> 32);
    }
    __syncthreads();
    return (int)(all & 0xffffff);
}

// A task's cluster fields: the plan's from the LDS rows, the fixed ones
// from the cluster record.
__device__ __forceinline__ RCluF r_pf_fields(const RArgs &a, const int *s_key, int k)
{
    const RClu &c = a.clu[k];
    RCluF f;
    f.eps2 = c.eps2;
    f.off = c.off;
    f.n = c.n;
    f.ntiles = c.ntiles;
    f.task0 = s_key[k];
    f.nact = s_key[R_PF_MAX + k];
    const int mw = s_key[2 * R_PF_MAX + k];
    f.mode = mw & 15;
    f.wj = mw >> 4;
    f.nslice = s_key[3 * R_PF_MAX + k];
    f.tps = f.jchunk = s_key[4 * R_PF_MAX + k];
    f.ipad = s_key[5 * R_PF_MAX + k];
    f.pofs = (long long)(unsigned)s_key[6 * R_PF_MAX + k] | (long long)s_key[7 * R_PF_MAX + k] << 32;
    return f;
}

// Persistent force kernel: a fixed grid of workgroups walks the step's task
// table, which holds tasks of either kind.  Measured choices: one kernel
// with both task kinds runs the tile tasks faster than a tile-only kernel
// with fewer registers and four waves per SIMD; a budget of 256 registers
// (two waves per SIMD) is faster than 168 (three waves), which the 6th-order
// tile loop fits only by giving up scheduling freedom and the 8th-order one
// only by spilling; the 8th-order kernel, left alone, would take 288
// registers (one wave per SIMD).
template <int ORDER, int MODE, bool PF, bool STAMP>
__device__ __forceinline__ void r_force_body(const RArgs &a)
{
    constexpr int Q = pr_words(ORDER);
    constexpr int L = lds_words(ORDER);
    constexpr int F = f_words(ORDER);
    constexpr int NC = F;
    constexpr int V = R_GI * NC;
    // the split body takes split records (FP32 words packed): LDS stride LS
    constexpr bool SR = MODE == SPLIT;
    constexpr int LS = SR ? split_words(ORDER) : L;
    constexpr int LST = SR ? split_stride(ORDER) : L;
    constexpr int NW = NT / 64;
    constexpr int FOLD = 16;  // j per FP32 partial sum of the split body
    __shared__ double sj[TJ * L];
    __shared__ double red[NW][V];
    __shared__ int s_key[R_CLU_LDS];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int npad = a.npad;
    const int nclu = a.nclu;
    // a single cluster's task fields are read together with the task count
    // (one round trip to memory instead of two)
    RCluF c0{};
    if (!PF && nclu == 1) c0 = r_clu_fields(a.clu[0]);
    const bool keys_lds = !PF && nclu > 1 && nclu <= R_CLU_LDS;
    int total;
    if constexpr (PF) {
        total = r_pf_plan<ORDER>(a, s_key);
    } else {
        total = a.glob->total_tasks;
        if (keys_lds) {
            for (int k = tid; k < nclu; k += NT) s_key[k] = a.clu[k].task0;
            __syncthreads();
        }
        // the step's force stamp, once the loads above are issued (see
        // r_compact_kernel)
        if constexpr (STAMP) r_stamp_glob(a, 2, 1);
    }

    // Static round-robin over the task table: no queue counter (a shared
    // counter is a same-address atomic per task).  Tasks of one step have
    // near-equal work (the plan shares the budget by nact * n), and with a
    // grid that is a multiple of 8 task t runs on XCD t % 8, so slices that
    // are a multiple of 8 apart share an L2.
    // PF: workgroup 0 does the step's bookkeeping only; the tasks go to the others
    const int tfirst = PF ? (int)blockIdx.x - 1 : (int)blockIdx.x, tstride = PF ? (int)gridDim.x - 1 : (int)gridDim.x;
    for (int task = tfirst; task < total; task += tstride) {
        const RCluF c = PF          ? r_pf_fields(a, s_key, r_find_lds(s_key, nclu, task))
                        : nclu == 1 ? c0
                                    : r_clu_fields(a.clu[keys_lds ? r_find_lds(s_key, nclu, task)
                                                                  : r_find_cluster(a.clu, nclu, task)]);
        const int local = task - c.task0;
        const int grp = local / c.nslice, slice = local - grp * c.nslice;
        const double eps2 = c.eps2;
        const double *src = a.jpr;

        if (c.mode == 1) {
            // ---- i in lanes (NI per lane), j-tiles of the slice through LDS.
            // The 4 waves are wi_n rows of i by wj columns of j: wave
            // (wi, wjx) takes i rows wi and the tile's j quarter/half wjx.
            constexpr int NI = r_ni(MODE);
            const int wj = c.wj, wi_n = NT / 64 / wj;
            const int wi = wave % wi_n, wjx = wave / wi_n;
            const int GS = 64 * wi_n * NI;
            int ii[NI];
            double pi[NI][Q];
#pragma unroll
            for (int k = 0; k < NI; k++) {
                ii[k] = grp * GS + k * (64 * wi_n) + wi * 64 + lane;
                const int ip = c.off + min(ii[k], c.nact - 1);
#pragma unroll
                for (int w = 0; w < Q; w++) pi[k][w] = a.ipr[(size_t)w * npad + ip];
                if constexpr (SR) {
                    double t[LS];
                    to_split<ORDER>(pi[k], t);
#pragma unroll
                    for (int w = 0; w < LS; w++) pi[k][w] = t[w];
                }
            }
            // a wave whose lanes are all past the active list only helps
            // stage the tiles; its SIMD slots go to the other waves
            const bool busy = grp * GS + wi * 64 < c.nact;
            const int jlo = wjx * (TJ / wj), jhi = jlo + TJ / wj;
            double A[NI][NC];
#pragma unroll
            for (int k = 0; k < NI; k++)
#pragma unroll
                for (int q = 0; q < NC; q++) A[k][q] = 0.0;
            const int tb = c.off / TJ;
            const int t0 = slice * c.tps, t1 = min(t0 + c.tps, c.ntiles);
            double jr[Q];
#pragma unroll
            for (int w = 0; w < Q; w++) jr[w] = src[(size_t)w * npad + (tb + t0) * TJ + tid];
            for (int t = t0; t < t1; t++) {
                __syncthreads();
                if constexpr (SR) {
                    double tq[LS];
                    to_split<ORDER>(jr, tq);
#pragma unroll
                    for (int w = 0; w < LS; w++) sj[tid * LST + w] = tq[w];
                } else {
#pragma unroll
                    for (int w = 0; w < Q; w++) sj[tid * L + w] = jr[w];
                }
                __syncthreads();
                if (t + 1 < t1) {
#pragma unroll
                    for (int w = 0; w < Q; w++) jr[w] = src[(size_t)w * npad + (tb + t + 1) * TJ + tid];
                }
                if (!busy) continue;
                if constexpr (MODE == EXACT) {
                    const int jn = min(TJ, c.off + c.n - (c.off + t * TJ));
                    for (int jj = 0; jj < jn; jj++) {
                        double q[L];
#pragma unroll
                        for (int w = 0; w < L; w++) q[w] = sj[jj * L + w];
                        body_exact<ORDER, false>(pi[0], q, eps2, false, 0, A[0]);
                    }
                } else if constexpr (MODE == FP64) {
#pragma unroll 2
                    for (int jj = jlo; jj < jhi; jj++) {
                        double q[L];
#pragma unroll
                        for (int w = 0; w < L; w++) q[w] = sj[jj * L + w];
#pragma unroll
                        for (int k = 0; k < NI; k++) body_fp64<ORDER, false>(pi[k], q, eps2, false, 0, A[k]);
                    }
                } else {
                    for (int jb = jlo; jb < jhi; jb += FOLD) {
                        float Af[NI][n_f32(ORDER)];
#pragma unroll
                        for (int k = 0; k < NI; k++)
#pragma unroll
                            for (int q = 0; q < n_f32(ORDER); q++) Af[k][q] = 0.0f;
#pragma unroll 2
                        for (int jj = jb; jj < jb + FOLD; jj++) {
                            double q[LS];
#pragma unroll
                            for (int w = 0; w < LS; w++) q[w] = sj[jj * LST + w];
#pragma unroll
                            for (int k = 0; k < NI; k++) body_split<ORDER, false>(pi[k], q, eps2, false, A[k], Af[k]);
                        }
#pragma unroll
                        for (int k = 0; k < NI; k++) split_fold<ORDER>(A[k], Af[k]);
                    }
                }
            }
            if (wj > 1) {
                // j columns 1.. hand their sums to column 0 through the
                // tile buffer, one i per lane at a time; column 0 adds them
                // in column order
                static_assert((NT / 64 - 1) * NC * 64 <= TJ * L, "tile buffer too small for the column sums");
#pragma unroll
                for (int k = 0; k < NI; k++) {
                    __syncthreads();
                    if (wjx > 0) {
#pragma unroll
                        for (int q = 0; q < NC; q++) sj[(((wjx - 1) * wi_n + wi) * NC + q) * 64 + lane] = A[k][q];
                    }
                    __syncthreads();
                    if (wjx == 0) {
                        for (int r = 1; r < wj; r++) {
#pragma unroll
                            for (int q = 0; q < NC; q++) A[k][q] += sj[(((r - 1) * wi_n + wi) * NC + q) * 64 + lane];
                        }
                    }
                }
            }
            double *part = a.part + c.pofs;
            if (wjx == 0) {
#pragma unroll
                for (int k = 0; k < NI; k++)
                    if (ii[k] < c.nact) {
#pragma unroll
                        for (int q = 0; q < NC; q++) part[((size_t)slice * c.ipad + ii[k]) * NC + q] = A[k][q];
                    }
            }
        } else if constexpr (MODE != EXACT) {  // EXACT always plans mode 1
            // ---- j in lanes, R_GI i-particles per group
            double pi[R_GI][Q];
#pragma unroll
            for (int g = 0; g < R_GI; g++) {
                const int ip = c.off + min(grp * R_GI + g, c.nact - 1);
#pragma unroll
                for (int w = 0; w < Q; w++) pi[g][w] = a.ipr[(size_t)w * npad + ip];
                if constexpr (SR) {
                    double t[LS];
                    to_split<ORDER>(pi[g], t);
#pragma unroll
                    for (int w = 0; w < LS; w++) pi[g][w] = t[w];
                }
            }
            double A[R_GI][NC];
#pragma unroll
            for (int g = 0; g < R_GI; g++)
#pragma unroll
                for (int q = 0; q < NC; q++) A[g][q] = 0.0;
            const int j0 = c.off + slice * c.jchunk;
            const int j1 = min(j0 + c.jchunk, c.off + c.ntiles * TJ);
            float Af[R_GI][n_f32(ORDER)];
#pragma unroll
            for (int g = 0; g < R_GI; g++)
#pragma unroll
                for (int qq = 0; qq < n_f32(ORDER); qq++) Af[g][qq] = 0.0f;
            int nfold = 0;
            for (int j = j0 + tid; j < j1; j += NT) {
                double q[Q];
#pragma unroll
                for (int w = 0; w < Q; w++) q[w] = src[(size_t)w * npad + j];
                if constexpr (SR) {
                    double t[LS];
                    to_split<ORDER>(q, t);
#pragma unroll
                    for (int w = 0; w < LS; w++) q[w] = t[w];
                }
#pragma unroll
                for (int g = 0; g < R_GI; g++) {
                    if constexpr (MODE == FP64) body_fp64<ORDER, false>(pi[g], q, eps2, false, 0, A[g]);
                    else body_split<ORDER, false>(pi[g], q, eps2, false, A[g], Af[g]);
                }
                if constexpr (MODE == SPLIT) {
                    if (++nfold == JFOLD) {
                        nfold = 0;
#pragma unroll
                        for (int g = 0; g < R_GI; g++) split_fold<ORDER>(A[g], Af[g]);
                    }
                }
            }
            if constexpr (MODE == SPLIT) {
#pragma unroll
                for (int g = 0; g < R_GI; g++) split_fold<ORDER>(A[g], Af[g]);
            }
            __syncthreads();  // red reuse across tasks
#pragma unroll
            for (int g = 0; g < R_GI; g++)
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const double x = wave_sum(A[g][q]);
                    if (lane == 0) red[wave][g * NC + q] = x;
                }
            __syncthreads();
            double tot = 0.0;
            if (tid < V) {
#pragma unroll
                for (int w = 0; w < NW; w++) tot += red[w][tid];
            }
            if (tid < V) {
                const int g = tid / NC, q = tid - g * NC;
                a.part[c.pofs + ((size_t)slice * c.ipad + grp * R_GI + g) * NC + q] = tot;
            }
        }
    }
}

// At least two waves per SIMD: the 8th-order body alone would take 288 VGPRs
// (one wave); held to 256 with a few spills it runs 4 % faster.  The
// 6th-order FP64 body is held to exactly two: left free, the compiler fits
// it in 168 VGPRs for three waves, which made the long steps of the
// 6th-order integration 7 % slower (and the whole C2 run 1-2 %).  The other
// bodies keep the compiler's choice (4th order: 4 waves).
//
// STAMP: the timed instance writes the step's force stamp (hhr_set_timing
// mode 2).  PF: an ensemble of at most R_PF_MAX clusters, planned by the
// force workgroups themselves (r_pf_plan); workgroup 0, launched beside
// the grid of task workgroups, does the step's bookkeeping instead of
// tasks and advances the step counter, which no other workgroup of this
// launch reads.  At 4th order the kernel is held to the four waves per SIMD
// that the compiler gives it without the plan.
__host__ __device__ constexpr int r_max_waves(int order, int mode) { return order == 6 && mode == FP64 ? 2 : 8; }
__host__ __device__ constexpr int r_min_waves(int order, bool pf) { return pf && order == 4 ? 4 : 2; }

template <int ORDER, int MODE, bool STAMP, bool PF>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(r_min_waves(ORDER, PF), r_max_waves(ORDER, MODE)))) void r_force_kernel(RArgs a)
{
    if constexpr (PF) {
        if (blockIdx.x == 0) {
            if constexpr (STAMP) r_stamp_glob(a, 2, 0);
            r_plan_body<ORDER, true>(a);
            return;
        }
    }
    r_force_body<ORDER, MODE, PF, STAMP>(a);
}

// Cluster of an active-list position: the last cluster whose act0 <= pos
// (idle clusters share act0 with the next active one, so they are never
// the last such cluster for a valid position).
__device__ int r_find_active(const RClu *clu, int nclu, int pos)
{
    int lo = 0, hi = nclu - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (clu[mid].act0 <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sum every active particle's slice partials (RLX lanes per particle, fixed
// DPP tree) and apply the corrector.  Grid-stride over the active list of
// all clusters, so the launch shape is fixed (graph-capturable).  Short
// lists take a whole wave per particle (one round of loads per lane), long
// ones 8 lanes per particle; the order of additions is fixed by the slice
// count and the active count of the step.
template <int RLX>
__device__ __forceinline__ double r_lane_tree(double x)
{
    static_assert(RLX == 8 || RLX == 64, "lane groups of 8 or a whole wave");
    if constexpr (RLX == 8) return lane_tree8(x);
    else return wave_sum(x);
}

template <int ORDER, int RLX>
__device__ __forceinline__ void r_reduce_body(const RArgs &a, const int *s_key, bool keys_lds, int nclu, int total,
                                              int nb, const RCluR &c0, int ip_first, double &tacc, int &kacc)
{
    constexpr int NC = f_words(ORDER), RW = r_words(ORDER);
    constexpr int PB = NT / RLX;  // particles per workgroup pass
    const int l = threadIdx.x % RLX;
    for (int base = blockIdx.x * PB; base < total; base += gridDim.x * PB) {
        const int pos = base + threadIdx.x / RLX;
        const bool live = pos < total;
        const int k = !live || nclu == 1 ? 0 : keys_lds ? r_find_lds(s_key, nclu, pos) : r_find_active(a.clu, nclu, pos);
        const RCluR c = nclu == 1 ? c0 : r_clu_rfields(a.clu[k]);
        const int ii = pos - c.act0;
        // the particle's index (one cluster: read at kernel start for the
        // first pass) and its record, loaded before the slice sums so that
        // both are in flight together
        const int ip = live && l == 0 ? (nclu == 1 && base == (int)blockIdx.x * PB ? ip_first : a.ilist[c.off + ii]) : 0;
        RRec<RW> rr;
        if (live && l == 0) {
#pragma unroll
            for (int w = 0; w < RW; w++) rr.v[w] = r_used(ORDER, w) ? a.rp[(size_t)w * a.npad + ip] : 0.0;
        }
        double S[NC];
#pragma unroll
        for (int q = 0; q < NC; q++) S[q] = 0.0;
        if (live) {
            const double *part = a.part + c.pofs;
#pragma unroll 4
            for (int sl = l; sl < c.nslice; sl += RLX) {
#pragma unroll
                for (int q = 0; q < NC; q++) S[q] += part[((size_t)sl * c.ipad + ii) * NC + q];
            }
        }
#pragma unroll
        for (int q = 0; q < NC; q++) S[q] = r_lane_tree<RLX>(S[q]);
        if (live && l == 0) {
            const double tn = r_finish<ORDER>(a, c, ip, rr, S);
            if (kacc != k && kacc >= 0) atomicMin(&a.clu[kacc].tmin_slot[nb][blockIdx.x % R_SLOTS], r_bits(tacc));
            if (kacc != k) tacc = r_inf();
            kacc = k;
            tacc = fmin(tacc, tn);
        }
    }
}

template <int ORDER, bool STAMP>
__global__ __launch_bounds__(NT) void r_reduce_correct_kernel(RArgs a)
{
    __shared__ int s_key[R_CLU_LDS];
    const int nclu = a.nclu;
    // a single cluster's fields, and this thread's first active-list entry
    // (for 64 and for RL lanes per particle; its list starts at position
    // 0), are read together with the step's counts
    RCluR c0{};
    int ip64 = 0, ipRL = 0;
    if (nclu == 1) {
        c0 = r_clu_rfields(a.clu[0]);
        ip64 = a.ilist[min((int)(blockIdx.x * (NT / 64) + threadIdx.x / 64), a.npad - 1)];
        ipRL = a.ilist[min((int)(blockIdx.x * (NT / RL) + threadIdx.x / RL), a.npad - 1)];
    }
    const int total = a.glob->total_active;
    const int nb = a.glob->step & 1;  // the step's plan has advanced the counter: next buffers
    // the step's stamp, and the stores below, come after the loads above
    // are issued (see r_compact_kernel)
    if constexpr (STAMP) r_stamp_glob(a, 3, 1);
    // an ensemble planned in its force launch: clear the step's append
    // counters now that every force workgroup has read them (those of the
    // other parity than the advanced step counter's)
    if (nclu > 1 && nclu <= R_PF_MAX && blockIdx.x == 0) {
        for (int k = threadIdx.x; k < nclu; k += NT) a.clu[k].acnt[nb ^ 1] = 0;
    }
    const bool wide = total <= R_WIDE_MAX;
    if (blockIdx.x * (wide ? NT / 64 : NT / RL) >= total) return;  // whole workgroup idle this step
    const bool keys_lds = nclu > 1 && nclu <= R_CLU_LDS;
    if (keys_lds) {
        for (int k = threadIdx.x; k < nclu; k += NT) s_key[k] = a.clu[k].act0;
        __syncthreads();
    }
    // each lane keeps the minimum of the new block times it produced for
    // one cluster and flushes it when the cluster changes
    double tacc = r_inf();
    int kacc = -1;
    if (wide) r_reduce_body<ORDER, 64>(a, s_key, keys_lds, nclu, total, nb, c0, ip64, tacc, kacc);
    else r_reduce_body<ORDER, RL>(a, s_key, keys_lds, nclu, total, nb, c0, ipRL, tacc, kacc);
    // one atomic per workgroup when all its lanes worked for one cluster
    __shared__ int s_k[NT / 64];
    __shared__ double s_t[NT / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kref = __shfl(kacc, 0, 64);
    const bool wuni = __all(kacc < 0 || kacc == kref);
    if (wuni) {
        const double m = r_wave_min(tacc);
        if (lane == 0) {
            s_k[wave] = kref;
            s_t[wave] = m;
        }
    } else {
        if (lane == 0) s_k[wave] = -2;
        if (kacc >= 0) atomicMin(&a.clu[kacc].tmin_slot[nb][blockIdx.x % R_SLOTS], r_bits(tacc));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int kb = -1;
        bool buni = true;
        for (int w = 0; w < NT / 64; w++) {
            if (s_k[w] == -2) continue;  // flushed by its lanes
            if (s_k[w] < 0) continue;    // no particles
            if (kb < 0) kb = s_k[w];
            else if (s_k[w] != kb) buni = false;
        }
        if (buni) {
            double m = r_inf();
            for (int w = 0; w < NT / 64; w++)
                if (s_k[w] >= 0) m = fmin(m, s_t[w]);
            if (kb >= 0 && m < r_inf()) atomicMin(&a.clu[kb].tmin_slot[nb][blockIdx.x % R_SLOTS], r_bits(m));
        } else {
            for (int w = 0; w < NT / 64; w++)
                if (s_k[w] >= 0 && s_t[w] < r_inf())
                    atomicMin(&a.clu[s_k[w]].tmin_slot[nb][blockIdx.x % R_SLOTS], r_bits(s_t[w]));
        }
    }
}

// ------------------------------------------------------ two-launch step
// A block step of one small cluster (n <= R_FUSED_MAX) in two launches
// instead of three, for the FP64 and split variants.  Short block steps
// are bound by the latency of dependent memory round trips, not by
// arithmetic, so both kernels issue every load that does not depend on the
// block time in their first round trip.
//
// r_fused_force_kernel: the n slots are cut into S = ceil(n / (JPL NT))
// j-slices of JPL NT slots, JPL per lane; workgroup b works on slice b % S
// for the groups of R_GI active particles b / S, b / S + wps, ... (wps
// workgroups per slice).  Each workgroup loads its lanes' j records and all n block
// times at once, builds the active list in LDS itself (a redundant scan:
// cheaper than a launch boundary at these sizes), predicts its i records
// into LDS, and for each group sums the lanes' interactions in a fixed
// order into the slice's partial force at the particle's slot.  Nothing in
// this launch writes the records, so every workgroup predicts from the
// corrected state of the previous step.  (r_fused_islice_kernel below is
// the variant without the scan for the smallest clusters.)
//
// r_fused_correct_kernel: one thread per slot, whose block time, record
// and slice partials are all loaded in one round trip.  Active slots sum
// their S partials in slice order and apply the corrector; every slot's
// new (or unchanged) block time goes into the next step's minimum and each
// workgroup adds its active count to the step's counter.  The last
// workgroup to finish (an atomic count, as in the compaction) does the
// step's bookkeeping of r_plan_body for one cluster.  No workgroup reads
// what another one of the same launch writes, except the counters, which
// it reads as atomics.
//
// A one-launch step, whose last force workgroup corrects after a
// device-scope release/acquire around its count, measured slower on
// MI355X (19.3 us against 9.8 + 6.5 us at N = 1024, 4th order): every
// workgroup's release writes back its XCD's L2, and 512 of those cost
// more than a launch boundary.
constexpr int R_FUSED_MAX = 4096;
constexpr int R_FUSED_TV = R_FUSED_MAX / NT;  // block times per thread in the scan
constexpr int R_FUSED_SLICES = R_FUSED_MAX / NT;  // at most, with one j per lane

// Block time, step limit and collapse test of one cluster, from the slot
// minima of both parities loaded together with the step counter.
struct RFusedStep {
    double tmin, dtlim;
    int step, p;
    bool live, collapsed, stepping;
};

__device__ __forceinline__ RFusedStep r_fused_step(const RArgs &a, const RClu &c)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long none = 0x7ff0000000000000ULL;
    const unsigned long long b0 = lane < R_SLOTS ? c.tmin_slot[0][lane] : none;
    const unsigned long long b1 = lane < R_SLOTS ? c.tmin_slot[1][lane] : none;
    RFusedStep f;
    f.step = a.glob->step;
    f.p = f.step & 1;
    f.tmin = r_wave_min(__longlong_as_double((long long)(f.p ? b1 : b0)));
    f.live = f.tmin <= c.tend;
    f.dtlim = f.live ? r_dtlim(f.tmin, c.dtmax) : 0.0;
    f.collapsed = f.live && (!(f.tmin > c.tsys) || f.dtlim <= 0.0);
    f.stepping = f.live && !f.collapsed;
    return f;
}

// The step's bookkeeping of r_plan_body for one cluster, by one thread
// after every other workgroup of the corrector launch is done: counters,
// the cluster's time and step limit, the cleared parity buffers of this
// step, and the step counter.
__device__ void r_fused_bookkeep(const RArgs &a, RClu &c, const RFusedStep &st, int nact, int n, int S)
{
    const int p = st.p;
    a.glob->cdone[p] = 0;
    c.acnt[p] = 0;
    c.mode = 0;
    c.ntask = 0;
    c.tnext = st.tmin;
#pragma unroll
    for (int q = 0; q < R_SLOTS; q++) c.tmin_slot[p][q] = r_bits(r_inf());
    c.nact = st.stepping ? nact : 0;
    if (st.collapsed) {
        c.tend = c.tsys;
        atomicExch(&a.glob->error, 1);
    }
    const int ngrp = (nact + R_GI - 1) / R_GI;
    const bool live_n = st.stepping && nact > 0;
    if (live_n) {
        c.dtlim = st.dtlim;
        c.num_step += nact;
        c.num_bstep += 1;
        c.tsys = st.tmin;
        c.mode = 2;
        c.ngrp = ngrp;
        c.nslice = S;
        c.jchunk = (n + S - 1) / S;
        c.ntask = ngrp * S;
    }
    c.task0 = 0;
    c.act0 = 0;
    c.pofs = 0;
    a.glob->total_tasks = live_n ? ngrp * S : 0;
    a.glob->nactive = live_n ? 1 : 0;
    a.glob->total_active = live_n ? nact : 0;
    if (a.steplog) a.steplog[a.glob->step % a.steplog_cap] = live_n ? nact : 0;
    a.glob->step += 1;
}

template <int ORDER, int MODE, int JPL, bool STAMP>
__global__ __launch_bounds__(NT) void r_fused_force_kernel(RArgs a, int n, int off, int S, int wps)
{
    static_assert(MODE != EXACT, "EXACT sums j in index order: three-launch step");
    constexpr int W = jp_words(ORDER), Q = pr_words(ORDER);
    constexpr int NC = f_words(ORDER);
    constexpr int V = R_GI * NC;
    constexpr int NW = NT / 64;
    constexpr int SW = MODE == SPLIT ? split_words(ORDER) : Q;  // words of the body's records
    __shared__ int s_list[R_FUSED_MAX];
    __shared__ int s_wcnt[R_FUSED_TV][NW];
    __shared__ double s_pi[SW][NT];
    __shared__ double red[2][NW][V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slice = blockIdx.x % S, g0 = blockIdx.x / S;  // the grid is wps * S workgroups
    RClu &c = a.clu[0];
    const double eps2 = c.eps2;

    // first round trip: step and block-time minima, all block times, this
    // lane's j records
    double tv[R_FUSED_TV];
#pragma unroll
    for (int k = 0; k < R_FUSED_TV; k++) {
        const int i = k * NT + tid;
        tv[k] = i < n ? a.tn[off + i] : r_inf();
    }
    double rj[JPL][W];
#pragma unroll
    for (int u = 0; u < JPL; u++) {
        const int j = (slice * JPL + u) * NT + tid;
#pragma unroll
        for (int w = 0; w < W; w++) rj[u][w] = j < n ? a.rp[(size_t)w * a.npad + off + j] : 0.0;
    }
    const RFusedStep st = r_fused_step(a, c);
    // the step's first stamp, also the force stamp, once the loads above
    // are issued (see r_compact_kernel)
    if constexpr (STAMP) r_stamp_glob(a, 0, 0);

    // active list in slot order
    unsigned long long amask[R_FUSED_TV];
#pragma unroll
    for (int k = 0; k < R_FUSED_TV; k++) {
        amask[k] = __ballot(st.stepping && tv[k] == st.tmin);
        if (lane == 0) s_wcnt[k][wave] = __popcll(amask[k]);
    }
    __syncthreads();
    int nact = 0;
    const unsigned long long below = lane == 0 ? 0ULL : (~0ULL) >> (64 - lane);
#pragma unroll
    for (int k = 0; k < R_FUSED_TV; k++) {
        int before = nact;
        for (int w = 0; w < NW; w++) {
            if (w < wave) before += s_wcnt[k][w];
            nact += s_wcnt[k][w];
        }
        if ((amask[k] >> lane) & 1ULL) s_list[before + __popcll(amask[k] & below)] = k * NT + tid;
    }

    // this lane's j, predicted to the block time
    double qj[JPL][SW];
#pragma unroll
    for (int u = 0; u < JPL; u++) {
        double q[Q];
        predict_words<ORDER, false>(rj[u], st.tmin, q);
        if constexpr (MODE == SPLIT) {
            to_split<ORDER>(q, qj[u]);
        } else {
#pragma unroll
            for (int w = 0; w < Q; w++) qj[u][w] = q[w];
        }
    }
    const int nj = min(JPL, max(0, (n - slice * JPL * NT - tid + NT - 1) / NT));  // this lane's j
    const int ngrp = (nact + R_GI - 1) / R_GI;
    const int mygrp = g0 < ngrp && g0 < wps ? (ngrp - g0 + wps - 1) / wps : 0;
    int rb = 0;  // red buffer
    for (int k0 = 0; k0 < mygrp; k0 += NT / R_GI) {
        const int kn = min(NT / R_GI, mygrp - k0);
        __syncthreads();  // s_list written; s_pi free
        // predicted i records of this round's groups, one particle per thread
        if (tid < kn * R_GI) {
            const int lp = min((g0 + (k0 + tid / R_GI) * wps) * R_GI + tid % R_GI, nact - 1);
            const int ip = off + s_list[lp];
            double r[W], q[Q];
#pragma unroll
            for (int w = 0; w < W; w++) r[w] = a.rp[(size_t)w * a.npad + ip];
            predict_words<ORDER, false>(r, st.tmin, q);
            if constexpr (MODE == SPLIT) {
                double t[SW];
                to_split<ORDER>(q, t);
#pragma unroll
                for (int w = 0; w < SW; w++) s_pi[w][tid] = t[w];
            } else {
#pragma unroll
                for (int w = 0; w < Q; w++) s_pi[w][tid] = q[w];
            }
        }
        __syncthreads();
        for (int kk = 0; kk < kn; kk++) {
            const int grp = g0 + (k0 + kk) * wps;
            double A[R_GI][NC];
#pragma unroll
            for (int g = 0; g < R_GI; g++)
#pragma unroll
                for (int q = 0; q < NC; q++) A[g][q] = 0.0;
#pragma unroll
            for (int g = 0; g < R_GI; g++) {
                double pi[SW];
#pragma unroll
                for (int w = 0; w < SW; w++) pi[w] = s_pi[w][kk * R_GI + g];
                float Af[n_f32(ORDER)];
#pragma unroll
                for (int q = 0; q < n_f32(ORDER); q++) Af[q] = 0.0f;
#pragma unroll
                for (int u = 0; u < JPL; u++) {
                    if (u < nj) {
                        if constexpr (MODE == FP64) body_fp64<ORDER, false>(pi, qj[u], eps2, false, 0, A[g]);
                        else body_split<ORDER, false>(pi, qj[u], eps2, false, A[g], Af);
                    }
                }
                if constexpr (MODE == SPLIT) split_fold<ORDER>(A[g], Af);
            }
            // workgroup sums in a fixed order: DPP tree per wave, waves
            // 0..3; the two red buffers alternate, one barrier per group
#pragma unroll
            for (int g = 0; g < R_GI; g++)
#pragma unroll
                for (int q = 0; q < NC; q++) {
                    const double x = wave_sum(A[g][q]);
                    if (lane == 0) red[rb][wave][g * NC + q] = x;
                }
            __syncthreads();
            if (tid < V) {
                double tot = 0.0;
#pragma unroll
                for (int w = 0; w < NW; w++) tot += red[rb][w][tid];
                const int g = tid / NC, q = tid - g * NC;
                const int lp = grp * R_GI + g;
                if (lp < nact) a.part[((size_t)slice * R_FUSED_MAX + s_list[lp]) * NC + q] = tot;
            }
            rb ^= 1;
        }
    }
}

template <int ORDER, bool STAMP>
__global__ __launch_bounds__(NT) void r_fused_correct_kernel(RArgs a, int n, int off, int S)
{
    constexpr int NC = f_words(ORDER), RW = r_words(ORDER);
    constexpr int NW = NT / 64;
    __shared__ double s_wmin[NW];
    __shared__ int s_wcnt[NW];
    __shared__ int s_last;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    RClu &c = a.clu[0];
    const int i = blockIdx.x * NT + tid;
    const bool in = i < n;
    // one round trip: the slot's block time, record and slice partials are
    // loaded before the block time is known (inactive slots discard them)
    const double told = in ? a.tn[off + i] : r_inf();
    RRec<RW> rr;
#pragma unroll
    for (int k = 0; k < RW; k++) rr.v[k] = in && r_used(ORDER, k) ? a.rp[(size_t)k * a.npad + off + i] : 0.0;
    double f[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) f[q] = 0.0;
    if (in)
        for (int sl = 0; sl < S; sl++)
#pragma unroll
            for (int q = 0; q < NC; q++) f[q] += a.part[((size_t)sl * R_FUSED_MAX + i) * NC + q];
    const double etapow = c.etapow, dtmin = c.dtmin;
    const RFusedStep st = r_fused_step(a, c);
    // the stamp once the loads above are issued (see r_compact_kernel)
    if constexpr (STAMP) {
        if (blockIdx.x == 0) r_stamp(a, 3, st.step, wall_clock64());
    }
    const bool act = st.stepping && told == st.tmin;
    double tnew = told;
    if (act) {  // r_finish on the words in registers
        r_correct<ORDER>(rr, f, etapow, st.dtlim);
        double dt = rr.v[R_DT];
        if (dtmin > 0.0 && dt < dtmin) {
            dt = dtmin;
            rr.v[R_DT] = dt;
            atomicAdd((unsigned long long *)&c.dtmin_hits, 1ULL);
        }
        tnew = rr.v[7] + dt;
        a.tn[off + i] = tnew;
#pragma unroll
        for (int k = 0; k < RW; k++)
            if (r_used(ORDER, k)) a.rp[(size_t)k * a.npad + off + i] = rr.v[k];
        if (!isfinite(rr.v[0] + rr.v[1] + rr.v[2] + rr.v[4] + rr.v[5] + rr.v[6])) atomicExch(&a.glob->error, 2);
    }
    // this workgroup's active count, and the minimum of its new block times
    // into the next step's minimum (every slot of the cluster is covered by
    // one thread)
    double m = r_wave_min(tnew);
    const int cnt = __popcll(__ballot(act));
    if (lane == 0) {
        s_wmin[wave] = m;
        s_wcnt[wave] = cnt;
    }
    __syncthreads();
    // the last workgroup does the step's bookkeeping (r_plan_body for one
    // cluster); every wave of a workgroup has read the step counter and the
    // block-time minima, and the workgroup's count and minimum have been
    // performed, before the workgroup counts itself done
    if (tid == 0) {
        int tot = s_wcnt[0];
#pragma unroll
        for (int w = 1; w < NW; w++) {
            m = fmin(m, s_wmin[w]);
            tot += s_wcnt[w];
        }
        if (m < r_inf()) atomicMin(&c.tmin_slot[st.p ^ 1][blockIdx.x % R_SLOTS], r_bits(m));
        if (tot) atomicAdd(&c.acnt[st.p], tot);
        __builtin_amdgcn_s_waitcnt(0);
        s_last = atomicAdd(&a.glob->cdone[st.p], 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && tid == 0) {
        const int nact = __hip_atomic_load(&c.acnt[st.p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        r_fused_bookkeep(a, c, st, nact, n, S);
    }
}

// Force launch of the two-launch step for clusters of at most 1024 or so
// particles, with no dependent round trip: the slots are cut into Si
// i-slices of NT slots (one per lane) and Sj j-slices of JPL NT slots;
// workgroup b takes the combination (b mod Si, (b / Si) mod Sj) and the
// groups b / (Si Sj), b / (Si Sj) + wpc, ... of the active particles of
// its i-slice.  Every load (its i-slice's block times and records, its
// lanes' j records) is issued before the block time is known; the active
// lanes of the i-slice put their predicted records in LDS in slot order.
// A group's i are taken from one i-slice, so active particles crowded
// into few slots leave workgroups of the other i-slices idle; the kernel
// of the global active list above balances those steps.
template <int ORDER, int MODE, int JPL, bool STAMP>
__global__ __launch_bounds__(NT) void r_fused_islice_kernel(RArgs a, int n, int off, int Si, int Sj, int wpc)
{
    static_assert(MODE != EXACT, "EXACT sums j in index order: three-launch step");
    constexpr int W = jp_words(ORDER), Q = pr_words(ORDER);
    constexpr int NC = f_words(ORDER);
    constexpr int V = R_GI * NC;
    constexpr int NW = NT / 64;
    constexpr int SW = MODE == SPLIT ? split_words(ORDER) : Q;
    __shared__ double s_pi[SW][NT];
    __shared__ int s_slot[NT];
    __shared__ int s_wcnt[NW];
    __shared__ double red[2][NW][V];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int combo = blockIdx.x % (Si * Sj), g0 = blockIdx.x / (Si * Sj);
    const int si = combo % Si, sj = combo / Si;
    RClu &c = a.clu[0];
    const double eps2 = c.eps2;

    // the only round trip
    const int islot = si * NT + tid;
    const bool iin = islot < n;
    const double ti = iin ? a.tn[off + islot] : r_inf();
    double ri[W];
#pragma unroll
    for (int w = 0; w < W; w++) ri[w] = iin ? a.rp[(size_t)w * a.npad + off + islot] : 0.0;
    double rj[JPL][W];
#pragma unroll
    for (int u = 0; u < JPL; u++) {
        const int j = (sj * JPL + u) * NT + tid;
#pragma unroll
        for (int w = 0; w < W; w++) rj[u][w] = j < n ? a.rp[(size_t)w * a.npad + off + j] : 0.0;
    }
    const RFusedStep st = r_fused_step(a, c);
    // the step's first stamp, also the force stamp, once the loads above
    // are issued (see r_compact_kernel)
    if constexpr (STAMP) r_stamp_glob(a, 0, 0);

    // the i-slice's active particles, predicted, in slot order
    const bool act = st.stepping && ti == st.tmin;
    const unsigned long long mask = __ballot(act);
    if (lane == 0) s_wcnt[wave] = __popcll(mask);
    if (act) {
        double q[Q];
        predict_words<ORDER, false>(ri, st.tmin, q);
        if constexpr (MODE == SPLIT) {
            double t[SW];
            to_split<ORDER>(q, t);
#pragma unroll
            for (int w = 0; w < SW; w++) ri[w] = t[w];
        } else {
#pragma unroll
            for (int w = 0; w < Q; w++) ri[w] = q[w];
        }
    }
    double qj[JPL][SW];
#pragma unroll
    for (int u = 0; u < JPL; u++) {
        double q[Q];
        predict_words<ORDER, false>(rj[u], st.tmin, q);
        if constexpr (MODE == SPLIT) {
            to_split<ORDER>(q, qj[u]);
        } else {
#pragma unroll
            for (int w = 0; w < Q; w++) qj[u][w] = q[w];
        }
    }
    const int nj = min(JPL, max(0, (n - sj * JPL * NT - tid + NT - 1) / NT));  // this lane's j
    __syncthreads();
    int pos = 0, na = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < wave) pos += s_wcnt[w];
        na += s_wcnt[w];
    }
    if (act) {
        pos += __popcll(mask & (lane == 0 ? 0ULL : (~0ULL) >> (64 - lane)));
#pragma unroll
        for (int w = 0; w < SW; w++) s_pi[w][pos] = ri[w];
        s_slot[pos] = islot;
    }
    __syncthreads();
    const int ngrp = (na + R_GI - 1) / R_GI;
    int rb = 0;  // red buffer
    for (int grp = g0; grp < ngrp; grp += wpc) {
        double A[R_GI][NC];
#pragma unroll
        for (int g = 0; g < R_GI; g++)
#pragma unroll
            for (int q = 0; q < NC; q++) A[g][q] = 0.0;
#pragma unroll
        for (int g = 0; g < R_GI; g++) {
            const int lp = min(grp * R_GI + g, na - 1);
            double pi[SW];
#pragma unroll
            for (int w = 0; w < SW; w++) pi[w] = s_pi[w][lp];
            float Af[n_f32(ORDER)];
#pragma unroll
            for (int q = 0; q < n_f32(ORDER); q++) Af[q] = 0.0f;
#pragma unroll
            for (int u = 0; u < JPL; u++) {
                if (u < nj) {
                    if constexpr (MODE == FP64) body_fp64<ORDER, false>(pi, qj[u], eps2, false, 0, A[g]);
                    else body_split<ORDER, false>(pi, qj[u], eps2, false, A[g], Af);
                }
            }
            if constexpr (MODE == SPLIT) split_fold<ORDER>(A[g], Af);
        }
        // workgroup sums in a fixed order: DPP tree per wave, waves
        // 0..3; the two red buffers alternate, one barrier per group
#pragma unroll
        for (int g = 0; g < R_GI; g++)
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const double x = wave_sum(A[g][q]);
                if (lane == 0) red[rb][wave][g * NC + q] = x;
            }
        __syncthreads();
        if (tid < V) {
            double tot = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) tot += red[rb][w][tid];
            const int g = tid / NC, q = tid - g * NC;
            const int lp = grp * R_GI + g;
            if (lp < na) a.part[((size_t)sj * R_FUSED_MAX + s_slot[lp]) * NC + q] = tot;
        }
        rb ^= 1;
    }
}

}  // namespace hh
