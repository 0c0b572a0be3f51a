// Device code of the Hermite force + predictor back end for gfx950 (CDNA4).
//
// Layout in HBM (one context = one Gravity instance, npad = nbody rounded up
// to a multiple of 256, padded slots hold a massless particle far away):
//   jst  : corrected j-particles, component-major (SoA): jst[w*npad + j] for
//          the 14/20/26 words pos[3] mass vel[3] tlast acc[3] jrk[3]
//          snp[3] crk[3] d4a[3] d5a[3] of the TITAN GParticle
//          (SRC/hermite6-titan.h:20-29).
//   jpr  : predicted j-particles, SoA: jpr[q*npad + j] for the 7/10/13 words
//          pos[3] mass vel[3] acc[3] jrk[3] (SRC/hermite6-titan.h:31-36).
//   part : per-j-slice partial forces part[(s*NC + c)*ipad + i].
// This file is compiled with -ffp-contract=off: every fused multiply-add is
// written explicitly, so the EXACT variant keeps the reference's rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hh {

enum Mode { FP64 = 0, SPLIT = 1, EXACT = 2 };

constexpr int TJ = 256;     // j per LDS tile = threads per workgroup
constexpr int NT = 256;     // threads per workgroup
constexpr double FAR = 1.0e10;  // position of padding particles

// Order codes: 4, 6, 8; 9 is the double-double 8th-order build (positions
// as hi + lo, SRC/hermite8dd.h); 0 and 1 are the stand-alone potential over
// corrected positions (1 with hi + lo positions).
//   j-store (jst) words: order 9 appends posL (26-28) to the order-8 words.
//   predictions (jpr): order 9 is posH m vel acc jrk posL (16 words).
__host__ __device__ constexpr int jp_words(int o) { return o == 4 ? 14 : o == 6 ? 20 : o == 8 ? 26 : 29; }
__host__ __device__ constexpr int pr_words(int o)
{
    return o == 4 ? 7 : o == 6 ? 10 : o == 8 ? 13 : o == 9 ? 16 : o == 1 ? 7 : 4;
}
__host__ __device__ constexpr int f_words(int o) { return o == 4 ? 6 : o == 6 ? 9 : (o == 8 || o == 9) ? 12 : 0; }
__host__ __device__ constexpr int p_words(int o) { return o == 4 ? 24 : o == 6 ? 28 : o == 8 ? 36 : 40; }
// words per j record in LDS, even so records can be read with 16-byte loads
__host__ __device__ constexpr int lds_words(int o)
{
    return o == 4 ? 8 : o == 6 ? 10 : o == 8 ? 14 : o == 9 ? 16 : o == 1 ? 8 : 4;
}
// Source word of record word w for the tile kernels: the double-double
// potential (order 1) reads posH, mass and posL (j-store words 26-28).
__host__ __device__ constexpr int src_word(int o, int w) { return (o == 1 && w >= 4) ? 22 + w : w; }

// ------------------------------------------------------------------ rsqrt
// 1/sqrt(x) to FP64 accuracy: an FP32 v_rsq_f32 seed y0 refined with the
// second-order truncation of (1-h)^(-1/2) = 1 + h/2 + 3h^2/8 + ..., where
// h = 1 - x*y0^2 (|h| ~ 2^-23, so the dropped 5h^3/16 term is ~2^-70).
// y0 has a 24-bit significand, so y0*y0 is exact in FP64 and h carries a
// single rounding.  Valid for FLT_MIN < x < FLT_MAX.
__device__ __forceinline__ double rsqrt64(double x)
{
    const float xf = (float)x;
    const double y = (double)__builtin_amdgcn_rsqf(xf);
    const double y2 = y * y;
    const double h = __builtin_fma(-x, y2, 1.0);
    const double p = __builtin_fma(h, 0.375, 0.5);
    return __builtin_fma(y * h, p, y);
}

// --------------------------------------------------------------- predictor
// Taylor prediction to tsys (Horner form of SRC/hermite6-titan.cu:153-227,
// SRC/hermite8-titan.cu:20-134, generic SRC/hermite{4,6,8}.h GPredictor).
// EXACT keeps the generic back end's unfused x + dt*(...) rounding.
template <bool EX>
__device__ __forceinline__ double madd(double a, double b, double c)
{
    if constexpr (EX) return c + a * b;
    else return __builtin_fma(a, b, c);
}

// Source word w of slot i: the dirty staging record or the SoA store.
template <int ORDER, bool EX>
__device__ __forceinline__ void predict_words(const double (&r)[jp_words(ORDER)], double tsys, double (&q)[pr_words(ORDER)])
{
    const double dt = tsys - r[7];
    const double dt2 = (1. / 2.) * dt;
    const double dt3 = (1. / 3.) * dt;
    const double dt4 = (1. / 4.) * dt;
    const double dt5 = (1. / 5.) * dt;
    const double dt6 = (1. / 6.) * dt;
    const double dt7 = (1. / 7.) * dt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double x = r[0 + k], v = r[4 + k], a = r[8 + k], j = r[11 + k];
        if constexpr (ORDER == 4) {
            q[0 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, j, a), v), x);
            q[4 + k] = madd<EX>(dt, madd<EX>(dt2, j, a), v);
        } else if constexpr (ORDER == 6) {
            const double s = r[14 + k], c = r[17 + k];
            q[0 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, madd<EX>(dt5, c, s), j), a), v), x);
            q[4 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, c, s), j), a), v);
            q[7 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, c, s), j), a);
        } else {
            const double s = r[14 + k], c = r[17 + k], d4 = r[20 + k], d5 = r[23 + k];
            // double-double: posH is kept and the displacement goes to posL
            // (SRC/hermite8dd.h:291-311)
            const double xs = ORDER == 9 ? r[26 + k] : x;
            const double xp = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, madd<EX>(dt5, madd<EX>(dt6, madd<EX>(dt7, d5, d4), c), s), j), a), v), xs);
            if constexpr (ORDER == 9) {
                q[0 + k] = x;
                q[13 + k] = xp;
            } else {
                q[0 + k] = xp;
            }
            q[4 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, madd<EX>(dt5, madd<EX>(dt6, d5, d4), c), s), j), a), v);
            q[7 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, madd<EX>(dt5, d5, d4), c), s), j), a);
            q[10 + k] = madd<EX>(dt, madd<EX>(dt2, madd<EX>(dt3, madd<EX>(dt4, d5, d4), c), s), j);
        }
    }
    q[3] = r[3];
}

// One thread per slot: slots below ndirty take their record from the
// uploaded AoS staging buffer and write it into the SoA store (the upload of
// SRC/hermite6-titan.cu:270 fused with the predict kernel :246-266); all
// slots are then predicted to tsys.  predict == false only scatters.
template <int ORDER, bool EX, bool PREDICT>
__global__ __launch_bounds__(256) void predict_kernel(int npad, int ndirty, const double *__restrict__ stage,
                                                      double *__restrict__ jst, double *__restrict__ jpr,
                                                      double tsys)
{
    constexpr int W = jp_words(ORDER), Q = pr_words(ORDER);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    double r[W];
    if (i < ndirty) {
#pragma unroll
        for (int w = 0; w < W; w++) r[w] = stage[(size_t)i * W + w];
#pragma unroll
        for (int w = 0; w < W; w++) jst[(size_t)w * npad + i] = r[w];
    } else {
        if constexpr (!PREDICT) return;
#pragma unroll
        for (int w = 0; w < W; w++) r[w] = jst[(size_t)w * npad + i];
    }
    if constexpr (PREDICT) {
        double q[Q];
        predict_words<ORDER, EX>(r, tsys, q);
#pragma unroll
        for (int w = 0; w < Q; w++) jpr[(size_t)w * npad + i] = q[w];
    }
}

// Fill the padding slots [nbody, npad) of the SoA store with massless
// particles at (FAR, FAR, FAR) and everything else zero.
template <int ORDER>
__global__ void pad_kernel(int nbody, int npad, double *__restrict__ jst, double *__restrict__ jpr)
{
    constexpr int W = jp_words(ORDER), Q = pr_words(ORDER);
    const int i = nbody + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
#pragma unroll
    for (int w = 0; w < W; w++) jst[(size_t)w * npad + i] = (w < 3) ? FAR : 0.0;
#pragma unroll
    for (int w = 0; w < Q; w++) jpr[(size_t)w * npad + i] = (w < 3) ? FAR : 0.0;
}

// --------------------------------------------------------- force / potential
struct TileArgs {
    const double *__restrict__ src;  // SoA j source (jpr for forces, jst for the potential)
    int npad;                        // component stride of src
    int nbody;                       // real j count
    int ntiles;                      // ceil(nbody / TJ)
    int tiles_per_slice;
    int nslice;
    int nact;                        // i count (i-particles are src slots [0, nact))
    int ipad;                        // i-blocks * i per block
    double eps2;
    int pot_mask_r2;                 // potential self-exclusion by r2 > eps2 (order 4)
    double *__restrict__ part;       // slice partials part[(s*NC + c)*ipad + i] (nslice > 1)
    double *__restrict__ fout;       // [nact][F] (ORDER > 0) or [nact] (ORDER 0)
    double *__restrict__ pout;       // fused potential [nact] or null
};

// Accumulator count: force words (+1 with the fused potential); ORDER 0 is
// the stand-alone potential.
__host__ __device__ constexpr int n_acc(int order, bool pot) { return order <= 1 ? 1 : f_words(order) + (pot ? 1 : 0); }

// ---- interaction bodies -------------------------------------------------
// q: j record (pos[3] m vel[3] acc[3] jrk[3]); p: i record; A: accumulators
// (acc[3] jrk[3] snp[3] crk[3] [pot]).  self: j is the i particle itself.

// EXACT: operation order of the generic CPU back end, IEEE division and
// square root (SRC/hermite4.h:347-357, SRC/hermite6.h:332-360,
// SRC/hermite8.h:385-425; potential SRC/hermite4.h:373-384,
// SRC/hermite6.h:376-386).
template <int ORDER, bool POT>
__device__ __forceinline__ void body_exact(const double *p, const double *q, double eps2, bool self, int mask_r2, double *A)
{
    if constexpr (ORDER <= 1) {
        // order 1: (posHj - posHi) + (posLj - posLi), SRC/hermite8dd.h:391-413
        const double dx = ORDER == 1 ? (q[0] - p[0]) + (q[4] - p[4]) : q[0] - p[0];
        const double dy = ORDER == 1 ? (q[1] - p[1]) + (q[5] - p[5]) : q[1] - p[1];
        const double dz = ORDER == 1 ? (q[2] - p[2]) + (q[6] - p[6]) : q[2] - p[2];
        const double r2 = eps2 + (dx * dx + dy * dy + dz * dz);
        const double mj = (mask_r2 ? (r2 > eps2) : !self) ? q[3] : 0.0;
        A[0] -= mj / __builtin_sqrt(r2);
        return;
    } else {
        double dr[3], dv[3], da[3], dj[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            // double-double: (posHj - posHi) + (posLj - posLi), SRC/hermite8dd.h:358
            dr[k] = ORDER == 9 ? (q[0 + k] - p[0 + k]) + (q[13 + k] - p[13 + k]) : q[0 + k] - p[0 + k];
            dv[k] = q[4 + k] - p[4 + k];
        }
        const double r2 = eps2 + (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        if constexpr (POT) {
            const double mj = self ? 0.0 : q[3];
            A[f_words(ORDER)] -= mj / __builtin_sqrt(r2);
        }
        // the double-double build skips r2 == eps2 (SRC/hermite8dd.h:364)
        if (ORDER == 9 && r2 == eps2) return;
        const double drdv = dr[0] * dv[0] + dr[1] * dv[1] + dr[2] * dv[2];
        const double ri2 = 1.0 / r2;
        const double sq = __builtin_sqrt(ri2);
        const double mri3 = q[3] * ri2 * sq;
        if constexpr (ORDER == 4) {
            const double alpha = -3.0 * drdv * ri2;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                A[k] = A[k] + mri3 * dr[k];
                A[3 + k] = A[3 + k] + (mri3 * dv[k] + alpha * (mri3 * dr[k]));
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) da[k] = q[7 + k] - p[7 + k];
            const double dvdv = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const double drda = dr[0] * da[0] + dr[1] * da[1] + dr[2] * da[2];
            const double alpha = drdv * ri2;
            const double beta = (dvdv + drda) * ri2 + alpha * alpha;
            if constexpr (ORDER == 6) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    A[k] = A[k] + mri3 * dr[k];
                    const double tmp = dv[k] + (-3.0 * alpha) * dr[k];
                    A[3 + k] = A[3 + k] + mri3 * tmp;
                    A[6 + k] = A[6 + k] + mri3 * (da[k] + (-6.0 * alpha) * tmp + (-3.0 * beta) * dr[k]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 3; k++) dj[k] = q[10 + k] - p[10 + k];
                const double dvda = dv[0] * da[0] + dv[1] * da[1] + dv[2] * da[2];
                const double drdj = dr[0] * dj[0] + dr[1] * dj[1] + dr[2] * dj[2];
                const double gamma = (3.0 * dvda + drdj) * ri2 + alpha * (3.0 * beta - 4.0 * alpha * alpha);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    A[k] = A[k] + mri3 * dr[k];
                    const double t1 = dv[k] + (-3.0 * alpha) * dr[k];
                    A[3 + k] = A[3 + k] + mri3 * t1;
                    const double t2 = da[k] + (-6.0 * alpha) * t1 + (-3.0 * beta) * dr[k];
                    A[6 + k] = A[6 + k] + mri3 * t2;
                    const double t3 = dj[k] + (-9.0 * alpha) * t2 + (-9.0 * beta) * t1 + (-3.0 * gamma) * dr[k];
                    A[9 + k] = A[9 + k] + mri3 * t3;
                }
            }
        }
    }
}

// FP64: same chain rule with explicit FMAs and the refined rsqrt
// (the TITAN pp_interact of SRC/hermite{4,6,8}-titan.cu reorganised for
// fused multiply-adds).
template <int ORDER, bool POT>
__device__ __forceinline__ void body_fp64(const double *p, const double *q, double eps2, bool self, int mask_r2, double *A)
{
    // relative position; hi + lo positions for the double-double orders
    constexpr int LO = ORDER == 9 ? 13 : 4;
    constexpr bool DD = ORDER == 9 || ORDER == 1;
    const double dx = DD ? (q[0] - p[0]) + (q[LO] - p[LO]) : q[0] - p[0];
    const double dy = DD ? (q[1] - p[1]) + (q[LO + 1] - p[LO + 1]) : q[1] - p[1];
    const double dz = DD ? (q[2] - p[2]) + (q[LO + 2] - p[LO + 2]) : q[2] - p[2];
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const double rinv = rsqrt64(r2);
    if constexpr (ORDER <= 1) {
        const bool keep = mask_r2 ? (r2 > eps2) : !self;
        A[0] = __builtin_fma(keep ? -q[3] : 0.0, rinv, A[0]);
        return;
    } else {
        const double dvx = q[4] - p[4], dvy = q[5] - p[5], dvz = q[6] - p[6];
        const double rinv2 = rinv * rinv;
        const double mr = q[3] * rinv;
        // the double-double build skips r2 == eps2 (SRC/hermite8dd.h:364)
        const double mr3 = (ORDER == 9 && r2 == eps2) ? 0.0 : mr * rinv2;
        if constexpr (POT) A[f_words(ORDER)] = __builtin_fma(self ? 0.0 : -1.0, mr, A[f_words(ORDER)]);
        const double drdv = __builtin_fma(dz, dvz, __builtin_fma(dy, dvy, dx * dvx));
        A[0] = __builtin_fma(mr3, dx, A[0]);
        A[1] = __builtin_fma(mr3, dy, A[1]);
        A[2] = __builtin_fma(mr3, dz, A[2]);
        if constexpr (ORDER == 4) {
            const double a3 = -3.0 * (drdv * rinv2);
            A[3] = __builtin_fma(mr3, __builtin_fma(a3, dx, dvx), A[3]);
            A[4] = __builtin_fma(mr3, __builtin_fma(a3, dy, dvy), A[4]);
            A[5] = __builtin_fma(mr3, __builtin_fma(a3, dz, dvz), A[5]);
        } else {
            const double dax = q[7] - p[7], day = q[8] - p[8], daz = q[9] - p[9];
            const double alpha = drdv * rinv2;
            const double dvdv_drda = __builtin_fma(dz, daz, __builtin_fma(dy, day, __builtin_fma(dx, dax,
                                     __builtin_fma(dvz, dvz, __builtin_fma(dvy, dvy, dvx * dvx)))));
            const double beta = __builtin_fma(dvdv_drda, rinv2, alpha * alpha);
            const double a3 = -3.0 * alpha;
            const double tx = __builtin_fma(a3, dx, dvx), ty = __builtin_fma(a3, dy, dvy), tz = __builtin_fma(a3, dz, dvz);
            A[3] = __builtin_fma(mr3, tx, A[3]);
            A[4] = __builtin_fma(mr3, ty, A[4]);
            A[5] = __builtin_fma(mr3, tz, A[5]);
            const double a6 = 2.0 * a3, b3 = -3.0 * beta;
            const double ux = __builtin_fma(b3, dx, __builtin_fma(a6, tx, dax));
            const double uy = __builtin_fma(b3, dy, __builtin_fma(a6, ty, day));
            const double uz = __builtin_fma(b3, dz, __builtin_fma(a6, tz, daz));
            A[6] = __builtin_fma(mr3, ux, A[6]);
            A[7] = __builtin_fma(mr3, uy, A[7]);
            A[8] = __builtin_fma(mr3, uz, A[8]);
            if constexpr (ORDER >= 8) {
                const double djx = q[10] - p[10], djy = q[11] - p[11], djz = q[12] - p[12];
                const double dvda = __builtin_fma(dvz, daz, __builtin_fma(dvy, day, dvx * dax));
                const double drdj = __builtin_fma(dz, djz, __builtin_fma(dy, djy, dx * djx));
                // gamma = (3 dv.da + dr.dj)/r^2 + alpha (3 beta - 4 alpha^2)
                const double g0 = __builtin_fma(3.0, dvda, drdj);
                const double g1 = __builtin_fma(-4.0 * alpha, alpha, 3.0 * beta);
                const double gamma = __builtin_fma(g0, rinv2, alpha * g1);
                const double a9 = 3.0 * a3, b9 = 3.0 * b3, g3 = -3.0 * gamma;
                const double cx = __builtin_fma(g3, dx, __builtin_fma(b9, tx, __builtin_fma(a9, ux, djx)));
                const double cy = __builtin_fma(g3, dy, __builtin_fma(b9, ty, __builtin_fma(a9, uy, djy)));
                const double cz = __builtin_fma(g3, dz, __builtin_fma(b9, tz, __builtin_fma(a9, uz, djz)));
                A[9] = __builtin_fma(mr3, cx, A[9]);
                A[10] = __builtin_fma(mr3, cy, A[10]);
                A[11] = __builtin_fma(mr3, cz, A[11]);
            }
        }
    }
}

// SPLIT: per-family precision.  A relative error of 1e-7 in the
// acceleration, the level of an FP32 interaction body, makes the 6th- and
// 8th-order correctors collapse (their reconstructed derivatives scale the
// force differences by (2/dt)^5 and (2/dt)^7, SRC/hermite6.h:168-170,
// SRC/hermite8.h:203-206), while the jerk tolerates 1e-6 at 6th order and
// snap and crackle 1e-5.  So the acceleration family (and at 8th order the
// jerk) is evaluated and accumulated in FP64: FP64 r^2, m r^-3 from the
// v_rsq_f32 seed y0 corrected to first order in h = 1 - r^2 y0^2 (relative
// error <= (15/8) h^2 ~ 2^-42).  The rest of the chain rule runs in FP32:
// relative velocity, acceleration and jerk from the FP32-rounded
// predictions (the split record below), relative position and the scalars
// rounded from FP64.  Force words [split_lo, F) go to the FP32 partial sums
// Af, which the caller folds into FP64 every few j; everything else goes to
// the FP64 accumulators A.
__host__ __device__ constexpr int split_lo(int o) { return o == 8 ? 6 : 3; }
__host__ __device__ constexpr int n_f32(int o) { return (o == 4 || o == 6 || o == 8) ? f_words(o) - split_lo(o) : 1; }

// Record of the split body (j in LDS or registers, i in registers): pos[3]
// and mass in FP64, then the velocity (FP64 at 8th order, where the jerk is
// FP64 arithmetic), the acceleration and the jerk rounded to FP32, two to a
// double word: order 4 {x y z m (vx vy) (vz -)}, order 6 {x y z m (vx vy)
// (vz ax) (ay az) -}, order 8 {x y z m vx vy vz (ax ay) (az jx) (jy jz)}.
// The FP32 words are rounded once per record instead of once per pair.
__host__ __device__ constexpr int split_words(int o) { return o == 4 ? 6 : o == 6 ? 8 : 10; }
// LDS stride of split records: an odd number of 16-byte words, so the
// staging stores of consecutive lanes fall in different banks
__host__ __device__ constexpr int split_stride(int o) { return o == 6 ? 10 : split_words(o); }

__device__ __forceinline__ double pack2f(float lo, float hi)
{
    return __longlong_as_double((long long)(((unsigned long long)__float_as_uint(hi) << 32) | __float_as_uint(lo)));
}
__device__ __forceinline__ float lo_f(double d) { return __uint_as_float((unsigned)__double_as_longlong(d)); }
__device__ __forceinline__ float hi_f(double d)
{
    return __uint_as_float((unsigned)((unsigned long long)__double_as_longlong(d) >> 32));
}

// Prediction record (pos m vel acc jrk, FP64) -> split record.
template <int ORDER>
__device__ __forceinline__ void to_split(const double *r, double *s)
{
    s[0] = r[0];
    s[1] = r[1];
    s[2] = r[2];
    s[3] = r[3];
    if constexpr (ORDER == 4) {
        s[4] = pack2f((float)r[4], (float)r[5]);
        s[5] = pack2f((float)r[6], 0.0f);
    } else if constexpr (ORDER == 6) {
        s[4] = pack2f((float)r[4], (float)r[5]);
        s[5] = pack2f((float)r[6], (float)r[7]);
        s[6] = pack2f((float)r[8], (float)r[9]);
        s[7] = 0.0;
    } else {
        s[4] = r[4];
        s[5] = r[5];
        s[6] = r[6];
        s[7] = pack2f((float)r[7], (float)r[8]);
        s[8] = pack2f((float)r[9], (float)r[10]);
        s[9] = pack2f((float)r[11], (float)r[12]);
    }
}

// p, q: split records of the i and the j particle.
template <int ORDER, bool POT>
__device__ __forceinline__ void body_split(const double *p, const double *q, double eps2, bool self, double *A, float *Af)
{
    static_assert(ORDER == 4 || ORDER == 6 || ORDER == 8, "split-precision body for the force orders only");
    const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const float yf = __builtin_amdgcn_rsqf((float)r2);
    const double y = (double)yf;
    const double y2 = y * y;  // exact: y has 24 significant bits
    const double h = __builtin_fma(-r2, y2, 1.0);
    // m r^-3 = m y^3 (1 - h)^(-3/2) to first order in h: relative error
    // <= (15/8) h^2 ~ 2^-42 (and m r^-1 = m y (1 - h)^(-1/2) for the
    // potential, <= (3/8) h^2)
    const double my = q[3] * y;
    const double my3 = my * y2;
    const double mr3 = __builtin_fma(my3, 1.5 * h, my3);
    if constexpr (POT) {
        const double mr = __builtin_fma(my, 0.5 * h, my);
        A[f_words(ORDER)] = __builtin_fma(self ? 0.0 : -1.0, mr, A[f_words(ORDER)]);
    }
    A[0] = __builtin_fma(mr3, dx, A[0]);
    A[1] = __builtin_fma(mr3, dy, A[1]);
    A[2] = __builtin_fma(mr3, dz, A[2]);
    const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
    // r^-2 = y^2 (1 - h)^(-1) to first order, rounded to FP32 for the FP32
    // chain (yf*yf alone, 2^-22 off, raised the 6th-order N = 65536 run's
    // |dE/E| at t = 1 from 2.4e-12 to 4.1e-12)
    const float ri2f = (float)__builtin_fma(y2, h, y2), mr3f = (float)mr3;
    if constexpr (ORDER == 4) {
        const float dvx = lo_f(q[4]) - lo_f(p[4]), dvy = hi_f(q[4]) - hi_f(p[4]), dvz = lo_f(q[5]) - lo_f(p[5]);
        const float drdv = __builtin_fmaf(dzf, dvz, __builtin_fmaf(dyf, dvy, dxf * dvx));
        const float a3 = -3.0f * (drdv * ri2f);
        Af[0] = __builtin_fmaf(mr3f, __builtin_fmaf(a3, dxf, dvx), Af[0]);
        Af[1] = __builtin_fmaf(mr3f, __builtin_fmaf(a3, dyf, dvy), Af[1]);
        Af[2] = __builtin_fmaf(mr3f, __builtin_fmaf(a3, dzf, dvz), Af[2]);
    } else if constexpr (ORDER == 6) {
        const float dvx = lo_f(q[4]) - lo_f(p[4]), dvy = hi_f(q[4]) - hi_f(p[4]), dvz = lo_f(q[5]) - lo_f(p[5]);
        const float dax = hi_f(q[5]) - hi_f(p[5]), day = lo_f(q[6]) - lo_f(p[6]), daz = hi_f(q[6]) - hi_f(p[6]);
        const float drdv = __builtin_fmaf(dzf, dvz, __builtin_fmaf(dyf, dvy, dxf * dvx));
        const float alpha = drdv * ri2f;
        const float s = __builtin_fmaf(dzf, daz, __builtin_fmaf(dyf, day, __builtin_fmaf(dxf, dax,
                        __builtin_fmaf(dvz, dvz, __builtin_fmaf(dvy, dvy, dvx * dvx)))));
        const float beta = __builtin_fmaf(s, ri2f, alpha * alpha);
        const float a3 = -3.0f * alpha;
        const float tx = __builtin_fmaf(a3, dxf, dvx), ty = __builtin_fmaf(a3, dyf, dvy), tz = __builtin_fmaf(a3, dzf, dvz);
        Af[0] = __builtin_fmaf(mr3f, tx, Af[0]);
        Af[1] = __builtin_fmaf(mr3f, ty, Af[1]);
        Af[2] = __builtin_fmaf(mr3f, tz, Af[2]);
        const float a6 = 2.0f * a3, b3 = -3.0f * beta;
        Af[3] = __builtin_fmaf(mr3f, __builtin_fmaf(b3, dxf, __builtin_fmaf(a6, tx, dax)), Af[3]);
        Af[4] = __builtin_fmaf(mr3f, __builtin_fmaf(b3, dyf, __builtin_fmaf(a6, ty, day)), Af[4]);
        Af[5] = __builtin_fmaf(mr3f, __builtin_fmaf(b3, dzf, __builtin_fmaf(a6, tz, daz)), Af[5]);
    } else {
        // the jerk in FP64 as well
        const double rinv2 = __builtin_fma(y2, h, y2);
        const double dvx = q[4] - p[4], dvy = q[5] - p[5], dvz = q[6] - p[6];
        const double drdv = __builtin_fma(dz, dvz, __builtin_fma(dy, dvy, dx * dvx));
        const double alpha = drdv * rinv2;
        const double a3 = -3.0 * alpha;
        const double tx = __builtin_fma(a3, dx, dvx), ty = __builtin_fma(a3, dy, dvy), tz = __builtin_fma(a3, dz, dvz);
        A[3] = __builtin_fma(mr3, tx, A[3]);
        A[4] = __builtin_fma(mr3, ty, A[4]);
        A[5] = __builtin_fma(mr3, tz, A[5]);
        const float dvxf = (float)dvx, dvyf = (float)dvy, dvzf = (float)dvz;
        const float txf = (float)tx, tyf = (float)ty, tzf = (float)tz, alphaf = (float)alpha;
        const float dax = lo_f(q[7]) - lo_f(p[7]), day = hi_f(q[7]) - hi_f(p[7]), daz = lo_f(q[8]) - lo_f(p[8]);
        const float djx = hi_f(q[8]) - hi_f(p[8]), djy = lo_f(q[9]) - lo_f(p[9]), djz = hi_f(q[9]) - hi_f(p[9]);
        const float s = __builtin_fmaf(dzf, daz, __builtin_fmaf(dyf, day, __builtin_fmaf(dxf, dax,
                        __builtin_fmaf(dvzf, dvzf, __builtin_fmaf(dvyf, dvyf, dvxf * dvxf)))));
        const float beta = __builtin_fmaf(s, ri2f, alphaf * alphaf);
        const float a6 = -6.0f * alphaf, b3 = -3.0f * beta;
        const float ux = __builtin_fmaf(b3, dxf, __builtin_fmaf(a6, txf, dax));
        const float uy = __builtin_fmaf(b3, dyf, __builtin_fmaf(a6, tyf, day));
        const float uz = __builtin_fmaf(b3, dzf, __builtin_fmaf(a6, tzf, daz));
        Af[0] = __builtin_fmaf(mr3f, ux, Af[0]);
        Af[1] = __builtin_fmaf(mr3f, uy, Af[1]);
        Af[2] = __builtin_fmaf(mr3f, uz, Af[2]);
        const float dvda = __builtin_fmaf(dvzf, daz, __builtin_fmaf(dvyf, day, dvxf * dax));
        const float drdj = __builtin_fmaf(dzf, djz, __builtin_fmaf(dyf, djy, dxf * djx));
        // gamma = (3 dv.da + dr.dj)/r^2 + alpha (3 beta - 4 alpha^2)
        const float g0 = __builtin_fmaf(3.0f, dvda, drdj);
        const float g1 = __builtin_fmaf(-4.0f * alphaf, alphaf, 3.0f * beta);
        const float gamma = __builtin_fmaf(g0, ri2f, alphaf * g1);
        const float a9 = -9.0f * alphaf, b9 = 3.0f * b3, g3 = -3.0f * gamma;
        Af[3] = __builtin_fmaf(mr3f, __builtin_fmaf(g3, dxf, __builtin_fmaf(b9, txf, __builtin_fmaf(a9, ux, djx))), Af[3]);
        Af[4] = __builtin_fmaf(mr3f, __builtin_fmaf(g3, dyf, __builtin_fmaf(b9, tyf, __builtin_fmaf(a9, uy, djy))), Af[4]);
        Af[5] = __builtin_fmaf(mr3f, __builtin_fmaf(g3, dzf, __builtin_fmaf(b9, tzf, __builtin_fmaf(a9, uz, djz))), Af[5]);
    }
}

// Adds the FP32 partial sums of the split body to the FP64 accumulators and
// clears them.
template <int ORDER>
__device__ __forceinline__ void split_fold(double *A, float *Af)
{
#pragma unroll
    for (int c = 0; c < n_f32(ORDER); c++) {
        A[split_lo(ORDER) + c] += (double)Af[c];
        Af[c] = 0.0f;
    }
}

// SPLIT: j per FP32 partial sum before it is folded into the FP64
// accumulators (a template parameter FOLD of tile_kernel, tuned per order;
// JFOLD consecutive j of a lane in the j-parallel kernels).
constexpr int JFOLD = 8;

// ---- i-parallel tile kernel ----------------------------------------------
// Workgroup (iblock, slice): NT threads each own NI i-particles
// (i = iblock*NT*NI + k*NT + tid); the slice's j-tiles are staged through LDS
// as j-major records (all lanes then read the same address: LDS broadcast),
// the next tile's global loads in flight while the current tile is used.
// With nslice > 1 every workgroup writes its partial sums; the last one to
// finish an i-block (device-scope counter) adds the slices in slice order,
// so the result does not depend on scheduling.  ORDER 0 is the potential.
template <int ORDER, int MODE, int NI, bool POT, int FOLD>
__global__ __launch_bounds__(NT) void tile_kernel(TileArgs a)
{
    static_assert(TJ % FOLD == 0, "FOLD must divide the tile");
    constexpr int Q = pr_words(ORDER);
    constexpr int L = lds_words(ORDER);
    constexpr int NC = n_acc(ORDER, POT);
    // i words actually used: pos (+ vel, acc, jrk); mass not needed
    constexpr int QI = ORDER == 0 ? 3 : Q;  // order 1: posH, mass (unused), posL
    // the split body takes split records (FP32 words packed): LDS stride LS
    constexpr bool SR = MODE == SPLIT && ORDER > 1;
    constexpr int LS = SR ? split_words(ORDER) : L;
    constexpr int LST = SR ? split_stride(ORDER) : L;
    __shared__ double sj[TJ * L];

    const int tid = threadIdx.x;
    const int iblock = blockIdx.x, slice = blockIdx.y;
    const int npad = a.npad;

    double pi[NI][QI];
    int ii[NI];
#pragma unroll
    for (int k = 0; k < NI; k++) {
        ii[k] = iblock * NT * NI + k * NT + tid;
        const int idx = ii[k] < a.nact ? ii[k] : a.nact - 1;
#pragma unroll
        for (int w = 0; w < QI; w++) pi[k][w] = a.src[(size_t)src_word(ORDER, w) * npad + idx];
        if constexpr (SR) {
            double t[LS];
            to_split<ORDER>(pi[k], t);
#pragma unroll
            for (int w = 0; w < LS; w++) pi[k][w] = t[w];
        }
    }
    double A[NI][NC];
#pragma unroll
    for (int k = 0; k < NI; k++)
#pragma unroll
        for (int c = 0; c < NC; c++) A[k][c] = 0.0;

    const int t0 = slice * a.tiles_per_slice;
    const int t1 = min(t0 + a.tiles_per_slice, a.ntiles);

    double jr[Q];
    if (t0 < t1) {
#pragma unroll
        for (int w = 0; w < Q; w++) jr[w] = a.src[(size_t)src_word(ORDER, w) * npad + t0 * TJ + tid];
    }
    // a wave whose lanes are all past the i range only helps stage tiles
    const bool busy = iblock * NT * NI + (tid & ~63) < a.nact;

    for (int t = t0; t < t1; t++) {
        __syncthreads();
        if constexpr (SR) {
            double t[LS];
            to_split<ORDER>(jr, t);
#pragma unroll
            for (int w = 0; w < LS; w++) sj[tid * LST + w] = t[w];
        } else {
#pragma unroll
            for (int w = 0; w < Q; w++) sj[tid * L + w] = jr[w];
        }
        __syncthreads();
        if (t + 1 < t1) {
#pragma unroll
            for (int w = 0; w < Q; w++) jr[w] = a.src[(size_t)src_word(ORDER, w) * npad + (t + 1) * TJ + tid];
        }
        if (!busy) continue;
        const int jbase = t * TJ;
        if constexpr (MODE == EXACT) {
            const int jn = min(TJ, a.nbody - jbase);
            for (int jj = 0; jj < jn; jj++) {
                double q[L];
#pragma unroll
                for (int w = 0; w < L; w++) q[w] = sj[jj * L + w];
#pragma unroll
                for (int k = 0; k < NI; k++)
                    body_exact<ORDER, POT>(pi[k], q, a.eps2, jbase + jj == ii[k], a.pot_mask_r2, A[k]);
            }
        } else if constexpr (MODE == FP64) {
#pragma unroll 2
            for (int jj = 0; jj < TJ; jj++) {
                double q[L];
#pragma unroll
                for (int w = 0; w < L; w++) q[w] = sj[jj * L + w];
#pragma unroll
                for (int k = 0; k < NI; k++)
                    body_fp64<ORDER, POT>(pi[k], q, a.eps2, jbase + jj == ii[k], a.pot_mask_r2, A[k]);
            }
        } else {
            static_assert(ORDER > 1, "the potential runs in FP64 for every variant");
            for (int jb = 0; jb < TJ; jb += FOLD) {
                float Af[NI][n_f32(ORDER)];
#pragma unroll
                for (int k = 0; k < NI; k++)
#pragma unroll
                    for (int c = 0; c < n_f32(ORDER); c++) Af[k][c] = 0.0f;
#pragma unroll 2
                for (int jj = jb; jj < jb + FOLD; jj++) {
                    double q[LS];
#pragma unroll
                    for (int w = 0; w < LS; w++) q[w] = sj[jj * LST + w];
#pragma unroll
                    for (int k = 0; k < NI; k++)
                        body_split<ORDER, POT>(pi[k], q, a.eps2, jbase + jj == ii[k], A[k], Af[k]);
                }
#pragma unroll
                for (int k = 0; k < NI; k++) split_fold<ORDER>(A[k], Af[k]);
            }
        }
    }

    constexpr int F = ORDER <= 1 ? 1 : f_words(ORDER);
    if (a.nslice == 1) {
#pragma unroll
        for (int k = 0; k < NI; k++) {
            if (ii[k] < a.nact) {
#pragma unroll
                for (int c = 0; c < F; c++) a.fout[(size_t)ii[k] * F + c] = A[k][c];
                if constexpr (POT) a.pout[ii[k]] = A[k][F];
            }
        }
        return;
    }
    // several slices: partial sums, added by slice_reduce_kernel.  No
    // cross-workgroup handshake here: on gfx950 an agent-scope fence is an L2
    // write-back plus invalidate, far dearer than the next kernel boundary.
    const size_t ipad = (size_t)a.ipad;
#pragma unroll
    for (int k = 0; k < NI; k++)
        if (ii[k] < a.nact) {
#pragma unroll
            for (int c = 0; c < NC; c++) a.part[((size_t)slice * NC + c) * ipad + ii[k]] = A[k][c];
        }
}

// Sum of the j-slice partials part[(s*nc + c)*ipad + i] over s, for i in
// [0, ni), written as fout[i*f + c] (c < f) or pout[i] (c == f).  Each sum
// is split over RL = 8 lanes (lane l takes s = l, l + 8, ...) and the lanes
// are combined by a fixed DPP tree (lane_tree8), so the order of additions,
// and the result, depend only on nslice.  (The reference reduces its NJBLOCK
// partials with a warp-shuffle tree, SRC/hermite6-titan.cu:441-459.)
constexpr int RL = 8;

// Cross-lane moves through DPP: the data stay in the vector ALU (a shuffle,
// ds_bpermute, goes through the LDS crossbar and back).  A double moves as
// its two 32-bit halves; lanes of rows that ROW_MASK leaves out get old.
// Controls (GFX9 DPP): quad permutations 1,0,3,2 and 2,3,0,1, the mirror of
// each half row (lane i <-> 7 - i within 8) and of each row (i <-> 15 - i
// within 16), and the broadcast of lane 15 (31) of a row into the rows that
// follow.
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double old, double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned long long o = (unsigned long long)__double_as_longlong(old);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, ROW_MASK, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), CTRL,
                                                              ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The value of lane 63, in every lane.
__device__ __forceinline__ double lane63(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over each aligned group of 8 lanes, in every lane of the group: lane
// pairs, then quads, then the two quads (every lane adds the same two
// numbers at each stage, so all get the same sum).
__device__ __forceinline__ double lane_tree8(double x)
{
    x += dpp_f64<DPP_QUAD_XOR1, 0xf>(0.0, x);
    x += dpp_f64<DPP_QUAD_XOR2, 0xf>(0.0, x);
    x += dpp_f64<DPP_ROW_HALF_MIRROR, 0xf>(0.0, x);
    return x;
}

__global__ __launch_bounds__(256) void slice_reduce_kernel(const double *__restrict__ part, int nslice, int nc,
                                                           int f, int ipad, int ni, double *__restrict__ fout,
                                                           double *__restrict__ pout)
{
    const int t = blockIdx.x * (blockDim.x / RL) + threadIdx.x / RL;  // one (c, i) per RL lanes
    const int l = threadIdx.x % RL;
    const bool live = t < ni * nc;
    const int c = live ? t / ni : 0, i = live ? t - c * ni : 0;
    double s = 0.0;
    if (live)
        for (int sl = l; sl < nslice; sl += RL) s += part[((size_t)sl * nc + c) * ipad + i];
    s = lane_tree8(s);
    if (live && l == 0) {
        if (c < f) fout[(size_t)i * f + c] = s;
        else pout[i] = s;
    }
}

// ---- j-parallel kernel for short active lists ----------------------------
// For nact much smaller than a workgroup the i-parallel kernel leaves most
// lanes idle and serialises a whole j-tile per wave.  Here lanes take j
// (coalesced SoA loads straight from HBM, no LDS staging) and every lane
// holds a group of GI i-particles (uniform across the workgroup, read with
// scalar-friendly broadcast loads).  Per-lane sums are reduced across the
// wave with DPP moves (wave_sum) and across the 4 waves through LDS; the j-slices
// are added by slice_reduce_kernel (deterministic).  FP64 and SPLIT bodies only: the EXACT variant must sum j
// in index order and uses tile_kernel with one slice.
struct JParArgs {
    const double *__restrict__ src;
    int npad;
    int nact;
    int ngroups;     // ceil(nact / GI)
    int jchunk;      // j per workgroup, a multiple of NT
    int nslice;
    double eps2;
    double *__restrict__ part;   // part[(slice*NC + c)*ipad + i], ipad = ngroups*GI
    double *__restrict__ fout;
    double *__restrict__ pout;
};

// Sum over the 64 lanes of a wave, in every lane: groups of 8 (lane_tree8),
// rows of 16 (mirror), rows 0 + 1 and 2 + 3 (lane 15 of a row broadcast into
// rows 1 and 3), then the two halves (lane 31 into rows 2 and 3): the total
// (r3 + r2) + (r1 + r0) lands in lane 63 and is broadcast from there.  The
// order of the additions is fixed.
__device__ __forceinline__ double wave_sum(double x)
{
    x = lane_tree8(x);
    x += dpp_f64<DPP_ROW_MIRROR, 0xf>(0.0, x);
    x += dpp_f64<DPP_ROW_BCAST15, 0xa>(0.0, x);
    x += dpp_f64<DPP_ROW_BCAST31, 0xc>(0.0, x);
    return lane63(x);
}

template <int ORDER, int MODE, int GI, bool POT>
__global__ __launch_bounds__(NT) void jpar_kernel(JParArgs a)
{
    static_assert(MODE != EXACT, "EXACT sums j in index order: use tile_kernel");
    constexpr int Q = pr_words(ORDER);
    constexpr int NC = n_acc(ORDER, POT);
    constexpr int V = GI * NC;
    constexpr int F = f_words(ORDER);
    constexpr int NW = NT / 64;
    static_assert(V <= NT, "too many values per group");
    __shared__ double red[NW][V];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int group = blockIdx.x, slice = blockIdx.y;
    const int npad = a.npad;

    double pi[GI][Q];
    int ii[GI];
#pragma unroll
    for (int g = 0; g < GI; g++) {
        ii[g] = group * GI + g;
        const int idx = ii[g] < a.nact ? ii[g] : a.nact - 1;
#pragma unroll
        for (int w = 0; w < Q; w++) pi[g][w] = a.src[(size_t)w * npad + idx];
        if constexpr (MODE == SPLIT) {
            double t[split_words(ORDER)];
            to_split<ORDER>(pi[g], t);
#pragma unroll
            for (int w = 0; w < split_words(ORDER); w++) pi[g][w] = t[w];
        }
    }
    double A[GI][NC];
#pragma unroll
    for (int g = 0; g < GI; g++)
#pragma unroll
        for (int c = 0; c < NC; c++) A[g][c] = 0.0;

    const int j0 = slice * a.jchunk;
    const int j1 = min(j0 + a.jchunk, npad);
    float Af[GI][n_f32(ORDER)];
#pragma unroll
    for (int g = 0; g < GI; g++)
#pragma unroll
        for (int c = 0; c < n_f32(ORDER); c++) Af[g][c] = 0.0f;
    int nfold = 0;
    for (int j = j0 + tid; j < j1; j += NT) {
        double q[Q];
#pragma unroll
        for (int w = 0; w < Q; w++) q[w] = a.src[(size_t)w * npad + j];
        if constexpr (MODE == SPLIT) {
            double t[split_words(ORDER)];
            to_split<ORDER>(q, t);
#pragma unroll
            for (int w = 0; w < split_words(ORDER); w++) q[w] = t[w];
        }
#pragma unroll
        for (int g = 0; g < GI; g++) {
            if constexpr (MODE == FP64) body_fp64<ORDER, POT>(pi[g], q, a.eps2, j == ii[g], 0, A[g]);
            else body_split<ORDER, POT>(pi[g], q, a.eps2, j == ii[g], A[g], Af[g]);
        }
        if constexpr (MODE == SPLIT) {
            if (++nfold == JFOLD) {
                nfold = 0;
#pragma unroll
                for (int g = 0; g < GI; g++) split_fold<ORDER>(A[g], Af[g]);
            }
        }
    }
    if constexpr (MODE == SPLIT) {
#pragma unroll
        for (int g = 0; g < GI; g++) split_fold<ORDER>(A[g], Af[g]);
    }

    // workgroup sum in a fixed order: DPP tree per wave, then waves 0..3
#pragma unroll
    for (int g = 0; g < GI; g++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double x = wave_sum(A[g][c]);
            if (lane == 0) red[wave][g * NC + c] = x;
        }
    __syncthreads();
    double tot = 0.0;
    if (tid < V) {
#pragma unroll
        for (int w = 0; w < NW; w++) tot += red[w][tid];
    }
    const int gi = tid / NC, ci = tid - gi * NC;
    if (tid >= V) return;
    const int i = group * GI + gi;
    if (a.nslice == 1) {
        if (i < a.nact) {
            if (ci < F) a.fout[(size_t)i * F + ci] = tot;
            else a.pout[i] = tot;
        }
        return;
    }
    a.part[((size_t)slice * NC + ci) * ((size_t)a.ngroups * GI) + i] = tot;
}

}  // namespace hh
