/*
 * CPU FP64 oracle for the Hermite force + predictor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the
 * smoke check in __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it.  The product (hermite_amd/) never links or calls it.
 *
 * It restates, in plain C, the reference's generic FP64 back end and the
 * per-order physics it is driven by.  Every function cites the reference
 * file:line it follows.  The expression order of every floating-point
 * statement follows the reference so that, when this file and the reference
 * are both compiled with -ffp-contract=off, results are bit-identical
 * (checked by tests/test_oracle_ref.py against oracle/_ref).
 *
 * Record layouts (all doubles, word index):
 *   Particle  : id(int64 bits) 0, mass 1, tlast 2, dt 3, pos 4-6, vel 7-9,
 *               acc 10-12, jrk 13-15, snp 16-18, crk 19-21,
 *               d4a 22-24, d5a 25-27, d6a 28-30, d7a 31-33
 *               stride 24 / 28 / 36 words for order 4 / 6 / 8, which is the
 *               reference's Particle (SRC/hermite4.h:21-31 = 192 B,
 *               SRC/hermite6.h:24-35 = 224 B, SRC/hermite8.h:31-45 = 288 B).
 *   Predictor : pos 0-2, mass 3, vel 4-6, acc 7-9, jrk 10-12
 *               stride 7 / 10 / 13 (the TITAN GPredictor,
 *               SRC/hermite6-titan.h:31-36, SRC/hermite8-titan.h:33-39).
 *   Force     : acc 0-2, jrk 3-5, snp 6-8, crk 9-11; stride 6 / 9 / 12
 *               (SRC/hermite4.h:4-14, SRC/hermite6.h:4-17, SRC/hermite8.h:8-24).
 *
 * Order code 9 is the double-double 8th-order build (EIGHTHDD,
 * SRC/compile.cpp:47-48, SRC/hermite8dd.h), positions held as hi + lo:
 *   Particle  : id 0, mass 1, tlast 2, dt 3, posH 4-6, posL 7-9, vel 10-12,
 *               acc 13-15, jrk 16-18, snp 19-21, crk 22-24, d4a 25-27,
 *               d5a 28-30, d6a 31-33, d7a 34-36, pad 37-39 (40 words = 320 B,
 *               SRC/hermite8dd.h:44-78)
 *   Predictor : posH 0-2, mass 3, posL 4-6, id 7, vel 8-10, acc 11-13,
 *               jrk 14-16 (17 words, SRC/hermite8dd.h:283-312)
 *   Force     : as order 8.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { W_ID = 0, W_MASS = 1, W_TLAST = 2, W_DT = 3, W_POS = 4, W_VEL = 7,
       W_ACC = 10, W_JRK = 13, W_SNP = 16, W_CRK = 19, W_D4A = 22, W_D5A = 25,
       W_D6A = 28, W_D7A = 31 };

/* double-double 8th order: word offsets of its Particle record */
enum { D_POSH = 4, D_POSL = 7, D_VEL = 10, D_ACC = 13, D_JRK = 16, D_SNP = 19, D_CRK = 22,
       D_D4A = 25, D_D5A = 28, D_D6A = 31, D_D7A = 34 };

int orc_pstride(int order) { return order == 4 ? 24 : order == 6 ? 28 : order == 8 ? 36 : 40; }
int orc_predstride(int order) { return order == 4 ? 7 : order == 6 ? 10 : order == 8 ? 13 : 17; }
int orc_fstride(int order) { return order == 4 ? 6 : order == 6 ? 9 : 12; }
static int vel_word(int order) { return order == 9 ? D_VEL : W_VEL; }

/* dvec3 dot product (x*v.x + y*v.y + z*v.z), SRC/vector3.h:121-123 */
static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* pow_one_nth_quant<N>: 2^floor(log2(x)/N) by exponent-bit arithmetic,
 * SRC/pownth.h:3-19.  The shift keeps the sign bit (x > 0 assumed). */
double orc_pow_one_nth_quant(int n, double x)
{
    uint64_t l;
    memcpy(&l, &x, 8);
    const uint64_t dec = 1023 % n;
    const uint64_t inc = 1023 - 1023 / n;
    l >>= 52;
    l -= dec;
    l /= (uint64_t)n;
    l += inc;
    l <<= 52;
    double d;
    memcpy(&d, &l, 8);
    return d;
}

/* InvForDt: h = dt/2 and hinv = 2/dt exactly for power-of-two dt,
 * SRC/pownth.h:21-34. */
void orc_invfordt(double dt, double *h, double *hinv)
{
    uint64_t l;
    memcpy(&l, &dt, 8);
    l -= (uint64_t)1 << 52;
    memcpy(h, &l, 8);
    l = ((uint64_t)2046 << 52) - l;
    memcpy(hinv, &l, 8);
}

/* ------------------------------------------------------------------ */
/* Predictor: Gravity::GPredictor constructors of the generic back end  */
/* SRC/hermite4.h:297-306, SRC/hermite6.h:277-289, SRC/hermite8.h:325-340 */
/* ------------------------------------------------------------------ */
static void predict_one(int order, const double *p, double tsys, double *q)
{
    const double dt = tsys - p[W_TLAST];
    const double dt2 = (1. / 2.) * dt;
    const double dt3 = (1. / 3.) * dt;
    const double dt4 = (1. / 4.) * dt;
    const double dt5 = (1. / 5.) * dt;
    const double dt6 = (1. / 6.) * dt;
    const double dt7 = (1. / 7.) * dt;
    for (int k = 0; k < 3; k++) {
        const double x = p[W_POS + k], v = p[W_VEL + k], a = p[W_ACC + k];
        const double j = p[W_JRK + k], s = p[W_SNP + k], c = p[W_CRK + k];
        if (order == 4) {
            q[0 + k] = x + dt * (v + dt2 * (a + dt3 * j));
            q[4 + k] = v + dt * (a + dt2 * (j));
        } else if (order == 6) {
            q[0 + k] = x + dt * (v + dt2 * (a + dt3 * (j + dt4 * (s + dt5 * (c)))));
            q[4 + k] = v + dt * (a + dt2 * (j + dt3 * (s + dt4 * (c))));
            q[7 + k] = a + dt * (j + dt2 * (s + dt3 * (c)));
        } else {
            const double d4 = p[W_D4A + k], d5 = p[W_D5A + k];
            q[0 + k] = x + dt * (v + dt2 * (a + dt3 * (j + dt4 * (s + dt5 * (c + dt6 * (d4 + dt7 * (d5)))))));
            q[4 + k] = v + dt * (a + dt2 * (j + dt3 * (s + dt4 * (c + dt5 * (d4 + dt6 * (d5))))));
            q[7 + k] = a + dt * (j + dt2 * (s + dt3 * (c + dt4 * (d4 + dt5 * (d5)))));
            q[10 + k] = j + dt * (s + dt2 * (c + dt3 * (d4 + dt4 * (d5))));
        }
    }
    q[3] = p[W_MASS];
}

/* GPredictor of the double-double build (SRC/hermite8dd.h:291-311): posH
 * kept, the Taylor displacement added to posL. */
static void predict_one_dd(const double *p, double tsys, double *q)
{
    const double dt = tsys - p[W_TLAST];
    const double dt2 = (1. / 2.) * dt;
    const double dt3 = (1. / 3.) * dt;
    const double dt4 = (1. / 4.) * dt;
    const double dt5 = (1. / 5.) * dt;
    const double dt6 = (1. / 6.) * dt;
    const double dt7 = (1. / 7.) * dt;
    for (int k = 0; k < 3; k++) {
        const double xl = p[D_POSL + k], v = p[D_VEL + k], a = p[D_ACC + k], j = p[D_JRK + k];
        const double sn = p[D_SNP + k], c = p[D_CRK + k], d4 = p[D_D4A + k], d5 = p[D_D5A + k];
        q[0 + k] = p[D_POSH + k];
        q[4 + k] = xl + dt * (v + dt2 * (a + dt3 * (j + dt4 * (sn + dt5 * (c + dt6 * (d4 + dt7 * (d5)))))));
        q[8 + k] = v + dt * (a + dt2 * (j + dt3 * (sn + dt4 * (c + dt5 * (d4 + dt6 * (d5))))));
        q[11 + k] = a + dt * (j + dt2 * (sn + dt3 * (c + dt4 * (d4 + dt5 * (d5)))));
        q[14 + k] = j + dt * (sn + dt2 * (c + dt3 * (d4 + dt4 * (d5))));
    }
    q[3] = p[W_MASS];
    q[7] = p[W_ID];
}

void orc_predict(int order, int n, const double *ptcl, double tsys, double *pred)
{
    const int ps = orc_pstride(order), qs = orc_predstride(order);
    for (int i = 0; i < n; i++) {
        if (order == 9) predict_one_dd(ptcl + (size_t)i * ps, tsys, pred + (size_t)i * qs);
        else predict_one(order, ptcl + (size_t)i * ps, tsys, pred + (size_t)i * qs);
    }
}

/* ------------------------------------------------------------------ */
/* Force on the first nact predicted particles, j summed in index order */
/* SRC/hermite4.h:334-361, SRC/hermite6.h:317-366, SRC/hermite8.h:368-432 */
/* ------------------------------------------------------------------ */
static void force_one(int order, int nj, const double *pred, int i, double eps2, double *f)
{
    const int qs = orc_predstride(order);
    const double *pi = pred + (size_t)i * qs;
    double acc[3] = {0, 0, 0}, jrk[3] = {0, 0, 0}, snp[3] = {0, 0, 0}, crk[3] = {0, 0, 0};
    for (int j = 0; j < nj; j++) {
        const double *pj = pred + (size_t)j * qs;
        double dr[3], dv[3], da[3], dj[3];
        for (int k = 0; k < 3; k++) {
            dr[k] = pj[0 + k] - pi[0 + k];
            dv[k] = pj[4 + k] - pi[4 + k];
        }
        const double r2 = eps2 + dot3(dr, dr);
        const double drdv = dot3(dr, dv);
        const double ri2 = 1.0 / r2;
        const double mri3 = pj[3] * ri2 * sqrt(ri2);
        if (order == 4) {
            const double alpha = -3.0 * drdv * ri2;
            for (int k = 0; k < 3; k++) {
                acc[k] = acc[k] + mri3 * dr[k];
                jrk[k] = jrk[k] + (mri3 * dv[k] + alpha * (mri3 * dr[k]));
            }
            continue;
        }
        for (int k = 0; k < 3; k++) da[k] = pj[7 + k] - pi[7 + k];
        const double dvdv = dot3(dv, dv);
        const double drda = dot3(dr, da);
        const double alpha = drdv * ri2;
        const double beta = (dvdv + drda) * ri2 + alpha * alpha;
        if (order == 6) {
            for (int k = 0; k < 3; k++) {
                acc[k] = acc[k] + mri3 * dr[k];
                const double tmp = dv[k] + (-3.0 * alpha) * dr[k];
                jrk[k] = jrk[k] + mri3 * tmp;
                snp[k] = snp[k] + mri3 * (da[k] + (-6.0 * alpha) * tmp + (-3.0 * beta) * dr[k]);
            }
            continue;
        }
        for (int k = 0; k < 3; k++) dj[k] = pj[10 + k] - pi[10 + k];
        const double dvda = dot3(dv, da);
        const double drdj = dot3(dr, dj);
        const double gamma = (3.0 * dvda + drdj) * ri2 + alpha * (3.0 * beta - 4.0 * alpha * alpha);
        for (int k = 0; k < 3; k++) {
            acc[k] = acc[k] + mri3 * dr[k];
            const double tmp1 = dv[k] + (-3.0 * alpha) * dr[k];
            jrk[k] = jrk[k] + mri3 * tmp1;
            const double tmp2 = da[k] + (-6.0 * alpha) * tmp1 + (-3.0 * beta) * dr[k];
            snp[k] = snp[k] + mri3 * tmp2;
            const double tmp3 = dj[k] + (-9.0 * alpha) * tmp2 + (-9.0 * beta) * tmp1 + (-3.0 * gamma) * dr[k];
            crk[k] = crk[k] + mri3 * tmp3;
        }
    }
    for (int k = 0; k < 3; k++) {
        f[0 + k] = acc[k];
        f[3 + k] = jrk[k];
        if (order >= 6) f[6 + k] = snp[k];
        if (order == 8) f[9 + k] = crk[k];
    }
}

/* Force of the double-double build (SRC/hermite8dd.h:339-389): relative
 * position (posHj - posHi) + (posLj - posLi), pairs with r2 == eps2 skipped
 * (the self-pair), otherwise the 8th-order chain of SRC/hermite8.h. */
static void force_one_dd(int nj, const double *pred, int i, double eps2, double *f)
{
    const double *pi = pred + (size_t)i * 17;
    double acc[3] = {0, 0, 0}, jrk[3] = {0, 0, 0}, snp[3] = {0, 0, 0}, crk[3] = {0, 0, 0};
    for (int j = 0; j < nj; j++) {
        const double *pj = pred + (size_t)j * 17;
        double dr[3], dv[3], da[3], dj[3];
        for (int k = 0; k < 3; k++) {
            dr[k] = (pj[0 + k] - pi[0 + k]) + (pj[4 + k] - pi[4 + k]);
            dv[k] = pj[8 + k] - pi[8 + k];
            da[k] = pj[11 + k] - pi[11 + k];
            dj[k] = pj[14 + k] - pi[14 + k];
        }
        const double r2 = eps2 + dot3(dr, dr);
        if (r2 == eps2) continue;
        const double drdv = dot3(dr, dv);
        const double dvdv = dot3(dv, dv);
        const double drda = dot3(dr, da);
        const double dvda = dot3(dv, da);
        const double drdj = dot3(dr, dj);
        const double ri2 = 1.0 / r2;
        const double mri3 = pj[3] * ri2 * sqrt(ri2);
        const double alpha = drdv * ri2;
        const double beta = (dvdv + drda) * ri2 + alpha * alpha;
        const double gamma = (3.0 * dvda + drdj) * ri2 + alpha * (3.0 * beta - 4.0 * alpha * alpha);
        for (int k = 0; k < 3; k++) {
            acc[k] = acc[k] + mri3 * dr[k];
            const double tmp1 = dv[k] + (-3.0 * alpha) * dr[k];
            jrk[k] = jrk[k] + mri3 * tmp1;
            const double tmp2 = da[k] + (-6.0 * alpha) * tmp1 + (-3.0 * beta) * dr[k];
            snp[k] = snp[k] + mri3 * tmp2;
            const double tmp3 = dj[k] + (-9.0 * alpha) * tmp2 + (-9.0 * beta) * tmp1 + (-3.0 * gamma) * dr[k];
            crk[k] = crk[k] + mri3 * tmp3;
        }
    }
    for (int k = 0; k < 3; k++) {
        f[0 + k] = acc[k];
        f[3 + k] = jrk[k];
        f[6 + k] = snp[k];
        f[9 + k] = crk[k];
    }
}

void orc_force(int order, int nj, const double *pred, int nact, double eps2, double *force)
{
    const int fs = orc_fstride(order);
    for (int i = 0; i < nact; i++) {
        if (order == 9) force_one_dd(nj, pred, i, eps2, force + (size_t)i * fs);
        else force_one(order, nj, pred, i, eps2, force + (size_t)i * fs);
    }
}

/* Same force on an arbitrary i-list (indices into pred), for tests that
 * need forces on a subset without the prefix convention. */
void orc_force_list(int order, int nj, const double *pred, int ni, const int *ilist,
                    double eps2, double *force)
{
    const int fs = orc_fstride(order);
    for (int ii = 0; ii < ni; ii++) {
        if (order == 9) force_one_dd(nj, pred, ilist[ii], eps2, force + (size_t)ii * fs);
        else force_one(order, nj, pred, ilist[ii], eps2, force + (size_t)ii * fs);
    }
}

/* Potential over corrected positions: order 4 excludes j by r2 > eps2
 * (SRC/hermite4.h:363-387), orders 6/8 by index j != i
 * (SRC/hermite6.h:368-388, SRC/hermite8.h:434-454). */
void orc_potential(int order, int n, const double *ptcl, double eps2, double *pot)
{
    const int ps = orc_pstride(order);
    for (int i = 0; i < n; i++) {
        const double *pi = ptcl + (size_t)i * ps;
        double acc = 0.0;
        for (int j = 0; j < n; j++) {
            const double *pj = ptcl + (size_t)j * ps;
            double dr[3];
            /* double-double: (posHj - posHi) + (posLj - posLi), SRC/hermite8dd.h:391-413 */
            for (int k = 0; k < 3; k++)
                dr[k] = order == 9 ? (pj[D_POSH + k] - pi[D_POSH + k]) + (pj[D_POSL + k] - pi[D_POSL + k])
                                   : pj[W_POS + k] - pi[W_POS + k];
            const double r2 = eps2 + dot3(dr, dr);
            double mj;
            if (order == 4) mj = (r2 > eps2) ? pj[W_MASS] : 0.0;
            else mj = (j != i) ? pj[W_MASS] : 0.0;
            acc -= mj / sqrt(r2);
        }
        pot[i] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* Particle physics: corrector, derivative reconstruction, timesteps    */
/* ------------------------------------------------------------------ */
static inline double norm2(const double *a) { return dot3(a, a); }

/* aarseth_step_quant: SRC/hermite4.h:78-94, SRC/hermite6.h:87-108,
 * SRC/hermite8.h:112-135.  p points to a Particle record. */
/* The DD record without posL is the 8th-order record: view and back. */
static void dd_to_8(const double *d, double *e)
{
    memcpy(e, d, 4 * 8);
    memcpy(e + W_POS, d + D_POSH, 3 * 8);
    memcpy(e + W_VEL, d + D_VEL, 27 * 8);
}
static void dd_from_8(const double *e, double *d)
{
    memcpy(d, e, 4 * 8);
    memcpy(d + D_POSH, e + W_POS, 3 * 8);
    memcpy(d + D_VEL, e + W_VEL, 27 * 8);
}

double orc_aarseth_step_quant(int order, const double *p, double etapow)
{
    if (order == 9) {
        double e[36];
        dd_to_8(p, e);
        return orc_aarseth_step_quant(8, e, etapow);
    }
    const double s0 = norm2(p + W_ACC), s1 = norm2(p + W_JRK), s2 = norm2(p + W_SNP);
    if (order == 4) {
        const double s3 = norm2(p + W_CRK);
        const double u = sqrt(s0 * s2) + s1;
        const double l = sqrt(s1 * s3) + s2;
        return orc_pow_one_nth_quant(2, etapow * (u / l));
    } else if (order == 6) {
        const double s3 = norm2(p + W_CRK), s4 = norm2(p + W_D4A), s5 = norm2(p + W_D5A);
        const double u = sqrt(s0 * s2) + s1;
        const double l = sqrt(s3 * s5) + s4;
        return orc_pow_one_nth_quant(6, etapow * (u / l));
    } else {
        const double s5 = norm2(p + W_D5A), s6 = norm2(p + W_D6A), s7 = norm2(p + W_D7A);
        const double u = sqrt(s0 * s2) + s1;
        const double l = sqrt(s5 * s7) + s6;
        return orc_pow_one_nth_quant(10, etapow * (u / l));
    }
}

/* assign_force: SRC/hermite4.h:67-70, SRC/hermite6.h:75-79, SRC/hermite8.h:91-96 */
void orc_assign_force(int order, double *p, const double *f)
{
    if (order == 9) {
        for (int k = 0; k < 3; k++) {
            p[D_ACC + k] = f[0 + k];
            p[D_JRK + k] = f[3 + k];
            p[D_SNP + k] = f[6 + k];
            p[D_CRK + k] = f[9 + k];
        }
        return;
    }
    for (int k = 0; k < 3; k++) {
        p[W_ACC + k] = f[0 + k];
        p[W_JRK + k] = f[3 + k];
        if (order >= 6) p[W_SNP + k] = f[6 + k];
        if (order == 8) p[W_CRK + k] = f[9 + k];
    }
}

/* init_dt: SRC/hermite4.h:72-76, SRC/hermite6.h:81-85, SRC/hermite8.h:98-110 */
void orc_init_dt(int order, double *p, double eta_s, double dtmax)
{
    if (order == 9) {  /* as 8th order, SRC/hermite8dd.h:134-147 */
        double e[36];
        dd_to_8(p, e);
        orc_init_dt(8, e, eta_s, dtmax);
        p[W_DT] = e[W_DT];
        return;
    }
    double dtnat;
    if (order == 4) {
        dtnat = eta_s * sqrt(norm2(p + W_ACC) / norm2(p + W_JRK));
    } else if (order == 6) {
        dtnat = eta_s * sqrt(sqrt(norm2(p + W_ACC) / norm2(p + W_SNP)));
    } else {
        const double s0 = norm2(p + W_ACC), s1 = norm2(p + W_JRK);
        const double s2 = norm2(p + W_SNP), s3 = norm2(p + W_CRK);
        const double u = sqrt(s0 * s2) + s1;
        const double l = sqrt(s1 * s3) + s2;
        dtnat = eta_s * sqrt(u / l);
    }
    double dt = 0.25 * dtmax;
    while (dt > dtnat) dt *= 0.5;
    p[W_DT] = dt;
}

/* Hermite corrector with derivative reconstruction and the quantised
 * Aarseth step: SRC/hermite4.h:118-153, SRC/hermite6.h:136-184,
 * SRC/hermite8.h:166-222. */
/* dd_accumulate (SRC/hermite8dd.h:4-24, the second form): adds dx to the
 * pair xh + xl and renormalises. */
static void dd_accumulate(double *xh, double *xl, double dx)
{
    const double yh = *xh + dx;
    const double res = dx - (yh - *xh);
    const double yl = *xl + res;
    const double zh = yh + yl;
    const double zl = yl - (zh - yh);
    *xh = zh;
    *xl = zl;
}

/* Corrector of the double-double build (SRC/hermite8dd.h:187-238): the
 * 8th-order corrector with h = 0.5 dt, hinv = 2 / dt and the position
 * increment added to posH + posL by dd_accumulate. */
static void correct_dd(double *p, const double *f, double etapow, double dtlim)
{
    const double dt = p[W_DT];
    const double h = 0.5 * dt;
    const double hinv = 2.0 / dt;
    double *vel = p + D_VEL, *acc = p + D_ACC, *jrk = p + D_JRK, *snp = p + D_SNP, *crk = p + D_CRK;
    double *d4a = p + D_D4A, *d5a = p + D_D5A, *d6a = p + D_D6A, *d7a = p + D_D7A;
    const double *facc = f, *fjrk = f + 3, *fsnp = f + 6, *fcrk = f + 9;
    double Ap[3], Am[3], Jp[3], Jm[3], Sp[3], Sm[3], Cp[3], Cm[3];
    for (int k = 0; k < 3; k++) {
        Ap[k] = facc[k] + acc[k];
        Am[k] = facc[k] - acc[k];
        Jp[k] = (fjrk[k] + jrk[k]) * h;
        Jm[k] = (fjrk[k] - jrk[k]) * h;
        Sp[k] = (fsnp[k] + snp[k]) * (h * h);
        Sm[k] = (fsnp[k] - snp[k]) * (h * h);
        Cp[k] = (fcrk[k] + crk[k]) * (h * h * h);
        Cm[k] = (fcrk[k] - crk[k]) * (h * h * h);
    }
    for (int k = 0; k < 3; k++) {
        const double vel1 = vel[k] + h * (Ap[k] + (-3. / 7.) * Jm[k] + (2. / 21) * Sp[k] - (1. / 105.) * Cm[k]);
        const double dpos = h * ((vel[k] + vel1) + h * ((-3. / 7.) * Am[k] + (2. / 21) * Jp[k] - (1. / 105.) * Sm[k]));
        dd_accumulate(p + D_POSH + k, p + D_POSL + k, dpos);
        vel[k] = vel1;
        acc[k] = facc[k];
        jrk[k] = fjrk[k];
        snp[k] = fsnp[k];
        crk[k] = fcrk[k];
    }
    p[W_TLAST] += dt;
    const double hinv2 = hinv * hinv;
    const double hinv3 = hinv2 * hinv;
    const double hinv4 = hinv2 * hinv2;
    const double hinv5 = hinv2 * hinv3;
    const double hinv6 = hinv3 * hinv3;
    const double hinv7 = hinv3 * hinv4;
    const double third = 1.0 / 3.;
    for (int k = 0; k < 3; k++) {
        d4a[k] = (hinv4 * 24. / 32.) * (-5. * Jm[k] + 5. * Sp[k] - Cm[k]);
        d5a[k] = (hinv5 * 120. / 32.) * (21. * Am[k] - 21. * Jp[k] + 8. * Sm[k] - Cp[k]);
        d6a[k] = (hinv6 * 720. / 32.) * (Jm[k] - Sp[k] + Cm[k] * third);
        d7a[k] = (hinv7 * 5040. / 32.) * (-5. * Am[k] + 5. * Jp[k] - 2. * Sm[k] + Cp[k] * third);
    }
    const double h2 = 0.5 * h, h3 = (1. / 3.) * h;
    for (int k = 0; k < 3; k++) {
        d4a[k] = d4a[k] + h * (d5a[k] + h2 * (d6a[k] + h3 * d7a[k]));
        d5a[k] = d5a[k] + h * (d6a[k] + h2 * d7a[k]);
        d6a[k] = d6a[k] + h * d7a[k];
    }
    const double dtq = orc_aarseth_step_quant(9, p, etapow);
    p[W_DT] = dtq < dtlim ? dtq : dtlim;
}

void orc_correct(int order, double *p, const double *f, double etapow, double dtlim)
{
    if (order == 9) {
        correct_dd(p, f, etapow, dtlim);
        return;
    }
    double h, hinv;
    orc_invfordt(p[W_DT], &h, &hinv);
    double *pos = p + W_POS, *vel = p + W_VEL, *acc = p + W_ACC, *jrk = p + W_JRK;
    double *snp = p + W_SNP, *crk = p + W_CRK, *d4a = p + W_D4A, *d5a = p + W_D5A;
    double *d6a = p + W_D6A, *d7a = p + W_D7A;
    const double *facc = f, *fjrk = f + 3, *fsnp = f + 6, *fcrk = f + 9;
    double Ap[3], Am[3], Jp[3], Jm[3], Sp[3], Sm[3], Cp[3], Cm[3];
    for (int k = 0; k < 3; k++) {
        Ap[k] = facc[k] + acc[k];
        Am[k] = facc[k] - acc[k];
        Jp[k] = (fjrk[k] + jrk[k]) * h;
        Jm[k] = (fjrk[k] - jrk[k]) * h;
        if (order >= 6) {
            Sp[k] = (fsnp[k] + snp[k]) * (h * h);
            Sm[k] = (fsnp[k] - snp[k]) * (h * h);
        }
        if (order == 8) {
            Cp[k] = (fcrk[k] + crk[k]) * (h * h * h);
            Cm[k] = (fcrk[k] - crk[k]) * (h * h * h);
        }
    }
    if (order == 4) {
        for (int k = 0; k < 3; k++) {
            const double vel1 = vel[k] + h * (Ap[k] - (1. / 3.) * Jm[k]);
            pos[k] = pos[k] + h * ((vel[k] + vel1) + h * (-(1. / 3.) * Am[k]));
            vel[k] = vel1;
            acc[k] = facc[k];
            jrk[k] = fjrk[k];
        }
        p[W_TLAST] += p[W_DT];
        for (int k = 0; k < 3; k++) {
            snp[k] = (0.5 * hinv * hinv) * Jm[k];
            crk[k] = (1.5 * hinv * hinv * hinv) * (Jp[k] - Am[k]);
            snp[k] = snp[k] + h * crk[k];
        }
    } else if (order == 6) {
        for (int k = 0; k < 3; k++) {
            const double vel1 = vel[k] + h * (Ap[k] - 0.4 * Jm[k] + (1. / 15) * Sp[k]);
            pos[k] = pos[k] + h * ((vel[k] + vel1) + h * (-0.4 * Am[k] + (1. / 15) * Jp[k]));
            vel[k] = vel1;
            acc[k] = facc[k];
            jrk[k] = fjrk[k];
            snp[k] = fsnp[k];
        }
        p[W_TLAST] += p[W_DT];
        const double hinv2 = hinv * hinv;
        const double hinv3 = hinv2 * hinv;
        const double hinv4 = hinv2 * hinv2;
        const double hinv5 = hinv2 * hinv3;
        for (int k = 0; k < 3; k++) {
            crk[k] = (0.75 * hinv3) * (-5. * Am[k] + 5. * Jp[k] - Sm[k]);
            d4a[k] = (1.5 * hinv4) * (-Jm[k] + Sp[k]);
            d5a[k] = (7.5 * hinv5) * (3 * Am[k] - 3 * Jp[k] + Sm[k]);
            crk[k] = crk[k] + h * (d4a[k] + 0.5 * h * d5a[k]);
            d4a[k] = d4a[k] + h * d5a[k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            const double vel1 = vel[k] + h * (Ap[k] + (-3. / 7.) * Jm[k] + (2. / 21) * Sp[k] - (1. / 105.) * Cm[k]);
            pos[k] = pos[k] + h * ((vel[k] + vel1) + h * ((-3. / 7.) * Am[k] + (2. / 21) * Jp[k] - (1. / 105.) * Sm[k]));
            vel[k] = vel1;
            acc[k] = facc[k];
            jrk[k] = fjrk[k];
            snp[k] = fsnp[k];
            crk[k] = fcrk[k];
        }
        p[W_TLAST] += p[W_DT];
        const double hinv2 = hinv * hinv;
        const double hinv3 = hinv2 * hinv;
        const double hinv4 = hinv2 * hinv2;
        const double hinv5 = hinv2 * hinv3;
        const double hinv6 = hinv3 * hinv3;
        const double hinv7 = hinv3 * hinv4;
        /* vector / scalar multiplies by the rounded reciprocal,
         * SRC/vector3.h:130-133 */
        const double third = 1.0 / 3.;
        for (int k = 0; k < 3; k++) {
            d4a[k] = (hinv4 * 24. / 32.) * (-5. * Jm[k] + 5. * Sp[k] - Cm[k]);
            d5a[k] = (hinv5 * 120. / 32.) * (21. * Am[k] - 21. * Jp[k] + 8. * Sm[k] - Cp[k]);
            d6a[k] = (hinv6 * 720. / 32.) * (Jm[k] - Sp[k] + Cm[k] * third);
            d7a[k] = (hinv7 * 5040. / 32.) * (-5. * Am[k] + 5. * Jp[k] - 2. * Sm[k] + Cp[k] * third);
        }
        const double h2 = 0.5 * h, h3 = (1. / 3.) * h;
        for (int k = 0; k < 3; k++) {
            d4a[k] = d4a[k] + h * (d5a[k] + h2 * (d6a[k] + h3 * d7a[k]));
            d5a[k] = d5a[k] + h * (d6a[k] + h2 * d7a[k]);
            d6a[k] = d6a[k] + h * d7a[k];
        }
    }
    const double dtq = orc_aarseth_step_quant(order, p, etapow);
    p[W_DT] = dtq < dtlim ? dtq : dtlim;
}

/* ------------------------------------------------------------------ */
/* Block-timestep driver: NbodySystem of SRC/nbodysystem.h (TITAN path, */
/* no FAST_OMP_SYNC).                                                   */
/* ------------------------------------------------------------------ */
typedef struct orc_sys {
    int order, n, ps, fs, qs;
    double eps2, eta, etapow, eta_s, dtmax, tsys;
    double init_energy, prev_energy;
    long num_step, num_bstep;
    double *ptcl, *force, *pred, *pot, *dtbuf, *tmp;
} orc_sys;

static void swap_rec(double *a, double *b, double *tmp, int ps)
{
    memcpy(tmp, a, ps * 8);
    memcpy(a, b, ps * 8);
    memcpy(b, tmp, ps * 8);
}

/* sort_ptcl(nact, dtlim): repeated std::partition with predicate
 * p.dt < dt, dt halving from dtlim (SRC/nbodysystem.h:486-504).  The
 * partition is the two-ended swap partition that libstdc++ uses for
 * bidirectional iterators. */
static void sys_sort_ptcl(orc_sys *s, int nact, double dtlim)
{
    int beg = 0, end = nact;
    double dt = dtlim;
    while (end != beg) {
        int first = beg, last = end;
        for (;;) {
            for (;;) {
                if (first == last) goto done;
                if (s->ptcl[(size_t)first * s->ps + W_DT] < dt) ++first;
                else break;
            }
            --last;
            for (;;) {
                if (first == last) goto done;
                if (!(s->ptcl[(size_t)last * s->ps + W_DT] < dt)) --last;
                else break;
            }
            swap_rec(s->ptcl + (size_t)first * s->ps, s->ptcl + (size_t)last * s->ps, s->tmp, s->ps);
            ++first;
        }
    done:
        end = first;
        dt *= 0.5;
    }
    for (int i = 0; i < nact; i++)
        s->dtbuf[i] = s->ptcl[(size_t)i * s->ps + W_DT] + s->ptcl[(size_t)i * s->ps + W_TLAST];
}

/* sort_ptcl_dtcache: SRC/nbodysystem.h:505-555, comparing dt bit patterns
 * and halving by subtracting one from the exponent field. */
static void sys_sort_ptcl_dtcache(orc_sys *s, int nact, double dtlim)
{
    int beg = 0, end = nact;
    int64_t dtl;
    memcpy(&dtl, &dtlim, 8);
    while (end != beg) {
        int i = beg, j = end;
        for (;;) {
            for (;;) {
                int64_t di;
                if (i == j) goto brk;
                memcpy(&di, &s->dtbuf[i], 8);
                if (di != dtl) ++i;
                else break;
            }
            --j;
            for (;;) {
                int64_t dj;
                if (i == j) goto brk;
                memcpy(&dj, &s->dtbuf[j], 8);
                if (dj == dtl) --j;
                else break;
            }
            {
                const double t = s->dtbuf[i];
                s->dtbuf[i] = s->dtbuf[j];
                s->dtbuf[j] = t;
            }
            swap_rec(s->ptcl + (size_t)i * s->ps, s->ptcl + (size_t)j * s->ps, s->tmp, s->ps);
            ++i;
        }
    brk:
        end = i;
        dtl -= (int64_t)1 << 52;
    }
    for (int i = 0; i < nact; i++)
        s->dtbuf[i] = s->ptcl[(size_t)i * s->ps + W_DT] + s->ptcl[(size_t)i * s->ps + W_TLAST];
}

/* set_run_params: etapow = eta^(2(order-3)), SRC/nbodysystem.h:460-479 */
orc_sys *orc_sys_new(int order, int n, const double *ptcl_in, double tsys, double eps2,
                     double eta, double eta_s, double dtmax)
{
    orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
    s->order = order;
    s->n = n;
    s->ps = orc_pstride(order);
    s->fs = orc_fstride(order);
    s->qs = orc_predstride(order);
    s->eps2 = eps2;
    s->eta = eta;
    s->eta_s = eta_s;
    s->dtmax = dtmax;
    s->tsys = tsys;
    s->etapow = pow(eta, (double)(2 * ((order == 9 ? 8 : order) - 3)));
    s->ptcl = (double *)malloc((size_t)n * s->ps * 8);
    s->force = (double *)calloc((size_t)n * s->fs, 8);
    s->pred = (double *)calloc((size_t)n * s->qs, 8);
    s->pot = (double *)calloc((size_t)n, 8);
    s->dtbuf = (double *)calloc((size_t)n, 8);
    s->tmp = (double *)malloc((size_t)s->ps * 8);
    memcpy(s->ptcl, ptcl_in, (size_t)n * s->ps * 8);
    return s;
}

void orc_sys_free(orc_sys *s)
{
    if (!s) return;
    free(s->ptcl); free(s->force); free(s->pred); free(s->pot); free(s->dtbuf); free(s->tmp);
    free(s);
}

/* calc_energy: SRC/nbodysystem.h:633-650 */
double orc_sys_energy(orc_sys *s)
{
    orc_potential(s->order, s->n, s->ptcl, s->eps2, s->pot);
    double ke = 0.0, pe = 0.0;
    for (int i = 0; i < s->n; i++) {
        const double *p = s->ptcl + (size_t)i * s->ps;
        ke += p[W_MASS] * norm2(p + vel_word(s->order));
        pe += p[W_MASS] * s->pot[i];
    }
    ke *= 0.5;
    pe *= 0.5;
    return ke + pe;
}

/* init_step: SRC/nbodysystem.h:565-631, including the bootstrap force
 * pass for order > 4. */
void orc_sys_init_step(orc_sys *s)
{
    const int n = s->n;
    for (int i = 0; i < n; i++) s->ptcl[(size_t)i * s->ps + W_TLAST] = s->tsys;
    s->init_energy = s->prev_energy = orc_sys_energy(s);
    if (s->order > 4) {
        orc_predict(s->order, n, s->ptcl, s->tsys, s->pred);
        orc_force(s->order, n, s->pred, n, s->eps2, s->force);
        for (int i = 0; i < n; i++)
            orc_assign_force(s->order, s->ptcl + (size_t)i * s->ps, s->force + (size_t)i * s->fs);
    }
    orc_predict(s->order, n, s->ptcl, s->tsys, s->pred);
    orc_force(s->order, n, s->pred, n, s->eps2, s->force);
    for (int i = 0; i < n; i++) {
        double *p = s->ptcl + (size_t)i * s->ps;
        orc_assign_force(s->order, p, s->force + (size_t)i * s->fs);
        orc_init_dt(s->order, p, s->eta_s, s->dtmax);
    }
    sys_sort_ptcl(s, n, s->dtmax);
    s->init_energy = orc_sys_energy(s);
}

/* calc_dtlim: largest power of two <= dtmax dividing tnext,
 * SRC/nbodysystem.h:652-661 (floor() replaces the int cast). */
double orc_calc_dtlim(double tnext, double dtmax)
{
    double dtlim = dtmax;
    double sv = tnext / dtmax;
    while (sv != floor(sv)) {
        sv *= 2.0;
        dtlim *= 0.5;
    }
    return dtlim;
}

/* count_nact: SRC/nbodysystem.h:663-690 */
static int sys_count_nact(const orc_sys *s, double tnext)
{
    int nact;
    for (nact = 0; nact < s->n; nact++)
        if (s->dtbuf[nact] != tnext) break;
    return nact;
}

/* integrate_one_block: SRC/nbodysystem.h:692-765.  Returns nact. */
int orc_sys_integrate_one_block(orc_sys *s)
{
    const double tnext = s->ptcl[W_TLAST] + s->ptcl[W_DT];
    const double dtlim = orc_calc_dtlim(tnext, s->dtmax);
    const int nact = sys_count_nact(s, tnext);
    orc_predict(s->order, s->n, s->ptcl, tnext, s->pred);
    orc_force(s->order, s->n, s->pred, nact, s->eps2, s->force);
    for (int i = 0; i < nact; i++) {
        double *p = s->ptcl + (size_t)i * s->ps;
        orc_correct(s->order, p, s->force + (size_t)i * s->fs, s->etapow, dtlim);
        s->dtbuf[i] = p[W_DT];
    }
    sys_sort_ptcl_dtcache(s, nact, dtlim);
    s->num_step += nact;
    s->num_bstep++;
    s->tsys = tnext;
    return nact;
}

/* Block step without the dt-level permutation: the active set is every
 * particle with tlast + dt == min_i(tlast_i + dt_i), taken in index order,
 * and particles stay in place, so j is always summed in the original index
 * order.  Same physics as orc_sys_integrate_one_block (corrector, quantised
 * step, dtlim); this is the schedule the GPU-resident driver implements.
 * Returns nact. */
int orc_sys_integrate_one_block_inplace(orc_sys *s, int *ilist)
{
    double tnext = s->ptcl[W_TLAST] + s->ptcl[W_DT];
    for (int i = 1; i < s->n; i++) {
        const double *p = s->ptcl + (size_t)i * s->ps;
        const double t = p[W_TLAST] + p[W_DT];
        if (t < tnext) tnext = t;
    }
    const double dtlim = orc_calc_dtlim(tnext, s->dtmax);
    int nact = 0;
    for (int i = 0; i < s->n; i++) {
        const double *p = s->ptcl + (size_t)i * s->ps;
        if (p[W_TLAST] + p[W_DT] == tnext) ilist[nact++] = i;
    }
    orc_predict(s->order, s->n, s->ptcl, tnext, s->pred);
    orc_force_list(s->order, s->n, s->pred, nact, ilist, s->eps2, s->force);
    for (int k = 0; k < nact; k++)
        orc_correct(s->order, s->ptcl + (size_t)ilist[k] * s->ps, s->force + (size_t)k * s->fs, s->etapow, dtlim);
    s->num_step += nact;
    s->num_bstep++;
    s->tsys = tnext;
    return nact;
}

void orc_sys_integrate_until_inplace(orc_sys *s, double tend)
{
    int *ilist = (int *)malloc(sizeof(int) * (size_t)s->n);
    while (s->tsys < tend) orc_sys_integrate_one_block_inplace(s, ilist);
    free(ilist);
}

/* integrate_one_dtmax: SRC/nbodysystem.h:904-914 */
void orc_sys_integrate_one_dtmax(orc_sys *s)
{
    const double tt = s->tsys + s->dtmax;
    while (s->tsys < tt) orc_sys_integrate_one_block(s);
}

double orc_sys_tsys(const orc_sys *s) { return s->tsys; }
double orc_sys_init_energy(const orc_sys *s) { return s->init_energy; }
long orc_sys_num_step(const orc_sys *s) { return s->num_step; }
long orc_sys_num_bstep(const orc_sys *s) { return s->num_bstep; }
double *orc_sys_ptcl(orc_sys *s) { return s->ptcl; }
double *orc_sys_pot(orc_sys *s) { return s->pot; }
double *orc_sys_dtbuf(orc_sys *s) { return s->dtbuf; }

/* ------------------------------------------------------------------ */
/* Long-double brute force of the interaction chain, for small N.       */
/* Independent of the expression order above: it evaluates the closed  */
/* form derivatives of m r / |r|^3 term by term.                        */
/* ------------------------------------------------------------------ */
void orc_force_ld(int order, int nj, const double *pred, int nact, double eps2, double *force)
{
    const int qs = orc_predstride(order), fs = orc_fstride(order);
    for (int i = 0; i < nact; i++) {
        const double *pi = pred + (size_t)i * qs;
        long double F[12] = {0};
        for (int j = 0; j < nj; j++) {
            const double *pj = pred + (size_t)j * qs;
            long double r[3], v[3], a[3] = {0, 0, 0}, jj[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                r[k] = (long double)pj[k] - pi[k];
                v[k] = (long double)pj[4 + k] - pi[4 + k];
                if (order >= 6) a[k] = (long double)pj[7 + k] - pi[7 + k];
                if (order == 8) jj[k] = (long double)pj[10 + k] - pi[10 + k];
            }
            const long double r2 = eps2 + r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const long double rinv2 = 1.0L / r2;
            const long double m3 = pj[3] * rinv2 * sqrtl(rinv2);
            const long double A = (r[0] * v[0] + r[1] * v[1] + r[2] * v[2]) * rinv2;
            const long double B = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] +
                                   r[0] * a[0] + r[1] * a[1] + r[2] * a[2]) * rinv2 + A * A;
            const long double G = (3 * (v[0] * a[0] + v[1] * a[1] + v[2] * a[2]) +
                                   r[0] * jj[0] + r[1] * jj[1] + r[2] * jj[2]) * rinv2 +
                                  A * (3 * B - 4 * A * A);
            for (int k = 0; k < 3; k++) {
                const long double a0 = m3 * r[k];
                const long double j0 = m3 * v[k] - 3 * A * a0;
                const long double s0 = m3 * a[k] - 6 * A * j0 - 3 * B * a0;
                const long double c0 = m3 * jj[k] - 9 * A * s0 - 9 * B * j0 - 3 * G * a0;
                F[k] += a0;
                F[3 + k] += j0;
                F[6 + k] += s0;
                F[9 + k] += c0;
            }
        }
        for (int c = 0; c < fs; c++) force[(size_t)i * fs + c] = (double)F[c];
    }
}
