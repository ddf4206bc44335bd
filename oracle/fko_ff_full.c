/*
 * fko_ff_full.c -- CPU oracle for the whole FF self-energy evaluation: the
 * static CH sum, the SX term, sigma 1-4 and the delta-function terms (TEST
 * INFRASTRUCTURE ONLY; never linked into the product library).
 *
 * Variant 0 (canonical) restates the Fortran program
 * ffkernel.deslippe.best.f90, with schdt_matrix from the zgemm contraction of
 * ffkernel.deslippe.best.f90.back1:402-418 (the .f90 leaves it zero because
 * that block is commented out):
 *   static CH          :251-267   (frequency achs_freq; Fortran reads 1 = 0 here)
 *   SX (asxDtemp)      :278-386   occupied bands; I_epsR for wx >= 0, I_epsA
 *                                 for wx < 0, linear interpolation between
 *                                 the grid points around |wx|
 *   sigma 1 / 3        :488-494   all frequencies but the last
 *   sigma 2 / 4        :495-516   all frequencies but the first
 *   delta terms        :529-613
 *   accumulation       :624-627
 * The SX and delta terms are evaluated as the Fortran does, as direct sums
 * over (my_igp, ig) of interpolated I_eps values.
 *
 * Variant 1 restates the C++ port FFKer_cori/ffKernel.cpp as its main
 * (:485-500) calls it, with one thread, including its departures from the
 * Fortran:
 *   - wx = freqevalmin - ekq + freqevalstep for every iw in the SX and
 *     delta kernels (:140, :224); sigma 1 uses (iw-1) with a 0-based iw
 *     (:202);
 *   - strict interval tests and the nFreq-2 fall-backs of compute_fact
 *     (:30-78), SX always from I_epsR (:148);
 *   - sigma 1 over every frequency, written into asxDtemp (main passes
 *     asxDtemp as achDtemp, :495); achDtemp stays zero;
 *   - sigma 2 / 4 through calculate_schDt_lin3 (:10-27): freqevalmin and
 *     freqevalstep truncated to int, no abs in the sigma-2 ratio,
 *     (schDt_left + schDt_lin2) * (x * intfact) for sigma 4;
 *   - sch2Di / schDi_cor / schDi_corb reset inside the loops (:191, :201,
 *     :222-223), so ach2Dtemp receives nb copies of the last sigma-2 value and
 *     achDtemp_cor only the delta terms;
 *   - schDt_matrix is never computed by the C++ program (zeros, :422);
 *     the caller passes the matrix to use (NULL = zeros).
 * Variant 1 exists to pin the arithmetic the two share against the compiled
 * C++ functions (oracle/ref_ff_harness.cpp); the product follows variant 0.
 *
 * Per-term arithmetic is FP64 (-ffp-contract=off); every sum over bands,
 * columns, G-vectors or frequencies is accumulated in long double.
 */
#include "fko.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef fko_z z;
typedef struct { long double re, im; } zl;

static inline z zmk(double re, double im) { z r; r.re = re; r.im = im; return r; }
static inline z zadd(z a, z b) { return zmk(a.re + b.re, a.im + b.im); }
static inline z zsub(z a, z b) { return zmk(a.re - b.re, a.im - b.im); }
static inline z zmul(z a, z b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline z zconj(z a) { return zmk(a.re, -a.im); }
static inline z zscale(z a, double s) { return zmk(a.re * s, a.im * s); }
/* a / b as a conj(b) / (b conj(b)) (FFKer_cori/Complex.h:208-218) */
static inline z zdiv(z a, z b)
{
    z n = zmul(a, zconj(b));
    double d = b.re * b.re + b.im * b.im;
    return zmk(n.re / d, n.im / d);
}
/* i * s * a */
static inline z zimul(double s, z a) { return zmul(zmk(0.0, s), a); }
static inline void zl_add(zl *acc, z v) { acc->re += v.re; acc->im += v.im; }
static inline void zl_sub(zl *acc, z v) { acc->re -= v.re; acc->im -= v.im; }
static inline z zl_z(zl a) { return zmk((double)a.re, (double)a.im); }

/* sum_my_igp vcoul[igp] sum_ig aqsn[n1][ig] conj(aqsm[n1][igp]) t(ig), with
 *   t = (X0 fa + X1 fb) s                      (Y0 == NULL)
 *   t = ((X0 - Y0) fa + (X1 - Y1) fb) s        (otherwise)
 * X0/X1/Y0/Y1: I_eps planes [ngpown][ncouls] of two neighbouring frequencies.
 * Skipped columns (igp < 0) contribute zero. */
static zl interp_sum(const fko_ff_full_problem *p, const int32_t *igp_of, int n1, const z *X0, const z *X1,
                     const z *Y0, const z *Y1, double fa, double fb, double s)
{
    const int C = p->ncouls;
    const z *an = p->aqsn + (size_t)n1 * C;
    zl tot = {0, 0};
    for (int my = 0; my < p->ngpown; ++my) {
        const int igp = igp_of[my];
        if (igp < 0) continue;
        const z cm = zconj(p->aqsm[(size_t)n1 * C + igp]);
        const size_t o = (size_t)my * C;
        zl col = {0, 0};
        for (int ig = 0; ig < C; ++ig) {
            z t;
            if (Y0)
                t = zadd(zscale(zsub(X0[o + ig], Y0[o + ig]), fa), zscale(zsub(X1[o + ig], Y1[o + ig]), fb));
            else
                t = zadd(zscale(X0[o + ig], fa), zscale(X1[o + ig], fb));
            if (s != 1.0) t = zscale(t, s);
            zl_add(&col, zmul(zmul(an[ig], cm), t));
        }
        tot.re += col.re * p->vcoul[igp];
        tot.im += col.im * p->vcoul[igp];
    }
    return tot;
}

/* Fortran interval search (:286-297, :537-546): the last 0-based i in
 * [0, nFreq-2] with dF[i] <= w < dF[i+1]; none -> nFreq-2 (the program's
 * nFreq+3 clamped to nFreq-1, 1-based). */
static int find_f90(const double *dF, int nF, double w)
{
    int f = -1;
    for (int i = 0; i < nF - 1; ++i)
        if (w >= dF[i] && w < dF[i + 1]) f = i;
    return f < 0 ? nF - 2 : f;
}

/* compute_fact's search (FFKer_cori/ffKernel.cpp:33-38): strict tests,
 * result 0 kept if nothing matches. */
static int find_cpp(const double *dF, int nF, double w)
{
    int f = 0;
    for (int i = 0; i < nF - 1; ++i)
        if (w > dF[i] && w < dF[i + 1]) f = i;
    return f;
}

static const z *plane(const z *I, const fko_ff_full_problem *p, int f)
{
    return I + (size_t)f * p->ngpown * p->ncouls;
}

static double clamp4(double x)
{
    if (x < 1e-4) x = 1e-4;
    if (x > 1e4) x = 1e4;
    return x;
}

/* ---- variant 0: ffkernel.deslippe.best.f90 ------------------------------ */
static void run_f90(const fko_ff_full_problem *p, const int32_t *igp_of, const z *M, fko_ff_full_out *o)
{
    const int nb = p->nb, nF = p->nfreq, ne = p->nfreqeval;
    const double *dF = p->dFreqGrid;
    zl *asx = calloc((size_t)ne, sizeof(zl)), *ach = calloc((size_t)ne, sizeof(zl));
    zl *cor = calloc((size_t)ne, sizeof(zl)), *corb = calloc((size_t)ne, sizeof(zl));
    zl *a2 = calloc((size_t)ne, sizeof(zl));
    zl *schDi = calloc((size_t)ne, sizeof(zl)), *schDi_cor = calloc((size_t)ne, sizeof(zl));
    zl *schDi_corb = calloc((size_t)ne, sizeof(zl)), *sch2Di = calloc((size_t)ne, sizeof(zl));
    for (int n1 = 0; n1 < nb; ++n1) {
        const int occ_band = n1 < p->nv;                                    /* :452 */
        /* SX, :278-386 */
        if (occ_band) {
            for (int iw = 0; iw < ne; ++iw) {
                const double wx = p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;  /* :280 */
                zl s;
                if (wx >= 0.0) {
                    const int f = find_f90(dF, nF, wx);
                    const double d = dF[f + 1] - dF[f];
                    const double fa = (dF[f + 1] - wx) / d, fb = (wx - dF[f]) / d;          /* :319-320 */
                    s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1), NULL, NULL,
                                   fa, fb, 1.0);
                } else {
                    const int f = find_f90(dF, nF, -wx);
                    const double d = dF[f + 1] - dF[f];
                    const double fa = (dF[f + 1] + wx) / d, fb = (-dF[f] - wx) / d;         /* :351-352 */
                    s = interp_sum(p, igp_of, n1, plane(p->I_epsA, p, f), plane(p->I_epsA, p, f + 1), NULL, NULL,
                                   fa, fb, 1.0);
                }
                asx[iw].re += s.re * p->occ;                                 /* :384 */
                asx[iw].im += s.im * p->occ;
            }
        }
        /* sigma 1-4, :458-517 */
        memset(schDi, 0, sizeof(zl) * ne);
        memset(schDi_cor, 0, sizeof(zl) * ne);
        memset(schDi_corb, 0, sizeof(zl) * ne);
        memset(sch2Di, 0, sizeof(zl) * ne);
        for (int f = 0; f < nF; ++f) {
            const z schDt = M[(size_t)n1 * nF + f];
            const double zb = dF[f];
            const z coh = zsub(zmk(zb, 0.0), p->dFreqBrd[f]);
            const z cor_e = occ_band ? zsub(zmk(-zb, 0.0), p->dFreqBrd[f]) : coh;
            double zr = 0, zlft = 0;
            z left = zmk(0, 0), avg = left, lin = left, lin2 = left;
            if (f != 0) {                                                   /* :477-485 */
                zr = zb;
                zlft = dF[f - 1];
                left = M[(size_t)n1 * nF + f - 1];
                avg = zscale(zadd(schDt, left), 0.5);
                lin = zsub(schDt, left);
                lin2 = zmk(lin.re / (zr - zlft), lin.im / (zr - zlft));
            }
            if (f != nF - 1) {                                              /* :488-494 */
                for (int iw = 0; iw < ne; ++iw) {
                    const double wx = p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;
                    const z num = zimul(p->pref[f], schDt);
                    zl_sub(&schDi[iw], zdiv(num, zsub(zmk(wx, 0.0), coh)));
                    zl_sub(&schDi_corb[iw], zdiv(num, zsub(zmk(wx, 0.0), cor_e)));
                }
            }
            if (f != 0) {                                                   /* :495-516 */
                for (int iw = 0; iw < ne; ++iw) {
                    const double wx = p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;
                    double it = clamp4(fabs((wx - zr) / (wx - zlft)));
                    it = -log(it);
                    zl_sub(&sch2Di[iw], zscale(zimul(p->pref_zb, avg), it));
                    z lin3;
                    if (occ_band) {
                        double it2 = clamp4(fabs((wx + zr) / (wx + zlft)));
                        it2 = log(it2);
                        const double wxp = -p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;  /* :509 */
                        lin3 = zscale(zadd(left, zscale(lin2, wxp - zlft)), it2);
                    } else {
                        lin3 = zscale(zadd(left, zscale(lin2, wx - zlft)), it);   /* :511 */
                    }
                    lin3 = zadd(lin3, lin);
                    zl_sub(&schDi_cor[iw], zimul(p->pref_zb, lin3));
                }
            }
        }
        /* delta-function terms, :529-613 */
        for (int iw = 0; iw < ne; ++iw) {
            double wx = p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;
            if (wx >= 0.0) {
                const int f = find_f90(dF, nF, wx);
                const double d = dF[f + 1] - dF[f];
                const double fa = -0.5 * (dF[f + 1] - wx) / d, fb = -0.5 * (wx - dF[f]) / d;   /* :543-544 */
                const zl s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1),
                                        plane(p->I_epsA, p, f), plane(p->I_epsA, p, f + 1), fa, fb, 1.0);
                sch2Di[iw].re += s.re;                                      /* :573 */
                sch2Di[iw].im += s.im;
                if (!occ_band) {                                            /* :574 */
                    schDi_cor[iw].re += s.re;
                    schDi_cor[iw].im += s.im;
                }
            } else if (occ_band) {
                wx = -wx;                                                   /* :576 */
                const int f = find_f90(dF, nF, wx);
                const double d = dF[f + 1] - dF[f];
                const double fa = (dF[f + 1] - wx) / d, fb = (wx - dF[f]) / d;                 /* :587-588 */
                const zl s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1),
                                        plane(p->I_epsA, p, f), plane(p->I_epsA, p, f + 1), fa, fb, -0.5);
                schDi_cor[iw].re += s.re;                                   /* :611 */
                schDi_cor[iw].im += s.im;
            }
        }
        for (int iw = 0; iw < ne; ++iw) {                                   /* :624-627 */
            ach[iw].re += schDi[iw].re;           ach[iw].im += schDi[iw].im;
            cor[iw].re += schDi_cor[iw].re;       cor[iw].im += schDi_cor[iw].im;
            corb[iw].re += schDi_corb[iw].re;     corb[iw].im += schDi_corb[iw].im;
            a2[iw].re += sch2Di[iw].re;           a2[iw].im += sch2Di[iw].im;
        }
    }
    for (int iw = 0; iw < ne; ++iw) {
        o->asxDtemp[iw] = zl_z(asx[iw]);
        o->achDtemp[iw] = zl_z(ach[iw]);
        o->achDtemp_cor[iw] = zl_z(cor[iw]);
        o->achDtemp_corb[iw] = zl_z(corb[iw]);
        o->ach2Dtemp[iw] = zl_z(a2[iw]);
    }
    free(asx); free(ach); free(cor); free(corb); free(a2);
    free(schDi); free(schDi_cor); free(schDi_corb); free(sch2Di);
}

/* ---- variant 1: FFKer_cori/ffKernel.cpp, one thread ---------------------- */
static void run_cpp(const fko_ff_full_problem *p, const int32_t *igp_of, const z *M, fko_ff_full_out *o)
{
    const int nb = p->nb, nF = p->nfreq, ne = p->nfreqeval;
    const double *dF = p->dFreqGrid;
    const int fmin_i = (int)p->freqevalmin, fstep_i = (int)p->freqevalstep;   /* int parameters, :10 */
    zl *asx = calloc((size_t)ne, sizeof(zl)), *cor = calloc((size_t)ne, sizeof(zl));
    z *sch2Di = calloc((size_t)ne, sizeof(z));
    /* asxDtemp_Kernel, :133-153 */
    for (int n1 = 0; n1 < p->nv; ++n1) {
        for (int iw = 0; iw < ne; ++iw) {
            const double wx = p->freqevalmin - p->ekq[n1] + p->freqevalstep;
            int f;
            double fa, fb;
            if (wx > 0.0) {                                                 /* compute_fact, :32-41 */
                f = find_cpp(dF, nF, wx);
                if (f == 0) f = nF - 2;
                fa = (dF[f + 1] - wx) / (dF[f + 1] - dF[f]);
                fb = (wx - dF[f]) / (dF[f + 1] - dF[f]);
            } else {                                                        /* :42-52 */
                f = find_cpp(dF, nF, -wx);
                if (f == 0) f = nF - 2;
                fa = (dF[f + 1] + wx) / (dF[f + 1] - dF[f]);
                fb = (-dF[f] - wx) / (dF[f + 1] - dF[f]);
            }
            const zl s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1), NULL, NULL,
                                    fa, fb, 1.0);
            asx[iw].re += s.re * p->occ;
            asx[iw].im += s.im * p->occ;
        }
    }
    /* achDtemp_Kernel, :155-210, writing sigma 1 into asxDtemp */
    for (int n1 = 0; n1 < nb; ++n1) {
        const int occ_band = n1 < p->nv;
        for (int f = 0; f < nF; ++f) {
            const z schDt = M[(size_t)n1 * nF + f];
            const double zb = dF[f];
            const z coh = zsub(zmk(zb, 0.0), p->dFreqBrd[f]);
            if (f != 0) {
                const double zr = zb, zlft = dF[f - 1];
                const z left = M[(size_t)n1 * nF + f - 1];
                const z avg = zscale(zadd(schDt, left), 0.5);
                const z lin = zsub(schDt, left);
                const z lin2 = zmk(lin.re / (zr - zlft), lin.im / (zr - zlft));
                for (int iw = 0; iw < ne; ++iw) {
                    /* calculate_schDt_lin3, :10-27 */
                    const double base = (double)fmin_i - p->ekq[n1] + (double)((iw - 1) * fstep_i);
                    double it = clamp4((base - zr) / (base - zlft));
                    it = -log(it);
                    sch2Di[iw] = zsub(zmk(0, 0), zscale(zmul(zmk(0.0, p->pref_zb), avg), it));
                    z lin3;
                    if (occ_band) {
                        double it2 = clamp4(fabs((base + zr) / (base + zlft)));
                        it2 = log(it2);
                        const double basep = (double)(-fmin_i) - p->ekq[n1] + (double)((iw - 1) * fstep_i);
                        lin3 = zscale(zadd(left, lin2), (basep - zlft) * it2);
                    } else {
                        lin3 = zscale(zadd(left, lin2), (base - zlft) * it);
                    }
                    (void)lin3;   /* schDi_cor is reset by achDtemp_cor_Kernel before it is read (:223) */
                }
            }
            for (int iw = 0; iw < ne; ++iw) {                               /* :199-206 */
                const double wx = p->freqevalmin - p->ekq[n1] + (double)(iw - 1) * p->freqevalstep;
                const z num = zmul(zmk(0.0, p->pref[f]), schDt);
                const z q = zdiv(num, zmk(wx - coh.re, 0.0 - coh.im));
                asx[iw].re -= q.re;
                asx[iw].im -= q.im;
            }
        }
    }
    /* achDtemp_cor_Kernel, :212-246 */
    for (int n1 = 0; n1 < nb; ++n1) {
        const int occ_band = n1 < p->nv;
        for (int iw = 0; iw < ne; ++iw) {
            const double wx = p->freqevalmin - p->ekq[n1] + p->freqevalstep;
            if (wx > 0.0) {
                if (!occ_band) {                                            /* schDttt_corKernel1, :248-275 */
                    const int f = find_cpp(dF, nF, wx);
                    const double fa = -0.5 * (dF[f + 1] - wx) / (dF[f + 1] - dF[f]);
                    const double fb = -0.5 * (wx - dF[f]) / (dF[f + 1] - dF[f]);
                    const zl s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1),
                                            plane(p->I_epsA, p, f), plane(p->I_epsA, p, f + 1), fa, fb, 1.0);
                    cor[iw].re += s.re;
                    cor[iw].im += s.im;
                }
            } else if (occ_band) {                                          /* schDttt_corKernel2, :277-300 */
                const double w = -wx;
                int f = find_cpp(dF, nF, w);
                if (f == 0) f = nF - 2;
                const double fa = (dF[f + 1] - w) / (dF[f + 1] - dF[f]);
                const double fb = (w - dF[f]) / (dF[f + 1] - dF[f]);
                const zl s = interp_sum(p, igp_of, n1, plane(p->I_epsR, p, f), plane(p->I_epsR, p, f + 1),
                                        plane(p->I_epsA, p, f), plane(p->I_epsA, p, f + 1), fa, fb, -0.5);
                cor[iw].re += s.re;
                cor[iw].im += s.im;
            }
        }
    }
    for (int iw = 0; iw < ne; ++iw) {
        o->asxDtemp[iw] = zl_z(asx[iw]);
        o->achDtemp[iw] = zmk(0, 0);
        o->achDtemp_cor[iw] = zl_z(cor[iw]);
        o->achDtemp_corb[iw] = zmk(0, 0);
        o->ach2Dtemp[iw] = zmk((double)((long double)nb * sch2Di[iw].re), (double)((long double)nb * sch2Di[iw].im));
    }
    free(asx); free(cor); free(sch2Di);
}

int fko_ff_full(const fko_ff_full_problem *p, int variant, int nthreads, fko_ff_full_out *o)
{
    if (!p || !o || p->nb < 1 || p->ncouls < 1 || p->ngpown < 1 || p->nfreq < 2 || p->nfreqeval < 1) return -1;
    if (p->nv < 0 || p->nv > p->nb || (variant != 0 && variant != 1)) return -1;
    if (p->achs_freq < 0 || p->achs_freq >= p->nfreq) return -1;
    if (!o->asxDtemp || !o->achDtemp || !o->achDtemp_cor || !o->achDtemp_corb || !o->ach2Dtemp) return -1;
    const int nb = p->nb, nF = p->nfreq;
    int32_t *igp_of = malloc(sizeof(int32_t) * (size_t)p->ngpown);
    z *M = calloc((size_t)nb * nF, sizeof(z));
    if (!igp_of || !M) { free(igp_of); free(M); return -1; }
    int rc = fko_resolve_igp(p->ncouls, p->ngpown, p->inv_igp_index, p->indinv, igp_of);
    /* static CH (:251-267) and schdt_matrix (.back1:402-418) from I_epsR */
    z achs = zmk(0, 0);
    if (rc == 0)
        rc = fko_ff(nb, p->ncouls, p->ngpown, nF, p->aqsm, p->aqsn, p->I_epsR, p->vcoul, p->inv_igp_index,
                    p->indinv, p->achs_freq, nthreads, M, NULL, &achs);
    if (rc == 0) {
        if (p->schDt_in) memcpy(M, p->schDt_in, sizeof(z) * (size_t)nb * nF);
        else if (variant == 1) memset(M, 0, sizeof(z) * (size_t)nb * nF);
        o->achsDtemp = achs;
        if (o->schDt) memcpy(o->schDt, M, sizeof(z) * (size_t)nb * nF);
        if (variant == 0) run_f90(p, igp_of, M, o);
        else run_cpp(p, igp_of, M, o);
    }
    free(igp_of);
    free(M);
    return rc;
}
