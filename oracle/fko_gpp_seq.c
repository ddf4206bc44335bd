/*
 * fko_gpp_seq.c -- CPU oracle for the gppsum = 1 form of the GPP band loop
 * (TEST INFRASTRUCTURE ONLY; never linked into the product library).
 *
 * Restates gppKerSeq.cpp (the reference's sequential GPP program with the
 * gppsum argument) for gppsum = 1, term by term:
 *   main loop                  gppKerSeq.cpp:383-472
 *   occupied bands             flagOCC_solver, gppsum branch :90-139
 *   unoccupied bands           noflagOCC_solver, gppsum branch :206-242,
 *                              called per 512-wide ig block (:438-450)
 * The reference stores aqsmtemp/aqsntemp as [ncouls][nb] and I_eps_array /
 * wtilde_array as [ncouls][ngpown]; this restatement takes the ABI layouts
 * ([nb][ncouls], [ngpown][ncouls]), which hold the same values.
 *
 * What the reference computes, per band n1, column my_igp (igp from the
 * column map), frequency iw, with igmax = igp:
 *   occupied (n1 < nvband): for ig = 0, 1, ... < igmax, stopping at the
 *     first ig with |wtilde^2 I_eps| < to1 (:99, a `break`):
 *       near (|wx - wtilde| < 0.5 or |delw|^2 < to1, delw = (wx - wtilde)
 *       * (0.5 / wtilde)): sch = 0, ssx = -Omega2 / (4 wtilde^2 (1 + delw))
 *       (0 if |wtilde| <= to1); else sch = wtilde I_eps / (wx - wtilde),
 *       ssx = Omega2 / (wx^2 - wtilde^2)                           :108-119
 *       terms m1 x + m2 conj(x), m1 = aqsn[ig] conj(aqsm[igp]),
 *       m2 = conj(aqsm[ig]) aqsm[igp]; only m1 x at ig = igp - 1   :125-134
 *   unoccupied: blocks igbeg = 0, 512, ... < igmax; each sums
 *       ig in [igbeg, min(igbeg + 511, igmax) clipped to igmax - 1), i.e.
 *       every ig < igmax except ig = 511 mod 512 and ig = igmax - 1; a block
 *       that ends exactly at igmax - 1 adds the diagonal ig = igmax with the
 *       m1 term only (:209-241).  delw = wtilde / (wx - wtilde), and the
 *       term is kept where |wdiff|^2 < limittwo or |delw|^2 > limitone
 *       (the near-pole points), sch = delw I_eps; terms m1 sch + m2 conj(sch)
 *   achtemp[iw] += 0.5 scht vcoul[igp]; asxtemp[iw] += ssxt occ vcoul[igp];
 *   acht_n1_loc[n1] += 0.5 scht(iw = 2) vcoul[igp]                  :452-469
 * The occupied-band constants are the functions' own locals in the
 * reference (to1 1e-6, gamma 0.5, limitone 1/(4 to1), limittwo 0.25); here
 * to1 / limittwo / limitone come from the problem and gamma =
 * sqrt(limittwo), which are those values for the reference's problem.
 *
 * Arithmetic: std::complex<double> products expand to (ac - bd, ad + bc)
 * (compiled with -ffp-contract=off, as the reference's x86-64 build has no
 * FMA); std::complex division is libgcc's __divdc3, whose result for
 * operands in the normal range is Smith's algorithm (R. L. Smith,
 * "Algorithm 116: Complex division", CACM 5(8), 1962), restated in zdiv;
 * std::abs of a complex is hypot.  Sums over ig, bands and columns are
 * long double in a fixed order, so the oracle's own rounding stays far
 * below the 1e-12 parity bar.
 */
#include "fko.h"

#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef fko_z z;

static inline z zmk(double re, double im) { z r; r.re = re; r.im = im; return r; }
static inline z zmul(z a, z b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline z zconj(z a) { return zmk(a.re, -a.im); }
static inline z zadd(z a, z b) { return zmk(a.re + b.re, a.im + b.im); }
static inline double zabs(z a) { return hypot(a.re, a.im); }

/* a / b by Smith's algorithm, the order of libgcc's __divdc3 */
static inline z zdiv(z a, z b)
{
    double ratio, denom;
    if (fabs(b.re) < fabs(b.im)) {
        ratio = b.re / b.im;
        denom = b.re * ratio + b.im;
        return zmk((a.re * ratio + a.im) / denom, (a.im * ratio - a.re) / denom);
    }
    ratio = b.im / b.re;
    denom = b.im * ratio + b.re;
    return zmk((a.im * ratio + a.re) / denom, (a.im - a.re * ratio) / denom);
}

typedef struct { long double re, im; } zl;
static inline void zl_add(zl *acc, z v) { acc->re += v.re; acc->im += v.im; }

/* Index of the first ig < igmax where the occupied loop stops (:99), or igmax. */
static int occupied_stop(const z *wt_row, const z *ie_row, int igmax, double to1)
{
    for (int ig = 0; ig < igmax; ++ig) {
        z w2 = zmul(wt_row[ig], wt_row[ig]);
        z om = zmul(w2, ie_row[ig]);
        if (zabs(om) < to1) return ig;
    }
    return igmax;
}

/* ig belongs to the unoccupied sum of column igp (the block bounds :438-450,
 * :209); 2 = the diagonal term (m1 only), 1 = both terms, 0 = not summed. */
int fko_gppsum_unocc_kind(int ig, int igp)
{
    if (ig == igp) return (igp > 0 && igp % 512 == 0) ? 2 : 0;
    if (ig < 0 || ig >= igp - 1) return 0;
    return ig % 512 == 511 ? 0 : 1;
}

int fko_gpp_gppsum(int nb, int nv, int ncouls, int ngpown, int nw,
                   const double *wx, const z *aqsm, const z *aqsn,
                   const z *I_eps, const z *wtilde, const double *vcoul,
                   const int32_t *inv_igp_index, const int32_t *indinv,
                   double to1, double limittwo, double limitone, double occ,
                   int nthreads, z *achtemp, z *asxtemp, z *acht_n1_loc)
{
    if (nb < 0 || nv < 0 || nv > nb || ncouls <= 0 || ngpown <= 0 || nw <= 0) return -1;
    if (acht_n1_loc && nw < 3) return -1;
    int32_t *igp_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)ngpown);
    if (!igp_of) return -1;
    if (fko_resolve_igp(ncouls, ngpown, inv_igp_index, indinv, igp_of)) { free(igp_of); return -1; }
    const double gamma = sqrt(limittwo);
    const size_t ncol = (size_t)ngpown;
    zl *col_ach = (zl *)calloc(ncol * nw, sizeof(zl));
    zl *col_asx = (zl *)calloc(ncol * nw, sizeof(zl));
    zl *col_n1 = (zl *)calloc(ncol * (size_t)(nb > 0 ? nb : 1), sizeof(zl));
    if (!col_ach || !col_asx || !col_n1) {
        free(col_ach); free(col_asx); free(col_n1); free(igp_of); return -1;
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
        const int igp = igp_of[my_igp];
        if (igp < 0) continue;
        const int igmax = igp;                                           /* :420 */
        const z *wt_row = wtilde + (size_t)my_igp * ncouls;
        const z *ie_row = I_eps + (size_t)my_igp * ncouls;
        const int stop = nv > 0 ? occupied_stop(wt_row, ie_row, igmax, to1) : 0;
        for (int n1 = 0; n1 < nb; ++n1) {
            const z *am = aqsm + (size_t)n1 * ncouls;
            const z *an = aqsn + (size_t)n1 * ncouls;
            const z mygpvar1 = zconj(am[igp]);
            const z mygpvar2 = am[igp];                                  /* :89, :204 */
            for (int iw = 0; iw < nw; ++iw) {
                const double wxt = wx[iw];
                zl scht = {0, 0}, ssxt = {0, 0};
                if (n1 < nv) {
                    /* flagOCC_solver, gppsum branch :92-138 */
                    for (int ig = 0; ig < stop; ++ig) {
                        const z wt = wt_row[ig], ie = ie_row[ig];
                        const z wt2 = zmul(wt, wt);
                        const z Omega2 = zmul(wt2, ie);
                        const z m1 = zmul(an[ig], mygpvar1);
                        const z m2 = zmul(zconj(am[ig]), mygpvar2);
                        const z half = zdiv(zmk(0.5, 0.0), wt);
                        const z wdiff = zmk(wxt - wt.re, -wt.im);
                        const z delw = zmul(wdiff, half);
                        const double ad = zabs(delw);
                        const double delw2 = ad * ad;
                        z sch, ssx;
                        if (zabs(wdiff) < gamma || delw2 < to1) {
                            sch = zmk(0.0, 0.0);
                            if (zabs(wt) > to1) {
                                const z w4 = zmk(4.0 * wt2.re, 4.0 * wt2.im);
                                const z den = zmul(w4, zmk(1.0 + delw.re, delw.im));
                                ssx = zdiv(zmk(-Omega2.re, -Omega2.im), den);
                            } else {
                                ssx = zmk(0.0, 0.0);
                            }
                        } else {
                            sch = zdiv(zmul(wt, ie), wdiff);
                            ssx = zdiv(Omega2, zmk(wxt * wxt - wt2.re, -wt2.im));
                        }
                        z sa, ha;
                        if (ig != igp - 1) {
                            sa = zadd(zmul(m1, ssx), zmul(m2, zconj(ssx)));
                            ha = zadd(zmul(m1, sch), zmul(m2, zconj(sch)));
                        } else {
                            sa = zmul(m1, ssx);
                            ha = zmul(m1, sch);
                        }
                        zl_add(&ssxt, sa);
                        zl_add(&scht, ha);
                    }
                } else {
                    /* noflagOCC_solver, gppsum branch :209-241, all blocks */
                    const int hi = igp + 1 < ncouls ? igp + 1 : ncouls;
                    for (int ig = 0; ig < hi; ++ig) {
                        const int kind = fko_gppsum_unocc_kind(ig, igp);
                        if (!kind) continue;
                        const z wt = wt_row[ig];
                        const z m1 = zmul(an[ig], mygpvar1);
                        const z wdiff = zmk(wxt - wt.re, -wt.im);
                        const z delw = zdiv(wt, wdiff);
                        const double delw2 = delw.re * delw.re - delw.im * (-delw.im);
                        const double wdiffr = wdiff.re * wdiff.re - wdiff.im * (-wdiff.im);
                        const double mult = (fabs(wdiffr) < limittwo || delw2 > limitone) ? 1.0 : 0.0;
                        const z dI = zmul(delw, ie_row[ig]);
                        const z sch = zmk(dI.re * mult, dI.im * mult);
                        z ha = zmul(m1, sch);
                        if (kind == 1) ha = zadd(ha, zmul(zmul(zconj(am[ig]), mygpvar2), zconj(sch)));
                        zl_add(&scht, ha);
                    }
                }
                zl *ach = col_ach + (size_t)my_igp * nw + iw;
                ach->re += 0.5L * scht.re * vcoul[igp];
                ach->im += 0.5L * scht.im * vcoul[igp];
                if (n1 < nv) {
                    zl *asx = col_asx + (size_t)my_igp * nw + iw;
                    asx->re += ssxt.re * (long double)occ * vcoul[igp];
                    asx->im += ssxt.im * (long double)occ * vcoul[igp];
                }
                if (iw == 2) {
                    zl *c = col_n1 + (size_t)my_igp * nb + n1;
                    c->re += 0.5L * scht.re * vcoul[igp];
                    c->im += 0.5L * scht.im * vcoul[igp];
                }
            }
        }
    }
    for (int iw = 0; iw < nw; ++iw) {
        long double ar = 0, ai = 0, sr = 0, si = 0;
        for (size_t c = 0; c < ncol; ++c) {
            ar += col_ach[c * nw + iw].re; ai += col_ach[c * nw + iw].im;
            sr += col_asx[c * nw + iw].re; si += col_asx[c * nw + iw].im;
        }
        achtemp[iw] = zmk((double)ar, (double)ai);
        asxtemp[iw] = zmk((double)sr, (double)si);
    }
    if (acht_n1_loc)
        for (int n1 = 0; n1 < nb; ++n1) {
            long double r = 0, i = 0;
            for (size_t c = 0; c < ncol; ++c) { r += col_n1[c * nb + n1].re; i += col_n1[c * nb + n1].im; }
            acht_n1_loc[n1] = zmk((double)r, (double)i);
        }
    free(col_ach); free(col_asx); free(col_n1); free(igp_of);
    return 0;
}
