/*
 * fko_gpp.c -- CPU oracle for the GPP band sum (TEST INFRASTRUCTURE ONLY;
 * never linked into the product library).
 *
 * Restates BGW-Kernels/gppKer.cpp:19-65 (ssxt_scht_solver),
 * :130-159 (flagOCC_solver) and :303-439 (the n1 x my_igp x ig x iw loop)
 * term by term.  Per-term arithmetic is FP64 in the reference's evaluation
 * order (std::complex<double> products expand to (ac - bd, ad + bc); this
 * file is compiled with -ffp-contract=off so no fused multiply-add changes
 * a rounding), so a term here is bit-identical to the reference's term.
 */
#include "fko.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef fko_z z;

static inline z zmk(double re, double im) { z r; r.re = re; r.im = im; return r; }
static inline z zmul(z a, z b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline z zconj(z a) { return zmk(a.re, -a.im); }
static inline z zscale(z a, double s) { return zmk(a.re * s, a.im * s); }
/* real(a * conj(a)) as std::complex evaluates it: a.re*a.re - a.im*(-a.im) */
static inline double znorm2(z a) { return a.re * a.re - a.im * (-a.im); }

typedef struct { long double re, im; } zl;
static inline void zl_add(zl *acc, z v) { acc->re += v.re; acc->im += v.im; }

/* Column map, BGW-Kernels/gppKer.cpp:322-328.  indinv has ncouls+1 entries
 * (the reference reads indinv[ncouls] when inv_igp_index hits ncouls and then
 * overrides the value, :323-325). */
int fko_resolve_igp(int ncouls, int ngpown, const int32_t *inv_igp_index,
                    const int32_t *indinv, int32_t *igp_out)
{
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
        int indigp = inv_igp_index[my_igp];
        if (indigp < 0 || indigp > ncouls) return -1;   /* reference reads out of bounds */
        int igp = indinv[indigp];
        if (indigp == ncouls) igp = ncouls - 1;
        if (igp > ncouls || igp < 0) { igp_out[my_igp] = -1; continue; }  /* skipped, :328 */
        if (igp == ncouls) return -1;                    /* reference reads aqsm[n1][ncouls] */
        igp_out[my_igp] = igp;
    }
    return 0;
}

/* ssxt_scht_solver, BGW-Kernels/gppKer.cpp:19-65, for one point.
 * The ssxcutoff test (:60-61) compares abs(wxt) < 0.00, which is never true,
 * so it is omitted. */
int fko_gpp_point(double wxt, z wtilde, z I_eps, double to1, double limittwo,
                  double limitone, z *ssx_out, z *sch_out)
{
    /* wtilde2 = std::pow(wtilde, 2): libstdc++ evaluates (1,0) * (w*w),
     * which equals w*w bit for bit for finite w (flagOCC_solver :148). */
    z wtilde2 = zmul(wtilde, wtilde);
    z Omega2 = zmul(wtilde2, I_eps);                       /* :149 */

    z wdiff = zmk(wxt - wtilde.re, -wtilde.im);            /* :30 */
    z cden = wdiff;
    double rden = 1.0 / znorm2(cden);                      /* :33 */
    z delw = zscale(zmul(wtilde, zconj(cden)), rden);      /* :34 */
    double delwr = znorm2(delw);                           /* :35 */
    double wdiffr = znorm2(wdiff);                         /* :36 */
    z sch, ssx;
    int branch;
    if (wdiffr > limittwo && delwr < limitone) {           /* :38-45 */
        sch = zmul(delw, I_eps);
        double wxt2 = wxt * wxt;                            /* pow(wxt,2), exact-rounded */
        z cd = zmk(wxt2, 0.0);
        double rd = 1.0 / znorm2(cd);
        ssx = zscale(zmul(Omega2, zconj(cd)), rd);
        branch = 1;
    } else if (delwr > to1) {                               /* :46-53 */
        sch = zmk(0.0, 0.0);
        z d5 = zmk(delw.re + 0.5, delw.im);
        z cd = zmul(zscale(wtilde2, 4.0), d5);
        double rd = 1.0 / znorm2(cd);
        z nO = zmk(-Omega2.re, -Omega2.im);
        ssx = zmul(zscale(zmul(nO, zconj(cd)), rd), delw);
        branch = 2;
    } else {                                                /* :54-58 */
        sch = zmk(0.0, 0.0);
        ssx = zmk(0.0, 0.0);
        branch = 3;
    }
    *ssx_out = ssx;
    *sch_out = sch;
    return branch;
}

/* One (n1, my_igp) column of the band loop: returns sch_array[iw] (already
 * scaled by 0.5, :346/:411) and ssx_array[iw] (:345) for every iw.
 * Accumulators are long double when ext, else double in reference order. */
static void gpp_column(int n1, int my_igp, int igp, int occupied, int nw,
                       const double *wx, int ncouls, const z *aqsm,
                       const z *aqsn, const z *I_eps, const z *wtilde,
                       double to1, double limittwo, double limitone,
                       int mask_on, int ext, zl *sch_array, zl *ssx_array)
{
    const z *wt_row = wtilde + (size_t)my_igp * ncouls;
    const z *ie_row = I_eps + (size_t)my_igp * ncouls;
    const z *an_row = aqsn + (size_t)n1 * ncouls;
    z aqsm_igp = aqsm[(size_t)n1 * ncouls + igp];
    z mygpvar1 = zconj(aqsm_igp);

    for (int iw = 0; iw < nw; ++iw) { sch_array[iw].re = sch_array[iw].im = 0; ssx_array[iw].re = ssx_array[iw].im = 0; }

    if (occupied) {
        /* flagOCC_solver per iw, :339-347 -> :145-158.  In mask-off mode the
         * occupied bands' sch is replaced by the unmasked sum below (the
         * GPU-variant semantics, gppKer_gpuComplex.cpp:140-141,148-206). */
        for (int iw = 0; iw < nw; ++iw) {
            double wxt = wx[iw];
            zl ssxt = {0, 0}, scht = {0, 0};
            double ssxt_d_re = 0, ssxt_d_im = 0, scht_d_re = 0, scht_d_im = 0;
            for (int ig = 0; ig < ncouls; ++ig) {
                z ssx, sch;
                fko_gpp_point(wxt, wt_row[ig], ie_row[ig], to1, limittwo, limitone, &ssx, &sch);
                z matngmatmgp = zmul(an_row[ig], mygpvar1);            /* :152 */
                z ssxa = zmul(matngmatmgp, ssx);                        /* :63 */
                z scha = zmul(matngmatmgp, sch);                        /* :64 */
                if (ext) { zl_add(&ssxt, ssxa); zl_add(&scht, scha); }
                else { ssxt_d_re += ssxa.re; ssxt_d_im += ssxa.im; scht_d_re += scha.re; scht_d_im += scha.im; }
            }
            if (!ext) { ssxt.re = ssxt_d_re; ssxt.im = ssxt_d_im; scht.re = scht_d_re; scht.im = scht_d_im; }
            ssx_array[iw].re += ssxt.re;                                /* :345 */
            ssx_array[iw].im += ssxt.im;
            if (mask_on) {
                if (ext) { sch_array[iw].re += 0.5L * scht.re; sch_array[iw].im += 0.5L * scht.im; }
                else {
                    z s = zscale(zmk((double)scht.re, (double)scht.im), 0.5); /* :346 */
                    sch_array[iw].re = (double)sch_array[iw].re + s.re;
                    sch_array[iw].im = (double)sch_array[iw].im + s.im;
                }
            }
        }
        if (mask_on) return;
    }

    /* Unoccupied bands (and every band in mask-off mode): cache-blocked loop
     * :386-413, igblk = 512, blocks -> iw -> ig. */
    const int igblk = 512;
    for (int igbeg = 0; igbeg < ncouls; igbeg += igblk) {
        int igend = igbeg + igblk < ncouls ? igbeg + igblk : ncouls;
        for (int iw = 0; iw < nw; ++iw) {
            double wxt = wx[iw];
            zl scht = {0, 0};
            double scht_re = 0, scht_im = 0;
            for (int ig = igbeg; ig < igend; ++ig) {
                z wt = wt_row[ig];
                z wdiff = zmk(wxt - wt.re, -wt.im);                     /* :396 */
                z cden = wdiff;
                double rden = znorm2(cden);                              /* :398 */
                rden = 1.0 / rden;                                       /* :399 */
                z delw = zscale(zmul(wt, zconj(cden)), rden);            /* :400 */
                double delwr = znorm2(delw);                             /* :401 */
                double wdiffr = znorm2(wdiff);                           /* :402 */
                z scha = zmul(zmul(zmul(mygpvar1, an_row[ig]), delw), ie_row[ig]); /* :404 */
                if (!mask_on || (wdiffr > limittwo && delwr < limitone)) {        /* :406 */
                    if (ext) zl_add(&scht, scha);
                    else { scht_re += scha.re; scht_im += scha.im; }
                }
            }
            if (ext) { sch_array[iw].re += 0.5L * scht.re; sch_array[iw].im += 0.5L * scht.im; }
            else {
                z s = zscale(zmk(scht_re, scht_im), 0.5);                /* :411 */
                sch_array[iw].re = (double)sch_array[iw].re + s.re;
                sch_array[iw].im = (double)sch_array[iw].im + s.im;
            }
        }
    }
}

int fko_gpp(int nb, int nv, int ncouls, int ngpown, int nw,
            const double *wx, const z *aqsm, const z *aqsn,
            const z *I_eps, const z *wtilde, const double *vcoul,
            const int32_t *inv_igp_index, const int32_t *indinv,
            double to1, double limittwo, double limitone, double occ,
            int mask_on, int extended, int nthreads,
            int band_lo, int band_hi,
            z *achtemp, z *asxtemp, z *acht_n1_loc)
{
    if (nb < 0 || nv < 0 || ncouls <= 0 || ngpown <= 0 || nw <= 0) return -1;
    if (band_lo < 0 || band_hi > nb || band_lo > band_hi) return -1;
    if (acht_n1_loc && nw < 3) return -1;
    int32_t *igp_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)ngpown);
    if (!igp_of) return -1;
    if (fko_resolve_igp(ncouls, ngpown, inv_igp_index, indinv, igp_of)) { free(igp_of); return -1; }

    if (!extended) {
        /* Reference order and precision, one thread (:303-439 with
         * OMP_NUM_THREADS=1): per-thread partials are one array. */
        double ach_re[64] = {0}, ach_im[64] = {0}, asx_re[64] = {0}, asx_im[64] = {0};
        if (nw > 64) { free(igp_of); return -1; }
        double *n1loc = (double *)calloc((size_t)2 * (nb > 0 ? nb : 1), sizeof(double));
        zl sch_array[64], ssx_array[64];
        for (int n1 = band_lo; n1 < band_hi; ++n1) {
            int occupied = n1 < nv;
            for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
                int igp = igp_of[my_igp];
                if (igp < 0) continue;
                gpp_column(n1, my_igp, igp, occupied, nw, wx, ncouls, aqsm, aqsn, I_eps, wtilde,
                           to1, limittwo, limitone, mask_on, 0, sch_array, ssx_array);
                if (occupied)
                    for (int iw = 0; iw < nw; ++iw) {            /* :416-419 */
                        z s = zscale(zscale(zmk((double)ssx_array[iw].re, (double)ssx_array[iw].im), occ), vcoul[igp]);
                        asx_re[iw] += s.re; asx_im[iw] += s.im;
                    }
                for (int iw = 0; iw < nw; ++iw) {                /* :421-422 */
                    z s = zscale(zmk((double)sch_array[iw].re, (double)sch_array[iw].im), vcoul[igp]);
                    ach_re[iw] += s.re; ach_im[iw] += s.im;
                }
                if (nw >= 3) {                                   /* :424 */
                    z s = zscale(zmk((double)sch_array[2].re, (double)sch_array[2].im), vcoul[igp]);
                    n1loc[2 * n1] += s.re; n1loc[2 * n1 + 1] += s.im;
                }
            }
        }
        for (int iw = 0; iw < nw; ++iw) {
            achtemp[iw] = zmk(ach_re[iw], ach_im[iw]);
            asxtemp[iw] = zmk(asx_re[iw], asx_im[iw]);
        }
        if (acht_n1_loc)
            for (int n1 = 0; n1 < nb; ++n1) acht_n1_loc[n1] = zmk(n1loc[2 * n1], n1loc[2 * n1 + 1]);
        free(n1loc);
        free(igp_of);
        return 0;
    }

    /* Extended precision, deterministic for any thread count: each my_igp
     * column accumulates its bands in order into its own long double slots,
     * then columns are summed in my_igp order. */
    size_t ncol = (size_t)ngpown;
    zl *col_ach = (zl *)calloc(ncol * nw, sizeof(zl));
    zl *col_asx = (zl *)calloc(ncol * nw, sizeof(zl));
    zl *col_n1 = acht_n1_loc ? (zl *)calloc(ncol * (size_t)(nb > 0 ? nb : 1), sizeof(zl)) : NULL;
    if (!col_ach || !col_asx || (acht_n1_loc && !col_n1)) {
        free(col_ach); free(col_asx); free(col_n1); free(igp_of); return -1;
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
        int igp = igp_of[my_igp];
        if (igp < 0) continue;
        zl *sch_array = (zl *)malloc(sizeof(zl) * nw);
        zl *ssx_array = (zl *)malloc(sizeof(zl) * nw);
        zl *ach = col_ach + (size_t)my_igp * nw;
        zl *asx = col_asx + (size_t)my_igp * nw;
        for (int n1 = band_lo; n1 < band_hi; ++n1) {
            int occupied = n1 < nv;
            gpp_column(n1, my_igp, igp, occupied, nw, wx, ncouls, aqsm, aqsn, I_eps, wtilde,
                       to1, limittwo, limitone, mask_on, 1, sch_array, ssx_array);
            for (int iw = 0; iw < nw; ++iw) {
                if (occupied) {
                    asx[iw].re += ssx_array[iw].re * (long double)occ * vcoul[igp];
                    asx[iw].im += ssx_array[iw].im * (long double)occ * vcoul[igp];
                }
                ach[iw].re += sch_array[iw].re * vcoul[igp];
                ach[iw].im += sch_array[iw].im * vcoul[igp];
            }
            if (col_n1) {
                col_n1[(size_t)my_igp * nb + n1].re += sch_array[2].re * vcoul[igp];
                col_n1[(size_t)my_igp * nb + n1].im += sch_array[2].im * vcoul[igp];
            }
        }
        free(sch_array);
        free(ssx_array);
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
    if (acht_n1_loc) {
        for (int n1 = 0; n1 < nb; ++n1) {
            long double r = 0, i = 0;
            for (size_t c = 0; c < ncol; ++c) { r += col_n1[c * nb + n1].re; i += col_n1[c * nb + n1].im; }
            acht_n1_loc[n1] = zmk((double)r, (double)i);
        }
    }
    free(col_ach); free(col_asx); free(col_n1); free(igp_of);
    return 0;
}

/* Static-CH remainder of the Fortran GPP program,
 * gppkernel.deslippe.09.30.2014.f90:192-230:
 *   schstemp = sum_ig -aqsntemp(ig,n1) * I_eps_array(ig,my_igp) * CONJG(aqsmtemp(igp,n1))
 *   achstemp += schstemp * vcoul(igp) * 0.5            for every band n1
 * over the columns the map keeps (:201-204; the C++ map of
 * BGW-Kernels/gppKer.cpp:322-328, as everywhere here).  Terms in FP64 in the
 * Fortran's order, sums in long double.  (The C++ reduce_achstemp,
 * BGW-Kernels/gppKer.cpp:67-128, never writes its output and indexes a
 * zero-length array; it has no observable result to restate.) */
int fko_gpp_achstemp(int nb, int ncouls, int ngpown, const z *aqsm, const z *aqsn, const z *I_eps,
                     const double *vcoul, const int32_t *inv_igp_index, const int32_t *indinv, int nthreads,
                     z *achstemp)
{
    if (nb < 0 || ncouls <= 0 || ngpown <= 0 || !achstemp) return -1;
    int32_t *igp_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)ngpown);
    if (!igp_of) return -1;
    if (fko_resolve_igp(ncouls, ngpown, inv_igp_index, indinv, igp_of)) { free(igp_of); return -1; }
    zl *col = (zl *)calloc((size_t)ngpown, sizeof(zl));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
        const int igp = igp_of[my_igp];
        if (igp < 0) continue;
        const z *ie = I_eps + (size_t)my_igp * ncouls;
        zl acc = {0, 0};
        for (int n1 = 0; n1 < nb; ++n1) {
            const z g1 = zconj(aqsm[(size_t)n1 * ncouls + igp]);
            const z *an = aqsn + (size_t)n1 * ncouls;
            zl s = {0, 0};
            for (int ig = 0; ig < ncouls; ++ig) {
                const z t = zmul(zmul(an[ig], ie[ig]), g1);
                s.re -= t.re;
                s.im -= t.im;
            }
            const double v = vcoul[igp] * 0.5;
            acc.re += s.re * v;
            acc.im += s.im * v;
        }
        col[my_igp] = acc;
    }
    long double r = 0, i = 0;
    for (int my_igp = 0; my_igp < ngpown; ++my_igp) { r += col[my_igp].re; i += col[my_igp].im; }
    *achstemp = zmk((double)r, (double)i);
    free(col);
    free(igp_of);
    return 0;
}
