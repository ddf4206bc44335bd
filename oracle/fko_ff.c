/*
 * fko_ff.c -- CPU oracle for the FF frequency contraction (TEST
 * INFRASTRUCTURE ONLY; never linked into the product library).
 *
 * The contraction is the intended math of
 * ffkernel.deslippe.best.f90.back1:402-418:
 *   tempz(my_igp,ifreq,n1) = sum_ig I_epsR(ig,my_igp,ifreq) * aqsntemp(ig,n1)
 *       (zgemm('t','n', ngpown*nfreq, number_bands, ncouls, ...), :404 --
 *        'T' is a plain transpose, no conjugation; the reference passes
 *        lda = ngpown*nfreq where the layout needs ncouls, we use the layout)
 *   schdt_matrix(ifreq,n1) += sum_my_igp tempz(my_igp,ifreq,n1) *
 *                                        CONJG(aqsmtemp(igp,n1))   (:408-417)
 * in C order: I_eps[nfreq][ngpown][ncouls], aqs*[nb][ncouls],
 * schDt[nb][nfreq].  BLAS rounding is implementation-defined, so this
 * oracle evaluates every product and sum in long double.
 *
 * achsDtemp follows FFKer_cori/ffKernel.cpp:107-131 per term (FP64 products
 * in that order), accumulated in long double; the Fortran original is
 * ffkernel.deslippe.best.f90:251-267 (frequency index 1 in Fortran = 0 here;
 * the C++ port reads index 1, :122 -- selected by achs_freq).
 */
#include "fko.h"

#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef fko_z z;

int fko_ff(int nb, int ncouls, int ngpown, int nfreq,
           const z *aqsm, const z *aqsn, const z *I_eps,
           const double *vcoul, const int32_t *inv_igp_index,
           const int32_t *indinv, int achs_freq, int nthreads,
           z *schDt, z *schDv, z *achsDtemp)
{
    if (nb < 0 || ncouls <= 0 || ngpown <= 0 || nfreq <= 0) return -1;
    if (achsDtemp && (achs_freq < 0 || achs_freq >= nfreq)) return -1;
    int32_t *igp_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)ngpown);
    if (!igp_of) return -1;
    if (fko_resolve_igp(ncouls, ngpown, inv_igp_index, indinv, igp_of)) { free(igp_of); return -1; }

    long double *achs_part = (long double *)calloc((size_t)2 * (nb > 0 ? nb : 1), sizeof(long double));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int n1 = 0; n1 < nb; ++n1) {
        const z *an = aqsn + (size_t)n1 * ncouls;
        for (int f = 0; f < nfreq; ++f) {
            long double sr = 0, si = 0, vr = 0, vi = 0;
            for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
                int igp = igp_of[my_igp];
                if (igp < 0) continue;
                const z *ie = I_eps + ((size_t)f * ngpown + my_igp) * ncouls;
                long double tr = 0, ti = 0;
                for (int ig = 0; ig < ncouls; ++ig) {
                    long double ar = ie[ig].re, ai = ie[ig].im, br = an[ig].re, bi = an[ig].im;
                    tr += ar * br - ai * bi;
                    ti += ar * bi + ai * br;
                }
                long double cr = aqsm[(size_t)n1 * ncouls + igp].re;
                long double ci = -(long double)aqsm[(size_t)n1 * ncouls + igp].im;
                long double pr = tr * cr - ti * ci, pi = tr * ci + ti * cr;
                sr += pr; si += pi;
                vr += pr * vcoul[igp]; vi += pi * vcoul[igp];
            }
            schDt[(size_t)n1 * nfreq + f].re = (double)sr;
            schDt[(size_t)n1 * nfreq + f].im = (double)si;
            if (schDv) {
                schDv[(size_t)n1 * nfreq + f].re = (double)vr;
                schDv[(size_t)n1 * nfreq + f].im = (double)vi;
            }
        }
        if (achsDtemp) {
            /* FFKer_cori/ffKernel.cpp:113-126 per band */
            long double ar = 0, ai = 0;
            for (int my_igp = 0; my_igp < ngpown; ++my_igp) {
                int igp = igp_of[my_igp];
                if (igp < 0) continue;
                const z *ie = I_eps + ((size_t)achs_freq * ngpown + my_igp) * ncouls;
                z cm = aqsm[(size_t)n1 * ncouls + igp];
                double cmr = cm.re, cmi = -cm.im;
                long double sr = 0, si = 0;
                for (int ig = 0; ig < ncouls; ++ig) {
                    double pr = an[ig].re * cmr - an[ig].im * cmi;   /* aqsn * conj(aqsm) */
                    double pi = an[ig].re * cmi + an[ig].im * cmr;
                    double tr = pr * ie[ig].re - pi * ie[ig].im;     /* * I_epsR */
                    double ti = pr * ie[ig].im + pi * ie[ig].re;
                    sr -= tr; si -= ti;                              /* schsDtemp - ... */
                }
                long double s = (long double)(vcoul[igp] * 0.5);
                ar += sr * s; ai += si * s;
            }
            achs_part[2 * n1] = ar;
            achs_part[2 * n1 + 1] = ai;
        }
    }
    if (achsDtemp) {
        long double r = 0, i = 0;
        for (int n1 = 0; n1 < nb; ++n1) { r += achs_part[2 * n1]; i += achs_part[2 * n1 + 1]; }
        achsDtemp->re = (double)r;
        achsDtemp->im = (double)i;
    }
    free(achs_part);
    free(igp_of);
    return 0;
}
