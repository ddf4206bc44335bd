/*
 * fko.h -- CPU oracle for the GPP / FF self-energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is never linked into, loaded by or
 * called from the product library (fortkernels-_amd).  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py use it.
 *
 * Every function restates the reference's arithmetic in plain C (FP64 per
 * term, exactly as the reference rounds it; -ffp-contract=off), citing the
 * reference file:line it follows.  Accumulations can run in the reference's
 * own order and precision (double, one thread) or in extended precision
 * (long double, fixed order, any thread count) so that the oracle's own
 * rounding error stays far below the 1e-12 parity bar.
 *
 * Pinning (what checks each function):
 *   fko_gpp          : closed forms of the constant-input case, golden vectors
 *                      written by the compiled reference (oracle/_ref:
 *                      BGW-Kernels/gppKer.cpp main and its flagOCC_solver /
 *                      ssxt_scht_solver on seeded inputs), brute-force
 *                      branch cases.
 *   fko_gpp_point    : golden vectors of the compiled ssxt_scht_solver.
 *   fko_gpp_gppsum   : golden vectors of the compiled gppKerSeq.cpp
 *                      flagOCC_solver / noflagOCC_solver (gppsum = 1) and
 *                      of its program's printed achtemp; the closed form of
 *                      constant inputs; brute force of the block bounds.
 *   fko_gpp_achstemp : the compiled achsDtemp_Kernel (FFKer_cori/ffKernel.cpp:
 *                      107-131), the same sum, run on the GPP inputs
 *                      (golden vectors); closed form of constant inputs.
 *   fko_ff           : closed form (constant inputs), numpy/BLAS matmul on
 *                      small shapes, golden vectors of the compiled
 *                      achsDtemp_Kernel (FFKer_cori/ffKernel.cpp:107-131).
 *   fko_ff_full      : variant 1 against the compiled C++ port's kernel
 *                      functions on seeded inputs (golden vectors); variant
 *                      0 by closed forms of frequency-linear inputs (exact
 *                      interpolation, telescoping log sums), the
 *                      occupied/unoccupied symmetries, and the parts it
 *                      shares with variant 1.
 *   fko_fill_*       : known-answer values of the counter hash (tests).
 */
#ifndef FKO_H
#define FKO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } fko_z;

/* ---- counter-based synthetic inputs (same hash as the device fill) ---- */
uint64_t fko_hash(uint64_t seed, uint64_t stream, uint64_t index);
double   fko_uniform(uint64_t seed, uint64_t stream, uint64_t index,
                     double lo, double hi);
void fko_fill_real(double *out, int64_t n, uint64_t seed, uint64_t stream,
                   double lo, double hi);
void fko_fill_complex(fko_z *out, int64_t n, uint64_t seed, uint64_t stream,
                      double re_lo, double re_hi, double im_lo, double im_hi);

/* ---- GPP ---------------------------------------------------------------- */

/* Per-point pole / SX arithmetic of ssxt_scht_solver
 * (BGW-Kernels/gppKer.cpp:19-65) for one (ig, igp, iw): returns the branch
 * taken (1 = mask, 2 = near-pole, 3 = zero) and writes ssx and sch (before
 * the matngmatmgp factor).  Omega2 = wtilde^2 * I_eps (:148-149). */
int fko_gpp_point(double wxt, fko_z wtilde, fko_z I_eps,
                  double to1, double limittwo, double limitone,
                  fko_z *ssx, fko_z *sch);

/* Whole GPP band loop, BGW-Kernels/gppKer.cpp:303-439.
 *   mask_on   : 1 = canonical (masked unoccupied sum, :406-407);
 *               0 = GPU-variant semantics (gppKer_gpuComplex.cpp:171-178:
 *                   every term added, occupied bands' SX still computed).
 *   extended  : 1 = long double accumulators, fixed order, nthreads any;
 *               0 = double accumulators in the reference's order; nthreads
 *                   is then forced to 1 so the order is the reference's
 *                   single-thread order.
 *   band_lo/hi: restrict n1 to [band_lo, band_hi) (bounded CPU samples);
 *               pass 0, nb for the whole loop.
 * Outputs are overwritten.  acht_n1_loc may be NULL.
 * Returns 0, or -1 on bad sizes / an index map the reference would read out
 * of bounds. */
int fko_gpp(int nb, int nv, int ncouls, int ngpown, int nw,
            const double *wx, const fko_z *aqsm, const fko_z *aqsn,
            const fko_z *I_eps, const fko_z *wtilde, const double *vcoul,
            const int32_t *inv_igp_index, const int32_t *indinv,
            double to1, double limittwo, double limitone, double occ,
            int mask_on, int extended, int nthreads,
            int band_lo, int band_hi,
            fko_z *achtemp, fko_z *asxtemp, fko_z *acht_n1_loc);

/* Static-CH remainder of the Fortran GPP program
 * (gppkernel.deslippe.09.30.2014.f90:192-230), summed over every band:
 *   achstemp = sum_n1 sum_my_igp 0.5 vcoul[igp] *
 *              (- sum_ig aqsn[n1][ig] I_eps[my_igp][ig] conj(aqsm[n1][igp]))
 * Long-double sums.  Returns 0 or -1. */
int fko_gpp_achstemp(int nb, int ncouls, int ngpown, const fko_z *aqsm, const fko_z *aqsn,
                     const fko_z *I_eps, const double *vcoul, const int32_t *inv_igp_index,
                     const int32_t *indinv, int nthreads, fko_z *achstemp);

/* The gppsum = 1 form of the band loop, gppKerSeq.cpp:76-261,383-472 (the
 * reference's sequential program with its fifth argument gppsum = 1): the
 * ig sum runs over ig < igp with the occupied bands' break at the first
 * |wtilde^2 I_eps| < to1, the mirrored term m2 conj(x) beside every m1 x
 * (m2 = conj(aqsm[ig]) aqsm[igp]), the unoccupied bands keep only their
 * near-pole terms and skip ig = 511 mod 512 and igp - 1 (block bounds),
 * with the diagonal ig = igp when igp is a positive multiple of 512.  Sums
 * are long double in a fixed order.  acht_n1_loc may be NULL (needs nw >= 3).
 * Returns 0, or -1 on bad sizes / an out-of-range index map. */
int fko_gpp_gppsum(int nb, int nv, int ncouls, int ngpown, int nw,
                   const double *wx, const fko_z *aqsm, const fko_z *aqsn,
                   const fko_z *I_eps, const fko_z *wtilde, const double *vcoul,
                   const int32_t *inv_igp_index, const int32_t *indinv,
                   double to1, double limittwo, double limitone, double occ,
                   int nthreads, fko_z *achtemp, fko_z *asxtemp, fko_z *acht_n1_loc);

/* Which ig enter the unoccupied gppsum = 1 sum of column igp: 1 = both
 * terms, 2 = the diagonal (m1 term only), 0 = none (gppKerSeq.cpp:209-241
 * with the 512-wide blocks of :438-450). */
int fko_gppsum_unocc_kind(int ig, int igp);

/* Resolve the column map exactly as BGW-Kernels/gppKer.cpp:322-328:
 * igp_out[my_igp] = igp, or -1 when the reference skips the column.
 * Returns -1 if an index would be read out of bounds by the reference. */
int fko_resolve_igp(int ncouls, int ngpown, const int32_t *inv_igp_index,
                    const int32_t *indinv, int32_t *igp_out);

/* ---- FF contraction ----------------------------------------------------- */

/* Intended math of ffkernel.deslippe.best.f90.back1:402-418 (zgemm + conj
 * contraction) in C order, I_eps[nfreq][ngpown][ncouls]:
 *   schDt[n1][f] = sum_my_igp conj(aqsm[n1][igp]) *
 *                  sum_ig I_eps[f][my_igp][ig] * aqsn[n1][ig]
 * and the vcoul-weighted twin used by the static CH and SX terms:
 *   schDv[n1][f] = sum_my_igp vcoul[igp] conj(aqsm[n1][igp]) * (same)
 * The contraction itself has no skip rule (:411-415); a column that
 * fko_resolve_igp marks skipped contributes zero, which is the Fortran
 * CH/SX loops' "cycle" rule (:257,:330).
 * achsDtemp (static CH, ffkernel.deslippe.best.f90:251-267; C++
 * FFKer_cori/ffKernel.cpp:107-131) uses frequency index achs_freq:
 *   achsDtemp = sum_n1 sum_my_igp 0.5 vcoul[igp] *
 *               (- sum_ig aqsn[n1][ig] conj(aqsm[n1][igp]) I_eps[achs_freq][my_igp][ig])
 * Extended-precision accumulation.  schDv, achsDtemp may be NULL. */
int fko_ff(int nb, int ncouls, int ngpown, int nfreq,
           const fko_z *aqsm, const fko_z *aqsn, const fko_z *I_eps,
           const double *vcoul, const int32_t *inv_igp_index,
           const int32_t *indinv, int achs_freq, int nthreads,
           fko_z *schDt, fko_z *schDv, fko_z *achsDtemp);

/* ---- whole FF evaluation (fko_ff_full.c) -------------------------------- */

/* Inputs of the FF program (ffkernel.deslippe.best.f90:111-206): the
 * frequency grid, its weights and broadening, band energies and the
 * evaluation-frequency axis wx = freqevalmin - ekq[n1] + iw freqevalstep
 * (iw 0-based).  schDt_in: [nb][nfreq] or NULL (then schdt_matrix is the
 * I_epsR contraction; variant 1 uses zeros, as the C++ program does). */
typedef struct {
    int nb, nv, ncouls, ngpown, nfreq, nfreqeval;
    const fko_z *aqsm, *aqsn;            /* [nb][ncouls] */
    const fko_z *I_epsR, *I_epsA;        /* [nfreq][ngpown][ncouls] */
    const double *vcoul;                 /* [ncouls] */
    const int32_t *inv_igp_index;        /* [ngpown] */
    const int32_t *indinv;               /* [ncouls+1] */
    const double *ekq;                   /* [nb] */
    const double *dFreqGrid;             /* [nfreq] */
    const double *pref;                  /* [nfreq] */
    const fko_z *dFreqBrd;               /* [nfreq] */
    double pref_zb, freqevalmin, freqevalstep, occ;
    int achs_freq;
    const fko_z *schDt_in;
} fko_ff_full_problem;

typedef struct {
    fko_z achsDtemp;
    fko_z *asxDtemp, *achDtemp, *achDtemp_cor, *achDtemp_corb, *ach2Dtemp;  /* [nfreqeval] */
    fko_z *schDt;                        /* [nb][nfreq] or NULL: the matrix used */
} fko_ff_full_out;

/* variant 0: the Fortran program (canonical); 1: the C++ port
 * FFKer_cori/ffKernel.cpp with one thread.  Returns 0 or -1. */
int fko_ff_full(const fko_ff_full_problem *p, int variant, int nthreads, fko_ff_full_out *o);

#ifdef __cplusplus
}
#endif
#endif
