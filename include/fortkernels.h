/*
 * fortkernels.h -- C ABI of the MI355X (gfx950) GPP / FF self-energy kernels.
 *
 * A drop-in for the BerkeleyGW mini-app loops of rahulgayatri23/FortKernels-:
 *   gpp_kernel()  replaces the GPP band loop of BGW-Kernels/gppKer.cpp:303-439
 *                 (the canonical CPU semantics), i.e. the work of
 *                 flagOCC_solver / ssxt_scht_solver (:19-65, :130-159) and of
 *                 the GPU-style achtemp_kernel / tillNvband_kernel
 *                 (gppKer_gpuComplex.cpp:127-206) and the OpenMP-target
 *                 kernels (BGW-Kernels/gppKer_gpuComplexGCCTarget.cpp:355-410).
 *   ff_kernel()   replaces the FF frequency contraction, zgemm + conj
 *                 contraction of ffkernel.deslippe.best.f90.back1:402-418,
 *                 and the static-CH sum achsDtemp_Kernel
 *                 (FFKer_cori/ffKernel.cpp:107-131; Fortran
 *                 ffkernel.deslippe.best.f90:251-267).
 *   ff_kernel_full()  the whole FF evaluation of that Fortran program
 *                 (static CH, SX, sigma 1-4, delta terms; :207-631).
 *
 * Conventions (the reference's functions are void and accumulate into
 * caller-zeroed outputs; these entry points instead):
 *   - return FK_OK (0) or a negative fk_status; fk_last_error() gives a
 *     message for the calling thread; the library never exits;
 *   - OVERWRITE their outputs;
 *   - take plain pointers and sizes; complex values are interleaved
 *     {re, im} pairs, the memory layout of std::complex<double>, GPUComplex
 *     and HIP double2;
 *   - hold no global mutable state: concurrent calls on distinct streams
 *     are safe; calls sharing a workspace must be serialised by the caller.
 */
#ifndef FORTKERNELS_H
#define FORTKERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FK_ABI_VERSION 3

typedef struct { double re, im; } fk_cdouble;  /* complex128 */
typedef struct { float re, im; } fk_cfloat;    /* complex64 (mixed-precision storage) */

typedef enum {
    FK_OK = 0,
    FK_ERR_ARG = -1,          /* NULL pointer, bad size or option */
    FK_ERR_INDEX = -2,        /* column map the reference would read out of bounds */
    FK_ERR_UNSUPPORTED = -3,  /* option combination not available */
    FK_ERR_HIP = -4,          /* HIP runtime error (message has the HIP error string) */
    FK_ERR_NOMEM = -5,        /* device allocation failed */
    FK_ERR_WORKSPACE = -6     /* caller workspace smaller than *_workspace_size() */
} fk_status;

typedef enum {
    FK_MEM_HOST = 0,    /* bulk inputs and outputs are host memory; the call
                           copies them, runs, copies back and synchronises */
    FK_MEM_DEVICE = 1   /* bulk inputs and outputs are device memory of the
                           current HIP device; the call is asynchronous on
                           `stream` (outputs valid after the stream syncs) */
} fk_memory;

typedef enum {
    FK_PREC_FP64 = 0,   /* complex128 inputs, FP64 arithmetic throughout */
    FK_PREC_MIXED = 1   /* gpp_kernel: complex64 inputs (fk_cfloat).
                           config 0: the band contraction on int8 MFMA by an
                           exact-accumulation split (relative error ~1e-9
                           against FP64 on the same values); configs 1..4:
                           FP32 MFMA with FP32 accumulation (~1e-6).  Pole
                           and SX factors and every cross-band / cross-tile
                           sum are FP64.
                           ff_kernel: the same complex128 inputs; the
                           contraction runs in split BF16 (three slices per
                           real, six slice products on BF16 MFMA, FP32
                           accumulation over an ig chunk) and is contracted
                           and summed in FP64 (relative error ~5e-7 against
                           FP64); its configs are listed at ff_kernel. */
} fk_precision;

typedef struct {
    int32_t memory;          /* fk_memory */
    int32_t precision;       /* fk_precision */
    int32_t mask;            /* 1: canonical near-pole mask on the achtemp sum
                                (BGW-Kernels/gppKer.cpp:406-407);
                                0: unmasked, the GPU-variant semantics
                                (gppKer_gpuComplex.cpp:171-178) */
    int32_t device;          /* HIP device ordinal for FK_MEM_HOST (-1 = current) */
    void *stream;            /* hipStream_t; NULL = the null stream */
    void *workspace;         /* device scratch; NULL = allocated per call
                                (stream-ordered) */
    size_t workspace_bytes;
    int32_t config;          /* kernel tile configuration; 0 = choose by shape */
    float *kernel_ms;        /* if non-NULL: the call times its dominant kernel
                                (the fused GEMM kernel; the int8 GEMM kernel of
                                the GPP mixed path; the split-BF16 GEMM kernel
                                of the FF mixed path) with HIP events,
                                synchronises, and stores the elapsed
                                milliseconds here */
} fk_options;

/* Zero-initialised options with the canonical defaults (mask on, FP64,
 * host memory, current device, null stream). */
fk_options fk_default_options(void);

/* ------------------------------------------------------------------------
 * GPP: achtemp / asxtemp of the band loop.
 *   achtemp[iw] = sum_{n1<nb} sum_{my_igp} 0.5 vcoul[igp] sum_ig
 *                 mask * conj(aqsm[n1][igp]) aqsn[n1][ig] delw I_eps[my_igp][ig]
 *   asxtemp[iw] = sum_{n1<nv} sum_{my_igp} occ vcoul[igp] sum_ig
 *                 aqsn[n1][ig] conj(aqsm[n1][igp]) ssx(three-way branch)
 * with wdiff = wx[iw] - wtilde, delw = wtilde conj(wdiff)/|wdiff|^2,
 * mask = |wdiff|^2 > limittwo && |delw|^2 < limitone, and
 * igp = indinv[inv_igp_index[my_igp]] (indigp == ncouls -> igp = ncouls-1;
 * igp > ncouls or < 0 -> column skipped), BGW-Kernels/gppKer.cpp:322-328.
 * ------------------------------------------------------------------------ */
typedef struct {
    int32_t number_bands;        /* nb */
    int32_t nvband;              /* nv <= nb occupied bands */
    int32_t ncouls;              /* C, G-vectors */
    int32_t ngpown;              /* G, columns owned (ncouls / nodes_per_group) */
    int32_t nfreq;               /* nw frequencies, 1..16 (reference: 3) */
    const double *wx;            /* [nfreq] HOST: wx_array (already clamped) */
    const void *aqsmtemp;        /* [nb][ncouls]  fk_cdouble (FP64) | fk_cfloat (MIXED) */
    const void *aqsntemp;        /* [nb][ncouls]  same element type */
    const void *I_eps_array;     /* [ngpown][ncouls] same element type */
    const void *wtilde_array;    /* [ngpown][ncouls] same element type */
    const double *vcoul;         /* [ncouls] */
    const int32_t *inv_igp_index;/* [ngpown]   HOST */
    const int32_t *indinv;       /* [ncouls+1] HOST */
    double to1;                  /* 1e-6 */
    double limittwo;             /* 0.25 */
    double limitone;             /* 1/(4 to1) */
    double occ;                  /* 1.0 */
    int32_t semantics;           /* fk_gpp_semantics; 0 (zero-initialised) = canonical */
} fk_gpp_problem;

/* Which program's band loop gpp_kernel evaluates.
 * FK_GPP_CANONICAL: BGW-Kernels/gppKer.cpp:303-439 (the formulas above).
 * FK_GPP_GPPSUM: the sequential program gppKerSeq.cpp run with its fifth
 *   argument gppsum = 1 (:76-261, 383-472).  With igp the column's G-vector:
 *   every ig sum runs over ig <= igp and adds beside each term
 *   m1 x (m1 = aqsn[n1][ig] conj(aqsm[n1][igp])) its mirror m2 conj(x)
 *   (m2 = conj(aqsm[n1][ig]) aqsm[n1][igp]);
 *   occupied bands (flagOCC_solver :90-139): ig < igp up to the first ig
 *     with |wtilde^2 I_eps| < to1; delw = (wx - wtilde) (0.5/wtilde); where
 *     |wx - wtilde| < sqrt(limittwo) or |delw|^2 < to1: sch = 0,
 *     ssx = -Omega2 / (4 wtilde^2 (1 + delw)) (0 if |wtilde| <= to1); else
 *     sch = wtilde I_eps / (wx - wtilde), ssx = Omega2 / (wx^2 - wtilde^2);
 *     no mirror at ig = igp - 1;
 *   unoccupied bands (noflagOCC_solver :206-242 over 512-wide blocks):
 *     ig < igp - 1 except ig = 511 mod 512, plus ig = igp (m1 term only)
 *     when igp is a positive multiple of 512; sch = delw I_eps with
 *     delw = wtilde / (wx - wtilde), kept only where |wdiff|^2 < limittwo
 *     or |delw|^2 > limitone;
 *   achtemp[iw] = sum 0.5 vcoul[igp] scht, asxtemp[iw] = sum occ vcoul[igp] ssxt.
 *   FP64 only (precision FK_PREC_FP64, mask 1, config 0), no acht_n1_loc or
 *   achstemp, not in batches (FK_ERR_UNSUPPORTED otherwise). */
typedef enum {
    FK_GPP_CANONICAL = 0,
    FK_GPP_GPPSUM = 1
} fk_gpp_semantics;

/* Replaces BGW-Kernels/gppKer.cpp:303-439 (and gppKer_gpuComplex.cpp:148
 * achtemp_kernel + :127 tillNvband_kernel).  achtemp, asxtemp: [nfreq].
 * acht_n1_loc: [number_bands] or NULL -- the per-band achtemp at iw = 2
 * (:424; needs nfreq >= 3). */
int gpp_kernel(const fk_gpp_problem *problem, const fk_options *options,
               fk_cdouble *achtemp, fk_cdouble *asxtemp, fk_cdouble *acht_n1_loc);

/* Device scratch bytes gpp_kernel (and gpp_kernel_ex) needs for this
 * problem and options, for achtemp, asxtemp and achstemp.  The per-band
 * output acht_n1_loc needs more (its pole factors at iw = 2, G x ncouls
 * complex doubles, and a second contraction): see
 * gpp_kernel_ex_workspace_size. */
size_t gpp_kernel_workspace_size(const fk_gpp_problem *problem, const fk_options *options);

/* gpp_kernel with its outputs in one struct and one more output, the
 * static-CH remainder of the Fortran GPP program
 * (gppkernel.deslippe.09.30.2014.f90:192-230), summed over every band:
 *   achstemp = sum_n1 sum_my_igp 0.5 vcoul[igp] *
 *              (- sum_ig aqsn[n1][ig] I_eps[my_igp][ig] conj(aqsm[n1][igp]))
 * over the columns the map keeps.  It comes out of the GPP kernels' own
 * tiles (B = the band-summed coefficient product, achstemp = -sum B I_eps),
 * so it costs one complex product per element of the epilogue.  The C++
 * program's reduce_achstemp (BGW-Kernels/gppKer.cpp:67-128) never writes its
 * output, and its program drops the value; the Fortran's sum is the one
 * computed.  achstemp may be NULL (then this is gpp_kernel). */
typedef struct {
    fk_cdouble *achtemp;       /* [nfreq] */
    fk_cdouble *asxtemp;       /* [nfreq] */
    fk_cdouble *acht_n1_loc;   /* [number_bands] or NULL */
    fk_cdouble *achstemp;      /* one value or NULL */
} fk_gpp_outputs;

int gpp_kernel_ex(const fk_gpp_problem *problem, const fk_options *options, const fk_gpp_outputs *outputs);

/* Device scratch bytes for a gpp_kernel_ex call with these outputs (only
 * which pointers are NULL matters; outputs may be NULL for achtemp and
 * asxtemp alone).  A caller workspace of at least this size serves the whole
 * call, acht_n1_loc included; a smaller one of at least
 * gpp_kernel_workspace_size() serves the rest, and acht_n1_loc then takes a
 * stream-ordered allocation of its own for the call.  No counterpart in
 * the reference, which allocates nothing on a device. */
size_t gpp_kernel_ex_workspace_size(const fk_gpp_problem *problem, const fk_options *options,
                                    const fk_gpp_outputs *outputs);

/* gpp_kernel that allocates nothing.  No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * workspace is the library's own device array of
 * gpp_kernel_noalloc_max_workspace() bytes (64 MiB), one per GPU, held until
 * the process ends; the call calls no hipMalloc, hipFree or allocating
 * library routine.  Same arguments, precisions, configurations and results
 * as gpp_kernel, limited to: options->memory = FK_MEM_DEVICE,
 * options->workspace = NULL, options->kernel_ms = NULL, acht_n1_loc = NULL,
 * and problems whose gpp_kernel_workspace_size() fits the array (else
 * FK_ERR_WORKSPACE; 512/2/32768/20 in FP64 needs 15 MB).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
int gpp_kernel_noalloc(const fk_gpp_problem *problem, const fk_options *options,
                       fk_cdouble *achtemp, fk_cdouble *asxtemp, fk_cdouble *acht_n1_loc);
size_t gpp_kernel_noalloc_max_workspace(void);

/* ------------------------------------------------------------------------
 * Batched GPP: many independent problems in one launch.  No counterpart in
 * the reference, whose program runs one problem's band loop
 * (BGW-Kernels/gppKer.cpp:303-439) per process; the batch computes exactly
 * what gpp_kernel computes for each problem.
 *
 * gpp_batch_create lays out one device workspace for all problems, builds
 * a tile table over every problem's (my_igp, ig) tiles (largest band count
 * first), picks the tile configuration (options->config > 0, or, with 0,
 * the fastest of all configurations timed on this batch -- create then
 * runs the batch 2 x 4 times, so the inputs must be valid), and captures
 * the launches of one run into a graph: one coefficient kernel, one
 * persistent kernel whose workgroups take tiles from a shared counter, one
 * reduction.  Inputs are device memory (FK_MEM_DEVICE) and are read again
 * on every run, so new values in the same arrays need no new batch.
 * Problems: nfreq <= 16; the arrays must outlive the batch.
 *
 * options->precision = FK_PREC_MIXED takes complex64 problems.  With
 * options->config = 0 it runs gpp_kernel's mixed path, the int8-split
 * contraction, for every problem: its operand-slice, int8 GEMM and FP64
 * pole kernels each as one launch over all problems' blocks, then one
 * reduction (error below 1e-8 against FP64 on the same values; problems of
 * at most 16000 bands).  With config 1..4 it runs the persistent tile kernel
 * with FP32 MFMA in that configuration (error ~1e-6).
 *
 * gpp_batch_run launches the graph on `stream` and writes achtemp[i],
 * asxtemp[i] (device pointers, [nfreq_i] each); elapsed_ms (optional)
 * receives the graph's duration and synchronises.  Calls that share one
 * batch must be serialised by the caller.
 * gpp_kernel_batched = create + run + synchronise + destroy, with config 2
 * when options->config = 0 and the precision is FP64.
 * ------------------------------------------------------------------------ */
typedef struct fk_gpp_batch fk_gpp_batch;

int gpp_batch_create(int32_t n, const fk_gpp_problem *const *problems, const fk_options *options,
                     fk_gpp_batch **batch);
int gpp_batch_run(fk_gpp_batch *batch, fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp, void *stream,
                  float *elapsed_ms);
/* Tile configuration the batch uses (1-based, as fk_options.config; 0 for
 * the int8-split contraction). */
int gpp_batch_config(const fk_gpp_batch *batch);
size_t gpp_batch_workspace_bytes(const fk_gpp_batch *batch);
void gpp_batch_destroy(fk_gpp_batch *batch);
int gpp_kernel_batched(int32_t n, const fk_gpp_problem *const *problems, const fk_options *options,
                       fk_cdouble *const *achtemp, fk_cdouble *const *asxtemp);

/* ------------------------------------------------------------------------
 * FF contraction:
 *   schDt[n1][f] = sum_my_igp conj(aqsm[n1][igp]) sum_ig I_eps[f][my_igp][ig] aqsn[n1][ig]
 *   schDv[n1][f] = the same with each column weighted by vcoul[igp]
 *   achsDtemp    = -0.5 * sum_n1 schDv[n1][achs_freq]
 * (ffkernel.deslippe.best.f90.back1:402-418; :251-267).  Skipped columns
 * (same rule as GPP) contribute zero.
 * ------------------------------------------------------------------------ */
typedef struct {
    int32_t number_bands;        /* nb */
    int32_t ncouls;              /* C */
    int32_t ngpown;              /* G */
    int32_t nfreq;               /* nFreq */
    const void *aqsmtemp;        /* [nb][ncouls] fk_cdouble */
    const void *aqsntemp;        /* [nb][ncouls] fk_cdouble */
    const void *I_eps_array;     /* [nfreq][ngpown][ncouls] fk_cdouble (I_epsR or I_epsA) */
    const double *vcoul;         /* [ncouls] */
    const int32_t *inv_igp_index;/* [ngpown]   HOST */
    const int32_t *indinv;       /* [ncouls+1] HOST */
    int32_t achs_freq;           /* frequency of the static-CH sum: 0 as the
                                    Fortran original, 1 as the C++ port (:122) */
} fk_ff_problem;

/* Replaces the zgemm + contraction of ffkernel.deslippe.best.f90.back1:402-418
 * and achsDtemp_Kernel (FFKer_cori/ffKernel.cpp:107).  schDt: [nb][nfreq];
 * schDv: [nb][nfreq] or NULL; achsDtemp: one value or NULL.
 * options->config: FK_PREC_FP64: 0 = default, 1 = 64x64 tiles, 2 = 64x128
 * tiles, 3 = 16x64 tiles staged through LDS, the ig range split over
 * workgroups (for 16 bands or fewer); FK_PREC_MIXED (split
 * BF16): ig chunk of 0 = 1024 (default), 1 = 256, 2 = 512, 3 = 1024 (a
 * smaller chunk: less FP32 rounding, more workgroups and partials). */
int ff_kernel(const fk_ff_problem *problem, const fk_options *options,
              fk_cdouble *schDt, fk_cdouble *schDv, fk_cdouble *achsDtemp);

size_t ff_kernel_workspace_size(const fk_ff_problem *problem, const fk_options *options);

/* ------------------------------------------------------------------------
 * Batched FF: many ff_kernel problems (uneven sizes) in one graph launch.
 * No counterpart in the reference, whose program contracts one problem per
 * process (ffkernel.deslippe.best.f90.back1:402-418, FFKer_cori/ffKernel.cpp:
 * 107-131); each problem gets exactly what ff_kernel computes for it.
 * ff_batch_create lays out one device allocation for every problem's
 * workspace and results, picks the configuration (options->config > 0, or
 * with 0 the fastest of the precision's configurations timed on this batch
 * -- create then runs the batch twice per configuration, so the inputs must
 * be valid) and captures every problem's launches into one graph.  Inputs
 * are device memory (FK_MEM_DEVICE), read again at every run; the arrays
 * must outlive the batch.  ff_batch_run launches the graph on `stream`, then
 * copies problem i's results to schDt[i] ([nb_i][nfreq_i]), schDv[i] and
 * achsDtemp[i] (device pointers; schDv / achsDtemp, or entries of them, may
 * be NULL); elapsed_ms (optional) receives the graph's duration and
 * synchronises.  Calls that share one batch must be serialised.
 * ------------------------------------------------------------------------ */
typedef struct fk_ff_batch fk_ff_batch;

int ff_batch_create(int32_t n, const fk_ff_problem *const *problems, const fk_options *options,
                    fk_ff_batch **batch);
int ff_batch_run(fk_ff_batch *batch, fk_cdouble *const *schDt, fk_cdouble *const *schDv,
                 fk_cdouble *const *achsDtemp, void *stream, float *elapsed_ms);
/* Configuration the batch uses (as fk_options.config). */
int ff_batch_config(const fk_ff_batch *batch);
size_t ff_batch_workspace_bytes(const fk_ff_batch *batch);
void ff_batch_destroy(fk_ff_batch *batch);

/* ------------------------------------------------------------------------
 * Whole FF self-energy evaluation: the Fortran program
 * ffkernel.deslippe.best.f90:251-631, with schdt_matrix from the contraction
 * of ffkernel.deslippe.best.f90.back1:402-418.  For the evaluation axis
 *   wx[n1][iw] = freqevalmin - ekq[n1] + iw * freqevalstep    (iw 0-based)
 * and S = schDt (I_epsR contraction, no vcoul), VR / VA = the vcoul-weighted
 * contractions of I_epsR / I_epsA (ff_kernel's schDv):
 *   achsDtemp         -0.5 sum_n1 VR[n1][achs_freq]                  :251-267
 *   asxDtemp[iw]      occ * sum over occupied bands of VR interpolated at
 *                     wx (wx >= 0) or VA interpolated at -wx (wx < 0) :278-386
 *   achDtemp[iw]      sigma 1: sum_{f<nfreq-1} -i pref[f] S[f] /
 *                              (wx - dFreqGrid[f] + dFreqBrd[f])        :488-494
 *   achDtemp_corb[iw] sigma 3: the same with -dFreqGrid[f] for occupied bands
 *   ach2Dtemp[iw]     sigma 2 (log-weighted averages of neighbouring S,
 *                     f >= 1) + delta terms (VR - VA interpolated)     :495-613
 *   achDtemp_cor[iw]  sigma 4 + delta terms
 * summed over bands n1 (occupied: n1 < nvband).  Interval search, its
 * fall-back to the last interval, the 1e-4..1e4 clamps of the log ratios
 * and the band/frequency ranges of each term follow the Fortran.
 *
 * semantics = FK_FF_CPP_PORT reproduces instead the C++ port
 * FFKer_cori/ffKernel.cpp:10-300 as its main calls it (:485-500):
 *   asxDtemp[iw]      SX from VR only (:148), at wx' = freqevalmin - ekq[n1]
 *                     + freqevalstep for every iw (:140), with strict
 *                     interval tests and the nfreq-2 fall-back of
 *                     compute_fact (:30-78); plus sigma 1 over every
 *                     frequency at freqevalmin - ekq[n1] + (iw-1)
 *                     freqevalstep (:199-206; main passes asxDtemp as
 *                     achDtemp, :495)
 *   achDtemp[iw], achDtemp_corb[iw]   zero
 *   achDtemp_cor[iw]  the delta terms at wx' (:212-300; sigma 4 is reset
 *                     before it is read, :222-223)
 *   ach2Dtemp[iw]     nb times the sigma-2 value of the last band and
 *                     frequency (sch2Di is reset per term, :191), its ratio
 *                     taken without abs on int(freqevalmin) - ekq +
 *                     (iw-1) int(freqevalstep) (calculate_schDt_lin3, :10-27)
 * FK_FF_CPP_PORT_ZERO_SCHDT evaluates the same with schDt_matrix = 0, as the
 * C++ program leaves it (:421-422), so that sigma 1 and ach2Dtemp vanish;
 * the schDt output still holds the contraction.  achs_freq is independent
 * (1 as the C++ program).
 *
 * I_epsA is contracted only over the bands and frequencies the SX and delta
 * terms read.
 * ------------------------------------------------------------------------ */
typedef enum {
    FK_FF_FORTRAN = 0,               /* ffkernel.deslippe.best.f90 (default) */
    FK_FF_CPP_PORT = 1,              /* FFKer_cori/ffKernel.cpp arithmetic */
    FK_FF_CPP_PORT_ZERO_SCHDT = 2    /* the same, schDt_matrix = 0 */
} fk_ff_semantics;

typedef struct {
    int32_t number_bands;        /* nb */
    int32_t nvband;              /* occupied bands, 0..nb */
    int32_t ncouls;              /* C */
    int32_t ngpown;              /* G */
    int32_t nfreq;               /* nFreq >= 2 grid frequencies */
    int32_t nfreqeval;           /* evaluation frequencies >= 1 */
    const void *aqsmtemp;        /* [nb][ncouls] fk_cdouble */
    const void *aqsntemp;        /* [nb][ncouls] fk_cdouble */
    const void *I_epsR_array;    /* [nfreq][ngpown][ncouls] fk_cdouble */
    const void *I_epsA_array;    /* [nfreq][ngpown][ncouls] fk_cdouble */
    const double *vcoul;         /* [ncouls] */
    const int32_t *inv_igp_index;/* [ngpown]   HOST */
    const int32_t *indinv;       /* [ncouls+1] HOST */
    const double *ekq;           /* [nb]       HOST band energies */
    const double *dFreqGrid;     /* [nfreq]    HOST, strictly increasing */
    const double *pref;          /* [nfreq]    HOST quadrature weights */
    const fk_cdouble *dFreqBrd;  /* [nfreq]    HOST broadening */
    double pref_zb;              /* 0.5/3.14 in the program */
    double freqevalmin;          /* 0 */
    double freqevalstep;         /* 0.5 */
    double occ;                  /* 1 */
    int32_t achs_freq;           /* 0 as the Fortran */
    int32_t semantics;           /* fk_ff_semantics; 0 = the Fortran */
} fk_ff_full_problem;

typedef struct {
    fk_cdouble *achsDtemp;       /* one value, or NULL */
    fk_cdouble *asxDtemp;        /* [nfreqeval] */
    fk_cdouble *achDtemp;        /* [nfreqeval] */
    fk_cdouble *achDtemp_cor;    /* [nfreqeval] */
    fk_cdouble *achDtemp_corb;   /* [nfreqeval] */
    fk_cdouble *ach2Dtemp;       /* [nfreqeval] */
    fk_cdouble *schDt;           /* [nb][nfreq] schdt_matrix, or NULL */
} fk_ff_full_outputs;

/* Replaces the band loop of ffkernel.deslippe.best.f90:207-631 (the static
 * CH, SX and CH terms) and the C++ port's achsDtemp_Kernel, asxDtemp_Kernel,
 * achDtemp_Kernel and achDtemp_cor_Kernel (FFKer_cori/ffKernel.cpp:107-254).
 * Output pointers follow options->memory.  options->precision: FK_PREC_FP64
 * or FK_PREC_MIXED (both contractions in split BF16, as ff_kernel; the
 * evaluation is FP64 either way); options->config as ff_kernel;
 * options->kernel_ms: the whole call's device time (both contractions and
 * the evaluation). */
int ff_kernel_full(const fk_ff_full_problem *problem, const fk_options *options,
                   const fk_ff_full_outputs *outputs);

size_t ff_kernel_full_workspace_size(const fk_ff_full_problem *problem, const fk_options *options);

/* ------------------------------------------------------------------------
 * Instance files: one GPP or FF problem in one binary file, for inputs dumped
 * from a BerkeleyGW run and for lists of batched problems.  No counterpart in
 * the reference, whose programs synthesise their inputs
 * (BGW-Kernels/gppKer.cpp:274-291, FFKer_cori/ffKernel.cpp:395-450).
 *
 * Layout (little-endian): a 256-byte header -- char magic[8] = "FKINST01",
 * int32 kind (1 GPP, 2 FF), int32 complex_bytes (16, or 8 for a complex64
 * GPP problem), int32 sizes[8], double scalars[8], int64 offsets[16] (file
 * offsets of the arrays, 64-byte aligned), zero padding -- then the arrays
 * in the ABI's C layouts (a Fortran column-major dump of the same arrays is
 * already in that order):
 *   GPP sizes  number_bands, nvband, ncouls, ngpown, nfreq
 *       scalars to1, limittwo, limitone, occ
 *       arrays  wx[nfreq] f64, inv_igp_index[ngpown] i32, indinv[ncouls+1]
 *               i32, vcoul[ncouls] f64, aqsmtemp, aqsntemp [nb][ncouls],
 *               I_eps_array, wtilde_array [ngpown][ncouls]
 *   FF  sizes  number_bands, nvband, ncouls, ngpown, nfreq, nfreqeval,
 *              achs_freq
 *       scalars pref_zb, freqevalmin, freqevalstep, occ
 *       arrays  ekq[nb], dFreqGrid[nfreq], pref[nfreq] f64, dFreqBrd[nfreq]
 *               c128, inv_igp_index, indinv, vcoul, aqsmtemp, aqsntemp,
 *               I_epsR_array, I_epsA_array [nfreq][ngpown][ncouls] c128
 * fk_instance_open maps the file read-only and validates it; the problem it
 * returns points into the mapping (host memory, FK_MEM_HOST) until
 * fk_instance_close.  An instance list is a text file with one instance
 * path per line (relative paths are relative to the list; '#' starts a
 * comment line).
 * ------------------------------------------------------------------------ */
typedef enum { FK_INSTANCE_GPP = 1, FK_INSTANCE_FF = 2 } fk_instance_kind_t;
typedef struct fk_instance fk_instance;

/* complex64 != 0: the problem's complex arrays are fk_cfloat */
int fk_gpp_instance_write(const char *path, const fk_gpp_problem *problem, int32_t complex64);
int fk_ff_instance_write(const char *path, const fk_ff_full_problem *problem);
int fk_instance_open(const char *path, fk_instance **instance);
int fk_instance_kind(const fk_instance *instance);
int32_t fk_instance_complex_bytes(const fk_instance *instance);
const fk_gpp_problem *fk_instance_gpp(const fk_instance *instance);   /* NULL unless a GPP file */
const fk_ff_full_problem *fk_instance_ff(const fk_instance *instance); /* NULL unless an FF file */
void fk_instance_close(fk_instance *instance);
/* paths: count malloc'd strings, released by fk_instance_list_free */
int fk_instance_list_read(const char *path, char ***paths, int32_t *count);
void fk_instance_list_free(char **paths, int32_t count);

/* ------------------------------------------------------------------------
 * Utilities
 * ------------------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char *fk_last_error(void);

/* Library ABI version (FK_ABI_VERSION) and the gfx target it was built for. */
int fk_abi_version(void);
const char *fk_build_target(void);

/* Deterministic synthetic inputs on the device: element j of a complex
 * array takes re = U(re_lo,re_hi) at counter 2j and im = U(im_lo,im_hi) at
 * 2j+1 of a SplitMix64 hash keyed by (seed, stream_id); real arrays use
 * counter j.  Bit-identical to the host generator the tests use.
 * as_complex64 != 0 stores fk_cfloat (rounded to nearest) instead of
 * fk_cdouble.  Asynchronous on `stream`.  No counterpart in the reference,
 * which fills constants (BGW-Kernels/gppKer.cpp:275-291). */
int fk_fill_complex(void *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                    double re_lo, double re_hi, double im_lo, double im_hi,
                    int32_t as_complex64, void *stream);
int fk_fill_real(double *dst, int64_t n, uint64_t seed, uint64_t stream_id,
                 double lo, double hi, void *stream);

/* The HIP complex helpers that replace GPUComplex / CustomComplex
 * (BGW-Kernels/GPUComplex.h:138-247, FFKer_cori/Complex.h:208-311,
 * FFKer_cori/CustomComplex.h:99-214), applied elementwise on host memory
 * either by the host instantiation (on_device = 0) or by a device kernel
 * (on_device = 1).  op: see fk_complex_op.  out[i] = op(a[i], b[i], s[i]);
 * FMA/FMS read out[i] as the accumulator.  Unused operands may be NULL. */
typedef enum {
    FK_COP_MUL = 0,        /* a*b                  GPUComplex_product          */
    FK_COP_CONJ = 1,       /* conj(a)              GPUComplex_conj             */
    FK_COP_ABS2 = 2,       /* |a|^2 (re), 0 (im)   GPUComplex_target abs       */
    FK_COP_ABS = 3,        /* |a| (re), 0 (im)     GPUComplex_abs              */
    FK_COP_SCALE = 4,      /* a*s                  GPUComplex_mult(a, b)       */
    FK_COP_FMA = 5,        /* out += a*b (out is read) GPUComplex_fma(out,a,b) */
    FK_COP_FMS = 6,        /* out -= a*b (out is read) GPUComplex_fms(out,a,b) */
    FK_COP_REAL_MINUS = 7, /* s - a                doubleMinusGPUComplex       */
    FK_COP_REAL_PLUS = 8,  /* s + a                doublePlusGPUComplex        */
    FK_COP_DIV = 9,        /* a/b                  GPUComplex_divide(a, b)     */
    FK_COP_DIV_REAL = 10,  /* a/s                  GPUComplex_divide(a, double)*/
    FK_COP_SQUARE = 11,    /* a*a                  GPUComplex_square           */
    FK_COP_ADD = 12,       /* a+b                                              */
    FK_COP_SUB = 13        /* a-b                  GPUComplex_minus            */
} fk_complex_op;
int fk_complex_apply(int32_t op, int64_t n, const fk_cdouble *a, const fk_cdouble *b,
                     const double *s, fk_cdouble *out, int32_t on_device);

/* Diagnostic with no counterpart in the reference: the raw accumulator of
 * one 16x16x4 MFMA (kind 0: FP64, kind 1: FP32) on fixed operands
 * A[i][k] = 1 + i + 16k, B[k][j] = 1 + 3j - k, where lane l supplies
 * A[l%16][l/16] and B[l/16][l%16]; out[4*lane + r] = accumulator register r
 * of lane `lane` (256 doubles, host memory).  Tests compare it with the
 * result layout the GEMM kernels assume. */
int fk_mfma_probe(int32_t kind, double *out);

#ifdef __cplusplus
}
#endif
#endif /* FORTKERNELS_H */
