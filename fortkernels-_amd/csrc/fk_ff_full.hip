// fk_ff_full.hip -- the whole FF self-energy evaluation on MI355X (gfx950).
//
// The Fortran program ffkernel.deslippe.best.f90 evaluates, per band n1 and
// evaluation frequency wx = freqevalmin - ekq(n1) + (iw-1) freqevalstep:
//   SX (:278-386)       occupied bands, sum_{my_igp,ig} vcoul aqsn conj(aqsm)
//                       times I_epsR (wx >= 0) or I_epsA (wx < 0) linearly
//                       interpolated between the grid points around |wx|;
//   sigma 1-4 (:458-517) on schdt_matrix(ifreq, n1), the I_epsR contraction
//                       of ffkernel.deslippe.best.f90.back1:402-418;
//   delta terms (:529-613) on I_epsR - I_epsA interpolated the same way.
// Every (my_igp, ig) sum there is linear in I_eps, so with the contractions
//   S[n1][f]  = sum_my_igp conj(aqsm[n1][igp]) sum_ig aqsn[n1][ig] I_epsR[f][my_igp][ig]
//   VR, VA    = the same weighted by vcoul[igp], for I_epsR and I_epsA
// (ff_kernel's schDt / schDv, two GEMMs on MFMA) the interpolated sums are
// fa VX[f] + fb VX[f+1], and the rest is O(nb * nfreqeval * nFreq) scalar
// work: one thread per (n1, iw) walks the grid and adds its terms, a
// workgroup sums 256 bands in a fixed tree, and a last kernel sums the
// workgroups in order.  I_epsA is contracted only over the bands and the
// frequencies the SX and delta terms read (the host walks the same
// evaluation axis and interval search the device does).
//
// wx is computed on the host without contraction into FMA, as the Fortran
// evaluates it, and handed to the device, so that the interval search and the
// sign tests see the very values the Fortran sees.
//
// FK_FF_CPP_PORT / FK_FF_CPP_PORT_ZERO_SCHDT evaluate instead what the C++
// port FFKer_cori/ffKernel.cpp computes as its main calls it (:485-500): the
// same contractions, a second form of the per-(n1, iw) kernel (cpp_terms)
// and its own host tables and I_epsA box.
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>
#include <cmath>

namespace fk {
namespace {

constexpr int kEpiThreads = 256;
constexpr int kTerms = 5;   // asxDtemp, achDtemp, achDtemp_cor, achDtemp_corb, ach2Dtemp

struct EpiArgs {
    const double2 *S;       // [nb][nF]   schdt_matrix (no vcoul)
    const double2 *VR;      // [nb][nF]   vcoul-weighted, I_epsR
    const double2 *VA;      // [nbA][nFA] vcoul-weighted, I_epsA, bands bA0.., frequencies fA0..
    // Fortran: wx = freqevalmin - ekq[n1] + iw freqevalstep; wxp the same
    // with -freqevalmin (:509).  C++ port: wx = freqevalmin - ekq[n1] +
    // freqevalstep for every iw (:140, :224), wxp = freqevalmin - ekq[n1] +
    // (iw-1) freqevalstep (:202), w2[iw] = int(freqevalmin) - ekq[nb-1] +
    // (iw-1) int(freqevalstep) (:10-12, the last band only).
    const double *wx;       // [nb][ne]
    const double *wxp;      // [nb][ne]
    const double *w2;       // [ne]
    const double *grid;     // [nF]
    const double *pref;     // [nF]
    const double2 *brd;     // [nF]
    double2 *part;          // [ne][nchunks][kTerms]
    int nb, nv, nF, ne, nchunks;
    int bA0, nbA, fA0, nFA;
    double occ, pref_zb;
};

// The interval search of :286-297 / :537-546: the last i in [0, nF-2] with
// grid[i] <= w < grid[i+1]; none -> nF-2 (the program's clamp to nFreq-1).
__host__ __device__ inline int find_interval(const double *grid, int nF, double w)
{
    int f = -1;
    for (int i = 0; i < nF - 1; ++i)
        if (w >= grid[i] && w < grid[i + 1]) f = i;
    return f < 0 ? nF - 2 : f;
}

// compute_fact's search (FFKer_cori/ffKernel.cpp:34-38): the last i with
// grid[i] < w < grid[i+1], else 0; the callers add the port's fall-backs.
__host__ __device__ inline int find_interval_strict(const double *grid, int nF, double w)
{
    int f = 0;
    for (int i = 0; i < nF - 1; ++i)
        if (w > grid[i] && w < grid[i + 1]) f = i;
    return f;
}

__device__ inline double clamp_ratio(double x)
{
    if (x < 1e-4) x = 1e-4;
    if (x > 1e4) x = 1e4;
    return x;
}

// i * s * a
__device__ inline double2 imul(double s, double2 a) { return make_double2(-s * a.y, s * a.x); }

__device__ inline double2 va_at(const EpiArgs &a, int n1, int f)
{
    // the host sized [bA0, bA0+nbA) x [fA0, fA0+nFA) to cover every read;
    // the clamps only keep an index in bounds
    const int b = min(max(n1 - a.bA0, 0), a.nbA - 1);
    const int g = min(max(f - a.fA0, 0), a.nFA - 1);
    return a.VA[(size_t)b * a.nFA + g];
}

// The C++ port's terms for one (n1, iw).  SX (asxDtemp_Kernel, :133-153):
// occupied bands, VR only, compute_fact's loop 1 with the nfreq-2 fall-back.
// Sigma 1 (achDtemp_Kernel, :196-206, written into asxDtemp by main's
// aliasing, :495) over every frequency; ZERO_S drops it with schDt_matrix.
// ach2Dtemp: every band adds sch2Di (:239), which holds the sigma-2 value of
// the last band and frequency (reset per term, :191), so the last band's
// thread adds nb copies.  Delta terms (achDtemp_cor_Kernel, :212-246 with
// schDttt_corKernel1/2, :248-300): unoccupied bands at wx > 0 without a
// fall-back (:55-58), occupied bands at wx <= 0 with it, scaled -0.5.
template <bool ZERO_S>
__device__ inline void cpp_terms(const EpiArgs &a, int n1, int iw, double2 *acc)
{
    const double *g = a.grid;
    const int nF = a.nF;
    const bool occb = n1 < a.nv;
    const double wx = a.wx[(size_t)n1 * a.ne + iw];
    const double2 *VR = a.VR + (size_t)n1 * nF;
    const double2 zero = make_double2(0.0, 0.0);
    double2 asx = zero, cor = zero, s2 = zero;
    if (occb) {
        int f = find_interval_strict(g, nF, wx > 0.0 ? wx : -wx);
        if (f == 0) f = nF - 2;
        const double d = g[f + 1] - g[f];
        double fa, fb;
        if (wx > 0.0) { fa = (g[f + 1] - wx) / d; fb = (wx - g[f]) / d; }
        else { fa = (g[f + 1] + wx) / d; fb = (-g[f] - wx) / d; }
        asx = cscale(cadd(cscale(VR[f], fa), cscale(VR[f + 1], fb)), a.occ);
    }
    if (!ZERO_S) {
        const double2 *S = a.S + (size_t)n1 * nF;
        const double w1 = a.wxp[(size_t)n1 * a.ne + iw];
        for (int f = 0; f < nF; ++f) {
            const double2 br = a.brd[f];
            const double2 coh = make_double2(g[f] - br.x, -br.y);
            asx = csub(asx, cdiv(imul(a.pref[f], S[f]), make_double2(w1 - coh.x, -coh.y)));
        }
        if (n1 == a.nb - 1) {
            const double zr = g[nF - 1], zl = g[nF - 2], base = a.w2[iw];
            const double2 avg = cscale(cadd(S[nF - 1], S[nF - 2]), 0.5);
            const double it = -log(clamp_ratio((base - zr) / (base - zl)));
            s2 = cscale(csub(zero, cscale(imul(a.pref_zb, avg), it)), (double)a.nb);
        }
    }
    if (wx > 0.0) {
        if (!occb) {
            const int f = find_interval_strict(g, nF, wx);
            const double d = g[f + 1] - g[f];
            const double fa = -0.5 * (g[f + 1] - wx) / d, fb = -0.5 * (wx - g[f]) / d;
            cor = cadd(cscale(csub(VR[f], va_at(a, n1, f)), fa), cscale(csub(VR[f + 1], va_at(a, n1, f + 1)), fb));
        }
    } else if (occb) {
        const double w = -wx;
        int f = find_interval_strict(g, nF, w);
        if (f == 0) f = nF - 2;
        const double d = g[f + 1] - g[f];
        const double fa = (g[f + 1] - w) / d, fb = (w - g[f]) / d;
        const double2 s = cadd(cscale(csub(VR[f], va_at(a, n1, f)), fa),
                               cscale(csub(VR[f + 1], va_at(a, n1, f + 1)), fb));
        cor = cscale(s, -0.5);
    }
    acc[0] = asx; acc[2] = cor; acc[4] = s2;   // achDtemp, achDtemp_corb: zero
}

// MODE: fk_ff_semantics
template <int MODE>
__global__ __launch_bounds__(kEpiThreads) void ff_epilogue_kernel(EpiArgs a)
{
    const int tid = threadIdx.x;
    const int iw = blockIdx.y;
    const int n1 = blockIdx.x * kEpiThreads + tid;
    double2 acc[kTerms];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) acc[k] = make_double2(0.0, 0.0);
    if (MODE != FK_FF_FORTRAN) {
        if (n1 < a.nb) cpp_terms<MODE == FK_FF_CPP_PORT_ZERO_SCHDT>(a, n1, iw, acc);
    } else if (n1 < a.nb) {
        const bool occb = n1 < a.nv;
        const double wx = a.wx[(size_t)n1 * a.ne + iw];
        const double *g = a.grid;
        const double2 *S = a.S + (size_t)n1 * a.nF;
        const double2 *VR = a.VR + (size_t)n1 * a.nF;
        double2 asx = make_double2(0.0, 0.0), ach = asx, cor = asx, corb = asx, s2 = asx;
        // SX, :278-386
        if (occb) {
            if (wx >= 0.0) {
                const int f = find_interval(g, a.nF, wx);
                const double d = g[f + 1] - g[f];
                const double fa = (g[f + 1] - wx) / d, fb = (wx - g[f]) / d;
                asx = cadd(cscale(VR[f], fa), cscale(VR[f + 1], fb));
            } else {
                const int f = find_interval(g, a.nF, -wx);
                const double d = g[f + 1] - g[f];
                const double fa = (g[f + 1] + wx) / d, fb = (-g[f] - wx) / d;
                asx = cadd(cscale(va_at(a, n1, f), fa), cscale(va_at(a, n1, f + 1), fb));
            }
            asx = cscale(asx, a.occ);
        }
        // sigma 1-4, :464-517
        const double wxp = a.wxp[(size_t)n1 * a.ne + iw];
        double2 left = S[0];
        for (int f = 0; f < a.nF; ++f) {
            const double2 sd = S[f];
            const double zb = g[f];
            const double2 br = a.brd[f];
            const double2 coh = make_double2(zb - br.x, -br.y);
            const double2 cre = occb ? make_double2(-zb - br.x, -br.y) : coh;
            if (f != a.nF - 1) {
                const double2 num = imul(a.pref[f], sd);
                ach = csub(ach, cdiv(num, make_double2(wx - coh.x, -coh.y)));
                corb = csub(corb, cdiv(num, make_double2(wx - cre.x, -cre.y)));
            }
            if (f != 0) {
                const double zr = zb, zl = g[f - 1];
                const double2 avg = cscale(cadd(sd, left), 0.5);
                const double2 lin = csub(sd, left);
                const double2 lin2 = cdiv(lin, zr - zl);
                const double it = -log(clamp_ratio(fabs((wx - zr) / (wx - zl))));
                s2 = csub(s2, cscale(imul(a.pref_zb, avg), it));
                double2 lin3;
                if (occb) {
                    const double it2 = log(clamp_ratio(fabs((wx + zr) / (wx + zl))));
                    lin3 = cscale(cadd(left, cscale(lin2, wxp - zl)), it2);
                } else {
                    lin3 = cscale(cadd(left, cscale(lin2, wx - zl)), it);
                }
                lin3 = cadd(lin3, lin);
                cor = csub(cor, imul(a.pref_zb, lin3));
            }
            left = sd;
        }
        // delta-function terms, :529-613
        if (wx >= 0.0) {
            const int f = find_interval(g, a.nF, wx);
            const double d = g[f + 1] - g[f];
            const double fa = -0.5 * (g[f + 1] - wx) / d, fb = -0.5 * (wx - g[f]) / d;
            const double2 s = cadd(cscale(csub(VR[f], va_at(a, n1, f)), fa),
                                   cscale(csub(VR[f + 1], va_at(a, n1, f + 1)), fb));
            s2 = cadd(s2, s);
            if (!occb) cor = cadd(cor, s);
        } else if (occb) {
            const double w = -wx;
            const int f = find_interval(g, a.nF, w);
            const double d = g[f + 1] - g[f];
            const double fa = (g[f + 1] - w) / d, fb = (w - g[f]) / d;
            const double2 s = cadd(cscale(csub(VR[f], va_at(a, n1, f)), fa),
                                   cscale(csub(VR[f + 1], va_at(a, n1, f + 1)), fb));
            cor = cadd(cor, cscale(s, -0.5));
        }
        acc[0] = asx; acc[1] = ach; acc[2] = cor; acc[3] = corb; acc[4] = s2;
    }
    // fixed-order tree over the workgroup's bands
    __shared__ double2 red[kTerms][kEpiThreads];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) red[k][tid] = acc[k];
    __syncthreads();
    for (int w = kEpiThreads / 2; w >= 1; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < kTerms; ++k) red[k][tid] = cadd(red[k][tid], red[k][tid + w]);
        }
        __syncthreads();
    }
    if (tid < kTerms) a.part[((size_t)iw * a.nchunks + blockIdx.x) * kTerms + tid] = red[tid][0];
}

// out_k[iw] = sum over the workgroups in order
__global__ void ff_epilogue_sum_kernel(const double2 *part, int nchunks, int ne, double2 *asx, double2 *ach,
                                       double2 *cor, double2 *corb, double2 *a2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * kTerms) return;
    const int iw = i / kTerms, k = i % kTerms;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < nchunks; ++c) s = cadd(s, part[((size_t)iw * nchunks + c) * kTerms + k]);
    double2 *dst[kTerms] = {asx, ach, cor, corb, a2};
    dst[k][iw] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int full_check(const fk_ff_full_problem *p, const fk_options *o, const fk_ff_full_outputs *out)
{
    if (!p || !o) { set_error("ff_kernel_full: NULL problem or options"); return FK_ERR_ARG; }
    if (p->number_bands < 1 || p->ncouls < 1 || p->ngpown < 1 || p->nfreq < 2 || p->nfreqeval < 1 ||
        p->nfreqeval > 65535 || p->nvband < 0 || p->nvband > p->number_bands) {
        set_error("ff_kernel_full: bad sizes (nb=%d nv=%d ncouls=%d ngpown=%d nfreq=%d nfreqeval=%d; need nb, "
                  "ncouls, ngpown >= 1, 1 <= nfreqeval <= 65535, nfreq >= 2, 0 <= nv <= nb)",
                  p->number_bands, p->nvband, p->ncouls, p->ngpown, p->nfreq, p->nfreqeval);
        return FK_ERR_ARG;
    }
    if (p->semantics < FK_FF_FORTRAN || p->semantics > FK_FF_CPP_PORT_ZERO_SCHDT) {
        set_error("ff_kernel_full: unknown semantics %d", p->semantics);
        return FK_ERR_ARG;
    }
    if (p->semantics != FK_FF_FORTRAN) {
        // the port takes freqevalmin and freqevalstep as int (:10) and forms
        // (iw-1) * freqevalstep in int
        const double lim = 2147483647.0;
        const bool ok = std::fabs(p->freqevalmin) < lim && std::fabs(p->freqevalstep) < lim &&
                        std::fabs((double)(int)p->freqevalstep) * std::max(1, p->nfreqeval - 2) < lim;
        if (!ok) {
            set_error("ff_kernel_full: the C++ port's semantics need int(freqevalmin), int(freqevalstep) and "
                      "(nfreqeval-2) * int(freqevalstep) within int32");
            return FK_ERR_ARG;
        }
    }
    if (!p->I_epsA_array || !p->ekq || !p->dFreqGrid || !p->pref || !p->dFreqBrd) {
        set_error("ff_kernel_full: NULL input array");
        return FK_ERR_ARG;
    }
    for (int f = 0; f + 1 < p->nfreq; ++f)
        if (!(p->dFreqGrid[f] < p->dFreqGrid[f + 1])) {
            set_error("ff_kernel_full: dFreqGrid must be strictly increasing (index %d)", f);
            return FK_ERR_ARG;
        }
    if (out && (!out->asxDtemp || !out->achDtemp || !out->achDtemp_cor || !out->achDtemp_corb || !out->ach2Dtemp)) {
        set_error("ff_kernel_full: NULL output array");
        return FK_ERR_ARG;
    }
    return FK_OK;
}

fk_ff_problem sub_problem(const fk_ff_full_problem *p, int nb, int nf, const void *aqsm, const void *aqsn,
                          const void *X, int achs_freq)
{
    fk_ff_problem q;
    q.number_bands = nb; q.ncouls = p->ncouls; q.ngpown = p->ngpown; q.nfreq = nf;
    q.aqsmtemp = aqsm; q.aqsntemp = aqsn; q.I_eps_array = X; q.vcoul = p->vcoul;
    q.inv_igp_index = p->inv_igp_index; q.indinv = p->indinv; q.achs_freq = achs_freq;
    return q;
}

struct FullPlan {
    int nb, nF, ne, nchunks;
    bool haveA;
    int bA0, nbA, fA0, nFA;
    std::vector<double> tab;            // wx, wxp, grid, pref, brd (re, im), w2
    FfLayout LR, LA;
    fk_options oA;                      // options of the I_epsA contraction
    size_t off_tab, off_S, off_VR, off_VA, off_part;
};

// The evaluation axes as the Fortran (or the C++ port) computes them, and the
// bands / frequencies of I_epsA the SX and delta terms read.
void full_axes(const fk_ff_full_problem *p, FullPlan &P)
{
#pragma clang fp contract(off)
    const int nb = p->number_bands, ne = p->nfreqeval, nF = p->nfreq;
    const bool port = p->semantics != FK_FF_FORTRAN;
    P.nb = nb; P.nF = nF; P.ne = ne;
    P.tab.assign((size_t)2 * nb * ne + 4 * (size_t)nF + ne, 0.0);
    double *wx = P.tab.data(), *wxp = wx + (size_t)nb * ne;
    int b_lo = nb, b_hi = -1, f_lo = nF, f_hi = -1;
    for (int n1 = 0; n1 < nb; ++n1) {
        const bool occb = n1 < p->nvband;
        for (int iw = 0; iw < ne; ++iw) {
            double w;
            int f = -1;
            if (!port) {
                w = p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;
                wxp[(size_t)n1 * ne + iw] = -p->freqevalmin - p->ekq[n1] + (double)iw * p->freqevalstep;
                if (w >= 0.0) f = find_interval(p->dFreqGrid, nF, w);          // delta terms
                else if (occb) f = find_interval(p->dFreqGrid, nF, -w);        // SX and delta terms
            } else {
                w = p->freqevalmin - p->ekq[n1] + p->freqevalstep;                // :140, :224
                wxp[(size_t)n1 * ne + iw] = p->freqevalmin - p->ekq[n1] + (double)(iw - 1) * p->freqevalstep;
                if (w > 0.0) {
                    if (!occb) f = find_interval_strict(p->dFreqGrid, nF, w);    // delta terms only
                } else if (occb) {
                    f = find_interval_strict(p->dFreqGrid, nF, -w);
                    if (f == 0) f = nF - 2;
                }
            }
            wx[(size_t)n1 * ne + iw] = w;
            if (f >= 0) {
                b_lo = std::min(b_lo, n1); b_hi = std::max(b_hi, n1);
                f_lo = std::min(f_lo, f); f_hi = std::max(f_hi, f + 1);
            }
        }
    }
    double *w2 = wxp + (size_t)nb * ne + 4 * (size_t)nF;
    if (port) {
        const int fmin = (int)p->freqevalmin, fstep = (int)p->freqevalstep;   // calculate_schDt_lin3, :10-12
        for (int iw = 0; iw < ne; ++iw)
            w2[iw] = (double)fmin - p->ekq[nb - 1] + (double)((iw - 1) * fstep);
    }
    double *t = wxp + (size_t)nb * ne;
    for (int f = 0; f < nF; ++f) {
        t[f] = p->dFreqGrid[f];
        t[nF + f] = p->pref[f];
        t[2 * nF + 2 * f] = p->dFreqBrd[f].re;
        t[2 * nF + 2 * f + 1] = p->dFreqBrd[f].im;
    }
    P.haveA = b_hi >= 0;
    P.bA0 = P.haveA ? b_lo : 0;
    P.nbA = P.haveA ? b_hi - b_lo + 1 : 1;
    P.fA0 = P.haveA ? f_lo : 0;
    P.nFA = P.haveA ? f_hi - f_lo + 1 : 1;
    P.nchunks = div_up(nb, kEpiThreads);
}

// The I_epsA contraction covers few bands at the program's axes (15 of
// 256): up to 16 bands it runs on the FP64 GEMM with 16-row tiles in either
// precision (exact, and 4x less MFMA work than a 64-row tile, 16x less than
// the split path's 256-row tile); otherwise as the caller asked.
fk_options a_options(const fk_options *o, int nbA)
{
    fk_options oa = *o;
    oa.kernel_ms = nullptr;
    if (nbA <= 16) {
        oa.precision = FK_PREC_FP64;
        oa.config = kFfConfigFewBands;
    }
    return oa;
}

void full_plan(const fk_ff_full_problem *p, const fk_options *o, Scratch &ws, FullPlan &P)
{
    full_axes(p, P);
    const fk_ff_problem pr = sub_problem(p, P.nb, P.nF, nullptr, nullptr, nullptr, p->achs_freq);
    P.LR = ff_plan(&pr, o, ws);
    if (P.haveA) {
        const fk_ff_problem pa = sub_problem(p, P.nbA, P.nFA, nullptr, nullptr, nullptr, 0);
        P.oA = a_options(o, P.nbA);
        P.LA = ff_plan(&pa, &P.oA, ws);
    }
    P.off_tab = ws.plan(sizeof(double) * P.tab.size());
    P.off_S = ws.plan(sizeof(double2) * (size_t)P.nb * P.nF);
    P.off_VR = ws.plan(sizeof(double2) * (size_t)P.nb * P.nF);
    P.off_VA = ws.plan(sizeof(double2) * (size_t)P.nbA * P.nFA);
    P.off_part = ws.plan(sizeof(double2) * (size_t)P.ne * P.nchunks * kTerms);
}

// Device pointers of the inputs (I_epsA: its first plane is frequency fA0 when
// A_is_slice, else frequency 0) and of the outputs.
struct FullIo {
    const double2 *aqsm, *aqsn, *IR, *IA;
    const double *vcoul;
    bool A_is_slice;
    double2 *achs, *asx, *ach, *cor, *corb, *a2, *schDt;
};

int full_launch(const fk_ff_full_problem *p, const fk_options *o, const FullPlan &P, Scratch &ws,
                const std::vector<int32_t> &cols, const FullIo &io, hipStream_t s)
{
    fk_options oi = *o;
    oi.kernel_ms = nullptr;
    const size_t C = (size_t)p->ncouls, GC = (size_t)p->ngpown * C;
    double *tab = static_cast<double *>(ws.at(P.off_tab));
    double2 *S = io.schDt ? io.schDt : static_cast<double2 *>(ws.at(P.off_S));
    double2 *VR = static_cast<double2 *>(ws.at(P.off_VR));
    double2 *VA = static_cast<double2 *>(ws.at(P.off_VA));
    double2 *part = static_cast<double2 *>(ws.at(P.off_part));
    KernelTimer timer(o->kernel_ms, s);
    FK_HIP(hipMemcpyAsync(tab, P.tab.data(), sizeof(double) * P.tab.size(), hipMemcpyHostToDevice, s));
    FK_TRY(timer.start());
    const fk_ff_problem pr = sub_problem(p, P.nb, P.nF, io.aqsm, io.aqsn, io.IR, p->achs_freq);
    FK_TRY(ff_launch(&pr, &oi, P.LR, ws, cols, io.aqsm, io.aqsn, io.IR, io.vcoul, S, VR, io.achs, 1.0, s));
    if (P.haveA) {
        const double2 *am = io.aqsm + (size_t)P.bA0 * C, *an = io.aqsn + (size_t)P.bA0 * C;
        const double2 *X = io.IA + (io.A_is_slice ? 0 : (size_t)P.fA0 * GC);
        const fk_ff_problem pa = sub_problem(p, P.nbA, P.nFA, am, an, X, 0);
        FK_TRY(ff_launch(&pa, &P.oA, P.LA, ws, cols, am, an, X, io.vcoul, nullptr, VA, nullptr, 1.0, s));
    }   // else no (n1, iw) reads VA
    EpiArgs a;
    a.S = S; a.VR = VR; a.VA = VA;
    a.wx = tab; a.wxp = tab + (size_t)P.nb * P.ne;
    a.grid = a.wxp + (size_t)P.nb * P.ne; a.pref = a.grid + P.nF;
    a.brd = reinterpret_cast<const double2 *>(a.pref + P.nF);
    a.w2 = a.pref + 3 * (size_t)P.nF;
    a.part = part;
    a.nb = P.nb; a.nv = p->nvband; a.nF = P.nF; a.ne = P.ne; a.nchunks = P.nchunks;
    a.bA0 = P.bA0; a.nbA = P.nbA; a.fA0 = P.fA0; a.nFA = P.nFA;
    a.occ = p->occ; a.pref_zb = p->pref_zb;
    const dim3 grid(P.nchunks, P.ne);
    switch (p->semantics) {
    case FK_FF_CPP_PORT: hipLaunchKernelGGL(ff_epilogue_kernel<FK_FF_CPP_PORT>, grid, dim3(kEpiThreads), 0, s, a); break;
    case FK_FF_CPP_PORT_ZERO_SCHDT:
        hipLaunchKernelGGL(ff_epilogue_kernel<FK_FF_CPP_PORT_ZERO_SCHDT>, grid, dim3(kEpiThreads), 0, s, a);
        break;
    default: hipLaunchKernelGGL(ff_epilogue_kernel<FK_FF_FORTRAN>, grid, dim3(kEpiThreads), 0, s, a); break;
    }
    FK_HIP(hipGetLastError());
    hipLaunchKernelGGL(ff_epilogue_sum_kernel, dim3(div_up(P.ne * kTerms, 64)), dim3(64), 0, s, part, P.nchunks,
                       P.ne, io.asx, io.ach, io.cor, io.corb, io.a2);
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());
    return timer.finish();
}

}  // namespace
}  // namespace fk

using namespace fk;

extern "C" {

size_t ff_kernel_full_workspace_size(const fk_ff_full_problem *p, const fk_options *o)
{
    fk_options d = fk_default_options();
    if (!o) o = &d;
    if (full_check(p, o, nullptr) != FK_OK) return 0;
    const fk_ff_problem pr = sub_problem(p, p->number_bands, p->nfreq, p->aqsmtemp, p->aqsntemp, p->I_epsR_array,
                                         p->achs_freq);
    if (ff_check(&pr, o) != FK_OK) return 0;
    Scratch ws;
    FullPlan P;
    full_plan(p, o, ws, P);
    return ws.total();
}

int ff_kernel_full(const fk_ff_full_problem *p, const fk_options *o, const fk_ff_full_outputs *out)
{
    clear_error();
    FK_TRY(full_check(p, o, out));
    if (!out) { set_error("ff_kernel_full: NULL outputs"); return FK_ERR_ARG; }
    {
        const fk_ff_problem pr = sub_problem(p, p->number_bands, p->nfreq, p->aqsmtemp, p->aqsntemp,
                                             p->I_epsR_array, p->achs_freq);
        FK_TRY(ff_check(&pr, o));
    }
    std::vector<int32_t> cols;
    FK_TRY(resolve_columns(p->ncouls, p->ngpown, p->inv_igp_index, p->indinv, cols));
    int prev_dev = -1;
    if (o->memory == FK_MEM_HOST && o->device >= 0) {
        FK_HIP(hipGetDevice(&prev_dev));
        FK_HIP(hipSetDevice(o->device));
    }
    hipStream_t s = static_cast<hipStream_t>(o->stream);
    int rc = FK_OK;
    {
        Scratch ws;
        FullPlan P;
        full_plan(p, o, ws, P);
        rc = ws.acquire(o->workspace, o->workspace_bytes, s);
        const size_t nbC = (size_t)p->number_bands * p->ncouls, GC = (size_t)p->ngpown * p->ncouls;
        const int ne = p->nfreqeval;
        const size_t nS = (size_t)p->number_bands * p->nfreq;
        if (rc == FK_OK && o->memory == FK_MEM_DEVICE) {
            FullIo io;
            io.aqsm = (const double2 *)p->aqsmtemp; io.aqsn = (const double2 *)p->aqsntemp;
            io.IR = (const double2 *)p->I_epsR_array; io.IA = (const double2 *)p->I_epsA_array;
            io.vcoul = p->vcoul; io.A_is_slice = false;
            io.achs = (double2 *)out->achsDtemp; io.asx = (double2 *)out->asxDtemp; io.ach = (double2 *)out->achDtemp;
            io.cor = (double2 *)out->achDtemp_cor; io.corb = (double2 *)out->achDtemp_corb;
            io.a2 = (double2 *)out->ach2Dtemp; io.schDt = (double2 *)out->schDt;
            rc = full_launch(p, o, P, ws, cols, io, s);
        } else if (rc == FK_OK) {
            // host memory: upload the inputs (I_epsA only over the frequencies
            // read), run, download
            double2 *dm = nullptr, *dn = nullptr, *dR = nullptr, *dA = nullptr, *dout = nullptr;
            double *dv = nullptr;
            auto alloc = [&](void **ptr, size_t bytes) {
                if (rc != FK_OK) return;
                if (hipMalloc(ptr, bytes) != hipSuccess) { set_error("hipMalloc(%zu) failed", bytes); rc = FK_ERR_NOMEM; }
            };
            auto up = [&](void *dst, const void *src, size_t bytes) {
                if (rc != FK_OK) return;
                hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
                if (e != hipSuccess) { set_error("upload: %s", hipGetErrorString(e)); rc = FK_ERR_HIP; }
            };
            const size_t nA = P.haveA ? (size_t)P.nFA * GC : 1;
            const size_t nout = 1 + (size_t)kTerms * ne + nS;
            alloc((void **)&dm, nbC * sizeof(double2));
            alloc((void **)&dn, nbC * sizeof(double2));
            alloc((void **)&dR, (size_t)p->nfreq * GC * sizeof(double2));
            alloc((void **)&dA, nA * sizeof(double2));
            alloc((void **)&dv, sizeof(double) * p->ncouls);
            alloc((void **)&dout, sizeof(double2) * nout);
            up(dm, p->aqsmtemp, nbC * sizeof(double2));
            up(dn, p->aqsntemp, nbC * sizeof(double2));
            up(dR, p->I_epsR_array, (size_t)p->nfreq * GC * sizeof(double2));
            if (P.haveA)
                up(dA, (const double2 *)p->I_epsA_array + (size_t)P.fA0 * GC, nA * sizeof(double2));
            up(dv, p->vcoul, sizeof(double) * p->ncouls);
            if (rc == FK_OK) {
                FullIo io;
                io.aqsm = dm; io.aqsn = dn; io.IR = dR; io.IA = dA; io.vcoul = dv; io.A_is_slice = true;
                io.achs = dout; io.asx = dout + 1; io.ach = io.asx + ne; io.cor = io.ach + ne;
                io.corb = io.cor + ne; io.a2 = io.corb + ne; io.schDt = io.a2 + ne;
                rc = full_launch(p, o, P, ws, cols, io, s);
            }
            if (rc == FK_OK) {
                std::vector<double2> h(nout);
                hipError_t e = hipMemcpyAsync(h.data(), dout, sizeof(double2) * nout, hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) {
                    set_error("download: %s", hipGetErrorString(e));
                    rc = FK_ERR_HIP;
                } else {
                    auto put = [&](fk_cdouble *dst, size_t off, size_t n) {
                        if (dst) std::copy(h.begin() + off, h.begin() + off + n, reinterpret_cast<double2 *>(dst));
                    };
                    put(out->achsDtemp, 0, 1);
                    put(out->asxDtemp, 1, ne);
                    put(out->achDtemp, 1 + ne, ne);
                    put(out->achDtemp_cor, 1 + 2 * (size_t)ne, ne);
                    put(out->achDtemp_corb, 1 + 3 * (size_t)ne, ne);
                    put(out->ach2Dtemp, 1 + 4 * (size_t)ne, ne);
                    put(out->schDt, 1 + 5 * (size_t)ne, nS);
                }
            } else {
                (void)hipStreamSynchronize(s);
            }
            for (void *ptr : {(void *)dm, (void *)dn, (void *)dR, (void *)dA, (void *)dv, (void *)dout})
                if (ptr) (void)hipFree(ptr);
        }
    }
    if (prev_dev >= 0) (void)hipSetDevice(prev_dev);
    return rc;
}

}  // extern "C"
