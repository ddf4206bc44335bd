// fk_gpp_sym.hip -- the gppsum = 1 form of the GPP band loop on MI355X.
//
// gppKerSeq.cpp (the reference's sequential program, fifth argument gppsum)
// sums, for gppsum = 1, per band n1, column my_igp (igp from the column map)
// and frequency, only over ig <= igp, and adds beside every term m1 x its
// mirror m2 conj(x), m1 = aqsn[n1][ig] conj(aqsm[n1][igp]),
// m2 = conj(aqsm[n1][ig]) aqsm[n1][igp]:
//   occupied bands (flagOCC_solver :90-139): ig < igp, stopping at the first
//     ig with |wtilde^2 I_eps| < to1; near-pole (|wx - wtilde| < gamma or
//     |delw|^2 < to1, delw = (wx - wtilde) (0.5 / wtilde)): sch = 0,
//     ssx = -Omega2 / (4 wtilde^2 (1 + delw)); else sch = wtilde I_eps /
//     (wx - wtilde), ssx = Omega2 / (wx^2 - wtilde^2); the mirror is left
//     out at ig = igp - 1;
//   unoccupied bands (noflagOCC_solver :206-242, called per 512-wide block,
//     :438-450): every ig < igp - 1 except ig = 511 mod 512, plus the
//     diagonal ig = igp (m1 term only) when igp is a positive multiple of
//     512; only the near-pole terms are kept (|wdiff|^2 < limittwo or
//     |delw|^2 > limitone, delw = wtilde / wdiff), sch = delw I_eps.
//
// Summed over bands the m1 and m2 factors are two GEMMs with one left
// operand, c[n1][my_igp] = 0.5 vcoul[igp] conj(aqsm[n1][igp]) (the
// coefficient matrix of the canonical path):
//   Bu  = sum_{n1 >= nv} c aqsn[n1][ig],   Bu' = sum_{n1 >= nv} c aqsm[n1][ig],
// and m2-sums are conj(Bu'), so a term is Bu x + conj(Bu' x).  The kernel
// runs both on FP64 MFMA over the unoccupied bands of a 64 x 64 tile of
// (my_igp, ig) (a tile with no ig <= igp of its rows is skipped), forms the
// occupied-band sums Bo, Bo' directly in the epilogue (nvband is small) and
// evaluates the factors per element and frequency in FP64.  Complex
// division is Smith's algorithm (R. L. Smith, "Algorithm 116: Complex
// division", CACM 5(8), 1962), the order of libgcc's __divdc3 that the
// reference's std::complex<double> calls, with IEEE quotients and no
// contraction; |z| is hypot.  Per-tile partials go to the reduction of the
// canonical path.
#include "fk_complex.h"
#include "fk_gpp_common.h"
#include "fk_internal.h"

#include <climits>

namespace fk {
namespace {

using namespace gpp;

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kSymTile = 64;     // workgroup tile: 2 x 2 waves of 32 x 32
constexpr int kSymNwc = 4;       // frequencies per epilogue pass

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// -a * b + c (the neg modifier on A)
__device__ __forceinline__ d4 mfma_nega(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

// std::complex<double> product and quotient as the reference rounds them.
__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
#pragma clang fp contract(off)
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 zdiv(double2 a, double2 b)
{
#pragma clang fp contract(off)
    double ratio, denom;
    if (fabs(b.x) < fabs(b.y)) {
        ratio = b.x / b.y;
        denom = b.x * ratio + b.y;
        return make_double2((a.x * ratio + a.y) / denom, (a.y * ratio - a.x) / denom);
    }
    ratio = b.y / b.x;
    denom = b.y * ratio + b.x;
    return make_double2((a.y * ratio + a.x) / denom, (a.y - a.x * ratio) / denom);
}

__device__ __forceinline__ double znorm2(double2 a)
{
#pragma clang fp contract(off)
    return a.x * a.x + a.y * a.y;
}

// ig of the unoccupied sum of column igp: 1 both terms, 2 diagonal, 0 none
__device__ __forceinline__ int unocc_kind(int ig, int igp)
{
    if (ig == igp) return (igp > 0 && (igp & 511) == 0) ? 2 : 0;
    if (ig >= igp - 1) return 0;
    return (ig & 511) == 511 ? 0 : 1;
}

// stop[m] = the first ig < igp with |wtilde^2 I_eps| < to1, or igp (the
// occupied loop's break, gppKerSeq.cpp:99); 0 for a skipped column.
__global__ __launch_bounds__(256) void gpp_sym_stop_kernel(const double2 *wtilde, const double2 *ieps, const int *igp,
                                                           int C, double to1, int *stop)
{
    __shared__ int found;
    const int m = blockIdx.x;
    const int g = igp[m];
    if (threadIdx.x == 0) found = INT_MAX;
    __syncthreads();
    const double2 *wr = wtilde + (size_t)m * C, *er = ieps + (size_t)m * C;
    for (int base = 0; base < g; base += 256) {
        const int ig = base + (int)threadIdx.x;
        if (ig < g) {
            const double2 w = wr[ig];
            const double2 om = zmul(zmul(w, w), er[ig]);
            if (hypot(om.x, om.y) < to1) atomicMin(&found, ig);
        }
        __syncthreads();
        // every thread reads `found` between the same two barriers: the
        // decision to stop is uniform
        const bool done = found != INT_MAX;
        __syncthreads();
        if (done) break;
    }
    if (threadIdx.x == 0) stop[m] = g < 0 ? 0 : min(found, g);
}

struct SymArgs {
    const double2 *c;        // [Kpad][ldc]: occupied rows 0..nv-1, zero rows to nvpad, unoccupied rows
    const double2 *aqsn;     // [nb][C]
    const double2 *aqsm;     // [nb][C]
    const double2 *I_eps;    // [G][C]
    const double2 *wtilde;   // [G][C]
    const int *igp;          // [G], -1 = skipped column
    const int *stop;         // [G] (read when nv > 0)
    double2 *partials;       // [tiles][nw][2]
    int nb, nv, C, G, nw, ldc, nvpad, nks, tiles_m, tiles_n;
    double limittwo, limitone, to1, gamma;
    double wx[kMaxFreq];
};

__device__ __forceinline__ void sym_tile_coords(int b, int tiles_m, int total, int &tm, int &tn)
{
    // tiles of one XCD (every kXcds-th workgroup) are consecutive in tm,
    // so they share an aqsn / aqsm column block in that XCD's L2
    const int q = total / kXcds, r = total % kXcds, x = b % kXcds, i = b / kXcds;
    const int t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    tm = t % tiles_m;
    tn = t / tiles_m;
}

__global__ __launch_bounds__(256, 1) void gpp_sym_kernel(SymArgs a)
{
    constexpr int MB = 2, NB = 2, WM = 32, WN = 32, NWAVES = 4;
    __shared__ double red[NWAVES][4 * kSymNwc];
    __shared__ int s_maxg;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    sym_tile_coords(blockIdx.x, a.tiles_m, a.tiles_m * a.tiles_n, tm, tn);
    const int m0 = tm * kSymTile, n0 = tn * kSymTile;
    const int rowbase = m0 + wm * WM, colbase = n0 + wn * WN;
    const int kl = lane >> 4, cl = lane & 15;
    const size_t tile = (size_t)tm + (size_t)tn * a.tiles_m;

    if (tid == 0) s_maxg = -1;
    __syncthreads();
    if (tid < kSymTile && m0 + tid < a.G) atomicMax(&s_maxg, a.igp[m0 + tid]);
    __syncthreads();
    if (n0 > s_maxg) {   // no ig <= igp in this tile: its partials are zero
        for (int k = tid; k < 4 * a.nw; k += 256)
            reinterpret_cast<double *>(a.partials + tile * a.nw * 2)[k] = 0.0;
        return;
    }

    // ---- GEMM over the unoccupied bands: Bu (right operand aqsn), Bu' (aqsm)
    d4 acc_re[MB][NB], acc_im[MB][NB], acc2_re[MB][NB], acc2_im[MB][NB];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            acc_re[i][j] = d4{0, 0, 0, 0};
            acc_im[i][j] = d4{0, 0, 0, 0};
            acc2_re[i][j] = d4{0, 0, 0, 0};
            acc2_im[i][j] = d4{0, 0, 0, 0};
        }
    const int nmax = a.C - 1;
    int colj[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) colj[j] = min(colbase + j * 16 + cl, nmax);
    const double2 *pa = a.c + (size_t)kl * a.ldc + rowbase + cl;
    struct Fr { double2 a[MB], b[NB], b2[NB]; };
    auto load = [&](Fr &f, int s) {
        s = min(s, a.nks - 1);
        const double2 *qa = pa + (size_t)(4 * s) * a.ldc;
#pragma unroll
        for (int i = 0; i < MB; ++i) f.a[i] = qa[i * 16];
        // coefficient row 4s + kl holds band kr - nvpad + nv (a zero row
        // past the last band makes the clamped band irrelevant)
        const int kr = 4 * s + kl;
        const int band = min(max(kr - a.nvpad + a.nv, 0), a.nb - 1);
        const double2 *qn = a.aqsn + (size_t)band * a.C, *qm = a.aqsm + (size_t)band * a.C;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            f.b[j] = qn[colj[j]];
            f.b2[j] = qm[colj[j]];
        }
    };
    auto mma = [&](const Fr &f) {
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                acc_re[i][j] = mfma(f.a[i].x, f.b[j].x, acc_re[i][j]);
                acc_im[i][j] = mfma(f.a[i].x, f.b[j].y, acc_im[i][j]);
                acc2_re[i][j] = mfma(f.a[i].x, f.b2[j].x, acc2_re[i][j]);
                acc2_im[i][j] = mfma(f.a[i].x, f.b2[j].y, acc2_im[i][j]);
            }
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                acc_re[i][j] = mfma_nega(f.a[i].y, f.b[j].y, acc_re[i][j]);
                acc_im[i][j] = mfma(f.a[i].y, f.b[j].x, acc_im[i][j]);
                acc2_re[i][j] = mfma_nega(f.a[i].y, f.b2[j].y, acc2_re[i][j]);
                acc2_im[i][j] = mfma(f.a[i].y, f.b2[j].x, acc2_im[i][j]);
            }
    };
    const int s_begin = a.nvpad / 4;
    if (s_begin < a.nks) {
        Fr f0, f1;
        load(f0, s_begin);
        load(f1, s_begin + 1);
        int s = s_begin;
        for (; s + 1 < a.nks; s += 2) {
            mma(f0);
            load(f0, s + 2);
            mma(f1);
            load(f1, s + 3);
        }
        if (s < a.nks) mma(f0);
    }

    // ---- epilogue ----------------------------------------------------------
    for (int w0 = 0; w0 < a.nw; w0 += kSymNwc) {
        double2 sach[kSymNwc], sasx[kSymNwc];
#pragma unroll
        for (int q = 0; q < kSymNwc; ++q) { sach[q] = make_double2(0.0, 0.0); sasx[q] = make_double2(0.0, 0.0); }
        // The 16 elements of a lane in a runtime loop (the body is too large
        // to repeat 16 times); a wave-uniform switch picks the element's
        // accumulator registers with constant indices.
#pragma unroll 1
        for (int el = 0; el < MB * NB * 4; ++el) {
                    const int i = el >> 3, j = (el >> 2) & 1, r = el & 3;
                    double2 Bu = make_double2(0.0, 0.0), Bp = make_double2(0.0, 0.0);
                    switch (el) {
#define FK_SYM_PICK(I, J, R)                                                        \
    case (I) * 8 + (J) * 4 + (R):                                                   \
        Bu = make_double2(acc_re[I][J][R], acc_im[I][J][R]);                        \
        Bp = make_double2(acc2_re[I][J][R], acc2_im[I][J][R]);                      \
        break;
                        FK_SYM_PICK(0, 0, 0) FK_SYM_PICK(0, 0, 1) FK_SYM_PICK(0, 0, 2) FK_SYM_PICK(0, 0, 3)
                        FK_SYM_PICK(0, 1, 0) FK_SYM_PICK(0, 1, 1) FK_SYM_PICK(0, 1, 2) FK_SYM_PICK(0, 1, 3)
                        FK_SYM_PICK(1, 0, 0) FK_SYM_PICK(1, 0, 1) FK_SYM_PICK(1, 0, 2) FK_SYM_PICK(1, 0, 3)
                        FK_SYM_PICK(1, 1, 0) FK_SYM_PICK(1, 1, 1) FK_SYM_PICK(1, 1, 2) FK_SYM_PICK(1, 1, 3)
#undef FK_SYM_PICK
                    default: break;
                    }
                    const int row = rowbase + i * 16 + 4 * r + (lane >> 4);
                    const int col = colbase + j * 16 + cl;
                    const int g = (row < a.G && col <= nmax) ? a.igp[row] : -1;
                    const int kind = g >= 0 ? unocc_kind(col, g) : 0;
                    const bool occ = g >= 0 && a.nv > 0 && col < a.stop[row];
                    if (kind || occ) {
                    const size_t e = (size_t)row * a.C + col;
                    const double2 wt = a.wtilde[e], ie = a.I_eps[e];
                    double2 Bo = make_double2(0.0, 0.0), Bpo = make_double2(0.0, 0.0);
                    if (occ) {
                        for (int n1 = 0; n1 < a.nv; ++n1) {
                            const double2 cn = a.c[(size_t)n1 * a.ldc + row];
                            Bo = cadd(Bo, cmul(cn, a.aqsn[(size_t)n1 * a.C + col]));
                            Bpo = cadd(Bpo, cmul(cn, a.aqsm[(size_t)n1 * a.C + col]));
                        }
                    }
                    const double2 wt2 = zmul(wt, wt);
                    const double2 Om = zmul(wt2, ie);
                    const double2 half = occ ? zdiv(make_double2(0.5, 0.0), wt) : make_double2(0.0, 0.0);
                    const bool mirror = col != g - 1;
#pragma unroll
                    for (int q = 0; q < kSymNwc; ++q) {
                        const int iw = min(w0 + q, a.nw - 1);
                        const bool live = w0 + q < a.nw;
                        const double wxt = a.wx[iw];
                        const double2 wdiff = make_double2(wxt - wt.x, -wt.y);
                        if (live && kind) {
                            const double2 delw = zdiv(wt, wdiff);
                            const double delw2 = znorm2(delw), wdiffr = znorm2(wdiff);
                            if (fabs(wdiffr) < a.limittwo || delw2 > a.limitone) {
                                const double2 sch = zmul(delw, ie);
                                double2 t = cmul(Bu, sch);
                                if (kind == 1) t = cadd(t, cconj(cmul(Bp, sch)));
                                sach[q] = cadd(sach[q], t);
                            }
                        }
                        if (live && occ) {
                            const double2 delw = zmul(wdiff, half);
                            const double ad = hypot(delw.x, delw.y);
                            double2 sch = make_double2(0.0, 0.0), ssx = make_double2(0.0, 0.0);
                            if (hypot(wdiff.x, wdiff.y) < a.gamma || ad * ad < a.to1) {
                                if (hypot(wt.x, wt.y) > a.to1) {
                                    const double2 den = zmul(make_double2(4.0 * wt2.x, 4.0 * wt2.y),
                                                             make_double2(1.0 + delw.x, delw.y));
                                    ssx = zdiv(make_double2(-Om.x, -Om.y), den);
                                }
                            } else {
                                sch = zdiv(zmul(wt, ie), wdiff);
                                ssx = zdiv(Om, make_double2(wxt * wxt - wt2.x, -wt2.y));
                            }
                            double2 th = cmul(Bo, sch), ts = cmul(Bo, ssx);
                            if (mirror) {
                                th = cadd(th, cconj(cmul(Bpo, sch)));
                                ts = cadd(ts, cconj(cmul(Bpo, ssx)));
                            }
                            sach[q] = cadd(sach[q], th);
                            sasx[q] = cadd(sasx[q], ts);
                        }
                    }
                    }
                }
        // wave reduction, then a fixed-order sum over the waves
#pragma unroll
        for (int q = 0; q < kSymNwc; ++q) {
            const double v0 = wave_sum(sach[q].x), v1 = wave_sum(sach[q].y);
            const double v2 = wave_sum(sasx[q].x), v3 = wave_sum(sasx[q].y);
            if (lane == 0) {
                red[wave][4 * q + 0] = v0;
                red[wave][4 * q + 1] = v1;
                red[wave][4 * q + 2] = v2;
                red[wave][4 * q + 3] = v3;
            }
        }
        __syncthreads();
        if (tid < 4 * kSymNwc) {
            const int q = tid >> 2, comp = tid & 3;
            const int iw = w0 + q;
            if (iw < a.nw) {
                double sum = red[0][tid];
#pragma unroll
                for (int w = 1; w < NWAVES; ++w) sum += red[w][tid];
                reinterpret_cast<double *>(a.partials + (tile * a.nw + iw) * 2)[comp] = sum;
            }
        }
        __syncthreads();
    }
}

}  // namespace

int gpp_sym_launch(const fk_gpp_problem *p, const double2 *c, int ldc, int Kpad, int nvpad, const int *d_igp,
                   int *d_stop, const double2 *aqsm, const double2 *aqsn, const double2 *ieps, const double2 *wt,
                   double2 *partials, int tiles_m, int tiles_n, float *kernel_ms, hipStream_t s)
{
    if (p->nvband > 0) {
        hipLaunchKernelGGL(gpp_sym_stop_kernel, dim3(p->ngpown), dim3(256), 0, s, wt, ieps, d_igp, p->ncouls, p->to1,
                           d_stop);
        FK_HIP(hipGetLastError());
    }
    SymArgs a;
    a.c = c; a.aqsn = aqsn; a.aqsm = aqsm; a.I_eps = ieps; a.wtilde = wt; a.igp = d_igp; a.stop = d_stop;
    a.partials = partials;
    a.nb = p->number_bands; a.nv = p->nvband; a.C = p->ncouls; a.G = p->ngpown; a.nw = p->nfreq;
    a.ldc = ldc; a.nvpad = nvpad; a.nks = Kpad / 4; a.tiles_m = tiles_m; a.tiles_n = tiles_n;
    a.limittwo = p->limittwo; a.limitone = p->limitone; a.to1 = p->to1; a.gamma = std::sqrt(p->limittwo);
    for (int iw = 0; iw < kMaxFreq; ++iw) a.wx[iw] = iw < p->nfreq ? p->wx[iw] : 0.0;
    KernelTimer timer(kernel_ms, s);
    FK_TRY(timer.start());
    hipLaunchKernelGGL(gpp_sym_kernel, dim3(tiles_m * tiles_n), dim3(256), 0, s, a);
    FK_HIP(hipGetLastError());
    FK_TRY(timer.stop());
    return timer.finish();
}

}  // namespace fk
